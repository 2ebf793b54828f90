// hmv_inputs.h -- where a work item's staged inputs come from (device code
// shared by the streaming and the H-matmat kernels).
#ifndef HMV_INPUTS_H
#define HMV_INPUTS_H

#include <hip/hip_runtime.h>
#include <limits.h>

#include "hmv_internal.h"

// Input code of item column c (x index >= 0, or -1 - t slot): from the
// item's run table in LDS when it has several runs (binary search over the
// runs' first columns), from its single run, or from its code table.
struct Inputs {
  HmvZSeg z0;        // the item's first segment
  int nr;            // segments; > 1: runs tabled in rcol/rsrc
  const int *rcol;   // LDS: first column of each run, INT_MAX past the nr runs
  const int *rsrc;   // LDS: code of each run's first column
};
// Code of column c from the run table (several runs).  The table is padded
// to HMV_RUN_MAX entries, so every step reads and selects without a branch
// (r + s <= HMV_RUN_MAX - 1 at every step).
__device__ inline int run_code(const Inputs &in, int c) {
  int r = 0;
#pragma unroll
  for (int s = HMV_RUN_MAX / 2; s > 0; s >>= 1) r += in.rcol[r + s] <= c ? s : 0;
  const int b = in.rsrc[r], d = c - in.rcol[r];
  return b >= 0 ? b + d : b - d;
}
__device__ inline int input_code(const Inputs &in, const int *__restrict__ codes, int c) {
  if (in.z0.kind == HMV_Z_CODES) return codes[in.z0.src + c];
  if (in.nr == 1) return in.z0.kind == HMV_Z_X ? in.z0.src + c : -1 - (in.z0.src + c);
  return run_code(in, c);
}

// Calls f(code_of) once with the item's column-to-code map, so that a loop
// over columns inside f branches on the input kind once, not per column.
template <typename F>
__device__ inline void with_codes(const Inputs &in, const int *__restrict__ codes, F &&f) {
  if (in.z0.kind == HMV_Z_CODES) {
    const int *cb = codes + in.z0.src;
    f([cb](int c) { return cb[c]; });
  } else if (in.nr == 1) {
    const int b = in.z0.kind == HMV_Z_X ? in.z0.src : -1 - in.z0.src;
    f([b](int c) { return b >= 0 ? b + c : b - c; });
  } else {
    f([&in](int c) { return run_code(in, c); });
  }
}

// Tables the run segments of item w in LDS (every thread of the workgroup
// calls it; a barrier follows when there is a table).
template <int NT>
__device__ inline Inputs load_inputs(const HmvWork &w, const HmvZSeg *__restrict__ zsegs,
                                     int *__restrict__ runs) {
  static_assert(HMV_RUN_MAX <= NT, "one thread per run");
  Inputs in;
  in.nr = w.zseg_end - w.zseg_begin;
  in.z0 = in.nr > 0 ? zsegs[w.zseg_begin] : HmvZSeg{HMV_Z_X, 0, 0, 0};
  in.rcol = runs;
  in.rsrc = runs + HMV_RUN_MAX;
  if (in.nr > 1) {
    const int tid = threadIdx.x;
    if (tid < in.nr) {
      const HmvZSeg zs = zsegs[w.zseg_begin + tid];
      runs[tid] = zs.col;
      runs[HMV_RUN_MAX + tid] = zs.kind == HMV_Z_X ? zs.src : -1 - zs.src;
    } else if (tid < HMV_RUN_MAX) {
      runs[tid] = INT_MAX;  // no column reaches a padding entry
    }
    __syncthreads();
  }
  return in;
}

#endif
