// hmv_kernels.hip -- H-matvec streaming kernel for gfx950.
//
// Replaces the reference's GPU leaf apply, c_hacapk_adot_body_lfmtx_batch_queue
// (HACApK_BATCH.c:47-255) with its vbatched atomic DGEMV launches
// (HACApK_BATCH.c:328-334), by two launches of one streaming kernel:
//   phase 1: t partials = V^T x  over the column-grouped rank vectors
//   phase 2: y = [A|U][x;t]      over row atoms (one writer per y row)
//
// Workgroup = 256 threads (4 wave64) owning one column-major block of
// rows_p x ncols.  Lane l handles rows [vr*(l%q), vr*(l%q)+vr) of columns
// l/q, l/q + m, ... (q = rows_p/vr lanes per column, m = 256/q columns per
// step), so one step of the workgroup reads m whole columns, i.e. one
// contiguous run of the arena, and every lane keeps its rows' sums in
// registers across all columns.  Inputs are staged in LDS once per 2048
// columns; the m column-group partials of each row are summed in a fixed
// order at the end, so results are deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hmv_engine.h"
#include "hmv_internal.h"

namespace {

__device__ inline float bf16_lo(uint32_t v) { return __uint_as_float(v << 16); }
__device__ inline float bf16_hi(uint32_t v) { return __uint_as_float(v & 0xffff0000u); }

// Accumulates columns [ca, cb) of one part (column-major, ld = rows_p) into
// the lane's row sums; zc points at the staged input of column ca.  DT is
// the part's storage type (HMV_DT_*).
template <typename Acc, int DT>
__device__ inline void accumulate(const char *__restrict__ part, int rows_p, int rq, int cg, int m,
                                  int ca, int cb, const Acc *__restrict__ zc, Acc &a0, Acc &a1) {
  if constexpr (DT == HMV_DT_F64) {
    const double *B = reinterpret_cast<const double *>(part) + rq;
#pragma unroll 8
    for (int c = ca + cg; c < cb; c += m)
      a0 = __builtin_fma((Acc)B[(size_t)c * rows_p], zc[c - ca], a0);
  } else if constexpr (DT == HMV_DT_F32) {
    const float *B = reinterpret_cast<const float *>(part) + 2 * rq;
#pragma unroll 8
    for (int c = ca + cg; c < cb; c += m) {
      const float2 v = *reinterpret_cast<const float2 *>(B + (size_t)c * rows_p);
      const Acc zz = zc[c - ca];
      a0 = __builtin_fma((Acc)v.x, zz, a0);
      a1 = __builtin_fma((Acc)v.y, zz, a1);
    }
  } else {
    const uint16_t *B = reinterpret_cast<const uint16_t *>(part) + 2 * rq;
#pragma unroll 8
    for (int c = ca + cg; c < cb; c += m) {
      const uint32_t v = *reinterpret_cast<const uint32_t *>(B + (size_t)c * rows_p);
      const Acc zz = zc[c - ca];
      a0 = __builtin_fma((Acc)bf16_lo(v), zz, a0);
      a1 = __builtin_fma((Acc)bf16_hi(v), zz, a1);
    }
  }
}

// Stages inputs of columns [c0, c0+cw) into z[0, cw).
template <typename Acc>
__device__ inline void stage(Acc *__restrict__ z, const HmvZSeg *__restrict__ zsegs,
                             const HmvWork &w, int c0, int cw, const double *__restrict__ x,
                             const int *__restrict__ zidx, const int *__restrict__ tdesc,
                             const Acc *__restrict__ tpart) {
  const int tid = threadIdx.x;
  int col = 0;
  for (int s = w.zseg_begin; s < w.zseg_end && col < c0 + cw; ++s) {
    const HmvZSeg zs = zsegs[s];
    const int lo = max(col, c0), hi = min(col + zs.ncols, c0 + cw);
    for (int c = lo + tid; c < hi; c += HMV_WG_THREADS) {
      const int j = c - col;
      Acc v;
      if (zs.kind == HMV_Z_X) {
        v = (Acc)x[zs.src + j];
      } else if (zs.kind == HMV_Z_XGATHER) {
        v = (Acc)x[zidx[zs.src + j]];
      } else {
        const int slot = zs.src + j;
        const int b = tdesc[2 * slot], np = tdesc[2 * slot + 1];
        v = tpart[b];
        for (int p = 1; p < np; ++p) v += tpart[b + p];
      }
      z[c - c0] = v;
    }
    col += zs.ncols;
  }
}

// PHASE 1: tpart[outidx[out_base + r]] = row sums.
// PHASE 2: y[out_base + r] = row sums + beta * y.
// DT0/DT1: storage types of the two column parts (fixed per arena precision).
template <typename Acc, int PHASE, int DT0, int DT1>
__global__ __launch_bounds__(HMV_WG_THREADS) void hmv_stream_kernel(
    const char *__restrict__ arena, const HmvWork *__restrict__ works,
    const HmvZSeg *__restrict__ zsegs, const double *__restrict__ x, const int *__restrict__ zidx,
    const int *__restrict__ tdesc, const Acc *__restrict__ tpart_in, const int *__restrict__ outidx,
    Acc *__restrict__ tpart_out, double *__restrict__ y, double beta, const int *__restrict__ skip) {
  if (skip && *skip) return;
  __shared__ __attribute__((aligned(16))) Acc z[HMV_ZCHUNK];
  __shared__ Acc red[2 * HMV_WG_THREADS];
  const HmvWork w = works[blockIdx.x];
  const int tid = threadIdx.x;
  constexpr int vr = (DT0 == HMV_DT_F64) ? 1 : 2;
  const int q = w.rows_p / vr;
  const int m = HMV_WG_THREADS / q;
  const bool active = tid < q * m;
  const int rq = tid % q, cg = tid / q;
  Acc a0 = Acc(0), a1 = Acc(0);
  const int ncols = w.ncol0 + w.ncol1;
  for (int c0 = 0; c0 < ncols; c0 += HMV_ZCHUNK) {
    const int cw = min(HMV_ZCHUNK, ncols - c0);
    if (c0 > 0) __syncthreads();
    stage<Acc>(z, zsegs, w, c0, cw, x, zidx, tdesc, tpart_in);
    __syncthreads();
    if (active) {
      const int e0 = min(c0 + cw, w.ncol0);
      if (c0 < e0) accumulate<Acc, DT0>(arena + w.off0, w.rows_p, rq, cg, m, c0, e0, z, a0, a1);
      const int s1 = max(c0, w.ncol0);
      if (s1 < c0 + cw)
        accumulate<Acc, DT1>(arena + w.off1, w.rows_p, rq, cg, m, s1 - w.ncol0, c0 + cw - w.ncol0,
                             z + (s1 - c0), a0, a1);
    }
  }
  if (active) {
    red[cg * w.rows_p + rq * vr] = a0;
    if constexpr (vr == 2) red[cg * w.rows_p + rq * vr + 1] = a1;
  }
  __syncthreads();
  for (int r = tid; r < w.rows; r += HMV_WG_THREADS) {
    Acc s = red[r];
    for (int g = 1; g < m; ++g) s += red[g * w.rows_p + r];
    if constexpr (PHASE == 1) {
      tpart_out[outidx[w.out_base + r]] = s;
    } else {
      const int row = w.out_base + r;
      const double v = (double)s;
      y[row] = beta == 0.0 ? v : __builtin_fma(beta, y[row], v);
    }
  }
}

template <typename Acc, int DT0, int DT1>
void launch_mode(const char *arena, const HmvWork *work1, int nwork1, const HmvWork *work2,
                 int nwork2, const HmvZSeg *zseg, const int *outidx, const int *zidx,
                 const int *tdesc, void *tpart, const double *x, double *y, double beta,
                 const int *skip, hipStream_t stream) {
  const dim3 block(HMV_WG_THREADS);
  // phase 1 reads only low-rank factors: both parts use DT1's type
  if (nwork1 > 0)
    hipLaunchKernelGGL((hmv_stream_kernel<Acc, 1, DT1, DT1>), dim3(nwork1), block, 0, stream, arena,
                       work1, zseg, x, zidx, tdesc, (const Acc *)nullptr, outidx, (Acc *)tpart,
                       (double *)nullptr, 0.0, skip);
  if (nwork2 > 0)
    hipLaunchKernelGGL((hmv_stream_kernel<Acc, 2, DT0, DT1>), dim3(nwork2), block, 0, stream, arena,
                       work2, zseg, x, zidx, tdesc, (const Acc *)tpart, outidx, (Acc *)nullptr, y,
                       beta, skip);
}

// Plain streaming kernels of known traffic, for bandwidth calibration.
__global__ __launch_bounds__(HMV_WG_THREADS) void hmv_read_kernel(const double2 *__restrict__ src,
                                                                   size_t n, double *__restrict__ out) {
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * HMV_WG_THREADS + threadIdx.x; i < n;
       i += (size_t)gridDim.x * HMV_WG_THREADS) {
    const double2 v = src[i];
    s += v.x + v.y;
  }
  if (s == 1.2345e-300) out[0] = s;  // keeps the loads alive; never true for the data used
}

__global__ __launch_bounds__(HMV_WG_THREADS) void hmv_copy_kernel(const double2 *__restrict__ src,
                                                                   double2 *__restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * HMV_WG_THREADS + threadIdx.x; i < n;
       i += (size_t)gridDim.x * HMV_WG_THREADS)
    dst[i] = src[i];
}

}  // namespace

extern "C" hipError_t hmv_launch_stream_test(int kind, const void *src, void *dst, size_t bytes,
                                             hipStream_t stream) {
  const size_t n = bytes / 16;
  const int grid = 256 * 32;
  if (kind == 0)
    hipLaunchKernelGGL(hmv_read_kernel, dim3(grid), dim3(HMV_WG_THREADS), 0, stream,
                       (const double2 *)src, n, (double *)dst);
  else
    hipLaunchKernelGGL(hmv_copy_kernel, dim3(grid), dim3(HMV_WG_THREADS), 0, stream,
                       (const double2 *)src, (double2 *)dst, n);
  return hipGetLastError();
}

// Launches phase 1 (nwork1 workgroups) and/or phase 2 (nwork2) of y = H x + beta*y.
// mode: 0 = FP64 arena, 1 = FP32 arena, 2 = BF16 low-rank + FP32 dense arena.
extern "C" hipError_t hmv_launch_apply(int mode, const char *arena, const HmvWork *work1,
                                       int nwork1, const HmvWork *work2, int nwork2,
                                       const HmvZSeg *zseg, const int *outidx, const int *zidx,
                                       const int *tdesc, void *tpart, const double *x, double *y,
                                       double beta, const int *skip, hipStream_t stream) {
  if (mode == 0)
    launch_mode<double, HMV_DT_F64, HMV_DT_F64>(arena, work1, nwork1, work2, nwork2, zseg, outidx,
                                                zidx, tdesc, tpart, x, y, beta, skip, stream);
  else if (mode == 1)
    launch_mode<float, HMV_DT_F32, HMV_DT_F32>(arena, work1, nwork1, work2, nwork2, zseg, outidx,
                                               zidx, tdesc, tpart, x, y, beta, skip, stream);
  else
    launch_mode<float, HMV_DT_F32, HMV_DT_BF16>(arena, work1, nwork1, work2, nwork2, zseg, outidx,
                                                zidx, tdesc, tpart, x, y, beta, skip, stream);
  return hipGetLastError();
}
