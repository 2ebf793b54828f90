// hmv_lanes.h -- per-lane loads of a work item's column-major block (device
// code shared by the single-vector and the k-right-hand-side streaming
// kernels).
//
// A lane owns VR consecutive rows of the columns it visits; its element of
// one column is one 8-byte load (FP64 entry, FP32 pair, BF16 quad) or one
// 4-byte load (BF16 pair), through a raw buffer resource over the part, so
// that a column past the part loads zero without touching memory.
#ifndef HMV_LANES_H
#define HMV_LANES_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hmv_internal.h"

namespace hmv_lanes {

typedef float float2_nt __attribute__((ext_vector_type(2)));

__device__ inline float bf16_lo(uint32_t v) { return __uint_as_float(v << 16); }
__device__ inline float bf16_hi(uint32_t v) { return __uint_as_float(v & 0xffff0000u); }

// Fused multiply-add in the accumulation type: one FP32 FMA for FP32
// accumulators (__builtin_fma is the FP64 builtin: on floats it converts
// all three operands, multiplies in FP64 and rounds back).
__device__ inline float acc_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double acc_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Buffer resource over one column part (raw, byte-addressed).  A load whose
// vector offset is at or past num_records returns zero without touching
// memory; lanes with nothing to load use kOffNone.
constexpr uint32_t kOffNone = 0x80000000u;
constexpr int kBufNT = 2;  // cache policy bits: nontemporal
__device__ inline __amdgpu_buffer_rsrc_t part_rsrc(const char *base, uint32_t nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)nbytes, 0x00020000);
}

// Bytes per stored entry, and rows per lane of a block whose first part is
// stored as DT0 (the second part uses the same rows): 8-byte lane loads for
// FP64 entries, FP32 pairs and BF16 quads (phase 1 of BF16 arenas); the
// BF16 second part of a phase-2 block of a BF16 arena loads 4-byte pairs.
template <int DT>
constexpr int dt_bytes() {
  return DT == HMV_DT_F64 ? 8 : DT == HMV_DT_F32 ? 4 : 2;
}
template <int PHASE, int DT0>
constexpr int rows_per_lane() {
  return DT0 == HMV_DT_F64 ? 1 : (DT0 == HMV_DT_F32 || PHASE == 2) ? 2 : 4;
}

typedef uint32_t uint2_nt __attribute__((ext_vector_type(2)));

// Per-lane element of one column part: VR consecutive rows of one column,
// accumulated into a[0..VR).
template <int DT, int VR>
struct Lane;
template <>
struct Lane<HMV_DT_F64, 1> {
  using raw = double;
  __device__ static raw load(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, kBufNT));
  }
  template <typename Acc>
  __device__ static void fma(raw v, Acc z, Acc (&a)[1]) {
    a[0] = acc_fma((Acc)v, z, a[0]);
  }
  template <typename Acc>
  __device__ static void unpack(raw v, Acc (&o)[1]) {
    o[0] = (Acc)v;
  }
};
template <>
struct Lane<HMV_DT_F32, 2> {
  using raw = float2_nt;
  __device__ static raw load(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff) {
    return __builtin_bit_cast(float2_nt, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, kBufNT));
  }
  template <typename Acc>
  __device__ static void fma(raw v, Acc z, Acc (&a)[2]) {
    a[0] = acc_fma((Acc)v.x, z, a[0]);
    a[1] = acc_fma((Acc)v.y, z, a[1]);
  }
  template <typename Acc>
  __device__ static void unpack(raw v, Acc (&o)[2]) {
    o[0] = (Acc)v.x;
    o[1] = (Acc)v.y;
  }
};
template <>
struct Lane<HMV_DT_BF16, 2> {
  using raw = uint32_t;
  __device__ static raw load(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, kBufNT);
  }
  template <typename Acc>
  __device__ static void fma(raw v, Acc z, Acc (&a)[2]) {
    a[0] = acc_fma((Acc)bf16_lo(v), z, a[0]);
    a[1] = acc_fma((Acc)bf16_hi(v), z, a[1]);
  }
  template <typename Acc>
  __device__ static void unpack(raw v, Acc (&o)[2]) {
    o[0] = (Acc)bf16_lo(v);
    o[1] = (Acc)bf16_hi(v);
  }
};
template <>
struct Lane<HMV_DT_BF16, 4> {
  using raw = uint2_nt;
  __device__ static raw load(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff) {
    return __builtin_bit_cast(uint2_nt, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, kBufNT));
  }
  template <typename Acc>
  __device__ static void fma(raw v, Acc z, Acc (&a)[4]) {
    a[0] = acc_fma((Acc)bf16_lo(v.x), z, a[0]);
    a[1] = acc_fma((Acc)bf16_hi(v.x), z, a[1]);
    a[2] = acc_fma((Acc)bf16_lo(v.y), z, a[2]);
    a[3] = acc_fma((Acc)bf16_hi(v.y), z, a[3]);
  }
  template <typename Acc>
  __device__ static void unpack(raw v, Acc (&o)[4]) {
    o[0] = (Acc)bf16_lo(v.x);
    o[1] = (Acc)bf16_hi(v.x);
    o[2] = (Acc)bf16_lo(v.y);
    o[3] = (Acc)bf16_hi(v.y);
  }
};


// One column part of a block: buffer resource, the lane's byte offset of its
// first column (column cg), and the byte stride of one step of m columns.
template <int DT, int VR>
struct Part {
  static constexpr int es = dt_bytes<DT>();
  __amdgpu_buffer_rsrc_t rs;
  uint32_t voff;
  int step;
  __device__ Part(const char *base, int rows_p, int ncols, int rq, int cg, int m) {
    rs = part_rsrc(base, (uint32_t)((size_t)rows_p * ncols * es));
    voff = (uint32_t)(rq * VR + cg * rows_p) * (uint32_t)es;
    step = m * rows_p * es;
  }
};

// Loads per batch: 32 bytes in flight per lane (4 loads of 8-byte lane
// elements, 8 of 4-byte ones), so that narrow lanes keep as many bytes in
// flight as FP64 ones.
template <int DT, int VR>
constexpr int batch_loads() {
  return sizeof(typename Lane<DT, VR>::raw) >= 8 ? 4 : 8;
}


}  // namespace hmv_lanes

#endif
