// hmv_kernels.hip -- H-matvec streaming kernels for gfx950.
//
// Replaces the reference's GPU leaf apply, c_hacapk_adot_body_lfmtx_batch_queue
// (HACApK_BATCH.c:47-255) with its vbatched atomic DGEMV launches
// (HACApK_BATCH.c:328-334), by two launches of one streaming kernel:
//   phase 1: t = V^T x   over the column-grouped V^T stacks
//   phase 2: y = [A|U][x;t] over the row-grouped panels (one writer per y row)
//
// Workgroup = 256 threads (4 wave64).  A segment's rows are handed to teams
// of LPR lanes (LPR = 4..64 within a wave, or 256 = the whole workgroup for
// long rows); each lane streams 16-byte vectors of its rows, multiplies by
// the segment's inputs staged in LDS, and the team reduces with lane shuffles.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hmv_engine.h"
#include "hmv_internal.h"

namespace {

template <typename T>
struct VecLoad;

template <>
struct VecLoad<double> {
  static constexpr int N = 2;
  __device__ static inline double fma_dot(const double *p, const double *z, double acc) {
    const double2 a = *reinterpret_cast<const double2 *>(p);
    const double2 b = *reinterpret_cast<const double2 *>(z);
    acc = __builtin_fma(a.x, b.x, acc);
    acc = __builtin_fma(a.y, b.y, acc);
    return acc;
  }
};

template <>
struct VecLoad<float> {
  static constexpr int N = 4;
  __device__ static inline float fma_dot(const float *p, const float *z, float acc) {
    const float4 a = *reinterpret_cast<const float4 *>(p);
    const float4 b = *reinterpret_cast<const float4 *>(z);
    acc = __builtin_fmaf(a.x, b.x, acc);
    acc = __builtin_fmaf(a.y, b.y, acc);
    acc = __builtin_fmaf(a.z, b.z, acc);
    acc = __builtin_fmaf(a.w, b.w, acc);
    return acc;
  }
};

// 16-bit brain float storage; the payload member gives the type its 2-byte
// size, so pointer arithmetic on bf16x* steps over whole entries.
struct bf16x {
  uint16_t bits;
};
static_assert(sizeof(bf16x) == 2, "bf16x must be 2 bytes");

template <>
struct VecLoad<bf16x> {
  static constexpr int N = 8;
  __device__ static inline float lo(uint32_t v) { return __uint_as_float(v << 16); }
  __device__ static inline float hi(uint32_t v) { return __uint_as_float(v & 0xffff0000u); }
  __device__ static inline float fma_dot(const bf16x *p, const float *z, float acc) {
    const uint4 a = *reinterpret_cast<const uint4 *>(p);
    const float4 b0 = *reinterpret_cast<const float4 *>(z);
    const float4 b1 = *reinterpret_cast<const float4 *>(z + 4);
    acc = __builtin_fmaf(lo(a.x), b0.x, acc);
    acc = __builtin_fmaf(hi(a.x), b0.y, acc);
    acc = __builtin_fmaf(lo(a.y), b0.z, acc);
    acc = __builtin_fmaf(hi(a.y), b0.w, acc);
    acc = __builtin_fmaf(lo(a.z), b1.x, acc);
    acc = __builtin_fmaf(hi(a.z), b1.y, acc);
    acc = __builtin_fmaf(lo(a.w), b1.z, acc);
    acc = __builtin_fmaf(hi(a.w), b1.w, acc);
    return acc;
  }
};

template <typename T>
struct StoreSize {
  static constexpr int value = sizeof(T);
};

// Rows [0,nrows) of a row-major block (row stride ld elements) times the
// staged inputs z[0, nvec*N); row sums added to accrows[r].
template <typename T, typename Acc, int LPR>
__device__ inline void team_rows(const char *__restrict__ base, int ld, int nvec, int nrows,
                                 const Acc *__restrict__ z, Acc *__restrict__ accrows,
                                 Acc *__restrict__ red) {
  constexpr int N = VecLoad<T>::N;
  constexpr int ES = StoreSize<T>::value;
  const int tid = threadIdx.x;
  if constexpr (LPR <= 64) {
    constexpr int TEAMS = HMV_WG_THREADS / LPR;
    const int team = tid / LPR, lane = tid % LPR;
    for (int r = team; r < nrows; r += TEAMS) {
      const T *row = reinterpret_cast<const T *>(base + (size_t)r * ld * ES);
      Acc s0 = 0, s1 = 0;
      int v = lane;
      for (; v + LPR < nvec; v += 2 * LPR) {
        s0 = VecLoad<T>::fma_dot(row + (size_t)v * N, z + v * N, s0);
        s1 = VecLoad<T>::fma_dot(row + (size_t)(v + LPR) * N, z + (v + LPR) * N, s1);
      }
      if (v < nvec) s0 = VecLoad<T>::fma_dot(row + (size_t)v * N, z + v * N, s0);
      Acc s = s0 + s1;
#pragma unroll
      for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPR);
      if (lane == 0) accrows[r] += s;
    }
  } else {
    // whole workgroup per row
    const int wave = tid >> 6, lane = tid & 63;
    for (int r = 0; r < nrows; ++r) {
      const T *row = reinterpret_cast<const T *>(base + (size_t)r * ld * ES);
      Acc s0 = 0, s1 = 0;
      int v = tid;
      for (; v + HMV_WG_THREADS < nvec; v += 2 * HMV_WG_THREADS) {
        s0 = VecLoad<T>::fma_dot(row + (size_t)v * N, z + v * N, s0);
        s1 = VecLoad<T>::fma_dot(row + (size_t)(v + HMV_WG_THREADS) * N, z + (v + HMV_WG_THREADS) * N, s1);
      }
      if (v < nvec) s0 = VecLoad<T>::fma_dot(row + (size_t)v * N, z + v * N, s0);
      Acc s = s0 + s1;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (lane == 0) red[wave] = s;
      __syncthreads();
      if (tid == 0) accrows[r] += (red[0] + red[1]) + (red[2] + red[3]);
      __syncthreads();
    }
  }
}

template <typename T, typename Acc>
__device__ inline void dispatch_lpr(int lpr, const char *base, int ld, int nvec, int nrows,
                                    const Acc *z, Acc *accrows, Acc *red) {
  switch (lpr) {
    case 4: team_rows<T, Acc, 4>(base, ld, nvec, nrows, z, accrows, red); break;
    case 8: team_rows<T, Acc, 8>(base, ld, nvec, nrows, z, accrows, red); break;
    case 16: team_rows<T, Acc, 16>(base, ld, nvec, nrows, z, accrows, red); break;
    case 32: team_rows<T, Acc, 32>(base, ld, nvec, nrows, z, accrows, red); break;
    case 64: team_rows<T, Acc, 64>(base, ld, nvec, nrows, z, accrows, red); break;
    default: team_rows<T, Acc, 256>(base, ld, nvec, nrows, z, accrows, red); break;
  }
}

__device__ inline int dt_bytes(int dt) { return dt == HMV_DT_F64 ? 8 : dt == HMV_DT_F32 ? 4 : 2; }

// PHASE 1: out[slotmap[out_base+i]] = rows of V^T times x.
// PHASE 2: y[out_base+i] = beta*y + sum over segments.
template <typename Acc, int PHASE>
__global__ __launch_bounds__(HMV_WG_THREADS) void hmv_stream_kernel(
    const char *__restrict__ arena, const HmvSeg *__restrict__ segs,
    const HmvWork *__restrict__ works, const double *__restrict__ x,
    const Acc *__restrict__ tin, const int *__restrict__ zidx,
    const int *__restrict__ slotmap, Acc *__restrict__ tout, double *__restrict__ y,
    double beta, const int *__restrict__ skip) {
  if (skip && *skip) return;
  __shared__ Acc acc[HMV_MAX_ROWS];
  __shared__ __attribute__((aligned(16))) Acc z[HMV_ZCHUNK + 16];
  __shared__ Acc red[4];
  const HmvWork w = works[blockIdx.x];
  const int tid = threadIdx.x;
  for (int i = tid; i < w.nout; i += HMV_WG_THREADS) acc[i] = Acc(0);
  for (int si = w.seg_begin; si < w.seg_end; ++si) {
    const HmvSeg s = segs[si];
    const int es = dt_bytes(s.dtype);
    const int vecn = 16 / es;
    for (int c0 = 0; c0 < s.ncols; c0 += HMV_ZCHUNK) {
      const int cw = min(HMV_ZCHUNK, s.ncols - c0);
      const int nvec = (cw + vecn - 1) / vecn;
      const int cwp = nvec * vecn;
      __syncthreads();
      for (int c = tid; c < cwp; c += HMV_WG_THREADS) {
        const int gc = c0 + c;
        Acc v = Acc(0);
        if (c < cw) {
          if (gc < s.nx) {
            v = (Acc)(PHASE == 1 ? x[s.xsrc + gc] : x[zidx[s.xsrc + gc]]);
          } else {
            v = tin[s.tsrc + (gc - s.nx)];
          }
        }
        z[c] = v;
      }
      __syncthreads();
      const char *base = arena + s.data_off + (long long)c0 * es;
      if constexpr (sizeof(Acc) == 8) {
        dispatch_lpr<double, Acc>(s.lpr, base, s.ld, nvec, s.nrows, z, acc + s.acc_row, red);
      } else {
        if (s.dtype == HMV_DT_F32)
          dispatch_lpr<float, Acc>(s.lpr, base, s.ld, nvec, s.nrows, z, acc + s.acc_row, red);
        else
          dispatch_lpr<bf16x, Acc>(s.lpr, base, s.ld, nvec, s.nrows, z, acc + s.acc_row, red);
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < w.nout; i += HMV_WG_THREADS) {
    if constexpr (PHASE == 1) {
      tout[slotmap[w.out_base + i]] = acc[i];
    } else {
      const int r = w.out_base + i;
      const double v = (double)acc[i];
      y[r] = beta == 0.0 ? v : __builtin_fma(beta, y[r], v);
    }
  }
}

}  // namespace

// Launches one apply y = H x + beta*y on `stream`.  t has length >= tsize
// elements of the accumulation type (double if acc64, float otherwise).
extern "C" hipError_t hmv_launch_apply(int acc64, const char *arena, const HmvSeg *seg1,
                                       const HmvWork *work1, int nwork1, const HmvSeg *seg2,
                                       const HmvWork *work2, int nwork2, const int *slotmap,
                                       const int *zidx, void *t, const double *x, double *y,
                                       double beta, const int *skip, hipStream_t stream) {
  const dim3 block(HMV_WG_THREADS);
  if (acc64) {
    if (nwork1 > 0)
      hipLaunchKernelGGL((hmv_stream_kernel<double, 1>), dim3(nwork1), block, 0, stream, arena,
                         seg1, work1, x, (const double *)nullptr, zidx, slotmap, (double *)t,
                         (double *)nullptr, 0.0, skip);
    if (nwork2 > 0)
      hipLaunchKernelGGL((hmv_stream_kernel<double, 2>), dim3(nwork2), block, 0, stream, arena,
                         seg2, work2, x, (const double *)t, zidx, slotmap, (double *)nullptr, y,
                         beta, skip);
  } else {
    if (nwork1 > 0)
      hipLaunchKernelGGL((hmv_stream_kernel<float, 1>), dim3(nwork1), block, 0, stream, arena,
                         seg1, work1, x, (const float *)nullptr, zidx, slotmap, (float *)t,
                         (double *)nullptr, 0.0, skip);
    if (nwork2 > 0)
      hipLaunchKernelGGL((hmv_stream_kernel<float, 2>), dim3(nwork2), block, 0, stream, arena,
                         seg2, work2, x, (const float *)t, zidx, slotmap, (float *)nullptr, y, beta,
                         skip);
  }
  return hipGetLastError();
}
