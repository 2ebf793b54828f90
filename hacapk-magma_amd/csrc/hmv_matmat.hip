// hmv_matmat.hip -- multi-right-hand-side H-matmat Y = H X + beta*Y on MFMA.
//
// X and Y hold k right-hand sides interleaved (row i's k values contiguous,
// N x k row-major); t likewise (slot s's k values at t[s*k]).  The apply runs
// over the same arena and work lists as the single-vector kernel, as
// 16-row tiles: one wave (64 threads) per (work item, 16-row tile).
//
// Per tile the wave stages inputs z[column][rhs] (k padded to KT*16) for a
// chunk of the block's columns in LDS, then walks the chunk four columns at
// a time: one 8-byte (or 4/2-byte) load per lane gives the 16 x 4 A fragment
// (lane l: row l%16 of column l/16 of the column-major block), and KT MFMAs
// (16x16x4; FP64, or FP32 for FP32/BF16 arenas) multiply it with the 4 x 16
// B fragments z[column][16 t + l%16] into KT 16 x 16 accumulators.  Rows and
// columns outside the block load zero through the buffer range check
// (out-of-range vector offset), inputs outside the chunk read a zero row.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hmv_engine.h"
#include "hmv_inputs.h"
#include "hmv_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kOff = 0x80000000u;  // out-of-range offset: the load returns 0
constexpr int kNT = 2;                  // nontemporal cache policy
constexpr int kZBytes = 16384;          // LDS bytes of staged inputs per wave

__device__ inline __amdgpu_buffer_rsrc_t rsrc(const char *base, uint32_t nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)nbytes, 0x00020000);
}

template <typename Acc>
struct Mfma;
// A fragment: lane l holds A[l%16][l/16]; B fragment: lane l holds
// B[l/16][l%16]; result register i of lane l is D[row(l, i)][l%16] (the two
// types order their result rows differently; measured on gfx950).
template <>
struct Mfma<double> {
  using acc = d4;
  __device__ static d4 op(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int i) { return (lane >> 4) + 4 * i; }
};
template <>
struct Mfma<float> {
  using acc = f4;
  __device__ static f4 op(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int i) { return 4 * (lane >> 4) + i; }
};

// One A entry of a part, widened to the accumulation type (0 when off).
template <int DT>
struct Entry;
template <>
struct Entry<HMV_DT_F64> {
  static constexpr int es = 8;
  __device__ static double load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, kNT));
  }
};
template <>
struct Entry<HMV_DT_F32> {
  static constexpr int es = 4;
  __device__ static float load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, kNT));
  }
};
template <>
struct Entry<HMV_DT_BF16> {
  static constexpr int es = 2;
  __device__ static float load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const uint32_t v = __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, kNT);
    return __uint_as_float(v << 16);
  }
};

// Stages z[cc][j] = input j of column c0+cc (cc < cw, j < KP; 0 for j >= k),
// and the zero row z[cw][*].  The chunk's column sources (x index or t slot
// of the item's runs or code table) are resolved once into `src` (LDS;
// element offsets, t slots flagged negative), then the wave gathers SB
// inputs per lane per round trip.
template <typename Acc, int KP>
__device__ inline void stage(Acc *__restrict__ z, long long *__restrict__ src, const Inputs &in,
                             int c0, int cw, int k, const double *__restrict__ x,
                             const int *__restrict__ codes, const double *__restrict__ t) {
  constexpr int SB = 16;
  const int lane = threadIdx.x;
  for (int cc = lane; cc < cw; cc += 64) {
    const int code = input_code(in, codes, c0 + cc);
    src[cc] = code >= 0 ? (long long)code * k : -1 - (long long)(-code - 1) * k;
  }
  __syncthreads();
  const int n = (cw + 1) * KP;
  for (int e0 = 0; e0 < n; e0 += 64 * SB) {
    double v[SB];
#pragma unroll
    for (int s = 0; s < SB; ++s) {
      const int e = e0 + s * 64 + lane;
      const int cc = e / KP, j = e % KP;
      v[s] = 0.0;
      if (e < n && cc < cw && j < k) {
        const long long o = src[cc];
        v[s] = o >= 0 ? x[o + j] : t[-1 - o + j];
      }
    }
#pragma unroll
    for (int s = 0; s < SB; ++s) {
      const int e = e0 + s * 64 + lane;
      if (e < n) z[e] = (Acc)v[s];
    }
  }
}

// Independent accumulator sets per tile, summed at the end: two sets for
// two RHS tiles measured fastest on MI355X (k = 32); one set otherwise.
template <int KT>
constexpr int acc_sets() {
  return KT == 2 ? 2 : 1;
}

// Accumulates the tile's product over part columns [lo, hi) (part-local;
// the chunk's staged inputs start at part column zbase).
template <typename Acc, int DT, int KT>
__device__ inline void part_mma(typename Mfma<Acc>::acc (&acc)[acc_sets<KT>()][KT],
                                const Acc *__restrict__ z, __amdgpu_buffer_rsrc_t rs, int rows_p,
                                int r, int ncols, int lo, int hi, int zbase, int cw) {
  using E = Entry<DT>;
  constexpr int KP = KT * 16;
  constexpr int U = 4;  // quads (4 columns) in flight
  constexpr int NA = acc_sets<KT>();
  const int lane = threadIdx.x;
  const int cq = lane >> 4, n16 = lane & 15;
  for (int q0 = lo; q0 < hi; q0 += 4 * U) {
    Acc a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int col = q0 + 4 * u + cq;
      const bool ok = r < rows_p && col < min(hi, ncols);
      a[u] = (Acc)E::load(rs, ok ? (uint32_t)((col * rows_p + r) * E::es) : kOff);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int zr = min(q0 + 4 * u + cq - zbase, cw);  // zero row past the chunk
#pragma unroll
      for (int tt = 0; tt < KT; ++tt)
        acc[u % NA][tt] = Mfma<Acc>::op(a[u], z[zr * KP + tt * 16 + n16], acc[u % NA][tt]);
    }
  }
}

// PHASE 1: t[outidx[out_base + row] * k + j];  PHASE 2: y[(out_base + row) * k + j].
template <typename Acc, int PHASE, int DT0, int DT1, int KT>
__global__ __launch_bounds__(64) void hmv_matmat_kernel(
    const char *__restrict__ arena, const HmvWork *__restrict__ works, const int2 *__restrict__ tiles,
    const HmvZSeg *__restrict__ zsegs, const double *__restrict__ x, const int *__restrict__ codes,
    const double *__restrict__ t_in, const int *__restrict__ outidx, double *__restrict__ t_out,
    double *__restrict__ y, int k, double beta, const int *__restrict__ skip) {
  if (skip && *skip) return;
  constexpr int KP = KT * 16;
  constexpr int CH = (kZBytes / (int)sizeof(Acc) / KP - 1) & ~3;  // columns per chunk (multiple of 4)
  static_assert(CH >= 4, "chunk holds at least one column quad");
  __shared__ __attribute__((aligned(16))) Acc z[(CH + 1) * KP];
  __shared__ long long zsrc[CH];
  __shared__ int runs[2 * HMV_RUN_MAX];
  const int2 tile = tiles[blockIdx.x];
  const HmvWork w = works[tile.x];
  const int lane = threadIdx.x;
  const int r = tile.y * 16 + (lane & 15);  // this lane's A row
  const int ncols = w.ncol0 + w.ncol1;
  const Inputs in = load_inputs<64>(w, zsegs, runs);
  const __amdgpu_buffer_rsrc_t rs0 =
      rsrc(arena + w.off0, (uint32_t)((size_t)w.rows_p * w.ncol0 * Entry<DT0>::es));
  const __amdgpu_buffer_rsrc_t rs1 =
      rsrc(arena + (w.ncol1 > 0 ? w.off1 : w.off0),
           (uint32_t)((size_t)w.rows_p * w.ncol1 * Entry<DT1>::es));
  constexpr int NA = acc_sets<KT>();
  typename Mfma<Acc>::acc acc[NA][KT];
#pragma unroll
  for (int sa = 0; sa < NA; ++sa)
#pragma unroll
    for (int tt = 0; tt < KT; ++tt) acc[sa][tt] = typename Mfma<Acc>::acc{0, 0, 0, 0};
  for (int c0 = 0; c0 < ncols; c0 += CH) {
    const int cw = min(CH, ncols - c0);
    if (c0 > 0) __syncthreads();
    stage<Acc, KP>(z, zsrc, in, c0, cw, k, x, codes, t_in);
    __syncthreads();
    if (c0 < w.ncol0)
      part_mma<Acc, DT0, KT>(acc, z, rs0, w.rows_p, r, w.ncol0, c0, min(c0 + cw, w.ncol0), c0, cw);
    if (c0 + cw > w.ncol0) {
      const int lo = max(c0, w.ncol0) - w.ncol0;
      part_mma<Acc, DT1, KT>(acc, z, rs1, w.rows_p, r, w.ncol1, lo, c0 + cw - w.ncol0,
                             c0 - w.ncol0, cw);
    }
  }
#pragma unroll
  for (int sa = 1; sa < NA; ++sa)
#pragma unroll
    for (int tt = 0; tt < KT; ++tt) acc[0][tt] += acc[sa][tt];
  // lane holds rows Mfma::row(lane, i), column lane%16 of each 16 x 16 tile
  const int j0 = lane & 15;
#pragma unroll
  for (int tt = 0; tt < KT; ++tt) {
    const int j = tt * 16 + j0;
    if (j >= k) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = tile.y * 16 + Mfma<Acc>::row(lane, i);
      if (row >= w.rows) continue;
      const double v = (double)acc[0][tt][i];
      if constexpr (PHASE == 1) {
        t_out[(size_t)outidx[w.out_base + row] * k + j] = v;
      } else {
        double *py = y + (size_t)(w.out_base + row) * k + j;
        *py = beta == 0.0 ? v : __builtin_fma(beta, *py, v);
      }
    }
  }
}

// t[slot*k + j] = sum over p of t[(base+p)*k + j], in chunk order.
__global__ __launch_bounds__(256) void hmv_tfinal_block_kernel(const HmvTFin *__restrict__ fin,
                                                               int nfin, int k,
                                                               double *__restrict__ t,
                                                               const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)nfin * k) return;
  const HmvTFin f = fin[i / k];
  const int j = (int)(i % k);
  double s = t[(size_t)f.base * k + j];
  for (int p = 1; p < f.np; ++p) s += t[(size_t)(f.base + p) * k + j];
  t[(size_t)f.slot * k + j] = s;
}

template <typename Acc, int PHASE, int DT0, int DT1>
void launch_tiles(int kt, const HmvLaunch &L, const int2 *tiles, int ntiles, const HmvWork *works,
                  const double *x, double *t, double *y, int k, double beta, const int *skip,
                  hipStream_t s) {
  if (ntiles <= 0) return;
  const double *t_in = PHASE == 2 ? t : nullptr;
  double *t_out = PHASE == 1 ? t : nullptr;
#define HMV_MM_LAUNCH(KT)                                                                         \
  hipLaunchKernelGGL((hmv_matmat_kernel<Acc, PHASE, DT0, DT1, KT>), dim3(ntiles), dim3(64), 0, s, \
                     L.arena, works, tiles, L.zseg, x, L.codes, t_in, L.outidx, t_out, y, k, beta,  \
                     skip)
  if (kt == 1)
    HMV_MM_LAUNCH(1);
  else if (kt == 2)
    HMV_MM_LAUNCH(2);
  else if (kt <= 4)
    HMV_MM_LAUNCH(4);
  else
    HMV_MM_LAUNCH(8);
#undef HMV_MM_LAUNCH
}

template <typename Acc, int DT0, int DT1>
void launch_block(const HmvLaunch &L, const int2 *tiles1, int ntiles1, const int2 *tiles2,
                  int ntiles2, int k, const double *x, double *t, double *y, double beta,
                  const int *skip, hipStream_t s) {
  const int kt = (k + 15) / 16;
  launch_tiles<Acc, 1, DT1, DT1>(kt, L, tiles1, ntiles1, L.work1, x, t, nullptr, k, 0.0, skip, s);
  if (L.ntfin > 0) {
    const int64_t n = (int64_t)L.ntfin * k;
    hipLaunchKernelGGL(hmv_tfinal_block_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       L.tfin, L.ntfin, k, t, skip);
  }
  launch_tiles<Acc, 2, DT0, DT1>(kt, L, tiles2, ntiles2, L.work2, x, t, y, k, beta, skip, s);
}

}  // namespace

// Y = H X + beta*Y for k (1..128) interleaved right-hand sides; t holds
// tpart_size * k doubles.
extern "C" hipError_t hmv_launch_block(const HmvLaunch *L, const int2 *tiles1, int ntiles1,
                                       const int2 *tiles2, int ntiles2, int k, const double *x,
                                       double *t, double *y, double beta, const int *skip,
                                       hipStream_t stream) {
  if (L->mode == 0)
    launch_block<double, HMV_DT_F64, HMV_DT_F64>(*L, tiles1, ntiles1, tiles2, ntiles2, k, x, t, y,
                                                 beta, skip, stream);
  else if (L->mode == 1)
    launch_block<float, HMV_DT_F32, HMV_DT_F32>(*L, tiles1, ntiles1, tiles2, ntiles2, k, x, t, y,
                                                beta, skip, stream);
  else if (L->mode == 2)
    launch_block<float, HMV_DT_F32, HMV_DT_BF16>(*L, tiles1, ntiles1, tiles2, ntiles2, k, x, t, y,
                                                 beta, skip, stream);
  else
    launch_block<float, HMV_DT_BF16, HMV_DT_BF16>(*L, tiles1, ntiles1, tiles2, ntiles2, k, x, t, y,
                                                  beta, skip, stream);
  return hipGetLastError();
}
