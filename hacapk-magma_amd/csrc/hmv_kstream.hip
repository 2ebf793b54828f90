// hmv_kstream.hip -- H-matmat for a few right-hand sides (k <= 8 with FP64
// factors, k <= 4 with FP32 / BF16 factors) on the vector units.
//
// At small k the multi-RHS product is bound by streaming the arena, not by
// arithmetic (2k flops per stored entry), and a 16-wide MFMA tile would be
// mostly empty.  This kernel is the single-vector streaming kernel
// (hmv_kernels.hip) with K accumulators per row instead of one: the same
// work items, lane layout (VR rows of a column group per lane), batched
// buffer loads, input runs and fixed-order reduction; the staged inputs of
// a column are its K values (zero past k) and each loaded entry feeds K FMAs.
// Vectors are row-interleaved as in the MFMA H-matmat (x[i*k + j], t[s*k + j],
// y[i*k + j]), so the phases share the t buffer and its chunk-partial sums.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hmv_engine.h"
#include "hmv_inputs.h"
#include "hmv_internal.h"
#include "hmv_lanes.h"

namespace {
using namespace hmv_lanes;

// LDS elements of one workgroup: the row reduction needs NT*VR*K values; the
// staging holds at least 16 KB per 256 threads (that of the single-vector
// kernel), i.e. kcols columns of K values and one zero column.
template <typename Acc, int NT, int VR, int K>
constexpr int kbuf_elems() {
  constexpr int red = NT * VR * K;
  constexpr int stage = (16384 / (int)sizeof(Acc)) * NT / HMV_WG_THREADS;
  return red > stage ? red : stage;
}
template <typename Acc, int NT, int VR, int K>
constexpr int kcols() {
  return kbuf_elems<Acc, NT, VR, K>() / K - 1;
}

// Staging of the inputs of item columns [c0, c0+cw) into z[c*K + j], j < K
// (0 for j >= k), and of the zero column at c = cw.
//
// kstage_rows: one thread per column, its K values loaded together.
template <typename Acc, int NT, int K>
__device__ inline void kstage_rows(Acc *__restrict__ z, const Inputs &in, int c0, int cw,
                                   const double *__restrict__ x, const int *__restrict__ codes,
                                   const double *__restrict__ t, int k) {
  for (int c = threadIdx.x; c < cw; c += NT) {
    const int code = input_code(in, codes, c0 + c);
    const double *src = code >= 0 ? x + (size_t)code * k : t + (size_t)(-code - 1) * k;
    double v[K];
#pragma unroll
    for (int j = 0; j < K; ++j) v[j] = j < k ? src[j] : 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j) z[c * K + j] = (Acc)v[j];
  }
  if ((int)threadIdx.x < K) z[cw * K + threadIdx.x] = Acc(0);
}
// kstage_cols: each wave takes 64 columns at a time; lane l finds the code
// of column l, then K lanes per column load its K values (one wave load
// covers 64/K whole columns, K*8 contiguous bytes each; B loads in flight
// per lane) and write 64 consecutive z slots.
template <typename Acc, int NT, int K, int B>
__device__ inline void kstage_cols(Acc *__restrict__ z, const Inputs &in, int c0, int cw,
                                   const double *__restrict__ x, const int *__restrict__ codes,
                                   const double *__restrict__ t, int k) {
  static_assert(B > 0 && K % B == 0, "whole load groups");
  constexpr int CPI = 64 / K;  // columns per wave load
  const int lane = threadIdx.x & 63, j = lane % K;
  for (int b = (threadIdx.x >> 6) * 64; b < cw; b += NT) {
    const int mine = b + lane < cw ? input_code(in, codes, c0 + b + lane) : 0;
#pragma unroll 1
    for (int i0 = 0; i0 < K; i0 += B) {
      double v[B];
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int src_lane = (i0 + i) * CPI + lane / K;
        const int code = __shfl(mine, src_lane);
        const double *src = code >= 0 ? x + (size_t)code * k : t + (size_t)(-code - 1) * k;
        v[i] = b + src_lane < cw && j < k ? src[j] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < B; ++i)
        if (b + (i0 + i) * CPI + lane / K < cw) z[(b + (i0 + i) * CPI) * K + lane] = (Acc)v[i];
    }
  }
  if ((int)threadIdx.x < K) z[cw * K + threadIdx.x] = Acc(0);
}
// kstage_pairs: kstage_cols with each lane taking two adjacent values in
// one 16-byte load (K/2 lanes per column, half the wave loads), for an even
// k and 16-byte aligned inputs (FP64; k = 8 at 327,680 unknowns: 1.068
// against 1.098 ms, k = 6: 1.031 against 1.073, products bit-equal).
typedef double kd2 __attribute__((ext_vector_type(2)));
template <int NT, int K, int B>
__device__ inline void kstage_pairs(double *__restrict__ z, const Inputs &in, int c0, int cw,
                                    const double *__restrict__ x, const int *__restrict__ codes,
                                    const double *__restrict__ t, int k) {
  constexpr int LPC = K / 2;     // lanes per column
  constexpr int CPI = 64 / LPC;  // columns per wave load
  static_assert(B > 0 && LPC % B == 0, "whole load groups");
  const int lane = threadIdx.x & 63, jl = 2 * (lane % LPC);
  for (int b = (threadIdx.x >> 6) * 64; b < cw; b += NT) {
    const int mine = b + lane < cw ? input_code(in, codes, c0 + b + lane) : 0;
#pragma unroll 1
    for (int i0 = 0; i0 < LPC; i0 += B) {
      kd2 v[B];
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int src_lane = (i0 + i) * CPI + lane / LPC;
        const int code = __shfl(mine, src_lane);
        const double *src = code >= 0 ? x + (size_t)code * k : t + (size_t)(-code - 1) * k;
        v[i] = b + src_lane < cw && jl < k ? *(const kd2 *)(src + jl) : kd2{0.0, 0.0};
      }
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int c = b + (i0 + i) * CPI + lane / LPC;
        if (c < cw) *(kd2 *)(z + c * K + jl) = v[i];
      }
    }
  }
  if ((int)threadIdx.x < K) z[cw * K + threadIdx.x] = 0.0;
}
// B = 0: kstage_rows; PAIRS: kstage_pairs (the launch checks that it
// applies); else kstage_cols.
template <typename Acc, int NT, int K, int B, bool PAIRS>
__device__ inline void kstage(Acc *__restrict__ z, const Inputs &in, int c0, int cw,
                              const double *__restrict__ x, const int *__restrict__ codes,
                              const double *__restrict__ t, int k) {
  if constexpr (B == 0)
    kstage_rows<Acc, NT, K>(z, in, c0, cw, x, codes, t, k);
  else if constexpr (PAIRS)
    kstage_pairs<NT, K, B / 2>(z, in, c0, cw, x, codes, t, k);
  else
    kstage_cols<Acc, NT, K, B>(z, in, c0, cw, x, codes, t, k);
}

// K FMAs per row of one loaded lane element: a[i][j] += e[i] * zc[j].
template <typename Acc, int DT, int VR, int K>
__device__ inline void kfma(typename Lane<DT, VR>::raw v, const Acc *__restrict__ zc,
                            Acc (&a)[VR][K]) {
  Acc e[VR];
  Lane<DT, VR>::unpack(v, e);
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const Acc zj = zc[j];
#pragma unroll
    for (int i = 0; i < VR; ++i) a[i][j] = acc_fma(e[i], zj, a[i][j]);
  }
}

// The lane's n columns of one part from byte offset vb (stride `step`), the
// j-th with staged inputs at column zb + j*m; loads in batches of U, the
// last batch padded with zero loads and the zero column zmax.
template <typename Acc, int DT, int VR, int K, int U>
__device__ inline void kaccumulate(__amdgpu_buffer_rsrc_t rs, uint32_t vb, int step, int n,
                                   const Acc *__restrict__ z, int zb, int m, int zmax,
                                   Acc (&a)[VR][K]) {
  using LN = Lane<DT, VR>;
  int j = 0;
  for (; j + U <= n; j += U) {
    typename LN::raw v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = LN::load(rs, vb, (j + u) * step);
#pragma unroll
    for (int u = 0; u < U; ++u) kfma<Acc, DT, VR, K>(v[u], z + (zb + (j + u) * m) * K, a);
  }
  if (j < n) {
    typename LN::raw v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = LN::load(rs, j + u < n ? vb : kOffNone, (j + u) * step);
#pragma unroll
    for (int u = 0; u < U; ++u)
      kfma<Acc, DT, VR, K>(v[u], z + min(zb + (j + u) * m, zmax) * K, a);
  }
}

// Loads per batch and columns loaded before the first staging, by K.  FP64
// lanes keep four loads in flight up to K = 4 (at 327,680 unknowns k = 2 /
// 3 take 0.85 / 0.91 ms against 0.88 / 0.94 with two); at K = 8 four loads
// need more than 64 VGPRs, and at 6 waves per SIMD the product is slower
// (1.25 against 1.20 ms).
template <int DT, int VR, int K>
constexpr int kbatch() {
  return DT == HMV_DT_F64 && K <= 4 ? 4 : K >= 2 || sizeof(typename Lane<DT, VR>::raw) < 8 ? 2 : 4;
}
template <int VR, int K>
constexpr int kprefetch() {
  return K >= 2 ? 1 : 2;
}
// Staging loads in flight per lane of kstage_cols, or 0 for kstage_rows.
// The column-cooperative staging pays from K = 8 (FP64 only; 1.23 -> 1.19 ms
// per H-matmat at 327,680 unknowns); at K <= 4 a thread's K*8 contiguous
// bytes already make a short burst and the shuffles cost more than they save.
template <typename Acc, int VR, int K>
constexpr int kstage_batch() {
  return K >= 8 ? 4 : 0;
}

template <typename Acc, int DT, int VR, int K>
__device__ inline void krun_cols(const Part<DT, VR> &P, int lo, int hi, int cg, int m, int zshift,
                                 int zmax, const Acc *__restrict__ z, Acc (&a)[VR][K]) {
  const int j0 = lo > cg ? (lo - cg + m - 1) / m : 0;
  const int j1 = hi > cg ? (hi - cg + m - 1) / m : 0;
  if (j1 > j0)
    kaccumulate<Acc, DT, VR, K, kbatch<DT, VR, K>()>(P.rs, P.voff + (uint32_t)(j0 * P.step),
                                                       P.step, j1 - j0, z, cg + j0 * m + zshift, m,
                                                       zmax, a);
}

// One work item w for k <= K right-hand sides.  PHASE 1: t[outidx[out_base +
// r]*k + j] = row sums; PHASE 2: y[(out_base + r)*k + j] = row sums + beta y.
template <typename Acc, int PHASE, int DT0, int DT1, int NT, int K, bool PAIRS>
__device__ inline void kstream_item(const HmvWork &w, Acc *__restrict__ z, int *__restrict__ runs,
                                    const char *__restrict__ arena,
                                    const HmvZSeg *__restrict__ zsegs,
                                    const double *__restrict__ x, const int *__restrict__ codes,
                                    const double *__restrict__ t_in,
                                    const int *__restrict__ outidx, double *__restrict__ t_out,
                                    double *__restrict__ y, int k, double beta) {
  constexpr int vr = rows_per_lane<PHASE, DT0>();
  constexpr int ZC = kcols<Acc, NT, vr, K>();
  const int tid = threadIdx.x;
  const int q = w.rows_p / vr;
  const int m = NT / q;
  const bool active = tid < q * m;
  const int rq = tid % q, cg = tid / q;
  Acc a[vr][K];
#pragma unroll
  for (int i = 0; i < vr; ++i)
#pragma unroll
    for (int j = 0; j < K; ++j) a[i][j] = Acc(0);
  const int ncols = w.ncol0 + w.ncol1;
  const Part<DT0, vr> P0(arena + w.off0, w.rows_p, w.ncol0, rq, cg, m);
  // the lane's first PF columns of part 0 are loaded before the inputs are
  // staged, so the first HBM round trip overlaps the gathers
  constexpr int PF = kprefetch<vr, K>();
  const int lim0 = min(w.ncol0, ZC);
  const int npf = active ? max(0, min(PF, (lim0 - cg + m - 1) / m)) : 0;
  using L0 = Lane<DT0, vr>;
  typename L0::raw pre[PF];
#pragma unroll
  for (int u = 0; u < PF; ++u) pre[u] = L0::load(P0.rs, u < npf ? P0.voff : kOffNone, u * P0.step);
  const Inputs in = load_inputs<NT>(w, zsegs, runs);
  for (int c0 = 0; c0 < ncols; c0 += ZC) {
    const int cw = min(ZC, ncols - c0);
    if (c0 > 0) __syncthreads();
    kstage<Acc, NT, K, kstage_batch<Acc, vr, K>(), PAIRS>(z, in, c0, cw, x, codes, t_in, k);
    __syncthreads();
    if (active) {
      int lo0 = c0;
      if (c0 == 0) {
#pragma unroll
        for (int u = 0; u < PF; ++u) kfma<Acc, DT0, vr, K>(pre[u], z + min(cg + u * m, cw) * K, a);
        lo0 = min(npf * m, w.ncol0);  // the lane's columns below npf*m are done
        lo0 = npf == PF ? lo0 : cw;   // a lane with fewer than PF columns here has none left
      }
      const int hi0 = min(c0 + cw, w.ncol0);
      if (lo0 < hi0) krun_cols<Acc, DT0, vr, K>(P0, lo0, hi0, cg, m, -c0, cw, z, a);
      if constexpr (DT0 != DT1) {
        if (w.ncol1 > 0 && c0 + cw > w.ncol0) {
          const Part<DT1, vr> P1(arena + w.off1, w.rows_p, w.ncol1, rq, cg, m);
          krun_cols<Acc, DT1, vr, K>(P1, max(c0, w.ncol0) - w.ncol0, c0 + cw - w.ncol0, cg, m,
                                     w.ncol0 - c0, cw, z, a);
        }
      }
    }
  }
  // fixed-order sum over the m column groups of every row
  Acc *red = z;
  __syncthreads();
  if (active) {
#pragma unroll
    for (int i = 0; i < vr; ++i)
#pragma unroll
      for (int j = 0; j < K; ++j) red[(cg * w.rows_p + rq * vr + i) * K + j] = a[i][j];
  }
  __syncthreads();
  for (int e = tid; e < w.rows * K; e += NT) {
    const int r = e / K, j = e % K;
    if (j >= k) continue;
    Acc s = red[r * K + j];
    for (int g = 1; g < m; ++g) s += red[(g * w.rows_p + r) * K + j];
    if constexpr (PHASE == 1) {
      t_out[(size_t)outidx[w.out_base + r] * k + j] = (double)s;
    } else {
      double *py = y + (size_t)(w.out_base + r) * k + j;
      const double v = (double)s;
      *py = beta == 0.0 ? v : __builtin_fma(beta, *py, v);
    }
  }
}

// The kernels are fitted to 64 VGPRs (8 waves per SIMD, as the
// single-vector kernel): the loads per batch, prefetched columns and
// staging loads in flight (kbatch, kprefetch, kstage_batch) are chosen so
// that every instantiation runs without spills.
template <typename Acc, int PHASE, int DT0, int DT1, int NT, int K, bool PAIRS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8, 8))) void
hmv_kstream_kernel(const char *__restrict__ arena, const HmvWork *__restrict__ works,
                   const HmvZSeg *__restrict__ zsegs, const double *__restrict__ x,
                   const int *__restrict__ codes, const double *__restrict__ t_in,
                   const int *__restrict__ outidx, double *__restrict__ t_out,
                   double *__restrict__ y, int k, double beta, const int *__restrict__ skip) {
  if (skip && *skip) return;
  constexpr int vr = rows_per_lane<PHASE, DT0>();
  __shared__ __attribute__((aligned(16))) Acc z[kbuf_elems<Acc, NT, vr, K>()];
  __shared__ int runs[2 * HMV_RUN_MAX];
  kstream_item<Acc, PHASE, DT0, DT1, NT, K, PAIRS>(works[blockIdx.x], z, runs, arena, zsegs, x,
                                                   codes, t_in, outidx, t_out, y, k, beta);
}

template <typename Acc, int PHASE, int DT0, int DT1, int K, bool PAIRS>
void klaunch_list2(const HmvLaunch &L, const HmvWork *work, int n, int nsmall, int nt_small,
                   const double *x, double *t, double *y, int k, double beta, const int *skip,
                   hipStream_t s) {
  const double *t_in = PHASE == 2 ? t : nullptr;
  double *t_out = PHASE == 1 ? t : nullptr;
  if (nsmall > 0) {
    if (nt_small == 64)
      hipLaunchKernelGGL((hmv_kstream_kernel<Acc, PHASE, DT0, DT1, 64, K, PAIRS>), dim3(nsmall),
                         dim3(64), 0, s, L.arena, work, L.zseg, x, L.codes, t_in, L.outidx, t_out,
                         y, k, beta, skip);
    else
      hipLaunchKernelGGL((hmv_kstream_kernel<Acc, PHASE, DT0, DT1, 128, K, PAIRS>), dim3(nsmall),
                         dim3(128), 0, s, L.arena, work, L.zseg, x, L.codes, t_in, L.outidx, t_out,
                         y, k, beta, skip);
  }
  if (n > nsmall)
    hipLaunchKernelGGL((hmv_kstream_kernel<Acc, PHASE, DT0, DT1, HMV_WG_THREADS, K, PAIRS>),
                       dim3(n - nsmall), dim3(HMV_WG_THREADS), 0, s, L.arena, work + nsmall, L.zseg,
                       x, L.codes, t_in, L.outidx, t_out, y, k, beta, skip);
}

// Paired staging (kstage_pairs) for the FP64 column-cooperative staging
// when k is even and x and t are 16-byte aligned.
template <typename Acc, int PHASE, int DT0, int DT1, int K>
void klaunch_list(const HmvLaunch &L, const HmvWork *work, int n, int nsmall, int nt_small,
                  const double *x, double *t, double *y, int k, double beta, const int *skip,
                  hipStream_t s) {
  constexpr int vr = rows_per_lane<PHASE, DT0>();
  if constexpr (sizeof(Acc) == 8 && kstage_batch<Acc, vr, K>() > 0) {
    if (k % 2 == 0 && ((uintptr_t)x | (uintptr_t)t) % 16 == 0) {
      klaunch_list2<Acc, PHASE, DT0, DT1, K, true>(L, work, n, nsmall, nt_small, x, t, y, k, beta,
                                                   skip, s);
      return;
    }
  }
  klaunch_list2<Acc, PHASE, DT0, DT1, K, false>(L, work, n, nsmall, nt_small, x, t, y, k, beta,
                                                skip, s);
}

template <typename Acc, int DT0, int DT1, int K>
void klaunch_phase(const HmvLaunch &L, int phase, const double *x, double *t, double *y, int k,
                   double beta, const int *skip, hipStream_t s) {
  if (phase == 1)  // phase 1 reads only low-rank factors: both parts use DT1's type
    klaunch_list<Acc, 1, DT1, DT1, K>(L, L.work1, L.nwork1, L.nwork1_small, L.nt1_small, x, t,
                                      nullptr, k, 0.0, skip, s);
  else
    klaunch_list<Acc, 2, DT0, DT1, K>(L, L.work2, L.nwork2, L.nwork2_small, L.nt2_small, x, t, y,
                                      k, beta, skip, s);
}

template <int K>
void klaunch_mode(const HmvLaunch &L, int phase, const double *x, double *t, double *y, int k,
                  double beta, const int *skip, hipStream_t s) {
  if (L.mode == 0)
    klaunch_phase<double, HMV_DT_F64, HMV_DT_F64, K>(L, phase, x, t, y, k, beta, skip, s);
  else if constexpr (K > HMV_KSTREAM_MAX_NARROW)
    return;  // not instantiated: hmv_launch_kstream rejects these k
  else if (L.mode == 1)
    klaunch_phase<float, HMV_DT_F32, HMV_DT_F32, K>(L, phase, x, t, y, k, beta, skip, s);
  else if (L.mode == 2)
    klaunch_phase<float, HMV_DT_F32, HMV_DT_BF16, K>(L, phase, x, t, y, k, beta, skip, s);
  else
    klaunch_phase<float, HMV_DT_BF16, HMV_DT_BF16, K>(L, phase, x, t, y, k, beta, skip, s);
}

}  // namespace

// One phase (1: t = V^T X into the t buffer; 2: Y = [A|U][X;T] + beta Y) of
// the H-matmat for 1 <= k <= 8 right-hand sides.
extern "C" hipError_t hmv_launch_kstream(const HmvLaunch *L, int phase, int k, const double *x,
                                         double *t, double *y, double beta, const int *skip,
                                         hipStream_t stream) {
  if (k < 1 || k > (L->mode == 0 ? HMV_KSTREAM_MAX_F64 : HMV_KSTREAM_MAX_NARROW))
    return hipErrorInvalidValue;
  if (k <= 1)
    klaunch_mode<1>(*L, phase, x, t, y, k, beta, skip, stream);
  else if (k <= 2)
    klaunch_mode<2>(*L, phase, x, t, y, k, beta, skip, stream);
  else if (k <= 4)
    klaunch_mode<4>(*L, phase, x, t, y, k, beta, skip, stream);
  else
    klaunch_mode<8>(*L, phase, x, t, y, k, beta, skip, stream);
  return hipGetLastError();
}
