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
#include "hmv_inputs.h"
#include "hmv_internal.h"
#include "hmv_lanes.h"

namespace {
using namespace hmv_lanes;

// Stages inputs of item columns [c0, c0+cw) into z[0, cw).  Threads take
// columns c0+tid, c0+tid+NT, ... in batches of K: all K codes, then all K
// values are fetched before any is waited for.
template <typename Acc, int NT>
__device__ inline void stage(Acc *__restrict__ z, const Inputs &in, int c0, int cw,
                             const double *__restrict__ x, const int *__restrict__ codes,
                             const double *__restrict__ t) {
  with_codes(in, codes, [&](auto code_of) {
    constexpr int K = 4;
    for (int cb = threadIdx.x; cb < cw; cb += K * NT) {
      // two code searches interleaved at a time: four would spill
      int code[K];
#pragma unroll 2
      for (int k = 0; k < K; ++k) code[k] = code_of(c0 + min(cb + k * NT, cw - 1));
      double val[K];
#pragma unroll
      for (int k = 0; k < K; ++k) val[k] = code[k] >= 0 ? x[code[k]] : t[-code[k] - 1];
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (cb + k * NT < cw) z[cb + k * NT] = (Acc)val[k];
    }
  });
}

// Accumulates the lane's n columns of one part, starting at byte offset vb
// (column stride `step` bytes); the input of the lane's j-th column is
// z[zb + j*m].  Loads go in batches of U issued together; the last, partial
// batch gives its missing columns the out-of-range offset (they load zero)
// and the zero slot z[zmax] as input.
template <typename Acc, int DT, int VR, int U>
__device__ inline void accumulate(__amdgpu_buffer_rsrc_t rs, uint32_t vb, int step, int n,
                                  const Acc *__restrict__ z, int zb, int m, int zmax,
                                  Acc (&a)[VR]) {
  using LN = Lane<DT, VR>;
  int j = 0;
  for (; j + U <= n; j += U) {
    typename LN::raw v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) v[k] = LN::load(rs, vb, (j + k) * step);
#pragma unroll
    for (int k = 0; k < U; ++k) LN::fma(v[k], z[zb + (j + k) * m], a);
  }
  if (j < n) {
    typename LN::raw v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) v[k] = LN::load(rs, j + k < n ? vb : kOffNone, (j + k) * step);
#pragma unroll
    for (int k = 0; k < U; ++k) LN::fma(v[k], z[min(zb + (j + k) * m, zmax)], a);
  }
}

// Lane's columns of local range [lo, hi) of a part: local column cg + j*m,
// j in [ceil((lo-cg)/m), ceil((hi-cg)/m)); zshift maps a local column to its
// staged slot (slot = local column + zshift).
template <typename Acc, int DT, int VR>
__device__ inline void run_cols(const Part<DT, VR> &P, int lo, int hi, int cg, int m, int zshift,
                                int zmax, const Acc *__restrict__ z, Acc (&a)[VR]) {
  const int j0 = lo > cg ? (lo - cg + m - 1) / m : 0;
  const int j1 = hi > cg ? (hi - cg + m - 1) / m : 0;
  if (j1 > j0)
    accumulate<Acc, DT, VR, batch_loads<DT, VR>()>(P.rs, P.voff + (uint32_t)(j0 * P.step), P.step,
                                                   j1 - j0, z, cg + j0 * m + zshift, m, zmax, a);
}

// PHASE 1: t[outidx[out_base + r]] = row sums.
// PHASE 2: y[out_base + r] = row sums + beta * y.
// DT0/DT1: storage types of the two column parts (fixed per arena precision).
// NT: threads per workgroup; inputs are staged NT*8 columns at a time.
// One work item w; z: LDS of ZCH + 4 accumulation-type slots (inputs of one
// chunk plus the zero slot; reused for the row reduction).
// Inputs staged per pass: HMV_ZCHUNK per 256 threads in FP64, twice that
// (up to HMV_ZCHUNK) for 4-byte accumulation types, so that a workgroup
// holds the same LDS bytes (at 32 waves per CU the z arrays of one CU take
// 128 KB either way).
template <typename Acc, int NT>
constexpr int zchunk() {
  return HMV_ZCHUNK * NT / HMV_WG_THREADS * (sizeof(Acc) == 4 && NT < HMV_WG_THREADS ? 2 : 1);
}

template <typename Acc, int PHASE, int DT0, int DT1, int NT>
__device__ inline void stream_item(
    const HmvWork &w, Acc *__restrict__ z, int *__restrict__ runs, const char *__restrict__ arena,
    const HmvZSeg *__restrict__ zsegs, const double *__restrict__ x, const int *__restrict__ codes,
    const double *__restrict__ t_in, const int *__restrict__ outidx, double *__restrict__ t_out,
    double *__restrict__ y, double beta) {
  constexpr int ZCH = zchunk<Acc, NT>();
  const int tid = threadIdx.x;
  constexpr int vr = rows_per_lane<PHASE, DT0>();
  static_assert(vr * NT <= ZCH, "reduction buffer fits in z");
  const int q = w.rows_p / vr;
  const int m = NT / q;
  const bool active = tid < q * m;
  const int rq = tid % q, cg = tid / q;
  Acc a[vr];
#pragma unroll
  for (int i = 0; i < vr; ++i) a[i] = Acc(0);
  const int ncols = w.ncol0 + w.ncol1;
  const Part<DT0, vr> P0(arena + w.off0, w.rows_p, w.ncol0, rq, cg, m);
  // The lane's first PF columns of part 0 do not depend on the staged
  // inputs: their loads are issued before staging, so the first HBM round
  // trip overlaps the gathers and the barrier.
  constexpr int PF = vr == 4 ? 2 : 4;  // 4-row lanes: two, so that nothing spills
  const int lim0 = min(w.ncol0, ZCH);
  const int npf = active ? max(0, min(PF, (lim0 - cg + m - 1) / m)) : 0;
  using L0 = Lane<DT0, vr>;
  typename L0::raw pre[PF];
#pragma unroll
  for (int k = 0; k < PF; ++k) pre[k] = L0::load(P0.rs, k < npf ? P0.voff : kOffNone, k * P0.step);
  const Inputs in = load_inputs<NT>(w, zsegs, runs);
  for (int c0 = 0; c0 < ncols; c0 += ZCH) {
    const int cw = min(ZCH, ncols - c0);
    if (c0 > 0) __syncthreads();
    stage<Acc, NT>(z, in, c0, cw, x, codes, t_in);
    if (tid == 0) z[cw] = Acc(0);
    __syncthreads();
    if (active) {
      int lo0 = c0;
      if (c0 == 0) {
#pragma unroll
        for (int k = 0; k < PF; ++k) L0::fma(pre[k], z[min(cg + k * m, cw)], a);
        lo0 = min(npf * m, w.ncol0);  // columns < npf*m are done (for this lane: j < npf)
        lo0 = npf == PF ? lo0 : cw;   // a lane with fewer than PF columns here has none left
      }
      const int hi0 = min(c0 + cw, w.ncol0);
      if (lo0 < hi0) run_cols<Acc, DT0, vr>(P0, lo0, hi0, cg, m, -c0, cw, z, a);
      // an item stored with one type has one part (ncol1 == 0)
      if constexpr (DT0 != DT1) {
        if (w.ncol1 > 0 && c0 + cw > w.ncol0) {
          const Part<DT1, vr> P1(arena + w.off1, w.rows_p, w.ncol1, rq, cg, m);
          run_cols<Acc, DT1, vr>(P1, max(c0, w.ncol0) - w.ncol0, c0 + cw - w.ncol0, cg, m,
                                 w.ncol0 - c0, cw, z, a);
        }
      }
    }
  }
  Acc *red = z;
  __syncthreads();
  if (active) {
#pragma unroll
    for (int i = 0; i < vr; ++i) red[cg * w.rows_p + rq * vr + i] = a[i];
  }
  __syncthreads();
  for (int r = tid; r < w.rows; r += NT) {
    Acc s = red[r];
    for (int g = 1; g < m; ++g) s += red[g * w.rows_p + r];
    if constexpr (PHASE == 1) {
      t_out[outidx[w.out_base + r]] = (double)s;
    } else {
      const int row = w.out_base + r;
      const double v = (double)s;
      y[row] = beta == 0.0 ? v : __builtin_fma(beta, y[row], v);
    }
  }
}

template <typename Acc, int PHASE, int DT0, int DT1, int NT = HMV_WG_THREADS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8, 8))) void
hmv_stream_kernel(
    const char *__restrict__ arena, const HmvWork *__restrict__ works,
    const HmvZSeg *__restrict__ zsegs, const double *__restrict__ x, const int *__restrict__ codes,
    const double *__restrict__ t_in, const int *__restrict__ outidx, double *__restrict__ t_out,
    double *__restrict__ y, double beta, const int *__restrict__ skip,
    const int *__restrict__ winst, const int *__restrict__ inst_done) {
  if (skip && *skip) return;
  // a finished instance of a multi-instance solve needs no more products
  if (inst_done && inst_done[winst[blockIdx.x]]) return;
  __shared__ __attribute__((aligned(16))) Acc z[zchunk<Acc, NT>() + 4];
  __shared__ int runs[2 * HMV_RUN_MAX];
  stream_item<Acc, PHASE, DT0, DT1, NT>(works[blockIdx.x], z, runs, arena, zsegs, x, codes, t_in,
                                        outidx, t_out, y, beta);
}

// t[slot] = sum of the chunk partials t[base + p], p = 0 .. np-1, in order.
__global__ __launch_bounds__(HMV_WG_THREADS) void hmv_tfinal_kernel(
    const HmvTFin *__restrict__ fin, int nfin, double *__restrict__ t, const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int i = blockIdx.x * HMV_WG_THREADS + threadIdx.x;
  if (i >= nfin) return;
  const HmvTFin f = fin[i];
  constexpr int K = 8;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = k < f.np ? t[f.base + k] : 0.0;
  double s = v[0];
#pragma unroll
  for (int k = 1; k < K; ++k)
    if (k < f.np) s += v[k];
  for (int p = K; p < f.np; ++p) s += t[f.base + p];
  t[f.slot] = s;
}

template <typename Acc, int PHASE, int DT0, int DT1>
void launch_list(const HmvLaunch &L, const HmvWork *work, int n, int nsmall, int nt_small,
                 const double *x, double *y, double beta, const int *skip,
                 const int *inst_done, hipStream_t stream) {
  const double *t_in = PHASE == 2 ? L.t : nullptr;
  double *t_out = PHASE == 1 ? L.t : nullptr;
  const int *winst = PHASE == 1 ? L.winst1 : L.winst2;
  if (!winst) inst_done = nullptr;
  if (nsmall > 0) {
    if (nt_small == 64)
      hipLaunchKernelGGL((hmv_stream_kernel<Acc, PHASE, DT0, DT1, 64>), dim3(nsmall), dim3(64), 0,
                         stream, L.arena, work, L.zseg, x, L.codes, t_in, L.outidx, t_out, y, beta,
                         skip, winst, inst_done);
    else
      hipLaunchKernelGGL((hmv_stream_kernel<Acc, PHASE, DT0, DT1, 128>), dim3(nsmall), dim3(128), 0,
                         stream, L.arena, work, L.zseg, x, L.codes, t_in, L.outidx, t_out, y, beta,
                         skip, winst, inst_done);
  }
  if (n > nsmall)
    hipLaunchKernelGGL((hmv_stream_kernel<Acc, PHASE, DT0, DT1>), dim3(n - nsmall),
                       dim3(HMV_WG_THREADS), 0, stream, L.arena, work + nsmall, L.zseg, x, L.codes,
                       t_in, L.outidx, t_out, y, beta, skip, winst ? winst + nsmall : nullptr,
                       inst_done);
}

template <typename Acc, int DT0, int DT1>
void launch_mode(const HmvLaunch &L, int phases, const double *x, double *y, double beta,
                 const int *skip, const int *inst_done, hipStream_t stream) {
  if (phases & HMV_PHASE1) {
    // phase 1 reads only low-rank factors: both parts use DT1's type
    launch_list<Acc, 1, DT1, DT1>(L, L.work1, L.nwork1, L.nwork1_small, L.nt1_small, x, nullptr,
                                  0.0, skip, inst_done, stream);
    if (L.ntfin > 0)
      hipLaunchKernelGGL(hmv_tfinal_kernel, dim3((L.ntfin + HMV_WG_THREADS - 1) / HMV_WG_THREADS),
                         dim3(HMV_WG_THREADS), 0, stream, L.tfin, L.ntfin, L.t, skip);
  }
  if (phases & HMV_PHASE2)
    launch_list<Acc, 2, DT0, DT1>(L, L.work2, L.nwork2, L.nwork2_small, L.nt2_small, x, y, beta,
                                  skip, inst_done, stream);
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

// FP64 MFMA rate: `rounds` x 8 independent 16x16x4 FP64 MFMAs per wave on
// register operands (eight accumulator chains, so that no MFMA waits for
// its predecessor's result).
typedef double hmv_d4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(64) void hmv_mfma_rate_kernel(int rounds, double *__restrict__ out) {
  hmv_d4 acc[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) acc[c] = hmv_d4{0, 0, 0, 0};
  const double a = 1.0 + threadIdx.x * 1e-9, b = 1e-3;
  for (int i = 0; i < rounds; ++i)
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  if (s == 1.2345e-300) out[0] = s;  // keeps the MFMAs alive; never true here
}

}  // namespace

extern "C" hipError_t hmv_launch_stream_test(int kind, const void *src, void *dst, size_t bytes,
                                             hipStream_t stream) {
  const size_t n = bytes / 16;
  const int grid = 256 * 32;
  if (kind == 0)
    hipLaunchKernelGGL(hmv_read_kernel, dim3(grid), dim3(HMV_WG_THREADS), 0, stream,
                       (const double2 *)src, n, (double *)dst);
  else if (kind == 1)
    hipLaunchKernelGGL(hmv_copy_kernel, dim3(grid), dim3(HMV_WG_THREADS), 0, stream,
                       (const double2 *)src, (double2 *)dst, n);
  else  // 4 waves per SIMD on 256 CUs; `bytes` counts MFMAs per wave
    hipLaunchKernelGGL(hmv_mfma_rate_kernel, dim3(HMV_MFMA_RATE_WAVES), dim3(64), 0, stream,
                       (int)(bytes / 8), (double *)dst);
  return hipGetLastError();
}

// Launches the phases selected by `phases` (HMV_PHASE1: t and the
// chunk-partial sums; HMV_PHASE2: y = [A|U][x;t] + beta*y) of one apply.
// inst_done (multi-instance arenas, may be null): workgroups of instances
// whose flag is set return at once.
extern "C" hipError_t hmv_launch_apply(const HmvLaunch *L, int phases, const double *x, double *y,
                                       double beta, const int *skip, const int *inst_done,
                                       hipStream_t stream) {
  if (L->mode == 0)
    launch_mode<double, HMV_DT_F64, HMV_DT_F64>(*L, phases, x, y, beta, skip, inst_done, stream);
  else if (L->mode == 1)
    launch_mode<float, HMV_DT_F32, HMV_DT_F32>(*L, phases, x, y, beta, skip, inst_done, stream);
  else if (L->mode == 2)
    launch_mode<float, HMV_DT_F32, HMV_DT_BF16>(*L, phases, x, y, beta, skip, inst_done, stream);
  else
    launch_mode<float, HMV_DT_BF16, HMV_DT_BF16>(*L, phases, x, y, beta, skip, inst_done, stream);
  return hipGetLastError();
}
