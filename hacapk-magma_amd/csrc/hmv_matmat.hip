// hmv_matmat.hip -- multi-right-hand-side H-matmat Y = H X + beta*Y on MFMA.
//
// X and Y hold k right-hand sides interleaved (row i's k values contiguous,
// N x k row-major); t likewise (slot s's k values at t[s*k]).  The apply runs
// over the same arena and work lists as the single-vector kernel, one wave
// (64 threads) per unit: a work item and a group of up to TW of its 16-row
// tiles.
//
// The wave walks the item's columns four at a time.  Per column quad, one
// 8-byte (or 4/2-byte) load per lane and tile gives the 16 x 4 A fragment
// (lane l: row l%16 of column l/16 of the column-major block), and KT loads
// per lane give the 4 x 16 B fragments x[column][16 tt + l%16] straight
// from the interleaved inputs: phase 1 from its item's x run, phase 2
// through a per-wave LDS table of the columns' input addresses (L2-resident
// inputs: the units of an XCD are neighbouring items, so they share
// columns).  Each (tile, RHS tile) pair then takes one
// MFMA (16x16x4; FP64, or FP32 for FP32/BF16 arenas) into its 16 x 16
// accumulator; the B fragment is shared by the group's tiles.  Loads run
// one group of U quads ahead of the MFMAs.  Columns past the block load
// A = 0 through the buffer range check and B = 0 from a zero row.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "hmv_engine.h"
#include "hmv_inputs.h"
#include "hmv_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kNT = 2;                  // nontemporal cache policy

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

// One A entry of a part as loaded (0 when off), widened to the accumulation
// type at its MFMA (so the load's wait sits at the MFMA).
template <int DT>
struct Entry;
template <>
struct Entry<HMV_DT_F64> {
  static constexpr int es = 8;
  using raw = double;
  __device__ static double load_raw(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, kNT));
  }
  __device__ static double widen(double v) { return v; }
};
template <>
struct Entry<HMV_DT_F32> {
  static constexpr int es = 4;
  using raw = float;
  __device__ static float load_raw(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, kNT));
  }
  __device__ static float widen(float v) { return v; }
};
template <>
struct Entry<HMV_DT_BF16> {
  static constexpr int es = 2;
  using raw = uint32_t;
  __device__ static uint32_t load_raw(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, kNT);
  }
  __device__ static float widen(uint32_t v) { return __uint_as_float(v << 16); }
};

// Tile-group width (tiles sharing one B fragment) by RHS tiles, and column
// quads per prefetch group by RHS tiles and active tiles (two groups of
// U * (NT + KT) loads per lane are live; measured on MI355X, deeper or
// wider groups do not gain).
// At KT = 8 (k > 64) two 16-row tiles share each B fragment with FP32
// MFMAs (k = 128: 4.14 against 4.52 ms at 327,680 unknowns), while FP64
// accumulators (twice the registers) are faster one tile per wave (5.2
// against 5.9 ms).
template <int KT, typename Acc>
struct Shape {  // default tiles per unit (TW) by RHS tile count and accumulator type
  static constexpr int TW = KT <= 1 ? 4 : KT <= 4 ? 2 : sizeof(Acc) == 4 ? 2 : 1;
  static constexpr int LOG_TW = TW == 4 ? 2 : TW == 2 ? 1 : 0;
};
template <int KT, int NT>
constexpr int group_quads() {
  return KT + NT <= 6 ? 4 : 2;
}

// Global-address-space inputs: loads through these are global loads, which
// count on vmcnt only (a generic pointer read from LDS would give flat
// loads, which also count on lgkmcnt and so serialise the LDS reads with
// the inputs in flight).
typedef const double __attribute__((address_space(1))) gdouble;

constexpr int kWin = 1024;  // input pointers tabled per window (LDS, per wave)
constexpr int kPad = 32;    // zero-input entries past a window (one group of columns)
// B source of columns past a window or part (never written)
__device__ __attribute__((aligned(16))) double g_zero_inputs[128];

// Right-hand side held by column j0 of RHS tile tt: 16 tt + j0 up to two
// tiles; from four tiles on the tiles go in pairs, tile 2s + b holding
// 32 s + 2 j0 + b, so that a lane's inputs of the two tiles of a pair are
// adjacent and one 16-byte load brings both (for k = 128 five loads per
// column quad instead of nine: one for A, four for B; 4.62 against 4.96 ms
// at 327,680 unknowns, 2.53 against 2.72 at k = 64, products bit-equal).
template <int KT>
constexpr bool bpaired() {
  return KT >= 4;
}
template <int KT>
__device__ inline int bcol(int tt, int j0) {
  return bpaired<KT>() ? 32 * (tt >> 1) + 2 * j0 + (tt & 1) : 16 * tt + j0;
}

// Per-wave constants of the B loads.  Lane l (j0 = l%16) points at its
// column's inputs + jl (jl = bcol(0, j0)); tile tt's input is jd[tt] =
// bcol(tt, j0) - jl past it.  EXACT (k == 16 KT, inputs 16-byte aligned):
// read at 16 tt, or paired tiles as 16-byte loads; otherwise each input is
// read on its own at min(bcol, k - 1) (inputs j >= k only feed accumulator
// columns that are not stored; the clamp keeps the read inside the
// column's k values).
template <int KT>
struct BLane {
  int j0;  // l % 16: the A row of the lane and its accumulator column
  int jl;
  int jd[KT];
};
typedef double gd2 __attribute__((ext_vector_type(2)));
typedef const gd2 __attribute__((address_space(1))) ggd2;

// One prefetch group: U column quads of A (NT tiles) and B (KT RHS tiles),
// as loaded (widened to Acc at the MFMA).
template <int DT, int KT, int NT, int U>
struct Group {
  typename Entry<DT>::raw a[U][NT];
  double b[U][KT];
};

// Where the B loads of a column come from: the window's pointer table
// (tbl[col - tbase]), or, for a part whose inputs are one x run (phase 1),
// straight from that run (xr + col k) with zeros past column hi.
struct BSrc {
  gdouble *const *tbl;
  int tbase;
  gdouble *xr, *zr;
  int hi;
};

// Loads the group of part columns [q0, q0 + 4U): lane l takes column
// q0 + 4u + l/16.  A: row 16 (tile0 + m) + l%16 at byte offset
// col * cs + r_es + 16 m es; a column past the part lies past the buffer
// range and loads 0, and rows past the block only feed accumulator rows
// that are not stored.  B: from the column's inputs as src gives them
// (zeros past the window or the part).

template <bool EXACT, bool RUN, int DT, int KT, int NT, int U>
__device__ inline void load_group(Group<DT, KT, NT, U> &g, __amdgpu_buffer_rsrc_t rs, uint32_t cs,
                                 uint32_t r_es, int q0, const BSrc &src, int k,
                                 const BLane<KT> &bl) {
  using E = Entry<DT>;
  const int cq = threadIdx.x >> 4;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int col = q0 + 4 * u + cq;
    const uint32_t cb = (uint32_t)col * cs + r_es;
#pragma unroll
    for (int m = 0; m < NT; ++m) g.a[u][m] = E::load_raw(rs, cb + (uint32_t)(m * 16 * E::es));
    gdouble *p;
    if constexpr (RUN)
      p = col < src.hi ? src.xr + (size_t)col * k : src.zr;
    else
      p = src.tbl[col - src.tbase] + bl.jl;
    if constexpr (EXACT && bpaired<KT>()) {
#pragma unroll
      for (int tt = 0; tt < KT; tt += 2) {
        const gd2 v = *(ggd2 *)(p + 32 * (tt >> 1));
        g.b[u][tt] = v.x;
        g.b[u][tt + 1] = v.y;
      }
    } else {
#pragma unroll
      for (int tt = 0; tt < KT; ++tt) g.b[u][tt] = EXACT ? p[16 * tt] : p[bl.jd[tt]];
    }
  }
}

template <typename Acc, int DT, int KT, int TW, int NT, int U>
__device__ inline void mma_group(typename Mfma<Acc>::acc (&acc)[TW][KT],
                                 const Group<DT, KT, NT, U> &g) {
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int m = 0; m < NT; ++m) {
      const Acc a = (Acc)Entry<DT>::widen(g.a[u][m]);
#pragma unroll
      for (int tt = 0; tt < KT; ++tt)
        acc[m][tt] = Mfma<Acc>::op(a, (Acc)g.b[u][tt], acc[m][tt]);
    }
}

// Tables the input pointers of part columns [w0, w1) (item columns cbase +
// ...) at tbl[0 ..), and zeros for the kPad entries that follow.
__device__ inline void fill_window(gdouble **__restrict__ tbl, const Inputs &in,
                                   const int *__restrict__ codes, int cbase, int w0, int w1,
                                   const double *__restrict__ x, const double *__restrict__ t,
                                   int k) {
  for (int i = (int)threadIdx.x; i < w1 - w0 + kPad; i += 64) {
    const double *p = g_zero_inputs;
    if (i < w1 - w0) {
      const int code = input_code(in, codes, cbase + w0 + i);
      p = code >= 0 ? x + (size_t)code * k : t + (size_t)(-1 - code) * k;
    }
    tbl[i] = (gdouble *)p;
  }
}

// Multiplies part columns [lo, hi) into the accumulators, U quads at a
// time, the next group's loads in flight while the current one multiplies.
template <typename Acc, bool EXACT, bool RUN, int DT, int KT, int TW, int NT>
__device__ inline void run_columns(typename Mfma<Acc>::acc (&acc)[TW][KT],
                                   __amdgpu_buffer_rsrc_t rs, uint32_t cs, uint32_t r_es, int lo,
                                   int hi, const BSrc &src, int k, const BLane<KT> &bl) {
  constexpr int U = group_quads<KT, NT>();
  constexpr int S = 4 * U;  // columns per group
  static_assert(S <= kPad, "a group past the window reads zero entries");
  Group<DT, KT, NT, U> g0, g1;
  load_group<EXACT, RUN>(g0, rs, cs, r_es, lo, src, k, bl);
  int q0 = lo;
  for (; q0 + 2 * S < hi; q0 += 2 * S) {  // g0 holds group q0; two more follow
    load_group<EXACT, RUN>(g1, rs, cs, r_es, q0 + S, src, k, bl);
    mma_group<Acc>(acc, g0);
    load_group<EXACT, RUN>(g0, rs, cs, r_es, q0 + 2 * S, src, k, bl);
    mma_group<Acc>(acc, g1);
  }
  const bool last2 = q0 + S < hi;  // one or two groups left
  if (last2) load_group<EXACT, RUN>(g1, rs, cs, r_es, q0 + S, src, k, bl);
  mma_group<Acc>(acc, g0);
  if (last2) mma_group<Acc>(acc, g1);
}

// Accumulates the product of NT tiles over one part (ncolp columns stored
// with DT at rs, item columns cbase + [0, ncolp)).  Inputs of one x run
// (phase 1) are addressed directly; otherwise kWin columns at a time, the
// window's input pointers tabled in LDS first.
template <typename Acc, bool EXACT, bool RUN, int DT, int KT, int TW, int NT>
__device__ inline void part_mma(typename Mfma<Acc>::acc (&acc)[TW][KT], __amdgpu_buffer_rsrc_t rs,
                                int rows_p, int tile0, int ncolp, int cbase, const Inputs &in,
                                const int *__restrict__ codes, gdouble **__restrict__ tbl,
                                const double *__restrict__ x, const double *__restrict__ t, int k,
                                const BLane<KT> &bl) {
  if (ncolp <= 0) return;
  const uint32_t cs = (uint32_t)rows_p * Entry<DT>::es;
  const uint32_t r_es = (uint32_t)(tile0 * 16 + bl.j0) * Entry<DT>::es;
  if constexpr (RUN) {
    const BSrc src{nullptr, 0, (gdouble *)(x + (size_t)(in.z0.src + cbase) * k) + bl.jl,
                   (gdouble *)g_zero_inputs + bl.jl, ncolp};
    run_columns<Acc, EXACT, true, DT, KT, TW, NT>(acc, rs, cs, r_es, 0, ncolp, src, k, bl);
  } else {
    for (int w0 = 0; w0 < ncolp; w0 += kWin) {
      const int w1 = min(ncolp, w0 + kWin);
      __syncthreads();  // the table's previous contents are read
      fill_window(tbl, in, codes, cbase, w0, w1, x, t, k);
      __syncthreads();
      const BSrc src{tbl, w0, nullptr, nullptr, w1};
      run_columns<Acc, EXACT, false, DT, KT, TW, NT>(acc, rs, cs, r_es, w0, w1, src, k, bl);
    }
  }
}

// Both parts of the item for NT tiles (the unit's tile count, fixed at
// compile time so that the accumulators stay in place); an item stored
// with one type has one part.
template <typename Acc, bool EXACT, bool RUN, int DT0, int DT1, int KT, int TW, int NT>
__device__ inline void unit_mma(typename Mfma<Acc>::acc (&acc)[TW][KT], const HmvWork &w,
                                __amdgpu_buffer_rsrc_t rs0, __amdgpu_buffer_rsrc_t rs1, int tile0,
                                const Inputs &in, const int *__restrict__ codes,
                                gdouble **__restrict__ tbl, const double *__restrict__ x,
                                const double *__restrict__ t, int k, const BLane<KT> &bl) {
  part_mma<Acc, EXACT, RUN, DT0, KT, TW, NT>(acc, rs0, w.rows_p, tile0, w.ncol0, 0, in, codes,
                                             tbl, x, t, k, bl);
  if constexpr (DT0 != DT1)
    part_mma<Acc, EXACT, RUN, DT1, KT, TW, NT>(acc, rs1, w.rows_p, tile0, w.ncol1, w.ncol0, in,
                                               codes, tbl, x, t, k, bl);
}

// One wave per unit {item, tile0 << 8 | nt}: tiles tile0 .. tile0+nt-1 of
// the item.  PHASE 1: t[outidx[out_base + row] * k + j];  PHASE 2:
// y[(out_base + row) * k + j].
template <typename Acc, int PHASE, int DT0, int DT1, int KT, int TW, bool EXACT>
__global__ __launch_bounds__(64) void hmv_matmat_kernel(
    const char *__restrict__ arena, const HmvWork *__restrict__ works, const int2 *__restrict__ units,
    const HmvZSeg *__restrict__ zsegs, const double *__restrict__ x, const int *__restrict__ codes,
    const double *__restrict__ t_in, const int *__restrict__ outidx, double *__restrict__ t_out,
    double *__restrict__ y, int k, double beta, const int *__restrict__ skip) {
  if (skip && *skip) return;
  static_assert(TW == 1 || TW == 2 || TW == 4, "tiles per unit");
  __shared__ int runs[2 * HMV_RUN_MAX];
  __shared__ gdouble *tbl[PHASE == 1 ? 1 : kWin + kPad];  // phase 1 reads its x run directly
  const int2 unit = units[blockIdx.x];
  const HmvWork w = works[unit.x];
  const int tile0 = unit.y >> 8, nt = unit.y & 255;
  const int lane = threadIdx.x;
  BLane<KT> bl;
  bl.j0 = lane & 15;
  bl.jl = bcol<KT>(0, bl.j0);
#pragma unroll
  for (int tt = 0; tt < KT; ++tt) bl.jd[tt] = min(bcol<KT>(tt, bl.j0), k - 1) - bl.jl;
  const Inputs in = load_inputs<64>(w, zsegs, runs);
  const __amdgpu_buffer_rsrc_t rs0 =
      rsrc(arena + w.off0, (uint32_t)((size_t)w.rows_p * w.ncol0 * Entry<DT0>::es));
  const __amdgpu_buffer_rsrc_t rs1 =
      rsrc(arena + (w.ncol1 > 0 ? w.off1 : w.off0),
           (uint32_t)((size_t)w.rows_p * w.ncol1 * Entry<DT1>::es));
  typename Mfma<Acc>::acc acc[TW][KT];
#pragma unroll
  for (int m = 0; m < TW; ++m)
#pragma unroll
    for (int tt = 0; tt < KT; ++tt) acc[m][tt] = typename Mfma<Acc>::acc{0, 0, 0, 0};
  if (nt >= TW)
    unit_mma<Acc, EXACT, PHASE == 1, DT0, DT1, KT, TW, TW>(acc, w, rs0, rs1, tile0, in, codes, tbl, x, t_in, k,
                                               bl);
  else if constexpr (TW == 2)
    unit_mma<Acc, EXACT, PHASE == 1, DT0, DT1, KT, TW, 1>(acc, w, rs0, rs1, tile0, in, codes, tbl, x, t_in, k,
                                              bl);
  else if constexpr (TW == 4) {
    if (nt == 1)
      unit_mma<Acc, EXACT, PHASE == 1, DT0, DT1, KT, TW, 1>(acc, w, rs0, rs1, tile0, in, codes, tbl, x, t_in,
                                                k, bl);
    else if (nt == 2)
      unit_mma<Acc, EXACT, PHASE == 1, DT0, DT1, KT, TW, 2>(acc, w, rs0, rs1, tile0, in, codes, tbl, x, t_in,
                                                k, bl);
    else
      unit_mma<Acc, EXACT, PHASE == 1, DT0, DT1, KT, TW, 3>(acc, w, rs0, rs1, tile0, in, codes, tbl, x, t_in,
                                                k, bl);
  }
  // lane holds rows Mfma::row(lane, i), column lane%16 of each 16 x 16 tile
  const int j0 = lane & 15;
#pragma unroll
  for (int m = 0; m < TW; ++m) {
    if (m >= nt) continue;
#pragma unroll
    for (int tt = 0; tt < KT; ++tt) {
      const int j = bcol<KT>(tt, j0);
      if (j >= k) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = (tile0 + m) * 16 + Mfma<Acc>::row(lane, i);
        if (row >= w.rows) continue;
        const double v = (double)acc[m][tt][i];
        if constexpr (PHASE == 1) {
          t_out[(size_t)outidx[w.out_base + row] * k + j] = v;
        } else {
          double *py = y + (size_t)(w.out_base + row) * k + j;
          *py = beta == 0.0 ? v : __builtin_fma(beta, *py, v);
        }
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
void launch_units(int kt, const HmvLaunch &L, const HmvUnits &U, const HmvWork *works,
                  const double *x, double *t, double *y, int k, double beta, const int *skip,
                  hipStream_t s) {
  const double *t_in = PHASE == 2 ? t : nullptr;
  double *t_out = PHASE == 1 ? t : nullptr;
  // EXACT reads the inputs in 16-byte pairs: x and t must be 16-byte aligned
  // (a caller's x may start anywhere; t is allocated by the engine)
  const bool exact = ((uintptr_t)x % 16 == 0) && ((uintptr_t)t % 16 == 0);
#define HMV_MM_GO(KT, TW, EX)                                                                  \
  hipLaunchKernelGGL((hmv_matmat_kernel<Acc, PHASE, DT0, DT1, KT, TW, EX>), dim3(n), dim3(64), 0, \
                     s, L.arena, works, U.u[PHASE - 1][w], L.zseg, x, L.codes, t_in, L.outidx,      \
                     t_out, y, k, beta, skip)
#define HMV_MM_LAUNCH(KT, TW)                                      \
  do {                                                             \
    const int w = TW == 4 ? 2 : TW == 2 ? 1 : 0, n = U.n[PHASE - 1][w]; \
    if (n > 0) {                                                   \
      if (exact && k == 16 * KT)                                   \
        HMV_MM_GO(KT, TW, true);                                   \
      else                                                         \
        HMV_MM_GO(KT, TW, false);                                  \
    }                                                              \
  } while (0)
  // tiles per unit: the tuned choice of the KT class when it is one of the
  // two instantiated for it, else the default (Shape)
  const int lt = U.log_tw[kt == 1 ? 0 : kt == 2 ? 1 : kt <= 4 ? 2 : 3];
  if (kt == 1) {
    if (lt == 1 || (lt < 0 && Shape<1, Acc>::TW == 2))
      HMV_MM_LAUNCH(1, 2);
    else
      HMV_MM_LAUNCH(1, 4);
  } else if (kt == 2) {
    if (lt == 0 || (lt < 0 && Shape<2, Acc>::TW == 1))
      HMV_MM_LAUNCH(2, 1);
    else
      HMV_MM_LAUNCH(2, 2);
  } else if (kt <= 4) {
    if (lt == 0 || (lt < 0 && Shape<4, Acc>::TW == 1))
      HMV_MM_LAUNCH(4, 1);
    else
      HMV_MM_LAUNCH(4, 2);
  } else {
    if (lt == 1 || (lt < 0 && Shape<8, Acc>::TW == 2))
      HMV_MM_LAUNCH(8, 2);
    else
      HMV_MM_LAUNCH(8, 1);
  }
#undef HMV_MM_LAUNCH
#undef HMV_MM_GO
}

template <typename Acc, int DT0, int DT1>
void launch_block(const HmvLaunch &L, const HmvUnits &U, int k, const double *x, double *t,
                  double *y, double beta, const int *skip, hipStream_t s) {
  const int kt = (k + 15) / 16;
  // few right-hand sides stream on the vector units (hmv_kstream.hip) unless
  // the arena's tuning sends their class to MFMA
  const bool vec = k <= (L.mode == 0 ? HMV_KSTREAM_MAX_F64 : HMV_KSTREAM_MAX_NARROW) &&
                   U.vec[k <= 1 ? 0 : k <= 2 ? 1 : k <= 4 ? 2 : 3];
  if (vec)
    (void)hmv_launch_kstream(&L, 1, k, x, t, nullptr, 0.0, skip, s);
  else
    launch_units<Acc, 1, DT1, DT1>(kt, L, U, L.work1, x, t, nullptr, k, 0.0, skip, s);
  if (L.ntfin > 0) {
    const int64_t n = (int64_t)L.ntfin * k;
    hipLaunchKernelGGL(hmv_tfinal_block_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       L.tfin, L.ntfin, k, t, skip);
  }
  if (vec)
    (void)hmv_launch_kstream(&L, 2, k, x, t, y, beta, skip, s);
  else
    launch_units<Acc, 2, DT0, DT1>(kt, L, U, L.work2, x, t, y, k, beta, skip, s);
}

}  // namespace

// Y = H X + beta*Y for k (1..128) interleaved right-hand sides; t holds
// tpart_size * k doubles.
extern "C" hipError_t hmv_launch_block(const HmvLaunch *L, const HmvUnits *units, int k,
                                       const double *x, double *t, double *y, double beta,
                                       const int *skip, hipStream_t stream) {
  if (L->mode == 0)
    launch_block<double, HMV_DT_F64, HMV_DT_F64>(*L, *units, k, x, t, y, beta, skip, stream);
  else if (L->mode == 1)
    launch_block<float, HMV_DT_F32, HMV_DT_F32>(*L, *units, k, x, t, y, beta, skip, stream);
  else if (L->mode == 2)
    launch_block<float, HMV_DT_F32, HMV_DT_BF16>(*L, *units, k, x, t, y, beta, skip, stream);
  else
    launch_block<float, HMV_DT_BF16, HMV_DT_BF16>(*L, *units, k, x, t, y, beta, skip, stream);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Static workspace of hmv_apply_block_static: the H-matmat's t buffer
// (tpart_size * k doubles) in a file-scope device array instead of memory
// allocated on first use.  Every device the library runs on holds its own
// copy until the process ends; host code never passes `scratch` itself (on
// the host it names a stand-in, not device memory): the device address of
// each device's copy comes from hipGetSymbolAddress, taken once per device
// with that device current, and only that address reaches the kernels.
// Nothing in it needs clearing between calls: phase 1 writes every t slot
// and chunk partial that the finalize and phase 2 read.  Calls that overlap
// on one device share the array and race.
static __device__ __align__(256) double scratch[HMV_STATIC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};
}  // namespace

namespace hmv {
// Device address of the current device's copy of `scratch`, looked up on
// the first call per device.
int static_scratch(double **out) {
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "static scratch: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("static scratch: device index " + std::to_string(dev) + " out of range");
    return HMV_ERR_DEVICE;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "static scratch: hipGetSymbolAddress");
    if (!p) {
      set_error("static scratch: no device address");
      return HMV_ERR_DEVICE;
    }
    g_scratch[dev] = static_cast<double *>(p);
  }
  *out = g_scratch[dev];
  return HMV_OK;
}
}  // namespace hmv

extern "C" int hmv_apply_block_static(hmv_arena a, int k, const double *x, double *y, double beta,
                                      void *stream) {
  if (!a || !x || !y || k < 1 || k > HMV_MAX_RHS) {
    hmv::set_error("hmv_apply_block_static: invalid argument (k must be 1..128)");
    return HMV_ERR_ARG;
  }
  const int64_t need = std::max<int64_t>(1, a->tpart_size) * k * (int64_t)sizeof(double);
  if (need > (int64_t)HMV_STATIC_SCRATCH_BYTES) {
    hmv::set_error("hmv_apply_block_static: the product needs " + std::to_string(need) +
                   " bytes of t buffer (k = " + std::to_string(k) +
                   "), the static scratch holds " + std::to_string(HMV_STATIC_SCRATCH_BYTES));
    return HMV_ERR_ARG;
  }
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hmv::hip_fail(e, "hmv_apply_block_static");
  if (dev != a->device) {
    hmv::set_error("hmv_apply_block_static: the arena lives on device " +
                   std::to_string(a->device) + ", the current device is " + std::to_string(dev));
    return HMV_ERR_ARG;
  }
  double *t = nullptr;
  int rc = hmv::static_scratch(&t);
  if (rc) return rc;
  HmvUnits u;
  hmv::fill_units(a, u);
  e = hmv_launch_block(&a->launch, &u, k, x, t, y, beta, nullptr, (hipStream_t)stream);
  if (e != hipSuccess) return hmv::hip_fail(e, "hmv_apply_block_static: launch");
  return HMV_OK;
}
