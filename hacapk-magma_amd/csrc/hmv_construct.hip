// hmv_construct.hip -- the H-matrix fill on the device: dense leaves by entry
// evaluation, low-rank leaves by ACA+, and the arena fill from device
// factors.
//
// The reference fills its leaves on the CPU (HACApK_generate,
// m_HACApK_base.f90:1005-1051: HACApK_acaplus per admissible block,
// :783-975, and the entry function HACApK_entry_ij per dense entry).  Here
// the host keeps the cluster tree and the block list (cheap), and the device
// evaluates every entry and runs every block's ACA+ in parallel: one
// workgroup per block, all of its lanes on the block's rows/columns, the
// scalar steps of ACA+ in the same order and arithmetic as the host
// restatement (the cross updates, the pivot searches with their first-index
// ties, and the overflow-safe norm, which stays sequential because its
// scaling depends on the order).  Floating-point contraction is off in this
// file, so entries and updates round as the host's do.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "hmv.h"
#include "hmv_engine.h"
#include "hmv_internal.h"
#include "hmv_plan.h"

#pragma clang fp contract(off)

namespace {

constexpr int kT = 256;
constexpr int kCapFirst = 64;  // factor columns reserved per block in the first pass

// ---- entry function --------------------------------------------------------

// m_ppohBEM_user_func.f90:42-43 (default-REAL literals in REAL(8) parameters)
__device__ constexpr double kPiF = (double)3.141592653589793f;
__device__ constexpr double kEps0F = (double)8.854187818f * 1e-12;
constexpr int kKernelPpohbem = 0;
constexpr int kKernelCentroid = 1;

struct Geo {
  const double *cent;   // nface x 3, original order
  const double *area;   // nface
  const double *nodes;  // nnode x 3
  const int *faces;     // nface x 3, 0-based
  const int *lod;       // cluster order -> original (1-based)
  int kernel;
};

__device__ inline void cross3(const double u[3], const double v[3], double w[3]) {
  w[0] = u[1] * v[2] - u[2] * v[1];
  w[1] = u[2] * v[0] - u[0] * v[2];
  w[2] = u[0] * v[1] - u[1] * v[0];
}
__device__ inline double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Signed sum of the analytic 1/r potential of a flat triangle at (x,y,z),
// face_integral (m_ppohBEM_user_func.f90:49-100) term by term.
__device__ double triangle_potential_sum(const double xs[3], const double ys[3],
                                         const double zs[3], double x, double y, double z) {
  double r[3];
  for (int i = 0; i < 3; ++i) {
    const double dx = xs[i] - x, dy = ys[i] - y, dz = zs[i] - z;
    r[i] = sqrt(dx * dx + dy * dy + dz * dz);
  }
  double u[3] = {xs[1] - xs[0], ys[1] - ys[0], zs[1] - zs[0]};
  double v[3] = {xs[2] - xs[1], ys[2] - ys[1], zs[2] - zs[1]};
  double w[3];
  cross3(u, v, w);
  const double wn = sqrt(dot3(w, w));
  w[0] /= wn;
  w[1] /= wn;
  w[2] /= wn;
  u[0] = x - xs[0];
  u[1] = y - ys[0];
  u[2] = z - zs[0];
  const double zp = dot3(u, w);
  const double ox = x - zp * w[0], oy = y - zp * w[1], oz = z - zp * w[2];
  const double zpabs = fabs(zp);
  double sum = 0.0;
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3;
    u[0] = xs[j] - ox;
    u[1] = ys[j] - oy;
    u[2] = zs[j] - oz;
    const double xj = sqrt(dot3(u, u));
    u[0] /= xj;
    u[1] /= xj;
    u[2] /= xj;
    cross3(w, u, v);
    const double xi = (xs[i] - ox) * u[0] + (ys[i] - oy) * u[1] + (zs[i] - oz) * u[2];
    const double yi = (xs[i] - ox) * v[0] + (ys[i] - oy) * v[1] + (zs[i] - oz) * v[2];
    const double dx = xj - xi, dy = -yi;
    const double t = sqrt(dx * dx + dy * dy);
    const double l = dx / t, m = dy / t;
    const double d = l * yi - m * xi;
    const double ti = l * xi + m * yi, tj = l * xj;
    const double theta = atan2(yi, xi);
    const double omega = theta - atan2(r[i] * d, zpabs * ti) + atan2(r[j] * d, zpabs * tj);
    const double q = log((r[j] + tj) / (r[i] + ti));
    const double g = d * q - zpabs * omega;
    sum = sum + g;
  }
  return sum;
}

// HACApK_entry_ij(i, j) for 1-based original indices: collocation at the
// centroid of face i, source face j (m_ppohBEM_matrix_element_ij.f90:49-76).
__device__ double entry(const Geo &g, int i, int j) {
  const int fi = i - 1, fj = j - 1;
  if (g.kernel == kKernelCentroid && fi != fj) {
    const double dx = g.cent[3 * fi] - g.cent[3 * fj];
    const double dy = g.cent[3 * fi + 1] - g.cent[3 * fj + 1];
    const double dz = g.cent[3 * fi + 2] - g.cent[3 * fj + 2];
    return g.area[fj] / (4.0 * M_PI * sqrt(dx * dx + dy * dy + dz * dz));
  }
  double xs[3], ys[3], zs[3];
  for (int a = 0; a < 3; ++a) {
    const int nn = g.faces[3 * fj + a];
    xs[a] = g.nodes[3 * nn];
    ys[a] = g.nodes[3 * nn + 1];
    zs[a] = g.nodes[3 * nn + 2];
  }
  const double s =
      triangle_potential_sum(xs, ys, zs, g.cent[3 * fi], g.cent[3 * fi + 1], g.cent[3 * fi + 2]);
  if (g.kernel == kKernelPpohbem) return fabs(s) / (4.0 * kPiF * kEps0F);
  return fabs(s) / (4.0 * M_PI);
}

// ---- leaves ----------------------------------------------------------------

struct CLeaf {
  int ltmtx, nstrtl, ndl, nstrtt, ndt;
  int cap;        // low rank: factor columns reserved at base
  double *base;   // dense: A (ndl rows of ndt); low rank: V (ndt x cap), then U (ndl x cap)
  double *pa;     // low rank scratch: reference column (ndl), then reference row (ndt)
  int *msk;       // low rank scratch: row mask (ndl), then column mask (ndt)
};

struct AcaParams {
  double eps, eps_factor;
  int kmax, kmin;
};

enum { kAcaOk = 0, kAcaOverflow = 1 };

__global__ __launch_bounds__(kT) void k_dense(const CLeaf *__restrict__ lv,
                                              const int *__restrict__ ids, Geo g) {
  const CLeaf L = lv[ids[blockIdx.x]];
  const int64_t n = (int64_t)L.ndl * L.ndt;
  for (int64_t e = threadIdx.x; e < n; e += kT) {
    const int il = (int)(e / L.ndt), it = (int)(e % L.ndt);
    L.base[e] = entry(g, g.lod[L.nstrtl - 1 + il], g.lod[L.nstrtt - 1 + it]);
  }
}

// ACA+ runs one wave per block: the block's scalar steps are sequential, so
// a wave keeps them cheap (reductions by lane shuffles, no workgroup
// barrier between waves) and many blocks run at once.
constexpr int kA = 64;

// Helpers called by every lane with the same arguments; each returns the
// same value in every lane, and ends with the lanes' writes visible.

__device__ inline void wave_sync() { __syncthreads(); }

// First index (1-based) of the largest (want_max) or smallest |vec|, as
// HACApK_maxabsvalloc_d / HACApK_minabsvalloc_d scan.
__device__ void abs_loc(const double *vec, int n, bool want_max, double &val, int &idx) {
  double bv = 0.0;
  int bi = n;  // none
  for (int a = threadIdx.x; a < n; a += kA) {
    const double x = fabs(vec[a]);
    if (bi == n || (want_max ? x > bv : x < bv)) {
      bv = x;
      bi = a;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(bv, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    const bool take =
        oi != n && (bi == n || (want_max ? ov > bv : ov < bv) || (ov == bv && oi < bi));
    if (take) {
      bv = ov;
      bi = oi;
    }
  }
  val = bv;
  idx = bi + 1;
}

// HACApK_unrm_d, bit for bit.  Its running scale zscl is the prefix maximum
// of the |entries| (those below 1e-30 skipped), so every entry's ratio
// (zscl/zza or zza/zscl) comes from a wave-wide prefix-max scan and is formed
// by the entry's own lane; only the accumulation of zzz, whose rounding order
// matters, runs entry by entry (uniformly in every lane).
__device__ inline double lane_double(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ double unrm(const double *za, int nd) {
  const int lane = threadIdx.x;
  double zscl = 0.0, zzz = 1.0;  // same value in every lane
  for (int c0 = 0; c0 < nd; c0 += 64) {
    const int i = c0 + lane;
    const double a = i < nd ? fabs(za[i]) : 0.0;
    const bool skip = i >= nd || a < 1.0e-30;
    double incl = skip ? 0.0 : a;  // inclusive prefix max within the chunk
    for (int off = 1; off < 64; off <<= 1) {
      const double o = __shfl_up(incl, off, 64);
      if (lane >= off && o > incl) incl = o;
    }
    double excl = __shfl_up(incl, 1, 64);
    if (lane == 0) excl = 0.0;
    const double m = excl > zscl ? excl : zscl;  // zscl before entry i
    const bool isnew = !skip && m < a;
    const double q = skip ? 0.0 : (isnew ? m / a : a / m);
    const unsigned long long newm = __ballot(isnew), skipm = __ballot(skip);
    const int nl = min(64, nd - c0);
    for (int l = 0; l < nl; ++l) {
      if ((skipm >> l) & 1ull) continue;
      const double ql = lane_double(q, l);
      if ((newm >> l) & 1ull)
        zzz = 1.0 + zzz * ql * ql;
      else
        zzz = zzz + ql * ql;
    }
    const double cm = lane_double(incl, nl - 1);
    if (cm > zscl) zscl = cm;
  }
  return zscl * sqrt(zzz);
}

// Residual column (col = true: ndl entries of column ip) or row (ndt entries
// of row ip) of a block after k crosses (HACApK_calc_vec): the entries of
// unmasked positions, minus the accumulated cross products (summed in cross
// order before the subtraction), masked positions zero once k > 0.
__device__ void calc_vec(const CLeaf &L, const Geo &g, bool col, int k, int ip, double *vec,
                         const double *U, const double *V) {
  const int n = col ? L.ndl : L.ndt;
  const int *m = col ? L.msk : L.msk + L.ndl;
  for (int a = threadIdx.x; a < n; a += kA) {
    double e;
    if (m[a] == 0)
      e = col ? entry(g, g.lod[a + L.nstrtl - 1], g.lod[ip + L.nstrtt - 2])
              : entry(g, g.lod[ip + L.nstrtl - 2], g.lod[a + L.nstrtt - 1]);
    else
      e = vec[a];
    if (k > 0) {
      double t = 0.0;
      for (int it = 0; it < k; ++it) {
        const double x = col ? U[(size_t)it * L.ndl + a] : V[(size_t)it * L.ndt + a];
        const double y = col ? V[(size_t)it * L.ndt + (ip - 1)] : U[(size_t)it * L.ndl + (ip - 1)];
        t = t + x * y;
      }
      e = e - t;
      if (m[a] == 1) e = 0.0;
    }
    vec[a] = e;
  }
  wave_sync();
}

__device__ void copy_vec(double *dst, const double *src, int n) {
  for (int a = threadIdx.x; a < n; a += kA) dst[a] = src[a];
  wave_sync();
}

__device__ void scale_vec(double *v, int n, double s) {
  for (int a = threadIdx.x; a < n; a += kA) v[a] = v[a] * s;
  wave_sync();
}

// y += x * s
__device__ void axpy_vec(double *y, const double *x, int n, double s) {
  for (int a = threadIdx.x; a < n; a += kA) y[a] = y[a] + x[a] * s;
  wave_sync();
}

__device__ void set_mask(int *m, int i1) {
  if (threadIdx.x == 0) m[i1 - 1] = 1;
  wave_sync();
}

// HACApK_acaplus (m_HACApK_base.f90:783-975) with param(61) = 1, for one
// block per workgroup; restates the host fill's ACA+ step for step.
__global__ __launch_bounds__(kA) void k_aca(const CLeaf *__restrict__ lv,
                                            const int *__restrict__ ids, Geo g, AcaParams p,
                                            int *__restrict__ kt_out, int *__restrict__ status) {
  const int id = ids[blockIdx.x];
  const CLeaf L = lv[id];
  const int ndl = L.ndl, ndt = L.ndt, cap = L.cap;
  double *V = L.base;                    // zab(ndt, cap)
  double *U = L.base + (size_t)ndt * cap;  // zaa(ndl, cap)
  double *pa_ref = L.pa, *pb_ref = L.pa + ndl;
  int *lrow_msk = L.msk, *lcol_msk = L.msk + ndl;
  auto ucol = [&](int k1) { return U + (size_t)(k1 - 1) * ndl; };
  auto vcol = [&](int k1) { return V + (size_t)(k1 - 1) * ndt; };

  for (int a = threadIdx.x; a < ndl + ndt; a += kA) L.msk[a] = 0;
  for (int a = threadIdx.x; a < ndl + ndt; a += kA) L.pa[a] = 0.0;
  wave_sync();

  const double eps = p.eps;
  const double ACA_EPS = p.eps_factor * eps;
  const double za_ACA_EPS = (double)1.0e-30f;
  const double piv_min = (double)1.0e-20f;
  int ntries = max(ndl, ndt) + 1, ntries_row = 6, ntries_col = 6;
  int k = 0;
  int st = kAcaOk;

  int j_ref = 1;
  calc_vec(L, g, true, k, j_ref, pa_ref, U, V);
  double colnorm = unrm(pa_ref, ndl);
  double rownorm;
  int i_ref;
  abs_loc(pa_ref, ndl, false, rownorm, i_ref);
  calc_vec(L, g, false, k, i_ref, pb_ref, U, V);
  rownorm = unrm(pb_ref, ndt);

  double apxnorm = 0.0;
  int lstop_aca = 0;
  int i = 0, j = 0;
  double col_maxval, row_maxval;
  while ((k < p.kmax) && (ntries_row > 0 || ntries_col > 0) && (ntries > 0)) {
    ntries = ntries - 1;
    if (k + 1 > cap) {
      st = kAcaOverflow;
      break;
    }
    double *pcol = ucol(k + 1);
    double *prow = vcol(k + 1);
    abs_loc(pa_ref, ndl, true, col_maxval, i);
    abs_loc(pb_ref, ndt, true, row_maxval, j);

    if (row_maxval > col_maxval) {
      if (j != j_ref)
        calc_vec(L, g, true, k, j, pcol, U, V);
      else
        copy_vec(pcol, pa_ref, ndl);
      abs_loc(pcol, ndl, true, col_maxval, i);
      if (col_maxval < ACA_EPS && k >= p.kmin) {
        lstop_aca = 1;
      } else {
        calc_vec(L, g, false, k, i, prow, U, V);
        const double piv = pcol[i - 1];
        wave_sync();  // every lane has the pivot before the column is scaled
        if (fabs(piv) > piv_min) {
          scale_vec(pcol, ndl, 1.0 / piv);
        } else {
          k = max(k - 1, 0);
          break;
        }
      }
    } else {
      if (i != i_ref)
        calc_vec(L, g, false, k, i, prow, U, V);
      else
        copy_vec(prow, pb_ref, ndt);
      abs_loc(prow, ndt, true, row_maxval, j);
      if (row_maxval < ACA_EPS && k >= p.kmin) {
        lstop_aca = 1;
      } else {
        calc_vec(L, g, true, k, j, pcol, U, V);
        const double piv = prow[j - 1];
        wave_sync();  // every lane has the pivot before the row is scaled
        if (fabs(piv) > piv_min) {
          scale_vec(prow, ndt, 1.0 / piv);
        } else {
          k = max(k - 1, 0);
          break;
        }
      }
    }
    set_mask(lrow_msk, i);
    set_mask(lcol_msk, j);

    if (i != i_ref) {
      const double zinv = -pcol[i_ref - 1];
      axpy_vec(pb_ref, prow, ndt, zinv);
      rownorm = unrm(pb_ref, ndt);
    }
    if (i == i_ref || rownorm < ACA_EPS) {
      if (i == i_ref) ntries_row = ntries_row + 1;
      if (ntries_row > 0) {
        rownorm = 0.0;
        i = i_ref;
        while (i != ((i_ref + ndl - 2) % ndl) + 1 && rownorm < za_ACA_EPS && ntries_row > 0) {
          if (lrow_msk[i - 1] == 0) {
            if (k + 1 > cap) {
              st = kAcaOverflow;
              break;
            }
            calc_vec(L, g, false, k + 1, i, pb_ref, U, V);
            rownorm = unrm(pb_ref, ndt);
            if (rownorm < ACA_EPS) set_mask(lrow_msk, i);
            ntries_row = ntries_row - 1;
          } else {
            rownorm = 0.0;
          }
          i = (i % ndl) + 1;
        }
        if (st != kAcaOk) break;
        i_ref = ((i + ndl - 2) % ndl) + 1;
      }
    }

    if (j != j_ref) {
      const double zinv = -prow[j_ref - 1];
      axpy_vec(pa_ref, pcol, ndl, zinv);
      colnorm = unrm(pa_ref, ndl);
    }
    if (j == j_ref || colnorm < ACA_EPS) {
      if (j == j_ref) ntries_col = ntries_col + 1;
      if (ntries_col > 0) {
        colnorm = 0.0;
        j = j_ref;
        while (j != ((j_ref + ndt - 2) % ndt) + 1 && colnorm < za_ACA_EPS && ntries_col > 0) {
          if (lcol_msk[j - 1] == 0) {
            if (k + 1 > cap) {
              st = kAcaOverflow;
              break;
            }
            calc_vec(L, g, true, k + 1, j, pa_ref, U, V);
            colnorm = unrm(pa_ref, ndl);
            if (colnorm < ACA_EPS) set_mask(lcol_msk, j);
            ntries_col = ntries_col - 1;
          } else {
            colnorm = 0.0;
          }
          j = (j % ndt) + 1;
        }
        if (st != kAcaOk) break;
        j_ref = ((j + ndt - 2) % ndt) + 1;
      }
    }

    if (colnorm < ACA_EPS && rownorm < ACA_EPS && k >= p.kmin) {
      lstop_aca = 1;
      k = k + 1;
    }
    if (lstop_aca == 0) {
      const double blknorm = unrm(pcol, ndl) * unrm(prow, ndt);
      if (k == 0) {
        apxnorm = blknorm;
      } else {
        if (blknorm < apxnorm * eps && rownorm < apxnorm * eps && colnorm < apxnorm * eps &&
            k >= p.kmin)
          lstop_aca = 1;
      }
    }
    if (lstop_aca == 1 && k >= p.kmin) break;
    k = k + 1;
  }
  if (st == kAcaOk && k > cap) st = kAcaOverflow;
  if (threadIdx.x == 0) {
    kt_out[id] = k;
    status[id] = st;
  }
}

// ---- arena fill from device factors ----------------------------------------

__device__ inline uint16_t to_bf16(double v) {
  const float f = (float)v;
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

__global__ __launch_bounds__(kT) void k_fill(const HmvFillOp *__restrict__ ops, char *arena) {
  const HmvFillOp o = ops[blockIdx.x];
  const int64_t n = (int64_t)o.nrows * o.ncols;
  char *dst = arena + o.dst;
  for (int64_t e = threadIdx.x; e < n; e += kT) {
    const int r = (int)(e % o.nrows), c = (int)(e / o.nrows);
    const double v = o.src[r * o.s_r + c * o.s_c];
    const int64_t i = (o.r0 + r) + (int64_t)o.ld * (o.c0 + c);
    if (o.dt == HMV_DT_F64)
      reinterpret_cast<double *>(dst)[i] = v;
    else if (o.dt == HMV_DT_F32)
      reinterpret_cast<float *>(dst)[i] = (float)v;
    else
      reinterpret_cast<uint16_t *>(dst)[i] = to_bf16(v);
  }
}

template <typename T>
hipError_t upload(T **dst, const T *src, size_t n) {
  *dst = nullptr;
  if (n == 0) return hipSuccess;
  hipError_t e = hipMalloc((void **)dst, sizeof(T) * n);
  if (e != hipSuccess) return e;
  return hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice);
}

}  // namespace

namespace hmv {

// Runs the arena fill records on the device (the image already zeroed).
hipError_t launch_fill(const FillOps &ops, char *arena) {
  constexpr size_t kBatch = 1 << 20;  // records per launch (grid size)
  HmvFillOp *d = nullptr;
  const size_t nb = std::min(ops.size(), kBatch);
  hipError_t e = hipMalloc((void **)&d, sizeof(HmvFillOp) * std::max<size_t>(1, nb));
  if (e != hipSuccess) return e;
  for (size_t o = 0; o < ops.size() && e == hipSuccess; o += kBatch) {
    const size_t n = std::min(kBatch, ops.size() - o);
    e = hipMemcpy(d, ops.data() + o, sizeof(HmvFillOp) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_fill, dim3((unsigned)n), dim3(kT), 0, nullptr, d, arena);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
  }
  (void)hipFree(d);
  return e;
}

}  // namespace hmv

// ---- public object ---------------------------------------------------------

struct hmv_dleaves_s {
  int nd = 0, nlf = 0;
  std::vector<CLeaf> lv;           // host copy of the device leaf table
  std::vector<int> kt, status;     // per leaf
  std::vector<hmv::RefLeaf> rec;   // leaf records with device factor pointers
  std::vector<void *> allocs;      // device allocations to free
  double seconds = 0.0;            // device fill time
  ~hmv_dleaves_s() {
    for (void *p : allocs) (void)hipFree(p);
  }
};

namespace {

hipError_t dalloc(hmv_dleaves_s *d, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 8));
  if (e == hipSuccess) d->allocs.push_back(*p);
  return e;
}

// Runs ACA+ on the listed low-rank leaves with the capacities in d->lv.
hipError_t run_aca(hmv_dleaves_s *d, const std::vector<int> &ids, const Geo &g,
                   const AcaParams &p, int *d_kt, int *d_st, CLeaf *d_lv) {
  if (ids.empty()) return hipSuccess;
  int *d_ids = nullptr;
  hipError_t e = upload(&d_ids, ids.data(), ids.size());
  if (e == hipSuccess &&
      (e = hipMemcpy(d_lv, d->lv.data(), sizeof(CLeaf) * d->lv.size(), hipMemcpyHostToDevice)) ==
          hipSuccess) {
    hipLaunchKernelGGL(k_aca, dim3((unsigned)ids.size()), dim3(kA), 0, nullptr, d_lv, d_ids, g, p,
                       d_kt, d_st);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
  }
  (void)hipFree(d_ids);
  return e;
}

}  // namespace

extern "C" {

int hmv_construct(int nnode, const double *nodes, int nface, const int *faces, const double *cent,
                  const double *area, const int *lod, int nlf, const void *specs, int stride,
                  const hmv_construct_params *cp, hmv_dleaves *out) {
  if (!out || !nodes || !faces || !cent || !area || !lod || !specs || !cp || nnode <= 0 ||
      nface <= 0 || nlf < 0 || stride < (int)sizeof(hmv::RefLeaf) ||
      (cp->kernel != kKernelPpohbem && cp->kernel != kKernelCentroid) || cp->kmax < 1 ||
      cp->kmin < 0) {
    hmv::set_error("hmv_construct: invalid argument");
    return HMV_ERR_ARG;
  }
  *out = nullptr;
  for (int64_t i = 0; i < 3 * (int64_t)nface; ++i)
    if (faces[i] < 0 || faces[i] >= nnode) {
      hmv::set_error("hmv_construct: face references a missing node");
      return HMV_ERR_ARG;
    }
  for (int i = 0; i < nface; ++i)
    if (lod[i] < 1 || lod[i] > nface) {
      hmv::set_error("hmv_construct: lod out of range");
      return HMV_ERR_ARG;
    }
  std::unique_ptr<hmv_dleaves_s> d(new hmv_dleaves_s());
  d->nd = nface;
  d->nlf = nlf;
  d->lv.resize(nlf);
  d->kt.assign(nlf, 0);
  d->status.assign(nlf, 0);
  std::vector<int> lr_ids, dense_ids;
  size_t pool = 0, scr = 0, msk = 0;
  for (int i = 0; i < nlf; ++i) {
    const auto *r = reinterpret_cast<const hmv::RefLeaf *>(static_cast<const char *>(specs) +
                                                          (size_t)stride * i);
    CLeaf &L = d->lv[i];
    L.ltmtx = r->ltmtx;
    L.nstrtl = r->nstrtl;
    L.ndl = r->ndl;
    L.nstrtt = r->nstrtt;
    L.ndt = r->ndt;
    if ((L.ltmtx != 1 && L.ltmtx != 2) || L.ndl <= 0 || L.ndt <= 0 || L.nstrtl < 1 ||
        L.nstrtt < 1 || (int64_t)L.nstrtl - 1 + L.ndl > nface ||
        (int64_t)L.nstrtt - 1 + L.ndt > nface) {
      hmv::set_error("hmv_construct: leaf " + std::to_string(i + 1) + " outside the matrix");
      return HMV_ERR_ARG;
    }
    if (L.ltmtx == 1) {
      L.cap = std::min(kCapFirst, std::min(cp->kmax, std::min(L.ndl, L.ndt)) + 1);
      pool += (size_t)(L.ndl + L.ndt) * L.cap;
      scr += (size_t)(L.ndl + L.ndt);
      msk += (size_t)(L.ndl + L.ndt);
      lr_ids.push_back(i);
    } else {
      L.cap = 0;
      pool += (size_t)L.ndl * L.ndt;
      dense_ids.push_back(i);
    }
  }
  hipError_t e;
  double *d_pool = nullptr, *d_scr = nullptr, *d_geo = nullptr;
  int *d_msk = nullptr, *d_faces = nullptr, *d_lod = nullptr, *d_kt = nullptr, *d_st = nullptr;
  CLeaf *d_lv = nullptr;
  const size_t geo_n = 3 * (size_t)nface + nface + 3 * (size_t)nnode;
  if ((e = dalloc(d.get(), (void **)&d_pool, sizeof(double) * pool)) != hipSuccess ||
      (e = hipMalloc((void **)&d_scr, sizeof(double) * std::max<size_t>(scr, 1))) != hipSuccess ||
      (e = hipMalloc((void **)&d_msk, sizeof(int) * std::max<size_t>(msk, 1))) != hipSuccess ||
      (e = hipMalloc((void **)&d_geo, sizeof(double) * geo_n)) != hipSuccess ||
      (e = hipMalloc((void **)&d_faces, sizeof(int) * 3 * (size_t)nface)) != hipSuccess ||
      (e = hipMalloc((void **)&d_lod, sizeof(int) * (size_t)nface)) != hipSuccess ||
      (e = hipMalloc((void **)&d_kt, sizeof(int) * std::max(nlf, 1))) != hipSuccess ||
      (e = hipMalloc((void **)&d_st, sizeof(int) * std::max(nlf, 1))) != hipSuccess ||
      (e = hipMalloc((void **)&d_lv, sizeof(CLeaf) * std::max(nlf, 1))) != hipSuccess) {
    (void)hipFree(d_scr);
    (void)hipFree(d_msk);
    (void)hipFree(d_geo);
    (void)hipFree(d_faces);
    (void)hipFree(d_lod);
    (void)hipFree(d_kt);
    (void)hipFree(d_st);
    (void)hipFree(d_lv);
    return hmv::hip_fail(e, "hmv_construct: device memory");
  }
  struct Tmp {  // freed on every exit
    std::vector<void *> p;
    ~Tmp() {
      for (void *q : p) (void)hipFree(q);
    }
  } tmp;
  tmp.p = {d_scr, d_msk, d_geo, d_faces, d_lod, d_kt, d_st, d_lv};
  if ((e = hipMemset(d_kt, 0, sizeof(int) * std::max(nlf, 1))) != hipSuccess ||
      (e = hipMemset(d_st, 0, sizeof(int) * std::max(nlf, 1))) != hipSuccess)
    return hmv::hip_fail(e, "hmv_construct: device memory");
  auto t0 = std::chrono::steady_clock::now();
  {
    size_t po = 0, so = 0, mo = 0;
    for (CLeaf &L : d->lv) {
      L.base = d_pool + po;
      if (L.ltmtx == 1) {
        po += (size_t)(L.ndl + L.ndt) * L.cap;
        L.pa = d_scr + so;
        L.msk = d_msk + mo;
        so += (size_t)(L.ndl + L.ndt);
        mo += (size_t)(L.ndl + L.ndt);
      } else {
        po += (size_t)L.ndl * L.ndt;
        L.pa = nullptr;
        L.msk = nullptr;
      }
    }
  }
  if ((e = hipMemcpy(d_geo, cent, sizeof(double) * 3 * (size_t)nface, hipMemcpyHostToDevice)) !=
          hipSuccess ||
      (e = hipMemcpy(d_geo + 3 * (size_t)nface, area, sizeof(double) * nface,
                     hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d_geo + 4 * (size_t)nface, nodes, sizeof(double) * 3 * (size_t)nnode,
                     hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d_faces, faces, sizeof(int) * 3 * (size_t)nface, hipMemcpyHostToDevice)) !=
          hipSuccess ||
      (e = hipMemcpy(d_lod, lod, sizeof(int) * (size_t)nface, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d_lv, d->lv.data(), sizeof(CLeaf) * d->lv.size(), hipMemcpyHostToDevice)) !=
          hipSuccess)
    return hmv::hip_fail(e, "hmv_construct: upload");
  Geo g{d_geo, d_geo + 3 * (size_t)nface, d_geo + 4 * (size_t)nface, d_faces, d_lod, cp->kernel};
  AcaParams ap{cp->aca_eps, cp->aca_eps_factor, cp->kmax, cp->kmin};
  // dense leaves
  if (!dense_ids.empty()) {
    int *d_ids = nullptr;
    if ((e = upload(&d_ids, dense_ids.data(), dense_ids.size())) == hipSuccess) {
      hipLaunchKernelGGL(k_dense, dim3((unsigned)dense_ids.size()), dim3(kT), 0, nullptr, d_lv,
                         d_ids, g);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    (void)hipFree(d_ids);
    if (e != hipSuccess) return hmv::hip_fail(e, "hmv_construct: dense fill");
  }
  // low-rank leaves, largest first; then the ones that needed more columns
  std::sort(lr_ids.begin(), lr_ids.end(), [&](int a, int b) {
    const int64_t sa = (int64_t)d->lv[a].ndl + d->lv[a].ndt, sb = (int64_t)d->lv[b].ndl + d->lv[b].ndt;
    return sa != sb ? sa > sb : a < b;
  });
  if ((e = run_aca(d.get(), lr_ids, g, ap, d_kt, d_st, d_lv)) != hipSuccess ||
      (e = hipMemcpy(d->kt.data(), d_kt, sizeof(int) * nlf, hipMemcpyDeviceToHost)) != hipSuccess ||
      (e = hipMemcpy(d->status.data(), d_st, sizeof(int) * nlf, hipMemcpyDeviceToHost)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_construct: ACA+");
  std::vector<int> again;
  size_t pool2 = 0;
  for (int i : lr_ids)
    if (d->status[i] == kAcaOverflow) {
      CLeaf &L = d->lv[i];
      L.cap = std::min(cp->kmax, std::min(L.ndl, L.ndt)) + 1;
      pool2 += (size_t)(L.ndl + L.ndt) * L.cap;
      again.push_back(i);
    }
  if (!again.empty()) {
    double *d_pool2 = nullptr;
    if ((e = dalloc(d.get(), (void **)&d_pool2, sizeof(double) * pool2)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_construct: device memory");
    size_t po = 0;
    for (int i : again) {
      d->lv[i].base = d_pool2 + po;
      po += (size_t)(d->lv[i].ndl + d->lv[i].ndt) * d->lv[i].cap;
    }
    if ((e = run_aca(d.get(), again, g, ap, d_kt, d_st, d_lv)) != hipSuccess ||
        (e = hipMemcpy(d->kt.data(), d_kt, sizeof(int) * nlf, hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(d->status.data(), d_st, sizeof(int) * nlf, hipMemcpyDeviceToHost)) !=
            hipSuccess)
      return hmv::hip_fail(e, "hmv_construct: ACA+");
    for (int i : again)
      if (d->status[i] != kAcaOk) {
        hmv::set_error("hmv_construct: ACA+ of leaf " + std::to_string(i + 1) +
                       " exceeded its rank bound");
        return HMV_ERR_STATE;
      }
  }
  auto t1 = std::chrono::steady_clock::now();
  d->seconds = std::chrono::duration<double>(t1 - t0).count();
  // leaf records with device pointers (a2 = a1 + a1size, as the reference)
  d->rec.resize(nlf);
  for (int i = 0; i < nlf; ++i) {
    const CLeaf &L = d->lv[i];
    hmv::RefLeaf &r = d->rec[i];
    std::memset(&r, 0, sizeof(r));
    r.ltmtx = L.ltmtx;
    r.nstrtl = L.nstrtl;
    r.ndl = L.ndl;
    r.nstrtt = L.nstrtt;
    r.ndt = L.ndt;
    r.a1 = L.base;
    if (L.ltmtx == 1) {
      r.kt = d->kt[i];
      r.a1size = (long long)sizeof(double) * L.ndt * L.cap;
      r.a2 = L.base + (size_t)L.ndt * L.cap;
    } else {
      r.kt = 0;
      r.a1size = (long long)sizeof(double) * L.ndl * L.ndt;
      r.a2 = nullptr;
    }
  }
  *out = d.release();
  return HMV_OK;
}

int hmv_dleaves_info(hmv_dleaves d, int *nlf, int *kt, double *seconds) {
  if (!d) {
    hmv::set_error("hmv_dleaves_info: null handle");
    return HMV_ERR_ARG;
  }
  if (nlf) *nlf = d->nlf;
  if (kt) std::copy(d->kt.begin(), d->kt.end(), kt);
  if (seconds) *seconds = d->seconds;
  return HMV_OK;
}

int hmv_dleaves_download(hmv_dleaves d, int leaf, double *a1) {
  if (!d || leaf < 0 || leaf >= d->nlf || !a1) {
    hmv::set_error("hmv_dleaves_download: invalid argument");
    return HMV_ERR_ARG;
  }
  const CLeaf &L = d->lv[leaf];
  hipError_t e;
  if (L.ltmtx == 2) {
    e = hipMemcpy(a1, L.base, sizeof(double) * L.ndl * L.ndt, hipMemcpyDeviceToHost);
  } else {
    const int k = d->kt[leaf];
    e = hipMemcpy(a1, L.base, sizeof(double) * L.ndt * k, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      e = hipMemcpy(a1 + (size_t)L.ndt * k, L.base + (size_t)L.ndt * L.cap,
                    sizeof(double) * L.ndl * k, hipMemcpyDeviceToHost);
  }
  if (e != hipSuccess) return hmv::hip_fail(e, "hmv_dleaves_download");
  return HMV_OK;
}

int hmv_arena_build_dleaves(hmv_dleaves d, int precision, const hmv_tuning *tuning,
                            hmv_arena *out) {
  if (!d || !out) {
    hmv::set_error("hmv_arena_build_dleaves: null argument");
    return HMV_ERR_ARG;
  }
  return hmv::build_from_device_leaves(d->nd, d->nlf, d->rec.data(), (int)sizeof(hmv::RefLeaf),
                                       precision, tuning, out);
}

int hmv_dleaves_free(hmv_dleaves d) {
  delete d;
  return HMV_OK;
}

}  // extern "C"
