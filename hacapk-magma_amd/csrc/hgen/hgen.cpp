// hgen.cpp -- host-side H-matrix construction (input generator).
//
// Restates the reference's construction so that leaf lists can be produced
// without its Fortran/MPI build:
//   * bisection cluster tree        m_HACApK_base.f90:1385-1445 (HACApK_generate_cbitree)
//   * bounding boxes                m_HACApK_base.f90:1344-1382 (HACApK_bndbox)
//   * block tree / admissibility    m_HACApK_base.f90:1172-1206, 1236-1282
//   * leaf ordering                 m_HACApK_base.f90:451-461 (row sort, then column sort)
//   * ACA+                          m_HACApK_base.f90:783-975 (HACApK_acaplus)
//   * calc_vec / adotsub_dsm        m_HACApK_base.f90:677-698, HACApK_lib.f90:107-129
//   * scaled norm / pivots          HACApK_lib.f90:31-81
//   * fill                          m_HACApK_base.f90:978-1051 (HACApK_fill_leafmtx_hyp)
//   * statistics                    m_HACApK_base.f90:1108-1169 (HACApK_chk_leafmtx)
//   * entry function                m_ppohBEM_user_func.f90:49-100 (face_integral),
//                                   m_ppohBEM_matrix_element_ij.f90:49-76
// The fill runs leaf-parallel under OpenMP; each leaf is computed by one
// thread with a fixed operation order, so results do not depend on the
// thread count.
#include "hmvgen.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// Leaf record in the reference's C layout (HACApK_MAGMA.h:12-20).
struct LeafRec {
  int ltmtx;
  int kt;
  int nstrtl, ndl;
  int nstrtt, ndt;
  long long a1size;
  double *a1;
  double *a2;
};
static_assert(sizeof(LeafRec) == 48, "leaf record must match stc_HACApK_leafmtx");

// ---------------------------------------------------------------------------
// Entry function.
// ---------------------------------------------------------------------------

// m_ppohBEM_user_func.f90:42-43: both constants are default-REAL literals
// assigned to REAL(8) parameters, so they carry single-precision values.
const double kPiF = (double)3.141592653589793f;
const double kEps0F = (double)8.854187818f * 1e-12;

inline void cross3(const double u[3], const double v[3], double w[3]) {
  w[0] = u[1] * v[2] - u[2] * v[1];
  w[1] = u[2] * v[0] - u[0] * v[2];
  w[2] = u[0] * v[1] - u[1] * v[0];
}

inline double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Signed sum of the analytic 1/r potential of a flat triangle at (x,y,z);
// follows face_integral (m_ppohBEM_user_func.f90:49-100) term by term.
// The returned value's magnitude is  int_T 1/|p-y| dS_y .
double triangle_potential_sum(const double xs[3], const double ys[3], const double zs[3],
                              double x, double y, double z) {
  double r[3];
  for (int i = 0; i < 3; ++i) {
    double dx = xs[i] - x, dy = ys[i] - y, dz = zs[i] - z;
    r[i] = std::sqrt(dx * dx + dy * dy + dz * dz);
  }
  double u[3] = {xs[1] - xs[0], ys[1] - ys[0], zs[1] - zs[0]};
  double v[3] = {xs[2] - xs[1], ys[2] - ys[1], zs[2] - zs[1]};
  double w[3];
  cross3(u, v, w);
  double wn = std::sqrt(dot3(w, w));
  w[0] /= wn;
  w[1] /= wn;
  w[2] /= wn;
  u[0] = x - xs[0];
  u[1] = y - ys[0];
  u[2] = z - zs[0];
  double zp = dot3(u, w);
  double ox = x - zp * w[0], oy = y - zp * w[1], oz = z - zp * w[2];
  double zpabs = std::fabs(zp);
  double sum = 0.0;
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3;
    u[0] = xs[j] - ox;
    u[1] = ys[j] - oy;
    u[2] = zs[j] - oz;
    double xj = std::sqrt(dot3(u, u));
    u[0] /= xj;
    u[1] /= xj;
    u[2] /= xj;
    cross3(w, u, v);
    double xi = (xs[i] - ox) * u[0] + (ys[i] - oy) * u[1] + (zs[i] - oz) * u[2];
    double yi = (xs[i] - ox) * v[0] + (ys[i] - oy) * v[1] + (zs[i] - oz) * v[2];
    double dx = xj - xi, dy = -yi;
    double t = std::sqrt(dx * dx + dy * dy);
    double l = dx / t, m = dy / t;
    double d = l * yi - m * xi;
    double ti = l * xi + m * yi, tj = l * xj;
    double theta = std::atan2(yi, xi);
    double omega = theta - std::atan2(r[i] * d, zpabs * ti) + std::atan2(r[j] * d, zpabs * tj);
    double q = std::log((r[j] + tj) / (r[i] + ti));
    double g = d * q - zpabs * omega;
    sum = sum + g;
  }
  return sum;
}

double face_integral_mode(int kernel, const double xs[3], const double ys[3], const double zs[3],
                          double x, double y, double z) {
  double s = triangle_potential_sum(xs, ys, zs, x, y, z);
  if (kernel == HMVGEN_KERNEL_PPOHBEM) return std::fabs(s) / (4.0 * kPiF * kEps0F);
  return std::fabs(s) / (4.0 * M_PI);
}

struct Bem {
  int kernel = HMVGEN_KERNEL_PPOHBEM;
  int nface = 0;
  std::vector<double> nodes;   // nnode x 3
  std::vector<int> faces;      // nface x 3, 0-based
  std::vector<double> cent;    // nface x 3
  std::vector<double> area;    // nface

  void tri(int f, double xs[3], double ys[3], double zs[3]) const {
    for (int a = 0; a < 3; ++a) {
      int n = faces[3 * f + a];
      xs[a] = nodes[3 * n + 0];
      ys[a] = nodes[3 * n + 1];
      zs[a] = nodes[3 * n + 2];
    }
  }

  // HACApK_entry_ij(i,j): collocation at the centroid of face i, source face j
  // (1-based original indices; m_ppohBEM_matrix_element_ij.f90:49-76).
  double entry(int i, int j) const {
    const int fi = i - 1, fj = j - 1;
    if (kernel == HMVGEN_KERNEL_CENTROID && fi != fj) {
      double dx = cent[3 * fi] - cent[3 * fj];
      double dy = cent[3 * fi + 1] - cent[3 * fj + 1];
      double dz = cent[3 * fi + 2] - cent[3 * fj + 2];
      return area[fj] / (4.0 * M_PI * std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    double xs[3], ys[3], zs[3];
    tri(fj, xs, ys, zs);
    return face_integral_mode(kernel, xs, ys, zs, cent[3 * fi], cent[3 * fi + 1], cent[3 * fi + 2]);
  }
};

// ---------------------------------------------------------------------------
// Small numeric helpers (HACApK_lib.f90).
// ---------------------------------------------------------------------------

// HACApK_unrm_d: overflow-safe 2-norm, entries below 1e-30 skipped.
double unrm(int nd, const double *za) {
  double zscl = 0.0, zzz = 1.0;
  for (int il = 0; il < nd; ++il) {
    double zza = std::fabs(za[il]);
    if (zza < 1.0e-30) continue;
    if (zscl < zza) {
      zzz = 1.0 + zzz * (zscl / zza) * (zscl / zza);
      zscl = zza;
    } else {
      zzz = zzz + (zza / zscl) * (zza / zscl);
    }
  }
  return zscl * std::sqrt(zzz);
}

// HACApK_maxabsvalloc_d: first index of max |za| (1-based).
void maxabsvalloc(const double *za, double &zz, int &il, int nd) {
  il = 1;
  zz = std::fabs(za[0]);
  for (int it = 2; it <= nd; ++it) {
    double a = std::fabs(za[it - 1]);
    if (a > zz) {
      il = it;
      zz = a;
    }
  }
}

// HACApK_minabsvalloc_d: first index of min |za| (1-based).
void minabsvalloc(const double *za, double &zz, int &il, int nd) {
  il = 1;
  zz = std::fabs(za[0]);
  for (int it = 2; it <= nd; ++it) {
    double a = std::fabs(za[it - 1]);
    if (a < zz) {
      il = it;
      zz = a;
    }
  }
}

// ---------------------------------------------------------------------------
// Cluster tree.
// ---------------------------------------------------------------------------

struct Cluster {
  int nstrt = 0, nsize = 0, ndpth = 0, nnson = 0;
  int son[2] = {-1, -1};
  double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
  double zwdth = 0.0;
};

struct Tree {
  std::vector<Cluster> cl;
  const double *g = nullptr;  // nofc x 3 (original order)
  int minsz = 15;

  double gv(int idx1, int d) const { return g[3 * (idx1 - 1) + d]; }

  // HACApK_generate_cbitree; lod is the 1-based index array of this node's
  // points (lod[0] is the node's first point).
  int build(int *lod, int nsrt, int nd, int ndpth) {
    int me = (int)cl.size();
    cl.emplace_back();
    cl[me].nstrt = nsrt;
    cl[me].nsize = nd;
    cl[me].ndpth = ndpth;
    if (nd <= minsz) {
      cl[me].nnson = 0;
      return me;
    }
    double zlmin[3], zlmax[3];
    for (int id = 0; id < 3; ++id) {
      zlmin[id] = gv(lod[0], id);
      zlmax[id] = zlmin[id];
      for (int il = 2; il <= nd; ++il) {
        double zg = gv(lod[il - 1], id);
        if (zg < zlmin[id])
          zlmin[id] = zg;
        else if (zlmax[id] < zg)
          zlmax[id] = zg;
      }
    }
    double zdiff = zlmax[0] - zlmin[0];
    int ncut = 0;
    for (int id = 0; id < 3; ++id) {
      double zidiff = zlmax[id] - zlmin[id];
      if (zidiff > zdiff) {
        zdiff = zidiff;
        ncut = id;
      }
    }
    double zlmid = (zlmax[ncut] + zlmin[ncut]) / 2;
    int nl = 1, nr = nd;
    while (nl < nr) {
      while (nl < nd && gv(lod[nl - 1], ncut) <= zlmid) ++nl;
      while (nr >= 1 && gv(lod[nr - 1], ncut) > zlmid) --nr;
      if (nl < nr) std::swap(lod[nl - 1], lod[nr - 1]);
    }
    cl[me].nnson = 2;
    int s0 = build(lod, nsrt, nl - 1, ndpth + 1);
    int s1 = build(lod + (nl - 1), nsrt + nl - 1, nd - nl + 1, ndpth + 1);
    cl[me].son[0] = s0;
    cl[me].son[1] = s1;
    return me;
  }

  // HACApK_bndbox (children first, then this node).
  void bndbox(int c, const int *lod) {
    Cluster &n = cl[c];
    if (n.nnson > 0) {
      bndbox(n.son[0], lod);
      bndbox(n.son[1], lod + cl[n.son[0]].nsize);
      for (int id = 0; id < 3; ++id) {
        n.bmin[id] = cl[n.son[0]].bmin[id];
        n.bmax[id] = cl[n.son[0]].bmax[id];
      }
      for (int s = 1; s < n.nnson; ++s)
        for (int id = 0; id < 3; ++id) {
          if (cl[n.son[s]].bmin[id] < n.bmin[id]) n.bmin[id] = cl[n.son[s]].bmin[id];
          if (n.bmax[id] < cl[n.son[s]].bmax[id]) n.bmax[id] = cl[n.son[s]].bmax[id];
        }
    } else {
      for (int id = 0; id < 3; ++id) {
        n.bmin[id] = gv(lod[0], id);
        n.bmax[id] = gv(lod[0], id);
      }
      for (int id = 0; id < 3; ++id)
        for (int il = 2; il <= n.nsize; ++il) {
          double v = gv(lod[il - 1], id);
          if (v < n.bmin[id]) n.bmin[id] = v;
          if (n.bmax[id] < v) n.bmax[id] = v;
        }
    }
    double zw = (n.bmax[0] - n.bmin[0]) * (n.bmax[0] - n.bmin[0]);
    for (int id = 1; id < 3; ++id) zw = zw + (n.bmax[id] - n.bmin[id]) * (n.bmax[id] - n.bmin[id]);
    n.zwdth = std::sqrt(zw);
  }
};

struct LeafSpec {
  int ltmtx, nstrtl, ndl, nstrtt, ndt;
};

struct BlockGen {
  const Tree *tr;
  int nleaf, nlmax;
  double zeta;
  std::vector<LeafSpec> leaves;
  int nlr = 0, ndense = 0, nh = 0;

  // HACApK_generate_leafmtx / HACApK_count_lntmx.
  void gen(int l, int t) {
    const Cluster &cl_ = tr->cl[l], &ct = tr->cl[t];
    int ndl = cl_.nsize, ndt = ct.nsize;
    double zs = 0.0;
    for (int id = 0; id < 3; ++id) {
      if (cl_.bmax[id] < ct.bmin[id])
        zs = zs + (ct.bmin[id] - cl_.bmax[id]) * (ct.bmin[id] - cl_.bmax[id]);
      else if (ct.bmax[id] < cl_.bmin[id])
        zs = zs + (cl_.bmin[id] - ct.bmax[id]) * (cl_.bmin[id] - ct.bmax[id]);
    }
    double zdistlt = std::sqrt(zs);
    if ((cl_.zwdth <= zeta * zdistlt || ct.zwdth <= zeta * zdistlt) &&
        (ndl >= nleaf && ndt >= nleaf && ndl <= nlmax && ndt <= nlmax)) {
      leaves.push_back({1, cl_.nstrt, ndl, ct.nstrt, ndt});
      ++nlr;
    } else if (cl_.nnson == 0 || ct.nnson == 0 || ndl <= nleaf || ndt <= nleaf) {
      leaves.push_back({2, cl_.nstrt, ndl, ct.nstrt, ndt});
      ++ndense;
    } else {
      ++nh;
      for (int il = 0; il < cl_.nnson; ++il)
        for (int it = 0; it < ct.nnson; ++it) gen(cl_.son[il], ct.son[it]);
    }
  }
};

// ---------------------------------------------------------------------------
// ACA+ (HACApK_acaplus, m_HACApK_base.f90:783-975).
// ---------------------------------------------------------------------------

struct Aca {
  const Bem *bem;
  const int *lod;  // 1-based original indices, lod[i-1]
  int ndl, ndt, nstrtl, nstrtt;
  int cap = 0;
  std::vector<double> U;  // zaa(ndl, cap)
  std::vector<double> V;  // zab(ndt, cap)
  std::vector<double> tmp;

  double *ucol(int k1) { return &U[(size_t)(k1 - 1) * ndl]; }  // zaa(:,k1)
  double *vcol(int k1) { return &V[(size_t)(k1 - 1) * ndt]; }  // zab(:,k1)

  void ensure(int k1) {
    if (k1 <= cap) return;
    int nc = std::max(k1, cap + 32);
    U.resize((size_t)nc * ndl, 0.0);
    V.resize((size_t)nc * ndt, 0.0);
    cap = nc;
  }

  double ent(int il1, int it1) const {  // block-local 1-based row/col
    return bem->entry(lod[il1 + nstrtl - 2], lod[it1 + nstrtt - 2]);
  }

  // HACApK_calc_vec(..., ld=1): column ip of the residual block.
  void calc_col(int k, int ip, double *vec, const int *lmsk) {
    for (int ii = 1; ii <= ndl; ++ii)
      if (lmsk[ii - 1] == 0) vec[ii - 1] = ent(ii, ip);
    if (k == 0) return;
    subtract(vec, U.data(), ndl, V.data(), ndt, ip, k);
    for (int ii = 0; ii < ndl; ++ii)
      if (lmsk[ii] == 1) vec[ii] = 0.0;
  }

  // HACApK_calc_vec(..., ld=0): row ip of the residual block.
  void calc_row(int k, int ip, double *vec, const int *lmsk) {
    for (int ii = 1; ii <= ndt; ++ii)
      if (lmsk[ii - 1] == 0) vec[ii - 1] = ent(ip, ii);
    if (k == 0) return;
    subtract(vec, V.data(), ndt, U.data(), ndl, ip, k);
    for (int ii = 0; ii < ndt; ++ii)
      if (lmsk[ii] == 1) vec[ii] = 0.0;
  }

  // HACApK_adotsub_dsm(vec, zaa, zab(ip,1:k)): vec -= zaa(:,1:k) * zab(ip,1:k),
  // the product accumulated column by column before the subtraction.
  void subtract(double *vec, const double *zaa, int lda, const double *zab, int ldb, int ip,
                int k) {
    tmp.assign(lda, 0.0);
    for (int it = 0; it < k; ++it) {
      double zu = zab[(size_t)it * ldb + (ip - 1)];
      const double *col = zaa + (size_t)it * lda;
      for (int ii = 0; ii < lda; ++ii) tmp[ii] = tmp[ii] + col[ii] * zu;
    }
    for (int ii = 0; ii < lda; ++ii) vec[ii] = vec[ii] - tmp[ii];
  }

  int run(const hmvgen_params &p) {
    const double eps = p.aca_eps;
    const double pACA_EPS = p.aca_eps_factor * eps;
    const int kmax = p.kmax;
    const int kmin = p.kmin;
    const double za_ACA_EPS = (double)1.0e-30f;
    const double piv_min = (double)1.0e-20f;
    const double znrm = 0.0;  // znrmmat = 0 for param(61) = 1
    double ACA_EPS = pACA_EPS;

    int ntries = std::max(ndl, ndt) + 1, ntries_row = 6, ntries_col = 6;
    std::vector<int> lrow_msk(ndl, 0), lcol_msk(ndt, 0);
    int k = 0;
    ensure(1);

    int j_ref = 1;
    std::vector<double> pa_ref(ndl, 0.0), pb_ref(ndt, 0.0);
    calc_col(k, j_ref, pa_ref.data(), lrow_msk.data());
    double colnorm = unrm(ndl, pa_ref.data());
    double rownorm;
    int i_ref;
    minabsvalloc(pa_ref.data(), rownorm, i_ref, ndl);
    calc_row(k, i_ref, pb_ref.data(), lcol_msk.data());
    rownorm = unrm(ndt, pb_ref.data());

    double apxnorm = 0.0;
    int lstop_aca = 0;
    int i = 0, j = 0;
    double col_maxval, row_maxval;
    while ((k < kmax) && (ntries_row > 0 || ntries_col > 0) && (ntries > 0)) {
      ntries = ntries - 1;
      ensure(k + 1);
      double *pcol = ucol(k + 1);
      double *prow = vcol(k + 1);
      col_maxval = 0.0;
      maxabsvalloc(pa_ref.data(), col_maxval, i, ndl);
      row_maxval = 0.0;
      maxabsvalloc(pb_ref.data(), row_maxval, j, ndt);

      if (row_maxval > col_maxval) {
        if (j != j_ref)
          calc_col(k, j, pcol, lrow_msk.data());
        else
          std::copy(pa_ref.begin(), pa_ref.end(), pcol);
        maxabsvalloc(pcol, col_maxval, i, ndl);
        if (col_maxval < ACA_EPS && k >= kmin) {
          lstop_aca = 1;
        } else {
          calc_row(k, i, prow, lcol_msk.data());
          double zinvmax;
          if (std::fabs(pcol[i - 1]) > piv_min) {
            zinvmax = 1.0 / pcol[i - 1];
          } else {
            k = std::max(k - 1, 0);
            break;
          }
          for (int a = 0; a < ndl; ++a) pcol[a] = pcol[a] * zinvmax;
        }
      } else {
        if (i != i_ref)
          calc_row(k, i, prow, lcol_msk.data());
        else
          std::copy(pb_ref.begin(), pb_ref.end(), prow);
        maxabsvalloc(prow, row_maxval, j, ndt);
        if (row_maxval < ACA_EPS && k >= kmin) {
          lstop_aca = 1;
        } else {
          calc_col(k, j, pcol, lrow_msk.data());
          double zinvmax;
          if (std::fabs(prow[j - 1]) > piv_min) {
            zinvmax = 1.0 / prow[j - 1];
          } else {
            k = std::max(k - 1, 0);
            break;
          }
          for (int a = 0; a < ndt; ++a) prow[a] = prow[a] * zinvmax;
        }
      }
      lrow_msk[i - 1] = 1;
      lcol_msk[j - 1] = 1;

      if (i != i_ref) {
        double zinv = -pcol[i_ref - 1];
        for (int a = 0; a < ndt; ++a) pb_ref[a] = pb_ref[a] + prow[a] * zinv;
        rownorm = unrm(ndt, pb_ref.data());
      }
      if (i == i_ref || rownorm < ACA_EPS) {
        if (i == i_ref) ntries_row = ntries_row + 1;
        if (ntries_row > 0) {
          rownorm = 0.0;
          i = i_ref;
          while (i != ((i_ref + ndl - 2) % ndl) + 1 && rownorm < za_ACA_EPS && ntries_row > 0) {
            if (lrow_msk[i - 1] == 0) {
              ensure(k + 1);
              calc_row(k + 1, i, pb_ref.data(), lcol_msk.data());
              rownorm = unrm(ndt, pb_ref.data());
              if (rownorm < ACA_EPS) lrow_msk[i - 1] = 1;
              ntries_row = ntries_row - 1;
            } else {
              rownorm = 0.0;
            }
            i = (i % ndl) + 1;
          }
          i_ref = ((i + ndl - 2) % ndl) + 1;
        }
      }

      if (j != j_ref) {
        double zinv = -prow[j_ref - 1];
        for (int a = 0; a < ndl; ++a) pa_ref[a] = pa_ref[a] + pcol[a] * zinv;
        colnorm = unrm(ndl, pa_ref.data());
      }
      if (j == j_ref || colnorm < ACA_EPS) {
        if (j == j_ref) ntries_col = ntries_col + 1;
        if (ntries_col > 0) {
          colnorm = 0.0;
          j = j_ref;
          while (j != ((j_ref + ndt - 2) % ndt) + 1 && colnorm < za_ACA_EPS && ntries_col > 0) {
            if (lcol_msk[j - 1] == 0) {
              ensure(k + 1);
              calc_col(k + 1, j, pa_ref.data(), lrow_msk.data());
              colnorm = unrm(ndl, pa_ref.data());
              if (colnorm < ACA_EPS) lcol_msk[j - 1] = 1;
              ntries_col = ntries_col - 1;
            } else {
              colnorm = 0.0;
            }
            j = (j % ndt) + 1;
          }
          j_ref = ((j + ndt - 2) % ndt) + 1;
        }
      }

      if (colnorm < ACA_EPS && rownorm < ACA_EPS && k >= kmin) {
        lstop_aca = 1;
        k = k + 1;
      }
      if (lstop_aca == 0) {
        double blknorm = unrm(ndl, pcol) * unrm(ndt, prow);
        if (k == 0) {
          if (p.aca_norm == 1)
            apxnorm = blknorm;
          else
            apxnorm = znrm;
        } else {
          if (blknorm < apxnorm * eps && rownorm < apxnorm * eps && colnorm < apxnorm * eps &&
              k >= kmin)
            lstop_aca = 1;
        }
      }
      if (lstop_aca == 1 && k >= kmin) break;
      k = k + 1;
    }
    return k;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// Public object.
// ---------------------------------------------------------------------------

struct hmvgen_hmatrix {
  Bem bem;
  hmvgen_params params;
  hmvgen_stats stats;
  std::vector<int> lod;            // 1-based
  std::vector<LeafRec> leaves;
  std::vector<std::unique_ptr<double[]>> store;
};

namespace {
int build_impl(int nnode, const double *nodes, int nface, const int *faces,
               const hmvgen_params *pin, hmvgen_hmatrix **out, bool fill);
}

extern "C" {

const char *hmvgen_last_error(void) { return g_err.c_str(); }

void hmvgen_default_params(hmvgen_params *p) {
  std::memset(p, 0, sizeof(*p));
  p->leaf_size = 15;
  p->max_leaf_frac = 1.0;
  p->eta = 2.0;
  p->aca_kind = 2;
  p->aca_norm = 1;
  p->kt_predict = 7;
  p->kmax = 1000;
  p->kmin = 1;
  p->aca_eps = 1.0e-6;
  p->aca_eps_factor = (double)1.0e-3f;  // st_ctl%param(72)=1.0e-3 is a REAL*4 literal
  p->kernel = HMVGEN_KERNEL_PPOHBEM;
  p->nthreads = 0;
}

int hmvgen_icosphere(int subdiv, double r, double cx, double cy, double cz, int *nnode,
                     int *nface, double *nodes, int *faces) {
  if (subdiv < 0 || subdiv > 9) return fail(-1, "icosphere: subdiv out of range");
  const int64_t nf = 20LL << (2 * subdiv);
  const int64_t nn = 10LL * (1LL << (2 * subdiv)) + 2;
  if (nnode) *nnode = (int)nn;
  if (nface) *nface = (int)nf;
  if (!nodes || !faces) return 0;
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<double> v = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                           0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
  std::vector<int> f = {0, 11, 5, 0, 5,  1, 0, 1, 7, 0, 7,  10, 0, 10, 11, 1, 5, 9, 5, 11,
                        4, 11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3,  4,  2, 3, 2, 6, 3,
                        6, 8,  3, 8, 9,  4, 9, 5, 2, 4, 11, 6, 2, 10, 8,  6, 7, 9, 8, 1};
  auto normalize = [&](int i) {
    double x = v[3 * i], y = v[3 * i + 1], z = v[3 * i + 2];
    double n = std::sqrt(x * x + y * y + z * z);
    v[3 * i] = x / n;
    v[3 * i + 1] = y / n;
    v[3 * i + 2] = z / n;
  };
  for (int i = 0; i < 12; ++i) normalize(i);
  for (int s = 0; s < subdiv; ++s) {
    std::map<std::pair<int, int>, int> mid;
    std::vector<int> nf2;
    nf2.reserve(f.size() * 4);
    auto midpoint = [&](int a, int b) {
      std::pair<int, int> key(std::min(a, b), std::max(a, b));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      int id = (int)(v.size() / 3);
      v.push_back(0.5 * (v[3 * a] + v[3 * b]));
      v.push_back(0.5 * (v[3 * a + 1] + v[3 * b + 1]));
      v.push_back(0.5 * (v[3 * a + 2] + v[3 * b + 2]));
      normalize(id);
      mid.emplace(key, id);
      return id;
    };
    for (size_t q = 0; q < f.size(); q += 3) {
      int a = f[q], b = f[q + 1], c = f[q + 2];
      int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      int tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      nf2.insert(nf2.end(), tri, tri + 12);
    }
    f.swap(nf2);
  }
  if ((int64_t)v.size() / 3 != nn || (int64_t)f.size() / 3 != nf)
    return fail(-2, "icosphere: internal size mismatch");
  for (int64_t i = 0; i < nn; ++i) {
    nodes[3 * i] = cx + r * v[3 * i];
    nodes[3 * i + 1] = cy + r * v[3 * i + 1];
    nodes[3 * i + 2] = cz + r * v[3 * i + 2];
  }
  std::memcpy(faces, f.data(), sizeof(int) * f.size());
  return 0;
}

int hmvgen_read_pbf(const char *path, int *nnode, int *nface, double *nodes, int *faces,
                    double *dble_para) {
  FILE *fp = std::fopen(path, "r");
  if (!fp) return fail(-1, std::string("read_pbf: cannot open ") + path);
  int nn = 0, nf = 0, nof = 0, nint = 0, ndble = 0;
  int rc = 0;
  if (std::fscanf(fp, "%d", &nn) != 1 || nn <= 0) rc = -2;
  std::vector<double> xyz;
  if (!rc) {
    xyz.resize(3 * (size_t)nn);
    for (int i = 0; i < 3 * nn && !rc; ++i)
      if (std::fscanf(fp, "%lf", &xyz[i]) != 1) rc = -3;
  }
  if (!rc && (std::fscanf(fp, "%d %d %d %d", &nf, &nof, &nint, &ndble) != 4 || nf <= 0 || nof != 3))
    rc = -4;
  std::vector<int> f;
  std::vector<double> dp;
  if (!rc) {
    f.resize(3 * (size_t)nf);
    for (int i = 0; i < 3 * nf && !rc; ++i)
      if (std::fscanf(fp, "%d", &f[i]) != 1) rc = -5;
    for (int i = 0; i < nf * nint && !rc; ++i) {
      int dummy;
      if (std::fscanf(fp, "%d", &dummy) != 1) rc = -6;
    }
    if (ndble > 0) {
      dp.resize((size_t)nf * ndble);
      for (int i = 0; i < nf * ndble && !rc; ++i)
        if (std::fscanf(fp, "%lf", &dp[i]) != 1) rc = -7;
    }
  }
  std::fclose(fp);
  if (rc) return fail(rc, std::string("read_pbf: malformed file ") + path);
  for (int v : f)
    if (v < 0 || v >= nn) return fail(-8, "read_pbf: node index out of range");
  if (nnode) *nnode = nn;
  if (nface) *nface = nf;
  if (nodes) std::memcpy(nodes, xyz.data(), sizeof(double) * xyz.size());
  if (faces) std::memcpy(faces, f.data(), sizeof(int) * f.size());
  if (dble_para) {
    for (int i = 0; i < nf; ++i) dble_para[i] = ndble > 0 ? dp[(size_t)i * ndble] : 0.0;
  }
  return 0;
}

namespace {

// One value as the reference's mesh files print it (input_sample.pbf):
// 12-digit mantissa and a signed three-digit exponent, e.g. -7.500000000000e-001.
void put_e3(FILE *fp, double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.12e", v);
  char *e = std::strchr(buf, 'e');
  if (!e) {  // inf / nan
    std::fputs(buf, fp);
    return;
  }
  const int ex = std::atoi(e + 1);
  *e = 0;
  std::fprintf(fp, "%se%c%03d", buf, ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
}

}  // namespace

int hmvgen_write_pbf(const char *path, int nnode, const double *nodes, int nface, const int *faces,
                     int ndble, const double *dble_para) {
  if (!path || nnode <= 0 || nface <= 0 || !nodes || !faces || ndble < 0 ||
      (ndble > 0 && !dble_para))
    return fail(-1, "write_pbf: invalid argument");
  for (int64_t i = 0; i < 3 * (int64_t)nface; ++i)
    if (faces[i] < 0 || faces[i] >= nnode) return fail(-1, "write_pbf: node index out of range");
  FILE *fp = std::fopen(path, "w");
  if (!fp) return fail(-1, std::string("write_pbf: cannot open ") + path);
  // Read_model_data (bem-bb-fw-HACApK-0.4.2.f90:411-640) reads, list-directed:
  // nond; nond coordinate triples; nofc, nodes per face (3), integer and
  // double parameters per face; the faces' 0-based node triples; the
  // integer parameters (none here); the double parameters, one face a line.
  std::fprintf(fp, "%6d\n", nnode);
  for (int i = 0; i < nnode; ++i) {
    for (int k = 0; k < 3; ++k) {
      if (k) std::fputc('\t', fp);
      put_e3(fp, nodes[3 * (size_t)i + k]);
    }
    std::fputc('\n', fp);
  }
  std::fprintf(fp, "%6d\n%6d\n%6d\n%6d\n", nface, 3, 0, ndble);
  for (int i = 0; i < nface; ++i)
    std::fprintf(fp, "%6d\t%6d\t%6d\n", faces[3 * (size_t)i], faces[3 * (size_t)i + 1],
                 faces[3 * (size_t)i + 2]);
  for (int i = 0; i < nface && ndble > 0; ++i) {
    for (int k = 0; k < ndble; ++k) {
      put_e3(fp, dble_para[(size_t)i * ndble + k]);
      std::fputc('\t', fp);
    }
    std::fputc('\n', fp);
  }
  const bool ok = std::ferror(fp) == 0;
  if (std::fclose(fp) != 0 || !ok) return fail(-1, std::string("write_pbf: write failed ") + path);
  return 0;
}

int hmvgen_write_vtk(const char *path, int nnode, const double *nodes, int nface, const int *faces,
                     const double *sol) {
  if (!path || nnode <= 0 || nface <= 0 || !nodes || !faces || !sol)
    return fail(-1, "write_vtk: invalid argument");
  FILE *fp = std::fopen(path, "w");
  if (!fp) return fail(-1, std::string("write_vtk: cannot open ") + path);
  // Print_result (bem-bb-fw-HACApK-0.4.2.f90:793-855): legacy ASCII
  // unstructured grid, points as f16.13, triangles (cell type 5) with their
  // 0-based node indices, the solution per face as f32.28.  The reference's
  // list-directed lines (cell types, CELL_DATA, SCALARS, LOOKUP_TABLE) carry
  // compiler-chosen blanks; here they are written with single spaces.
  std::fprintf(fp, "# vtk DataFile Version 2.0\n");
  std::fprintf(fp, "header : this data is results of BEM-BB\n");
  std::fprintf(fp, "ASCII\n");
  std::fprintf(fp, "DATASET UNSTRUCTURED_GRID\n");
  std::fprintf(fp, "POINTS %d float\n", nnode);
  for (int i = 0; i < nnode; ++i)
    std::fprintf(fp, "%16.13f %16.13f %16.13f\n", nodes[3 * (size_t)i], nodes[3 * (size_t)i + 1],
                 nodes[3 * (size_t)i + 2]);
  std::fprintf(fp, "CELLS %d %lld\n", nface, 4LL * nface);
  for (int i = 0; i < nface; ++i)
    std::fprintf(fp, "3 %d %d %d\n", faces[3 * (size_t)i], faces[3 * (size_t)i + 1],
                 faces[3 * (size_t)i + 2]);
  std::fprintf(fp, "CELL_TYPES %d\n", nface);
  for (int i = 0; i < nface; ++i) std::fprintf(fp, " 5\n");
  std::fprintf(fp, " CELL_DATA %d\n", nface);
  std::fprintf(fp, " SCALARS solve float 1\n");
  std::fprintf(fp, " LOOKUP_TABLE default\n");
  for (int i = 0; i < nface; ++i) std::fprintf(fp, "%32.28f\n", sol[i]);
  const bool ok = std::ferror(fp) == 0;
  if (std::fclose(fp) != 0 || !ok) return fail(-1, std::string("write_vtk: write failed ") + path);
  return 0;
}

int hmvgen_build(int nnode, const double *nodes, int nface, const int *faces,
                 const hmvgen_params *pin, hmvgen_hmatrix **out) {
  return build_impl(nnode, nodes, nface, faces, pin, out, true);
}

int hmvgen_build_blocks(int nnode, const double *nodes, int nface, const int *faces,
                        const hmvgen_params *pin, hmvgen_hmatrix **out) {
  return build_impl(nnode, nodes, nface, faces, pin, out, false);
}

}  // extern "C"

namespace {
int build_impl(int nnode, const double *nodes, int nface, const int *faces,
               const hmvgen_params *pin, hmvgen_hmatrix **out, bool fill) {
  if (!out) return fail(-1, "build: null output");
  *out = nullptr;
  if (nnode <= 0 || nface <= 0 || !nodes || !faces) return fail(-1, "build: empty mesh");
  hmvgen_params p;
  if (pin)
    p = *pin;
  else
    hmvgen_default_params(&p);
  if (p.aca_kind != 2) return fail(-2, "build: only ACA+ (param(60)=2) is implemented");
  if (p.aca_norm != 1) return fail(-2, "build: only param(61)=1 is implemented");
  // HACApK_chk_st_ctl (m_HACApK_base.f90:392-400)
  if (p.leaf_size < 2) p.leaf_size = 3;
  if (p.kmax < 1 || p.kmin < 0) return fail(-2, "build: invalid kmax/kmin");
  for (int i = 0; i < 3 * nface; ++i)
    if (faces[i] < 0 || faces[i] >= nnode) return fail(-3, "build: face references a missing node");

  auto t0 = std::chrono::steady_clock::now();
  std::unique_ptr<hmvgen_hmatrix> h(new hmvgen_hmatrix());
  h->params = p;
  std::memset(&h->stats, 0, sizeof(h->stats));
  Bem &bem = h->bem;
  bem.kernel = p.kernel;
  bem.nface = nface;
  bem.nodes.assign(nodes, nodes + 3 * (size_t)nnode);
  bem.faces.assign(faces, faces + 3 * (size_t)nface);
  bem.cent.resize(3 * (size_t)nface);
  bem.area.resize(nface);
  for (int fidx = 0; fidx < nface; ++fidx) {
    double xs[3], ys[3], zs[3];
    bem.tri(fidx, xs, ys, zs);
    // m_ppohBEM_bembb2hacapk.f90:102-104 (centre of balance)
    bem.cent[3 * fidx] = (xs[0] + xs[1] + xs[2]) / 3.0;
    bem.cent[3 * fidx + 1] = (ys[0] + ys[1] + ys[2]) / 3.0;
    bem.cent[3 * fidx + 2] = (zs[0] + zs[1] + zs[2]) / 3.0;
    double u[3] = {xs[1] - xs[0], ys[1] - ys[0], zs[1] - zs[0]};
    double v[3] = {xs[2] - xs[0], ys[2] - ys[0], zs[2] - zs[0]};
    double w[3];
    cross3(u, v, w);
    bem.area[fidx] = 0.5 * std::sqrt(dot3(w, w));
  }

  // Cluster tree over the centroids (nffc = 1).
  const int nd = nface;
  h->lod.resize(nd);
  for (int i = 0; i < nd; ++i) h->lod[i] = i + 1;
  Tree tr;
  tr.g = bem.cent.data();
  tr.minsz = p.leaf_size;
  tr.cl.reserve(4 * (size_t)nd / std::max(1, p.leaf_size) + 16);
  tr.build(h->lod.data(), 1, nd, 1);
  tr.bndbox(0, h->lod.data());

  BlockGen bg;
  bg.tr = &tr;
  bg.nleaf = p.leaf_size + 1;
  bg.nlmax = (int)(p.max_leaf_frac * nd);
  bg.zeta = p.eta;
  bg.gen(0, 0);
  std::sort(bg.leaves.begin(), bg.leaves.end(), [](const LeafSpec &a, const LeafSpec &b) {
    return a.nstrtl != b.nstrtl ? a.nstrtl < b.nstrtl : a.nstrtt < b.nstrtt;
  });
  auto t1 = std::chrono::steady_clock::now();

  const int nlf = (int)bg.leaves.size();
  h->leaves.resize(nlf);
  h->store.resize(nlf);
  int nthr = p.nthreads;
#ifdef _OPENMP
  if (nthr <= 0) nthr = omp_get_max_threads();
#else
  nthr = 1;
#endif
  int err = 0;
  // Largest blocks first so that the dynamic schedule of the fill ends
  // balanced (the block structure alone keeps leaf order).
  std::vector<int> order(nlf);
  for (int i = 0; i < nlf; ++i) order[i] = i;
  if (fill)
    std::sort(order.begin(), order.end(), [&](int a, int b) {
      int64_t sa = (int64_t)bg.leaves[a].ndl * bg.leaves[a].ndt;
      int64_t sb = (int64_t)bg.leaves[b].ndl * bg.leaves[b].ndt;
      return sa != sb ? sa > sb : a < b;
    });
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthr)
  for (int q = 0; q < nlf; ++q) {
    const int ip = order[q];
    const LeafSpec &ls = bg.leaves[ip];
    LeafRec &lr = h->leaves[ip];
    lr.ltmtx = ls.ltmtx;
    lr.nstrtl = ls.nstrtl;
    lr.ndl = ls.ndl;
    lr.nstrtt = ls.nstrtt;
    lr.ndt = ls.ndt;
    if (!fill) {  // block structure only: no factors
      lr.kt = 0;
      lr.a1 = nullptr;
      lr.a2 = nullptr;
      lr.a1size = 0;
      continue;
    }
    if (ls.ltmtx == 1) {
      Aca aca;
      aca.bem = &bem;
      aca.lod = h->lod.data();
      aca.ndl = ls.ndl;
      aca.ndt = ls.ndt;
      aca.nstrtl = ls.nstrtl;
      aca.nstrtt = ls.nstrtt;
      int kt = aca.run(p);
      if (kt < 0) kt = 0;
      if (kt > aca.cap) aca.ensure(kt);
      lr.kt = kt;
      size_t n1 = (size_t)ls.ndt * kt, n2 = (size_t)ls.ndl * kt;
      double *buf = new (std::nothrow) double[std::max<size_t>(1, n1 + n2)];
      if (!buf) {
#pragma omp atomic write
        err = 1;
        continue;
      }
      std::memcpy(buf, aca.V.data(), n1 * sizeof(double));
      std::memcpy(buf + n1, aca.U.data(), n2 * sizeof(double));
      h->store[ip].reset(buf);
      lr.a1 = buf;
      lr.a2 = buf + n1;
      lr.a1size = (long long)(n1 * sizeof(double));
    } else {
      lr.kt = 0;
      size_t n1 = (size_t)ls.ndt * ls.ndl;
      double *buf = new (std::nothrow) double[std::max<size_t>(1, n1)];
      if (!buf) {
#pragma omp atomic write
        err = 1;
        continue;
      }
      for (int il = 1; il <= ls.ndl; ++il) {
        int ill = il + ls.nstrtl - 1;
        for (int it = 1; it <= ls.ndt; ++it) {
          int itt = it + ls.nstrtt - 1;
          buf[(size_t)(il - 1) * ls.ndt + (it - 1)] = bem.entry(h->lod[ill - 1], h->lod[itt - 1]);
        }
      }
      h->store[ip].reset(buf);
      lr.a1 = buf;
      lr.a2 = nullptr;
      lr.a1size = (long long)(n1 * sizeof(double));
    }
  }
  if (err) return fail(-4, "build: out of host memory during fill");
  auto t2 = std::chrono::steady_clock::now();

  hmvgen_stats &s = h->stats;
  s.nd = nd;
  s.nclst = (int)tr.cl.size();
  s.nlf = nlf;
  s.nlfkt = bg.nlr;
  s.nlfdense = bg.ndense;
  s.nhblocks = bg.nh;
  int64_t ncpc = 0, mem8 = 0;
  int ktmax = 0;
  for (const LeafRec &lr : h->leaves) {
    if (lr.ltmtx == 1) {
      ncpc += (int64_t)(lr.ndl + lr.ndt) * p.kt_predict;
      mem8 += (int64_t)(lr.ndl + lr.ndt) * lr.kt;
      ktmax = std::max(ktmax, lr.kt);
    } else {
      ncpc += (int64_t)lr.ndl * lr.ndt;
      mem8 += (int64_t)lr.ndl * lr.ndt;
    }
  }
  s.ncpc = ncpc;
  s.ktmax = ktmax;
  s.mem_entries = mem8;
  // znn=real(memh)*8.0/1024.0/1024.0 evaluated in REAL*4 (m_HACApK_base.f90:1153-1154)
  float znn = (float)(mem8 + bg.nlr + bg.ndense) * 8.0f / 1024.0f / 1024.0f;
  s.mem_mbyte = (double)znn;
  s.time_tree = std::chrono::duration<double>(t1 - t0).count();
  s.time_fill = std::chrono::duration<double>(t2 - t1).count();
  *out = h.release();
  return 0;
}
}  // namespace

extern "C" {

void hmvgen_free(hmvgen_hmatrix *h) { delete h; }

const double *hmvgen_areas(const hmvgen_hmatrix *h) { return h ? h->bem.area.data() : nullptr; }

int hmvgen_get_stats(const hmvgen_hmatrix *h, hmvgen_stats *s) {
  if (!h || !s) return fail(-1, "get_stats: null argument");
  *s = h->stats;
  return 0;
}

const void *hmvgen_leaves(const hmvgen_hmatrix *h, int *nlf, int *stride) {
  if (!h) return nullptr;
  if (nlf) *nlf = (int)h->leaves.size();
  if (stride) *stride = (int)sizeof(LeafRec);
  return h->leaves.data();
}

const int *hmvgen_lod(const hmvgen_hmatrix *h) { return h ? h->lod.data() : nullptr; }

const double *hmvgen_centroids(const hmvgen_hmatrix *h) { return h ? h->bem.cent.data() : nullptr; }

double hmvgen_entry(const hmvgen_hmatrix *h, int i, int j) {
  if (!h || i < 1 || j < 1 || i > h->bem.nface || j > h->bem.nface) return NAN;
  return h->bem.entry(i, j);
}

double hmvgen_face_integral(int kernel, const double *xs, const double *ys, const double *zs,
                            double x, double y, double z) {
  return face_integral_mode(kernel, xs, ys, zs, x, y, z);
}

double hmvgen_util_unrm(int nd, const double *za) { return unrm(nd, za); }

int hmvgen_util_maxabs(int nd, const double *za, double *zz) {
  int il;
  double z;
  maxabsvalloc(za, z, il, nd);
  if (zz) *zz = z;
  return il;
}

int hmvgen_util_minabs(int nd, const double *za, double *zz) {
  int il;
  double z;
  minabsvalloc(za, z, il, nd);
  if (zz) *zz = z;
  return il;
}

void hmvgen_uniform(uint64_t seed, int64_t n, double *out) {
  uint64_t st = seed;
  for (int64_t i = 0; i < n; ++i) {
    st += 0x9E3779B97F4A7C15ULL;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    out[i] = 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
  }
}

}  // extern "C"
