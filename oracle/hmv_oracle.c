/*
 * hmv_oracle.c -- CPU oracle for the H-matvec path.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker or as the timed CPU baseline; the
 * engine (libhmv.so) never calls it.
 *
 * Functions and the reference code each one restates:
 *   oracle_adot_body_lfmtx   HACApK_MAGMA.c:5-57      (C loop apply, zau += H*zu)
 *   oracle_adot_lfmtx_hyp    m_HACApK_solve.f90:127-161, 203-253 (zero zau, apply over
 *                            storage-balanced leaf ranges with thread-private zaut), with
 *                            the per-leaf GEMVs of HACApK_MAGMA.c:114-199 (CPU BLAS path)
 *   oracle_setcutthread      m_HACApK_base.f90:584-611 (HACApK_setcutthread, ktp=0)
 *   oracle_bicgstab          m_HACApK_solve.f90:342-449 (HACApK_bicgstab_cax_lfmtx_hyp)
 *   oracle_gcrm              m_HACApK_solve.f90:506-571 (HACApK_gcrm_lfmtx)
 *   oracle_bicgstab_pipe     HACApK_SOLVER_PIPE.c:144-637 (c_hacapk_bicgstab_cax_lfmtx_pipe_,
 *                            pipelined BiCGSTAB, single process, no diagonal scaling)
 *   oracle_dotp              HACApK_lib.f90:26-29 (HACApK_dotp_d)
 *   oracle_dense             brute-force assembly of H from its leaves (for checks)
 *
 * Pinning (what fixes each function independently of this file):
 *   - adot_body_lfmtx / adot_lfmtx_hyp: against the dense matrix assembled from the
 *     same leaves (closed definition of the leaf apply), against each other, and
 *     against golden y=Hx fixtures in tests/golden.
 *   - bicgstab: against the residual history recorded by the reference at
 *     N=10,400 (ivb/job_cpu.sh.o78632:29-37), against the iterates of scipy's
 *     own BiCGSTAB (same method and shadow residual) on a small H-matrix, and
 *     the true-residual invariant on small dense systems.
 *   - gcrm: within one restart cycle GCR minimises ||r|| over the Krylov space,
 *     so its residual norms equal GMRES's (scipy) to rounding; convergence to the
 *     dense solution.  The restart step is the reference's own (see the
 *     function): the step after a restart equals its closed form built from a
 *     QR factorisation of the Krylov matrix (GMRES(m) residual, slot-1 scaling
 *     1 + beta1 with the old capap(1), projection on slots 2..m).
 *   - bicgstab_pipe: in exact arithmetic the pipelined recurrences produce the
 *     iterates of BiCGSTAB with the same shadow residual, so its residual norms
 *     equal oracle_bicgstab's (itself pinned above) while far above rounding;
 *     convergence to the dense solution.
 *   - setcutthread: leaf ranges tile 1..nlf (invariant test).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* stc_HACApK_leafmtx, HACApK_MAGMA.h:12-20 */
typedef struct {
  int ltmtx;
  int kt;
  int nstrtl, ndl;
  int nstrtt, ndt;
  long long a1size;
  double *a1;
  double *a2;
} oracle_leaf;

static const oracle_leaf *leaf_at(const void *base, int stride, int ip) {
  return (const oracle_leaf *)((const char *)base + (size_t)stride * (size_t)ip);
}

/* HACApK_MAGMA.c:5-57, loop order kept. */
void oracle_adot_body_lfmtx(double *zau, const void *leaves, int nlf, int stride,
                            const double *zu, double *zbu) {
  for (int ip = 0; ip < nlf; ip++) {
    const oracle_leaf *s = leaf_at(leaves, stride, ip);
    const int ndl = s->ndl, ndt = s->ndt, nstrtl = s->nstrtl, nstrtt = s->nstrtt;
    if (s->ltmtx == 1) {
      const double *a2 = (const double *)((const char *)s->a1 + s->a1size);
      const int kt = s->kt;
      for (int il = 0; il < kt; il++) {
        zbu[il] = 0.0;
        for (int it = 0; it < ndt; it++) zbu[il] += s->a1[it + (size_t)il * ndt] * zu[it + nstrtt - 1];
      }
      for (int il = 0; il < kt; il++)
        for (int it = 0; it < ndl; it++) zau[it + nstrtl - 1] += a2[it + (size_t)il * ndl] * zbu[il];
    } else if (s->ltmtx == 2) {
      for (int il = 0; il < ndl; il++)
        for (int it = 0; it < ndt; it++)
          zau[il + nstrtl - 1] += s->a1[it + (size_t)il * ndt] * zu[it + nstrtt - 1];
    }
  }
}

/* m_HACApK_base.f90:584-611 with ktp=0 (actual ranks), as called from
 * HACApK_generate (m_HACApK_use.f90:165-166).  lthr has nthr+1 entries,
 * 1-based leaf indices, thread t owns lthr[t]..lthr[t+1]-1. */
void oracle_setcutthread(const void *leaves, int nlf, int stride, int nthr, int *lthr) {
  int64_t mem8 = 0;
  for (int il = 0; il < nlf; il++) {
    const oracle_leaf *s = leaf_at(leaves, stride, il);
    mem8 += s->ltmtx == 1 ? (int64_t)(s->ndl + s->ndt) * s->kt : (int64_t)s->ndl * s->ndt;
  }
  for (int t = 0; t <= nthr; t++) lthr[t] = nlf + 1;
  lthr[0] = 1;
  if (nthr <= 1) return;
  const int64_t nth1_mem = mem8 / nthr;
  int64_t imem = 0;
  int ith = 1;
  for (int il = 1; il <= nlf; il++) {
    const oracle_leaf *s = leaf_at(leaves, stride, il - 1);
    imem += s->ltmtx == 1 ? (int64_t)(s->ndl + s->ndt) * s->kt : (int64_t)s->ndl * s->ndt;
    if (imem > nth1_mem * ith) {
      lthr[ith] = il;
      ith++;
      if (ith == nthr) break;
    }
  }
}

/* Per-leaf GEMVs of the CPU BLAS path (HACApK_MAGMA.c:140-143, 163-166, 184-187):
 * dgemv("T") dot products with four fixed partial sums, dgemv("N") as axpys. */
static double dot4(int n, const double *a, const double *b) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int i = 0;
  for (; i + 4 <= n; i += 4) {
    s0 += a[i] * b[i];
    s1 += a[i + 1] * b[i + 1];
    s2 += a[i + 2] * b[i + 2];
    s3 += a[i + 3] * b[i + 3];
  }
  for (; i < n; i++) s0 += a[i] * b[i];
  return (s0 + s1) + (s2 + s3);
}

static void leaf_apply_blas(const oracle_leaf *s, const double *zu, double *zaut, double *zbu) {
  const int ndl = s->ndl, ndt = s->ndt;
  const double *x = zu + (s->nstrtt - 1);
  double *y = zaut + (s->nstrtl - 1);
  if (s->ltmtx == 1) {
    const double *a2 = (const double *)((const char *)s->a1 + s->a1size);
    const int kt = s->kt;
    for (int il = 0; il < kt; il++) zbu[il] = dot4(ndt, s->a1 + (size_t)il * ndt, x);
    for (int il = 0; il < kt; il++) {
      const double b = zbu[il];
      const double *u = a2 + (size_t)il * ndl;
      for (int it = 0; it < ndl; it++) y[it] += u[it] * b;
    }
  } else if (s->ltmtx == 2) {
    for (int il = 0; il < ndl; il++) y[il] += dot4(ndt, s->a1 + (size_t)il * ndt, x);
  }
}

/* m_HACApK_solve.f90:127-161 + 203-253: zau = 0, then every thread applies its
 * storage-balanced leaf range into a private zaut and the partial vectors are
 * added in thread order (fixed order, so the result depends only on nthr). */
int oracle_adot_lfmtx_hyp(double *zau, const void *leaves, int nlf, int stride, int nd,
                          const double *zu, int nthr) {
  if (nthr < 1) nthr = 1;
  int *lthr = (int *)malloc(sizeof(int) * (nthr + 1));
  double *zaut = (double *)calloc((size_t)nd * nthr, sizeof(double));
  int *ls = (int *)malloc(sizeof(int) * nthr);
  int *le = (int *)malloc(sizeof(int) * nthr);
  if (!lthr || !zaut || !ls || !le) {
    free(lthr);
    free(zaut);
    free(ls);
    free(le);
    return -1;
  }
  oracle_setcutthread(leaves, nlf, stride, nthr, lthr);
  int ktmax = 1;
  for (int ip = 0; ip < nlf; ip++) {
    const oracle_leaf *s = leaf_at(leaves, stride, ip);
    if (s->ltmtx == 1 && s->kt > ktmax) ktmax = s->kt;
  }
#pragma omp parallel num_threads(nthr)
  {
#ifdef _OPENMP
    const int ith = omp_get_thread_num();
#else
    const int ith = 0;
#endif
    if (ith < nthr) {
      double *mine = zaut + (size_t)nd * ith;
      double *zbut = (double *)malloc(sizeof(double) * ktmax);
      int lo = nd, hi = 1;
      for (int ip = lthr[ith]; ip <= lthr[ith + 1] - 1; ip++) {
        const oracle_leaf *s = leaf_at(leaves, stride, ip - 1);
        if (s->nstrtl < lo) lo = s->nstrtl;
        if (s->nstrtl + s->ndl - 1 > hi) hi = s->nstrtl + s->ndl - 1;
        leaf_apply_blas(s, zu, mine, zbut);
      }
      ls[ith] = lo;
      le[ith] = hi;
      free(zbut);
    }
  }
  for (int i = 0; i < nd; i++) zau[i] = 0.0;
  for (int t = 0; t < nthr; t++) {
    const double *mine = zaut + (size_t)nd * t;
    for (int il = ls[t]; il <= le[t]; il++) zau[il - 1] = zau[il - 1] + mine[il - 1];
  }
  free(lthr);
  free(zaut);
  free(ls);
  free(le);
  return 0;
}

/* HACApK_dotp_d (HACApK_lib.f90:26-29), sequential sum. */
double oracle_dotp(int nd, const double *za, const double *zb) {
  double s = 0.0;
  for (int i = 0; i < nd; i++) s += za[i] * zb[i];
  return s;
}

/* HACApK_bicgstab_cax_lfmtx_hyp (m_HACApK_solve.f90:342-449), single rank.
 * u: initial guess in / solution out; hist[0..mstep] receives ||r||/||b|| after
 * each step (hist[0] = initial).  Returns the number of steps taken (nstp). */
int oracle_bicgstab(const void *leaves, int nlf, int stride, int nd, double *u,
                    const double *b, int mstep, double eps, int nthr, double *hist) {
  double *zr = (double *)malloc(sizeof(double) * nd);
  double *zshdw = (double *)malloc(sizeof(double) * nd);
  double *zp = (double *)malloc(sizeof(double) * nd);
  double *zt = (double *)malloc(sizeof(double) * nd);
  double *zkp = (double *)malloc(sizeof(double) * nd);
  double *zakp = (double *)malloc(sizeof(double) * nd);
  double *zkt = (double *)malloc(sizeof(double) * nd);
  double *zakt = (double *)malloc(sizeof(double) * nd);
  int nstp = 0;
  if (!zr || !zshdw || !zp || !zt || !zkp || !zakp || !zkt || !zakt) {
    nstp = -1;
    goto done;
  }
  double alpha = 0.0, beta = 0.0, zeta = 0.0;
  double zz = oracle_dotp(nd, b, b);
  double bnorm = sqrt(zz);
  for (int i = 0; i < nd; i++) {
    zp[i] = 0.0;
    zakp[i] = 0.0;
    zr[i] = b[i];
  }
  /* HACApK_adotsub_lfmtx_hyp: zshdw = A u; zr = zr - zshdw */
  oracle_adot_lfmtx_hyp(zshdw, leaves, nlf, stride, nd, u, nthr);
  for (int i = 0; i < nd; i++) zr[i] = zr[i] - zshdw[i];
  for (int i = 0; i < nd; i++) zshdw[i] = zr[i];
  double zrnorm = sqrt(oracle_dotp(nd, zr, zr));
  if (hist) hist[0] = zrnorm / bnorm;
  for (int in = 1; in <= mstep; in++) {
    if (zrnorm / bnorm < eps) break;
    for (int i = 0; i < nd; i++) zp[i] = zr[i] + beta * (zp[i] - zeta * zakp[i]);
    for (int i = 0; i < nd; i++) zkp[i] = zp[i];
    oracle_adot_lfmtx_hyp(zakp, leaves, nlf, stride, nd, zkp, nthr);
    double znom = oracle_dotp(nd, zshdw, zr);
    double zden = oracle_dotp(nd, zshdw, zakp);
    alpha = znom / zden;
    double znomold = znom;
    for (int i = 0; i < nd; i++) zt[i] = zr[i] - alpha * zakp[i];
    for (int i = 0; i < nd; i++) zkt[i] = zt[i];
    oracle_adot_lfmtx_hyp(zakt, leaves, nlf, stride, nd, zkt, nthr);
    znom = oracle_dotp(nd, zakt, zt);
    zden = oracle_dotp(nd, zakt, zakt);
    zeta = znom / zden;
    for (int i = 0; i < nd; i++) u[i] = u[i] + alpha * zkp[i] + zeta * zkt[i];
    for (int i = 0; i < nd; i++) zr[i] = zt[i] - zeta * zakt[i];
    beta = alpha / zeta * oracle_dotp(nd, zshdw, zr) / znomold;
    zrnorm = sqrt(oracle_dotp(nd, zr, zr));
    nstp = in;
    if (hist) hist[in] = zrnorm / bnorm;
  }
done:
  free(zr);
  free(zshdw);
  free(zp);
  free(zt);
  free(zkp);
  free(zakp);
  free(zkt);
  free(zakt);
  return nstp;
}

/* HACApK_gcrm_lfmtx (m_HACApK_solve.f90:506-571), single rank: GCR(m) with
 * mreset = m.  Follows the reference literally, including its restart step:
 * when the new direction goes to slot 1 (in a multiple of m), slot 1's
 * A-direction has already been overwritten by A r, so the loop over slots
 * 1..m projects A r on itself first (beta_1 = -(Ar,Ar)/capap(1)) and then on
 * the old slots 2..m.  u: initial guess in / solution out; hist[0..mstep]
 * receives ||r||/||b||.  Returns nstp (the reference's in at exit). */
int oracle_gcrm(const void *leaves, int nlf, int stride, int nd, double *u, const double *b,
                int mstep, int mreset, double eps, int nthr, double *hist) {
  if (mreset < 1) return -1;
  double *zr = (double *)malloc(sizeof(double) * nd);
  double *zar = (double *)malloc(sizeof(double) * nd);
  double *zp = (double *)malloc(sizeof(double) * (size_t)nd * mreset);
  double *zap = (double *)malloc(sizeof(double) * (size_t)nd * mreset);
  double *capap = (double *)malloc(sizeof(double) * mreset);
  int nstp = 0;
  if (!zr || !zar || !zp || !zap || !capap) {
    nstp = -1;
    goto done;
  }
  {
    double bnorm = sqrt(oracle_dotp(nd, b, b));
    oracle_adot_lfmtx_hyp(zar, leaves, nlf, stride, nd, u, nthr);
    for (int i = 0; i < nd; i++) zr[i] = b[i] - zar[i];
    for (int i = 0; i < nd; i++) zp[i] = zr[i];  /* zp(:,1) */
    double zrnorm2 = oracle_dotp(nd, zr, zr);
    double zrnorm = sqrt(zrnorm2);
    if (hist) hist[0] = zrnorm / bnorm;
    if (zrnorm / bnorm < eps) goto done;
    oracle_adot_lfmtx_hyp(zap, leaves, nlf, stride, nd, zp, nthr); /* zap(:,1) */
    int in;
    for (in = 1; in <= mstep; in++) {
      const int ik = (in - 1) % mreset; /* 0-based slot of zp(:,ik) */
      double *zq = zap + (size_t)ik * nd;
      double znom = oracle_dotp(nd, zq, zr);
      capap[ik] = oracle_dotp(nd, zq, zq);
      double alpha = znom / capap[ik];
      for (int i = 0; i < nd; i++) u[i] = u[i] + alpha * zp[(size_t)ik * nd + i];
      for (int i = 0; i < nd; i++) zr[i] = zr[i] - alpha * zq[i];
      zrnorm2 = oracle_dotp(nd, zr, zr);
      zrnorm = sqrt(zrnorm2);
      if (hist) hist[in] = zrnorm / bnorm;
      nstp = in;
      if (zrnorm / bnorm < eps || in == mstep) break;
      oracle_adot_lfmtx_hyp(zar, leaves, nlf, stride, nd, zr, nthr);
      const int ikn = in % mreset; /* 0-based slot of zp(:,ikn) */
      double *pn = zp + (size_t)ikn * nd, *apn = zap + (size_t)ikn * nd;
      for (int i = 0; i < nd; i++) pn[i] = zr[i];
      for (int i = 0; i < nd; i++) apn[i] = zar[i];
      for (int il = 0; il <= ik; il++) {
        const double *q = zap + (size_t)il * nd, *pl = zp + (size_t)il * nd;
        double beta = -oracle_dotp(nd, q, zar) / capap[il];
        for (int i = 0; i < nd; i++) pn[i] = pn[i] + beta * pl[i];
        for (int i = 0; i < nd; i++) apn[i] = apn[i] + beta * q[i];
      }
    }
  }
done:
  free(zr);
  free(zar);
  free(zp);
  free(zap);
  free(capap);
  return nstp;
}

/* HACApK_SOLVER_PIPE.c:144-637, one process.  The reference accumulates the
 * iterate in a device vector zx that it never initialises nor returns (it
 * copies u, the unchanged initial guess, back; :188,483-484,595); here
 * zx starts at u and is returned in u.  Every other recurrence is kept, in
 * the reference's order.  hist[0] initial, hist[step] after each step. */
int oracle_bicgstab_pipe(const void *leaves, int nlf, int stride, int nd, double *u,
                         const double *b, int mstep, double eps, int nthr, double *hist) {
  double *mem = (double *)malloc(sizeof(double) * (size_t)nd * 11);
  if (!mem) return -1;
  double *zr = mem, *zr0 = zr + nd, *zw = zr0 + nd, *zt = zw + nd, *zv = zt + nd;
  double *zp = zv + nd, *zs = zp + nd, *zz = zs + nd, *zq = zz + nd, *zy = zq + nd;
  double *zx = u;
  double alpha = 0.0, beta = 0.0, zeta = 0.0;
  const double bnorm = sqrt(oracle_dotp(nd, b, b));
  oracle_adot_lfmtx_hyp(zv, leaves, nlf, stride, nd, u, nthr);        /* zv = A u  */
  for (int i = 0; i < nd; i++) zr[i] = b[i] - zv[i];                  /* zr = b - zv */
  memcpy(zr0, zr, sizeof(double) * nd);
  oracle_adot_lfmtx_hyp(zw, leaves, nlf, stride, nd, zr, nthr);       /* zw = A zr */
  oracle_adot_lfmtx_hyp(zt, leaves, nlf, stride, nd, zw, nthr);       /* zt = A zw */
  double znorm = oracle_dotp(nd, zr, zr);
  double zrnorm = oracle_dotp(nd, zw, zr);
  alpha = znorm / zrnorm;
  zrnorm = sqrt(znorm);
  if (hist) hist[0] = zrnorm / bnorm;
  int nstp = 0;
  for (int step = 1; step <= mstep; step++) {
    if (zrnorm / bnorm < eps) break;
    if (step > 1) {
      for (int i = 0; i < nd; i++) zp[i] = zr[i] + beta * (zp[i] - zeta * zs[i]);
      for (int i = 0; i < nd; i++) zs[i] = zw[i] + beta * (zs[i] - zeta * zz[i]);
      for (int i = 0; i < nd; i++) zz[i] = zt[i] + beta * (zz[i] - zeta * zv[i]);
    } else {
      memcpy(zp, zr, sizeof(double) * nd);
      memcpy(zs, zw, sizeof(double) * nd);
      memcpy(zz, zt, sizeof(double) * nd);
    }
    for (int i = 0; i < nd; i++) zq[i] = zr[i] - alpha * zs[i];
    for (int i = 0; i < nd; i++) zy[i] = zw[i] - alpha * zz[i];
    const double zqy = oracle_dotp(nd, zy, zq), zden = oracle_dotp(nd, zy, zy);
    zeta = zqy / zden;
    oracle_adot_lfmtx_hyp(zv, leaves, nlf, stride, nd, zz, nthr);     /* zv = A zz */
    for (int i = 0; i < nd; i++) zx[i] = zx[i] + alpha * zp[i] + zeta * zq[i];
    for (int i = 0; i < nd; i++) zr[i] = zq[i] - zeta * zy[i];
    /* zw = zy - zeta*(zt - alpha*zv) (the code's dlacpy/daxpy/dlascl/daxpy) */
    for (int i = 0; i < nd; i++) zw[i] = zy[i] + (-zeta) * (zt[i] - alpha * zv[i]);
    const double znormold = znorm;
    znorm = oracle_dotp(nd, zr, zr0);
    const double zwnorm = oracle_dotp(nd, zw, zr0);
    const double zsnorm = oracle_dotp(nd, zs, zr0);
    const double zznorm = oracle_dotp(nd, zz, zr0);
    const double rr = oracle_dotp(nd, zr, zr);
    beta = (alpha / zeta) * (znorm / znormold);
    alpha = znorm / (zwnorm + beta * zsnorm - beta * zeta * zznorm);
    zrnorm = sqrt(rr);
    nstp = step;
    if (hist) hist[step] = zrnorm / bnorm;
    oracle_adot_lfmtx_hyp(zt, leaves, nlf, stride, nd, zw, nthr);     /* zt = A zw */
  }
  free(mem);
  return nstp;
}

/* Dense nd x nd (row-major) matrix assembled from the leaves. */
void oracle_dense(const void *leaves, int nlf, int stride, int nd, double *a) {
  memset(a, 0, sizeof(double) * (size_t)nd * nd);
  for (int ip = 0; ip < nlf; ip++) {
    const oracle_leaf *s = leaf_at(leaves, stride, ip);
    const int r0 = s->nstrtl - 1, c0 = s->nstrtt - 1;
    if (s->ltmtx == 1) {
      const double *a2 = (const double *)((const char *)s->a1 + s->a1size);
      for (int i = 0; i < s->ndl; i++)
        for (int j = 0; j < s->ndt; j++) {
          double v = 0.0;
          for (int l = 0; l < s->kt; l++) v += a2[i + (size_t)l * s->ndl] * s->a1[j + (size_t)l * s->ndt];
          a[(size_t)(r0 + i) * nd + (c0 + j)] += v;
        }
    } else if (s->ltmtx == 2) {
      for (int i = 0; i < s->ndl; i++)
        for (int j = 0; j < s->ndt; j++)
          a[(size_t)(r0 + i) * nd + (c0 + j)] += s->a1[j + (size_t)i * s->ndt];
    }
  }
}

/* Number of OpenMP threads available to the oracle. */
int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
