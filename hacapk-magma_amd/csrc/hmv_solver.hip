// hmv_solver.hip -- BiCGSTAB on the H-matvec engine, all vectors on device.
//
// Same iteration as the reference's Fortran driver
// HACApK_bicgstab_cax_lfmtx_hyp (m_HACApK_solve.f90:342-449) and its all-GPU
// C version c_hacapk_bicgstab_cax_lfmtx_gpu_ (HACApK_SOLVER.c:421-807):
//   p = r + beta (p - zeta v);  v = A p;  alpha = (r^,r)/(r^,v);
//   t = r - alpha v;  s = A t;  zeta = (s,t)/(s,s);
//   u += alpha p + zeta t;  r = t - zeta s;  beta = alpha/zeta (r^,r_new)/(r^,r)
// The reference copies p->kp and t->kt (no preconditioner) and returns every
// dot product to the host; here the scalars stay on the device, dots are
// fused (two per pass, deterministic two-level reduction), one iteration is
// captured as a hipGraph, and the host reads the residual every
// `check_every` iterations.  A device flag set on convergence turns the
// remaining launched iterations into no-ops, so the result equals stopping
// at the first converged iteration.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "hmv.h"
#include "hmv_engine.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 512;

struct SolveState {
  double alpha, beta, zeta;
  double rr;      // (r^, r) of the current residual
  double rnorm;   // ||r||
  double bnorm;   // ||b||
  double eps;
  double d0, d1;  // last reduced dot products
  int step;       // iterations completed
  int done;       // 1 once converged / stopped
  int status;     // HMV_SOLVE_*
  int maxit;
};

__device__ inline double block_sum(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

// block_sum's total (same order) on every thread of the block.
__device__ inline double block_sum_all(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

// Partial sums of (a,b) and, when c != null, (c,d); part[2*block + {0,1}].
__global__ __launch_bounds__(kThreads) void k_dots(int n, const double *__restrict__ a,
                                                    const double *__restrict__ b,
                                                    const double *__restrict__ c,
                                                    const double *__restrict__ d,
                                                    double *__restrict__ part,
                                                    const SolveState *st) {
  if (st && st->done) return;
  __shared__ double sh[4];
  double s0 = 0.0, s1 = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    s0 = __builtin_fma(a[i], b[i], s0);
    if (c) s1 = __builtin_fma(c[i], d[i], s1);
  }
  double r0 = block_sum(s0, sh);
  double r1 = c ? block_sum(s1, sh) : 0.0;
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0;
    part[2 * blockIdx.x + 1] = r1;
  }
}

// Fixed-order sum of the partials into st->d0, st->d1.
__device__ inline void reduce_parts(int nb, const double *part, SolveState *st, double *sh) {
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) {
    s0 += part[2 * i];
    s1 += part[2 * i + 1];
  }
  double r0 = block_sum(s0, sh);
  double r1 = block_sum(s1, sh);
  if (threadIdx.x == 0) {
    st->d0 = r0;
    st->d1 = r1;
  }
}

__device__ inline bool finite(double v) { return isfinite(v); }

// stage: 0 setup (d0 = (r,r), d1 = (b,b)), 1 alpha, 2 zeta, 3 beta,
// 4 residual replacement (d0 = (r^, r), d1 = (r, r) of the true residual)
__global__ __launch_bounds__(kThreads) void k_final(int nb, const double *__restrict__ part,
                                                    SolveState *st, int stage,
                                                    double *__restrict__ hist) {
  if (st->done && stage != 0 && stage != 4) return;
  __shared__ double sh[4];
  reduce_parts(nb, part, st, sh);
  if (threadIdx.x != 0) return;
  if (stage == 0) {
    // d0 = (r,r) (r^ = r initially), d1 = (b,b)
    st->rr = st->d0;
    st->rnorm = sqrt(st->d0);
    st->bnorm = sqrt(st->d1);
    st->step = 0;
    st->alpha = 0.0;
    st->beta = 0.0;
    st->zeta = 0.0;
    st->status = HMV_SOLVE_MAXITER;
    st->done = 0;
    if (hist) hist[0] = st->rnorm / st->bnorm;
    if (!finite(st->rnorm) || !finite(st->bnorm) || st->bnorm == 0.0) {
      st->status = HMV_SOLVE_NONFINITE;
      st->done = 1;
    } else if (st->rnorm / st->bnorm < st->eps) {
      st->status = HMV_SOLVE_CONVERGED;
      st->done = 1;
    } else if (st->maxit <= 0) {
      st->done = 1;
    }
  } else if (stage == 1) {
    const double zden = st->d0;
    st->alpha = st->rr / zden;
    if (zden == 0.0 || !finite(st->alpha)) {
      st->status = finite(zden) ? HMV_SOLVE_BREAKDOWN : HMV_SOLVE_NONFINITE;
      st->done = 1;
    }
  } else if (stage == 2) {
    const double znom = st->d0, zden = st->d1;
    // s = A t = 0 (both dots zero): t is the residual of u + alpha p; zeta = 0
    // takes that half step and stage 3 decides (converged, else breakdown)
    const bool half = zden == 0.0 && znom == 0.0;
    st->zeta = half ? 0.0 : znom / zden;
    if (!half && (zden == 0.0 || !finite(st->zeta))) {
      st->status = finite(zden) ? HMV_SOLVE_BREAKDOWN : HMV_SOLVE_NONFINITE;
      st->done = 1;
    }
  } else if (stage == 4) {
    // the iterated residual is replaced by the true one; the iteration goes
    // on (p and r^ are kept) unless the true residual has converged
    st->rr = st->d0;
    st->rnorm = sqrt(st->d1);
    st->done = 0;
    st->status = HMV_SOLVE_MAXITER;
    if (!finite(st->rnorm)) {
      st->status = HMV_SOLVE_NONFINITE;
      st->done = 1;
    } else if (st->rnorm / st->bnorm < st->eps) {
      st->status = HMV_SOLVE_CONVERGED;
      st->done = 1;
    } else if (st->step >= st->maxit) {
      st->done = 1;
    } else if (st->rr == 0.0) {
      st->status = HMV_SOLVE_BREAKDOWN;
      st->done = 1;
    }
  } else {
    const double rr_new = st->d0;
    st->beta = st->alpha / st->zeta * rr_new / st->rr;
    st->rr = rr_new;
    st->rnorm = sqrt(st->d1);
    st->step += 1;
    if (hist) hist[st->step] = st->rnorm / st->bnorm;
    if (!finite(st->rnorm)) {
      st->status = HMV_SOLVE_NONFINITE;
      st->done = 1;
    } else if (st->rnorm / st->bnorm < st->eps) {
      st->status = HMV_SOLVE_CONVERGED;
      st->done = 1;
    } else if (st->step >= st->maxit) {
      st->status = HMV_SOLVE_MAXITER;
      st->done = 1;
    } else if (st->rr == 0.0 || st->zeta == 0.0) {
      // zeta = 0: a half step that did not converge (beta = alpha / zeta fails)
      st->status = HMV_SOLVE_BREAKDOWN;
      st->done = 1;
    }
  }
}

// p = r + beta (p - zeta v)
__global__ __launch_bounds__(kThreads) void k_update_p(int n, const double *__restrict__ r,
                                                        const double *__restrict__ v,
                                                        double *__restrict__ p,
                                                        const SolveState *st) {
  if (st->done) return;
  const double beta = st->beta, zeta = st->zeta;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    p[i] = r[i] + beta * (p[i] - zeta * v[i]);
}

// alpha = rr / (r^, v) from the nb partial pairs of k_dots, summed in
// k_final's order (every block forms it alike; block 0 records it and flags
// a breakdown), then t = r - alpha v.
__global__ __launch_bounds__(kThreads) void k_update_t(int n, int nb, const double *__restrict__ r,
                                                        const double *__restrict__ v,
                                                        double *__restrict__ t,
                                                        const double *__restrict__ part,
                                                        SolveState *st) {
  if (st->done) return;
  __shared__ double sh[4];
  double s0 = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) s0 += part[2 * i];
  const double zden = block_sum_all(s0, sh);
  const double alpha = st->rr / zden;
  const bool bad = zden == 0.0 || !finite(alpha);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->d0 = zden;
    st->alpha = alpha;
    if (bad) {
      st->status = finite(zden) ? HMV_SOLVE_BREAKDOWN : HMV_SOLVE_NONFINITE;
      st->done = 1;
    }
  }
  if (bad) return;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    t[i] = r[i] - alpha * v[i];
}

// zeta = (s, t) / (s, s) from the nb partial pairs zpart of k_dots, summed
// in k_final's order (every block alike; block 0 records it and flags a
// breakdown), then u += alpha p + zeta t;  r = t - zeta s;  partials of
// (r^, r) and (r, r) into part (apart from zpart, which other blocks are
// still reading).
__global__ __launch_bounds__(kThreads) void k_update_ur(
    int n, int nb, const double *__restrict__ p, const double *__restrict__ t,
    const double *__restrict__ s, const double *__restrict__ rhat, double *__restrict__ u,
    double *__restrict__ r, const double *__restrict__ zpart, double *__restrict__ part,
    SolveState *st) {
  if (st->done) return;
  __shared__ double sh[4];
  double z0 = 0.0, z1 = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) {
    z0 += zpart[2 * i];
    z1 += zpart[2 * i + 1];
  }
  const double znom = block_sum_all(z0, sh), zden = block_sum_all(z1, sh);
  // s = A t = 0 (both dots zero): the half step u += alpha p, r = t is taken
  // with zeta = 0, and k_final's beta stage reports convergence or breakdown
  const bool half = zden == 0.0 && znom == 0.0;
  const double zeta = half ? 0.0 : znom / zden;
  const bool bad = !half && (zden == 0.0 || !finite(zeta));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->d0 = znom;
    st->d1 = zden;
    st->zeta = zeta;
    if (bad) {
      st->status = finite(zden) ? HMV_SOLVE_BREAKDOWN : HMV_SOLVE_NONFINITE;
      st->done = 1;
    }
  }
  if (bad) return;
  const double alpha = st->alpha;
  double s0 = 0.0, s1 = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    u[i] = (u[i] + alpha * p[i]) + zeta * t[i];
    const double ri = t[i] - zeta * s[i];
    r[i] = ri;
    s0 = __builtin_fma(rhat[i], ri, s0);
    s1 = __builtin_fma(ri, ri, s1);
  }
  double r0 = block_sum(s0, sh);
  double r1 = block_sum(s1, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0;
    part[2 * blockIdx.x + 1] = r1;
  }
}

// r = b - Ab (in place: y holds A u on entry), rhat = r, p = v = 0
__global__ __launch_bounds__(kThreads) void k_init(int n, const double *__restrict__ b,
                                                    double *__restrict__ r,
                                                    double *__restrict__ rhat,
                                                    double *__restrict__ p,
                                                    double *__restrict__ v, int zero_guess) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const double ri = zero_guess ? b[i] : b[i] - r[i];  // r holds A u unless u = 0
    r[i] = ri;
    rhat[i] = ri;
    p[i] = 0.0;
    v[i] = 0.0;
  }
}

// r = b - r (r holds A u), used for true residuals
__global__ __launch_bounds__(kThreads) void k_residual(int n, const double *__restrict__ b,
                                                        double *__restrict__ r) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    r[i] = b[i] - r[i];
}

// u += d ; d = 0
__global__ __launch_bounds__(kThreads) void k_add_clear(int n, double *__restrict__ u,
                                                         double *__restrict__ d) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    u[i] += d[i];
    d[i] = 0.0;
  }
}

__global__ void k_fill(int n, double *__restrict__ x, double v) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) x[i] = v;
}

int grid_for(int n) {
  int g = (n + kThreads - 1) / kThreads;
  if (g > kMaxBlocks) g = kMaxBlocks;
  return g < 1 ? 1 : g;
}

}  // namespace

struct HmvSolveWork {
  int n = 0;
  hipStream_t stream = nullptr;
  double *r = nullptr, *rhat = nullptr, *p = nullptr, *v = nullptr, *t = nullptr, *s = nullptr;
  double *x2 = nullptr, *r2 = nullptr;  // refinement: correction, outer residual
  double *part = nullptr;
  double *hist = nullptr;
  int hist_cap = 0;
  SolveState *st = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  const double *g_b = nullptr;  // vectors the graph was captured with
  double *g_u = nullptr;
  hmv_arena_s *g_arena = nullptr;
  double *g_hist = nullptr;
};

namespace hmv {

void free_solve_work(hmv_arena_s *a) {
  HmvSolveWork *w = a->solve;
  if (!w) return;
  if (w->exec) (void)hipGraphExecDestroy(w->exec);
  if (w->graph) (void)hipGraphDestroy(w->graph);
  if (w->stream) (void)hipStreamDestroy(w->stream);
  (void)hipFree(w->r);
  (void)hipFree(w->rhat);
  (void)hipFree(w->p);
  (void)hipFree(w->v);
  (void)hipFree(w->t);
  (void)hipFree(w->s);
  (void)hipFree(w->x2);
  (void)hipFree(w->r2);
  (void)hipFree(w->part);
  (void)hipFree(w->hist);
  (void)hipFree(w->st);
  delete w;
  a->solve = nullptr;
}

}  // namespace hmv

namespace {

int ensure_work(hmv_arena_s *a, int hist_need) {
  HmvSolveWork *w = a->solve;
  hipError_t e;
  if (!w) {
    w = new HmvSolveWork();
    a->solve = w;
    w->n = a->nd;
    const size_t nb = sizeof(double) * (size_t)a->nd;
    if ((e = hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&w->r, nb)) != hipSuccess || (e = hipMalloc(&w->rhat, nb)) != hipSuccess ||
        (e = hipMalloc(&w->p, nb)) != hipSuccess || (e = hipMalloc(&w->v, nb)) != hipSuccess ||
        (e = hipMalloc(&w->t, nb)) != hipSuccess || (e = hipMalloc(&w->s, nb)) != hipSuccess ||
        (e = hipMalloc(&w->x2, nb)) != hipSuccess || (e = hipMalloc(&w->r2, nb)) != hipSuccess ||
        (e = hipMalloc(&w->part, sizeof(double) * 4 * kMaxBlocks)) != hipSuccess ||
        (e = hipMalloc(&w->st, sizeof(SolveState))) != hipSuccess) {
      hmv::free_solve_work(a);
      return hmv::hip_fail(e, "solver workspace");
    }
  }
  if (hist_need > w->hist_cap) {
    (void)hipFree(w->hist);
    w->hist = nullptr;
    w->hist_cap = 0;
    if ((e = hipMalloc(&w->hist, sizeof(double) * hist_need)) != hipSuccess)
      return hmv::hip_fail(e, "solver history");
    w->hist_cap = hist_need;
    if (w->exec) {
      (void)hipGraphExecDestroy(w->exec);
      (void)hipGraphDestroy(w->graph);
      w->exec = nullptr;
      w->graph = nullptr;
    }
  }
  return HMV_OK;
}

// Enqueues one BiCGSTAB iteration on w->stream.
hipError_t enqueue_iteration(hmv_arena_s *a, HmvSolveWork *w, double *u) {
  const int n = w->n, g = grid_for(n);
  hipStream_t s = w->stream;
  const int *skip = &w->st->done;
  hipLaunchKernelGGL(k_update_p, dim3(g), dim3(kThreads), 0, s, n, w->r, w->v, w->p, w->st);
  hipError_t e = hmv::enqueue_apply(a, w->p, w->v, 0.0, skip, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_dots, dim3(g), dim3(kThreads), 0, s, n, w->rhat, w->v,
                     (const double *)nullptr, (const double *)nullptr, w->part, w->st);
  // alpha and zeta are formed inside k_update_t / k_update_ur; the (r^, r),
  // (r, r) partials go to the second half of the partial buffer
  double *part2 = w->part + 2 * (size_t)g;
  hipLaunchKernelGGL(k_update_t, dim3(g), dim3(kThreads), 0, s, n, g, w->r, w->v, w->t,
                     (const double *)w->part, w->st);
  if ((e = hmv::enqueue_apply(a, w->t, w->s, 0.0, skip, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_dots, dim3(g), dim3(kThreads), 0, s, n, w->s, w->t, w->s, w->s, w->part,
                     w->st);
  hipLaunchKernelGGL(k_update_ur, dim3(g), dim3(kThreads), 0, s, n, g, w->p, w->t, w->s, w->rhat,
                     u, w->r, (const double *)w->part, part2, w->st);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(kThreads), 0, s, g, (const double *)part2, w->st, 3,
                     w->hist);
  return hipGetLastError();
}

// Core BiCGSTAB; b, u device vectors.  hist_host may be null.
// zero_guess: u is known to be zero, so r = b without the product A u.
int run_bicgstab(hmv_arena_s *a, const double *b, double *u, double eps, int maxit,
                 int check_every, double *hist_host, hmv_solve_info *info, bool use_graph,
                 bool zero_guess) {
  int rc = ensure_work(a, maxit + 1);
  if (rc) return rc;
  HmvSolveWork *w = a->solve;
  const int n = w->n, g = grid_for(n);
  hipStream_t s = w->stream;
  hipError_t e;
  if (check_every < 1) check_every = 1;
  auto t0 = std::chrono::steady_clock::now();

  // r = b - A u; r^ = r; p = v = 0; rr = (r,r), bnorm = sqrt((b,b)) on device
  SolveState init{};
  init.eps = eps;
  init.maxit = maxit;
  SolveState host{};
  if ((e = hipMemcpyAsync(w->st, &init, sizeof(init), hipMemcpyHostToDevice, s)) != hipSuccess)
    return hmv::hip_fail(e, "bicgstab: state upload");
  if (!zero_guess && (e = hmv::enqueue_apply(a, u, w->r, 0.0, nullptr, s)) != hipSuccess)
    return hmv::hip_fail(e, "bicgstab: initial apply");
  hipLaunchKernelGGL(k_init, dim3(g), dim3(kThreads), 0, s, n, b, w->r, w->rhat, w->p, w->v,
                     zero_guess ? 1 : 0);
  hipLaunchKernelGGL(k_dots, dim3(g), dim3(kThreads), 0, s, n, w->r, w->r, b, b, w->part,
                     (const SolveState *)nullptr);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(kThreads), 0, s, g, w->part, w->st, 0, w->hist);
  if ((e = hipGetLastError()) != hipSuccess) return hmv::hip_fail(e, "bicgstab: setup launch");

  // graph of one iteration (re-captured when the vectors or arena change)
  if (use_graph && (w->g_u != u || w->g_arena != a || !w->exec)) {
    if (w->exec) {
      (void)hipGraphExecDestroy(w->exec);
      (void)hipGraphDestroy(w->graph);
      w->exec = nullptr;
      w->graph = nullptr;
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hmv::hip_fail(e, "bicgstab: sync");
    if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return hmv::hip_fail(e, "bicgstab: begin capture");
    hipError_t ec = enqueue_iteration(a, w, u);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(s, &graph);
    if (ec != hipSuccess) return hmv::hip_fail(ec, "bicgstab: capture");
    if (e != hipSuccess) return hmv::hip_fail(e, "bicgstab: end capture");
    w->graph = graph;
    if ((e = hipGraphInstantiate(&w->exec, graph, nullptr, nullptr, 0)) != hipSuccess)
      return hmv::hip_fail(e, "bicgstab: instantiate");
    w->g_u = u;
    w->g_arena = a;
  }

  int launched = 0;
  while (launched < maxit) {
    const int burst = std::min(check_every, maxit - launched);
    for (int k = 0; k < burst; ++k) {
      if (use_graph)
        e = hipGraphLaunch(w->exec, s);
      else
        e = enqueue_iteration(a, w, u);
      if (e != hipSuccess) return hmv::hip_fail(e, "bicgstab: iteration launch");
    }
    launched += burst;
    if ((e = hipMemcpyAsync(&host, w->st, sizeof(host), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
      return hmv::hip_fail(e, "bicgstab: state read");
    if (host.done) break;
  }
  if ((e = hipMemcpyAsync(&host, w->st, sizeof(host), hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return hmv::hip_fail(e, "bicgstab: state read");
  auto t1 = std::chrono::steady_clock::now();

  if (hist_host) {
    std::vector<double> hh(maxit + 1, NAN);
    if ((e = hipMemcpy(hh.data(), w->hist, sizeof(double) * (host.step + 1), hipMemcpyDeviceToHost)) !=
        hipSuccess)
      return hmv::hip_fail(e, "bicgstab: history");
    std::memcpy(hist_host, hh.data(), sizeof(double) * (maxit + 1));
  }
  if (info) {
    info->nstp = host.step;
    info->status = host.done ? host.status : HMV_SOLVE_MAXITER;
    info->relres = host.rnorm / host.bnorm;
    info->seconds = std::chrono::duration<double>(t1 - t0).count();
    info->outer = 0;
    info->inner_total = host.step;
  }
  return HMV_OK;
}

// ||b - A u|| / ||b|| with the arena `a` (device vectors), using w->r2.
int true_relres(hmv_arena_s *a, const double *b, const double *u, double *out) {
  HmvSolveWork *w = a->solve;
  const int n = w->n, g = grid_for(n);
  hipStream_t s = w->stream;
  hipError_t e;
  if ((e = hmv::enqueue_apply(a, u, w->r2, 0.0, nullptr, s)) != hipSuccess)
    return hmv::hip_fail(e, "true residual apply");
  hipLaunchKernelGGL(k_residual, dim3(g), dim3(kThreads), 0, s, n, b, w->r2);
  hipLaunchKernelGGL(k_dots, dim3(g), dim3(kThreads), 0, s, n, w->r2, w->r2, b, b, w->part,
                     (const SolveState *)nullptr);
  std::vector<double> part(2 * g);
  if ((e = hipMemcpyAsync(part.data(), w->part, sizeof(double) * 2 * g, hipMemcpyDeviceToHost, s)) !=
          hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return hmv::hip_fail(e, "true residual");
  double rr = 0.0, bb = 0.0;
  for (int i = 0; i < g; ++i) {
    rr += part[2 * i];
    bb += part[2 * i + 1];
  }
  *out = std::sqrt(rr) / std::sqrt(bb);
  return HMV_OK;
}

}  // namespace

extern "C" {

int hmv_bicgstab(hmv_arena a, const double *b, double *u, double eps, int maxit, int check_every,
                 double *hist, hmv_solve_info *info) {
  if (!a || !b || !u) {
    hmv::set_error("hmv_bicgstab: null argument");
    return HMV_ERR_ARG;
  }
  if (maxit < 0) {
    hmv::set_error("hmv_bicgstab: negative maxit");
    return HMV_ERR_ARG;
  }
  hmv_solve_info loc{};
  int rc = run_bicgstab(a, b, u, eps, maxit, check_every, hist, &loc, true, false);
  if (rc) return rc;
  rc = true_relres(a, b, u, &loc.true_relres);
  if (rc) return rc;
  if (info) *info = loc;
  return HMV_OK;
}

int hmv_bicgstab_ir(hmv_arena a64, hmv_arena alow, const double *b, double *u, double eps,
                    double inner_eps, int max_outer, int max_inner, int check_every, double *hist,
                    hmv_solve_info *info) {
  if (!a64 || !alow || !b || !u || a64->nd != alow->nd) {
    hmv::set_error("hmv_bicgstab_ir: invalid arguments");
    return HMV_ERR_ARG;
  }
  int rc = ensure_work(a64, 2);
  if (rc) return rc;
  rc = ensure_work(alow, max_inner + 1);
  if (rc) return rc;
  HmvSolveWork *w64 = a64->solve;
  HmvSolveWork *wl = alow->solve;
  const int n = a64->nd, g = grid_for(n);
  hipError_t e;
  auto t0 = std::chrono::steady_clock::now();
  double rel = 0.0;
  // outer residual r2 = b - A64 u (in w64->r2), relres
  rc = true_relres(a64, b, u, &rel);
  if (rc) return rc;
  if (hist) hist[0] = rel;
  int outer = 0, inner_total = 0, status = HMV_SOLVE_MAXITER;
  if (!std::isfinite(rel)) status = HMV_SOLVE_NONFINITE;
  while (status == HMV_SOLVE_MAXITER && rel >= eps && outer < max_outer) {
    // correction d (wl->x2) = 0; solve Alow d = r2 with the inner BiCGSTAB
    if ((e = hipStreamSynchronize(w64->stream)) != hipSuccess)
      return hmv::hip_fail(e, "ir: sync");
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(kThreads), 0, wl->stream, n, wl->x2, 0.0);
    hmv_solve_info in{};
    rc = run_bicgstab(alow, w64->r2, wl->x2, inner_eps, max_inner, check_every, nullptr, &in, true,
                      true);
    if (rc) return rc;
    inner_total += in.nstp;
    if ((e = hipStreamSynchronize(wl->stream)) != hipSuccess) return hmv::hip_fail(e, "ir: sync");
    hipLaunchKernelGGL(k_add_clear, dim3(g), dim3(kThreads), 0, w64->stream, n, u, wl->x2);
    ++outer;
    rc = true_relres(a64, b, u, &rel);
    if (rc) return rc;
    if (hist) hist[outer] = rel;
    if (!std::isfinite(rel)) status = HMV_SOLVE_NONFINITE;
    if (in.status == HMV_SOLVE_NONFINITE) status = HMV_SOLVE_NONFINITE;
  }
  if (status == HMV_SOLVE_MAXITER && rel < eps) status = HMV_SOLVE_CONVERGED;
  if ((e = hipStreamSynchronize(w64->stream)) != hipSuccess) return hmv::hip_fail(e, "ir: sync");
  auto t1 = std::chrono::steady_clock::now();
  if (info) {
    info->nstp = inner_total;
    info->status = status;
    info->relres = rel;
    info->true_relres = rel;
    info->seconds = std::chrono::duration<double>(t1 - t0).count();
    info->outer = outer;
    info->inner_total = inner_total;
  }
  return HMV_OK;
}

// Mixed-precision BiCGSTAB with reliable updates: the iteration runs with
// the low-precision arena `alow`; u, the vectors and the scalars stay FP64;
// whenever the iterated residual has dropped by `delta` since the last
// replacement (or the iteration believes it has converged), r = b - A64 u
// is recomputed with the FP64 arena and replaces the iterated residual
// without restarting the Krylov recurrence.
int hmv_bicgstab_ru(hmv_arena a64, hmv_arena alow, const double *b, double *u, double eps,
                    int maxit, int check_every, double delta, hmv_solve_info *info) {
  if (!a64 || !alow || !b || !u || a64->nd != alow->nd || maxit < 0 || !(delta > 0.0) ||
      delta >= 1.0) {
    hmv::set_error("hmv_bicgstab_ru: invalid arguments (0 < delta < 1)");
    return HMV_ERR_ARG;
  }
  int rc = ensure_work(alow, maxit + 1);
  if (rc) return rc;
  rc = ensure_work(a64, 2);
  if (rc) return rc;
  if (check_every < 1) check_every = 1;
  HmvSolveWork *w = alow->solve;
  const int n = w->n, g = grid_for(n);
  hipStream_t s = w->stream;
  hipError_t e;
  auto t0 = std::chrono::steady_clock::now();
  SolveState init{};
  init.eps = eps;
  init.maxit = maxit;
  SolveState host{};
  // r = b - A64 u, r^ = r, p = v = 0
  if ((e = hipMemcpyAsync(w->st, &init, sizeof(init), hipMemcpyHostToDevice, s)) != hipSuccess ||
      (e = hmv::enqueue_apply(a64, u, w->r, 0.0, nullptr, s)) != hipSuccess)
    return hmv::hip_fail(e, "bicgstab_ru: setup");
  hipLaunchKernelGGL(k_init, dim3(g), dim3(kThreads), 0, s, n, b, w->r, w->rhat, w->p, w->v, 0);
  hipLaunchKernelGGL(k_dots, dim3(g), dim3(kThreads), 0, s, n, w->r, w->r, b, b, w->part,
                     (const SolveState *)nullptr);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(kThreads), 0, s, g, w->part, w->st, 0, w->hist);
  if ((e = hipGetLastError()) != hipSuccess) return hmv::hip_fail(e, "bicgstab_ru: setup");
  if (w->g_u != u || w->g_arena != alow || !w->exec) {
    if (w->exec) {
      (void)hipGraphExecDestroy(w->exec);
      (void)hipGraphDestroy(w->graph);
      w->exec = nullptr;
      w->graph = nullptr;
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess ||
        (e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return hmv::hip_fail(e, "bicgstab_ru: capture");
    hipError_t ec = enqueue_iteration(alow, w, u);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(s, &graph);
    if (ec != hipSuccess) return hmv::hip_fail(ec, "bicgstab_ru: capture");
    if (e != hipSuccess) return hmv::hip_fail(e, "bicgstab_ru: end capture");
    w->graph = graph;
    if ((e = hipGraphInstantiate(&w->exec, graph, nullptr, nullptr, 0)) != hipSuccess)
      return hmv::hip_fail(e, "bicgstab_ru: instantiate");
    w->g_u = u;
    w->g_arena = alow;
  }
  auto read_state = [&]() -> hipError_t {
    hipError_t re = hipMemcpyAsync(&host, w->st, sizeof(host), hipMemcpyDeviceToHost, s);
    return re == hipSuccess ? hipStreamSynchronize(s) : re;
  };
  if ((e = read_state()) != hipSuccess) return hmv::hip_fail(e, "bicgstab_ru: state");
  double r_last = host.rnorm;
  int updates = 0;
  while (!host.done) {  // done: converged on the true residual, or stopped
    const int burst = std::min(check_every, std::max(1, maxit - host.step));
    for (int k = 0; k < burst; ++k)
      if ((e = hipGraphLaunch(w->exec, s)) != hipSuccess)
        return hmv::hip_fail(e, "bicgstab_ru: iteration launch");
    if ((e = read_state()) != hipSuccess) return hmv::hip_fail(e, "bicgstab_ru: state");
    const bool sloppy_done = host.done && host.status == HMV_SOLVE_CONVERGED;
    if (host.done && !sloppy_done) break;  // breakdown, non-finite or out of steps
    if (sloppy_done || host.rnorm < delta * r_last) {
      // reliable update: r = b - A64 u replaces the iterated residual
      if ((e = hmv::enqueue_apply(a64, u, w->r, 0.0, nullptr, s)) != hipSuccess)
        return hmv::hip_fail(e, "bicgstab_ru: replacement");
      hipLaunchKernelGGL(k_residual, dim3(g), dim3(kThreads), 0, s, n, b, w->r);
      hipLaunchKernelGGL(k_dots, dim3(g), dim3(kThreads), 0, s, n, w->rhat, w->r, w->r, w->r,
                         w->part, (const SolveState *)nullptr);
      hipLaunchKernelGGL(k_final, dim3(1), dim3(kThreads), 0, s, g, w->part, w->st, 4, w->hist);
      if ((e = read_state()) != hipSuccess) return hmv::hip_fail(e, "bicgstab_ru: state");
      ++updates;
      r_last = host.rnorm;
      if (host.done) break;
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  hmv_solve_info loc{};
  loc.nstp = host.step;
  loc.status = host.done ? host.status : HMV_SOLVE_MAXITER;
  loc.relres = host.rnorm / host.bnorm;
  loc.seconds = std::chrono::duration<double>(t1 - t0).count();
  loc.outer = updates;
  loc.inner_total = host.step;
  rc = true_relres(a64, b, u, &loc.true_relres);
  if (rc) return rc;
  if (loc.status == HMV_SOLVE_CONVERGED && !(loc.true_relres < eps)) loc.status = HMV_SOLVE_MAXITER;
  if (info) *info = loc;
  return HMV_OK;
}

}  // extern "C"
