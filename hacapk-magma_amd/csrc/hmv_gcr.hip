// hmv_gcr.hip -- GCR(m) on the H-matvec engine, all vectors and scalars on
// the device.
//
// Restates HACApK_gcrm_lfmtx (m_HACApK_solve.f90:506-571) step for step:
//   r = b - A u;  p_1 = r;  ap_1 = A p_1
//   step in (slot ik = (in-1) mod m):
//     alpha = (ap_ik, r) / (ap_ik, ap_ik);  capap_ik = (ap_ik, ap_ik)
//     u += alpha p_ik;  r -= alpha ap_ik;  stop on ||r||/||b|| < eps
//     ar = A r;  new slot ikn = in mod m:  p_ikn = r, ap_ikn = ar, then for
//     il = 1..ik:  beta_il = -(ap_il, ar) / capap_il;
//                  p_ikn += beta_il p_il;  ap_ikn += beta_il ap_il
// When ikn = 1 (in a multiple of m) slot 1 is among 1..ik and was just
// overwritten, so its beta is -(ar, ar)/capap_1 and it scales the new
// direction by (1 + beta_1), as in the reference.  All betas use the
// unmodified ar, so one pass computes them and one fused pass applies them.
// One iteration is captured as a hipGraph; a device flag set on convergence
// turns the remaining launched iterations into no-ops.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "hmv.h"
#include "hmv_engine.h"

namespace {

constexpr int kT = 256;
constexpr int kMaxB = 512;
constexpr int kMaxM = HMV_GCR_MAX_M;
static_assert(kMaxM >= 3, "the partial buffer holds the alpha pairs and the (r, r) partials");

struct GcrState {
  double alpha, bnorm, rnorm, eps;
  double capap[kMaxM];
  double beta[kMaxM];
  int step;  // completed steps (the reference's in)
  int ik;    // slot of the current direction (0-based)
  int ikn;   // slot of the next direction (0-based)
  int m, maxit, done, status, pad;
};

__device__ inline double blk_sum(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

// part[2*blk + {0,1}]: (a,b), (c,d) partials
__global__ __launch_bounds__(kT) void g_dots2(int n, const double *__restrict__ a,
                                              const double *__restrict__ b,
                                              const double *__restrict__ c,
                                              const double *__restrict__ d,
                                              double *__restrict__ part) {
  __shared__ double sh[4];
  double s0 = 0.0, s1 = 0.0;
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    s0 = __builtin_fma(a[i], b[i], s0);
    s1 = __builtin_fma(c[i], d[i], s1);
  }
  const double r0 = blk_sum(s0, sh), r1 = blk_sum(s1, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0;
    part[2 * blockIdx.x + 1] = r1;
  }
}

// (ap_ik, r), (ap_ik, ap_ik) for the slot ik = st->ikn of the first step
// (at setup; g_update_dir produces them for the later steps)
__global__ __launch_bounds__(kT) void g_dots_alpha(int n, const double *__restrict__ ap,
                                                   const double *__restrict__ r,
                                                   double *__restrict__ part, const GcrState *st) {
  if (st->done) return;
  __shared__ double sh[4];
  const double *q = ap + (size_t)st->ikn * n;  // the slot of this step
  double s0 = 0.0, s1 = 0.0;
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    s0 = __builtin_fma(q[i], r[i], s0);
    s1 = __builtin_fma(q[i], q[i], s1);
  }
  const double r0 = blk_sum(s0, sh), r1 = blk_sum(s1, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0;
    part[2 * blockIdx.x + 1] = r1;
  }
}

// alpha = (ap_ik, r) / (ap_ik, ap_ik) from the nb partial pairs of
// g_dots_alpha (every block forms it alike; block 0 records it, the slot
// and capap_ik, and flags a breakdown), then u += alpha p_ik;
// r -= alpha ap_ik;  (r, r) partials into rpart (apart is still being read
// by other blocks).
__global__ __launch_bounds__(kT) void g_update_ur(int n, int nb, const double *__restrict__ p,
                                                  const double *__restrict__ ap,
                                                  double *__restrict__ u, double *__restrict__ r,
                                                  const double *__restrict__ apart,
                                                  double *__restrict__ rpart, GcrState *st) {
  if (st->done) return;
  __shared__ double sh[4];
  double s0 = 0.0, s1 = 0.0;
  for (int b = threadIdx.x; b < nb; b += kT) {
    s0 += apart[2 * b];
    s1 += apart[2 * b + 1];
  }
  const double d0 = blk_sum(s0, sh), d1 = blk_sum(s1, sh);
  const int ik = st->ikn;  // the slot of this step
  const double alpha = d0 / d1;
  const bool bad = d1 == 0.0 || !isfinite(alpha);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->ik = ik;
    st->capap[ik] = d1;
    st->alpha = alpha;
    if (bad) {
      st->status = isfinite(d1) ? HMV_SOLVE_BREAKDOWN : HMV_SOLVE_NONFINITE;
      st->done = 1;
    }
  }
  if (bad) return;
  const double *pk = p + (size_t)ik * n, *q = ap + (size_t)ik * n;
  double s2 = 0.0;
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    u[i] = u[i] + alpha * pk[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    s2 = __builtin_fma(ri, ri, s2);
  }
  const double r0 = blk_sum(s2, sh);
  if (threadIdx.x == 0) rpart[blockIdx.x] = r0;
}

// partials of (v_il, ar) for il = 0..ik, v_il = ar for the slot just
// overwritten (il == ikn), else ap_il:  part[blk * kMaxM + il].  Block row
// blockIdx.y covers the kGroup slots from kGroup*blockIdx.y, so ar is read
// once per group.
constexpr int kGroup = 4;
__global__ __launch_bounds__(kT) void g_dots_beta(int n, const double *__restrict__ ap,
                                                  const double *__restrict__ ar,
                                                  double *__restrict__ part, const GcrState *st) {
  if (st->done) return;
  const int ik = st->ik, ikn = st->ikn, l0 = blockIdx.y * kGroup;
  if (l0 > ik) return;
  __shared__ double sh[4];
  const double *v[kGroup];
#pragma unroll
  for (int j = 0; j < kGroup; ++j) {
    const int l = min(l0 + j, ik);
    v[j] = l == ikn ? ar : ap + (size_t)l * n;
  }
  double s[kGroup] = {0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    const double a = ar[i];
#pragma unroll
    for (int j = 0; j < kGroup; ++j) s[j] = __builtin_fma(v[j][i], a, s[j]);
  }
#pragma unroll
  for (int j = 0; j < kGroup; ++j) {
    const double r0 = blk_sum(s[j], sh);
    if (threadIdx.x == 0 && l0 + j <= ik) part[(size_t)blockIdx.x * kMaxM + l0 + j] = r0;
  }
}

// stage 0: setup (d0 = (r,r), d1 = (b,b) from nb partial pairs); 2: ||r||
// (nb partials), step, stop; 3: betas from the per-slot partials, one wave
// per slot (lanes sum the blocks in a fixed order, then a fixed shuffle
// tree: deterministic)
__global__ __launch_bounds__(kT) void g_final(int nb, const double *__restrict__ part,
                                              GcrState *st, int stage, double *__restrict__ hist) {
  if (st->done && stage != 0) return;
  __shared__ double sh[4];
  if (stage == 3) {
    // wave w takes slots w, w + 4, ...; all their loads go out together
    constexpr int kW = kT / 64, kL = (kMaxM + kW - 1) / kW;
    const int ik = st->ik, lane = threadIdx.x & 63, w0 = threadIdx.x >> 6;
    double s[kL];
#pragma unroll
    for (int j = 0; j < kL; ++j) s[j] = 0.0;
    for (int b = lane; b < nb; b += 64)
#pragma unroll
      for (int j = 0; j < kL; ++j)
        if (w0 + j * kW <= ik) s[j] += part[(size_t)b * kMaxM + w0 + j * kW];
#pragma unroll
    for (int j = 0; j < kL; ++j) {
      const int l = w0 + j * kW;
      if (l > ik) break;
      double v = s[j];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) st->beta[l] = -v / st->capap[l];
    }
    return;
  }
  double s0 = 0.0, s1 = 0.0;
  if (stage == 0) {
    for (int b = threadIdx.x; b < nb; b += kT) {
      s0 += part[2 * b];
      s1 += part[2 * b + 1];
    }
  } else {
    for (int b = threadIdx.x; b < nb; b += kT) s0 += part[b];
  }
  const double d0 = blk_sum(s0, sh), d1 = blk_sum(s1, sh);
  if (threadIdx.x != 0) return;
  if (stage == 0) {
    st->rnorm = sqrt(d0);
    st->bnorm = sqrt(d1);
    st->step = 0;
    st->ik = 0;
    st->ikn = 0;
    st->done = 0;
    st->status = HMV_SOLVE_MAXITER;
    if (hist) hist[0] = st->rnorm / st->bnorm;
    if (!isfinite(st->rnorm) || !isfinite(st->bnorm) || st->bnorm == 0.0) {
      st->status = HMV_SOLVE_NONFINITE;
      st->done = 1;
    } else if (st->rnorm / st->bnorm < st->eps) {
      st->status = HMV_SOLVE_CONVERGED;
      st->done = 1;
    } else if (st->maxit <= 0) {
      st->done = 1;
    }
  } else {
    st->step += 1;
    st->rnorm = sqrt(d0);
    st->ikn = st->step % st->m;
    if (hist) hist[st->step] = st->rnorm / st->bnorm;
    if (!isfinite(st->rnorm)) {
      st->status = HMV_SOLVE_NONFINITE;
      st->done = 1;
    } else if (st->rnorm / st->bnorm < st->eps) {
      st->status = HMV_SOLVE_CONVERGED;
      st->done = 1;
    } else if (st->step >= st->maxit) {
      st->status = HMV_SOLVE_MAXITER;
      st->done = 1;
    }
  }
}

// p_ikn = c0 r + sum_{il <= ik, il != ikn} beta_il p_il;  ap_ikn likewise
// with ar;  c0 = 1 + beta_ikn when ikn <= ik (slot 1 at a restart), else 1
// ... and the next step's alpha partials (ap_ikn, r), (ap_ikn, ap_ikn) into
// part[2*blk + {0,1}], in g_dots_alpha's order (same grid, same per-thread
// sequence), so that the step needs no separate dot pass.
__global__ __launch_bounds__(kT) void g_update_dir(int n, const double *__restrict__ r,
                                                   const double *__restrict__ ar,
                                                   double *__restrict__ p, double *__restrict__ ap,
                                                   double *__restrict__ part, const GcrState *st) {
  if (st->done) return;
  __shared__ double sh[4];
  const int ik = st->ik, ikn = st->ikn;
  const double c0 = ikn <= ik ? 1.0 + st->beta[ikn] : 1.0;
  double *pn = p + (size_t)ikn * n, *apn = ap + (size_t)ikn * n;
  double s0 = 0.0, s1 = 0.0;
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    const double ri = r[i];
    double pv = c0 * ri, av = c0 * ar[i];
    for (int l = 0; l <= ik; ++l) {
      if (l == ikn) continue;
      const double bl = st->beta[l];
      pv = __builtin_fma(bl, p[(size_t)l * n + i], pv);
      av = __builtin_fma(bl, ap[(size_t)l * n + i], av);
    }
    pn[i] = pv;
    apn[i] = av;
    s0 = __builtin_fma(av, ri, s0);
    s1 = __builtin_fma(av, av, s1);
  }
  const double r0 = blk_sum(s0, sh), r1 = blk_sum(s1, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0;
    part[2 * blockIdx.x + 1] = r1;
  }
}

// r = b - ar (r = b when ar is null: a zero initial guess);  p_1 = r
__global__ __launch_bounds__(kT) void g_init(int n, const double *__restrict__ b,
                                             const double *__restrict__ ar,
                                             double *__restrict__ r, double *__restrict__ p) {
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    const double ri = ar ? b[i] - ar[i] : b[i];
    r[i] = ri;
    p[i] = ri;
  }
}

int grid_of(int n) {
  return std::max(1, std::min((n + kT - 1) / kT, kMaxB));
}

}  // namespace

// GCR workspace, kept with the arena between calls (the iteration graph is
// re-captured only when m, the history capacity or the solution vector change).
struct HmvGcrWork {
  int m = 0;
  size_t hist_cap = 0;
  hipStream_t s = nullptr;
  double *r = nullptr, *ar = nullptr, *p = nullptr, *ap = nullptr, *part = nullptr, *hist = nullptr;
  GcrState *st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double *g_u = nullptr;  // solution vector the graph was captured with
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    g_u = nullptr;
  }
  ~HmvGcrWork() {
    drop_graph();
    if (s) (void)hipStreamDestroy(s);
    for (double *q : {r, ar, p, ap, part, hist}) (void)hipFree(q);
    (void)hipFree(st);
  }
};

namespace hmv {
void free_gcr_work(hmv_arena_s *a) {
  delete a->gcr;
  a->gcr = nullptr;
}
}  // namespace hmv

namespace {

using GcrWork = HmvGcrWork;

// The arena's GCR workspace for m directions and maxit steps.
int gcr_work(hmv_arena_s *a, int m, int maxit, GcrWork *&out) {
  GcrWork *w = a->gcr;
  if (w && (w->m != m || w->hist_cap < (size_t)maxit + 1)) {
    hmv::free_gcr_work(a);
    w = nullptr;
  }
  if (!w) {
    w = new GcrWork();
    a->gcr = w;
    const int n = a->nd, g = grid_of(n);
    const size_t nb = sizeof(double) * (size_t)n;
    w->m = m;
    w->hist_cap = (size_t)maxit + 1;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&w->s, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&w->r, nb)) != hipSuccess || (e = hipMalloc(&w->ar, nb)) != hipSuccess ||
        (e = hipMalloc(&w->p, nb * m)) != hipSuccess || (e = hipMalloc(&w->ap, nb * m)) != hipSuccess ||
        (e = hipMalloc(&w->part, sizeof(double) * (size_t)g * kMaxM)) != hipSuccess ||
        (e = hipMalloc(&w->hist, sizeof(double) * w->hist_cap)) != hipSuccess ||
        (e = hipMalloc(&w->st, sizeof(GcrState))) != hipSuccess) {
      hmv::free_gcr_work(a);
      return hmv::hip_fail(e, "hmv_gcrm: workspace");
    }
  }
  out = w;
  return HMV_OK;
}

hipError_t gcr_iteration(hmv_arena_s *a, GcrWork &w, int n, int m, double *u) {
  const int g = grid_of(n);
  const int *skip = &w.st->done;
  // (r, r) partials at [2g, 3g) of the g * kMaxM partials, apart from the
  // alpha pairs at [0, 2g) that g_update_ur reads while writing them
  double *rpart = w.part + 2 * (size_t)g;
  // the alpha partials of this step come from the previous step's
  // g_update_dir (or, for the first step, from the setup's g_dots_alpha)
  hipLaunchKernelGGL(g_update_ur, dim3(g), dim3(kT), 0, w.s, n, g, w.p, w.ap, u, w.r,
                     (const double *)w.part, rpart, w.st);
  hipLaunchKernelGGL(g_final, dim3(1), dim3(kT), 0, w.s, g, (const double *)rpart, w.st, 2, w.hist);
  hipError_t e = hmv::enqueue_apply(a, w.r, w.ar, 0.0, skip, w.s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(g_dots_beta, dim3(g, (m + kGroup - 1) / kGroup), dim3(kT), 0, w.s, n, w.ap,
                     w.ar, w.part,
                     (const GcrState *)w.st);
  hipLaunchKernelGGL(g_final, dim3(1), dim3(kT), 0, w.s, g, w.part, w.st, 3, w.hist);
  hipLaunchKernelGGL(g_update_dir, dim3(g), dim3(kT), 0, w.s, n, w.r, w.ar, w.p, w.ap, w.part,
                     (const GcrState *)w.st);
  return hipGetLastError();
}

}  // namespace

namespace {
int gcrm_run(hmv_arena a, const double *b, double *u, double eps, int maxit, int m,
             int check_every, double *hist, hmv_solve_info *info, bool true_res,
             bool zero_guess = false);
}

extern "C" int hmv_gcrm(hmv_arena a, const double *b, double *u, double eps, int maxit, int m,
                        int check_every, double *hist, hmv_solve_info *info) {
  return gcrm_run(a, b, u, eps, maxit, m, check_every, hist, info, true);
}

namespace {
// true_res: recompute ||b - A u|| / ||b|| at the end (one more product).
// zero_guess: u is known to be zero, so r = b without the product A u.
int gcrm_run(hmv_arena a, const double *b, double *u, double eps, int maxit, int m,
             int check_every, double *hist, hmv_solve_info *info, bool true_res,
             bool zero_guess) {
  if (!a || !b || !u || maxit < 0 || m < 1 || m > kMaxM) {
    hmv::set_error("hmv_gcrm: invalid argument (1 <= m <= 32)");
    return HMV_ERR_ARG;
  }
  if (check_every < 1) check_every = 1;
  const int n = a->nd, g = grid_of(n);
  GcrWork *wp = nullptr;
  int rc = gcr_work(a, m, maxit, wp);
  if (rc) return rc;
  GcrWork &w = *wp;
  hipError_t e;
  auto t0 = std::chrono::steady_clock::now();
  GcrState init{};
  init.eps = eps;
  init.m = m;
  init.maxit = maxit;
  if ((e = hipMemcpyAsync(w.st, &init, sizeof(init), hipMemcpyHostToDevice, w.s)) != hipSuccess ||
      (!zero_guess && (e = hmv::enqueue_apply(a, u, w.ar, 0.0, nullptr, w.s)) != hipSuccess))
    return hmv::hip_fail(e, "hmv_gcrm: setup");
  hipLaunchKernelGGL(g_init, dim3(g), dim3(kT), 0, w.s, n, b, zero_guess ? nullptr : w.ar, w.r,
                     w.p);
  hipLaunchKernelGGL(g_dots2, dim3(g), dim3(kT), 0, w.s, n, w.r, w.r, b, b, w.part);
  hipLaunchKernelGGL(g_final, dim3(1), dim3(kT), 0, w.s, g, w.part, w.st, 0, w.hist);
  // ap_1 = A p_1 (skipped when already converged) and the first step's
  // alpha partials
  if ((e = hmv::enqueue_apply(a, w.p, w.ap, 0.0, &w.st->done, w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_gcrm: setup");
  hipLaunchKernelGGL(g_dots_alpha, dim3(g), dim3(kT), 0, w.s, n, w.ap, w.r, w.part,
                     (const GcrState *)w.st);
  if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_gcrm: setup");
  if (!w.exec || w.g_u != u) {
    w.drop_graph();
    if ((e = hipStreamBeginCapture(w.s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_gcrm: capture");
    hipError_t ec = gcr_iteration(a, w, n, m, u);
    e = hipStreamEndCapture(w.s, &w.graph);
    if (ec != hipSuccess) return hmv::hip_fail(ec, "hmv_gcrm: capture");
    if (e != hipSuccess ||
        (e = hipGraphInstantiate(&w.exec, w.graph, nullptr, nullptr, 0)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_gcrm: graph");
    w.g_u = u;
  }
  GcrState host{};
  int launched = 0;
  for (;;) {
    if ((e = hipMemcpyAsync(&host, w.st, sizeof(host), hipMemcpyDeviceToHost, w.s)) != hipSuccess ||
        (e = hipStreamSynchronize(w.s)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_gcrm: state");
    if (host.done || launched >= maxit) break;
    const int burst = std::min(check_every, maxit - launched);
    for (int i = 0; i < burst; ++i)
      if ((e = hipGraphLaunch(w.exec, w.s)) != hipSuccess)
        return hmv::hip_fail(e, "hmv_gcrm: launch");
    launched += burst;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (hist) {
    std::fill(hist, hist + maxit + 1, NAN);
    if ((e = hipMemcpy(hist, w.hist, sizeof(double) * ((size_t)host.step + 1),
                       hipMemcpyDeviceToHost)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_gcrm: history");
  }
  // true residual ||b - A u|| / ||b||
  double tr = NAN;
  if (true_res) {
    if ((e = hmv::enqueue_apply(a, u, w.ar, 0.0, nullptr, w.s)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_gcrm: true residual");
    hipLaunchKernelGGL(g_init, dim3(g), dim3(kT), 0, w.s, n, b, w.ar, w.r, w.p);
    hipLaunchKernelGGL(g_dots2, dim3(g), dim3(kT), 0, w.s, n, w.r, w.r, b, b, w.part);
    std::vector<double> part(2 * (size_t)g);
    if ((e = hipMemcpyAsync(part.data(), w.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                            w.s)) != hipSuccess ||
        (e = hipStreamSynchronize(w.s)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_gcrm: true residual");
    double rr = 0.0, bb = 0.0;
    for (int i = 0; i < g; ++i) {
      rr += part[2 * i];
      bb += part[2 * i + 1];
    }
    tr = std::sqrt(rr) / std::sqrt(bb);
  }
  if (info) {
    info->nstp = host.step;
    info->status = host.done ? host.status : HMV_SOLVE_MAXITER;
    info->relres = host.rnorm / host.bnorm;
    info->true_relres = tr;
    info->seconds = std::chrono::duration<double>(t1 - t0).count();
    info->outer = 0;
    info->inner_total = host.step;
  }
  return HMV_OK;
}
}  // namespace

namespace {

// r = b - v
__global__ __launch_bounds__(kT) void g_residual(int n, const double *__restrict__ b,
                                                 const double *__restrict__ v,
                                                 double *__restrict__ r) {
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) r[i] = b[i] - v[i];
}

// u += d;  d = 0
__global__ __launch_bounds__(kT) void g_add_clear(int n, double *__restrict__ u,
                                                  double *__restrict__ d) {
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    u[i] = u[i] + d[i];
    d[i] = 0.0;
  }
}

struct IrBufs {
  double *r = nullptr, *d = nullptr, *v = nullptr, *part = nullptr;
  hipStream_t s = nullptr;
  ~IrBufs() {
    for (double *q : {r, d, v, part}) (void)hipFree(q);
    if (s) (void)hipStreamDestroy(s);
  }
};

// r = b - A u and ||r|| / ||b|| (A: the arena a)
int residual(hmv_arena_s *a, IrBufs &w, const double *b, const double *u, double *rel) {
  const int n = a->nd, g = grid_of(n);
  hipError_t e;
  if ((e = hmv::enqueue_apply(a, u, w.v, 0.0, nullptr, w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_gcr_ir: residual");
  hipLaunchKernelGGL(g_residual, dim3(g), dim3(kT), 0, w.s, n, b, w.v, w.r);
  hipLaunchKernelGGL(g_dots2, dim3(g), dim3(kT), 0, w.s, n, w.r, w.r, b, b, w.part);
  std::vector<double> part(2 * (size_t)g);
  if ((e = hipMemcpyAsync(part.data(), w.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                          w.s)) != hipSuccess ||
      (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_gcr_ir: residual");
  double rr = 0.0, bb = 0.0;
  for (int i = 0; i < g; ++i) {
    rr += part[2 * i];
    bb += part[2 * i + 1];
  }
  *rel = std::sqrt(rr) / std::sqrt(bb);
  return HMV_OK;
}

}  // namespace

extern "C" int hmv_gcr_ir(hmv_arena a64, hmv_arena alow, const double *b, double *u, double eps,
                          double inner_eps, int m, int max_outer, int max_inner, int check_every,
                          double *hist, hmv_solve_info *info) {
  if (!a64 || !alow || !b || !u || a64->nd != alow->nd || m < 1 || m > kMaxM || max_outer < 0 ||
      max_inner < 0) {
    hmv::set_error("hmv_gcr_ir: invalid arguments");
    return HMV_ERR_ARG;
  }
  const int n = a64->nd, g = grid_of(n);
  const size_t nb = sizeof(double) * (size_t)n;
  IrBufs w;
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&w.s, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc(&w.r, nb)) != hipSuccess || (e = hipMalloc(&w.d, nb)) != hipSuccess ||
      (e = hipMalloc(&w.v, nb)) != hipSuccess ||
      (e = hipMalloc(&w.part, sizeof(double) * 2 * (size_t)g)) != hipSuccess ||
      (e = hipMemsetAsync(w.d, 0, nb, w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_gcr_ir: workspace");
  auto t0 = std::chrono::steady_clock::now();
  double rel = 0.0;
  int rc = residual(a64, w, b, u, &rel);
  if (rc) return rc;
  if (hist) hist[0] = rel;
  int outer = 0, inner_total = 0, status = HMV_SOLVE_MAXITER;
  if (!std::isfinite(rel)) status = HMV_SOLVE_NONFINITE;
  while (status == HMV_SOLVE_MAXITER && rel >= eps && outer < max_outer) {
    // solve Alow d = r to inner_eps with GCR(m), d starting at zero
    hmv_solve_info in{};
    rc = gcrm_run(alow, w.r, w.d, inner_eps, max_inner, m, check_every, nullptr, &in, false, true);
    if (rc) return rc;
    inner_total += in.nstp;
    if (in.status == HMV_SOLVE_NONFINITE) status = HMV_SOLVE_NONFINITE;
    hipLaunchKernelGGL(g_add_clear, dim3(g), dim3(kT), 0, w.s, n, u, w.d);
    ++outer;
    rc = residual(a64, w, b, u, &rel);
    if (rc) return rc;
    if (hist) hist[outer] = rel;
    if (!std::isfinite(rel)) status = HMV_SOLVE_NONFINITE;
  }
  if (status == HMV_SOLVE_MAXITER && rel < eps) status = HMV_SOLVE_CONVERGED;
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
