// hmv_pipe.hip -- pipelined BiCGSTAB on the H-matvec engine, all vectors and
// scalars on the device.
//
// Restates c_hacapk_bicgstab_cax_lfmtx_pipe_ (HACApK_SOLVER_PIPE.c:144-637):
//   v = A u;  r = b - v;  r0 = r;  w = A r;  t = A w
//   alpha = (r,r) / (w,r)
//   step:  p = r + beta (p - zeta s);  s = w + beta (s - zeta z);
//          z = t + beta (z - zeta v)        (first step: p, s, z = r, w, t)
//          q = r - alpha s;  y = w - alpha z;  zeta = (y,q) / (y,y)
//          v = A z
//          x += alpha p + zeta q;  r = q - zeta y;  w = y - zeta (t - alpha v)
//          beta = (alpha/zeta) (r,r0)/(r,r0)_old
//          alpha = (r,r0) / ((w,r0) + beta (s,r0) - beta zeta (z,r0))
//          ||r||;  t = A w
// The reference accumulates the iterate in a vector it never initialises
// nor returns (it hands back the initial guess, :188,483-484,595); here the
// iterate starts at u and is returned in u.  Each step is two fused vector
// kernels (with their dot products), two one-block scalar kernels and two
// H-matvecs, captured as one hipGraph; a device flag set on convergence turns
// the remaining launched steps into no-ops.
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
constexpr int kND = 5;  // dot products per partial record

struct PipeState {
  double alpha, beta, zeta, znorm, bnorm, rnorm, eps;
  int step, maxit, done, status;
};

__device__ inline double blk_sum(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

template <int N>
__device__ inline void put_partials(const double (&s)[N], double *__restrict__ part) {
  __shared__ double sh[4];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const double r = blk_sum(s[j], sh);
    if (threadIdx.x == 0) part[blockIdx.x * kND + j] = r;
  }
}

// r = b - v;  r0 = r
__global__ __launch_bounds__(kT) void p_init(int n, const double *__restrict__ b,
                                             const double *__restrict__ v, double *__restrict__ r,
                                             double *__restrict__ r0) {
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    const double ri = b[i] - v[i];
    r[i] = ri;
    r0[i] = ri;
  }
}

// partials of (r,r), (w,r), (b,b)
__global__ __launch_bounds__(kT) void p_dots_setup(int n, const double *__restrict__ r,
                                                   const double *__restrict__ w,
                                                   const double *__restrict__ b,
                                                   double *__restrict__ part) {
  double s[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    const double ri = r[i];
    s[0] = __builtin_fma(ri, ri, s[0]);
    s[1] = __builtin_fma(w[i], ri, s[1]);
    s[2] = __builtin_fma(b[i], b[i], s[2]);
  }
  put_partials<3>(s, part);
}

// p, s, z updates; q = r - alpha s;  y = w - alpha z;  partials (y,q), (y,y)
__global__ __launch_bounds__(kT) void p_update1(int n, const double *__restrict__ r,
                                                const double *__restrict__ w,
                                                const double *__restrict__ t,
                                                const double *__restrict__ v, double *__restrict__ p,
                                                double *__restrict__ s, double *__restrict__ z,
                                                double *__restrict__ q, double *__restrict__ y,
                                                double *__restrict__ part, const PipeState *st) {
  if (st->done) return;
  const bool first = st->step == 0;
  const double alpha = st->alpha, beta = st->beta, zeta = st->zeta;
  double d[2] = {0.0, 0.0};
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    const double ri = r[i], wi = w[i], ti = t[i];
    double pi, si, zi;
    if (first) {
      pi = ri;
      si = wi;
      zi = ti;
    } else {
      pi = ri + beta * (p[i] - zeta * s[i]);
      si = wi + beta * (s[i] - zeta * z[i]);
      zi = ti + beta * (z[i] - zeta * v[i]);
    }
    p[i] = pi;
    s[i] = si;
    z[i] = zi;
    const double qi = ri - alpha * si, yi = wi - alpha * zi;
    q[i] = qi;
    y[i] = yi;
    d[0] = __builtin_fma(yi, qi, d[0]);
    d[1] = __builtin_fma(yi, yi, d[1]);
  }
  put_partials<2>(d, part);
}

// x += alpha p + zeta q;  r = q - zeta y;  w = y - zeta (t - alpha v);
// partials (r,r0), (w,r0), (s,r0), (z,r0), (r,r)
__global__ __launch_bounds__(kT) void p_update2(
    int n, const double *__restrict__ p, const double *__restrict__ q, const double *__restrict__ y,
    const double *__restrict__ t, const double *__restrict__ v, const double *__restrict__ s,
    const double *__restrict__ z, const double *__restrict__ r0, double *__restrict__ x,
    double *__restrict__ r, double *__restrict__ w, double *__restrict__ part, const PipeState *st) {
  if (st->done) return;
  const double alpha = st->alpha, zeta = st->zeta;
  double d[kND] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    const double qi = q[i], yi = y[i];
    x[i] = (x[i] + alpha * p[i]) + zeta * q[i];
    const double ri = qi - zeta * yi;
    const double wi = yi + (-zeta) * (t[i] - alpha * v[i]);
    r[i] = ri;
    w[i] = wi;
    const double r0i = r0[i];
    d[0] = __builtin_fma(ri, r0i, d[0]);
    d[1] = __builtin_fma(wi, r0i, d[1]);
    d[2] = __builtin_fma(s[i], r0i, d[2]);
    d[3] = __builtin_fma(z[i], r0i, d[3]);
    d[4] = __builtin_fma(ri, ri, d[4]);
  }
  put_partials<kND>(d, part);
}

// stage 0: setup;  1: zeta;  2: beta, alpha, ||r||, step, stop
__global__ __launch_bounds__(kT) void p_final(int nb, const double *__restrict__ part, PipeState *st,
                                              int stage, double *__restrict__ hist) {
  if (st->done && stage != 0) return;
  __shared__ double sh[4];
  double d[kND];
  const int nd = stage == 0 ? 3 : stage == 1 ? 2 : kND;
  for (int j = 0; j < kND; ++j) {
    double s = 0.0;
    if (j < nd)
      for (int b = threadIdx.x; b < nb; b += kT) s += part[b * kND + j];
    d[j] = blk_sum(s, sh);
  }
  if (threadIdx.x != 0) return;
  if (stage == 0) {
    st->znorm = d[0];
    st->alpha = d[0] / d[1];
    st->beta = 0.0;
    st->zeta = 0.0;
    st->rnorm = sqrt(d[0]);
    st->bnorm = sqrt(d[2]);
    st->step = 0;
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
    } else if (!isfinite(st->alpha)) {
      st->status = HMV_SOLVE_BREAKDOWN;
      st->done = 1;
    }
  } else if (stage == 1) {
    // both dots zero: the half step (zeta = 0) is taken and stage 2 reports
    // convergence, or breakdown through beta = alpha / zeta
    const bool half = d[0] == 0.0 && d[1] == 0.0;
    st->zeta = half ? 0.0 : d[0] / d[1];
    if (!half && (!isfinite(st->zeta) || d[1] == 0.0)) {
      st->status = isfinite(d[1]) ? HMV_SOLVE_BREAKDOWN : HMV_SOLVE_NONFINITE;
      st->done = 1;
    }
  } else {
    const double znormold = st->znorm, alpha = st->alpha, zeta = st->zeta;
    const double znorm = d[0];
    const double beta = (alpha / zeta) * (znorm / znormold);
    st->znorm = znorm;
    st->beta = beta;
    st->alpha = znorm / (d[1] + beta * d[2] - beta * zeta * d[3]);
    st->rnorm = sqrt(d[4]);
    st->step += 1;
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
    } else if (!isfinite(st->alpha) || !isfinite(st->beta)) {
      st->status = HMV_SOLVE_BREAKDOWN;
      st->done = 1;
    }
  }
}

// ||b - v|| / ||b|| partials into part[blk * kND + {0, 1}]
__global__ __launch_bounds__(kT) void p_true_res(int n, const double *__restrict__ b,
                                                 const double *__restrict__ v,
                                                 double *__restrict__ part) {
  double d[2] = {0.0, 0.0};
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
    const double ri = b[i] - v[i];
    d[0] = __builtin_fma(ri, ri, d[0]);
    d[1] = __builtin_fma(b[i], b[i], d[1]);
  }
  put_partials<2>(d, part);
}

int grid_of(int n) { return std::max(1, std::min((n + kT - 1) / kT, kMaxB)); }

struct PipeWork {
  hipStream_t s = nullptr;
  double *mem = nullptr, *part = nullptr, *hist = nullptr;
  double *r = nullptr, *r0 = nullptr, *w = nullptr, *t = nullptr, *v = nullptr, *p = nullptr,
         *sv = nullptr, *z = nullptr, *q = nullptr, *y = nullptr;
  PipeState *st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~PipeWork() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(mem);
    (void)hipFree(part);
    (void)hipFree(hist);
    (void)hipFree(st);
  }
};

hipError_t pipe_step(hmv_arena_s *a, PipeWork &w, int n, double *x) {
  const int g = grid_of(n);
  const int *skip = &w.st->done;
  const PipeState *cst = w.st;
  hipLaunchKernelGGL(p_update1, dim3(g), dim3(kT), 0, w.s, n, w.r, w.w, w.t, w.v, w.p, w.sv, w.z,
                     w.q, w.y, w.part, cst);
  hipLaunchKernelGGL(p_final, dim3(1), dim3(kT), 0, w.s, g, w.part, w.st, 1, w.hist);
  hipError_t e = hmv::enqueue_apply(a, w.z, w.v, 0.0, skip, w.s);  // v = A z
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(p_update2, dim3(g), dim3(kT), 0, w.s, n, w.p, w.q, w.y, w.t, w.v, w.sv, w.z,
                     w.r0, x, w.r, w.w, w.part, cst);
  hipLaunchKernelGGL(p_final, dim3(1), dim3(kT), 0, w.s, g, w.part, w.st, 2, w.hist);
  e = hmv::enqueue_apply(a, w.w, w.t, 0.0, skip, w.s);  // t = A w
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace

extern "C" int hmv_bicgstab_pipe(hmv_arena a, const double *b, double *u, double eps, int maxit,
                                 int check_every, double *hist, hmv_solve_info *info) {
  if (!a || !b || !u || maxit < 0) {
    hmv::set_error("hmv_bicgstab_pipe: invalid argument");
    return HMV_ERR_ARG;
  }
  if (check_every < 1) check_every = 1;
  const int n = a->nd, g = grid_of(n);
  PipeWork w;
  const size_t nb = sizeof(double) * (size_t)n;
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&w.s, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc(&w.mem, nb * 10)) != hipSuccess ||
      (e = hipMalloc(&w.part, sizeof(double) * (size_t)g * kND)) != hipSuccess ||
      (e = hipMalloc(&w.hist, sizeof(double) * ((size_t)maxit + 1))) != hipSuccess ||
      (e = hipMalloc(&w.st, sizeof(PipeState))) != hipSuccess)
    return hmv::hip_fail(e, "hmv_bicgstab_pipe: workspace");
  double *m = w.mem;
  w.r = m;
  w.r0 = m + n;
  w.w = m + 2 * (size_t)n;
  w.t = m + 3 * (size_t)n;
  w.v = m + 4 * (size_t)n;
  w.p = m + 5 * (size_t)n;
  w.sv = m + 6 * (size_t)n;
  w.z = m + 7 * (size_t)n;
  w.q = m + 8 * (size_t)n;
  w.y = m + 9 * (size_t)n;
  auto t0 = std::chrono::steady_clock::now();
  PipeState init{};
  init.eps = eps;
  init.maxit = maxit;
  if ((e = hipMemcpyAsync(w.st, &init, sizeof(init), hipMemcpyHostToDevice, w.s)) != hipSuccess ||
      (e = hmv::enqueue_apply(a, u, w.v, 0.0, nullptr, w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_bicgstab_pipe: setup");
  hipLaunchKernelGGL(p_init, dim3(g), dim3(kT), 0, w.s, n, b, w.v, w.r, w.r0);
  if ((e = hmv::enqueue_apply(a, w.r, w.w, 0.0, nullptr, w.s)) != hipSuccess ||
      (e = hmv::enqueue_apply(a, w.w, w.t, 0.0, nullptr, w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_bicgstab_pipe: setup");
  hipLaunchKernelGGL(p_dots_setup, dim3(g), dim3(kT), 0, w.s, n, w.r, w.w, b, w.part);
  hipLaunchKernelGGL(p_final, dim3(1), dim3(kT), 0, w.s, g, w.part, w.st, 0, w.hist);
  if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(w.s)) != hipSuccess ||
      (e = hipStreamBeginCapture(w.s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_bicgstab_pipe: setup");
  hipError_t ec = pipe_step(a, w, n, u);
  e = hipStreamEndCapture(w.s, &w.graph);
  if (ec != hipSuccess) return hmv::hip_fail(ec, "hmv_bicgstab_pipe: capture");
  if (e != hipSuccess || (e = hipGraphInstantiate(&w.exec, w.graph, nullptr, nullptr, 0)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_bicgstab_pipe: graph");
  PipeState host{};
  int launched = 0;
  for (;;) {
    if ((e = hipMemcpyAsync(&host, w.st, sizeof(host), hipMemcpyDeviceToHost, w.s)) != hipSuccess ||
        (e = hipStreamSynchronize(w.s)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_bicgstab_pipe: state");
    if (host.done || launched >= maxit) break;
    const int burst = std::min(check_every, maxit - launched);
    for (int i = 0; i < burst; ++i)
      if ((e = hipGraphLaunch(w.exec, w.s)) != hipSuccess)
        return hmv::hip_fail(e, "hmv_bicgstab_pipe: launch");
    launched += burst;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (hist) {
    std::fill(hist, hist + maxit + 1, NAN);
    if ((e = hipMemcpy(hist, w.hist, sizeof(double) * ((size_t)host.step + 1),
                       hipMemcpyDeviceToHost)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_bicgstab_pipe: history");
  }
  // true residual ||b - A u|| / ||b||
  if ((e = hmv::enqueue_apply(a, u, w.v, 0.0, nullptr, w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_bicgstab_pipe: true residual");
  hipLaunchKernelGGL(p_true_res, dim3(g), dim3(kT), 0, w.s, n, b, w.v, w.part);
  std::vector<double> part((size_t)g * kND);
  if ((e = hipMemcpyAsync(part.data(), w.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                          w.s)) != hipSuccess ||
      (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_bicgstab_pipe: true residual");
  double rr = 0.0, bb = 0.0;
  for (int i = 0; i < g; ++i) {
    rr += part[(size_t)i * kND];
    bb += part[(size_t)i * kND + 1];
  }
  if (info) {
    info->nstp = host.step;
    info->status = host.done ? host.status : HMV_SOLVE_MAXITER;
    info->relres = host.rnorm / host.bnorm;
    info->true_relres = std::sqrt(rr) / std::sqrt(bb);
    info->seconds = std::chrono::duration<double>(t1 - t0).count();
    info->outer = 0;
    info->inner_total = host.step;
  }
  return HMV_OK;
}
