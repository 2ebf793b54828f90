// hmv_block_solver.hip -- k independent BiCGSTAB solves that share one
// H-matmat per step (block of right-hand sides).
//
// Each column j runs the iteration of m_HACApK_solve.f90:342-449 with its own
// scalars (alpha_j, zeta_j, beta_j, (r^,r)_j) and its own stopping test
// ||r_j|| / ||b_j|| < eps; a converged column is frozen (its updates are
// skipped) while the others continue.  Vectors are N x k row-interleaved
// (element i belongs to column i % k), k a power of two <= 128, so each
// thread of a 256-thread block stays on one column through its grid-stride
// loop; per-column dot products reduce in a fixed order (deterministic).  One
// iteration (two H-matmats and the fused vector kernels) is captured as a
// hipGraph; a device flag set when every column has stopped turns the
// remaining launched iterations into no-ops.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "hmv.h"
#include "hmv_engine.h"

namespace {

constexpr int kT = 256;
constexpr int kMaxB = 2048;  // 8 blocks (32 waves) per CU for the element-wise passes

struct ColState {
  double alpha, beta, zeta, rr, rnorm, bnorm;
  int step, done, status, pad;
};

struct BlockState {
  double eps;
  int maxit;
  int all_done;  // every column stopped
  int iter;      // iterations launched and not skipped
  int pad;
};

// Sums v over the threads of the block that share a column (tid % k);
// result valid in threads tid < k.
__device__ inline double col_sum(double v, int k, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  double s = 0.0;
  if ((int)threadIdx.x < k)
    for (int i = threadIdx.x; i < kT; i += k) s += sh[i];
  __syncthreads();
  return s;
}

// part[(blk * k + j) * 2 + {0,1}] = column-j partials of (a,b) and (c,d)
__global__ __launch_bounds__(kT) void kb_dots(int64_t n, int k, const double *__restrict__ a,
                                              const double *__restrict__ b,
                                              const double *__restrict__ c,
                                              const double *__restrict__ d,
                                              double *__restrict__ part, const BlockState *bs) {
  if (bs && bs->all_done) return;
  __shared__ double sh[kT];
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    s0 = __builtin_fma(a[i], b[i], s0);
    if (c) s1 = __builtin_fma(c[i], d[i], s1);
  }
  const double r0 = col_sum(s0, k, sh);
  const double r1 = col_sum(s1, k, sh);
  if ((int)threadIdx.x < k) {
    part[2 * ((size_t)blockIdx.x * k + threadIdx.x)] = r0;
    part[2 * ((size_t)blockIdx.x * k + threadIdx.x) + 1] = r1;
  }
}

// Scalar step of one system (column or instance) from its reduced dots:
// stage 0 setup (d0 = (r,r), d1 = (b,b)), 1 alpha, 2 zeta, 3 beta/convergence.
__device__ inline void col_step(ColState &c, double d0, double d1, const BlockState *bs, int stage) {
  if (stage == 0) {
    c.rr = d0;
    c.rnorm = sqrt(d0);
    c.bnorm = sqrt(d1);
    c.alpha = c.beta = c.zeta = 0.0;
    c.step = 0;
    c.done = 0;
    c.status = HMV_SOLVE_MAXITER;
    if (!isfinite(c.rnorm) || !isfinite(c.bnorm) || c.bnorm == 0.0) {
      c.status = HMV_SOLVE_NONFINITE;
      c.done = 1;
    } else if (c.rnorm / c.bnorm < bs->eps) {
      c.status = HMV_SOLVE_CONVERGED;
      c.done = 1;
    } else if (bs->maxit <= 0) {
      c.done = 1;
    }
  } else if (!c.done) {
    if (stage == 1) {
      c.alpha = c.rr / d0;
      if (d0 == 0.0 || !isfinite(c.alpha)) {
        c.status = isfinite(d0) ? HMV_SOLVE_BREAKDOWN : HMV_SOLVE_NONFINITE;
        c.done = 1;
        c.alpha = 0.0;
      }
    } else if (stage == 2) {
      // s = A t = 0 (both dots zero): the half step u += alpha p, r = t is
      // taken with zeta = 0; stage 3 reports convergence or breakdown
      const bool half = d1 == 0.0 && d0 == 0.0;
      c.zeta = half ? 0.0 : d0 / d1;
      if (!half && (d1 == 0.0 || !isfinite(c.zeta))) {
        c.status = isfinite(d1) ? HMV_SOLVE_BREAKDOWN : HMV_SOLVE_NONFINITE;
        c.done = 1;
        c.zeta = 0.0;
      }
    } else {
      c.beta = c.alpha / c.zeta * d0 / c.rr;
      c.rr = d0;
      c.rnorm = sqrt(d1);
      c.step += 1;
      if (!isfinite(c.rnorm)) {
        c.status = HMV_SOLVE_NONFINITE;
        c.done = 1;
      } else if (c.rnorm / c.bnorm < bs->eps) {
        c.status = HMV_SOLVE_CONVERGED;
        c.done = 1;
      } else if (c.step >= bs->maxit) {
        c.status = HMV_SOLVE_MAXITER;
        c.done = 1;
      } else if (c.rr == 0.0 || c.zeta == 0.0) {
        c.status = HMV_SOLVE_BREAKDOWN;
        c.done = 1;
      }
    }
  }
}

// Fixed-order sum of each column's block partials, then the column's
// scalar step.  A block's 2k partial values are contiguous: thread t sums
// value t % 2k over the blocks b = t / 2k + G i (G = 256 / 2k groups), so
// every load is coalesced, then the G group sums are added in order.
__global__ __launch_bounds__(kT) void kb_final(int nb, int k, const double *__restrict__ part,
                                               ColState *__restrict__ cs, BlockState *bs,
                                               int stage) {
  if (bs->all_done && stage != 0) return;
  __shared__ double red[kT];
  __shared__ int sdone;
  const int tid = threadIdx.x, nv = 2 * k, G = kT / nv;  // k is a power of two <= 128
  if (tid == 0) sdone = 1;
  const int v = tid % nv, q = tid / nv;
  double acc = 0.0;
#pragma unroll 8
  for (int b = q; b < nb; b += G) acc += part[(size_t)b * nv + v];
  red[tid] = acc;
  __syncthreads();
  if (tid < nv) {
    double sum = red[tid];
    for (int g = 1; g < G; ++g) sum += red[g * nv + tid];
    red[tid] = sum;
  }
  __syncthreads();
  if (tid < k) {
    ColState c = cs[tid];
    col_step(c, red[2 * tid], red[2 * tid + 1], bs, stage);
    cs[tid] = c;
    if (!c.done) atomicAnd(&sdone, 0);
  }
  __syncthreads();
  if (tid == 0) {
    if (stage == 0) bs->iter = 0;
    if (stage == 3) bs->iter += 1;
    bs->all_done = sdone;
  }
}

// Element updates skip frozen columns (alpha/zeta of a column that stopped
// mid-iteration are never applied).
__global__ __launch_bounds__(kT) void kb_update_p(int64_t n, int k, const double *__restrict__ r,
                                                  const double *__restrict__ v,
                                                  double *__restrict__ p,
                                                  const ColState *__restrict__ cs,
                                                  const BlockState *bs) {
  if (bs->all_done) return;
  const int j = threadIdx.x % k;
  if (cs[j].done) return;
  const double beta = cs[j].beta, zeta = cs[j].zeta;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    p[i] = r[i] + beta * (p[i] - zeta * v[i]);
}

__global__ __launch_bounds__(kT) void kb_update_t(int64_t n, int k, const double *__restrict__ r,
                                                  const double *__restrict__ v,
                                                  double *__restrict__ t,
                                                  const ColState *__restrict__ cs,
                                                  const BlockState *bs) {
  if (bs->all_done) return;
  const int j = threadIdx.x % k;
  if (cs[j].done) return;
  const double alpha = cs[j].alpha;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    t[i] = r[i] - alpha * v[i];
}

// u += alpha p + zeta t;  r = t - zeta s;  partials of (r^,r) and (r,r)
__global__ __launch_bounds__(kT) void kb_update_ur(
    int64_t n, int k, const double *__restrict__ p, const double *__restrict__ t,
    const double *__restrict__ s, const double *__restrict__ rhat, double *__restrict__ u,
    double *__restrict__ r, double *__restrict__ part, const ColState *__restrict__ cs,
    const BlockState *bs) {
  if (bs->all_done) return;
  __shared__ double sh[kT];
  const int j = threadIdx.x % k;
  const bool live = !cs[j].done;
  const double alpha = cs[j].alpha, zeta = cs[j].zeta;
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    double ri = r[i];
    if (live) {
      u[i] = (u[i] + alpha * p[i]) + zeta * t[i];
      ri = t[i] - zeta * s[i];
      r[i] = ri;
    }
    s0 = __builtin_fma(rhat[i], ri, s0);
    s1 = __builtin_fma(ri, ri, s1);
  }
  const double r0 = col_sum(s0, k, sh);
  const double r1 = col_sum(s1, k, sh);
  if ((int)threadIdx.x < k) {
    part[2 * ((size_t)blockIdx.x * k + threadIdx.x)] = r0;
    part[2 * ((size_t)blockIdx.x * k + threadIdx.x) + 1] = r1;
  }
}

// r = b - r (r holds A u), r^ = r, p = v = 0
__global__ __launch_bounds__(kT) void kb_init(int64_t n, const double *__restrict__ b,
                                              double *__restrict__ r, double *__restrict__ rhat,
                                              double *__restrict__ p, double *__restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const double ri = b[i] - r[i];
    r[i] = ri;
    rhat[i] = ri;
    p[i] = 0.0;
    v[i] = 0.0;
  }
}

__global__ __launch_bounds__(kT) void kb_residual(int64_t n, const double *__restrict__ b,
                                                  double *__restrict__ r) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    r[i] = b[i] - r[i];
}

int grid_of(int64_t n) {
  const int64_t g = (n + kT - 1) / kT;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, kMaxB));
}

struct BlockWork {
  int k = 0;
  int64_t n = 0;  // N * k
  hipStream_t s = nullptr;
  double *r = nullptr, *rhat = nullptr, *p = nullptr, *v = nullptr, *t = nullptr, *sv = nullptr;
  double *part = nullptr;
  ColState *cs = nullptr;
  BlockState *bs = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const double *graph_u = nullptr;  // the iterate the captured graph updates
  ~BlockWork() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (s) (void)hipStreamDestroy(s);
    for (double *q : {r, rhat, p, v, t, sv, part}) (void)hipFree(q);
    (void)hipFree(cs);
    (void)hipFree(bs);
  }
};

hipError_t enqueue_iteration(hmv_arena_s *a, BlockWork &w, double *u) {
  const int g = grid_of(w.n), k = w.k;
  const int *skip = &w.bs->all_done;
  hipLaunchKernelGGL(kb_update_p, dim3(g), dim3(kT), 0, w.s, w.n, k, w.r, w.v, w.p, w.cs, w.bs);
  hipError_t e = hmv::enqueue_apply_block(a, k, w.p, w.v, 0.0, skip, w.s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kb_dots, dim3(g), dim3(kT), 0, w.s, w.n, k, w.rhat, w.v,
                     (const double *)nullptr, (const double *)nullptr, w.part,
                     (const BlockState *)w.bs);
  hipLaunchKernelGGL(kb_final, dim3(1), dim3(kT), 0, w.s, g, k, w.part, w.cs, w.bs, 1);
  hipLaunchKernelGGL(kb_update_t, dim3(g), dim3(kT), 0, w.s, w.n, k, w.r, w.v, w.t, w.cs, w.bs);
  if ((e = hmv::enqueue_apply_block(a, k, w.t, w.sv, 0.0, skip, w.s)) != hipSuccess) return e;
  hipLaunchKernelGGL(kb_dots, dim3(g), dim3(kT), 0, w.s, w.n, k, w.sv, w.t, w.sv, w.sv, w.part,
                     (const BlockState *)w.bs);
  hipLaunchKernelGGL(kb_final, dim3(1), dim3(kT), 0, w.s, g, k, w.part, w.cs, w.bs, 2);
  hipLaunchKernelGGL(kb_update_ur, dim3(g), dim3(kT), 0, w.s, w.n, k, w.p, w.t, w.sv, w.rhat, u,
                     w.r, w.part, w.cs, w.bs);
  hipLaunchKernelGGL(kb_final, dim3(1), dim3(kT), 0, w.s, g, k, w.part, w.cs, w.bs, 3);
  return hipGetLastError();
}

// per-column ||b - A u|| / ||b|| into rel[k]; r (N x k) receives b - A u
int block_residual(hmv_arena_s *a, BlockWork &w, const double *b, const double *u, double *r,
                   std::vector<double> &rel) {
  const int g = grid_of(w.n), k = w.k;
  hipError_t e = hmv::enqueue_apply_block(a, k, u, r, 0.0, nullptr, w.s);
  if (e != hipSuccess) return hmv::hip_fail(e, "block true residual");
  hipLaunchKernelGGL(kb_residual, dim3(g), dim3(kT), 0, w.s, w.n, b, r);
  hipLaunchKernelGGL(kb_dots, dim3(g), dim3(kT), 0, w.s, w.n, k, r, r, b, b, w.part,
                     (const BlockState *)nullptr);
  std::vector<double> part(2 * (size_t)g * k);
  if ((e = hipMemcpyAsync(part.data(), w.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                          w.s)) != hipSuccess ||
      (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, "block true residual");
  rel.assign(k, 0.0);
  for (int j = 0; j < k; ++j) {
    double rr = 0.0, bb = 0.0;
    for (int bi = 0; bi < g; ++bi) {
      rr += part[2 * ((size_t)bi * k + j)];
      bb += part[2 * ((size_t)bi * k + j) + 1];
    }
    rel[j] = std::sqrt(rr) / std::sqrt(bb);
  }
  return HMV_OK;
}

__global__ __launch_bounds__(kT) void kb_fill_n(int64_t n, double *__restrict__ x, double v) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    x[i] = v;
}

// x = 0 in the stopped columns (element i is column i % k)
__global__ __launch_bounds__(kT) void kb_zero_cols(int64_t n, int k, double *__restrict__ x,
                                                   const int *__restrict__ stopped) {
  const int j = threadIdx.x % k;
  if (!stopped[j]) return;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    x[i] = 0.0;
}

// u += d
__global__ __launch_bounds__(kT) void kb_add_n(int64_t n, double *__restrict__ u,
                                               const double *__restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    u[i] += d[i];
}

// Buffers of a k-column block solve on arena a.
int block_setup(hmv_arena_s *a, BlockWork &w, int k, const char *who) {
  int rc = hmv::ensure_tk(a, k);
  if (rc) return rc;
  w.k = k;
  w.n = (int64_t)a->nd * k;
  const int g = grid_of(w.n);
  const size_t nb = sizeof(double) * (size_t)w.n;
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&w.s, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc(&w.r, nb)) != hipSuccess || (e = hipMalloc(&w.rhat, nb)) != hipSuccess ||
      (e = hipMalloc(&w.p, nb)) != hipSuccess || (e = hipMalloc(&w.v, nb)) != hipSuccess ||
      (e = hipMalloc(&w.t, nb)) != hipSuccess || (e = hipMalloc(&w.sv, nb)) != hipSuccess ||
      (e = hipMalloc(&w.part, sizeof(double) * 2 * g * k)) != hipSuccess ||
      (e = hipMalloc(&w.cs, sizeof(ColState) * k)) != hipSuccess ||
      (e = hipMalloc(&w.bs, sizeof(BlockState))) != hipSuccess)
    return hmv::hip_fail(e, (std::string(who) + ": workspace").c_str());
  return HMV_OK;
}

// k BiCGSTABs sharing one H-matmat per step, from u (zero_guess: u = 0 and
// r = b without a product), until every column has stopped or maxit
// iterations ran; cs receives the columns' final states, iters the
// iterations run.  The iteration graph is captured on the first call with a
// given u and reused by later calls.
int block_solve(hmv_arena_s *a, BlockWork &w, const double *b, double *u, double eps, int maxit,
                int check_every, bool zero_guess, std::vector<ColState> &cs, int &iters,
                const char *who) {
  const std::string name(who);
  const int g = grid_of(w.n), k = w.k;
  BlockState init{};
  init.eps = eps;
  init.maxit = maxit;
  hipError_t e = hipMemcpyAsync(w.bs, &init, sizeof(init), hipMemcpyHostToDevice, w.s);
  if (e != hipSuccess) return hmv::hip_fail(e, (name + ": setup").c_str());
  if (zero_guess)
    hipLaunchKernelGGL(kb_fill_n, dim3(g), dim3(kT), 0, w.s, w.n, w.r, 0.0);
  else if ((e = hmv::enqueue_apply_block(a, k, u, w.r, 0.0, nullptr, w.s)) != hipSuccess)
    return hmv::hip_fail(e, (name + ": setup").c_str());
  hipLaunchKernelGGL(kb_init, dim3(g), dim3(kT), 0, w.s, w.n, b, w.r, w.rhat, w.p, w.v);
  hipLaunchKernelGGL(kb_dots, dim3(g), dim3(kT), 0, w.s, w.n, k, w.r, w.r, b, b, w.part,
                     (const BlockState *)nullptr);
  hipLaunchKernelGGL(kb_final, dim3(1), dim3(kT), 0, w.s, g, k, w.part, w.cs, w.bs, 0);
  if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, (name + ": setup").c_str());
  if (!w.exec || w.graph_u != u) {
    if (w.exec) (void)hipGraphExecDestroy(w.exec);
    if (w.graph) (void)hipGraphDestroy(w.graph);
    w.exec = nullptr;
    w.graph = nullptr;
    if ((e = hipStreamBeginCapture(w.s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return hmv::hip_fail(e, (name + ": capture").c_str());
    hipError_t ec = enqueue_iteration(a, w, u);
    e = hipStreamEndCapture(w.s, &w.graph);
    if (ec != hipSuccess) return hmv::hip_fail(ec, (name + ": capture").c_str());
    if (e != hipSuccess ||
        (e = hipGraphInstantiate(&w.exec, w.graph, nullptr, nullptr, 0)) != hipSuccess)
      return hmv::hip_fail(e, (name + ": graph").c_str());
    w.graph_u = u;
  }
  BlockState host{};
  int launched = 0;
  while (launched < maxit) {
    const int burst = std::min(check_every, maxit - launched);
    for (int i = 0; i < burst; ++i)
      if ((e = hipGraphLaunch(w.exec, w.s)) != hipSuccess)
        return hmv::hip_fail(e, (name + ": launch").c_str());
    launched += burst;
    if ((e = hipMemcpyAsync(&host, w.bs, sizeof(host), hipMemcpyDeviceToHost, w.s)) != hipSuccess ||
        (e = hipStreamSynchronize(w.s)) != hipSuccess)
      return hmv::hip_fail(e, (name + ": state").c_str());
    if (host.all_done) break;
  }
  if ((e = hipMemcpyAsync(&host, w.bs, sizeof(host), hipMemcpyDeviceToHost, w.s)) != hipSuccess ||
      (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, (name + ": state").c_str());
  iters = host.iter;
  cs.resize(k);
  if ((e = hipMemcpy(cs.data(), w.cs, sizeof(ColState) * k, hipMemcpyDeviceToHost)) != hipSuccess)
    return hmv::hip_fail(e, (name + ": columns").c_str());
  return HMV_OK;
}

}  // namespace

extern "C" int hmv_block_bicgstab(hmv_arena a, int k, const double *b, double *u, double eps,
                                  int maxit, int check_every, hmv_column_info *cols,
                                  hmv_solve_info *info) {
  if (!a || !b || !u || k < 1 || k > HMV_MAX_RHS || (k & (k - 1)) != 0 || maxit < 0) {
    hmv::set_error("hmv_block_bicgstab: invalid argument (k must be a power of two <= 128)");
    return HMV_ERR_ARG;
  }
  if (check_every < 1) check_every = 1;
  const char *who = "hmv_block_bicgstab";
  BlockWork w;
  int rc = block_setup(a, w, k, who);
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  std::vector<ColState> cs;
  int iters = 0;
  if ((rc = block_solve(a, w, b, u, eps, maxit, check_every, false, cs, iters, who))) return rc;
  auto t1 = std::chrono::steady_clock::now();
  std::vector<double> rel;
  if ((rc = block_residual(a, w, b, u, w.r, rel)) != HMV_OK) return rc;
  int maxstep = 0, worst = HMV_SOLVE_CONVERGED;
  double maxrel = 0.0;
  for (int j = 0; j < k; ++j) {
    const int status = cs[j].done ? cs[j].status : HMV_SOLVE_MAXITER;
    if (cols) {
      cols[j].nstp = cs[j].step;
      cols[j].status = status;
      cols[j].relres = cs[j].rnorm / cs[j].bnorm;
      cols[j].true_relres = rel[j];
    }
    maxstep = std::max(maxstep, cs[j].step);
    maxrel = std::max(maxrel, rel[j]);
    if (status != HMV_SOLVE_CONVERGED) worst = status;
  }
  if (info) {
    info->nstp = iters;
    info->status = worst;
    info->relres = maxrel;
    info->true_relres = maxrel;
    info->seconds = std::chrono::duration<double>(t1 - t0).count();
    info->outer = 0;
    info->inner_total = maxstep;
  }
  return HMV_OK;
}

extern "C" int hmv_block_bicgstab_ir(hmv_arena a64, hmv_arena alow, int k, const double *b,
                                     double *u, double eps, double inner_eps, int max_outer,
                                     int max_inner, int check_every, hmv_column_info *cols,
                                     hmv_solve_info *info) {
  const char *who = "hmv_block_bicgstab_ir";
  if (!a64 || !alow || !b || !u || k < 1 || k > HMV_MAX_RHS || (k & (k - 1)) != 0 ||
      max_outer < 1 || max_inner < 1 || a64->nd != alow->nd ||
      !(inner_eps > 0.0 && inner_eps < 1.0)) {
    hmv::set_error(std::string(who) + ": invalid argument (arenas of one matrix, k a power of "
                   "two <= 128, 0 < inner_eps < 1)");
    return HMV_ERR_ARG;
  }
  if (check_every < 1) check_every = 1;
  int rc = hmv::ensure_tk(a64, k);
  if (rc) return rc;
  BlockWork w;  // inner solves over the low-precision arena
  if ((rc = block_setup(alow, w, k, who))) return rc;
  struct Buf {
    double *r2 = nullptr, *d = nullptr;
    int *stopped = nullptr;
    ~Buf() {
      (void)hipFree(r2);
      (void)hipFree(d);
      (void)hipFree(stopped);
    }
  } buf;
  const size_t nbytes = sizeof(double) * (size_t)w.n;
  hipError_t e;
  if ((e = hipMalloc(&buf.r2, nbytes)) != hipSuccess || (e = hipMalloc(&buf.d, nbytes)) != hipSuccess ||
      (e = hipMalloc(&buf.stopped, sizeof(int) * k)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_block_bicgstab_ir: workspace");
  const int g = grid_of(w.n);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<int> status(k, HMV_SOLVE_MAXITER), stopped(k, 0), steps(k, 0);
  std::vector<double> rel(k, 0.0), now_rel;
  std::vector<ColState> cs;
  int outer = 0, inner_total = 0;
  for (;;) {
    // FP64 residual of every column with the FP64 arena
    if ((rc = block_residual(a64, w, b, u, buf.r2, now_rel))) return rc;
    bool all = true;
    for (int j = 0; j < k; ++j) {
      if (stopped[j]) continue;
      rel[j] = now_rel[j];
      if (!std::isfinite(rel[j]))
        status[j] = HMV_SOLVE_NONFINITE;
      else if (rel[j] < eps)
        status[j] = HMV_SOLVE_CONVERGED;
      else {
        all = false;
        continue;
      }
      stopped[j] = 1;
    }
    if (all || outer >= max_outer) break;
    // inner solves A_low D = R on the columns still running (a zero column
    // stops at its setup)
    if ((e = hipMemcpyAsync(buf.stopped, stopped.data(), sizeof(int) * k, hipMemcpyHostToDevice,
                            w.s)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_block_bicgstab_ir: flags");
    hipLaunchKernelGGL(kb_zero_cols, dim3(g), dim3(kT), 0, w.s, w.n, k, buf.r2, buf.stopped);
    hipLaunchKernelGGL(kb_fill_n, dim3(g), dim3(kT), 0, w.s, w.n, buf.d, 0.0);
    int iters = 0;
    if ((rc = block_solve(alow, w, buf.r2, buf.d, inner_eps, max_inner, check_every, true, cs, iters,
                          who)))
      return rc;
    inner_total += iters;
    for (int j = 0; j < k; ++j) {
      if (stopped[j]) continue;
      steps[j] += cs[j].step;
      if (cs[j].done && cs[j].status == HMV_SOLVE_NONFINITE) {
        status[j] = HMV_SOLVE_NONFINITE;
        stopped[j] = 1;
      }
    }
    // U += D, with D = 0 in every stopped column
    if ((e = hipMemcpyAsync(buf.stopped, stopped.data(), sizeof(int) * k, hipMemcpyHostToDevice,
                            w.s)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_block_bicgstab_ir: flags");
    hipLaunchKernelGGL(kb_zero_cols, dim3(g), dim3(kT), 0, w.s, w.n, k, buf.d, buf.stopped);
    hipLaunchKernelGGL(kb_add_n, dim3(g), dim3(kT), 0, w.s, w.n, u, buf.d);
    if ((e = hipGetLastError()) != hipSuccess)
      return hmv::hip_fail(e, "hmv_block_bicgstab_ir: update");
    ++outer;
  }
  if ((e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_block_bicgstab_ir: sync");
  const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  int worst = HMV_SOLVE_CONVERGED;
  double maxrel = 0.0;
  for (int j = 0; j < k; ++j) {
    if (cols) {
      cols[j].nstp = steps[j];
      cols[j].status = status[j];
      cols[j].relres = rel[j];
      cols[j].true_relres = rel[j];
    }
    if (std::isfinite(rel[j])) maxrel = std::max(maxrel, rel[j]);
    if (status[j] != HMV_SOLVE_CONVERGED) worst = status[j];
  }
  if (info) {
    info->nstp = inner_total;
    info->status = worst;
    info->relres = maxrel;
    info->true_relres = maxrel;
    info->seconds = total;
    info->outer = outer;
    info->inner_total = inner_total;
  }
  return HMV_OK;
}

// ---- many instances: one solve per contiguous segment ---------------------
//
// Instances of a multi-instance arena occupy contiguous row segments.  The
// vector kernels run over a table of blocks that never cross a segment
// (block b covers [beg, end) of segment seg), so a block reads its
// instance's scalars once and its dot partials belong to one instance; the
// partials of an instance are summed in block order.

namespace {

struct SegBlock {
  int seg, beg, end, pad;
};

__device__ inline double blk_sum(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kT) void ks_dots(const SegBlock *__restrict__ tb,
                                              const double *__restrict__ a,
                                              const double *__restrict__ b,
                                              const double *__restrict__ c,
                                              const double *__restrict__ d,
                                              double *__restrict__ part, const BlockState *bs) {
  if (bs && bs->all_done) return;
  __shared__ double sh[4];
  const SegBlock B = tb[blockIdx.x];
  double s0 = 0.0, s1 = 0.0;
  for (int i = B.beg + threadIdx.x; i < B.end; i += kT) {
    s0 = __builtin_fma(a[i], b[i], s0);
    if (c) s1 = __builtin_fma(c[i], d[i], s1);
  }
  const double r0 = blk_sum(s0, sh);
  const double r1 = blk_sum(s1, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0;
    part[2 * blockIdx.x + 1] = r1;
  }
}

// One thread per instance: its blocks [bstart[s], bstart[s+1]) summed in order.
__global__ __launch_bounds__(kT) void ks_final(int nseg, const int *__restrict__ bstart,
                                               const double *__restrict__ part,
                                               ColState *__restrict__ cs, BlockState *bs,
                                               int stage, int *__restrict__ done_flags) {
  if (bs->all_done && stage != 0) return;
  __shared__ int sdone;
  if (threadIdx.x == 0) sdone = 1;
  __syncthreads();
  for (int s = threadIdx.x; s < nseg; s += kT) {
    double d0 = 0.0, d1 = 0.0;
    for (int b = bstart[s]; b < bstart[s + 1]; ++b) {
      d0 += part[2 * b];
      d1 += part[2 * b + 1];
    }
    ColState c = cs[s];
    col_step(c, d0, d1, bs, stage);
    cs[s] = c;
    done_flags[s] = c.done;
    if (!c.done) atomicAnd(&sdone, 0);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (stage == 0) bs->iter = 0;
    if (stage == 3) bs->iter += 1;
    bs->all_done = sdone;
  }
}

__global__ __launch_bounds__(kT) void ks_update_p(const SegBlock *__restrict__ tb,
                                                  const double *__restrict__ r,
                                                  const double *__restrict__ v,
                                                  double *__restrict__ p,
                                                  const ColState *__restrict__ cs,
                                                  const BlockState *bs) {
  if (bs->all_done) return;
  const SegBlock B = tb[blockIdx.x];
  if (cs[B.seg].done) return;
  const double beta = cs[B.seg].beta, zeta = cs[B.seg].zeta;
  for (int i = B.beg + threadIdx.x; i < B.end; i += kT) p[i] = r[i] + beta * (p[i] - zeta * v[i]);
}

__global__ __launch_bounds__(kT) void ks_update_t(const SegBlock *__restrict__ tb,
                                                  const double *__restrict__ r,
                                                  const double *__restrict__ v,
                                                  double *__restrict__ t,
                                                  const ColState *__restrict__ cs,
                                                  const BlockState *bs) {
  if (bs->all_done) return;
  const SegBlock B = tb[blockIdx.x];
  if (cs[B.seg].done) return;
  const double alpha = cs[B.seg].alpha;
  for (int i = B.beg + threadIdx.x; i < B.end; i += kT) t[i] = r[i] - alpha * v[i];
}

__global__ __launch_bounds__(kT) void ks_update_ur(
    const SegBlock *__restrict__ tb, const double *__restrict__ p, const double *__restrict__ t,
    const double *__restrict__ s, const double *__restrict__ rhat, double *__restrict__ u,
    double *__restrict__ r, double *__restrict__ part, const ColState *__restrict__ cs,
    const BlockState *bs) {
  if (bs->all_done) return;
  __shared__ double sh[4];
  const SegBlock B = tb[blockIdx.x];
  const bool live = !cs[B.seg].done;
  const double alpha = cs[B.seg].alpha, zeta = cs[B.seg].zeta;
  double s0 = 0.0, s1 = 0.0;
  for (int i = B.beg + threadIdx.x; i < B.end; i += kT) {
    double ri = r[i];
    if (live) {
      u[i] = (u[i] + alpha * p[i]) + zeta * t[i];
      ri = t[i] - zeta * s[i];
      r[i] = ri;
    }
    s0 = __builtin_fma(rhat[i], ri, s0);
    s1 = __builtin_fma(ri, ri, s1);
  }
  const double r0 = blk_sum(s0, sh);
  const double r1 = blk_sum(s1, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0;
    part[2 * blockIdx.x + 1] = r1;
  }
}

struct SegWork {
  int nseg = 0, nblk = 0;
  int64_t n = 0;
  hipStream_t s = nullptr;
  SegBlock *tb = nullptr;
  int *bstart = nullptr;
  double *r = nullptr, *rhat = nullptr, *p = nullptr, *v = nullptr, *t = nullptr, *sv = nullptr;
  double *part = nullptr;
  ColState *cs = nullptr;
  BlockState *bs = nullptr;
  int *flags = nullptr;  // per-instance done flags read by the H-matvec
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const double *graph_u = nullptr;  // the iterate the captured graph updates
  ~SegWork() {
    (void)hipFree(flags);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (s) (void)hipStreamDestroy(s);
    for (double *q : {r, rhat, p, v, t, sv, part}) (void)hipFree(q);
    (void)hipFree(tb);
    (void)hipFree(bstart);
    (void)hipFree(cs);
    (void)hipFree(bs);
  }
};

__global__ __launch_bounds__(kT) void kb_fill(int64_t n, double *__restrict__ x, double v) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    x[i] = v;
}

// u += d
__global__ __launch_bounds__(kT) void kb_add(int64_t n, double *__restrict__ u,
                                             const double *__restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    u[i] += d[i];
}

// r = 0 on the segments of stopped instances (their next inner solve then
// stops at its setup and their products are skipped)
__global__ __launch_bounds__(kT) void ks_zero_stopped(const SegBlock *__restrict__ tb,
                                                      double *__restrict__ r,
                                                      const int *__restrict__ stopped) {
  const SegBlock B = tb[blockIdx.x];
  if (!stopped[B.seg]) return;
  for (int i = B.beg + threadIdx.x; i < B.end; i += kT) r[i] = 0.0;
}

hipError_t seg_iteration(hmv_arena_s *a, SegWork &w, double *u) {
  const int nb = w.nblk;
  const int *skip = &w.bs->all_done;
  hipLaunchKernelGGL(ks_update_p, dim3(nb), dim3(kT), 0, w.s, w.tb, w.r, w.v, w.p, w.cs, w.bs);
  hipError_t e = hmv::enqueue_apply(a, w.p, w.v, 0.0, skip, w.s, w.flags);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(ks_dots, dim3(nb), dim3(kT), 0, w.s, w.tb, w.rhat, w.v,
                     (const double *)nullptr, (const double *)nullptr, w.part,
                     (const BlockState *)w.bs);
  hipLaunchKernelGGL(ks_final, dim3(1), dim3(kT), 0, w.s, w.nseg, w.bstart, w.part, w.cs, w.bs, 1,
                     w.flags);
  hipLaunchKernelGGL(ks_update_t, dim3(nb), dim3(kT), 0, w.s, w.tb, w.r, w.v, w.t, w.cs, w.bs);
  if ((e = hmv::enqueue_apply(a, w.t, w.sv, 0.0, skip, w.s, w.flags)) != hipSuccess) return e;
  hipLaunchKernelGGL(ks_dots, dim3(nb), dim3(kT), 0, w.s, w.tb, w.sv, w.t, w.sv, w.sv, w.part,
                     (const BlockState *)w.bs);
  hipLaunchKernelGGL(ks_final, dim3(1), dim3(kT), 0, w.s, w.nseg, w.bstart, w.part, w.cs, w.bs, 2,
                     w.flags);
  hipLaunchKernelGGL(ks_update_ur, dim3(nb), dim3(kT), 0, w.s, w.tb, w.p, w.t, w.sv, w.rhat, u,
                     w.r, w.part, w.cs, w.bs);
  hipLaunchKernelGGL(ks_final, dim3(1), dim3(kT), 0, w.s, w.nseg, w.bstart, w.part, w.cs, w.bs, 3,
                     w.flags);
  return hipGetLastError();
}

// Tables and buffers of the per-instance solver over arena a's segments:
// a block table whose blocks hold at most 16 elements per thread and never
// cross an instance; bstart[s] = first block of instance s.
int seg_setup(hmv_arena_s *a, SegWork &w, std::vector<int> &bstart, const char *who) {
  w.nseg = (int)a->seg.size() - 1;
  w.n = a->nd;
  constexpr int kChunk = 16 * kT;
  std::vector<SegBlock> tb;
  bstart.assign(w.nseg + 1, 0);
  for (int sgi = 0; sgi < w.nseg; ++sgi) {
    bstart[sgi] = (int)tb.size();
    for (int i = a->seg[sgi]; i < a->seg[sgi + 1]; i += kChunk)
      tb.push_back(SegBlock{sgi, i, std::min(a->seg[sgi + 1], i + kChunk), 0});
  }
  bstart[w.nseg] = (int)tb.size();
  w.nblk = (int)tb.size();
  const size_t nbytes = sizeof(double) * (size_t)w.n;
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&w.s, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc(&w.tb, sizeof(SegBlock) * tb.size())) != hipSuccess ||
      (e = hipMalloc(&w.bstart, sizeof(int) * bstart.size())) != hipSuccess ||
      (e = hipMalloc(&w.r, nbytes)) != hipSuccess || (e = hipMalloc(&w.rhat, nbytes)) != hipSuccess ||
      (e = hipMalloc(&w.p, nbytes)) != hipSuccess || (e = hipMalloc(&w.v, nbytes)) != hipSuccess ||
      (e = hipMalloc(&w.t, nbytes)) != hipSuccess || (e = hipMalloc(&w.sv, nbytes)) != hipSuccess ||
      (e = hipMalloc(&w.part, sizeof(double) * 2 * tb.size())) != hipSuccess ||
      (e = hipMalloc(&w.cs, sizeof(ColState) * w.nseg)) != hipSuccess ||
      (e = hipMalloc(&w.bs, sizeof(BlockState))) != hipSuccess ||
      (e = hipMalloc(&w.flags, sizeof(int) * w.nseg)) != hipSuccess ||
      (e = hipMemcpy(w.tb, tb.data(), sizeof(SegBlock) * tb.size(), hipMemcpyHostToDevice)) !=
          hipSuccess ||
      (e = hipMemcpy(w.bstart, bstart.data(), sizeof(int) * bstart.size(), hipMemcpyHostToDevice)) !=
          hipSuccess)
    return hmv::hip_fail(e, (std::string(who) + ": workspace").c_str());
  return HMV_OK;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// One BiCGSTAB per instance from u (zero_guess: u = 0, and r = b without a
// product), in lockstep until every instance has stopped or maxit
// iterations ran.  cs receives the instances' final states; done_at (may be
// null) the time since t0 at which each instance was first seen stopped;
// iters the iterations run.  The iteration graph is captured on the first
// call with a given u and reused by later calls.
int seg_solve(hmv_arena_s *a, SegWork &w, const double *b, double *u, double eps, int maxit,
              int check_every, bool zero_guess, std::chrono::steady_clock::time_point t0,
              std::vector<ColState> &cs, std::vector<double> *done_at, int &iters,
              const char *who) {
  const std::string name(who);
  BlockState init{};
  init.eps = eps;
  init.maxit = maxit;
  const int g = grid_of(w.n);
  hipError_t e = hipMemcpyAsync(w.bs, &init, sizeof(init), hipMemcpyHostToDevice, w.s);
  if (e != hipSuccess) return hmv::hip_fail(e, (name + ": setup").c_str());
  if (zero_guess)
    hipLaunchKernelGGL(kb_fill, dim3(g), dim3(kT), 0, w.s, w.n, w.r, 0.0);
  else if ((e = hmv::enqueue_apply(a, u, w.r, 0.0, nullptr, w.s)) != hipSuccess)
    return hmv::hip_fail(e, (name + ": setup").c_str());
  hipLaunchKernelGGL(kb_init, dim3(g), dim3(kT), 0, w.s, w.n, b, w.r, w.rhat, w.p, w.v);
  hipLaunchKernelGGL(ks_dots, dim3(w.nblk), dim3(kT), 0, w.s, w.tb, w.r, w.r, b, b, w.part,
                     (const BlockState *)nullptr);
  hipLaunchKernelGGL(ks_final, dim3(1), dim3(kT), 0, w.s, w.nseg, w.bstart, w.part, w.cs, w.bs, 0,
                     w.flags);
  if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, (name + ": setup").c_str());
  if (!w.exec || w.graph_u != u) {
    if (w.exec) (void)hipGraphExecDestroy(w.exec);
    if (w.graph) (void)hipGraphDestroy(w.graph);
    w.exec = nullptr;
    w.graph = nullptr;
    if ((e = hipStreamBeginCapture(w.s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return hmv::hip_fail(e, (name + ": capture").c_str());
    hipError_t ec = seg_iteration(a, w, u);
    e = hipStreamEndCapture(w.s, &w.graph);
    if (ec != hipSuccess) return hmv::hip_fail(ec, (name + ": capture").c_str());
    if (e != hipSuccess ||
        (e = hipGraphInstantiate(&w.exec, w.graph, nullptr, nullptr, 0)) != hipSuccess)
      return hmv::hip_fail(e, (name + ": graph").c_str());
    w.graph_u = u;
  }
  cs.resize(w.nseg);
  auto poll = [&]() -> hipError_t {
    hipError_t pe = hipMemcpyAsync(cs.data(), w.cs, sizeof(ColState) * w.nseg,
                                   hipMemcpyDeviceToHost, w.s);
    if (pe == hipSuccess) pe = hipStreamSynchronize(w.s);
    if (pe == hipSuccess && done_at) {
      const double now = seconds_since(t0);
      for (int sgi = 0; sgi < w.nseg; ++sgi)
        if (cs[sgi].done && (*done_at)[sgi] < 0.0) (*done_at)[sgi] = now;
    }
    return pe;
  };
  if ((e = poll()) != hipSuccess) return hmv::hip_fail(e, (name + ": state").c_str());
  BlockState host{};
  int launched = 0;
  while (launched < maxit) {
    const int burst = std::min(check_every, maxit - launched);
    for (int i = 0; i < burst; ++i)
      if ((e = hipGraphLaunch(w.exec, w.s)) != hipSuccess)
        return hmv::hip_fail(e, (name + ": launch").c_str());
    launched += burst;
    if ((e = hipMemcpyAsync(&host, w.bs, sizeof(host), hipMemcpyDeviceToHost, w.s)) != hipSuccess ||
        (e = poll()) != hipSuccess)
      return hmv::hip_fail(e, (name + ": state").c_str());
    if (host.all_done) break;
  }
  if ((e = hipMemcpyAsync(&host, w.bs, sizeof(host), hipMemcpyDeviceToHost, w.s)) != hipSuccess ||
      (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, (name + ": state").c_str());
  iters = host.iter;
  return HMV_OK;
}

// Per-instance ||b - A u||^2 and ||b||^2 (r receives b - A u), each summed
// over the instance's blocks in block order.
int seg_residuals(hmv_arena_s *a, SegWork &w, const double *b, const double *u, double *r,
                  const std::vector<int> &bstart, std::vector<double> &rr, std::vector<double> &bb,
                  const char *who) {
  const int g = grid_of(w.n);
  hipError_t e = hmv::enqueue_apply(a, u, r, 0.0, nullptr, w.s);
  if (e != hipSuccess) return hmv::hip_fail(e, (std::string(who) + ": residual").c_str());
  hipLaunchKernelGGL(kb_residual, dim3(g), dim3(kT), 0, w.s, w.n, b, r);
  hipLaunchKernelGGL(ks_dots, dim3(w.nblk), dim3(kT), 0, w.s, w.tb, r, r, b, b, w.part,
                     (const BlockState *)nullptr);
  std::vector<double> part(2 * (size_t)w.nblk);
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(part.data(), w.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                          w.s)) != hipSuccess ||
      (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, (std::string(who) + ": residual").c_str());
  rr.assign(w.nseg, 0.0);
  bb.assign(w.nseg, 0.0);
  for (int sgi = 0; sgi < w.nseg; ++sgi)
    for (int bi = bstart[sgi]; bi < bstart[sgi + 1]; ++bi) {
      rr[sgi] += part[2 * bi];
      bb[sgi] += part[2 * bi + 1];
    }
  return HMV_OK;
}

}  // namespace

extern "C" int hmv_multi_bicgstab(hmv_arena a, const double *b, double *u, double eps, int maxit,
                                  int check_every, hmv_column_info *inst, double *inst_seconds,
                                  hmv_solve_info *info) {
  if (!a || !b || !u || maxit < 0 || a->seg.size() < 2) {
    hmv::set_error("hmv_multi_bicgstab: invalid argument");
    return HMV_ERR_ARG;
  }
  if (check_every < 1) check_every = 1;
  const char *who = "hmv_multi_bicgstab";
  SegWork w;
  std::vector<int> bstart;
  int rc = seg_setup(a, w, bstart, who);
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  std::vector<ColState> cs;
  std::vector<double> done_at(w.nseg, -1.0);
  int iters = 0;
  if ((rc = seg_solve(a, w, b, u, eps, maxit, check_every, false, t0, cs, &done_at, iters, who)))
    return rc;
  const double total = seconds_since(t0);
  // true residuals per instance: r = b - A u
  std::vector<double> rr, bb;
  if ((rc = seg_residuals(a, w, b, u, w.r, bstart, rr, bb, who))) return rc;
  int maxstep = 0, worst = HMV_SOLVE_CONVERGED;
  double maxrel = 0.0;
  for (int sgi = 0; sgi < w.nseg; ++sgi) {
    const double rel = std::sqrt(rr[sgi]) / std::sqrt(bb[sgi]);
    const int status = cs[sgi].done ? cs[sgi].status : HMV_SOLVE_MAXITER;
    if (inst) {
      inst[sgi].nstp = cs[sgi].step;
      inst[sgi].status = status;
      inst[sgi].relres = cs[sgi].rnorm / cs[sgi].bnorm;
      inst[sgi].true_relres = rel;
    }
    if (inst_seconds) inst_seconds[sgi] = done_at[sgi] >= 0.0 ? done_at[sgi] : total;
    maxstep = std::max(maxstep, cs[sgi].step);
    if (std::isfinite(rel)) maxrel = std::max(maxrel, rel);
    if (status != HMV_SOLVE_CONVERGED) worst = status;
  }
  if (info) {
    info->nstp = iters;
    info->status = worst;
    info->relres = maxrel;
    info->true_relres = maxrel;
    info->seconds = total;
    info->outer = 0;
    info->inner_total = maxstep;
  }
  return HMV_OK;
}

extern "C" int hmv_multi_bicgstab_ir(hmv_arena a64, hmv_arena alow, const double *b, double *u,
                                     double eps, double inner_eps, int max_outer, int max_inner,
                                     int check_every, hmv_column_info *inst,
                                     double *inst_seconds, hmv_solve_info *info) {
  const char *who = "hmv_multi_bicgstab_ir";
  if (!a64 || !alow || !b || !u || max_outer < 1 || max_inner < 1 || a64->seg.size() < 2 ||
      a64->seg != alow->seg || a64->nd != alow->nd || !(inner_eps > 0.0 && inner_eps < 1.0)) {
    hmv::set_error(std::string(who) + ": invalid argument (two arenas of the same instances, "
                   "0 < inner_eps < 1)");
    return HMV_ERR_ARG;
  }
  if (check_every < 1) check_every = 1;
  SegWork w;  // inner solves over the low-precision arena
  std::vector<int> bstart;
  int rc = seg_setup(alow, w, bstart, who);
  if (rc) return rc;
  const int nseg = w.nseg;
  struct Buf {
    double *r2 = nullptr, *d = nullptr;
    int *stopped = nullptr;
    ~Buf() {
      (void)hipFree(r2);
      (void)hipFree(d);
      (void)hipFree(stopped);
    }
  } buf;
  const size_t nbytes = sizeof(double) * (size_t)w.n;
  hipError_t e;
  if ((e = hipMalloc(&buf.r2, nbytes)) != hipSuccess || (e = hipMalloc(&buf.d, nbytes)) != hipSuccess ||
      (e = hipMalloc(&buf.stopped, sizeof(int) * nseg)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_multi_bicgstab_ir: workspace");
  const int g = grid_of(w.n);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<int> status(nseg, HMV_SOLVE_MAXITER), stopped(nseg, 0), steps(nseg, 0);
  std::vector<double> rel(nseg, 0.0), done_at(nseg, -1.0), rr, bb;
  std::vector<ColState> cs;
  int outer = 0, inner_total = 0;
  for (;;) {
    // FP64 residual of every instance on the FP64 arena
    if ((rc = seg_residuals(a64, w, b, u, buf.r2, bstart, rr, bb, who))) return rc;
    const double now = seconds_since(t0);
    bool all = true;
    for (int s = 0; s < nseg; ++s) {
      if (stopped[s]) continue;
      rel[s] = std::sqrt(rr[s]) / std::sqrt(bb[s]);
      if (!std::isfinite(rel[s]))
        status[s] = HMV_SOLVE_NONFINITE;
      else if (rel[s] < eps)
        status[s] = HMV_SOLVE_CONVERGED;
      else {
        all = false;
        continue;
      }
      stopped[s] = 1;
      done_at[s] = now;
    }
    if (all || outer >= max_outer) break;
    // inner solves A_low d = r on the instances still running
    if ((e = hipMemcpyAsync(buf.stopped, stopped.data(), sizeof(int) * nseg, hipMemcpyHostToDevice,
                            w.s)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_multi_bicgstab_ir: flags");
    hipLaunchKernelGGL(ks_zero_stopped, dim3(w.nblk), dim3(kT), 0, w.s, w.tb, buf.r2, buf.stopped);
    hipLaunchKernelGGL(kb_fill, dim3(g), dim3(kT), 0, w.s, w.n, buf.d, 0.0);
    int iters = 0;
    if ((rc = seg_solve(alow, w, buf.r2, buf.d, inner_eps, max_inner, check_every, true, t0, cs,
                        nullptr, iters, who)))
      return rc;
    inner_total += iters;
    for (int s = 0; s < nseg; ++s) {
      if (stopped[s]) continue;
      steps[s] += cs[s].step;
      if (cs[s].done && cs[s].status == HMV_SOLVE_NONFINITE) {
        // the correction of this instance is not usable: it stops here
        status[s] = HMV_SOLVE_NONFINITE;
        stopped[s] = 1;
        done_at[s] = seconds_since(t0);
      }
    }
    // u += d, with d = 0 on every stopped instance (those stopped before this
    // solve never updated d; one stopped just now may hold a non-finite d)
    if ((e = hipMemcpyAsync(buf.stopped, stopped.data(), sizeof(int) * nseg, hipMemcpyHostToDevice,
                            w.s)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_multi_bicgstab_ir: flags");
    hipLaunchKernelGGL(ks_zero_stopped, dim3(w.nblk), dim3(kT), 0, w.s, w.tb, buf.d, buf.stopped);
    hipLaunchKernelGGL(kb_add, dim3(g), dim3(kT), 0, w.s, w.n, u, buf.d);
    if ((e = hipGetLastError()) != hipSuccess)
      return hmv::hip_fail(e, "hmv_multi_bicgstab_ir: update");
    ++outer;
  }
  if ((e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_multi_bicgstab_ir: sync");
  const double total = seconds_since(t0);
  int worst = HMV_SOLVE_CONVERGED;
  double maxrel = 0.0;
  for (int s = 0; s < nseg; ++s) {
    if (inst) {
      inst[s].nstp = steps[s];
      inst[s].status = status[s];
      inst[s].relres = rel[s];
      inst[s].true_relres = rel[s];
    }
    if (inst_seconds) inst_seconds[s] = done_at[s] >= 0.0 ? done_at[s] : total;
    if (std::isfinite(rel[s])) maxrel = std::max(maxrel, rel[s]);
    if (status[s] != HMV_SOLVE_CONVERGED) worst = status[s];
  }
  if (info) {
    info->nstp = inner_total;
    info->status = worst;
    info->relres = maxrel;
    info->true_relres = maxrel;
    info->seconds = total;
    info->outer = outer;
    info->inner_total = inner_total;
  }
  return HMV_OK;
}

// ---- many instances: GCR(m) per segment (inner solver of hmv_multi_gcr_ir) -
//
// One GCR(m) per instance (m_HACApK_solve.f90:506-571, as hmv_gcrm runs it),
// over the same block table.  The running instances take their steps
// together, so the slot of a step (ik) and of the next direction (ikn) are
// common to all of them; each instance keeps its own alpha, betas and
// (A p_l, A p_l).  A stopped instance's blocks skip every update, and its
// H-matvec workgroups return at once (per-instance done flags).

namespace {

constexpr int kGcrM = HMV_GCR_MAX_M;
constexpr int kGroup = 4;  // slots per block row of the beta dots

struct SgInst {
  double alpha, rnorm, bnorm;
  double capap[kGcrM];
  double beta[kGcrM];
  int step, done, status, pad;
};

struct SgGlobal {
  double eps;
  int step, ik, ikn, m, maxit, all_done, pad;
};

// Fixed-order sum over the blocks [b0, b1) of part[stride * b + off]: thread
// t takes blocks b0 + t, b0 + t + kT, ..., then the block sum.
__device__ inline double inst_sum(const double *__restrict__ part, int b0, int b1, int stride,
                                  int off, double *sh) {
  double s = 0.0;
  for (int b = b0 + (int)threadIdx.x; b < b1; b += kT) s += part[(size_t)stride * b + off];
  return blk_sum(s, sh);
}

// r = b, p_0 = b (zero initial guess)
__global__ __launch_bounds__(kT) void sg_init(int64_t n, const double *__restrict__ b,
                                              double *__restrict__ r, double *__restrict__ p) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    r[i] = b[i];
    p[i] = b[i];
  }
}

// (ap_ikn, r), (ap_ikn, ap_ikn) partials of each block (the first step's)
__global__ __launch_bounds__(kT) void sg_dots_alpha(const SegBlock *__restrict__ tb, int64_t n,
                                                    const double *__restrict__ ap,
                                                    const double *__restrict__ r,
                                                    double *__restrict__ part,
                                                    const SgGlobal *gs) {
  if (gs->all_done) return;
  __shared__ double sh[4];
  const SegBlock B = tb[blockIdx.x];
  const double *q = ap + (size_t)gs->ikn * n;
  double s0 = 0.0, s1 = 0.0;
  for (int i = B.beg + threadIdx.x; i < B.end; i += kT) {
    s0 = __builtin_fma(q[i], r[i], s0);
    s1 = __builtin_fma(q[i], q[i], s1);
  }
  const double r0 = blk_sum(s0, sh), r1 = blk_sum(s1, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0;
    part[2 * blockIdx.x + 1] = r1;
  }
}

// alpha = (ap_ik, r) / (ap_ik, ap_ik) of the block's instance, from its
// blocks' pairs in apart (every block of the instance forms it alike; the
// instance's first block records it and capap_ik, and flags a breakdown),
// then u += alpha p_ik, r -= alpha ap_ik and the block's (r, r) partial.
__global__ __launch_bounds__(kT) void sg_update_ur(const SegBlock *__restrict__ tb,
                                                   const int *__restrict__ bstart, int64_t n,
                                                   const double *__restrict__ p,
                                                   const double *__restrict__ ap,
                                                   double *__restrict__ u, double *__restrict__ r,
                                                   const double *__restrict__ apart,
                                                   double *__restrict__ rpart, SgInst *inst,
                                                   const SgGlobal *gs) {
  if (gs->all_done) return;
  __shared__ double sh[4];
  const SegBlock B = tb[blockIdx.x];
  const int s = B.seg;
  if (inst[s].done) return;  // the stage-2 pass skips stopped instances
  const int b0 = bstart[s], b1 = bstart[s + 1];
  const double d0 = inst_sum(apart, b0, b1, 2, 0, sh), d1 = inst_sum(apart, b0, b1, 2, 1, sh);
  const int ik = gs->ikn;  // the slot of this step
  const double alpha = d0 / d1;
  const bool bad = d1 == 0.0 || !isfinite(alpha);
  if ((int)blockIdx.x == b0 && threadIdx.x == 0) {
    inst[s].capap[ik] = d1;
    inst[s].alpha = alpha;
    if (bad) {
      inst[s].status = isfinite(d1) ? HMV_SOLVE_BREAKDOWN : HMV_SOLVE_NONFINITE;
      inst[s].done = 1;
    }
  }
  if (bad) return;
  const double *pk = p + (size_t)ik * n, *q = ap + (size_t)ik * n;
  double s2 = 0.0;
  for (int i = B.beg + threadIdx.x; i < B.end; i += kT) {
    u[i] = u[i] + alpha * pk[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    s2 = __builtin_fma(ri, ri, s2);
  }
  const double r0 = blk_sum(s2, sh);
  if (threadIdx.x == 0) rpart[blockIdx.x] = r0;
}

// Partials of (v_l, ar) for the slots l <= ik of the block row's group
// (v_l = ar for the slot just overwritten, l == ikn, else ap_l):
// part[blk * kGcrM + l].
__global__ __launch_bounds__(kT) void sg_dots_beta(const SegBlock *__restrict__ tb, int64_t n,
                                                   const double *__restrict__ ap,
                                                   const double *__restrict__ ar,
                                                   double *__restrict__ part,
                                                   const SgInst *__restrict__ inst,
                                                   const SgGlobal *gs) {
  if (gs->all_done) return;
  const SegBlock B = tb[blockIdx.x];
  if (inst[B.seg].done) return;
  const int ik = gs->ik, ikn = gs->ikn, l0 = blockIdx.y * kGroup;
  if (l0 > ik) return;
  __shared__ double sh[4];
  const double *v[kGroup];
#pragma unroll
  for (int j = 0; j < kGroup; ++j) {
    const int l = min(l0 + j, ik);
    v[j] = l == ikn ? ar : ap + (size_t)l * n;
  }
  double s[kGroup] = {0.0, 0.0, 0.0, 0.0};
  for (int i = B.beg + threadIdx.x; i < B.end; i += kT) {
    const double a = ar[i];
#pragma unroll
    for (int j = 0; j < kGroup; ++j) s[j] = __builtin_fma(v[j][i], a, s[j]);
  }
#pragma unroll
  for (int j = 0; j < kGroup; ++j) {
    const double r0 = blk_sum(s[j], sh);
    if (threadIdx.x == 0 && l0 + j <= ik) part[(size_t)blockIdx.x * kGcrM + l0 + j] = r0;
  }
}

// Per-instance scalar stages, one thread per instance (its blocks summed in
// block order):
//   0: setup from the (r, r), (b, b) pairs;
//   2: ||r|| from the (r, r) partials, the step and stopping tests, then the
//      common slots (ik = the slot just used, ikn = the next one);
//   3: beta_l = -(A p_l, A r) / capap_l for l <= ik (one thread per
//      instance and slot).
// done_flags mirrors each instance's done flag for the H-matvec.
__global__ __launch_bounds__(kT) void sg_final(int nseg, const int *__restrict__ bstart,
                                               const double *__restrict__ part, SgInst *inst,
                                               SgGlobal *gs, int stage,
                                               int *__restrict__ done_flags) {
  if (gs->all_done && stage != 0) return;
  __shared__ int sdone;
  if (threadIdx.x == 0) sdone = 1;
  __syncthreads();
  if (stage == 3) {
    const int ik = gs->ik;
    for (int e = threadIdx.x; e < nseg * (ik + 1); e += kT) {
      const int s = e / (ik + 1), l = e % (ik + 1);
      if (inst[s].done) continue;
      double v = 0.0;
      for (int b = bstart[s]; b < bstart[s + 1]; ++b) v += part[(size_t)b * kGcrM + l];
      inst[s].beta[l] = -v / inst[s].capap[l];
    }
    return;
  }
  for (int s = threadIdx.x; s < nseg; s += kT) {
    SgInst &c = inst[s];
    if (stage == 0) {
      double d0 = 0.0, d1 = 0.0;
      for (int b = bstart[s]; b < bstart[s + 1]; ++b) {
        d0 += part[2 * b];
        d1 += part[2 * b + 1];
      }
      c.rnorm = sqrt(d0);
      c.bnorm = sqrt(d1);
      c.alpha = 0.0;
      c.step = 0;
      c.done = 0;
      c.status = HMV_SOLVE_MAXITER;
      if (!isfinite(c.rnorm) || !isfinite(c.bnorm) || c.bnorm == 0.0) {
        c.status = HMV_SOLVE_NONFINITE;
        c.done = 1;
      } else if (c.rnorm / c.bnorm < gs->eps) {
        c.status = HMV_SOLVE_CONVERGED;
        c.done = 1;
      } else if (gs->maxit <= 0) {
        c.done = 1;
      }
    } else if (!c.done) {
      double d0 = 0.0;
      for (int b = bstart[s]; b < bstart[s + 1]; ++b) d0 += part[b];
      c.step += 1;
      c.rnorm = sqrt(d0);
      if (!isfinite(c.rnorm)) {
        c.status = HMV_SOLVE_NONFINITE;
        c.done = 1;
      } else if (c.rnorm / c.bnorm < gs->eps) {
        c.status = HMV_SOLVE_CONVERGED;
        c.done = 1;
      } else if (c.step >= gs->maxit) {
        c.status = HMV_SOLVE_MAXITER;
        c.done = 1;
      }
    }
    done_flags[s] = c.done;
    if (!c.done) atomicAnd(&sdone, 0);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (stage == 0) {
      gs->step = 0;
      gs->ik = 0;
      gs->ikn = 0;
    } else {
      gs->ik = gs->ikn;
      gs->step += 1;
      gs->ikn = gs->step % gs->m;
    }
    gs->all_done = sdone;
  }
}

// p_ikn = c0 r + sum_{l <= ik, l != ikn} beta_l p_l and ap_ikn likewise with
// ar (c0 = 1 + beta_ikn when the slot is reused, else 1), per instance, and
// the next step's alpha partials (ap_ikn, r), (ap_ikn, ap_ikn).
__global__ __launch_bounds__(kT) void sg_update_dir(const SegBlock *__restrict__ tb, int64_t n,
                                                    const double *__restrict__ r,
                                                    const double *__restrict__ ar,
                                                    double *__restrict__ p,
                                                    double *__restrict__ ap,
                                                    double *__restrict__ part,
                                                    const SgInst *__restrict__ inst,
                                                    const SgGlobal *gs) {
  if (gs->all_done) return;
  const SegBlock B = tb[blockIdx.x];
  const SgInst &c = inst[B.seg];
  if (c.done) return;
  __shared__ double sh[4];
  const int ik = gs->ik, ikn = gs->ikn;
  const double c0 = ikn <= ik ? 1.0 + c.beta[ikn] : 1.0;
  double *pn = p + (size_t)ikn * n, *apn = ap + (size_t)ikn * n;
  double s0 = 0.0, s1 = 0.0;
  for (int i = B.beg + threadIdx.x; i < B.end; i += kT) {
    const double ri = r[i];
    double pv = c0 * ri, av = c0 * ar[i];
    for (int l = 0; l <= ik; ++l) {
      if (l == ikn) continue;
      const double bl = c.beta[l];
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

struct SegGcrWork {
  int m = 0;
  double *r = nullptr, *ar = nullptr, *p = nullptr, *ap = nullptr, *part = nullptr;
  SgInst *inst = nullptr;
  SgGlobal *gs = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const double *graph_u = nullptr;
  ~SegGcrWork() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (double *q : {r, ar, p, ap, part}) (void)hipFree(q);
    (void)hipFree(inst);
    (void)hipFree(gs);
  }
};

int sg_setup(const SegWork &w, SegGcrWork &gw, int m, const char *who) {
  gw.m = m;
  const size_t nb = sizeof(double) * (size_t)w.n;
  hipError_t e;
  if ((e = hipMalloc(&gw.r, nb)) != hipSuccess || (e = hipMalloc(&gw.ar, nb)) != hipSuccess ||
      (e = hipMalloc(&gw.p, nb * m)) != hipSuccess || (e = hipMalloc(&gw.ap, nb * m)) != hipSuccess ||
      (e = hipMalloc(&gw.part, sizeof(double) * (size_t)w.nblk * kGcrM)) != hipSuccess ||
      (e = hipMalloc(&gw.inst, sizeof(SgInst) * w.nseg)) != hipSuccess ||
      (e = hipMalloc(&gw.gs, sizeof(SgGlobal))) != hipSuccess)
    return hmv::hip_fail(e, (std::string(who) + ": GCR workspace").c_str());
  return HMV_OK;
}

hipError_t sg_iteration(hmv_arena_s *a, const SegWork &w, SegGcrWork &gw, double *u) {
  const int nb = w.nblk;
  const int *skip = &gw.gs->all_done;
  // (r, r) partials at [2 nb, 3 nb) of the nb * kGcrM partials, apart from the
  // alpha pairs at [0, 2 nb) that sg_update_ur reads while writing them
  double *rpart = gw.part + 2 * (size_t)nb;
  hipLaunchKernelGGL(sg_update_ur, dim3(nb), dim3(kT), 0, w.s, w.tb, w.bstart, w.n, gw.p, gw.ap,
                     u, gw.r, (const double *)gw.part, rpart, gw.inst, (const SgGlobal *)gw.gs);
  hipLaunchKernelGGL(sg_final, dim3(1), dim3(kT), 0, w.s, w.nseg, w.bstart, (const double *)rpart,
                     gw.inst, gw.gs, 2, w.flags);
  hipError_t e = hmv::enqueue_apply(a, gw.r, gw.ar, 0.0, skip, w.s, w.flags);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(sg_dots_beta, dim3(nb, (gw.m + kGroup - 1) / kGroup), dim3(kT), 0, w.s, w.tb,
                     w.n, gw.ap, gw.ar, gw.part, (const SgInst *)gw.inst, (const SgGlobal *)gw.gs);
  hipLaunchKernelGGL(sg_final, dim3(1), dim3(kT), 0, w.s, w.nseg, w.bstart, (const double *)gw.part,
                     gw.inst, gw.gs, 3, w.flags);
  hipLaunchKernelGGL(sg_update_dir, dim3(nb), dim3(kT), 0, w.s, w.tb, w.n, gw.r, gw.ar, gw.p,
                     gw.ap, gw.part, (const SgInst *)gw.inst, (const SgGlobal *)gw.gs);
  return hipGetLastError();
}

// One GCR(m) per instance for A d = b from d = 0 (the caller zeroes d), in
// lockstep until every instance has stopped or maxit steps ran.  inst
// receives the instances' final states, iters the steps run.  The
// iteration graph is captured on the first call with a given d.
int sg_solve(hmv_arena_s *a, const SegWork &w, SegGcrWork &gw, const double *b, double *d,
             double eps, int maxit, int check_every, std::vector<SgInst> &inst, int &iters,
             const char *who) {
  const std::string name(who);
  SgGlobal init{};
  init.eps = eps;
  init.m = gw.m;
  init.maxit = maxit;
  const int g = grid_of(w.n);
  hipError_t e = hipMemcpyAsync(gw.gs, &init, sizeof(init), hipMemcpyHostToDevice, w.s);
  if (e != hipSuccess) return hmv::hip_fail(e, (name + ": setup").c_str());
  hipLaunchKernelGGL(sg_init, dim3(g), dim3(kT), 0, w.s, w.n, b, gw.r, gw.p);
  hipLaunchKernelGGL(ks_dots, dim3(w.nblk), dim3(kT), 0, w.s, w.tb, gw.r, gw.r, b, b, gw.part,
                     (const BlockState *)nullptr);
  hipLaunchKernelGGL(sg_final, dim3(1), dim3(kT), 0, w.s, w.nseg, w.bstart,
                     (const double *)gw.part, gw.inst, gw.gs, 0, w.flags);
  // ap_0 = A p_0 (the stopped instances' workgroups return) and the first
  // step's alpha partials
  if ((e = hmv::enqueue_apply(a, gw.p, gw.ap, 0.0, &gw.gs->all_done, w.s, w.flags)) != hipSuccess)
    return hmv::hip_fail(e, (name + ": setup").c_str());
  hipLaunchKernelGGL(sg_dots_alpha, dim3(w.nblk), dim3(kT), 0, w.s, w.tb, w.n, gw.ap, gw.r,
                     gw.part, (const SgGlobal *)gw.gs);
  if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, (name + ": setup").c_str());
  if (!gw.exec || gw.graph_u != d) {
    if (gw.exec) (void)hipGraphExecDestroy(gw.exec);
    if (gw.graph) (void)hipGraphDestroy(gw.graph);
    gw.exec = nullptr;
    gw.graph = nullptr;
    if ((e = hipStreamBeginCapture(w.s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return hmv::hip_fail(e, (name + ": capture").c_str());
    hipError_t ec = sg_iteration(a, w, gw, d);
    e = hipStreamEndCapture(w.s, &gw.graph);
    if (ec != hipSuccess) return hmv::hip_fail(ec, (name + ": capture").c_str());
    if (e != hipSuccess ||
        (e = hipGraphInstantiate(&gw.exec, gw.graph, nullptr, nullptr, 0)) != hipSuccess)
      return hmv::hip_fail(e, (name + ": graph").c_str());
    gw.graph_u = d;
  }
  SgGlobal host{};
  int launched = 0;
  for (;;) {
    if ((e = hipMemcpyAsync(&host, gw.gs, sizeof(host), hipMemcpyDeviceToHost, w.s)) != hipSuccess ||
        (e = hipStreamSynchronize(w.s)) != hipSuccess)
      return hmv::hip_fail(e, (name + ": state").c_str());
    if (host.all_done || launched >= maxit) break;
    const int burst = std::min(check_every, maxit - launched);
    for (int i = 0; i < burst; ++i)
      if ((e = hipGraphLaunch(gw.exec, w.s)) != hipSuccess)
        return hmv::hip_fail(e, (name + ": launch").c_str());
    launched += burst;
  }
  inst.resize(w.nseg);
  if ((e = hipMemcpyAsync(inst.data(), gw.inst, sizeof(SgInst) * w.nseg, hipMemcpyDeviceToHost,
                          w.s)) != hipSuccess ||
      (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, (name + ": state").c_str());
  iters = host.step;
  return HMV_OK;
}

}  // namespace

extern "C" int hmv_multi_gcr_ir(hmv_arena a64, hmv_arena alow, const double *b, double *u,
                                double eps, double inner_eps, int m, int max_outer, int max_inner,
                                int check_every, hmv_column_info *inst, double *inst_seconds,
                                hmv_solve_info *info) {
  const char *who = "hmv_multi_gcr_ir";
  if (!a64 || !alow || !b || !u || max_outer < 1 || max_inner < 1 || m < 1 || m > kGcrM ||
      a64->seg.size() < 2 || a64->seg != alow->seg || a64->nd != alow->nd ||
      !(inner_eps > 0.0 && inner_eps < 1.0)) {
    hmv::set_error(std::string(who) + ": invalid argument (two arenas of the same instances, "
                   "1 <= m <= 32, 0 < inner_eps < 1)");
    return HMV_ERR_ARG;
  }
  if (check_every < 1) check_every = 1;
  SegWork w;  // block table, stream and FP64 residuals
  std::vector<int> bstart;
  int rc = seg_setup(alow, w, bstart, who);
  if (rc) return rc;
  SegGcrWork gw;
  if ((rc = sg_setup(w, gw, m, who))) return rc;
  const int nseg = w.nseg;
  struct Buf {
    double *r2 = nullptr, *d = nullptr;
    int *stopped = nullptr;
    ~Buf() {
      (void)hipFree(r2);
      (void)hipFree(d);
      (void)hipFree(stopped);
    }
  } buf;
  const size_t nbytes = sizeof(double) * (size_t)w.n;
  hipError_t e;
  if ((e = hipMalloc(&buf.r2, nbytes)) != hipSuccess || (e = hipMalloc(&buf.d, nbytes)) != hipSuccess ||
      (e = hipMalloc(&buf.stopped, sizeof(int) * nseg)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_multi_gcr_ir: workspace");
  const int g = grid_of(w.n);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<int> status(nseg, HMV_SOLVE_MAXITER), stopped(nseg, 0), steps(nseg, 0);
  std::vector<double> rel(nseg, 0.0), done_at(nseg, -1.0), rr, bb;
  std::vector<SgInst> cs;
  int outer = 0, inner_total = 0;
  for (;;) {
    // FP64 residual of every instance on the FP64 arena
    if ((rc = seg_residuals(a64, w, b, u, buf.r2, bstart, rr, bb, who))) return rc;
    const double now = seconds_since(t0);
    bool all = true;
    for (int s = 0; s < nseg; ++s) {
      if (stopped[s]) continue;
      rel[s] = std::sqrt(rr[s]) / std::sqrt(bb[s]);
      if (!std::isfinite(rel[s]))
        status[s] = HMV_SOLVE_NONFINITE;
      else if (rel[s] < eps)
        status[s] = HMV_SOLVE_CONVERGED;
      else {
        all = false;
        continue;
      }
      stopped[s] = 1;
      done_at[s] = now;
    }
    if (all || outer >= max_outer) break;
    // inner solves A_low d = r on the instances still running (a stopped
    // instance's r is zero, so its inner solve stops at the setup)
    if ((e = hipMemcpyAsync(buf.stopped, stopped.data(), sizeof(int) * nseg, hipMemcpyHostToDevice,
                            w.s)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_multi_gcr_ir: flags");
    hipLaunchKernelGGL(ks_zero_stopped, dim3(w.nblk), dim3(kT), 0, w.s, w.tb, buf.r2, buf.stopped);
    hipLaunchKernelGGL(kb_fill, dim3(g), dim3(kT), 0, w.s, w.n, buf.d, 0.0);
    int iters = 0;
    if ((rc = sg_solve(alow, w, gw, buf.r2, buf.d, inner_eps, max_inner, check_every, cs, iters,
                       who)))
      return rc;
    inner_total += iters;
    for (int s = 0; s < nseg; ++s) {
      if (stopped[s]) continue;
      steps[s] += cs[s].step;
      if (cs[s].done && cs[s].status == HMV_SOLVE_NONFINITE) {
        // the correction of this instance is not usable: it stops here
        status[s] = HMV_SOLVE_NONFINITE;
        stopped[s] = 1;
        done_at[s] = seconds_since(t0);
      }
    }
    // u += d, with d = 0 on every stopped instance
    if ((e = hipMemcpyAsync(buf.stopped, stopped.data(), sizeof(int) * nseg, hipMemcpyHostToDevice,
                            w.s)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_multi_gcr_ir: flags");
    hipLaunchKernelGGL(ks_zero_stopped, dim3(w.nblk), dim3(kT), 0, w.s, w.tb, buf.d, buf.stopped);
    hipLaunchKernelGGL(kb_add, dim3(g), dim3(kT), 0, w.s, w.n, u, buf.d);
    if ((e = hipGetLastError()) != hipSuccess)
      return hmv::hip_fail(e, "hmv_multi_gcr_ir: update");
    ++outer;
  }
  if ((e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_multi_gcr_ir: sync");
  const double total = seconds_since(t0);
  int worst = HMV_SOLVE_CONVERGED;
  double maxrel = 0.0;
  for (int s = 0; s < nseg; ++s) {
    if (inst) {
      inst[s].nstp = steps[s];
      inst[s].status = status[s];
      inst[s].relres = rel[s];
      inst[s].true_relres = rel[s];
    }
    if (inst_seconds) inst_seconds[s] = done_at[s] >= 0.0 ? done_at[s] : total;
    if (std::isfinite(rel[s])) maxrel = std::max(maxrel, rel[s]);
    if (status[s] != HMV_SOLVE_CONVERGED) worst = status[s];
  }
  if (info) {
    info->nstp = inner_total;
    info->status = worst;
    info->relres = maxrel;
    info->true_relres = maxrel;
    info->seconds = total;
    info->outer = outer;
    info->inner_total = inner_total;
  }
  return HMV_OK;
}
