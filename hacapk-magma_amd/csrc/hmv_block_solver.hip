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

// per-column ||b - A u|| / ||b|| into rel[k] (uses w.r)
int true_relres(hmv_arena_s *a, BlockWork &w, const double *b, const double *u,
                std::vector<double> &rel) {
  const int g = grid_of(w.n), k = w.k;
  hipError_t e = hmv::enqueue_apply_block(a, k, u, w.r, 0.0, nullptr, w.s);
  if (e != hipSuccess) return hmv::hip_fail(e, "block true residual");
  hipLaunchKernelGGL(kb_residual, dim3(g), dim3(kT), 0, w.s, w.n, b, w.r);
  hipLaunchKernelGGL(kb_dots, dim3(g), dim3(kT), 0, w.s, w.n, k, w.r, w.r, b, b, w.part,
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

}  // namespace

extern "C" int hmv_block_bicgstab(hmv_arena a, int k, const double *b, double *u, double eps,
                                  int maxit, int check_every, hmv_column_info *cols,
                                  hmv_solve_info *info) {
  if (!a || !b || !u || k < 1 || k > HMV_MAX_RHS || (k & (k - 1)) != 0 || maxit < 0) {
    hmv::set_error("hmv_block_bicgstab: invalid argument (k must be a power of two <= 128)");
    return HMV_ERR_ARG;
  }
  int rc = hmv::ensure_tk(a, k);
  if (rc) return rc;
  if (check_every < 1) check_every = 1;
  BlockWork w;
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
    return hmv::hip_fail(e, "hmv_block_bicgstab: workspace");
  auto t0 = std::chrono::steady_clock::now();
  BlockState init{};
  init.eps = eps;
  init.maxit = maxit;
  if ((e = hipMemcpyAsync(w.bs, &init, sizeof(init), hipMemcpyHostToDevice, w.s)) != hipSuccess ||
      (e = hmv::enqueue_apply_block(a, k, u, w.r, 0.0, nullptr, w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_block_bicgstab: setup");
  hipLaunchKernelGGL(kb_init, dim3(g), dim3(kT), 0, w.s, w.n, b, w.r, w.rhat, w.p, w.v);
  hipLaunchKernelGGL(kb_dots, dim3(g), dim3(kT), 0, w.s, w.n, k, w.r, w.r, b, b, w.part,
                     (const BlockState *)nullptr);
  hipLaunchKernelGGL(kb_final, dim3(1), dim3(kT), 0, w.s, g, k, w.part, w.cs, w.bs, 0);
  if ((e = hipStreamSynchronize(w.s)) != hipSuccess ||
      (e = hipStreamBeginCapture(w.s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_block_bicgstab: capture");
  hipError_t ec = enqueue_iteration(a, w, u);
  e = hipStreamEndCapture(w.s, &w.graph);
  if (ec != hipSuccess) return hmv::hip_fail(ec, "hmv_block_bicgstab: capture");
  if (e != hipSuccess || (e = hipGraphInstantiate(&w.exec, w.graph, nullptr, nullptr, 0)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_block_bicgstab: graph");
  BlockState host{};
  int launched = 0;
  while (launched < maxit) {
    const int burst = std::min(check_every, maxit - launched);
    for (int i = 0; i < burst; ++i)
      if ((e = hipGraphLaunch(w.exec, w.s)) != hipSuccess)
        return hmv::hip_fail(e, "hmv_block_bicgstab: launch");
    launched += burst;
    if ((e = hipMemcpyAsync(&host, w.bs, sizeof(host), hipMemcpyDeviceToHost, w.s)) != hipSuccess ||
        (e = hipStreamSynchronize(w.s)) != hipSuccess)
      return hmv::hip_fail(e, "hmv_block_bicgstab: state");
    if (host.all_done) break;
  }
  if ((e = hipMemcpyAsync(&host, w.bs, sizeof(host), hipMemcpyDeviceToHost, w.s)) != hipSuccess ||
      (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_block_bicgstab: state");
  auto t1 = std::chrono::steady_clock::now();
  std::vector<ColState> cs(k);
  if ((e = hipMemcpy(cs.data(), w.cs, sizeof(ColState) * k, hipMemcpyDeviceToHost)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_block_bicgstab: columns");
  std::vector<double> rel;
  if ((rc = true_relres(a, w, b, u, rel)) != HMV_OK) return rc;
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
    info->nstp = host.iter;
    info->status = worst;
    info->relres = maxrel;
    info->true_relres = maxrel;
    info->seconds = std::chrono::duration<double>(t1 - t0).count();
    info->outer = 0;
    info->inner_total = maxstep;
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

}  // namespace

extern "C" int hmv_multi_bicgstab(hmv_arena a, const double *b, double *u, double eps, int maxit,
                                  int check_every, hmv_column_info *inst, double *inst_seconds,
                                  hmv_solve_info *info) {
  if (!a || !b || !u || maxit < 0 || a->seg.size() < 2) {
    hmv::set_error("hmv_multi_bicgstab: invalid argument");
    return HMV_ERR_ARG;
  }
  if (check_every < 1) check_every = 1;
  SegWork w;
  w.nseg = (int)a->seg.size() - 1;
  w.n = a->nd;
  // block table: chunks of at most 16 elements per thread, never across instances
  constexpr int kChunk = 16 * kT;
  std::vector<SegBlock> tb;
  std::vector<int> bstart(w.nseg + 1, 0);
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
    return hmv::hip_fail(e, "hmv_multi_bicgstab: workspace");
  auto t0 = std::chrono::steady_clock::now();
  BlockState init{};
  init.eps = eps;
  init.maxit = maxit;
  const int g = grid_of(w.n);
  if ((e = hipMemcpyAsync(w.bs, &init, sizeof(init), hipMemcpyHostToDevice, w.s)) != hipSuccess ||
      (e = hmv::enqueue_apply(a, u, w.r, 0.0, nullptr, w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_multi_bicgstab: setup");
  hipLaunchKernelGGL(kb_init, dim3(g), dim3(kT), 0, w.s, w.n, b, w.r, w.rhat, w.p, w.v);
  hipLaunchKernelGGL(ks_dots, dim3(w.nblk), dim3(kT), 0, w.s, w.tb, w.r, w.r, b, b, w.part,
                     (const BlockState *)nullptr);
  hipLaunchKernelGGL(ks_final, dim3(1), dim3(kT), 0, w.s, w.nseg, w.bstart, w.part, w.cs, w.bs, 0,
                     w.flags);
  if ((e = hipStreamSynchronize(w.s)) != hipSuccess ||
      (e = hipStreamBeginCapture(w.s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_multi_bicgstab: capture");
  hipError_t ec = seg_iteration(a, w, u);
  e = hipStreamEndCapture(w.s, &w.graph);
  if (ec != hipSuccess) return hmv::hip_fail(ec, "hmv_multi_bicgstab: capture");
  if (e != hipSuccess || (e = hipGraphInstantiate(&w.exec, w.graph, nullptr, nullptr, 0)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_multi_bicgstab: graph");
  // per-instance completion: the wall time of the burst in which it stopped
  std::vector<ColState> cs(w.nseg);
  std::vector<double> done_at(w.nseg, -1.0);
  auto poll = [&](double now) -> hipError_t {
    hipError_t pe = hipMemcpyAsync(cs.data(), w.cs, sizeof(ColState) * w.nseg,
                                   hipMemcpyDeviceToHost, w.s);
    if (pe == hipSuccess) pe = hipStreamSynchronize(w.s);
    if (pe == hipSuccess)
      for (int sgi = 0; sgi < w.nseg; ++sgi)
        if (cs[sgi].done && done_at[sgi] < 0.0) done_at[sgi] = now;
    return pe;
  };
  auto secs = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  if ((e = poll(secs())) != hipSuccess) return hmv::hip_fail(e, "hmv_multi_bicgstab: state");
  BlockState host{};
  int launched = 0;
  while (launched < maxit) {
    const int burst = std::min(check_every, maxit - launched);
    for (int i = 0; i < burst; ++i)
      if ((e = hipGraphLaunch(w.exec, w.s)) != hipSuccess)
        return hmv::hip_fail(e, "hmv_multi_bicgstab: launch");
    launched += burst;
    if ((e = hipMemcpyAsync(&host, w.bs, sizeof(host), hipMemcpyDeviceToHost, w.s)) != hipSuccess ||
        (e = poll(secs())) != hipSuccess)
      return hmv::hip_fail(e, "hmv_multi_bicgstab: state");
    if (host.all_done) break;
  }
  const double total = secs();
  // true residuals per instance: r = b - A u
  if ((e = hmv::enqueue_apply(a, u, w.r, 0.0, nullptr, w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_multi_bicgstab: true residual");
  hipLaunchKernelGGL(kb_residual, dim3(g), dim3(kT), 0, w.s, w.n, b, w.r);
  hipLaunchKernelGGL(ks_dots, dim3(w.nblk), dim3(kT), 0, w.s, w.tb, w.r, w.r, b, b, w.part,
                     (const BlockState *)nullptr);
  std::vector<double> part(2 * tb.size());
  if ((e = hipMemcpyAsync(part.data(), w.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                          w.s)) != hipSuccess ||
      (e = hipStreamSynchronize(w.s)) != hipSuccess)
    return hmv::hip_fail(e, "hmv_multi_bicgstab: true residual");
  int maxstep = 0, worst = HMV_SOLVE_CONVERGED;
  double maxrel = 0.0;
  for (int sgi = 0; sgi < w.nseg; ++sgi) {
    double rr = 0.0, bb = 0.0;
    for (int bi = bstart[sgi]; bi < bstart[sgi + 1]; ++bi) {
      rr += part[2 * bi];
      bb += part[2 * bi + 1];
    }
    const double rel = std::sqrt(rr) / std::sqrt(bb);
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
    info->nstp = host.iter;
    info->status = worst;
    info->relres = maxrel;
    info->true_relres = maxrel;
    info->seconds = total;
    info->outer = 0;
    info->inner_total = maxstep;
  }
  return HMV_OK;
}
