// hmv_api.cpp -- C ABI of the H-matvec engine: arena lifecycle, apply, and
// the reference-named entry points a Fortran driver links against.
#include <hip/hip_runtime.h>

#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hmv.h"
#include "hmv_engine.h"
#include "hmv_plan.h"

namespace hmv {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int hip_fail(hipError_t e, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return (e == hipErrorOutOfMemory) ? HMV_ERR_MEMORY : HMV_ERR_DEVICE;
}

// Host threads for the host-vector apply's copies.  The caller's arrays are
// pageable, so each vector crosses the CPU once (into or out of pinned
// staging); the copy is split over up to 8 threads (half the CPUs this
// process may use).  The workers sleep between calls and are joined at exit.
class HostPool {
 public:
  static HostPool &get() {
    static HostPool pool;
    return pool;
  }
  // fn(lo, hi) over [0, n) in nparts() contiguous parts; the caller runs
  // the first part itself.
  void parallel_for(size_t n, const std::function<void(size_t, size_t)> &fn) {
    const size_t parts = th_.size() + 1;
    if (parts == 1 || n < 4096) {
      fn(0, n);
      return;
    }
    std::lock_guard<std::mutex> one_caller(call_mu_);
    std::unique_lock<std::mutex> lk(mu_);
    fn_ = &fn;
    n_ = n;
    pending_ = (int)th_.size();
    ++gen_;
    lk.unlock();
    cv_.notify_all();
    fn(0, n / parts);
    lk.lock();
    done_cv_.wait(lk, [&] { return pending_ == 0; });
    fn_ = nullptr;
  }

 private:
  HostPool() {
    cpu_set_t set;
    int cpus = 1;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) cpus = CPU_COUNT(&set);
    const int workers = std::max(0, std::min(8, cpus / 2) - 1);
    for (int i = 0; i < workers; ++i) th_.emplace_back([this, i] { run(i + 1); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  void run(int part) {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      const std::function<void(size_t, size_t)> *fn = fn_;
      const size_t n = n_, parts = th_.size() + 1;
      lk.unlock();
      (*fn)(n * part / parts, n * (part + 1) / parts);
      lk.lock();
      if (--pending_ == 0) done_cv_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  std::mutex call_mu_, mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(size_t, size_t)> *fn_ = nullptr;
  size_t n_ = 0;
  int pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

int ensure_tk(hmv_arena_s *a, int k) {
  const int64_t need = std::max<int64_t>(1, a->tpart_size) * k;
  if (need <= a->tk_cap) return HMV_OK;
  (void)hipFree(a->d_tk);
  a->d_tk = nullptr;
  a->tk_cap = 0;
  hipError_t e = hipMalloc((void **)&a->d_tk, sizeof(double) * (size_t)need);
  if (e != hipSuccess) return hip_fail(e, "H-matmat t buffer");
  a->tk_cap = need;
  return HMV_OK;
}

void fill_units(const hmv_arena_s *a, HmvUnits &u) {
  for (int p = 0; p < 2; ++p)
    for (int w = 0; w < HMV_MM_WIDTHS; ++w) {
      u.u[p][w] = a->d_units[p][w];
      u.n[p][w] = a->nunits[p][w];
    }
  for (int c = 0; c < 4; ++c) {
    u.vec[c] = a->bt_vec[c];
    u.log_tw[c] = a->bt_log_tw[c];
  }
}

hipError_t enqueue_apply_block(hmv_arena_s *a, int k, const double *x, double *y, double beta,
                               const int *skip, hipStream_t s) {
  HmvUnits u;
  fill_units(a, u);
  return hmv_launch_block(&a->launch, &u, k, x, a->d_tk, y, beta, skip, s);
}

hipError_t enqueue_apply(hmv_arena_s *a, const double *x, double *y, double beta, const int *skip,
                         hipStream_t s, const int *inst_done) {
  return hmv_launch_apply(&a->launch, HMV_PHASE1 | HMV_PHASE2, x, y, beta, skip, inst_done, s);
}

}  // namespace hmv

using hmv::hip_fail;
using hmv::set_error;

namespace {

int check_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    set_error("no HIP device available (the engine has no CPU path)");
    return HMV_ERR_DEVICE;
  }
  return HMV_OK;
}

template <typename T>
hipError_t upload(T **dst, const std::vector<T> &src) {
  *dst = nullptr;
  if (src.empty()) return hipSuccess;
  hipError_t e = hipMalloc((void **)dst, sizeof(T) * src.size());
  if (e != hipSuccess) return e;
  return hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice);
}

void release(hmv_arena_s *a) {
  if (!a) return;
  hmv::free_solve_work(a);
  hmv::free_gcr_work(a);
  (void)hipFree(a->d_arena);
  (void)hipFree(a->d_work1);
  (void)hipFree(a->d_work2);
  (void)hipFree(a->d_tfin);
  (void)hipFree(a->d_winst1);
  (void)hipFree(a->d_winst2);
  for (auto &ph : a->d_units)
    for (int2 *u : ph) (void)hipFree(u);
  (void)hipFree(a->d_tk);
  (void)hipFree(a->d_zseg);
  (void)hipFree(a->d_outidx);
  (void)hipFree(a->d_codes);
  (void)hipFree(a->d_t);
  (void)hipFree(a->d_hx);
  (void)hipHostFree(a->h_x);
  (void)hipHostFree(a->h_y);
  if (a->hstream) (void)hipStreamDestroy(a->hstream);
  delete a;
}

hmv_tuning to_public(const hmv::PlanTuning &t) {
  hmv_tuning o{};
  o.p1_target_bytes = t.p1_target_bytes;
  o.p1_row_tile = t.p1_row_tile;
  o.p2_max_rows = t.p2_max_rows;
  o.p1_small_threads = t.p1_small_threads;
  o.p2_small_threads = t.p2_small_threads;
  return o;
}

// Public tuning (0 = the precision's default) -> plan tuning.
hmv::PlanTuning from_public(int prec, const hmv_tuning *t) {
  hmv::PlanTuning p = hmv::default_tuning(prec);
  if (!t) return p;
  if (t->p1_target_bytes != 0) p.p1_target_bytes = t->p1_target_bytes;
  if (t->p1_row_tile != 0) p.p1_row_tile = t->p1_row_tile;
  if (t->p2_max_rows != 0) p.p2_max_rows = t->p2_max_rows;
  if (t->p1_small_threads != 0) p.p1_small_threads = t->p1_small_threads;
  if (t->p2_small_threads != 0) p.p2_small_threads = t->p2_small_threads;
  return p;
}

int upload_plan(hmv::Plan &P, hmv_arena *out, std::chrono::steady_clock::time_point t0);

int build(int nd, int nlf, const void *leaves, int stride, int prec, const hmv::PlanTuning &tune,
          hmv_arena *out, const std::vector<int> *seg = nullptr) {
  if (!out) {
    set_error("hmv_arena_build: null output handle");
    return HMV_ERR_ARG;
  }
  *out = nullptr;
  int rc = check_device();
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  hmv::Plan P;
  std::string err;
  if (!hmv::build_plan(nd, nlf, leaves, stride, prec, tune, P, err)) {
    set_error("hmv_arena_build: " + err);
    return HMV_ERR_ARG;
  }
  if (seg) P.seg = *seg;
  return upload_plan(P, out, t0);
}

}  // namespace

namespace hmv {
int build_from_device_leaves(int nd, int nlf, const void *leaves, int stride, int prec,
                             const hmv_tuning *tuning, hmv_arena *out) {
  if (!out) {
    set_error("hmv_arena_build_dleaves: null output handle");
    return HMV_ERR_ARG;
  }
  *out = nullptr;
  int rc = check_device();
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  hmv::Plan P;
  std::string err;
  if (!hmv::build_plan_device(nd, nlf, leaves, stride, prec, from_public(prec, tuning), P, err)) {
    set_error("hmv_arena_build_dleaves: " + err);
    return HMV_ERR_ARG;
  }
  return upload_plan(P, out, t0);
}
}  // namespace hmv

namespace {

// Input code of the first column of item w: x index >= 0, or -1 - t slot.
int first_input(const hmv::Plan &P, const HmvWork &w) {
  if (w.zseg_begin >= w.zseg_end) return 0;
  const HmvZSeg &z = P.zseg[w.zseg_begin];
  if (z.kind == HMV_Z_CODES) return P.codes[z.src];
  return z.kind == HMV_Z_X ? z.src : -1 - z.src;
}

// H-matmat units of group width tw: (work item, up to tw consecutive
// 16-row tiles).  The items are sorted by locality (phase 1: first input
// column; phase 2: first row) and the units cut into G contiguous ranges,
// one per XCD; the list interleaves the ranges so that workgroup p, which
// the hardware sends to XCD p % G, runs a unit of range p % G.  Each XCD
// then walks neighbouring items together, and the k inputs of the columns
// they share stay in its L2.
std::vector<int2> matmat_units(const hmv::Plan &P, const std::vector<HmvWork> &wl, bool phase1,
                               int tw) {
  std::vector<std::pair<int64_t, int>> key(wl.size());
  for (size_t i = 0; i < wl.size(); ++i)
    key[i] = {phase1 ? (int64_t)first_input(P, wl[i]) : (int64_t)wl[i].out_base, (int)i};
  std::stable_sort(key.begin(), key.end(),
                   [](const std::pair<int64_t, int> &x, const std::pair<int64_t, int> &y) {
                     return x.first < y.first;
                   });
  std::vector<int2> sorted;
  for (const auto &kv : key) {
    const int T = (wl[kv.second].rows + 15) / 16;
    for (int m = 0; m < T; m += tw)
      sorted.push_back(make_int2(kv.second, m << 8 | std::min(tw, T - m)));
  }
  const size_t G = 8, n = sorted.size();  // MI355X: 8 XCDs
  const size_t q = n / G, r = n % G;      // the first r ranges take one extra unit
  std::vector<int2> out(n);
  size_t b = 0;
  for (size_t g = 0; g < G; ++g) {
    const size_t len = q + (g < r ? 1 : 0);
    for (size_t pos = 0; pos < len; ++pos) out[pos * G + g] = sorted[b + pos];
    b += len;
  }
  return out;
}

// Uploads a plan (built here or read from a file) as a device arena.
int upload_plan(hmv::Plan &P, hmv_arena *out, std::chrono::steady_clock::time_point t0) {
  const int nd = P.nd, prec = P.prec;
  const std::vector<int> *seg = &P.seg;
  hmv_arena_s *a = new hmv_arena_s();
  (void)hipGetDevice(&a->device);  // the current device (hmv_init succeeded)
  a->nd = nd;
  a->prec = prec;
  a->acc64 = prec == HMV_PREC_F64 ? 1 : 0;
  a->mode = prec;  // 0 FP64, 1 FP32, 2 BF16 U/V + FP32 A, 3 BF16 throughout
  a->tsize = P.tsize;
  a->tpart_size = P.tpart_size;
  a->nwork1 = (int)P.work1.size();
  a->nwork2 = (int)P.work2.size();
  a->ntfin = (int)P.tfin.size();
  hipError_t e;
  if ((e = hipMalloc(&a->d_arena, P.arena_size)) != hipSuccess ||
      (e = P.arena ? hipMemcpy(a->d_arena, P.arena.get(), P.arena_size, hipMemcpyHostToDevice)
                   : hipMemset(a->d_arena, 0, P.arena_size)) != hipSuccess ||
      (e = P.arena ? hipSuccess : hmv::launch_fill(P.fill, a->d_arena)) != hipSuccess ||
      (e = upload(&a->d_work1, P.work1)) != hipSuccess ||
      (e = upload(&a->d_work2, P.work2)) != hipSuccess ||
      (e = upload(&a->d_tfin, P.tfin)) != hipSuccess ||
      (e = upload(&a->d_zseg, P.zseg)) != hipSuccess ||
      (e = upload(&a->d_outidx, P.outidx)) != hipSuccess ||
      (e = upload(&a->d_codes, P.codes)) != hipSuccess) {
    release(a);
    return hip_fail(e, "hmv_arena_build: upload");
  }
  // multi-instance arenas: the instance of every work item (phase 1 by its
  // first input column, phase 2 by its first row; blocks never cross instances)
  if (seg && seg->size() > 2) {
    auto inst_of = [&](int row) {
      return (int)(std::upper_bound(seg->begin(), seg->end(), row) - seg->begin()) - 1;
    };
    std::vector<int> w1(P.work1.size()), w2(P.work2.size());
    for (size_t i = 0; i < P.work1.size(); ++i) w1[i] = inst_of(first_input(P, P.work1[i]));
    for (size_t i = 0; i < P.work2.size(); ++i) w2[i] = inst_of(P.work2[i].out_base);
    if ((e = upload(&a->d_winst1, w1)) != hipSuccess ||
        (e = upload(&a->d_winst2, w2)) != hipSuccess) {
      release(a);
      return hip_fail(e, "hmv_arena_build: instance tables");
    }
  }
  for (int ph = 0; ph < 2; ++ph)
    for (int w = 0; w < HMV_MM_WIDTHS; ++w) {
      const std::vector<int2> u = matmat_units(P, ph == 0 ? P.work1 : P.work2, ph == 0, 1 << w);
      a->nunits[ph][w] = (int)u.size();
      if ((e = upload(&a->d_units[ph][w], u)) != hipSuccess) {
        release(a);
        return hip_fail(e, "hmv_arena_build: H-matmat unit upload");
      }
    }
  size_t tb = (size_t)std::max<int64_t>(1, P.tpart_size) * sizeof(double);
  if ((e = hipMalloc((void **)&a->d_t, tb)) != hipSuccess) {
    release(a);
    return hip_fail(e, "hmv_arena_build: t vector");
  }
  if ((e = hipDeviceSynchronize()) != hipSuccess) {
    release(a);
    return hip_fail(e, "hmv_arena_build: sync");
  }
  HmvLaunch &L = a->launch;
  L.mode = a->mode;
  L.arena = a->d_arena;
  L.work1 = a->d_work1;
  L.nwork1 = a->nwork1;
  L.nwork1_small = P.nwork1_small;
  L.nt1_small = P.nt1_small;
  L.tfin = a->d_tfin;
  L.ntfin = a->ntfin;
  L.work2 = a->d_work2;
  L.nwork2 = a->nwork2;
  L.nwork2_small = P.nwork2_small;
  L.nt2_small = P.nt2_small;
  L.zseg = a->d_zseg;
  L.outidx = a->d_outidx;
  L.codes = a->d_codes;
  L.t = a->d_t;
  L.winst1 = a->d_winst1;
  L.winst2 = a->d_winst2;
  a->tuning.p1_target_bytes = P.tune.p1_target_bytes;
  a->tuning.p1_row_tile = P.tune.p1_row_tile;
  a->tuning.p2_max_rows = P.tune.p2_max_rows;
  a->tuning.p1_small_threads = P.tune.p1_small_threads;
  a->tuning.p2_small_threads = P.tune.p2_small_threads;
  a->seg = P.seg;
  a->nzseg = (int64_t)P.zseg.size();
  a->noutidx = (int64_t)P.outidx.size();
  a->ncodes = (int64_t)P.codes.size();
  a->arena_size = P.arena_size;
  a->nwork1_small = P.nwork1_small;
  a->nt1_small = P.nt1_small;
  a->nwork2_small = P.nwork2_small;
  a->nt2_small = P.nt2_small;
  a->pstats = P.st;
  auto t1 = std::chrono::steady_clock::now();
  hmv_arena_stats &s = a->stats;
  s.nd = nd;
  s.precision = prec;
  s.nlf = P.st.nlf;
  s.nlr = P.st.nlr;
  s.ndense = P.st.ndense;
  s.entries_lr = P.st.entries_lr;
  s.entries_dense = P.st.entries_dense;
  s.tslots = P.st.tslots;
  s.tparts = P.st.tparts;
  s.phase1_bytes = P.st.phase1_bytes;
  s.phase2_bytes = P.st.phase2_bytes;
  s.arena_bytes = P.st.arena_bytes;
  s.payload_bytes = P.st.payload_bytes;
  s.alg_bytes = P.st.alg_bytes;
  s.col_groups = P.st.col_groups;
  s.row_groups = P.st.row_groups;
  s.atoms = P.st.atoms;
  s.nwork_phase1 = a->nwork1;
  s.nwork_phase2 = a->nwork2;
  s.ktmax = P.st.ktmax;
  s.build_seconds = std::chrono::duration<double>(t1 - t0).count();
  *out = a;
  return HMV_OK;
}

// One block-diagonal leaf list of ninst instances: instance i's rows and
// columns shifted by the sizes of the instances before it; seg receives the
// ninst + 1 row offsets.  False (with the error set) on invalid input.
bool concat_instances(const char *who, int ninst, const int *nd, const int *nlf,
                      const void *const *leaves, int leaf_stride, std::vector<int> &seg,
                      std::vector<hmv::RefLeaf> &all) {
  if (ninst < 1 || !nd || !nlf || !leaves || leaf_stride < (int)sizeof(hmv::RefLeaf)) {
    set_error(std::string(who) + ": invalid argument");
    return false;
  }
  seg.assign(ninst + 1, 0);
  int64_t total_lf = 0;
  for (int i = 0; i < ninst; ++i) {
    if (nd[i] < 1 || nlf[i] < 1 || !leaves[i]) {
      set_error(std::string(who) + ": empty instance " + std::to_string(i));
      return false;
    }
    if ((int64_t)seg[i] + nd[i] > INT32_MAX) {
      set_error(std::string(who) + ": more than 2^31 unknowns");
      return false;
    }
    seg[i + 1] = seg[i] + nd[i];
    total_lf += nlf[i];
  }
  if (total_lf > INT32_MAX) {
    set_error(std::string(who) + ": too many leaves");
    return false;
  }
  all.clear();
  all.reserve((size_t)total_lf);
  for (int i = 0; i < ninst; ++i) {
    const char *p = static_cast<const char *>(leaves[i]);
    for (int l = 0; l < nlf[i]; ++l) {
      hmv::RefLeaf r;
      std::memcpy(&r, p + (size_t)l * leaf_stride, sizeof(r));
      if (r.nstrtl < 1 || r.nstrtt < 1 || r.nstrtl - 1 + r.ndl > nd[i] ||
          r.nstrtt - 1 + r.ndt > nd[i]) {
        set_error(std::string(who) + ": leaf outside its instance " + std::to_string(i));
        return false;
      }
      r.nstrtl += seg[i];
      r.nstrtt += seg[i];
      all.push_back(r);
    }
  }
  return true;
}

// Coordinate search of the repack tunables (small-workgroup threads of each
// phase, phase-1 bytes per workgroup, phase-2 rows, phase-1 row tile), one
// at a time, each candidate an arena built by make(tuning, &arena) and timed
// over `reps` applies (HIP events); a candidate must beat the current best
// by more than 1% (timing noise) to replace it.
template <typename Make>
int autotune_search(const char *who, int nd, int precision, int reps, Make make, hmv_tuning *best,
                    double *best_ms, double *default_ms) {
  int rc = check_device();
  if (rc) return rc;
  if (reps < 3) reps = 3;
  const size_t nb = sizeof(double) * (size_t)nd;
  double *dx = nullptr, *dy = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e;
  if ((e = hipMalloc(&dx, nb)) != hipSuccess || (e = hipMalloc(&dy, nb)) != hipSuccess ||
      (e = hipMemset(dx, 0, nb)) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
    rc = hip_fail(e, (std::string(who) + ": setup").c_str());
  }
  // time per apply of an arena built with t (the arena lives only for the measurement)
  auto measure = [&](const hmv::PlanTuning &t, double *ms) -> int {
    hmv_arena a = nullptr;
    int r = make(t, &a);
    if (r) return r;
    hipError_t he = hipSuccess;
    for (int i = 0; i < 2 && he == hipSuccess; ++i) he = hmv::enqueue_apply(a, dx, dy, 0.0, nullptr, s);
    if (he == hipSuccess) he = hipEventRecord(e0, s);
    for (int i = 0; i < reps && he == hipSuccess; ++i)
      he = hmv::enqueue_apply(a, dx, dy, 0.0, nullptr, s);
    if (he == hipSuccess) he = hipEventRecord(e1, s);
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    float m = 0.f;
    if (he == hipSuccess) he = hipEventElapsedTime(&m, e0, e1);
    release(a);
    if (he != hipSuccess) return hip_fail(he, (std::string(who) + ": timing").c_str());
    *ms = (double)m / reps;
    return HMV_OK;
  };
  auto same = [](const hmv::PlanTuning &a, const hmv::PlanTuning &b) {
    return a.p1_target_bytes == b.p1_target_bytes && a.p1_row_tile == b.p1_row_tile &&
           a.p2_max_rows == b.p2_max_rows && a.p1_small_threads == b.p1_small_threads &&
           a.p2_small_threads == b.p2_small_threads;
  };
  hmv::PlanTuning cur = hmv::default_tuning(precision);
  double cur_ms = 0.0;
  if (!rc) rc = measure(cur, &cur_ms);
  if (!rc && default_ms) *default_ms = cur_ms;
  auto search = [&](auto set, std::initializer_list<int64_t> values) {
    for (int64_t v : values) {
      if (rc) return;
      hmv::PlanTuning t = cur;
      set(t, v);
      if (same(t, cur)) continue;
      double ms = 0.0;
      rc = measure(t, &ms);
      if (!rc && ms < 0.99 * cur_ms) {
        cur = t;
        cur_ms = ms;
      }
    }
  };
  search([](hmv::PlanTuning &t, int64_t v) { t.p1_small_threads = (int)v; }, {-1, 64, 128});
  search([](hmv::PlanTuning &t, int64_t v) { t.p2_small_threads = (int)v; }, {-1, 64, 128});
  search([](hmv::PlanTuning &t, int64_t v) { t.p1_target_bytes = v; }, {48 << 10, 96 << 10, 192 << 10});
  search([](hmv::PlanTuning &t, int64_t v) { t.p2_max_rows = (int)v; }, {128, 256});
  search([](hmv::PlanTuning &t, int64_t v) { t.p1_row_tile = (int)v; }, {128, 256});
  if (e1) (void)hipEventDestroy(e1);
  if (e0) (void)hipEventDestroy(e0);
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(dx);
  (void)hipFree(dy);
  if (rc) return rc;
  *best = to_public(cur);
  if (best_ms) *best_ms = cur_ms;
  return HMV_OK;
}

}  // namespace

extern "C" {

const char *hmv_last_error(void) { return hmv::g_err.c_str(); }

int hmv_init(int device) {
  int rc = check_device();
  if (rc) return rc;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hmv_init");
  // the static workspace's device address, looked up once per device here
  // (hmv_apply_block_static looks it up itself if hmv_init was not called)
  double *scratch = nullptr;
  return hmv::static_scratch(&scratch);
}

int hmv_arena_build(int nd, int nlf, const void *leaves, int leaf_stride, int precision,
                    hmv_arena *out) {
  int rc = build(nd, nlf, leaves, leaf_stride, precision, from_public(precision, nullptr), out);
  if (!rc) (*out)->seg = {0, nd};
  return rc;
}

int hmv_arena_build_multi(int ninst, const int *nd, const int *nlf, const void *const *leaves,
                          int leaf_stride, int precision, const hmv_tuning *tuning,
                          hmv_arena *out) {
  if (!out) {
    set_error("hmv_arena_build_multi: null output handle");
    return HMV_ERR_ARG;
  }
  std::vector<int> seg;
  std::vector<hmv::RefLeaf> all;
  if (!concat_instances("hmv_arena_build_multi", ninst, nd, nlf, leaves, leaf_stride, seg, all))
    return HMV_ERR_ARG;
  return build(seg[ninst], (int)all.size(), all.data(), (int)sizeof(hmv::RefLeaf), precision,
               from_public(precision, tuning), out, &seg);
}

int hmv_arena_segments(hmv_arena a, int *offsets, int cap) {
  if (!a) {
    set_error("hmv_arena_segments: null arena");
    return HMV_ERR_ARG;
  }
  const int n = (int)a->seg.size() - 1;
  if (offsets)
    for (int i = 0; i <= n && i < cap; ++i) offsets[i] = a->seg[i];
  return n;
}

int hmv_default_tuning(int precision, hmv_tuning *t) {
  if (!t) {
    set_error("hmv_default_tuning: null argument");
    return HMV_ERR_ARG;
  }
  *t = to_public(hmv::default_tuning(precision));
  return HMV_OK;
}

int hmv_arena_build_tuned(int nd, int nlf, const void *leaves, int leaf_stride, int precision,
                          const hmv_tuning *tuning, hmv_arena *out) {
  int rc = build(nd, nlf, leaves, leaf_stride, precision, from_public(precision, tuning), out);
  if (!rc) (*out)->seg = {0, nd};
  return rc;
}

int hmv_arena_tuning_get(hmv_arena a, hmv_tuning *t) {
  if (!a || !t) {
    set_error("hmv_arena_tuning_get: null argument");
    return HMV_ERR_ARG;
  }
  *t = a->tuning;
  return HMV_OK;
}

int hmv_arena_stats_get(hmv_arena a, hmv_arena_stats *s) {
  if (!a || !s) {
    set_error("hmv_arena_stats_get: null argument");
    return HMV_ERR_ARG;
  }
  *s = a->stats;
  return HMV_OK;
}

int hmv_arena_free(hmv_arena a) {
  if (!a) return HMV_OK;
  release(a);
  return HMV_OK;
}

int64_t hmv_arena_tbytes(hmv_arena a) {
  if (!a) return -1;
  return std::max<int64_t>(1, a->tpart_size) * (int64_t)sizeof(double);
}

int hmv_apply(hmv_arena a, const double *x, double *y, double beta, void *stream) {
  if (!a || !x || !y) {
    set_error("hmv_apply: null argument");
    return HMV_ERR_ARG;
  }
  if (beta != 0.0 && beta != 1.0) {
    set_error("hmv_apply: beta must be 0 or 1");
    return HMV_ERR_ARG;
  }
  hipError_t e = hmv::enqueue_apply(a, x, y, beta, nullptr, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "hmv_apply: launch");
  return HMV_OK;
}

int hmv_apply_block(hmv_arena a, int k, const double *x, double *y, double beta, void *stream) {
  if (!a || !x || !y || k < 1 || k > HMV_MAX_RHS) {
    set_error("hmv_apply_block: invalid argument (k must be 1..128)");
    return HMV_ERR_ARG;
  }
  int rc = hmv::ensure_tk(a, k);
  if (rc) return rc;
  hipError_t e = hmv::enqueue_apply_block(a, k, x, y, beta, nullptr, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "hmv_apply_block");
  return HMV_OK;
}

int hmv_apply_block_timed(hmv_arena a, int k, const double *x, double *y, int reps, double *ms) {
  if (!a || !x || !y || k < 1 || k > HMV_MAX_RHS || reps < 1) {
    set_error("hmv_apply_block_timed: invalid argument");
    return HMV_ERR_ARG;
  }
  int rc = hmv::ensure_tk(a, k);
  if (rc) return rc;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hmv::enqueue_apply_block(a, k, x, y, 0.0, nullptr, s);
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int r = 0; r < reps && e == hipSuccess; ++r)
    e = hmv::enqueue_apply_block(a, k, x, y, 0.0, nullptr, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float m = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&m, e0, e1);
  if (e1) (void)hipEventDestroy(e1);
  if (e0) (void)hipEventDestroy(e0);
  if (s) (void)hipStreamDestroy(s);
  if (e != hipSuccess) return hip_fail(e, "hmv_apply_block_timed");
  if (ms) *ms = (double)m / reps;
  return HMV_OK;
}
}  // extern "C"

namespace {
// right-hand-side classes of the H-matmat path choices
int rhs_class(int k) { return k <= 1 ? 0 : k <= 2 ? 1 : k <= 4 ? 2 : 3; }
int tile_class(int k) {
  const int kt = (k + 15) / 16;
  return kt == 1 ? 0 : kt == 2 ? 1 : kt <= 4 ? 2 : 3;
}
int vec_limit(const hmv_arena_s *a) {
  return a->mode == 0 ? HMV_KSTREAM_MAX_F64 : HMV_KSTREAM_MAX_NARROW;
}
// tiles per unit instantiated for a tile class, the default first (the
// MFMA kernel's dispatch in hmv_matmat.hip)
std::vector<int> tw_choices(const hmv_arena_s *a, int tc) {
  if (tc == 0) return {4, 2};
  if (tc < 3) return {2, 1};
  return a->mode == 0 ? std::vector<int>{1, 2} : std::vector<int>{2, 1};  // 5-8 tiles: FP64 1, else 2
}
int tw_of(const hmv_arena_s *a, int tc) {
  return a->bt_log_tw[tc] < 0 ? tw_choices(a, tc)[0] : 1 << a->bt_log_tw[tc];
}
bool block_tuning_args(hmv_arena a, int k, const void *t, const char *who) {
  if (!a || !t || k < 1 || k > HMV_MAX_RHS) {
    set_error(std::string(who) + ": invalid argument (k must be 1..128)");
    return false;
  }
  return true;
}
}  // namespace

extern "C" {
int hmv_block_tuning_get(hmv_arena a, int k, hmv_block_tuning *t) {
  if (!block_tuning_args(a, k, t, "hmv_block_tuning_get")) return HMV_ERR_ARG;
  t->vector = k <= vec_limit(a) && a->bt_vec[rhs_class(k)] ? 1 : 0;
  t->tiles_per_unit = tw_of(a, tile_class(k));
  return HMV_OK;
}

int hmv_block_tuning_set(hmv_arena a, int k, const hmv_block_tuning *t) {
  if (!block_tuning_args(a, k, t, "hmv_block_tuning_set")) return HMV_ERR_ARG;
  if (t->vector && k > vec_limit(a)) {
    set_error("hmv_block_tuning_set: the vector path takes k <= " + std::to_string(vec_limit(a)) +
              " on this arena");
    return HMV_ERR_ARG;
  }
  const int tc = tile_class(k);
  const std::vector<int> ch = tw_choices(a, tc);
  if (t->tiles_per_unit != 0 && std::find(ch.begin(), ch.end(), t->tiles_per_unit) == ch.end()) {
    set_error("hmv_block_tuning_set: tiles_per_unit must be 0, " + std::to_string(ch[0]) + " or " +
              std::to_string(ch[1]) + " for k = " + std::to_string(k));
    return HMV_ERR_ARG;
  }
  if (k <= vec_limit(a)) a->bt_vec[rhs_class(k)] = t->vector ? 1 : 0;
  a->bt_log_tw[tc] = t->tiles_per_unit == 0 ? -1 : t->tiles_per_unit == 4 ? 2 : t->tiles_per_unit / 2;
  return HMV_OK;
}

int hmv_autotune_block(hmv_arena a, int k, int reps, hmv_block_tuning *best, double *best_ms,
                       double *default_ms) {
  if (!block_tuning_args(a, k, best, "hmv_autotune_block")) return HMV_ERR_ARG;
  if (reps < 1) {
    set_error("hmv_autotune_block: reps must be >= 1");
    return HMV_ERR_ARG;
  }
  const size_t count = (size_t)a->nd * k, bytes = sizeof(double) * count;
  // inputs uniform in [-1, 1) (a 64-bit LCG): all-zero inputs ran ~10%
  // faster than real data and would understate the times
  std::vector<double> hx(count);
  uint64_t st = 0x9e3779b97f4a7c15ull;
  for (double &v : hx) {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    v = (double)(st >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
  double *x = nullptr, *y = nullptr;
  hipError_t e;
  if ((e = hipMalloc(&x, bytes)) != hipSuccess || (e = hipMalloc(&y, bytes)) != hipSuccess ||
      (e = hipMemcpy(x, hx.data(), bytes, hipMemcpyHostToDevice)) != hipSuccess) {
    (void)hipFree(x);
    (void)hipFree(y);
    return hip_fail(e, "hmv_autotune_block: buffers");
  }
  // candidates: the defaults first, then every other choice of k's classes
  const int tc = tile_class(k);
  const std::vector<int> ch = tw_choices(a, tc);
  std::vector<hmv_block_tuning> cand;
  const bool vec_ok = k <= vec_limit(a);
  cand.push_back(hmv_block_tuning{vec_ok ? 1 : 0, ch[0]});
  if (vec_ok) cand.push_back(hmv_block_tuning{0, ch[0]});
  cand.push_back(hmv_block_tuning{0, ch[1]});
  int rc = HMV_OK, bi = -1;
  double bms = 0.0, dms = 0.0;
  for (size_t i = 0; i < cand.size() && rc == HMV_OK; ++i) {
    double ms = 0.0;
    if ((rc = hmv_block_tuning_set(a, k, &cand[i])) != HMV_OK ||
        (rc = hmv_apply_block_timed(a, k, x, y, reps, &ms)) != HMV_OK)
      break;
    if (i == 0) dms = ms;
    if (bi < 0 || ms < bms) {
      bi = (int)i;
      bms = ms;
    }
  }
  (void)hipFree(x);
  (void)hipFree(y);
  if (rc != HMV_OK) return rc;
  if ((rc = hmv_block_tuning_set(a, k, &cand[bi])) != HMV_OK) return rc;
  *best = cand[bi];
  if (best_ms) *best_ms = bms;
  if (default_ms) *default_ms = dms;
  return HMV_OK;
}

int hmv_apply_host(hmv_arena a, const double *zu, double *zau) {
  if (!a || !zu || !zau) {
    set_error("hmv_apply_host: null argument");
    return HMV_ERR_ARG;
  }
  hipError_t e;
  const size_t n = (size_t)a->nd, nb = sizeof(double) * n;
  if (!a->hstream) {  // created last: set only when every buffer exists
    if ((e = hipMalloc(&a->d_hx, nb)) != hipSuccess ||
        (e = hipHostMalloc((void **)&a->h_x, nb, hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostMalloc((void **)&a->h_y, nb, hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostGetDevicePointer((void **)&a->h_y_dev, a->h_y, 0)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&a->hstream, hipStreamNonBlocking)) != hipSuccess) {
      // leave nothing half set up: the next call starts again
      (void)hipFree(a->d_hx);
      (void)hipHostFree(a->h_x);
      (void)hipHostFree(a->h_y);
      a->d_hx = a->h_x = a->h_y = a->h_y_dev = nullptr;
      a->hstream = nullptr;
      return hip_fail(e, "hmv_apply_host: setup");
    }
  }
  // one copy over PCIe, from pinned staging: zu up; the apply's phase 2
  // writes H zu straight into pinned host memory as its rows finish (0.48
  // against 0.52 ms with a download after the apply at 163,840 unknowns).
  // The CPU's passes over the caller's arrays (zu into staging, zau += H zu
  // out of it) are split over the pool's threads.
  hmv::HostPool &pool = hmv::HostPool::get();
  double *hx = a->h_x, *hy = a->h_y;
  pool.parallel_for(n, [&](size_t lo, size_t hi) {
    std::memcpy(hx + lo, zu + lo, sizeof(double) * (hi - lo));
  });
  if ((e = hipMemcpyAsync(a->d_hx, hx, nb, hipMemcpyHostToDevice, a->hstream)) != hipSuccess)
    return hip_fail(e, "hmv_apply_host: copy x");
  if ((e = hmv::enqueue_apply(a, a->d_hx, a->h_y_dev, 0.0, nullptr, a->hstream)) != hipSuccess)
    return hip_fail(e, "hmv_apply_host: launch");
  if ((e = hipStreamSynchronize(a->hstream)) != hipSuccess)
    return hip_fail(e, "hmv_apply_host: sync");
  pool.parallel_for(n, [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) zau[i] += hy[i];
  });
  return HMV_OK;
}

int hmv_device_malloc(void **p, size_t bytes) {
  if (!p) {
    set_error("hmv_device_malloc: null pointer");
    return HMV_ERR_ARG;
  }
  *p = nullptr;
  int rc = check_device();
  if (rc) return rc;
  hipError_t e = hipMalloc(p, bytes ? bytes : 1);
  if (e != hipSuccess) return hip_fail(e, "hmv_device_malloc");
  return HMV_OK;
}

int hmv_device_free(void *p) {
  hipError_t e = hipFree(p);
  if (e != hipSuccess) return hip_fail(e, "hmv_device_free");
  return HMV_OK;
}

int hmv_memcpy_htod(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "hmv_memcpy_htod");
  return HMV_OK;
}

int hmv_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hmv_memcpy_dtoh");
  return HMV_OK;
}

int hmv_device_synchronize(void) {
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "hmv_device_synchronize");
  return HMV_OK;
}

int hmv_device_info(char *name, int name_len, int *compute_units) {
  int rc = check_device();
  if (rc) return rc;
  int dev = 0;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return hip_fail(e, "hmv_device_info");
  if (name && name_len > 0) {
    std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (compute_units) *compute_units = prop.multiProcessorCount;
  return HMV_OK;
}

int hmv_apply_timed(hmv_arena a, const double *x, double *y, int reps, double *ms_phase1,
                    double *ms_phase2, double *ms_apply) {
  if (!a || !x || !y || reps < 1) {
    set_error("hmv_apply_timed: invalid argument");
    return HMV_ERR_ARG;
  }
  hipStream_t s = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "hmv_apply_timed: stream");
  std::vector<hipEvent_t> ev(3 * (size_t)reps, nullptr);
  for (auto &v : ev)
    if ((e = hipEventCreate(&v)) != hipSuccess) break;
  double t1 = 0.0, t2 = 0.0, ta = 0.0;
  if (e == hipSuccess) {
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
      hipEvent_t e0 = ev[3 * r], e1 = ev[3 * r + 1], e2 = ev[3 * r + 2];
      e = hipEventRecord(e0, s);
      // phase 1 = t (rank-vector products and the chunk-partial sums)
      if (e == hipSuccess && a->nwork1 > 0)
        e = hmv_launch_apply(&a->launch, HMV_PHASE1, x, y, 0.0, nullptr, nullptr, s);
      if (e == hipSuccess) e = hipEventRecord(e1, s);
      if (e == hipSuccess) e = hmv_launch_apply(&a->launch, HMV_PHASE2, x, y, 0.0, nullptr, nullptr, s);
      if (e == hipSuccess) e = hipEventRecord(e2, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
      float m1 = 0.f, m2 = 0.f;
      e = hipEventElapsedTime(&m1, ev[3 * r], ev[3 * r + 1]);
      if (e == hipSuccess) e = hipEventElapsedTime(&m2, ev[3 * r + 1], ev[3 * r + 2]);
      t1 += m1;
      t2 += m2;
      ta += (double)m1 + (double)m2;
    }
  }
  for (auto &v : ev)
    if (v) (void)hipEventDestroy(v);
  (void)hipStreamDestroy(s);
  if (e != hipSuccess) return hip_fail(e, "hmv_apply_timed");
  if (ms_phase1) *ms_phase1 = a->nwork1 > 0 ? t1 / reps : 0.0;
  if (ms_phase2) *ms_phase2 = t2 / reps;
  if (ms_apply) *ms_apply = ta / reps;
  return HMV_OK;
}

int hmv_autotune(int nd, int nlf, const void *leaves, int leaf_stride, int precision, int reps,
                 hmv_tuning *best, double *best_ms, double *default_ms) {
  if (!leaves || !best || nd <= 0 || nlf <= 0) {
    set_error("hmv_autotune: invalid argument");
    return HMV_ERR_ARG;
  }
  auto make = [&](const hmv::PlanTuning &t, hmv_arena *a) {
    return build(nd, nlf, leaves, leaf_stride, precision, t, a);
  };
  return autotune_search("hmv_autotune", nd, precision, reps, make, best, best_ms, default_ms);
}

int hmv_autotune_multi(int ninst, const int *nd, const int *nlf, const void *const *leaves,
                       int leaf_stride, int precision, int reps, hmv_tuning *best, double *best_ms,
                       double *default_ms) {
  if (!best) {
    set_error("hmv_autotune_multi: invalid argument");
    return HMV_ERR_ARG;
  }
  std::vector<int> seg;
  std::vector<hmv::RefLeaf> all;
  if (!concat_instances("hmv_autotune_multi", ninst, nd, nlf, leaves, leaf_stride, seg, all))
    return HMV_ERR_ARG;
  auto make = [&](const hmv::PlanTuning &t, hmv_arena *a) {
    return build(seg[ninst], (int)all.size(), all.data(), (int)sizeof(hmv::RefLeaf), precision, t,
                 a, &seg);
  };
  return autotune_search("hmv_autotune_multi", seg[ninst], precision, reps, make, best, best_ms,
                         default_ms);
}

int hmv_stream_test(int kind, const void *src, void *dst, size_t bytes, int reps, double *ms) {
  if (kind < 0 || kind > 2 || (kind != 2 && !src) || (kind != 0 && !dst) || reps < 1 ||
      bytes < 16 || (kind == 2 && (bytes % 8 != 0 || bytes / 8 > (size_t)1 << 24))) {
    set_error("hmv_stream_test: invalid argument");
    return HMV_ERR_ARG;
  }
  hipStream_t s = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "hmv_stream_test: stream");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if ((e = hipEventCreate(&e0)) == hipSuccess && (e = hipEventCreate(&e1)) == hipSuccess)
    e = hmv_launch_stream_test(kind, src, dst, bytes, s);  // warm
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = hmv_launch_stream_test(kind, src, dst, bytes, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  if (e != hipSuccess) return hip_fail(e, "hmv_stream_test");
  if (ms) *ms = (double)t / reps;
  return HMV_OK;
}

struct hmv_plan_s {
  hmv::Plan P;
};

int hmv_plan_create_tuned(int nd, int nlf, const void *leaves, int leaf_stride, int precision,
                          const hmv_tuning *tuning, hmv_plan *out) {
  if (!out) {
    set_error("hmv_plan_create: null output");
    return HMV_ERR_ARG;
  }
  *out = nullptr;
  hmv_plan p = new hmv_plan_s();
  std::string err;
  if (!hmv::build_plan(nd, nlf, leaves, leaf_stride, precision, from_public(precision, tuning),
                       p->P, err)) {
    delete p;
    set_error("hmv_plan_create: " + err);
    return HMV_ERR_ARG;
  }
  *out = p;
  return HMV_OK;
}

int hmv_plan_save(hmv_plan p, const char *path) {
  if (!p || !path) {
    set_error("hmv_plan_save: null argument");
    return HMV_ERR_ARG;
  }
  std::string err;
  if (!hmv::save_plan_file(p->P, path, err)) {
    set_error("hmv_plan_save: " + err);
    return HMV_ERR_ARG;
  }
  return HMV_OK;
}

int hmv_plan_load(const char *path, hmv_plan *out) {
  if (!path || !out) {
    set_error("hmv_plan_load: null argument");
    return HMV_ERR_ARG;
  }
  *out = nullptr;
  hmv_plan p = new hmv_plan_s();
  std::string err;
  if (!hmv::load_plan_file(path, p->P, err)) {
    delete p;
    set_error("hmv_plan_load: " + err);
    return HMV_ERR_ARG;
  }
  *out = p;
  return HMV_OK;
}

int hmv_arena_from_plan(hmv_plan p, hmv_arena *out) {
  if (!p || !out) {
    set_error("hmv_arena_from_plan: null argument");
    return HMV_ERR_ARG;
  }
  *out = nullptr;
  int rc = check_device();
  if (rc) return rc;
  std::string err;
  if (!hmv::validate_plan(p->P, err)) {
    set_error("hmv_arena_from_plan: " + err);
    return HMV_ERR_ARG;
  }
  return upload_plan(p->P, out, std::chrono::steady_clock::now());
}

int hmv_arena_save(hmv_arena a, const char *path) {
  if (!a || !path) {
    set_error("hmv_arena_save: null argument");
    return HMV_ERR_ARG;
  }
  hmv_plan_s p;
  std::string err;
  if (!hmv::download_plan(a, p.P, err)) {
    set_error("hmv_arena_save: " + err);
    return HMV_ERR_DEVICE;
  }
  return hmv_plan_save(&p, path);
}

int hmv_arena_load(const char *path, hmv_arena *out) {
  hmv_plan p = nullptr;
  int rc = hmv_plan_load(path, &p);
  if (rc) return rc;
  rc = hmv_arena_from_plan(p, out);
  hmv_plan_free(p);
  return rc;
}

int hmv_plan_create(int nd, int nlf, const void *leaves, int leaf_stride, int precision,
                    hmv_plan *out) {
  return hmv_plan_create_tuned(nd, nlf, leaves, leaf_stride, precision, nullptr, out);
}

int hmv_plan_view(hmv_plan p, hmv_plan_tables *t) {
  if (!p || !t) {
    set_error("hmv_plan_view: null argument");
    return HMV_ERR_ARG;
  }
  const hmv::Plan &P = p->P;
  t->arena = P.arena.get();
  t->arena_bytes = (int64_t)P.arena_size;
  t->work1 = P.work1.data();
  t->nwork1 = (int64_t)P.work1.size();
  t->work2 = P.work2.data();
  t->nwork2 = (int64_t)P.work2.size();
  t->zseg = P.zseg.data();
  t->nzseg = (int64_t)P.zseg.size();
  t->outidx = P.outidx.data();
  t->noutidx = (int64_t)P.outidx.size();
  t->codes = P.codes.data();
  t->ncodes = (int64_t)P.codes.size();
  t->tfin = P.tfin.data();
  t->ntfin = (int64_t)P.tfin.size();
  t->tsize = P.tsize;
  t->tpart_size = P.tpart_size;
  t->work_bytes = (int)sizeof(HmvWork);
  t->zseg_bytes = (int)sizeof(HmvZSeg);
  t->tfin_bytes = (int)sizeof(HmvTFin);
  t->nwork1_small = P.nwork1_small;
  t->nt1_small = P.nt1_small;
  t->nwork2_small = P.nwork2_small;
  t->nt2_small = P.nt2_small;
  return HMV_OK;
}

int hmv_plan_free(hmv_plan p) {
  delete p;
  return HMV_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Reference-named entry points.
// ---------------------------------------------------------------------------
static_assert(sizeof(hmv_leaf) == 48 && offsetof(hmv_leaf, a1size) == 24 &&
                  offsetof(hmv_leaf, a1) == 32,
              "stc_HACApK_leafmtx layout (HACApK_MAGMA.h:12-20)");
static_assert(offsetof(hmv_ref_leafmtxp, st_lf_stride) == 16 &&
                  offsetof(hmv_ref_leafmtxp, gflops) == 40 &&
                  offsetof(hmv_ref_leafmtxp, gpu_vectors) == 48 &&
                  offsetof(hmv_ref_leafmtxp, num_batch) == 96 &&
                  offsetof(hmv_ref_leafmtxp, batch_tables) == 112 &&
                  offsetof(hmv_ref_leafmtxp, num_streamed) == 272 &&
                  offsetof(hmv_ref_leafmtxp, streamed_tables) == 280 &&
                  offsetof(hmv_ref_leafmtxp, mgpu_tables) == 328 &&
                  offsetof(hmv_ref_leafmtxp, mpi_rank) == 528 &&
                  offsetof(hmv_ref_leafmtxp, st_lf) == 536 && sizeof(hmv_ref_leafmtxp) == 544,
              "stc_HACApK_leafmtxp of a HAVE_MAGMA build (HACApK_MAGMA.h:28-106)");
static_assert(offsetof(hmv_ref_leafmtxp_plain, st_lf) == 24 &&
                  sizeof(hmv_ref_leafmtxp_plain) == 32,
              "stc_HACApK_leafmtxp without HAVE_MAGMA (m_HACApK_base.f90:64-76,234)");

namespace {
std::mutex g_tab_mu;
std::map<const void *, hmv_arena> g_tab;

// Struct layout of the caller's build: -1 = not read yet (HMV_REF_LAYOUT
// from the environment at the first call), else HMV_REF_LAYOUT_*.
std::atomic<int> g_layout{-1};

int ref_layout() {
  int l = g_layout.load();
  if (l < 0) {
    const char *env = std::getenv("HMV_REF_LAYOUT");
    int want = (env && std::strcmp(env, "plain") == 0) ? HMV_REF_LAYOUT_PLAIN : HMV_REF_LAYOUT_MAGMA;
    g_layout.compare_exchange_strong(l, want);
    l = g_layout.load();
  }
  return l;
}

// What the entry points read from the caller's st_HACApK_leafmtxp.
struct RefView {
  int nd, nlf, stride;
  const void *st_lf;
  int *mpi_rank;  // null in the plain layout
};

RefView ref_view(hmv_ref_leafmtxp *H) {
  if (ref_layout() == HMV_REF_LAYOUT_PLAIN) {
    auto *P = reinterpret_cast<hmv_ref_leafmtxp_plain *>(H);
    return {P->nd, P->nlf, P->st_lf_stride, P->st_lf, nullptr};
  }
  return {H->nd, H->nlf, H->st_lf_stride, H->st_lf, &H->mpi_rank};
}

// FP64 arena of the caller's leaves (nd from the caller's argument where
// the reference takes one); mpi_rank = 0, as the reference's lfcpy stores
// the rank of its one process.
int build_bound(hmv_ref_leafmtxp *H, int nd, hmv_arena *a) {
  const RefView v = ref_view(H);
  int rc = hmv_arena_build(nd, v.nlf, v.st_lf, v.stride, HMV_PREC_F64, a);
  if (rc == HMV_OK && v.mpi_rank) *v.mpi_rank = 0;
  return rc;
}

hmv_arena lookup(const void *key) {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  auto it = g_tab.find(key);
  return it == g_tab.end() ? nullptr : it->second;
}

void report(const char *fn) {
  std::fprintf(stderr, "%s: %s\n", fn, hmv_last_error());
  std::fflush(stderr);
}

// zau += H*zu on host vectors through the arena bound to H; the first call
// builds an FP64 arena and binds it.
void apply_bound(const char *name, double *zau, hmv_ref_leafmtxp *H, double *zu) {
  if (!zau || !H || !zu) {
    set_error(std::string(name) + ": null argument");
    report(name);
    return;
  }
  hmv_arena a = lookup(H);
  if (!a) {
    if (build_bound(H, ref_view(H).nd, &a) != HMV_OK) {
      report(name);
      return;
    }
    std::lock_guard<std::mutex> lk(g_tab_mu);
    g_tab[H] = a;
  }
  if (hmv_apply_host(a, zu, zau) != HMV_OK) report(name);
}
}  // namespace

extern "C" {

int hmv_ref_set_layout(int layout) {
  if (layout != HMV_REF_LAYOUT_MAGMA && layout != HMV_REF_LAYOUT_PLAIN) {
    set_error("hmv_ref_set_layout: unknown layout " + std::to_string(layout));
    return HMV_ERR_ARG;
  }
  const int prev = ref_layout();
  g_layout.store(layout);
  return prev;
}

int hmv_ref_get_layout(void) { return ref_layout(); }

void magma_init_(void) {
  if (hmv_init(0) != HMV_OK) report("magma_init_");
}

void magma_finalize_(void) {}

void c_hacapk_adot_body_lfcpy_batch_sorted_(int *nd, hmv_ref_leafmtxp *H) {
  if (!nd || !H) return;
  c_hacapk_adot_body_lfdel_batch_(H);
  hmv_arena a = nullptr;
  if (build_bound(H, *nd, &a) != HMV_OK) {
    report("c_hacapk_adot_body_lfcpy_batch_sorted_");
    return;
  }
  std::lock_guard<std::mutex> lk(g_tab_mu);
  g_tab[H] = a;
}

void c_hacapk_adot_body_lfmtx_batch_(double *zau, hmv_ref_leafmtxp *H, double *zu, double *zbu,
                                     double *time_batch, double *time_set, double *time_copy) {
  (void)zbu;
  (void)time_set;
  (void)time_copy;
  hmv_arena a = lookup(H);
  if (!a) {
    set_error("c_hacapk_adot_body_lfmtx_batch_: arena not built (call lfcpy_batch_sorted first)");
    report("c_hacapk_adot_body_lfmtx_batch_");
    return;
  }
  auto t0 = std::chrono::steady_clock::now();
  if (hmv_apply_host(a, zu, zau) != HMV_OK) report("c_hacapk_adot_body_lfmtx_batch_");
  auto t1 = std::chrono::steady_clock::now();
  if (time_batch) *time_batch += std::chrono::duration<double>(t1 - t0).count();
}

void c_hacapk_adot_body_lfmtx_(double *zau, hmv_ref_leafmtxp *H, double *zu, double *zbu) {
  (void)zbu;
  apply_bound("c_hacapk_adot_body_lfmtx_", zau, H, zu);
}

void c_hacapk_adot_body_lfmtx_gpu_(double *zau, hmv_ref_leafmtxp *H, double *zu, double *zbu) {
  (void)zbu;
  apply_bound("c_hacapk_adot_body_lfmtx_gpu_", zau, H, zu);
}

void c_hacapk_adot_body_lfcpy_gpu_(int *nd, hmv_ref_leafmtxp *H) {
  c_hacapk_adot_body_lfcpy_batch_sorted_(nd, H);
}

void c_hacapk_adot_body_lfdel_gpu_(hmv_ref_leafmtxp *H) { c_hacapk_adot_body_lfdel_batch_(H); }

void c_hacapk_adot_body_lfdel_batch_(hmv_ref_leafmtxp *H) {
  hmv_arena a = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = g_tab.find(H);
    if (it != g_tab.end()) {
      a = it->second;
      g_tab.erase(it);
    }
  }
  if (a) hmv_arena_free(a);
}

}  // extern "C"

namespace {

using SolveFn = int (*)(hmv_arena, const double *, double *, double, int, int, double *,
                        hmv_solve_info *);

// The reference's GPU solver entry points on host vectors: param[82] = max
// steps, param[90] = tolerance; the arena built by lfcpy_batch_sorted_ is
// reused, else one is built for the call.
void ref_solve(const char *name, const char *label, SolveFn solve, hmv_ref_leafmtxp *H,
               hmv_ref_lcontrol *C, double *u_cpu, double *b_cpu, double *param, int *nd,
               int *nstp, int *lrtrn) {
  if (lrtrn) *lrtrn = HMV_ERR_ARG;
  if (!H || !param || !nd || !u_cpu || !b_cpu) return;
  const int mstep = (int)param[82];
  const double eps = param[90];
  const int verbose = (C && C->param) ? (int)C->param[0] : (int)param[0];
  // The reference rebuilds the arena per call (HACApK_SOLVER.c:521,777); an
  // arena built earlier through lfcpy_batch_sorted_ is reused instead.
  hmv_arena a = lookup(H);
  bool own = false;
  if (!a) {
    if (build_bound(H, *nd, &a) != HMV_OK) {
      report(name);
      return;
    }
    own = true;
  }
  const size_t nb = sizeof(double) * (size_t)*nd;
  double *du = nullptr, *db = nullptr;
  int rc = HMV_OK;
  hmv_solve_info info{};
  if (hipMalloc(&du, nb) != hipSuccess || hipMalloc(&db, nb) != hipSuccess ||
      hipMemcpy(du, u_cpu, nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db, b_cpu, nb, hipMemcpyHostToDevice) != hipSuccess) {
    set_error(std::string(name) + ": device vectors");
    rc = HMV_ERR_DEVICE;
  }
  if (rc == HMV_OK) rc = solve(a, db, du, eps, mstep, 1, nullptr, &info);
  if (rc == HMV_OK && hipMemcpy(u_cpu, du, nb, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error(std::string(name) + ": copy back");
    rc = HMV_ERR_DEVICE;
  }
  if (rc != HMV_OK) report(name);
  if (nstp) *nstp = info.nstp;
  if (lrtrn) *lrtrn = rc == HMV_OK ? info.status : rc;
  if (verbose > 0 && rc == HMV_OK) {
    std::printf(" %s (hmv): steps=%d relres=%.2e true relres=%.2e time=%.5e s\n", label, info.nstp,
                info.relres, info.true_relres, info.seconds);
    std::fflush(stdout);
  }
  (void)hipFree(du);
  (void)hipFree(db);
  if (own) hmv_arena_free(a);
}

}  // namespace

extern "C" {

void c_hacapk_bicgstab_cax_lfmtx_gpu_(hmv_ref_leafmtxp *H, hmv_ref_lcontrol *C, double *u_cpu,
                                      double *b_cpu, double *param, int *nd, int *nstp,
                                      int *lrtrn) {
  ref_solve("c_hacapk_bicgstab_cax_lfmtx_gpu_", "BiCGSTAB", hmv_bicgstab, H, C, u_cpu, b_cpu, param,
            nd, nstp, lrtrn);
}

void c_hacapk_bicgstab_cax_lfmtx_pipe_(hmv_ref_leafmtxp *H, hmv_ref_lcontrol *C, double *u_cpu,
                                       double *b_cpu, double *param, int *nd, int *nstp,
                                       int *lrtrn) {
  ref_solve("c_hacapk_bicgstab_cax_lfmtx_pipe_", "pipelined BiCGSTAB", hmv_bicgstab_pipe, H, C,
            u_cpu, b_cpu, param, nd, nstp, lrtrn);
}

void c_hacapk_bicgstab_cax_lfmtx_pipe2_(hmv_ref_leafmtxp *H, hmv_ref_lcontrol *C, double *u_cpu,
                                        double *b_cpu, double *param, int *nd, int *nstp,
                                        int *lrtrn) {
  ref_solve("c_hacapk_bicgstab_cax_lfmtx_pipe2_", "pipelined BiCGSTAB", hmv_bicgstab_pipe, H, C,
            u_cpu, b_cpu, param, nd, nstp, lrtrn);
}

void c_hacapk_bicgstab_cax_lfmtx_flat_(hmv_ref_leafmtxp *H, hmv_ref_lcontrol *C, double *u_cpu,
                                       double *b_cpu, double *param, int *nd, int *nstp,
                                       int *lrtrn) {
  ref_solve("c_hacapk_bicgstab_cax_lfmtx_flat_", "BiCGSTAB", hmv_bicgstab, H, C, u_cpu, b_cpu, param,
            nd, nstp, lrtrn);
}

void c_hacapk_bicgstab_cax_lfmtx_mgpu1_(hmv_ref_leafmtxp *H, hmv_ref_lcontrol *C, double *u_cpu,
                                        double *b_cpu, double *param, int *nd, int *nstp,
                                        int *lrtrn) {
  // lpmd(2) = number of ranks (HACApK_SOLVER.c:837 reads lpmd[1])
  if (C && C->param) {
    const int *lpmd = (const int *)((const char *)C->param + C->lpmd_offset);
    if (lpmd[1] > 1) {
      set_error("c_hacapk_bicgstab_cax_lfmtx_mgpu1_: multi-rank solve not built (one GPU per "
                "process, one process)");
      report("c_hacapk_bicgstab_cax_lfmtx_mgpu1_");
      if (lrtrn) *lrtrn = HMV_ERR_ARG;
      return;
    }
  }
  ref_solve("c_hacapk_bicgstab_cax_lfmtx_mgpu1_", "BiCGSTAB", hmv_bicgstab, H, C, u_cpu, b_cpu,
            param, nd, nstp, lrtrn);
}

}  // extern "C"
