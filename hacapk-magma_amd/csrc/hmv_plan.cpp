// hmv_plan.cpp -- repack of the leaf list into the engine arena (host side).
//
// Replaces the reference's arena build c_hacapk_adot_body_lfcpy_batch_sorted_
// (HACApK_BATCH_SORT.c:86-858): instead of thousands of per-leaf uploads into
// padded per-leaf GEMV operands sorted into batches of 5000, every leaf entry
// is copied once into a host image of the arena, laid out as one contiguous
// column-major block per workgroup of the two apply phases, and the image goes
// to the device in one copy.
#include "hmv_plan.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <utility>

namespace hmv {

int dtype_size(int dt) { return dt == HMV_DT_F64 ? 8 : dt == HMV_DT_F32 ? 4 : 2; }
int row_vector(int dt) { return dt == HMV_DT_F64 ? 1 : dt == HMV_DT_F32 ? 2 : 4; }
int row_vector_p2(int dt) { return dt == HMV_DT_F64 ? 1 : 2; }

namespace {

int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

uint16_t to_bf16(double v) {
  // round to nearest even of the FP32 value
  float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

inline void store(char *dst, int dt, int64_t i, double v) {
  if (dt == HMV_DT_F64) {
    reinterpret_cast<double *>(dst)[i] = v;
  } else if (dt == HMV_DT_F32) {
    reinterpret_cast<float *>(dst)[i] = (float)v;
  } else {
    reinterpret_cast<uint16_t *>(dst)[i] = to_bf16(v);
  }
}

template <typename F>
void parallel_for(int64_t n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 64 || nt == 1) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  std::atomic<int64_t> next(0);
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&]() {
      for (;;) {
        int64_t i = next.fetch_add(1);
        if (i >= n) break;
        f(i);
      }
    });
  for (auto &t : th) t.join();
}

const double *u_factor(const RefLeaf *l) {
  return reinterpret_cast<const double *>(reinterpret_cast<const char *>(l->a1) + l->a1size);
}

}  // namespace

PlanTuning default_tuning(int prec) {
  PlanTuning t;
  if (prec == kF32) t.p1_small_threads = 128;
  if (prec == kBF16 || prec == kBF16All) {
    t.p1_small_threads = 128;
    t.p2_small_threads = 64;
  }
  return t;
}

namespace {
bool plan_impl(int nd, int nlf, const void *leaves, int stride, int prec, const PlanTuning &tune,
               Plan &P, std::string &err, bool device);
}

bool build_plan(int nd, int nlf, const void *leaves, int stride, int prec,
                const PlanTuning &tune, Plan &P, std::string &err) {
  return plan_impl(nd, nlf, leaves, stride, prec, tune, P, err, false);
}

bool build_plan_device(int nd, int nlf, const void *leaves, int stride, int prec,
                       const PlanTuning &tune, Plan &P, std::string &err) {
  return plan_impl(nd, nlf, leaves, stride, prec, tune, P, err, true);
}

namespace {
bool plan_impl(int nd, int nlf, const void *leaves, int stride, int prec, const PlanTuning &tune,
               Plan &P, std::string &err, bool device) {
  P = Plan();
  auto nt_ok = [](int v) { return v == -1 || v == 64 || v == 128; };
  if (!nt_ok(tune.p1_small_threads) || !nt_ok(tune.p2_small_threads)) {
    err = "invalid tuning: small-workgroup threads must be -1, 64 or 128";
    return false;
  }
  const bool timing = std::getenv("HMV_PLAN_TIMING") != nullptr;
  auto tp = std::chrono::steady_clock::now();
  auto lap = [&](const char *what) {
    if (!timing) return;
    auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "plan %-12s %.3f s\n", what, std::chrono::duration<double>(now - tp).count());
    tp = now;
  };
  P.nd = nd;
  P.prec = prec;
  P.tune.p1_target_bytes = tune.p1_target_bytes;
  P.tune.p1_row_tile = tune.p1_row_tile;
  P.tune.p2_max_rows = tune.p2_max_rows;
  P.tune.p1_small_threads = tune.p1_small_threads;
  P.tune.p2_small_threads = tune.p2_small_threads;
  P.seg = {0, nd};
  if (nd <= 0) {
    err = "nd must be positive";
    return false;
  }
  if (nlf < 0 || (nlf > 0 && !leaves)) {
    err = "invalid leaf array";
    return false;
  }
  if (stride < (int)sizeof(RefLeaf) && nlf > 0) {
    err = "leaf stride smaller than the leaf record";
    return false;
  }
  if (prec != kF64 && prec != kF32 && prec != kBF16 && prec != kBF16All) {
    err = "unknown precision";
    return false;
  }
  if (tune.p1_target_bytes < 1024 || tune.p1_row_tile < 1 || tune.p1_row_tile > HMV_MAX_ROWS ||
      tune.p2_max_rows < 1 || tune.p2_max_rows > HMV_MAX_ROWS) {
    err = "invalid tuning parameters";
    return false;
  }
  const int dt_dense = prec == kF64 ? HMV_DT_F64 : prec == kBF16All ? HMV_DT_BF16 : HMV_DT_F32;
  const int dt_lr = prec == kF64 ? HMV_DT_F64 : prec == kF32 ? HMV_DT_F32 : HMV_DT_BF16;
  const bool merged = dt_dense == dt_lr;
  const int es_d = dtype_size(dt_dense), es_l = dtype_size(dt_lr);
  // rows per lane: phase-1 blocks hold only low-rank factors (dt_lr), phase-2
  // blocks follow their first part's type (dt_dense)
  const int vr = row_vector(dt_lr), vr2 = row_vector_p2(dt_dense);

  std::vector<const RefLeaf *> L(nlf);
  PlanStats &st = P.st;
  st.nlf = nlf;
  for (int i = 0; i < nlf; ++i) {
    const RefLeaf *l = reinterpret_cast<const RefLeaf *>(static_cast<const char *>(leaves) +
                                                         (size_t)stride * i);
    L[i] = l;
    if (l->ltmtx != 1 && l->ltmtx != 2) {
      err = "leaf " + std::to_string(i + 1) + ": ltmtx must be 1 or 2";
      return false;
    }
    if (l->ndl <= 0 || l->ndt <= 0 || l->nstrtl < 1 || l->nstrtt < 1 ||
        (int64_t)l->nstrtl - 1 + l->ndl > nd || (int64_t)l->nstrtt - 1 + l->ndt > nd) {
      err = "leaf " + std::to_string(i + 1) + ": block outside the matrix";
      return false;
    }
    if (l->ltmtx == 1) {
      if (l->kt < 0) {
        err = "leaf " + std::to_string(i + 1) + ": negative rank";
        return false;
      }
      if (l->kt > 0 && !l->a1) {
        err = "leaf " + std::to_string(i + 1) + ": missing factors";
        return false;
      }
      ++st.nlr;
      if (l->kt == 0) ++st.nkt0;
      st.entries_lr += (int64_t)l->kt * (l->ndl + l->ndt);
      st.tslots += l->kt;
      st.ktmax = std::max(st.ktmax, l->kt);
    } else {
      if (!l->a1) {
        err = "leaf " + std::to_string(i + 1) + ": missing dense block";
        return false;
      }
      ++st.ndense;
      st.entries_dense += (int64_t)l->ndl * l->ndt;
    }
    st.max_ndt = std::max(st.max_ndt, l->ndt);
  }
  if (st.tslots > INT32_MAX / 4) {
    err = "too many low-rank slots for 32-bit indexing";
    return false;
  }

  // ---- groups -------------------------------------------------------------
  struct RowGroup {
    int s, n;  // 0-based first row, rows
    std::vector<int> dense, lr;
    int wd = 0, k = 0;
    int64_t tbase = 0;
  };
  struct ColGroup {
    int s, n;
    std::vector<int> lr;
    int r = 0;
    int np = 1;  // column chunks (t partials per rank vector)
  };
  // group index by (first row or column, size), numbered in order of first
  // appearance
  auto gkey = [](int s, int n) { return (uint64_t)(uint32_t)s << 32 | (uint32_t)n; };
  std::unordered_map<uint64_t, int> rkey, ckey;
  rkey.reserve((size_t)nlf);
  ckey.reserve((size_t)nlf);
  std::vector<RowGroup> rg;
  std::vector<ColGroup> cg;
  for (int i = 0; i < nlf; ++i) {
    const RefLeaf *l = L[i];
    const uint64_t rk = gkey(l->nstrtl - 1, l->ndl);
    auto it = rkey.find(rk);
    int r;
    if (it == rkey.end()) {
      r = (int)rg.size();
      rkey.emplace(rk, r);
      rg.push_back(RowGroup{l->nstrtl - 1, l->ndl, {}, {}});
    } else {
      r = it->second;
    }
    if (l->ltmtx == 2) {
      rg[r].dense.push_back(i);
      rg[r].wd += l->ndt;
    } else if (l->kt > 0) {
      rg[r].lr.push_back(i);
      rg[r].k += l->kt;
      const uint64_t ck = gkey(l->nstrtt - 1, l->ndt);
      auto jt = ckey.find(ck);
      int c;
      if (jt == ckey.end()) {
        c = (int)cg.size();
        ckey.emplace(ck, c);
        cg.push_back(ColGroup{l->nstrtt - 1, l->ndt, {}});
      } else {
        c = jt->second;
      }
      cg[c].lr.push_back(i);
      cg[c].r += l->kt;
    }
  }
  st.row_groups = (int64_t)rg.size();
  st.col_groups = (int64_t)cg.size();
  lap("groups");

  // t slots ordered by row group (ascending first row), leaves in leaf order
  std::vector<int> rg_order(rg.size());
  std::iota(rg_order.begin(), rg_order.end(), 0);
  std::sort(rg_order.begin(), rg_order.end(), [&](int a, int b) {
    return rg[a].s != rg[b].s ? rg[a].s < rg[b].s : rg[a].n < rg[b].n;
  });
  std::vector<int64_t> leaf_tbase(nlf, -1);
  std::vector<int> leaf_cg(nlf, -1);
  int64_t tcur = 0;
  for (int r : rg_order) {
    rg[r].tbase = tcur;
    for (int i : rg[r].lr) {
      leaf_tbase[i] = tcur;
      tcur += L[i]->kt;
    }
  }
  P.tsize = tcur;
  for (int c = 0; c < (int)cg.size(); ++c)
    for (int i : cg[c].lr) leaf_cg[i] = c;

  // ---- phase-1 pieces -----------------------------------------------------
  struct Piece {
    int c;       // column group
    int r0, nr;  // rank-vector rows
    int k;       // column chunk index
    int c0, cw;  // columns
    int64_t off, bytes;
  };
  std::vector<Piece> pieces;
  for (int c = 0; c < (int)cg.size(); ++c) {
    ColGroup &g = cg[c];
    const int64_t n = g.n;
    int rs = std::min(g.r, tune.p1_row_tile);
    if (round_up(rs, vr) * n * es_l > tune.p1_target_bytes)
      rs = (int)std::max<int64_t>(1, tune.p1_target_bytes / (n * es_l));
    int cw = g.n;
    if (rs < std::min(g.r, 8)) {
      // long rank vectors: chunk the columns; chunk sums are t partials
      rs = std::min(g.r, 16);
      cw = (int)std::max<int64_t>(256, tune.p1_target_bytes / (round_up(rs, vr) * es_l));
      cw = (int)std::max<int64_t>(cw, (n + HMV_MAX_TPARTS - 1) / HMV_MAX_TPARTS);
      if (cw >= g.n) cw = g.n;
    }
    g.np = (int)((n + cw - 1) / cw);
    for (int r0 = 0; r0 < g.r; r0 += rs) {
      const int nr = std::min(rs, g.r - r0);
      for (int k = 0; k < g.np; ++k) {
        Piece p{c, r0, nr, k, k * cw, (int)std::min<int64_t>(cw, n - (int64_t)k * cw), 0, 0};
        pieces.push_back(p);
      }
    }
  }
  // t buffer: the tsize slots, then the partials of rank vectors whose
  // column group is chunked (np > 1).  tout[s] = where chunk 0 of slot s
  // lands (chunk k at tout[s] + k).
  std::vector<int64_t> tout((size_t)P.tsize, 0);
  int64_t pcur = P.tsize;
  for (int r : rg_order)
    for (int i : rg[r].lr)
      for (int k = 0; k < L[i]->kt; ++k) {
        const int64_t s = leaf_tbase[i] + k;
        const int np = cg[leaf_cg[i]].np;
        if (np == 1) {
          tout[s] = s;
        } else {
          tout[s] = pcur;
          P.tfin.push_back(HmvTFin{(int)s, (int)pcur, np, 0});
          pcur += np;
        }
      }
  P.tpart_size = pcur;
  st.tparts = pcur;
  if (pcur > INT32_MAX) {
    err = "too many t entries for 32-bit input codes";
    return false;
  }
  // rank vectors of each column group, in group order: (leaf, l), in one
  // flat array where group c's rows start at cg_first[c]
  std::vector<int64_t> cg_first(cg.size() + 1, 0);
  for (size_t c = 0; c < cg.size(); ++c) cg_first[c + 1] = cg_first[c] + cg[c].r;
  std::vector<std::pair<int, int>> cg_flat((size_t)cg_first[cg.size()]);
  parallel_for((int64_t)cg.size(), [&](int64_t c) {
    std::pair<int, int> *o = cg_flat.data() + cg_first[c];
    for (int i : cg[c].lr)
      for (int k = 0; k < L[i]->kt; ++k) *o++ = {i, k};
  });
  auto cg_rows = [&](int c) { return cg_flat.data() + cg_first[c]; };

  // ---- phase-2 atoms ------------------------------------------------------
  std::vector<int> bnd;
  bnd.reserve(2 * rg.size() + 2);
  bnd.push_back(0);
  bnd.push_back(nd);
  for (auto &g : rg) {
    bnd.push_back(g.s);
    bnd.push_back(g.s + g.n);
  }
  std::sort(bnd.begin(), bnd.end());
  bnd.erase(std::unique(bnd.begin(), bnd.end()), bnd.end());
  st.atoms = (int64_t)bnd.size() - 1;
  std::vector<std::pair<int, int>> ranges;  // row slices of at most p2_max_rows rows
  for (size_t a = 0; a + 1 < bnd.size(); ++a)
    for (int r = bnd[a]; r < bnd[a + 1]; r += tune.p2_max_rows)
      ranges.push_back({r, std::min(bnd[a + 1], r + tune.p2_max_rows)});
  std::vector<int> by_start(rg.size());
  std::iota(by_start.begin(), by_start.end(), 0);
  std::sort(by_start.begin(), by_start.end(), [&](int a, int b) { return rg[a].s < rg[b].s; });
  struct Slice {
    int a, b;
    std::vector<int> cov;  // covering row groups, by first row then size (outer first)
    int ncol0 = 0, ncol1 = 0;
    int64_t off0 = -1, off1 = -1, bytes = 0;
  };
  std::vector<Slice> slices(ranges.size());
  {
    std::vector<int> active;
    size_t next = 0;
    for (size_t q = 0; q < ranges.size(); ++q) {
      const int a = ranges[q].first, b = ranges[q].second;
      while (next < by_start.size() && rg[by_start[next]].s < b) active.push_back(by_start[next++]);
      active.erase(std::remove_if(active.begin(), active.end(),
                                  [&](int gi) { return rg[gi].s + rg[gi].n <= a; }),
                   active.end());
      Slice &S = slices[q];
      S.a = a;
      S.b = b;
      for (int gi : active)
        if (rg[gi].s <= a && rg[gi].s + rg[gi].n >= b && (rg[gi].wd + rg[gi].k) > 0) S.cov.push_back(gi);
      std::sort(S.cov.begin(), S.cov.end(), [&](int x, int y) {
        return rg[x].s != rg[y].s ? rg[x].s < rg[y].s : rg[x].n > rg[y].n;
      });
      for (int gi : S.cov) {
        if (merged) {
          S.ncol0 += rg[gi].wd + rg[gi].k;
        } else {
          S.ncol0 += rg[gi].wd;
          S.ncol1 += rg[gi].k;
        }
      }
    }
  }
  lap("atoms");

  // ---- arena layout -------------------------------------------------------
  int64_t off = 0;
  for (auto &p : pieces) {
    p.off = off;
    p.bytes = round_up(p.nr, vr) * (int64_t)p.cw * es_l;
    off = round_up(off + p.bytes, HMV_ARENA_ALIGN);
    st.phase1_bytes += p.bytes;
  }
  for (auto &S : slices) {
    const int64_t rows_p = round_up(S.b - S.a, vr2);
    if (S.ncol0 > 0) {
      S.off0 = off;
      S.bytes += rows_p * S.ncol0 * es_d;
      off = round_up(off + rows_p * S.ncol0 * es_d, HMV_ARENA_ALIGN);
    }
    if (S.ncol1 > 0) {
      S.off1 = off;
      S.bytes += rows_p * S.ncol1 * es_l;
      off = round_up(off + rows_p * S.ncol1 * es_l, HMV_ARENA_ALIGN);
    }
    st.phase2_bytes += S.bytes;
  }
  st.arena_bytes = std::max<int64_t>(off, HMV_ARENA_ALIGN);
  st.payload_bytes = st.entries_lr * es_l + st.entries_dense * es_d;
  st.alg_bytes = st.payload_bytes + 16LL * nd;
  P.arena_size = (size_t)st.arena_bytes;
  char *A = nullptr;
  if (!device) {
    try {
      // left uninitialised: every block (with its padding) is zeroed by the
      // thread that fills it, so the image is first touched in parallel
      P.arena.reset(new char[(size_t)st.arena_bytes]);
    } catch (...) {
      err = "out of host memory for the arena image";
      return false;
    }
    A = P.arena.get();
    if (off < st.arena_bytes) std::memset(A + off, 0, (size_t)(st.arena_bytes - off));
  }
  lap("layout");

  // ---- tables -------------------------------------------------------------
  auto tcode = [&](int64_t slot) { return (int)(-1 - slot); };
  struct Item {
    int64_t bytes;
    int64_t key;  // locality: first x column (phase 1) or first row (phase 2)
    HmvWork w;
  };
  std::vector<Item> it1, it2;
  it1.reserve(pieces.size());
  for (auto &p : pieces) {
    const ColGroup &g = cg[p.c];
    HmvWork w{};
    w.off0 = p.off;
    w.off1 = -1;
    w.rows = p.nr;
    w.rows_p = (int)round_up(p.nr, vr);
    w.ncol0 = p.cw;
    w.ncol1 = 0;
    w.dt0 = (short)dt_lr;
    w.dt1 = (short)dt_lr;
    w.zseg_begin = (int)P.zseg.size();
    P.zseg.push_back(HmvZSeg{HMV_Z_X, p.cw, g.s + p.c0, 0});
    w.zseg_end = (int)P.zseg.size();
    w.out_base = (int)P.outidx.size();
    for (int r = 0; r < p.nr; ++r) {
      const auto &rv = cg_rows(p.c)[p.r0 + r];
      const int64_t s = leaf_tbase[rv.first] + rv.second;
      P.outidx.push_back((int)(tout[s] + p.k));
    }
    it1.push_back({p.bytes, (int64_t)g.s + p.c0, w});
  }
  it2.reserve(slices.size());
  for (auto &S : slices) {
    HmvWork w{};
    w.off0 = S.off0;
    w.off1 = S.off1;
    w.rows = S.b - S.a;
    w.rows_p = (int)round_up(S.b - S.a, vr2);
    w.ncol0 = S.ncol0;
    w.ncol1 = S.ncol1;
    w.dt0 = (short)dt_dense;
    w.dt1 = (short)dt_lr;
    w.zseg_begin = (int)P.zseg.size();
    // the block's inputs as runs of consecutive x entries (a dense leaf's
    // columns) and t slots (a row group's rank vectors), joined when they
    // continue each other
    std::vector<HmvZSeg> runs;
    int col = 0;
    auto add_run = [&](int kind, int64_t src, int n) {
      if (n <= 0) return;
      if (!runs.empty() && runs.back().kind == kind && runs.back().src + runs.back().ncols == src)
        runs.back().ncols += n;
      else
        runs.push_back(HmvZSeg{kind, n, (int)src, col});
      col += n;
    };
    auto dense_runs = [&](const RowGroup &g) {
      for (int i : g.dense) add_run(HMV_Z_X, L[i]->nstrtt - 1, L[i]->ndt);
    };
    auto t_runs = [&](const RowGroup &g) { add_run(HMV_Z_T, g.tbase, (int)g.k); };
    if (merged) {
      for (int gi : S.cov) {
        dense_runs(rg[gi]);
        t_runs(rg[gi]);
      }
    } else {
      for (int gi : S.cov) dense_runs(rg[gi]);
      for (int gi : S.cov) t_runs(rg[gi]);
    }
    if (col != S.ncol0 + S.ncol1) {
      err = "internal: input runs do not match the block columns";
      return false;
    }
    if ((int)runs.size() <= HMV_RUN_MAX) {
      P.zseg.insert(P.zseg.end(), runs.begin(), runs.end());
    } else {
      // too many runs for the workgroup's run table: one code per column
      const int64_t cbase = (int64_t)P.codes.size();
      if (cbase + col > INT32_MAX) {
        err = "too many input codes for 32-bit indexing";
        return false;
      }
      for (const HmvZSeg &r : runs)
        for (int c = 0; c < r.ncols; ++c)
          P.codes.push_back(r.kind == HMV_Z_X ? r.src + c : tcode((int64_t)r.src + c));
      P.zseg.push_back(HmvZSeg{HMV_Z_CODES, col, (int)cbase, 0});
    }
    w.zseg_end = (int)P.zseg.size();
    w.out_base = S.a;
    it2.push_back({S.bytes, S.a, w});
  }
  lap("tables");

  // ---- fill ---------------------------------------------------------------
  if (device) {
    // The same copies as the host fill below, as records for the device (the
    // image is zeroed there first).  A piece's consecutive rank vectors of
    // one leaf are one record (V's rows lie ndt apart).  The records are
    // counted per piece and per slice, then written in parallel at their
    // offsets.
    const size_t npc = pieces.size(), nsl = slices.size();
    std::vector<int64_t> at(npc + nsl + 1, 0);
    parallel_for((int64_t)npc, [&](int64_t pi) {
      const Piece &p = pieces[pi];
      const std::pair<int, int> *rows = cg_rows(p.c);
      int64_t n = 0;
      for (int r = 0; r < p.nr; ++r)
        if (r == 0 || rows[p.r0 + r].first != rows[p.r0 + r - 1].first) ++n;
      at[pi + 1] = n;
    });
    parallel_for((int64_t)nsl, [&](int64_t si) {
      int64_t n = 0;
      for (int gi : slices[si].cov) n += (int64_t)(rg[gi].dense.size() + rg[gi].lr.size());
      at[npc + si + 1] = n;
    });
    for (size_t q = 0; q < npc + nsl; ++q) at[q + 1] += at[q];
    P.fill.resize((size_t)at[npc + nsl]);
    HmvFillOp *F = P.fill.data();
    parallel_for((int64_t)npc, [&](int64_t pi) {
      const Piece &p = pieces[pi];
      const std::pair<int, int> *rows = cg_rows(p.c);
      const int rows_p = (int)round_up(p.nr, vr);
      HmvFillOp *o = F + at[pi];
      for (int r = 0; r < p.nr;) {
        const auto &rv = rows[p.r0 + r];
        int m = 1;  // rank vectors rv.second .. rv.second + m - 1 of one leaf
        while (r + m < p.nr && rows[p.r0 + r + m].first == rv.first) ++m;
        const RefLeaf *l = L[rv.first];
        *o++ = HmvFillOp{p.off, l->a1 + (size_t)rv.second * l->ndt + p.c0, l->ndt, 1, rows_p, r,
                         0, m, p.cw, dt_lr};
        r += m;
      }
    });
    parallel_for((int64_t)nsl, [&](int64_t si) {
      const Slice &S = slices[si];
      const int rows = S.b - S.a;
      const int rows_p = (int)round_up(rows, vr2);
      HmvFillOp *o = F + at[npc + si];
      int c0 = 0, c1 = 0;
      auto put_dense = [&](const RowGroup &g, int64_t off, int dt, int &col) {
        for (int i : g.dense) {
          const RefLeaf *l = L[i];
          *o++ = HmvFillOp{off, l->a1 + (size_t)(S.a - g.s) * l->ndt, l->ndt, 1, rows_p, 0, col,
                           rows, l->ndt, dt};
          col += l->ndt;
        }
      };
      auto put_u = [&](const RowGroup &g, int64_t off, int dt, int &col) {
        for (int i : g.lr) {  // rank > 0 only
          const RefLeaf *l = L[i];
          *o++ = HmvFillOp{off, u_factor(l) + (S.a - g.s), 1, l->ndl, rows_p, 0, col, rows, l->kt,
                           dt};
          col += l->kt;
        }
      };
      for (int gi : S.cov) {
        const RowGroup &g = rg[gi];
        if (merged) {
          put_dense(g, S.off0, dt_dense, c0);
          put_u(g, S.off0, dt_lr, c0);
        } else {
          put_dense(g, S.off0, dt_dense, c0);
          put_u(g, S.off1, dt_lr, c1);
        }
      }
    });
  } else {
  parallel_for((int64_t)pieces.size(), [&](int64_t pi) {
    const Piece &p = pieces[pi];
    char *dst = A + p.off;
    const int rows_p = (int)round_up(p.nr, vr);
    std::memset(dst, 0, (size_t)(round_up(p.off + p.bytes, HMV_ARENA_ALIGN) - p.off));
    for (int r = 0; r < p.nr; ++r) {
      const auto &rv = cg_rows(p.c)[p.r0 + r];
      const RefLeaf *l = L[rv.first];
      const double *v = l->a1 + (size_t)rv.second * l->ndt + p.c0;  // V(:, l) from column c0
      for (int j = 0; j < p.cw; ++j) store(dst, dt_lr, r + (int64_t)rows_p * j, v[j]);
    }
  });
  parallel_for((int64_t)slices.size(), [&](int64_t si) {
    const Slice &S = slices[si];
    const int rows = S.b - S.a;
    const int64_t rows_p = round_up(rows, vr2);
    if (S.off0 >= 0)
      std::memset(A + S.off0, 0, (size_t)(round_up(S.off0 + rows_p * S.ncol0 * es_d, HMV_ARENA_ALIGN) - S.off0));
    if (S.off1 >= 0)
      std::memset(A + S.off1, 0, (size_t)(round_up(S.off1 + rows_p * S.ncol1 * es_l, HMV_ARENA_ALIGN) - S.off1));
    char *d0 = S.off0 >= 0 ? A + S.off0 : nullptr;
    char *d1 = S.off1 >= 0 ? A + S.off1 : nullptr;
    int64_t c0 = 0, c1 = 0;
    auto put_dense = [&](const RowGroup &g, char *dst, int dt, int64_t &col) {
      for (int i : g.dense) {
        const RefLeaf *l = L[i];
        for (int r = 0; r < rows; ++r) {
          const double *arow = l->a1 + (size_t)(S.a + r - g.s) * l->ndt;  // row of A
          for (int j = 0; j < l->ndt; ++j) store(dst, dt, r + rows_p * (col + j), arow[j]);
        }
        col += l->ndt;
      }
    };
    auto put_u = [&](const RowGroup &g, char *dst, int dt, int64_t &col) {
      for (int i : g.lr) {
        const RefLeaf *l = L[i];
        const double *u = u_factor(l);
        for (int k = 0; k < l->kt; ++k)
          for (int r = 0; r < rows; ++r)
            store(dst, dt, r + rows_p * (col + k), u[(S.a + r - g.s) + (size_t)k * l->ndl]);
        col += l->kt;
      }
    };
    for (int gi : S.cov) {
      const RowGroup &g = rg[gi];
      if (merged) {
        put_dense(g, d0, dt_dense, c0);
        put_u(g, d0, dt_lr, c0);
      } else {
        put_dense(g, d0, dt_dense, c0);
        put_u(g, d1, dt_lr, c1);
      }
    }
  });
  }
  lap("fill");

  // Dispatch order.  The hardware sends workgroup b to XCD b % G; give each
  // XCD one contiguous G-th of the locality-sorted list (its gathers then
  // stay in its own L2), largest workgroups first within each XCD so its
  // in-order stream ends balanced.
  auto order = [&](std::vector<Item> &items, std::vector<HmvWork> &out) {
    const int G = std::max(1, tune.xcd_groups);
    std::stable_sort(items.begin(), items.end(),
                     [](const Item &x, const Item &y) { return x.key < y.key; });
    const size_t n = items.size();
    std::vector<std::vector<Item>> chunk(G);
    // the first n % G chunks take one extra item, so that workgroup b of the
    // interleaved list always belongs to chunk b % G
    const size_t q = n / G, r = n % G;
    size_t b = 0;
    for (int k = 0; k < G; ++k) {
      const size_t e = b + q + ((size_t)k < r ? 1 : 0);
      chunk[k].assign(items.begin() + b, items.begin() + e);
      b = e;
      std::stable_sort(chunk[k].begin(), chunk[k].end(),
                       [](const Item &x, const Item &y) { return x.bytes > y.bytes; });
    }
    out.clear();
    out.reserve(n);
    for (size_t pos = 0; out.size() < n; ++pos)
      for (int k = 0; k < G; ++k)
        if (pos < chunk[k].size()) out.push_back(chunk[k][pos].w);
  };
  // Short blocks (at most nt lanes per column) first, in their own dispatch
  // order, when the phase runs them in nt-thread workgroups.
  auto order_split = [&](std::vector<Item> &items, int nt, int vrw, std::vector<HmvWork> &out,
                         int &nsmall, int &nt_small) {
    nsmall = 0;
    nt_small = 0;
    if (nt == 64 || nt == 128) {
      std::vector<Item> small, large;
      for (auto &it : items) (it.w.rows_p / vrw <= nt ? small : large).push_back(it);
      if (!small.empty()) {
        std::vector<HmvWork> a, b;
        order(small, a);
        order(large, b);
        nsmall = (int)a.size();
        nt_small = nt;
        out = std::move(a);
        out.insert(out.end(), b.begin(), b.end());
        return;
      }
    }
    order(items, out);
  };
  order_split(it1, tune.p1_small_threads, vr, P.work1, P.nwork1_small, P.nt1_small);
  order_split(it2, tune.p2_small_threads, vr2, P.work2, P.nwork2_small, P.nt2_small);
  lap("order");
  return true;
}
}  // namespace

}  // namespace hmv
