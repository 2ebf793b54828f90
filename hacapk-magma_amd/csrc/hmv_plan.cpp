// hmv_plan.cpp -- repack of the leaf list into the engine arena (host side).
//
// Replaces the reference's arena build c_hacapk_adot_body_lfcpy_batch_sorted_
// (HACApK_BATCH_SORT.c:86-858): instead of thousands of per-leaf uploads into
// padded per-leaf GEMV operands sorted into batches of 5000, every leaf is
// copied once into a host image of the arena, grouped for the two phases of
// the apply, and the image goes to the device in one copy.
#include "hmv_plan.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <thread>
#include <utility>

namespace hmv {

int dtype_size(int dt) { return dt == HMV_DT_F64 ? 8 : dt == HMV_DT_F32 ? 4 : 2; }

namespace {

int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

uint16_t to_bf16(double v) {
  // round-to-nearest-even of the float value
  float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

void store(char *dst, int dt, int64_t i, double v) {
  if (dt == HMV_DT_F64) {
    reinterpret_cast<double *>(dst)[i] = v;
  } else if (dt == HMV_DT_F32) {
    reinterpret_cast<float *>(dst)[i] = (float)v;
  } else {
    reinterpret_cast<uint16_t *>(dst)[i] = to_bf16(v);
  }
}

int choose_lpr(int ncols, int vecw, int nrows, int bias) {
  int cw = std::min(ncols, HMV_ZCHUNK);
  int nvec = (cw + vecw - 1) / vecw;
  int lpr = 4;
  while (lpr < 64 && lpr < nvec) lpr <<= 1;
  for (int b = 0; b < bias && lpr < 64; ++b) lpr <<= 1;
  // long rows and too few rows to keep the four waves busy: whole workgroup per row
  if (nvec > 64 && nrows < HMV_WG_THREADS / 64) lpr = 256;
  return lpr;
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

}  // namespace

bool build_plan(int nd, int nlf, const void *leaves, int stride, int prec,
                const PlanTuning &tune, Plan &P, std::string &err) {
  P = Plan();
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
  if (prec != kF64 && prec != kF32 && prec != kBF16) {
    err = "unknown precision";
    return false;
  }
  const int dt_dense = prec == kF64 ? HMV_DT_F64 : HMV_DT_F32;
  const int dt_lr = prec == kF64 ? HMV_DT_F64 : prec == kF32 ? HMV_DT_F32 : HMV_DT_BF16;
  const bool merged = dt_dense == dt_lr;
  const int es_d = dtype_size(dt_dense), es_l = dtype_size(dt_lr);
  const int vec_d = 16 / es_d, vec_l = 16 / es_l;

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
  if (st.tslots > INT32_MAX / 2) {
    err = "too many low-rank slots for 32-bit indexing";
    return false;
  }

  // ---- groups -------------------------------------------------------------
  std::map<std::pair<int, int>, int> rkey, ckey;
  struct RowGroup {
    int s, n;  // 0-based first row, rows
    std::vector<int> dense, lr;
    int wd = 0, k = 0;
    int64_t tbase = 0;
    int64_t zoff = 0;
    int64_t off_d = -1, off_u = -1;  // byte offsets (dense/merged panel, U panel)
    int64_t beg = -1, end = -1;      // byte extent of both panels incl. padding
    int ld_d = 0, ld_u = 0;
  };
  struct ColGroup {
    int s, n;
    std::vector<int> lr;
    int r = 0;
    int64_t off = -1, end = -1;  // byte extent incl. alignment padding
    int ld = 0;
    int64_t row0 = 0;  // first global V^T row
  };
  std::vector<RowGroup> rg;
  std::vector<ColGroup> cg;
  for (int i = 0; i < nlf; ++i) {
    const RefLeaf *l = L[i];
    auto rk = std::make_pair(l->nstrtl - 1, l->ndl);
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
      auto ck = std::make_pair(l->nstrtt - 1, l->ndt);
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
  lap("groups");
  st.row_groups = (int64_t)rg.size();
  st.col_groups = (int64_t)cg.size();

  // t slots ordered by row group (ascending first row), leaves in leaf order
  std::vector<int> rg_order(rg.size());
  std::iota(rg_order.begin(), rg_order.end(), 0);
  std::sort(rg_order.begin(), rg_order.end(), [&](int a, int b) {
    return rg[a].s != rg[b].s ? rg[a].s < rg[b].s : rg[a].n < rg[b].n;
  });
  std::vector<int64_t> leaf_tbase(nlf, -1);
  int64_t tcur = 0;
  for (int r : rg_order) {
    rg[r].tbase = tcur;
    for (int i : rg[r].lr) {
      leaf_tbase[i] = tcur;
      tcur += L[i]->kt;
    }
  }
  P.tsize = tcur;

  // ---- arena layout -------------------------------------------------------
  int64_t off = 0;
  int64_t vrows = 0;
  for (auto &c : cg) {
    c.ld = (int)round_up(c.n, vec_l);
    c.off = off;
    c.row0 = vrows;
    vrows += c.r;
    off = round_up(off + (int64_t)c.r * c.ld * es_l, HMV_ARENA_ALIGN);
    c.end = off;
  }
  int64_t zcur = 0;
  for (int r : rg_order) {
    RowGroup &g = rg[r];
    g.zoff = zcur;
    zcur += g.wd;
    g.beg = off;
    if (merged) {
      int nc = g.wd + g.k;
      if (nc > 0) {
        g.ld_d = (int)round_up(nc, vec_d);
        g.off_d = off;
        off = round_up(off + (int64_t)g.n * g.ld_d * es_d, HMV_ARENA_ALIGN);
      }
    } else {
      if (g.wd > 0) {
        g.ld_d = (int)round_up(g.wd, vec_d);
        g.off_d = off;
        off = round_up(off + (int64_t)g.n * g.ld_d * es_d, HMV_ARENA_ALIGN);
      }
      if (g.k > 0) {
        g.ld_u = (int)round_up(g.k, vec_l);
        g.off_u = off;
        off = round_up(off + (int64_t)g.n * g.ld_u * es_l, HMV_ARENA_ALIGN);
      }
    }
    g.end = off;
  }
  if (zcur > INT32_MAX || vrows > INT32_MAX) {
    err = "problem too large for 32-bit gather indices";
    return false;
  }
  st.arena_bytes = std::max<int64_t>(off, HMV_ARENA_ALIGN);
  st.payload_bytes = st.entries_lr * es_l + st.entries_dense * es_d;
  st.alg_bytes = st.payload_bytes + 16LL * nd;
  try {
    // left uninitialised: every block (with its padding) is zeroed by the
    // thread that fills it, so the image is first touched in parallel
    P.arena.reset(new char[(size_t)st.arena_bytes]);
    P.arena_size = (size_t)st.arena_bytes;
    P.slotmap.assign((size_t)vrows, 0);
    P.zidx.assign((size_t)zcur, 0);
  } catch (...) {
    err = "out of host memory for the arena image";
    return false;
  }
  char *A = P.arena.get();
  if (off < st.arena_bytes) std::memset(A + off, 0, (size_t)(st.arena_bytes - off));
  lap("alloc");

  // ---- fill: V^T stacks ---------------------------------------------------
  parallel_for((int64_t)cg.size(), [&](int64_t ci) {
    ColGroup &c = cg[ci];
    std::memset(A + c.off, 0, (size_t)(c.end - c.off));
    int64_t row = 0;
    for (int i : c.lr) {
      const RefLeaf *l = L[i];
      for (int k = 0; k < l->kt; ++k, ++row) {
        char *dst = A + c.off + (row * c.ld) * es_l;
        const double *src = l->a1 + (size_t)k * l->ndt;
        for (int j = 0; j < l->ndt; ++j) store(dst, dt_lr, j, src[j]);
        P.slotmap[c.row0 + row] = (int)(leaf_tbase[i] + k);
      }
    }
  });

  lap("fill V");
  // ---- fill: row panels ---------------------------------------------------
  parallel_for((int64_t)rg.size(), [&](int64_t ri) {
    RowGroup &g = rg[ri];
    if (g.end > g.beg) std::memset(A + g.beg, 0, (size_t)(g.end - g.beg));
    int64_t zc = g.zoff;
    for (int i : g.dense) {
      const RefLeaf *l = L[i];
      for (int j = 0; j < l->ndt; ++j) P.zidx[zc++] = l->nstrtt - 1 + j;
    }
    if (g.off_d >= 0) {
      for (int row = 0; row < g.n; ++row) {
        char *dst = A + g.off_d + ((int64_t)row * g.ld_d) * es_d;
        int64_t col = 0;
        for (int i : g.dense) {
          const RefLeaf *l = L[i];
          const double *src = l->a1 + (size_t)row * l->ndt;  // row `row` of A
          for (int j = 0; j < l->ndt; ++j) store(dst, dt_dense, col + j, src[j]);
          col += l->ndt;
        }
        if (merged) {
          for (int i : g.lr) {
            const RefLeaf *l = L[i];
            const double *u = reinterpret_cast<const double *>(
                reinterpret_cast<const char *>(l->a1) + l->a1size);
            for (int k = 0; k < l->kt; ++k) store(dst, dt_dense, col + k, u[row + (size_t)k * l->ndl]);
            col += l->kt;
          }
        }
      }
    }
    if (g.off_u >= 0) {
      for (int row = 0; row < g.n; ++row) {
        char *dst = A + g.off_u + ((int64_t)row * g.ld_u) * es_l;
        int64_t col = 0;
        for (int i : g.lr) {
          const RefLeaf *l = L[i];
          const double *u = reinterpret_cast<const double *>(
              reinterpret_cast<const char *>(l->a1) + l->a1size);
          for (int k = 0; k < l->kt; ++k) store(dst, dt_lr, col + k, u[row + (size_t)k * l->ndl]);
          col += l->kt;
        }
      }
    }
  });

  lap("fill panels");
  // ---- phase-1 work: pieces of the V^T stacks -----------------------------
  struct Item {
    int64_t bytes;
    HmvWork w;
  };
  std::vector<Item> it1;
  for (auto &c : cg) {
    if (c.r == 0) continue;
    int64_t rowbytes = (int64_t)c.ld * es_l;
    int rpp = (int)std::max<int64_t>(1, std::min<int64_t>(HMV_MAX_ROWS, tune.p1_target_bytes / rowbytes));
    for (int r0 = 0; r0 < c.r; r0 += rpp) {
      int nr = std::min(rpp, c.r - r0);
      HmvSeg s{};
      s.data_off = c.off + (int64_t)r0 * rowbytes;
      s.ld = c.ld;
      s.ncols = c.n;
      s.nrows = nr;
      s.acc_row = 0;
      s.nx = c.n;
      s.xsrc = c.s;
      s.tsrc = 0;
      s.dtype = (short)dt_lr;
      s.lpr = (short)choose_lpr(c.n, vec_l, nr, tune.lpr_bias);
      HmvWork w{};
      w.seg_begin = (int)P.seg1.size();
      w.seg_end = w.seg_begin + 1;
      w.out_base = (int)(c.row0 + r0);
      w.nout = nr;
      P.seg1.push_back(s);
      it1.push_back({(int64_t)nr * rowbytes, w});
    }
  }

  // ---- phase-2 work: atoms of the row partition ---------------------------
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
  const int natoms = (int)bnd.size() - 1;
  st.atoms = natoms;
  // per-atom work estimate (payload bytes of the panel rows that cover it)
  std::vector<int64_t> atom_bytes(natoms, 0);
  for (auto &g : rg) {
    int a0 = (int)(std::lower_bound(bnd.begin(), bnd.end(), g.s) - bnd.begin());
    int64_t per_row = (merged ? (int64_t)(g.wd + g.k) * es_d : (int64_t)g.wd * es_d + (int64_t)g.k * es_l);
    for (int a = a0; a < natoms && bnd[a] < g.s + g.n; ++a)
      atom_bytes[a] += per_row * (bnd[a + 1] - bnd[a]);
  }
  // merge consecutive atoms into workgroups
  std::vector<std::pair<int, int>> ranges;  // [row begin, row end)
  {
    int a = 0;
    while (a < natoms) {
      int b = a;
      int64_t bytes = 0;
      while (b < natoms) {
        int rows_if = bnd[b + 1] - bnd[a];
        if (b > a && (rows_if > tune.p2_max_rows || bytes + atom_bytes[b] > tune.p2_target_bytes)) break;
        if (b == a && bnd[b + 1] - bnd[b] > tune.p2_max_rows) break;
        bytes += atom_bytes[b];
        ++b;
      }
      if (b == a) {
        // a single atom taller than the row limit: split it into row slices
        int r = bnd[a];
        while (r < bnd[a + 1]) {
          int e = std::min(bnd[a + 1], r + tune.p2_max_rows);
          ranges.push_back({r, e});
          r = e;
        }
        ++a;
      } else {
        ranges.push_back({bnd[a], bnd[b]});
        a = b;
      }
    }
  }
  // row groups sorted by first row for the overlap search; groups that
  // contain a row form a chain, found by scanning groups that start at or
  // before the range and end after it.
  std::vector<int> by_start(rg.size());
  std::iota(by_start.begin(), by_start.end(), 0);
  std::sort(by_start.begin(), by_start.end(), [&](int a, int b) { return rg[a].s < rg[b].s; });
  // For each range collect overlapping groups.  Use an interval stabbing
  // structure: process ranges in order with an active set.
  std::vector<Item> it2;
  it2.reserve(ranges.size());
  {
    std::vector<int> active;
    size_t next = 0;
    for (auto &rr : ranges) {
      const int a = rr.first, b = rr.second;
      while (next < by_start.size() && rg[by_start[next]].s < b) active.push_back(by_start[next++]);
      // drop groups that ended at or before a
      active.erase(std::remove_if(active.begin(), active.end(),
                                  [&](int gi) { return rg[gi].s + rg[gi].n <= a; }),
                   active.end());
      HmvWork w{};
      w.seg_begin = (int)P.seg2.size();
      w.out_base = a;
      w.nout = b - a;
      int64_t bytes = 0;
      // deterministic order: groups by first row, then size
      std::vector<int> cov;
      for (int gi : active)
        if (rg[gi].s < b && rg[gi].s + rg[gi].n > a) cov.push_back(gi);
      std::sort(cov.begin(), cov.end(), [&](int x, int y) {
        return rg[x].s != rg[y].s ? rg[x].s < rg[y].s : rg[x].n > rg[y].n;
      });
      for (int gi : cov) {
        const RowGroup &g = rg[gi];
        int lo = std::max(a, g.s), hi = std::min(b, g.s + g.n);
        if (g.off_d >= 0) {
          HmvSeg s{};
          s.data_off = g.off_d + (int64_t)(lo - g.s) * g.ld_d * es_d;
          s.ld = g.ld_d;
          s.ncols = merged ? g.wd + g.k : g.wd;
          s.nrows = hi - lo;
          s.acc_row = lo - a;
          s.nx = g.wd;
          s.xsrc = (int)g.zoff;
          s.tsrc = (int)g.tbase;
          s.dtype = (short)dt_dense;
          s.lpr = (short)choose_lpr(s.ncols, vec_d, s.nrows, tune.lpr_bias);
          P.seg2.push_back(s);
          bytes += (int64_t)s.nrows * s.ld * es_d;
        }
        if (g.off_u >= 0) {
          HmvSeg s{};
          s.data_off = g.off_u + (int64_t)(lo - g.s) * g.ld_u * es_l;
          s.ld = g.ld_u;
          s.ncols = g.k;
          s.nrows = hi - lo;
          s.acc_row = lo - a;
          s.nx = 0;
          s.xsrc = 0;
          s.tsrc = (int)g.tbase;
          s.dtype = (short)dt_lr;
          s.lpr = (short)choose_lpr(s.ncols, vec_l, s.nrows, tune.lpr_bias);
          P.seg2.push_back(s);
          bytes += (int64_t)s.nrows * s.ld * es_l;
        }
      }
      w.seg_end = (int)P.seg2.size();
      it2.push_back({bytes, w});
    }
  }

  lap("work lists");
  // largest workgroups first so the dispatcher's in-order launch ends balanced
  auto by_size = [](const Item &x, const Item &y) { return x.bytes > y.bytes; };
  std::stable_sort(it1.begin(), it1.end(), by_size);
  std::stable_sort(it2.begin(), it2.end(), by_size);
  for (auto &i : it1) P.work1.push_back(i.w);
  for (auto &i : it2) P.work2.push_back(i.w);
  return true;
}

}  // namespace hmv
