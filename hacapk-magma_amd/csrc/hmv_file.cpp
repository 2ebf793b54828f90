// hmv_file.cpp -- versioned binary file of a repacked arena, and the index
// checks that make a plan from a file safe to upload.
//
// The reference has no such file: it rebuilds the H-matrix and the device
// copy for every run (HACApK_SOLVER.c:521,777; param(12) "write down the
// H-matrix", m_HACApK_base.f90:300, is declared and never used).  Here the
// repack (construction -> apply interface) can be written once and loaded
// by later runs without the leaves.
//
// Layout (little-endian): a 512-byte header, then the sections, each at a
// 64-byte-aligned offset.  The header holds the format version, the sizes
// of the record types, the plan's scalars and statistics, and for every
// section its offset, element count, byte count and a 64-bit checksum
// (FNV-1a over 8-byte words, then the tail bytes).
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "hmv.h"
#include "hmv_engine.h"
#include "hmv_internal.h"
#include "hmv_plan.h"

namespace hmv {

namespace {

constexpr char kMagic[8] = {'H', 'M', 'V', 'A', 'R', 'E', 'N', 'A'};
constexpr uint32_t kVersion = 1;
constexpr int kSections = 8;  // arena, work1, work2, zseg, outidx, codes, tfin, seg
constexpr uint64_t kAlign = 64;
const char *const kSectionName[kSections] = {"arena", "work1", "work2", "zseg",
                                             "outidx", "codes", "tfin", "seg"};

struct Section {
  uint64_t offset, count, bytes, checksum;
};

struct FileHeader {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  uint32_t work_bytes, zseg_bytes, tfin_bytes, pad0;
  int32_t nd, prec;
  int64_t tsize, tpart_size;
  int32_t nwork1_small, nt1_small, nwork2_small, nt2_small;
  int64_t p1_target_bytes;
  int32_t p1_row_tile, p2_max_rows, p1_small_threads, p2_small_threads;
  // PlanStats
  int64_t nlf, nlr, ndense, nkt0, entries_lr, entries_dense, tslots, tparts;
  int64_t col_groups, row_groups, atoms, arena_bytes, payload_bytes, alg_bytes;
  int64_t phase1_bytes, phase2_bytes;
  int32_t ktmax, max_ndt;
  Section sec[kSections];
  char reserved[512 - 8 - 6 * 4 - 2 * 4 - 2 * 8 - 4 * 4 - 8 - 4 * 4 - 16 * 8 - 2 * 4 -
                kSections * 32];
};
static_assert(sizeof(FileHeader) == 512, "file header is 512 bytes");
static_assert(sizeof(HmvWork) == 48 && sizeof(HmvZSeg) == 16 && sizeof(HmvTFin) == 16,
              "record layouts of the file");

uint64_t checksum(const void *p, size_t n) {
  const unsigned char *b = static_cast<const unsigned char *>(p);
  uint64_t h = 0xcbf29ce484222325ull;
  const uint64_t prime = 0x100000001b3ull;
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    std::memcpy(&w, b + i, 8);
    h = (h ^ w) * prime;
  }
  for (; i < n; ++i) h = (h ^ b[i]) * prime;
  return h;
}

uint64_t align_up(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct Blob {
  const void *data;
  uint64_t count, elem;
};

bool write_file(const Plan &P, const char *path, std::string &err) {
  FileHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.header_bytes = sizeof(FileHeader);
  h.work_bytes = sizeof(HmvWork);
  h.zseg_bytes = sizeof(HmvZSeg);
  h.tfin_bytes = sizeof(HmvTFin);
  h.nd = P.nd;
  h.prec = P.prec;
  h.tsize = P.tsize;
  h.tpart_size = P.tpart_size;
  h.nwork1_small = P.nwork1_small;
  h.nt1_small = P.nt1_small;
  h.nwork2_small = P.nwork2_small;
  h.nt2_small = P.nt2_small;
  h.p1_target_bytes = P.tune.p1_target_bytes;
  h.p1_row_tile = P.tune.p1_row_tile;
  h.p2_max_rows = P.tune.p2_max_rows;
  h.p1_small_threads = P.tune.p1_small_threads;
  h.p2_small_threads = P.tune.p2_small_threads;
  const PlanStats &s = P.st;
  h.nlf = s.nlf;
  h.nlr = s.nlr;
  h.ndense = s.ndense;
  h.nkt0 = s.nkt0;
  h.entries_lr = s.entries_lr;
  h.entries_dense = s.entries_dense;
  h.tslots = s.tslots;
  h.tparts = s.tparts;
  h.col_groups = s.col_groups;
  h.row_groups = s.row_groups;
  h.atoms = s.atoms;
  h.arena_bytes = s.arena_bytes;
  h.payload_bytes = s.payload_bytes;
  h.alg_bytes = s.alg_bytes;
  h.phase1_bytes = s.phase1_bytes;
  h.phase2_bytes = s.phase2_bytes;
  h.ktmax = s.ktmax;
  h.max_ndt = s.max_ndt;
  const Blob blobs[kSections] = {
      {P.arena.get(), P.arena_size, 1},
      {P.work1.data(), P.work1.size(), sizeof(HmvWork)},
      {P.work2.data(), P.work2.size(), sizeof(HmvWork)},
      {P.zseg.data(), P.zseg.size(), sizeof(HmvZSeg)},
      {P.outidx.data(), P.outidx.size(), sizeof(int)},
      {P.codes.data(), P.codes.size(), sizeof(int)},
      {P.tfin.data(), P.tfin.size(), sizeof(HmvTFin)},
      {P.seg.data(), P.seg.size(), sizeof(int)},
  };
  uint64_t off = align_up(sizeof(FileHeader));
  for (int i = 0; i < kSections; ++i) {
    Section &sc = h.sec[i];
    sc.offset = off;
    sc.count = blobs[i].count;
    sc.bytes = blobs[i].count * blobs[i].elem;
    sc.checksum = checksum(blobs[i].data, sc.bytes);
    off = align_up(off + sc.bytes);
  }
  std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(path, "wb"), std::fclose);
  if (!f) {
    err = std::string("cannot open ") + path + " for writing";
    return false;
  }
  static const char zeros[kAlign] = {};
  bool ok = std::fwrite(&h, sizeof(h), 1, f.get()) == 1;
  uint64_t pos = sizeof(h);
  for (int i = 0; i < kSections && ok; ++i) {
    ok = std::fwrite(zeros, 1, h.sec[i].offset - pos, f.get()) == h.sec[i].offset - pos;
    pos = h.sec[i].offset;
    if (ok && h.sec[i].bytes > 0)
      ok = std::fwrite(blobs[i].data, 1, h.sec[i].bytes, f.get()) == h.sec[i].bytes;
    pos += h.sec[i].bytes;
  }
  if (!ok || std::fflush(f.get()) != 0) {
    err = std::string("write to ") + path + " failed";
    return false;
  }
  return true;
}

template <typename T>
bool read_section(FILE *f, const Section &sc, uint64_t fsize, std::vector<T> &out, const char *name,
                  std::string &err) {
  if (sc.bytes != sc.count * sizeof(T) || sc.offset % kAlign != 0 || sc.offset > fsize ||
      sc.bytes > fsize - sc.offset) {
    err = std::string("section ") + name + ": size or offset out of the file";
    return false;
  }
  out.resize(sc.count);
  if (sc.bytes > 0 && (std::fseek(f, (long)sc.offset, SEEK_SET) != 0 ||
                       std::fread(out.data(), 1, sc.bytes, f) != sc.bytes)) {
    err = std::string("section ") + name + ": read failed";
    return false;
  }
  if (checksum(out.data(), sc.bytes) != sc.checksum) {
    err = std::string("section ") + name + ": checksum mismatch";
    return false;
  }
  return true;
}

bool read_file(const char *path, Plan &P, std::string &err) {
  std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(path, "rb"), std::fclose);
  if (!f) {
    err = std::string("cannot open ") + path;
    return false;
  }
  if (std::fseek(f.get(), 0, SEEK_END) != 0) {
    err = "cannot size the file";
    return false;
  }
  const long fend = std::ftell(f.get());
  if (fend < 0 || std::fseek(f.get(), 0, SEEK_SET) != 0) {
    err = "cannot size the file";
    return false;
  }
  const uint64_t fsize = (uint64_t)fend;
  FileHeader h;
  if (fsize < sizeof(h) || std::fread(&h, sizeof(h), 1, f.get()) != 1) {
    err = "file shorter than the header";
    return false;
  }
  if (std::memcmp(h.magic, kMagic, 8) != 0) {
    err = "not an arena file (bad magic)";
    return false;
  }
  if (h.version != kVersion) {
    err = "unsupported arena file version " + std::to_string(h.version) + " (this library reads " +
          std::to_string(kVersion) + ")";
    return false;
  }
  if (h.header_bytes != sizeof(FileHeader) || h.work_bytes != sizeof(HmvWork) ||
      h.zseg_bytes != sizeof(HmvZSeg) || h.tfin_bytes != sizeof(HmvTFin)) {
    err = "record sizes of the file differ from this library's";
    return false;
  }
  P.nd = h.nd;
  P.prec = h.prec;
  P.tsize = h.tsize;
  P.tpart_size = h.tpart_size;
  P.nwork1_small = h.nwork1_small;
  P.nt1_small = h.nt1_small;
  P.nwork2_small = h.nwork2_small;
  P.nt2_small = h.nt2_small;
  P.tune.p1_target_bytes = h.p1_target_bytes;
  P.tune.p1_row_tile = h.p1_row_tile;
  P.tune.p2_max_rows = h.p2_max_rows;
  P.tune.p1_small_threads = h.p1_small_threads;
  P.tune.p2_small_threads = h.p2_small_threads;
  PlanStats &s = P.st;
  s.nlf = h.nlf;
  s.nlr = h.nlr;
  s.ndense = h.ndense;
  s.nkt0 = h.nkt0;
  s.entries_lr = h.entries_lr;
  s.entries_dense = h.entries_dense;
  s.tslots = h.tslots;
  s.tparts = h.tparts;
  s.col_groups = h.col_groups;
  s.row_groups = h.row_groups;
  s.atoms = h.atoms;
  s.arena_bytes = h.arena_bytes;
  s.payload_bytes = h.payload_bytes;
  s.alg_bytes = h.alg_bytes;
  s.phase1_bytes = h.phase1_bytes;
  s.phase2_bytes = h.phase2_bytes;
  s.ktmax = h.ktmax;
  s.max_ndt = h.max_ndt;
  // arena image
  const Section &sa = h.sec[0];
  if (sa.offset % kAlign != 0 || sa.offset > fsize || sa.bytes > fsize - sa.offset ||
      sa.count != sa.bytes) {
    err = "section arena: size or offset out of the file";
    return false;
  }
  try {
    P.arena.reset(new char[sa.bytes > 0 ? sa.bytes : 1]);
  } catch (...) {
    err = "out of host memory for the arena image";
    return false;
  }
  P.arena_size = sa.bytes;
  if (sa.bytes > 0 && (std::fseek(f.get(), (long)sa.offset, SEEK_SET) != 0 ||
                       std::fread(P.arena.get(), 1, sa.bytes, f.get()) != sa.bytes)) {
    err = "section arena: read failed";
    return false;
  }
  if (checksum(P.arena.get(), sa.bytes) != sa.checksum) {
    err = "section arena: checksum mismatch";
    return false;
  }
  return read_section(f.get(), h.sec[1], fsize, P.work1, kSectionName[1], err) &&
         read_section(f.get(), h.sec[2], fsize, P.work2, kSectionName[2], err) &&
         read_section(f.get(), h.sec[3], fsize, P.zseg, kSectionName[3], err) &&
         read_section(f.get(), h.sec[4], fsize, P.outidx, kSectionName[4], err) &&
         read_section(f.get(), h.sec[5], fsize, P.codes, kSectionName[5], err) &&
         read_section(f.get(), h.sec[6], fsize, P.tfin, kSectionName[6], err) &&
         read_section(f.get(), h.sec[7], fsize, P.seg, kSectionName[7], err);
}

bool fail(std::string &err, const std::string &msg) {
  err = msg;
  return false;
}

}  // namespace

bool validate_plan(const Plan &P, std::string &err) {
  if (P.nd <= 0) return fail(err, "nd must be positive");
  if (P.prec != HMV_PREC_F64 && P.prec != HMV_PREC_F32 && P.prec != HMV_PREC_BF16 &&
      P.prec != HMV_PREC_BF16_ALL)
    return fail(err, "unknown precision");
  const int dt_dense = P.prec == HMV_PREC_F64        ? HMV_DT_F64
                       : P.prec == HMV_PREC_BF16_ALL ? HMV_DT_BF16
                                                     : HMV_DT_F32;
  const int dt_lr = P.prec == HMV_PREC_F64   ? HMV_DT_F64
                    : P.prec == HMV_PREC_F32 ? HMV_DT_F32
                                             : HMV_DT_BF16;
  if (P.tsize < 0 || P.tpart_size < P.tsize || P.tpart_size > INT32_MAX)
    return fail(err, "t buffer sizes out of range");
  if (P.seg.size() < 2 || P.seg.front() != 0 || P.seg.back() != P.nd)
    return fail(err, "instance offsets must run from 0 to nd");
  for (size_t i = 1; i < P.seg.size(); ++i)
    if (P.seg[i] <= P.seg[i - 1]) return fail(err, "instance offsets must increase");
  const int64_t nz = (int64_t)P.zseg.size(), nout = (int64_t)P.outidx.size();
  const int64_t ncodes = (int64_t)P.codes.size();
  for (int64_t i = 0; i < ncodes; ++i) {
    const int c = P.codes[i];
    if (c >= 0 ? c >= P.nd : (int64_t)(-(int64_t)c - 1) >= P.tsize)
      return fail(err, "input code " + std::to_string(i) + " out of range");
  }
  for (int64_t i = 0; i < nout; ++i)
    if (P.outidx[i] < 0 || P.outidx[i] >= P.tpart_size)
      return fail(err, "phase-1 output index " + std::to_string(i) + " out of range");
  for (const HmvTFin &t : P.tfin)
    if (t.slot < 0 || t.slot >= P.tsize || t.base < 0 || t.np < 2 ||
        (int64_t)t.base + t.np > P.tpart_size)
      return fail(err, "partial-sum record out of range");
  auto check_list = [&](const std::vector<HmvWork> &wl, int phase, int nsmall, int nt_small) {
    const std::string ph = "phase-" + std::to_string(phase) + " block ";
    if (nsmall < 0 || nsmall > (int64_t)wl.size()) return fail(err, ph + "count of short blocks");
    if (nsmall > 0 ? (nt_small != 64 && nt_small != 128) : nt_small != 0)
      return fail(err, ph + "threads of short blocks");
    for (size_t i = 0; i < wl.size(); ++i) {
      const HmvWork &w = wl[i];
      const std::string id = ph + std::to_string(i);
      const int want0 = phase == 1 ? dt_lr : dt_dense;
      if (w.dt0 != want0 || w.dt1 != dt_lr) return fail(err, id + ": storage types");
      const int vr = phase == 1 ? row_vector(w.dt0) : row_vector_p2(w.dt0);
      const int nt = (int64_t)i < nsmall ? nt_small : HMV_WG_THREADS;
      if (w.rows < 1 || w.rows > HMV_MAX_ROWS || w.rows_p < w.rows || w.rows_p % vr != 0 ||
          w.rows_p - w.rows >= vr || w.rows_p / vr > nt)
        return fail(err, id + ": rows");
      // one storage type (phase 1, or dense and low rank alike): one part
      if (w.ncol0 < 0 || w.ncol1 < 0 || ((phase == 1 || w.dt0 == w.dt1) && w.ncol1 != 0))
        return fail(err, id + ": columns");
      auto part_ok = [&](int64_t off, int ncol, int dt) {
        if (ncol == 0) return true;
        const int64_t bytes = (int64_t)w.rows_p * ncol * dtype_size(dt);
        // the kernels address a part with 32-bit offsets and use 2^31 as
        // their out-of-range offset
        return off >= 0 && off % 16 == 0 && bytes < (int64_t)0x80000000LL &&
               off + bytes <= (int64_t)P.arena_size;
      };
      if (!part_ok(w.off0, w.ncol0, w.dt0) || !part_ok(w.off1, w.ncol1, w.dt1))
        return fail(err, id + ": block outside the arena image");
      const int ncols = w.ncol0 + w.ncol1;
      if (w.zseg_begin < 0 || w.zseg_end < w.zseg_begin || w.zseg_end > nz)
        return fail(err, id + ": input segments out of range");
      const int ns = w.zseg_end - w.zseg_begin;
      if (phase == 1 && ns != 1) return fail(err, id + ": one input run expected");
      if (ns == 0 && ncols > 0) return fail(err, id + ": inputs missing");
      const HmvZSeg *zs = P.zseg.data() + w.zseg_begin;
      if (ns == 1 && zs[0].kind == HMV_Z_CODES) {
        if (phase == 1) return fail(err, id + ": phase-1 inputs come from x");
        if (zs[0].ncols != ncols || zs[0].src < 0 || (int64_t)zs[0].src + ncols > ncodes)
          return fail(err, id + ": code table out of range");
      } else if (ns > 0) {
        if (ns > HMV_RUN_MAX) return fail(err, id + ": more runs than the run table holds");
        int64_t col = 0;
        for (int k = 0; k < ns; ++k) {
          const HmvZSeg &z = zs[k];
          if (z.ncols <= 0 || z.col != col || z.src < 0) return fail(err, id + ": input runs");
          if (z.kind == HMV_Z_X) {
            if ((int64_t)z.src + z.ncols > P.nd) return fail(err, id + ": x run out of range");
          } else if (z.kind == HMV_Z_T && phase == 2) {
            if ((int64_t)z.src + z.ncols > P.tsize) return fail(err, id + ": t run out of range");
          } else {
            return fail(err, id + ": input run kind");
          }
          col += z.ncols;
        }
        if (col != ncols) return fail(err, id + ": input runs do not cover the columns");
      }
      if (phase == 1 ? (w.out_base < 0 || (int64_t)w.out_base + w.rows > nout)
                     : (w.out_base < 0 || (int64_t)w.out_base + w.rows > P.nd))
        return fail(err, id + ": outputs out of range");
    }
    return true;
  };
  return check_list(P.work1, 1, P.nwork1_small, P.nt1_small) &&
         check_list(P.work2, 2, P.nwork2_small, P.nt2_small);
}

bool save_plan_file(const Plan &P, const char *path, std::string &err) {
  return write_file(P, path, err);
}

bool load_plan_file(const char *path, Plan &P, std::string &err) {
  if (!read_file(path, P, err)) return false;
  return validate_plan(P, err);
}

// Copies an arena's device tables back into a plan (for saving).
bool download_plan(const hmv_arena_s *a, Plan &P, std::string &err) {
  P.nd = a->nd;
  P.prec = a->prec;
  P.tsize = a->tsize;
  P.tpart_size = a->tpart_size;
  P.nwork1_small = a->nwork1_small;
  P.nt1_small = a->nt1_small;
  P.nwork2_small = a->nwork2_small;
  P.nt2_small = a->nt2_small;
  P.tune.p1_target_bytes = a->tuning.p1_target_bytes;
  P.tune.p1_row_tile = a->tuning.p1_row_tile;
  P.tune.p2_max_rows = a->tuning.p2_max_rows;
  P.tune.p1_small_threads = a->tuning.p1_small_threads;
  P.tune.p2_small_threads = a->tuning.p2_small_threads;
  P.st = a->pstats;
  P.seg = a->seg;
  try {
    P.arena.reset(new char[a->arena_size > 0 ? a->arena_size : 1]);
  } catch (...) {
    err = "out of host memory for the arena image";
    return false;
  }
  P.arena_size = a->arena_size;
  P.work1.resize(a->nwork1);
  P.work2.resize(a->nwork2);
  P.zseg.resize(a->nzseg);
  P.outidx.resize(a->noutidx);
  P.codes.resize(a->ncodes);
  P.tfin.resize(a->ntfin);
  auto get = [](void *dst, const void *src, size_t bytes) {
    return bytes == 0 || hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
  };
  if (!get(P.arena.get(), a->d_arena, P.arena_size) ||
      !get(P.work1.data(), a->d_work1, sizeof(HmvWork) * P.work1.size()) ||
      !get(P.work2.data(), a->d_work2, sizeof(HmvWork) * P.work2.size()) ||
      !get(P.zseg.data(), a->d_zseg, sizeof(HmvZSeg) * P.zseg.size()) ||
      !get(P.outidx.data(), a->d_outidx, sizeof(int) * P.outidx.size()) ||
      !get(P.codes.data(), a->d_codes, sizeof(int) * P.codes.size()) ||
      !get(P.tfin.data(), a->d_tfin, sizeof(HmvTFin) * P.tfin.size())) {
    err = "device to host copy failed";
    return false;
  }
  return true;
}

}  // namespace hmv
