// hmv_plan.h -- host-side repack of a leaf list into the engine's arena.
#ifndef HMV_PLAN_H
#define HMV_PLAN_H

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "hmv_internal.h"

namespace hmv {

// A leaf in the reference's record layout (stc_HACApK_leafmtx,
// HACApK_MAGMA.h:12-20); a2 is located at (char*)a1 + a1size.
struct RefLeaf {
  int ltmtx;
  int kt;
  int nstrtl, ndl;
  int nstrtt, ndt;
  long long a1size;
  double *a1;
  double *a2;
};

// Storage precision of an arena.
enum Precision { kF64 = 0, kF32 = 1, kBF16 = 2, kBF16All = 3 };

struct PlanStats {
  int64_t nlf = 0, nlr = 0, ndense = 0, nkt0 = 0;
  int64_t entries_lr = 0;     // sum_LR kt*(ndl+ndt)
  int64_t entries_dense = 0;  // sum_D ndl*ndt
  int64_t tslots = 0;         // sum_LR kt
  int64_t tparts = 0;         // t buffer entries (slots + chunk partials)
  int64_t col_groups = 0, row_groups = 0, atoms = 0;
  int64_t arena_bytes = 0;    // bytes of the arena including padding
  int64_t payload_bytes = 0;  // bytes of factor/dense entries at storage precision
  int64_t alg_bytes = 0;      // payload + x read once + y written once (8 B each)
  int64_t phase1_bytes = 0;   // arena bytes read by phase 1 (incl. row padding)
  int64_t phase2_bytes = 0;   // arena bytes read by phase 2 (incl. row padding)
  int ktmax = 0;
  int max_ndt = 0;
};

// Allocator whose resize leaves plain records uninitialised (they are all
// written afterwards, in parallel).
template <typename T>
struct NoInitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <typename U>
  NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
  template <typename U>
  void construct(U *p) noexcept {
    ::new (static_cast<void *>(p)) U;
  }
  template <typename U, typename... A>
  void construct(U *p, A &&...a) {
    ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
  }
};
using FillOps = std::vector<HmvFillOp, NoInitAlloc<HmvFillOp>>;

struct Plan {
  int nd = 0;
  int prec = kF64;
  int64_t tsize = 0;               // t slots (one per rank vector)
  int64_t tpart_size = 0;          // t buffer entries: tsize slots, then chunk partials
  std::unique_ptr<char[]> arena;   // device image of the arena (arena_size bytes)
  size_t arena_size = 0;
  std::vector<HmvWork> work1, work2;
  int nwork1_small = 0, nwork2_small = 0;  // short blocks at the front of each list
  int nt1_small = 0, nt2_small = 0;        // their workgroup threads (0: none)
  std::vector<HmvZSeg> zseg;       // input sources of both phases
  std::vector<int> outidx;         // phase-1 output row -> t partial slot
  std::vector<int> codes;          // phase-2 column input codes (HMV_Z_CODES)
  std::vector<HmvTFin> tfin;       // chunked rank vectors: t slot = sum of partials
  PlanStats st;
  struct PlanTuningRec {           // the tunables the plan was built with
    int64_t p1_target_bytes = 0;
    int p1_row_tile = 0, p2_max_rows = 0, p1_small_threads = 0, p2_small_threads = 0;
  } tune;
  std::vector<int> seg;            // instance row offsets (ninst + 1): {0, nd} for one H-matrix
  // device-fill plans (leaf factors in device memory): no host image; the
  // copies that make it, run on the device at upload
  FillOps fill;
};

// Tunables of the repack (chosen by the autotuner or left at defaults).
struct PlanTuning {
  int64_t p1_target_bytes = 96 * 1024;  // arena bytes per phase-1 workgroup
  int p1_row_tile = 256;                // max rank vectors per phase-1 workgroup
  int p2_max_rows = HMV_MAX_ROWS;       // max rows per phase-2 workgroup
  int p1_small_threads = -1;            // 64/128: short phase-1 blocks in smaller workgroups
  int p2_small_threads = -1;            // same for phase 2; -1: off
  int xcd_groups = 8;                   // dispatcher round-robin width (XCDs)
};
// Defaults of a storage precision (measured on MI355X: the narrow types
// gain from more, smaller workgroups on short blocks; FP64 does not).
PlanTuning default_tuning(int prec);

// Builds the plan.  `leaves` points at nlf records spaced `stride` bytes.
// Returns false and fills `err` on invalid input.
bool build_plan(int nd, int nlf, const void *leaves, int stride, int prec,
                const PlanTuning &tune, Plan &out, std::string &err);
// Same, for leaves whose a1/a2 point to device memory: the arena image is
// not built on the host; out.fill receives the copies that build it.
bool build_plan_device(int nd, int nlf, const void *leaves, int stride, int prec,
                       const PlanTuning &tune, Plan &out, std::string &err);

// Checks every index of a plan against the buffers it addresses (work
// blocks inside the arena image, input runs, codes and outputs inside x, t
// and y, partial sums inside t), so that a plan read from a file cannot make
// a kernel read or write out of bounds.  Returns false with `err` set.
bool validate_plan(const Plan &P, std::string &err);

// Versioned arena file (hmv_file.cpp): write a plan, and read one back
// (checksums and validate_plan included).
bool save_plan_file(const Plan &P, const char *path, std::string &err);
bool load_plan_file(const char *path, Plan &P, std::string &err);

int dtype_size(int dt);
// Rows handled per lane (each lane loads `row_vector(dt)` consecutive rows):
// phase-1 blocks by their low-rank type, phase-2 blocks by their first
// part's type (two-row lanes for the narrow types).
int row_vector(int dt);
int row_vector_p2(int dt);

}  // namespace hmv

#endif
