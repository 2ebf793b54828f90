// hmv_engine.h -- internal arena object shared by the API and solver files.
#ifndef HMV_ENGINE_H
#define HMV_ENGINE_H

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "hmv.h"
#include "hmv_internal.h"
#include "hmv_plan.h"

struct HmvSolveWork;

// H-matmat work units: one wave per (work item, group of up to 1 << w
// consecutive 16-row tiles), unit = {item, first tile << 8 | tiles}; one
// list per phase and width w < HMV_MM_WIDTHS.
#define HMV_MM_WIDTHS 3
struct HmvUnits {
  const int2 *u[2][HMV_MM_WIDTHS];
  int n[2][HMV_MM_WIDTHS];
  // Path choices per right-hand-side class (set by hmv_block_tuning_set /
  // hmv_autotune_block): vec[c] = 1 runs k <= 8 in class c (k <= 1, 2, 4,
  // 8) on the vector units; log_tw[c] = log2 of the 16-row tiles per MFMA
  // unit for KT class c (1, 2, 3-4, 5-8 tiles of 16 right-hand sides), -1
  // for the kernel's default.
  int vec[4];
  int log_tw[4];
};

struct hmv_arena_s {
  int device = 0;
  int nd = 0;
  int prec = 0;
  int acc64 = 1;
  int mode = 0;  // kernel specialisation: 0 FP64, 1 FP32, 2 BF16 low rank + FP32 dense
  int64_t tsize = 0;       // t slots
  int64_t tpart_size = 0;  // t buffer entries (slots + chunk partials)
  char *d_arena = nullptr;
  HmvWork *d_work1 = nullptr, *d_work2 = nullptr;
  int nwork1 = 0, nwork2 = 0;
  HmvTFin *d_tfin = nullptr;  // chunked rank vectors
  int ntfin = 0;
  HmvZSeg *d_zseg = nullptr;
  int *d_outidx = nullptr, *d_codes = nullptr;
  double *d_t = nullptr;   // t slots + chunk partials (FP64 in every mode)
  HmvLaunch launch{};      // device tables of one apply
  int2 *d_units[2][HMV_MM_WIDTHS] = {};  // H-matmat units per phase and tile-group width
  int nunits[2][HMV_MM_WIDTHS] = {};
  double *d_tk = nullptr;  // t buffer of the H-matmat (tpart_size * k doubles)
  int64_t tk_cap = 0;      // its capacity in doubles
  int bt_vec[4] = {1, 1, 1, 1};      // H-matmat path choices (HmvUnits::vec, log_tw)
  int bt_log_tw[4] = {-1, -1, -1, -1};
  std::vector<int> seg;    // instance row offsets (ninst + 1) of a multi-instance arena
  int *d_winst1 = nullptr, *d_winst2 = nullptr;  // instance of each work item (multi only)
  hmv_tuning tuning{};     // the tuning the arena was built with
  // hmv_apply_host: device x, pinned host staging of x and y, y's device
  // address (the apply writes H zu there directly), its stream (made on the
  // first call)
  double *d_hx = nullptr, *h_x = nullptr, *h_y = nullptr, *h_y_dev = nullptr;
  hipStream_t hstream = nullptr;
  HmvSolveWork *solve = nullptr;             // solver workspace (lazy)
  struct HmvGcrWork *gcr = nullptr;          // GCR(m) workspace (lazy)
  hmv_arena_stats stats{};
  // what a saved arena file needs besides the device tables
  int64_t nzseg = 0, noutidx = 0, ncodes = 0;
  size_t arena_size = 0;
  int nwork1_small = 0, nt1_small = 0, nwork2_small = 0, nt2_small = 0;
  hmv::PlanStats pstats{};
};

namespace hmv {
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);
// Enqueues y = H x + beta*y; `skip` (device int, may be null) makes every
// workgroup return early when *skip != 0.
hipError_t enqueue_apply(hmv_arena_s *a, const double *x, double *y, double beta, const int *skip,
                         hipStream_t s, const int *inst_done = nullptr);
void free_solve_work(hmv_arena_s *a);
void free_gcr_work(hmv_arena_s *a);
// The arena's H-matmat unit lists and path choices.
void fill_units(const hmv_arena_s *a, HmvUnits &u);
// Enqueues Y = H X + beta*Y for k interleaved right-hand sides (1..128).
hipError_t enqueue_apply_block(hmv_arena_s *a, int k, const double *x, double *y, double beta,
                               const int *skip, hipStream_t s);
// Runs an arena's fill records on the device (image zeroed beforehand).
hipError_t launch_fill(const FillOps &ops, char *arena);
// Arena from leaves whose factors are in device memory (hmv_construct).
int build_from_device_leaves(int nd, int nlf, const void *leaves, int stride, int prec,
                             const hmv_tuning *tuning, hmv_arena *out);
// Copies an arena's device tables back into a host plan (for saving).
bool download_plan(const hmv_arena_s *a, Plan &P, std::string &err);
// Makes the H-matmat t buffer hold k right-hand sides.
int ensure_tk(hmv_arena_s *a, int k);
// Device address of the current device's static workspace (hmv_matmat.hip).
int static_scratch(double **out);
}  // namespace hmv

#define HMV_MFMA_RATE_WAVES (256 * 4 * 4)  // waves of the MFMA-rate calibration
extern "C" hipError_t hmv_launch_stream_test(int kind, const void *src, void *dst, size_t bytes,
                                             hipStream_t stream);
extern "C" hipError_t hmv_launch_apply(const HmvLaunch *L, int phases, const double *x, double *y,
                                       double beta, const int *skip, const int *inst_done,
                                       hipStream_t stream);
// H-matmats with at most this many right-hand sides run on the vector units
// (hmv_kstream.hip), larger ones on MFMA (hmv_matmat.hip): measured at
// 327,680 unknowns, the vector kernel is the faster one up to k = 8 with
// FP64 factors and up to k = 4 with FP32 / BF16 factors (at k = 8 their
// MFMA kernel, one 16-wide tile of right-hand sides, takes 0.92 ms against
// 1.25 ms).
#define HMV_KSTREAM_MAX_F64 8
#define HMV_KSTREAM_MAX_NARROW 4
extern "C" hipError_t hmv_launch_kstream(const HmvLaunch *L, int phase, int k, const double *x,
                                         double *t, double *y, double beta, const int *skip,
                                         hipStream_t stream);
extern "C" hipError_t hmv_launch_block(const HmvLaunch *L, const HmvUnits *units, int k,
                                       const double *x, double *t, double *y, double beta,
                                       const int *skip, hipStream_t stream);

#endif
