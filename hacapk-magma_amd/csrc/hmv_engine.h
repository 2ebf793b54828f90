// hmv_engine.h -- internal arena object shared by the API and solver files.
#ifndef HMV_ENGINE_H
#define HMV_ENGINE_H

#include <hip/hip_runtime.h>

#include <string>

#include "hmv.h"
#include "hmv_internal.h"

struct HmvSolveWork;

struct hmv_arena_s {
  int device = 0;
  int nd = 0;
  int prec = 0;
  int acc64 = 1;
  int64_t tsize = 0;
  char *d_arena = nullptr;
  HmvSeg *d_seg1 = nullptr, *d_seg2 = nullptr;
  HmvWork *d_work1 = nullptr, *d_work2 = nullptr;
  int nwork1 = 0, nwork2 = 0;
  int *d_slotmap = nullptr, *d_zidx = nullptr;
  void *d_t = nullptr;
  double *d_hx = nullptr, *d_hy = nullptr;  // scratch for hmv_apply_host
  HmvSolveWork *solve = nullptr;             // solver workspace (lazy)
  hmv_arena_stats stats{};
};

namespace hmv {
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);
// Enqueues y = H x + beta*y; `skip` (device int, may be null) makes every
// workgroup return early when *skip != 0.
hipError_t enqueue_apply(hmv_arena_s *a, const double *x, double *y, double beta, const int *skip,
                         hipStream_t s);
void free_solve_work(hmv_arena_s *a);
}  // namespace hmv

extern "C" hipError_t hmv_launch_apply(int acc64, const char *arena, const HmvSeg *seg1,
                                       const HmvWork *work1, int nwork1, const HmvSeg *seg2,
                                       const HmvWork *work2, int nwork2, const int *slotmap,
                                       const int *zidx, void *t, const double *x, double *y,
                                       double beta, const int *skip, hipStream_t stream);

#endif
