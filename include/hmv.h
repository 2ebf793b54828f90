/*
 * hmv.h -- MI355X H-matrix matvec engine, its Krylov solvers and device
 * construction (C ABI).
 *
 * Plain pointers and sizes only.  Device pointers are hipMalloc'd memory of
 * the current device; `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  Every function returns 0 on success and a negative code on
 * failure; hmv_last_error() gives the message of the calling thread's last
 * failure.  There is no CPU fallback: without a usable gfx950 device every
 * compute entry point fails with HMV_ERR_DEVICE.
 *
 * Vectors are in the cluster (permuted) order of the leaves, as in the
 * reference (m_HACApK_use.f90:285,469 permute with lod on the host side).
 *
 * Two groups of entry points:
 *   1. an explicit-array API (hmv_*), and
 *   2. reference-named entry points with the reference's argument lists, for
 *      a Fortran driver that links this library in place of the MAGMA build
 *      (declared at the end of this file).
 */
#ifndef HMV_H
#define HMV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HMV_OK 0
#define HMV_ERR_ARG (-1)
#define HMV_ERR_DEVICE (-2)
#define HMV_ERR_MEMORY (-3)
#define HMV_ERR_STATE (-4)

/* Storage precision of an arena. */
#define HMV_PREC_F64 0   /* factors and dense blocks FP64, FP64 accumulation       */
#define HMV_PREC_F32 1   /* factors and dense blocks FP32, FP32 accumulation       */
#define HMV_PREC_BF16 2  /* low-rank factors BF16, dense blocks FP32, FP32 accum.  */
#define HMV_PREC_BF16_ALL 3  /* low-rank factors and dense blocks BF16, FP32 accum. */

/* Leaf record, identical to stc_HACApK_leafmtx (HACApK_MAGMA.h:12-20) and to
 * the Fortran st_HACApK_leafmtx as seen from C.  ltmtx 1 = low rank
 * (a1 = V: ndt x kt, a2 = U: ndl x kt, column-major, block = U V^T),
 * 2 = dense (a1 = ndt x ndl column-major, i.e. the block row-major).
 * nstrtl/nstrtt are 1-based.  The engine reads a2 at (char*)a1 + a1size,
 * like HACApK_MAGMA.c:27. */
typedef struct {
  int ltmtx;
  int kt;
  int nstrtl, ndl;
  int nstrtt, ndt;
  long long a1size;
  double *a1;
  double *a2;
} hmv_leaf;

typedef struct hmv_arena_s *hmv_arena;

typedef struct {
  int nd;                 /* unknowns                                          */
  int precision;          /* HMV_PREC_*                                        */
  int64_t nlf, nlr, ndense;
  int64_t entries_lr;     /* sum_LR kt*(ndl+ndt)                               */
  int64_t entries_dense;  /* sum_D ndl*ndt                                     */
  int64_t tslots;         /* sum_LR kt (rank vectors)                          */
  int64_t tparts;         /* t buffer entries: slots + column-chunk partials   */
  int64_t phase1_bytes;   /* arena bytes read by the V^T x phase                */
  int64_t phase2_bytes;   /* arena bytes read by the [A|U] phase                */
  int64_t arena_bytes;    /* device bytes of factors incl. alignment padding   */
  int64_t payload_bytes;  /* bytes of the stored entries                        */
  int64_t alg_bytes;      /* payload + 16*nd (x read once, y written once)      */
  int64_t col_groups, row_groups, atoms;
  int64_t nwork_phase1, nwork_phase2;
  int ktmax;
  int pad;
  double build_seconds;   /* host repack + upload                               */
} hmv_arena_stats;

/* Selects the device for the calling thread (hipSetDevice). */
int hmv_init(int device);

/* Repacks nlf leaves (records spaced leaf_stride bytes apart, stc_HACApK_leafmtx
 * layout) of an nd x nd H-matrix into a device-resident arena. */
int hmv_arena_build(int nd, int nlf, const void *leaves, int leaf_stride, int precision,
                    hmv_arena *out);

/* Repack tunables for hmv_arena_build_tuned (used by the autotuner); 0 in a
 * field = the precision's default (96 KiB, 256, 256, and per precision).
 *   p1_target_bytes   arena bytes per phase-1 workgroup (rank-vector rows x columns)
 *   p1_row_tile       most rank vectors per phase-1 workgroup
 *   p2_max_rows       most rows per phase-2 workgroup
 *   p1/p2_small_threads  64 or 128: blocks needing at most that many lanes per
 *                     column run in workgroups of that many threads (more
 *                     workgroups in flight for short blocks); -1: all blocks
 *                     run in 256-thread workgroups. */
typedef struct {
  int64_t p1_target_bytes;
  int p1_row_tile;
  int p2_max_rows;
  int p1_small_threads;
  int p2_small_threads;
} hmv_tuning;
/* Defaults of a precision (what hmv_arena_build uses). */
int hmv_default_tuning(int precision, hmv_tuning *t);
int hmv_arena_build_tuned(int nd, int nlf, const void *leaves, int leaf_stride, int precision,
                          const hmv_tuning *tuning, hmv_arena *out);
/* The tuning an arena was built with (defaults resolved). */
int hmv_arena_tuning_get(hmv_arena a, hmv_tuning *t);
/* Autotuner: searches the tunables one at a time (small-workgroup threads
 * of each phase, phase-1 bytes per workgroup, phase-2 rows), building an
 * arena for each candidate on the current device and timing `reps` applies
 * (HIP events).  *best receives the fastest tuning, *best_ms its time per
 * apply, *default_ms the defaults' time.  Needs device memory for one arena
 * at a time. */
int hmv_autotune(int nd, int nlf, const void *leaves, int leaf_stride, int precision, int reps,
                 hmv_tuning *best, double *best_ms, double *default_ms);

/* Path of the H-matmat (hmv_apply_block) for k right-hand sides, per
 * arena: vector = 1 streams k <= 8 (FP64 arenas) or k <= 4 (FP32 / BF16) on
 * the vector units, else MFMA; tiles_per_unit = 16-row tiles one MFMA wave
 * takes per work unit (1, 2 or 4; the choices for ceil(k/16) = 1 are 2 and 4,
 * else 1 and 2; 0 = the default).  A setting holds for every k of the same
 * class: the vector choice per k <= 1, 2, 4, 8, the tiles per unit per
 * ceil(k/16) = 1, 2, 3-4, 5-8.  Results of the paths agree to rounding.  No
 * reference counterpart. */
typedef struct {
  int vector;
  int tiles_per_unit;
} hmv_block_tuning;
int hmv_block_tuning_get(hmv_arena a, int k, hmv_block_tuning *t);
int hmv_block_tuning_set(hmv_arena a, int k, const hmv_block_tuning *t);
/* Autotuner of the H-matmat path for k right-hand sides: times the
 * default path and every other choice of k's classes (`reps` products each
 * after one untimed, HIP events, inputs uniform in [-1, 1)), sets the fastest on the
 * arena and returns it; *best_ms and *default_ms are the times per product.
 * Needs device memory for 2 x nd x k doubles. */
int hmv_autotune_block(hmv_arena a, int k, int reps, hmv_block_tuning *best, double *best_ms,
                       double *default_ms);

/* Many independent instances in one arena: instance i (nd[i] unknowns,
 * nlf[i] leaves at leaves[i], all with the same record stride) occupies rows
 * and columns [off_i, off_i + nd[i]) of one block-diagonal H-matrix, off_i =
 * nd[0] + ... + nd[i-1].  Every apply covers all instances in one launch per
 * phase; vectors are the instances' vectors concatenated.  tuning may be
 * NULL.  No reference counterpart (the reference holds one H-matrix). */
int hmv_arena_build_multi(int ninst, const int *nd, const int *nlf, const void *const *leaves,
                          int leaf_stride, int precision, const hmv_tuning *tuning,
                          hmv_arena *out);
/* Autotuner over a multi-instance arena: as hmv_autotune, each candidate
 * built as hmv_arena_build_multi builds it (one launch per phase over every
 * instance) and timed on the whole block-diagonal apply.  No reference
 * counterpart. */
int hmv_autotune_multi(int ninst, const int *nd, const int *nlf, const void *const *leaves,
                       int leaf_stride, int precision, int reps, hmv_tuning *best, double *best_ms,
                       double *default_ms);
/* Number of instances of an arena (1 unless built by hmv_arena_build_multi);
 * offsets (may be NULL) receives up to cap of the ninst+1 row offsets. */
int hmv_arena_segments(hmv_arena a, int *offsets, int cap);

int hmv_arena_stats_get(hmv_arena a, hmv_arena_stats *s);
int hmv_arena_free(hmv_arena a);

/* y = H*x + beta*y on device vectors (beta 0: y is overwritten, NaNs in y
 * ignored).  Asynchronous on `stream`. */
int hmv_apply(hmv_arena a, const double *x, double *y, double beta, void *stream);

/* H-matmat: Y = H X + beta*Y for k right-hand sides (1 <= k <= HMV_MAX_RHS)
 * on device arrays stored row-interleaved (N x k row-major: row i's k values
 * contiguous).  Few right-hand sides run on the vector units, more on MFMA
 * (FP64, or FP32 for FP32/BF16 arenas; hmv_block_tuning); each column
 * equals hmv_apply of that column to rounding.  x needs 8-byte alignment;
 * a 16-byte aligned x (as hipMalloc returns) lets the kernels read inputs
 * in 16-byte pairs, with the same result.  No reference counterpart (the
 * reference applies one vector per call). */
#define HMV_MAX_RHS 128
int hmv_apply_block(hmv_arena a, int k, const double *x, double *y, double beta, void *stream);
/* H-matmat that allocates nothing: as hmv_apply_block (same kernels and
 * result), with the product's t buffer in a static device workspace of
 * HMV_STATIC_SCRATCH_BYTES per device instead of memory hmv_apply_block
 * allocates on its first call with a larger k.  No reference counterpart:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops).  Needs max(1, tparts) * k * 8 bytes of t buffer
 * (tparts from hmv_arena_stats_get); a larger product fails with
 * HMV_ERR_ARG.  The arena's device must be current.  Each device holds its
 * workspace until the process ends; every call on a device shares it, so
 * calls that overlap on one device (other streams or threads) race. */
#define HMV_STATIC_SCRATCH_BYTES ((size_t)256 << 20)
int hmv_apply_block_static(hmv_arena a, int k, const double *x, double *y, double beta,
                           void *stream);
/* Average time of `reps` H-matmats (HIP events on an internal stream). */
int hmv_apply_block_timed(hmv_arena a, int k, const double *x, double *y, int reps, double *ms);

/* zau += H*zu on host vectors (the accumulate semantics of
 * c_hacapk_adot_body_lfmtx_batch_, HACApK_BATCH.c:257-271); synchronous.
 * zu goes up through pinned staging and the apply writes H*zu into pinned
 * host memory, both held by the arena from the first call on; calls on one
 * arena from several threads at once race on that staging. */
int hmv_apply_host(hmv_arena a, const double *zu, double *zau);

/* Bytes of the device t-partial buffer owned by the arena. */
int64_t hmv_arena_tbytes(hmv_arena a);

/* ---- BiCGSTAB ---------------------------------------------------------- */

#define HMV_SOLVE_CONVERGED 0
#define HMV_SOLVE_MAXITER 1
#define HMV_SOLVE_BREAKDOWN 2
#define HMV_SOLVE_NONFINITE 3

typedef struct {
  int nstp;            /* iterations taken (nstp of the reference)          */
  int status;          /* HMV_SOLVE_*                                       */
  double relres;       /* recursive ||r||/||b|| at exit                     */
  double true_relres;  /* ||b - A u||/||b|| recomputed after the solve      */
  double seconds;      /* wall time of the solve (device-synchronised)      */
  int outer;           /* refinement steps (mixed-precision solve), else 0  */
  int inner_total;     /* inner iterations summed over refinement steps     */
} hmv_solve_info;

/* Unpreconditioned BiCGSTAB (m_HACApK_solve.f90:342-449) on device vectors:
 * u (initial guess in, solution out), b.  Stops when ||r||/||b|| < eps or
 * after maxit steps.  hist (host, may be NULL) receives maxit+1 relative
 * residuals (hist[0] initial).  check_every: residual read-back period in
 * iterations (1 = every iteration, like the reference). */
int hmv_bicgstab(hmv_arena a, const double *b, double *u, double eps, int maxit, int check_every,
                 double *hist, hmv_solve_info *info);

/* Mixed-precision iterative refinement: the FP64 residual r = b - A64 u, an
 * inner BiCGSTAB with the low-precision arena `alow` to inner_eps, u += d;
 * repeats until ||b - A64 u||/||b|| < eps or max_outer steps.  check_every:
 * residual read-back period of the inner solver.
 * hist (host, may be NULL) receives max_outer+1 outer relative residuals. */
int hmv_bicgstab_ir(hmv_arena a64, hmv_arena alow, const double *b, double *u, double eps,
                    double inner_eps, int max_outer, int max_inner, int check_every, double *hist,
                    hmv_solve_info *info);

/* Mixed-precision iterative refinement with GCR(m) inner solves: as
 * hmv_bicgstab_ir, with each correction solved by GCR(m) (hmv_gcrm) on the
 * low-precision arena.  A short inner solve is mostly GCR's first cycle,
 * which minimises the residual over the Krylov space: fewer low-precision
 * products per decade than BiCGSTAB's two per step.  No reference
 * counterpart. */
int hmv_gcr_ir(hmv_arena a64, hmv_arena alow, const double *b, double *u, double eps,
               double inner_eps, int m, int max_outer, int max_inner, int check_every,
               double *hist, hmv_solve_info *info);

/* Mixed-precision BiCGSTAB with reliable updates: one BiCGSTAB whose
 * products use the low-precision arena `alow`, with u, vectors and scalars
 * in FP64; whenever the iterated residual has fallen by `delta` (0..1) since
 * the last replacement, or the iteration reports convergence, the true
 * residual b - A64 u (FP64 arena) replaces it without restarting.  Stops on
 * ||b - A64 u|| / ||b|| < eps.  info->outer = number of replacements. */
int hmv_bicgstab_ru(hmv_arena a64, hmv_arena alow, const double *b, double *u, double eps,
                    int maxit, int check_every, double delta, hmv_solve_info *info);

/* Pipelined BiCGSTAB (c_hacapk_bicgstab_cax_lfmtx_pipe_,
 * HACApK_SOLVER_PIPE.c:144-637) on device vectors u (initial guess in,
 * solution out), b: the same recurrences, with the iterate started at u and
 * returned (the reference never initialises nor returns it).  Arguments and
 * stopping as for hmv_bicgstab. */
int hmv_bicgstab_pipe(hmv_arena a, const double *b, double *u, double eps, int maxit,
                      int check_every, double *hist, hmv_solve_info *info);

/* GCR(m) (HACApK_gcrm_lfmtx, m_HACApK_solve.f90:506-571) on device vectors
 * u (initial guess in, solution out), b: m search directions kept, restart
 * every m steps with the reference's own restart step (the oldest slot is
 * overwritten by the new direction before the projections, so its
 * coefficient scales that direction).  The reference reads m from
 * param(87), default 8.  Stops when ||r||/||b|| < eps or after maxit steps;
 * hist (host, may be NULL) receives maxit+1 relative residuals (NaN after
 * the last step); check_every as for hmv_bicgstab. */
#define HMV_GCR_MAX_M 32
int hmv_gcrm(hmv_arena a, const double *b, double *u, double eps, int maxit, int m,
             int check_every, double *hist, hmv_solve_info *info);

/* Block BiCGSTAB: k independent solves A u_j = b_j (k a power of two,
 * 1..128) sharing one H-matmat per step; b, u device arrays N x k
 * row-interleaved (u holds the initial guesses).  Column j stops at
 * ||r_j||/||b_j|| < eps and is frozen while the others continue.  cols (k
 * entries, may be NULL) receives per-column step counts, status and the
 * recursive and true relative residuals; info the iterations run, the worst
 * status, the largest true residual and the wall time.  No reference
 * counterpart (the reference solves one right-hand side per call). */
typedef struct {
  int nstp;
  int status;
  double relres;
  double true_relres;
} hmv_column_info;
int hmv_block_bicgstab(hmv_arena a, int k, const double *b, double *u, double eps, int maxit,
                       int check_every, hmv_column_info *cols, hmv_solve_info *info);

/* Mixed-precision refinement of a block solve: a64 and alow are FP64 and
 * low-precision arenas of one H-matrix.  Each outer step computes every
 * running column's FP64 residual R = B - A64 U with an FP64 H-matmat; a
 * column with ||r_j||/||b_j|| < eps stops (converged); the others solve
 * A_low D = R by block BiCGSTAB (hmv_block_bicgstab's iteration, each
 * column to inner_eps, at most max_inner iterations) and take U += D.  At
 * most max_outer steps.  cols receives per column the inner iterations
 * summed over the outer steps, the status and the final FP64 relative
 * residual; info the outer steps, the inner iterations summed, the worst
 * status and residual and the wall time.  No reference counterpart. */
int hmv_block_bicgstab_ir(hmv_arena a64, hmv_arena alow, int k, const double *b, double *u,
                          double eps, double inner_eps, int max_outer, int max_inner,
                          int check_every, hmv_column_info *cols, hmv_solve_info *info);

/* One BiCGSTAB per instance of a multi-instance arena (hmv_arena_build_multi),
 * all advancing together: one H-matvec launch per phase covers every
 * instance, scalars and stopping are per instance (a finished instance is
 * frozen), one hipGraph per iteration.  b, u: the instances' vectors
 * concatenated.  inst (ninst entries, may be NULL) receives per-instance
 * steps, status and residuals; inst_seconds (may be NULL) the wall time at
 * which each instance was seen finished (residual read-back every
 * check_every iterations).  No reference counterpart. */
int hmv_multi_bicgstab(hmv_arena a, const double *b, double *u, double eps, int maxit,
                       int check_every, hmv_column_info *inst, double *inst_seconds,
                       hmv_solve_info *info);

/* Mixed-precision refinement per instance: a64 and alow are FP64 and
 * low-precision arenas of the same instances (hmv_arena_build_multi with
 * the same leaves).  Each outer step computes every running instance's FP64
 * residual r_i = b_i - A64_i u_i on a64; an instance with
 * ||r_i||/||b_i|| < eps stops (converged); the others solve A_low d = r by
 * one BiCGSTAB per instance in lockstep (hmv_multi_bicgstab's iteration,
 * one hipGraph per iteration, each instance to inner_eps, at most
 * max_inner iterations; stopped instances skip their products) and take
 * u += d.  At most max_outer steps.  inst receives per instance the inner
 * iterations summed over the outer steps, the status and the final FP64
 * relative residual; inst_seconds the time at which each instance was seen
 * converged; info the outer steps, the inner iterations summed over the
 * outer steps, the worst status and residual and the wall time.  No
 * reference counterpart. */
int hmv_multi_bicgstab_ir(hmv_arena a64, hmv_arena alow, const double *b, double *u, double eps,
                          double inner_eps, int max_outer, int max_inner, int check_every,
                          hmv_column_info *inst, double *inst_seconds, hmv_solve_info *info);

/* The same refinement per instance with GCR(m) inner solves (1 <= m <=
 * HMV_GCR_MAX_M): one GCR(m) per running instance in lockstep, the steps of
 * hmv_gcrm (m_HACApK_solve.f90:506-571) with each instance's own alpha,
 * betas and projections, one hipGraph per step, each instance to
 * inner_eps, at most max_inner steps.  inst's nstp counts the inner steps
 * (one product each).  No reference counterpart. */
int hmv_multi_gcr_ir(hmv_arena a64, hmv_arena alow, const double *b, double *u, double eps,
                     double inner_eps, int m, int max_outer, int max_inner, int check_every,
                     hmv_column_info *inst, double *inst_seconds, hmv_solve_info *info);

const char *hmv_last_error(void);

/* ---- device plumbing (buffers, copies, timing) ------------------------- */
int hmv_device_malloc(void **p, size_t bytes);
int hmv_device_free(void *p);
int hmv_memcpy_htod(void *dst, const void *src, size_t bytes);
int hmv_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int hmv_device_synchronize(void);
/* Device name and compute units of the current device. */
int hmv_device_info(char *name, int name_len, int *compute_units);
/* Runs `reps` applies y = H x on an internal stream and returns the average
 * duration of each phase launch and of the whole apply, measured with HIP
 * events recorded on that stream around every launch (milliseconds).
 * ms_phase1 is 0 when the arena has no low-rank leaves. */
int hmv_apply_timed(hmv_arena a, const double *x, double *y, int reps, double *ms_phase1,
                    double *ms_phase2, double *ms_apply);

/* Bandwidth calibration: `reps` launches of a plain streaming kernel over
 * `bytes` bytes (kind 0: read-only sum of src; kind 1: copy src -> dst);
 * *ms receives the average launch time (HIP events).  Kind 2 calibrates
 * the FP64 MFMA rate instead: `bytes` (a multiple of 8, at most 2^27) FP64
 * 16x16x4 MFMAs per wave on register operands, in HMV_MFMA_RATE_WAVES = 4096
 * waves (4 per SIMD); src is unused, dst receives nothing (8 bytes).
 * Flops per launch: 2048 * bytes * 4096. */
int hmv_stream_test(int kind, const void *src, void *dst, size_t bytes, int reps, double *ms);

/* ---- host-side plan inspection (no device needed) ----------------------
 * The repack that hmv_arena_build uploads, exposed for inspection and for
 * CPU-side validation of the arena's indexing.  Tables use the engine's
 * internal record layouts: work = {int64 off0, off1; int rows, rows_p,
 * ncol0, ncol1; int16 dt0, dt1; int zseg_begin, zseg_end, out_base}
 * (48 bytes); zseg = {int kind, ncols, src, col} (16 bytes): a run of a
 * block's input columns starting at block column col, from x[src..] (kind
 * 0) or t[src..] (kind 2), or (kind 1, alone) from codes[src..], one int per
 * column (>= 0: x index; < 0: t slot -1-code), used past 64 runs; tfin =
 * {int slot, base, np, pad} (16 bytes): t[slot] = t[base] + ... +
 * t[base+np-1] for rank vectors whose column group is cut into np chunks.
 * The t buffer holds tsize slots followed by those partials (tpart_size
 * entries in all). */
typedef struct hmv_plan_s *hmv_plan;
typedef struct {
  const void *arena;   int64_t arena_bytes;
  const void *work1;   int64_t nwork1;
  const void *work2;   int64_t nwork2;
  const void *zseg;    int64_t nzseg;
  const int *outidx;   int64_t noutidx;
  const int *codes;    int64_t ncodes;
  const void *tfin;    int64_t ntfin;
  int64_t tsize, tpart_size;
  int work_bytes, zseg_bytes, tfin_bytes;
  int nwork1_small, nt1_small;  /* short blocks first in work1, run with nt1_small threads */
  int nwork2_small, nt2_small;
} hmv_plan_tables;
int hmv_plan_create(int nd, int nlf, const void *leaves, int leaf_stride, int precision,
                    hmv_plan *out);
int hmv_plan_create_tuned(int nd, int nlf, const void *leaves, int leaf_stride, int precision,
                          const hmv_tuning *tuning, hmv_plan *out);
int hmv_plan_view(hmv_plan p, hmv_plan_tables *t);
int hmv_plan_free(hmv_plan p);

/* ---- device construction -----------------------------------------------
 * The fill of HACApK_generate on the device (m_HACApK_base.f90:1005-1051):
 * every dense leaf's entries (HACApK_entry_ij) and every low-rank leaf's
 * ACA+ factors (HACApK_acaplus, :783-975, param(61) = 1), one workgroup per
 * leaf.  The block structure comes from the host (hmvgen_build_blocks: tree,
 * admissibility, leaf list; specs are its leaf records, kt and factors
 * ignored).  Inputs are host arrays: the mesh (nodes nnode x 3, faces
 * nface x 3, 0-based), centroids and areas per face (original order), lod
 * (cluster order -> 1-based original index).  kernel: 0 = the ppohBEM
 * analytic integral for every pair, 1 = centroid 1/(4 pi r) off the diagonal
 * (as hmvgen's HMVGEN_KERNEL_*).  The factors stay in device memory. */
typedef struct {
  int kernel;
  int kmax;              /* param(63) */
  int kmin;              /* param(64) */
  int pad;
  double aca_eps;        /* param(71) */
  double aca_eps_factor; /* param(72) */
} hmv_construct_params;
typedef struct hmv_dleaves_s *hmv_dleaves;
int hmv_construct(int nnode, const double *nodes, int nface, const int *faces, const double *cent,
                  const double *area, const int *lod, int nlf, const void *specs, int stride,
                  const hmv_construct_params *p, hmv_dleaves *out);
/* nlf, the ranks per leaf (nlf entries; 0 for dense leaves; may be NULL) and
 * the device fill's wall time in seconds. */
int hmv_dleaves_info(hmv_dleaves d, int *nlf, int *kt, double *seconds);
/* One leaf's factors to host memory in the reference's layout: dense
 * ndt x ndl (a1), low rank V (ndt x kt) followed by U (ndl x kt). */
int hmv_dleaves_download(hmv_dleaves d, int leaf, double *a1);
/* Arena straight from the device factors: the repack runs on the host over
 * the leaf shapes and ranks, the copies into the arena image on the device. */
int hmv_arena_build_dleaves(hmv_dleaves d, int precision, const hmv_tuning *tuning,
                            hmv_arena *out);
int hmv_dleaves_free(hmv_dleaves d);

/* ---- arena files -------------------------------------------------------
 * A versioned binary file of the repack (arena image and tables, format
 * version HMV_ARENA_FILE_VERSION, checksummed sections), so that later runs
 * upload it without the leaves.  No reference counterpart: the reference
 * rebuilds the H-matrix and its device copy every run (param(12) "write
 * down the H-matrix", m_HACApK_base.f90:300, is declared and unused).
 * Loading checks the checksums and every index of the tables against the
 * buffers it addresses; a file that fails is rejected with HMV_ERR_ARG. */
#define HMV_ARENA_FILE_VERSION 1
int hmv_plan_save(hmv_plan p, const char *path);        /* host only */
int hmv_plan_load(const char *path, hmv_plan *out);     /* host only */
int hmv_arena_from_plan(hmv_plan p, hmv_arena *out);    /* uploads a plan */
int hmv_arena_save(hmv_arena a, const char *path);      /* copies the device tables back */
int hmv_arena_load(const char *path, hmv_arena *out);   /* hmv_plan_load + hmv_arena_from_plan */

/* ---- reference-named entry points --------------------------------------
 * The C view of the caller's st_HACApK_leafmtxp.  Every reference call
 * site of these entry points but one is compiled only under HAVE_MAGMA or
 * HAVE_MAGMA_BATCH (m_HACApK_use.f90:317-455, m_HACApK_solve.f90:97-103,
 * 372-435, 652-741, 850-966), and those flags put the MAGMA block between
 * st_lf_stride and st_lf (m_HACApK_base.f90:64-235, C mirror
 * HACApK_MAGMA.h:28-106).  hmv_ref_leafmtxp is that layout (544 bytes,
 * st_lf at byte 536, magma_int_t = int), the default.  A driver built
 * without either flag reaches only c_hacapk_adot_body_lfmtx_
 * (m_HACApK_solve.f90:102) and passes hmv_ref_leafmtxp_plain (st_lf at byte
 * 24); select it with hmv_ref_set_layout(HMV_REF_LAYOUT_PLAIN) or the
 * environment variable HMV_REF_LAYOUT=plain, read at the first call.
 *
 * The engine reads nd, nlf, st_lf_stride and st_lf (and each leaf through
 * st_lf_stride and a1size, as HACApK_MAGMA.c:27 does), writes mpi_rank = 0
 * when it builds an arena (the reference's lfcpy stores the MPI rank there,
 * HACApK_BATCH_SORT.c:98, and m_HACApK_solve.f90:739 reads it), and leaves
 * every other MAGMA field untouched.  Its device state lives in a side table
 * keyed by the struct's address. */
typedef struct {
  int nd;
  int nlf;
  int nlfkt;
  int ktmax;
  int st_lf_stride;
  /* MAGMA block, HACApK_MAGMA.h:34-102 (byte offsets in comments) */
  int m, n, gn, max_block;        /* 20 .. 32                                   */
  double gflops;                  /* 40                                         */
  void *gpu_vectors[6];           /* 48: mtx1_gpu mtx2_gpu zu_gpu zau_gpu zau_pin zbu_gpu */
  int num_batch, total_size_y, transA; /* 96 .. 104                             */
  void *batch_tables[20];         /* 112: d_A_array .. d_inc, batch_order .. batch_size */
  int num_streamed, num_streamed_t;    /* 272, 276                              */
  void *streamed_tables[6];       /* 280: h_A_array_streamed .. h_lda_streamed  */
  void *mgpu_tables[25];          /* 328: d_A_mgpu .. total_size_y_mgpu, iwork  */
  int mpi_rank;                   /* 528                                        */
  hmv_leaf *st_lf;                /* 536: first word of the Fortran descriptor  */
} hmv_ref_leafmtxp;

/* st_HACApK_leafmtxp of a build without HAVE_MAGMA / HAVE_MAGMA_BATCH
 * (m_HACApK_base.f90:64-76 and :234). */
typedef struct {
  int nd;
  int nlf;
  int nlfkt;
  int ktmax;
  int st_lf_stride;
  hmv_leaf *st_lf;
} hmv_ref_leafmtxp_plain;

#define HMV_REF_LAYOUT_MAGMA 0  /* hmv_ref_leafmtxp (default)   */
#define HMV_REF_LAYOUT_PLAIN 1  /* hmv_ref_leafmtxp_plain       */
/* Selects the struct layout the reference-named entry points read; returns
 * the previous layout, or HMV_ERR_ARG for an unknown value.  No reference
 * counterpart (the reference fixes it at compile time). */
int hmv_ref_set_layout(int layout);
int hmv_ref_get_layout(void);

typedef struct {
  int lf_umpi;
  int lpmd_offset;
  int lod_offset;
  int lsp_offset;
  int lnp_offset;
  int lthr_offset;
  double *param;
} hmv_ref_lcontrol;

/* Replace magma_init_ / magma_finalize_ (HACApK_MAGMA.c:69-75), which a
 * HAVE_MAGMA driver calls once around its run (m_HACApK_use.f90:61,74):
 * magma_init_ selects device 0 of the process (hmv_init) and prints the
 * error if there is none; magma_finalize_ does nothing (an arena bound by
 * lfcpy lives until lfdel). */
void magma_init_(void);
void magma_finalize_(void);

/* Replaces c_hacapk_adot_body_lfcpy_batch_sorted_ (HACApK_BATCH_SORT.c:86). */
void c_hacapk_adot_body_lfcpy_batch_sorted_(int *nd, hmv_ref_leafmtxp *st_leafmtxp);
/* Replaces c_hacapk_adot_body_lfmtx_batch_ (HACApK_BATCH.c:257): zau += H*zu
 * on host vectors; zbu is unused; the timers are incremented (seconds). */
void c_hacapk_adot_body_lfmtx_batch_(double *zau, hmv_ref_leafmtxp *st_leafmtxp, double *zu,
                                     double *zbu, double *time_batch, double *time_set,
                                     double *time_copy);
/* Replaces c_hacapk_adot_body_lfdel_batch_ (HACApK_BATCH.c:812). */
void c_hacapk_adot_body_lfdel_batch_(hmv_ref_leafmtxp *st_leafmtxp);
/* Replace the reference's host applies c_hacapk_adot_body_lfmtx_
 * (HACApK_MAGMA.c:5-57, scalar loops) and c_hacapk_adot_body_lfmtx_gpu_
 * (HACApK_MAGMA.c:79-227, per-leaf dgemv on the host despite its name),
 * which the Fortran driver calls in its non-batch builds
 * (m_HACApK_solve.f90:100,102) and in its timing loop (:664): zau += H*zu on
 * host vectors, zbu unused.  The first call builds an FP64 arena and binds
 * it to st_leafmtxp (reused by later calls and by the solvers; freed by
 * lfdel_batch_, which a caller must call after changing the leaves). */
void c_hacapk_adot_body_lfmtx_(double *zau, hmv_ref_leafmtxp *st_leafmtxp, double *zu,
                               double *zbu);
void c_hacapk_adot_body_lfmtx_gpu_(double *zau, hmv_ref_leafmtxp *st_leafmtxp, double *zu,
                                   double *zbu);
/* Replace c_hacapk_adot_body_lfcpy_gpu_ / lfdel_gpu_ (HACApK_MAGMA.c:231,369),
 * the per-leaf device copies of the non-batch MAGMA build
 * (m_HACApK_solve.f90:379,434,681-683): the same as lfcpy_batch_sorted_ and
 * lfdel_batch_. */
void c_hacapk_adot_body_lfcpy_gpu_(int *nd, hmv_ref_leafmtxp *st_leafmtxp);
void c_hacapk_adot_body_lfdel_gpu_(hmv_ref_leafmtxp *st_leafmtxp);
/* Replaces c_hacapk_bicgstab_cax_lfmtx_gpu_ (HACApK_SOLVER.c:421): host u, b;
 * param[82] = max steps, param[90] = tolerance (Fortran param(83), param(91)). */
void c_hacapk_bicgstab_cax_lfmtx_gpu_(hmv_ref_leafmtxp *st_leafmtxp, hmv_ref_lcontrol *st_ctl,
                                      double *u_cpu, double *b_cpu, double *param, int *nd,
                                      int *nstp, int *lrtrn);
/* Replaces c_hacapk_bicgstab_cax_lfmtx_pipe_ (HACApK_SOLVER_PIPE.c:144): the
 * pipelined BiCGSTAB (hmv_bicgstab_pipe) with the same host arguments; u
 * receives the solution. */
void c_hacapk_bicgstab_cax_lfmtx_pipe_(hmv_ref_leafmtxp *st_leafmtxp, hmv_ref_lcontrol *st_ctl,
                                       double *u_cpu, double *b_cpu, double *param, int *nd,
                                       int *nstp, int *lrtrn);
/* Replaces c_hacapk_bicgstab_cax_lfmtx_pipe2_ (HACApK_SOLVER_PIPE.c:640), the
 * variant the PIPE_BICG driver calls (m_HACApK_use.f90:418): the same
 * recurrences with the MPI all-gather overlapped with the dot products,
 * which one GPU does not have; it runs hmv_bicgstab_pipe. */
void c_hacapk_bicgstab_cax_lfmtx_pipe2_(hmv_ref_leafmtxp *st_leafmtxp, hmv_ref_lcontrol *st_ctl,
                                        double *u_cpu, double *b_cpu, double *param, int *nd,
                                        int *nstp, int *lrtrn);
/* Replaces c_hacapk_bicgstab_cax_lfmtx_flat_ (HACApK_SOLVER.c:254), the
 * BiCGSTAB on host vectors with one GPU matvec per product
 * (m_HACApK_use.f90:400): the device BiCGSTAB, same arguments as _gpu_. */
void c_hacapk_bicgstab_cax_lfmtx_flat_(hmv_ref_leafmtxp *st_leafmtxp, hmv_ref_lcontrol *st_ctl,
                                       double *u_cpu, double *b_cpu, double *param, int *nd,
                                       int *nstp, int *lrtrn);
/* Replaces c_hacapk_bicgstab_cax_lfmtx_mgpu1_ (HACApK_SOLVER.c:815,
 * m_HACApK_use.f90:349) for one rank (lpmd(2) = 1): the device BiCGSTAB.
 * With more ranks it prints an error and sets *lrtrn = HMV_ERR_ARG. */
void c_hacapk_bicgstab_cax_lfmtx_mgpu1_(hmv_ref_leafmtxp *st_leafmtxp, hmv_ref_lcontrol *st_ctl,
                                        double *u_cpu, double *b_cpu, double *param, int *nd,
                                        int *nstp, int *lrtrn);

#ifdef __cplusplus
}
#endif
#endif /* HMV_H */
