// hmv_internal.h -- shared definitions of the H-matvec engine (host + device).
//
// Arena layout (one arena per storage precision).  Every workgroup of either
// phase owns ONE contiguous, column-major block of `rows_p` rows (rows padded
// to the row-vector width) by `ncol0 + ncol1` columns: column c holds the
// rows' coefficients for input z[c].  Columns [0,ncol0) are stored with
// dtype dt0, columns [ncol0, ncol0+ncol1) with dt1 (the second part exists
// only for BF16 arenas, where dense blocks stay FP32).
//
//   Phase 1 (t = V^T x): low-rank leaves are grouped by column cluster
//     (nstrtt, ndt); a group's rank vectors are the rows, its x segment the
//     inputs.  Long groups are cut into column chunks whose partial sums
//     land in separate t-partial slots.
//
//   Phase 2 (y = [A | U] [x; t]): rows are cut into atoms, maximal row
//     ranges covered by the same set of row clusters; a workgroup owns one
//     atom (at most 256 rows) and all panel columns of every row cluster
//     covering it: dense columns (inputs gathered from x) and U columns
//     (inputs: sums of the t partials).  Each y row has exactly one writer.
//     The input of every column is given by a code (x index or t slot), so
//     staging is one flat gather.
//
//   Between the phases, rank vectors whose column group was cut into chunks
//   get their t entry as the sum of the chunk partials in chunk order (one
//   small launch over the HmvTFin list); every other rank vector's phase-1
//   workgroup writes its t entry directly.
//
// A workgroup stages its inputs z in LDS (zsegs describe where each column's
// input comes from), then each lane accumulates `VR` fixed rows over a
// strided set of columns with 8- or 4-byte loads that are contiguous across
// the workgroup, and one fixed-order LDS reduction produces the outputs.
// Work lists are ordered so that the workgroups the dispatcher sends to one
// XCD (every 8th) cover one contiguous eighth of the rows (or columns), which
// keeps each XCD's gathers in its own L2.
#ifndef HMV_INTERNAL_H
#define HMV_INTERNAL_H

#include <stdint.h>

#define HMV_DT_F64 0
#define HMV_DT_F32 1
#define HMV_DT_BF16 2

#define HMV_WG_THREADS 256
#define HMV_MAX_ROWS 256      // rows owned by one workgroup
#define HMV_ZCHUNK 2048       // inputs staged in LDS per pass
#define HMV_ARENA_ALIGN 256   // byte alignment of every block in the arena

// Sources of staged inputs.  A work item's inputs are either runs (kinds X
// and T, at most HMV_RUN_MAX of them, in column order) or one CODES segment.
#define HMV_Z_X 0      // z[col + c] = x[src + c]
#define HMV_Z_CODES 1  // z[c] from code = codes[src + c]:
                       //   code >= 0: x[code];  code < 0: t[-code - 1]
#define HMV_Z_T 2      // z[col + c] = t[src + c]
#define HMV_RUN_MAX 64
#define HMV_MAX_TPARTS 32  // t partials per rank vector (column chunks of a group)

// One workgroup's block.
struct HmvWork {
  long long off0;   // byte offset of the dt0 part (column-major, ld = rows_p)
  long long off1;   // byte offset of the dt1 part (or -1)
  int rows;         // rows produced
  int rows_p;       // rows stored per column (multiple of the row-vector width)
  int ncol0;        // columns stored with dt0
  int ncol1;        // columns stored with dt1
  short dt0, dt1;   // HMV_DT_*
  int zseg_begin;   // input sources [zseg_begin, zseg_end), in column order
  int zseg_end;
  int out_base;     // phase 1: offset into outidx; phase 2: first y row
};

// t[slot] = t[base] + t[base + 1] + ... + t[base + np - 1]  (np >= 2; the
// partials live past the t slots, base >= number of slots).
struct HmvTFin {
  int slot;
  int base;
  int np;
  int pad;
};

// Everything one apply launches over (device pointers).  Each phase's work
// list holds its short blocks first (nwork_small of them, run with nt_small
// threads per workgroup), then the rest (256 threads).
struct HmvLaunch {
  int mode;  // 0 FP64, 1 FP32, 2 BF16 low rank + FP32 dense
  const char *arena;
  const struct HmvWork *work1;
  int nwork1, nwork1_small, nt1_small;
  const struct HmvTFin *tfin;
  int ntfin;
  const struct HmvWork *work2;
  int nwork2, nwork2_small, nt2_small;
  const struct HmvZSeg *zseg;
  const int *outidx;
  const int *codes;
  double *t;
  const int *winst1, *winst2;  // instance of each work item (multi-instance arenas; else null)
};
#define HMV_PHASE1 1  // phase 1 and the chunk-partial sums
#define HMV_PHASE2 2

// A run of consecutive input columns with one source.
struct HmvZSeg {
  int kind;   // HMV_Z_*
  int ncols;
  int src;
  int col;    // first column of the run within its work item
};

// One copy of the arena fill when the leaf factors live in device memory:
// dst part (column-major, ld rows) at byte offset `dst`, rows r0.. and
// columns c0.. of it, from src[r * s_r + c * s_c] (doubles), converted to dt.
struct HmvFillOp {
  long long dst;
  const double *src;
  long long s_r, s_c;
  int ld, r0, c0, nrows, ncols, dt;
};

#endif