// hmv_internal.h -- shared definitions of the H-matvec engine (host + device).
//
// Arena layout (one per precision):
//
//   Phase 1 (t = V^T x):  low-rank leaves are grouped by column cluster
//     (nstrtt, ndt).  The V factors of one group are concatenated column by
//     column, i.e. the group's stacked V^T is one row-major block whose rows
//     all multiply the same x segment.  Each V^T row produces one t entry.
//
//   Phase 2 (y = [A | U] [x; t]):  leaves are grouped by row cluster
//     (nstrtl, ndl).  A row group's panel is row-major with the dense blocks'
//     rows first (inputs gathered from x) followed by the U factors' rows
//     (inputs: the group's contiguous t segment).  Every y row is produced by
//     exactly one workgroup, which sums the panel rows of all row clusters
//     that contain it, so y needs no atomics and the summation order is fixed.
//
// Both phases run the same streaming kernel: a workgroup walks a list of
// row-major segments, stages the segment's input vector in LDS, computes the
// row dot products with 16-byte loads and team (lanes-per-row) shuffles,
// accumulates per-row sums in LDS and writes its outputs once.
#ifndef HMV_INTERNAL_H
#define HMV_INTERNAL_H

#include <stdint.h>

#define HMV_DT_F64 0
#define HMV_DT_F32 1
#define HMV_DT_BF16 2

#define HMV_WG_THREADS 256
#define HMV_MAX_ROWS 256      // rows accumulated by one workgroup
#define HMV_ZCHUNK 2048       // inputs staged in LDS per pass
#define HMV_ARENA_ALIGN 256   // byte alignment of every block in the arena

// One row-major block processed by a workgroup.
struct HmvSeg {
  long long data_off;  // byte offset of row 0, column 0 in the arena
  int ld;              // row stride in elements (multiple of the 16-byte vector)
  int ncols;           // columns actually used
  int nrows;           // rows in this segment
  int acc_row;         // first row of the workgroup's accumulator it adds to
  int nx;              // columns [0,nx) read x; columns [nx,ncols) read t
  int xsrc;            // phase 1: x offset of column 0; phase 2: offset into zidx
  int tsrc;            // t index of column nx
  short dtype;         // HMV_DT_*
  short lpr;           // lanes per row: 4,8,16,32,64 or 256
  int pad0, pad1;
};

// One workgroup's work.
struct HmvWork {
  int seg_begin, seg_end;  // range in the segment table
  int out_base;            // phase 1: offset into slotmap; phase 2: first y row
  int nout;                // outputs written (<= HMV_MAX_ROWS)
};

#endif
