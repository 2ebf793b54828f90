/*
 * hmvgen.h -- CPU input generator for the H-matrix engine (C ABI).
 *
 * Builds the leaf list that the engine consumes: mesh -> collocation points
 * -> bisection cluster tree -> block tree (eta-admissibility) -> ACA+ /
 * dense fill.  This is the input side of the H-matvec path; it restates the
 * reference's construction (m_HACApK_base.f90) on the host so that the
 * engine can be fed without the reference's Fortran/MPI build.
 *
 * Leaves come out in the reference's own record layout (stc_HACApK_leafmtx,
 * HACApK_MAGMA.h:12-20): a1 = V (ndt x kt, column-major) or the dense block
 * (ndt x ndl, column-major, i.e. row-major A), a2 = U (ndl x kt,
 * column-major) stored directly after a1 so that a2 == (char*)a1 + a1size
 * exactly as HACApK_MAGMA.c:27 expects.
 *
 * All functions return 0 on success and a negative code on failure; the
 * message of the last failure is available from hmvgen_last_error().
 */
#ifndef HMVGEN_H
#define HMVGEN_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel (matrix entry) modes. */
#define HMVGEN_KERNEL_PPOHBEM   0  /* face_integral for every pair, |.|/(4*pi*eps0) with the
                                      reference's single-precision PI/EPSILON_0 literals
                                      (m_ppohBEM_user_func.f90:42-100)                     */
#define HMVGEN_KERNEL_CENTROID  1  /* 1/(4*pi*|x-y|): area_j/(4*pi*|c_i-c_j|) off-diagonal,
                                      analytic flat-triangle self term on the diagonal      */

/* Construction parameters; field comments give the reference's param() slot
 * (defaults in m_HACApK_base.f90:294-313). */
typedef struct {
  int    leaf_size;        /* param(21): max points in a leaf cluster (15)            */
  double max_leaf_frac;    /* param(22): max LR block size as fraction of N (1.0)     */
  double eta;              /* param(51): admissibility parameter (2.0)                */
  int    aca_kind;         /* param(60): 2 = ACA+ (only ACA+ is implemented)          */
  int    aca_norm;         /* param(61): 1 = relative to first cross (MREM)           */
  int    kt_predict;       /* param(62): predicted rank used by the storage balance (7)*/
  int    kmax;             /* param(63): max rank (1000)                              */
  int    kmin;             /* param(64): min rank (1)                                 */
  double aca_eps;          /* param(71): ACA tolerance                                */
  double aca_eps_factor;   /* param(72): absolute-pivot factor, ACA_EPS=factor*eps    */
  int    kernel;           /* HMVGEN_KERNEL_*                                         */
  int    nthreads;         /* OpenMP threads for the fill (<=0: runtime default)      */
} hmvgen_params;

/* Statistics of a built H-matrix, the figures the reference prints
 * (m_HACApK_base.f90:432-437, 1108-1169). */
typedef struct {
  int     nd;              /* unknowns                                  */
  int     nclst;           /* clusters in the bisection tree            */
  int     nlf;             /* leaves                                    */
  int     nlfkt;           /* low-rank leaves  (lnmtx(1))               */
  int     nlfdense;        /* dense leaves     (lnmtx(2))               */
  int     nhblocks;        /* non-leaf blocks  (lnmtx(3))               */
  int     ktmax;           /* max ACA rank                              */
  int     pad0;
  int64_t ncpc;            /* storage estimate with kt=param(62)         */
  int64_t mem_entries;     /* sum_LR (ndl+ndt)*kt + sum_D ndl*ndt        */
  double  mem_mbyte;       /* (mem_entries + nlf)*8/2^20 in REAL*4, as printed */
  double  time_tree;       /* seconds: tree + block tree                 */
  double  time_fill;       /* seconds: ACA+ and dense fill               */
} hmvgen_stats;

typedef struct hmvgen_hmatrix hmvgen_hmatrix;

void hmvgen_default_params(hmvgen_params *p);

/* Icosphere of 20*4^subdiv triangles on a sphere of radius r centred at c.
 * Call with nodes/faces NULL to query the sizes. */
int hmvgen_icosphere(int subdiv, double r, double cx, double cy, double cz,
                     int *nnode, int *nface, double *nodes, int *faces);

/* Reads a ppohBEM .pbf/.pbs mesh (format of bem-bb-fw-HACApK-0.4.2.f90:411-640).
 * Call with nodes/faces NULL to query the sizes.  dble_para receives the first
 * double parameter per face when non-NULL and present. */
int hmvgen_read_pbf(const char *path, int *nnode, int *nface,
                    double *nodes, int *faces, double *dble_para);

/* Writes a mesh in the same format (read back by Read_model_data,
 * bem-bb-fw-HACApK-0.4.2.f90:411-640): nodes nnode x 3, faces nface x 3
 * (0-based), dble_para nface x ndble (face-major; NULL when ndble = 0). */
int hmvgen_write_pbf(const char *path, int nnode, const double *nodes, int nface, const int *faces,
                     int ndble, const double *dble_para);

/* Writes the solution per face as a legacy ASCII VTK unstructured grid, the
 * file Print_result writes (bem-bb-fw-HACApK-0.4.2.f90:793-855). */
int hmvgen_write_vtk(const char *path, int nnode, const double *nodes, int nface, const int *faces,
                     const double *sol);

/* Builds the H-matrix for piecewise-constant collocation on the triangles. */
int hmvgen_build(int nnode, const double *nodes, int nface, const int *faces,
                 const hmvgen_params *p, hmvgen_hmatrix **out);

/* The tree and the leaf list only (HACApK_generate without the fill): every
 * leaf record has kt = 0 and no factors.  For a fill done elsewhere (the
 * engine's device construction, hmv_construct). */
int hmvgen_build_blocks(int nnode, const double *nodes, int nface, const int *faces,
                        const hmvgen_params *p, hmvgen_hmatrix **out);

void hmvgen_free(hmvgen_hmatrix *h);

/* Face areas (original order), as the entry function uses them. */
const double *hmvgen_areas(const hmvgen_hmatrix *h);

int  hmvgen_get_stats(const hmvgen_hmatrix *h, hmvgen_stats *s);

/* Leaf records (stc_HACApK_leafmtx layout), in the reference's order
 * (sorted by first row, then first column; m_HACApK_base.f90:451-461). */
const void *hmvgen_leaves(const hmvgen_hmatrix *h, int *nlf, int *stride);

/* lod(1:nd): cluster order -> original (1-based) face index. */
const int *hmvgen_lod(const hmvgen_hmatrix *h);

/* Centroids of the faces in original order, nface x 3. */
const double *hmvgen_centroids(const hmvgen_hmatrix *h);

/* Matrix entry A(i,j) for original 1-based indices (HACApK_entry_ij). */
double hmvgen_entry(const hmvgen_hmatrix *h, int i, int j);

/* Analytic potential of the flat triangle (xs,ys,zs) at the field point
 * (x,y,z), scaled as the kernel mode scales it: the reference's
 * face_integral, |int_T 1/r|/(4 pi eps0) with its literals, in PPOHBEM mode;
 * |int_T 1/r|/(4 pi) in CENTROID mode. */
double hmvgen_face_integral(int kernel, const double *xs, const double *ys,
                            const double *zs, double x, double y, double z);

/* The ACA helper routines used by the construction (HACApK_lib.f90:31-81),
 * exported so that tests can compare them with the reference's own:
 * scaled 2-norm, and first index (1-based) of max / min |za|. */
double hmvgen_util_unrm(int nd, const double *za);
int hmvgen_util_maxabs(int nd, const double *za, double *zz);
int hmvgen_util_minabs(int nd, const double *za, double *zz);

/* Deterministic i.i.d. uniform[-1,1) stream (splitmix64). */
void hmvgen_uniform(uint64_t seed, int64_t n, double *out);

const char *hmvgen_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* HMVGEN_H */
