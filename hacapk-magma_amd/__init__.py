"""Host-side Python mirror of the H-matvec engine (ctypes over the C ABI).

The compute path is libhmv.so (HIP, gfx950); this module only loads it,
describes its structs and moves host arrays in and out.  There is no CPU
fallback: every compute call raises HmvError when the library or the GPU is
missing.

Layers exposed here:
  * mesh helpers            icosphere(), merge(), read_pbf(), write_pbf(),
                            write_vtk()                               (libhmvgen)
  * H-matrix construction   HParams, HMatrix                          (libhmvgen)
  * engine                  Arena (arena build, apply, H-matmat, BiCGSTAB,
                            pipelined BiCGSTAB, GCR(m), block BiCGSTAB), MultiArena,
                            bicgstab_ir, bicgstab_ru, autotune             (libhmv)
  * reference-named calls   adot_body_lfcpy_batch_sorted, adot_body_lfmtx_batch,
                            adot_body_lfdel_batch, bicgstab_cax_lfmtx_gpu,
                            bicgstab_cax_lfmtx_pipe
                            (HACApK_BATCH_SORT.c:86, HACApK_BATCH.c:257,812,
                             HACApK_SOLVER.c:421, HACApK_SOLVER_PIPE.c:144) and solve()
                            (m_HACApK_use.f90:190-262)
"""
import ctypes as C
import os
import time

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIBDIR = os.path.join(_PKG, "lib")

PREC = {"f64": 0, "f32": 1, "bf16": 2, "bf16_all": 3}
KERNEL = {"ppohbem": 0, "centroid": 1}
SOLVE_STATUS = {0: "converged", 1: "maxiter", 2: "breakdown", 3: "nonfinite"}


class HmvError(RuntimeError):
    pass


# ---------------------------------------------------------------------------
# structs
# ---------------------------------------------------------------------------
def _leaf_ints(ptr, nlf, stride):
    """(ltmtx, kt, nstrtl, ndl, nstrtt, ndt) of nlf leaf records of `stride`
    bytes at address ptr, as an (nlf, 6) int64 table (the records' first six
    int32 fields, read through a strided view)."""
    if nlf == 0:
        return np.zeros((0, 6), np.int64)
    raw = np.ctypeslib.as_array((C.c_int32 * (nlf * stride // 4)).from_address(ptr))
    return raw.reshape(nlf, stride // 4)[:, :6].astype(np.int64)


class LeafRec(C.Structure):
    """stc_HACApK_leafmtx (HACApK_MAGMA.h:12-20)."""
    _fields_ = [("ltmtx", C.c_int), ("kt", C.c_int), ("nstrtl", C.c_int), ("ndl", C.c_int),
                ("nstrtt", C.c_int), ("ndt", C.c_int), ("a1size", C.c_longlong),
                ("a1", C.POINTER(C.c_double)), ("a2", C.POINTER(C.c_double))]


class _GenParams(C.Structure):
    _fields_ = [("leaf_size", C.c_int), ("max_leaf_frac", C.c_double), ("eta", C.c_double),
                ("aca_kind", C.c_int), ("aca_norm", C.c_int), ("kt_predict", C.c_int),
                ("kmax", C.c_int), ("kmin", C.c_int), ("aca_eps", C.c_double),
                ("aca_eps_factor", C.c_double), ("kernel", C.c_int), ("nthreads", C.c_int)]


class _GenStats(C.Structure):
    _fields_ = [("nd", C.c_int), ("nclst", C.c_int), ("nlf", C.c_int), ("nlfkt", C.c_int),
                ("nlfdense", C.c_int), ("nhblocks", C.c_int), ("ktmax", C.c_int), ("pad0", C.c_int),
                ("ncpc", C.c_int64), ("mem_entries", C.c_int64), ("mem_mbyte", C.c_double),
                ("time_tree", C.c_double), ("time_fill", C.c_double)]


class _ArenaStats(C.Structure):
    _fields_ = [("nd", C.c_int), ("precision", C.c_int), ("nlf", C.c_int64), ("nlr", C.c_int64),
                ("ndense", C.c_int64), ("entries_lr", C.c_int64), ("entries_dense", C.c_int64),
                ("tslots", C.c_int64), ("tparts", C.c_int64), ("phase1_bytes", C.c_int64),
                ("phase2_bytes", C.c_int64), ("arena_bytes", C.c_int64), ("payload_bytes", C.c_int64),
                ("alg_bytes", C.c_int64), ("col_groups", C.c_int64), ("row_groups", C.c_int64),
                ("atoms", C.c_int64), ("nwork_phase1", C.c_int64), ("nwork_phase2", C.c_int64),
                ("ktmax", C.c_int), ("pad", C.c_int), ("build_seconds", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("nstp", C.c_int), ("status", C.c_int), ("relres", C.c_double),
                ("true_relres", C.c_double), ("seconds", C.c_double), ("outer", C.c_int),
                ("inner_total", C.c_int)]

    def as_dict(self):
        d = {f[0]: getattr(self, f[0]) for f in self._fields_}
        d["status_name"] = SOLVE_STATUS.get(self.status, "?")
        return d


class _PlanTables(C.Structure):
    _fields_ = [("arena", C.c_void_p), ("arena_bytes", C.c_int64), ("work1", C.c_void_p),
                ("nwork1", C.c_int64), ("work2", C.c_void_p), ("nwork2", C.c_int64),
                ("zseg", C.c_void_p), ("nzseg", C.c_int64), ("outidx", C.c_void_p),
                ("noutidx", C.c_int64), ("codes", C.c_void_p), ("ncodes", C.c_int64),
                ("tfin", C.c_void_p), ("ntfin", C.c_int64),
                ("tsize", C.c_int64), ("tpart_size", C.c_int64),
                ("work_bytes", C.c_int), ("zseg_bytes", C.c_int), ("tfin_bytes", C.c_int),
                ("nwork1_small", C.c_int), ("nt1_small", C.c_int),
                ("nwork2_small", C.c_int), ("nt2_small", C.c_int)]


class ColumnInfo(C.Structure):
    """hmv_column_info: one column of a block solve."""
    _fields_ = [("nstp", C.c_int), ("status", C.c_int), ("relres", C.c_double),
                ("true_relres", C.c_double)]

    def as_dict(self):
        d = {f[0]: getattr(self, f[0]) for f in self._fields_}
        d["status_name"] = SOLVE_STATUS.get(self.status, "?")
        return d


class Tuning(C.Structure):
    """hmv_tuning: repack tunables (0 = the precision's default)."""
    _fields_ = [("p1_target_bytes", C.c_int64), ("p1_row_tile", C.c_int), ("p2_max_rows", C.c_int),
                ("p1_small_threads", C.c_int), ("p2_small_threads", C.c_int)]

    @classmethod
    def from_dict(cls, d):
        t = cls()
        for k, v in (d or {}).items():
            setattr(t, k, int(v))
        return t

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class BlockTuning(C.Structure):
    """hmv_block_tuning: the H-matmat path for a class of k."""
    _fields_ = [("vector", C.c_int), ("tiles_per_unit", C.c_int)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


WORK_DTYPE = np.dtype([("off0", "<i8"), ("off1", "<i8"), ("rows", "<i4"), ("rows_p", "<i4"),
                       ("ncol0", "<i4"), ("ncol1", "<i4"), ("dt0", "<i2"), ("dt1", "<i2"),
                       ("zseg_begin", "<i4"), ("zseg_end", "<i4"), ("out_base", "<i4")])
ZSEG_DTYPE = np.dtype([("kind", "<i4"), ("ncols", "<i4"), ("src", "<i4"), ("col", "<i4")])
TFIN_DTYPE = np.dtype([("slot", "<i4"), ("base", "<i4"), ("np", "<i4"), ("pad", "<i4")])


class RefLeafmtxp(C.Structure):
    """hmv_ref_leafmtxp: C view of st_HACApK_leafmtxp in a HAVE_MAGMA /
    HAVE_MAGMA_BATCH build (m_HACApK_base.f90:64-235, HACApK_MAGMA.h:28-106),
    the layout every GPU call site of the reference passes (544 bytes)."""
    _fields_ = [("nd", C.c_int), ("nlf", C.c_int), ("nlfkt", C.c_int), ("ktmax", C.c_int),
                ("st_lf_stride", C.c_int),
                ("m", C.c_int), ("n", C.c_int), ("gn", C.c_int), ("max_block", C.c_int),
                ("gflops", C.c_double), ("gpu_vectors", C.c_void_p * 6),
                ("num_batch", C.c_int), ("total_size_y", C.c_int), ("transA", C.c_int),
                ("batch_tables", C.c_void_p * 20),
                ("num_streamed", C.c_int), ("num_streamed_t", C.c_int),
                ("streamed_tables", C.c_void_p * 6), ("mgpu_tables", C.c_void_p * 25),
                ("mpi_rank", C.c_int), ("st_lf", C.c_void_p)]


class RefLeafmtxpPlain(C.Structure):
    """hmv_ref_leafmtxp_plain: st_HACApK_leafmtxp of a build without
    HAVE_MAGMA / HAVE_MAGMA_BATCH (m_HACApK_base.f90:64-76, 234)."""
    _fields_ = [("nd", C.c_int), ("nlf", C.c_int), ("nlfkt", C.c_int), ("ktmax", C.c_int),
                ("st_lf_stride", C.c_int), ("st_lf", C.c_void_p)]


REF_LAYOUT = {"magma": 0, "plain": 1}


class RefLcontrol(C.Structure):
    """stc_HACApK_lcontrol (HACApK_MAGMA.h:108-121)."""
    _fields_ = [("lf_umpi", C.c_int), ("lpmd_offset", C.c_int), ("lod_offset", C.c_int),
                ("lsp_offset", C.c_int), ("lnp_offset", C.c_int), ("lthr_offset", C.c_int),
                ("param", C.POINTER(C.c_double))]


# ---------------------------------------------------------------------------
# library loading
# ---------------------------------------------------------------------------
_gen = None
_hmv = None

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def _load(name):
    path = os.path.join(_LIBDIR, name)
    if not os.path.exists(path):
        raise HmvError("%s is not built (run hacapk-magma_amd/build.py)" % path)
    return C.CDLL(path)


def gen_lib():
    global _gen
    if _gen is None:
        L = _load("libhmvgen.so")
        L.hmvgen_default_params.argtypes = [C.POINTER(_GenParams)]
        L.hmvgen_icosphere.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                       _I, _I, _D, _I]
        L.hmvgen_read_pbf.argtypes = [C.c_char_p, _I, _I, _D, _I, _D]
        L.hmvgen_write_pbf.argtypes = [C.c_char_p, C.c_int, _D, C.c_int, _I, C.c_int, _D]
        L.hmvgen_build_blocks.argtypes = [C.c_int, _D, C.c_int, _I, C.POINTER(_GenParams),
                                          C.POINTER(C.c_void_p)]
        L.hmvgen_areas.argtypes = [C.c_void_p]
        L.hmvgen_areas.restype = _D
        L.hmvgen_write_vtk.argtypes = [C.c_char_p, C.c_int, _D, C.c_int, _I, _D]
        L.hmvgen_build.argtypes = [C.c_int, _D, C.c_int, _I, C.POINTER(_GenParams),
                                   C.POINTER(C.c_void_p)]
        L.hmvgen_free.argtypes = [C.c_void_p]
        L.hmvgen_get_stats.argtypes = [C.c_void_p, C.POINTER(_GenStats)]
        L.hmvgen_leaves.argtypes = [C.c_void_p, _I, _I]
        L.hmvgen_leaves.restype = C.c_void_p
        L.hmvgen_lod.argtypes = [C.c_void_p]
        L.hmvgen_lod.restype = _I
        L.hmvgen_centroids.argtypes = [C.c_void_p]
        L.hmvgen_centroids.restype = _D
        L.hmvgen_entry.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.hmvgen_entry.restype = C.c_double
        L.hmvgen_face_integral.argtypes = [C.c_int, _D, _D, _D, C.c_double, C.c_double, C.c_double]
        L.hmvgen_face_integral.restype = C.c_double
        L.hmvgen_uniform.argtypes = [C.c_uint64, C.c_int64, _D]
        L.hmvgen_util_unrm.argtypes = [C.c_int, _D]
        L.hmvgen_util_unrm.restype = C.c_double
        L.hmvgen_util_maxabs.argtypes = [C.c_int, _D, _D]
        L.hmvgen_util_minabs.argtypes = [C.c_int, _D, _D]
        L.hmvgen_last_error.restype = C.c_char_p
        _gen = L
    return _gen


def hmv_lib():
    global _hmv
    if _hmv is None:
        L = _load("libhmv.so")
        vp = C.c_void_p
        L.hmv_init.argtypes = [C.c_int]
        L.hmv_arena_build.argtypes = [C.c_int, C.c_int, vp, C.c_int, C.c_int, C.POINTER(vp)]
        L.hmv_arena_build_tuned.argtypes = [C.c_int, C.c_int, vp, C.c_int, C.c_int,
                                            C.POINTER(Tuning), C.POINTER(vp)]
        L.hmv_default_tuning.argtypes = [C.c_int, C.POINTER(Tuning)]
        L.hmv_apply_block.argtypes = [vp, C.c_int, vp, vp, C.c_double, vp]
        L.hmv_apply_block_static.argtypes = [vp, C.c_int, vp, vp, C.c_double, vp]
        L.hmv_bicgstab_ru.argtypes = [vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, C.c_double,
                                      C.POINTER(SolveInfo)]
        L.hmv_arena_build_multi.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                            C.POINTER(vp), C.c_int, C.c_int, C.POINTER(Tuning),
                                            C.POINTER(vp)]
        L.hmv_arena_segments.argtypes = [vp, C.POINTER(C.c_int), C.c_int]
        L.hmv_autotune_multi.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(vp), C.c_int, C.c_int, C.c_int,
                                         C.POINTER(Tuning), _D, _D]
        L.hmv_block_bicgstab_ir.argtypes = [vp, vp, C.c_int, vp, vp, C.c_double, C.c_double,
                                            C.c_int, C.c_int, C.c_int, C.POINTER(ColumnInfo),
                                            C.POINTER(SolveInfo)]
        L.hmv_multi_bicgstab_ir.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, C.c_int,
                                            C.c_int, C.c_int, C.POINTER(ColumnInfo), _D,
                                            C.POINTER(SolveInfo)]
        L.hmv_multi_gcr_ir.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.POINTER(ColumnInfo), _D,
                                       C.POINTER(SolveInfo)]
        L.hmv_multi_bicgstab.argtypes = [vp, vp, vp, C.c_double, C.c_int, C.c_int,
                                         C.POINTER(ColumnInfo), _D, C.POINTER(SolveInfo)]
        L.hmv_apply_block_timed.argtypes = [vp, C.c_int, vp, vp, C.c_int, _D]
        L.hmv_block_tuning_get.argtypes = [vp, C.c_int, C.POINTER(BlockTuning)]
        L.hmv_block_tuning_set.argtypes = [vp, C.c_int, C.POINTER(BlockTuning)]
        L.hmv_autotune_block.argtypes = [vp, C.c_int, C.c_int, C.POINTER(BlockTuning), _D, _D]
        L.hmv_block_bicgstab.argtypes = [vp, C.c_int, vp, vp, C.c_double, C.c_int, C.c_int,
                                         C.POINTER(ColumnInfo), C.POINTER(SolveInfo)]
        L.hmv_arena_tuning_get.argtypes = [vp, C.POINTER(Tuning)]
        L.hmv_autotune.argtypes = [C.c_int, C.c_int, vp, C.c_int, C.c_int, C.c_int, C.POINTER(Tuning),
                                   _D, _D]
        L.hmv_arena_stats_get.argtypes = [vp, C.POINTER(_ArenaStats)]
        L.hmv_arena_free.argtypes = [vp]
        L.hmv_arena_tbytes.argtypes = [vp]
        L.hmv_arena_tbytes.restype = C.c_int64
        L.hmv_apply.argtypes = [vp, vp, vp, C.c_double, vp]
        L.hmv_apply_host.argtypes = [vp, _D, _D]
        L.hmv_gcrm.argtypes = [vp, vp, vp, C.c_double, C.c_int, C.c_int, C.c_int, _D,
                               C.POINTER(SolveInfo)]
        L.hmv_bicgstab.argtypes = [vp, vp, vp, C.c_double, C.c_int, C.c_int, _D,
                                   C.POINTER(SolveInfo)]
        L.hmv_gcr_ir.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                 C.c_int, _D, C.POINTER(SolveInfo)]
        L.hmv_bicgstab_ir.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_int, _D, C.POINTER(SolveInfo)]
        L.hmv_last_error.restype = C.c_char_p
        L.hmv_device_malloc.argtypes = [C.POINTER(vp), C.c_size_t]
        L.hmv_device_free.argtypes = [vp]
        L.hmv_memcpy_htod.argtypes = [vp, vp, C.c_size_t]
        L.hmv_memcpy_dtoh.argtypes = [vp, vp, C.c_size_t]
        L.hmv_device_info.argtypes = [C.c_char_p, C.c_int, _I]
        L.hmv_apply_timed.argtypes = [vp, vp, vp, C.c_int, _D, _D, _D]
        L.hmv_stream_test.argtypes = [C.c_int, vp, vp, C.c_size_t, C.c_int, _D]
        L.hmv_plan_create.argtypes = [C.c_int, C.c_int, vp, C.c_int, C.c_int, C.POINTER(vp)]
        L.hmv_plan_view.argtypes = [vp, C.POINTER(_PlanTables)]
        L.hmv_plan_create_tuned.argtypes = [C.c_int, C.c_int, vp, C.c_int, C.c_int, C.POINTER(Tuning),
                                            C.POINTER(vp)]
        L.hmv_plan_free.argtypes = [vp]
        L.hmv_plan_save.argtypes = [vp, C.c_char_p]
        L.hmv_plan_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.hmv_arena_from_plan.argtypes = [vp, C.POINTER(vp)]
        L.hmv_arena_save.argtypes = [vp, C.c_char_p]
        L.hmv_arena_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.hmv_ref_set_layout.argtypes = [C.c_int]
        L.hmv_ref_get_layout.argtypes = []
        L.magma_init_.argtypes = []
        L.magma_init_.restype = None
        L.magma_finalize_.argtypes = []
        L.magma_finalize_.restype = None
        # the caller's st_HACApK_leafmtxp goes as void* (either layout)
        L.c_hacapk_adot_body_lfcpy_batch_sorted_.argtypes = [_I, vp]
        L.c_hacapk_adot_body_lfcpy_batch_sorted_.restype = None
        L.c_hacapk_adot_body_lfmtx_batch_.argtypes = [_D, vp, _D, _D, _D, _D, _D]
        L.c_hacapk_adot_body_lfmtx_batch_.restype = None
        L.c_hacapk_adot_body_lfdel_batch_.argtypes = [vp]
        L.c_hacapk_adot_body_lfdel_batch_.restype = None
        for n in ("c_hacapk_adot_body_lfmtx_", "c_hacapk_adot_body_lfmtx_gpu_"):
            getattr(L, n).argtypes = [_D, vp, _D, _D]
            getattr(L, n).restype = None
        L.c_hacapk_adot_body_lfcpy_gpu_.argtypes = [_I, vp]
        L.c_hacapk_adot_body_lfcpy_gpu_.restype = None
        L.c_hacapk_adot_body_lfdel_gpu_.argtypes = [vp]
        L.c_hacapk_adot_body_lfdel_gpu_.restype = None
        L.c_hacapk_bicgstab_cax_lfmtx_gpu_.argtypes = [vp, C.POINTER(RefLcontrol), _D, _D, _D, _I,
                                                       _I, _I]
        L.c_hacapk_bicgstab_cax_lfmtx_gpu_.restype = None
        L.c_hacapk_bicgstab_cax_lfmtx_pipe_.argtypes = L.c_hacapk_bicgstab_cax_lfmtx_gpu_.argtypes
        L.c_hacapk_bicgstab_cax_lfmtx_pipe_.restype = None
        L.c_hacapk_bicgstab_cax_lfmtx_pipe2_.argtypes = L.c_hacapk_bicgstab_cax_lfmtx_gpu_.argtypes
        L.c_hacapk_bicgstab_cax_lfmtx_pipe2_.restype = None
        for n in ("c_hacapk_bicgstab_cax_lfmtx_flat_", "c_hacapk_bicgstab_cax_lfmtx_mgpu1_"):
            getattr(L, n).argtypes = L.c_hacapk_bicgstab_cax_lfmtx_gpu_.argtypes
            getattr(L, n).restype = None
        L.hmv_bicgstab_pipe.argtypes = [vp, vp, vp, C.c_double, C.c_int, C.c_int, _D,
                                        C.POINTER(SolveInfo)]
        L.hmv_construct.argtypes = [C.c_int, _D, C.c_int, _I, _D, _D, _I, C.c_int, vp, C.c_int,
                                    C.POINTER(_ConstructParams), C.POINTER(vp)]
        L.hmv_dleaves_info.argtypes = [vp, _I, _I, _D]
        L.hmv_dleaves_download.argtypes = [vp, C.c_int, _D]
        L.hmv_arena_build_dleaves.argtypes = [vp, C.c_int, C.POINTER(Tuning), C.POINTER(vp)]
        L.hmv_dleaves_free.argtypes = [vp]
        _hmv = L
    return _hmv


def _check_gen(rc, what):
    if rc != 0:
        raise HmvError("%s: %s" % (what, gen_lib().hmvgen_last_error().decode()))


def _check(rc, what):
    if rc != 0:
        raise HmvError("%s: %s" % (what, hmv_lib().hmv_last_error().decode()))


def _dptr(a):
    return a.ctypes.data_as(_D)


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
def icosphere(subdiv, radius=1.0, center=(0.0, 0.0, 0.0)):
    """Icosahedron subdivided `subdiv` times onto a sphere: (nodes (n,3), faces (m,3) 0-based)."""
    L = gen_lib()
    nn, nf = C.c_int(), C.c_int()
    _check_gen(L.hmvgen_icosphere(subdiv, radius, *center, C.byref(nn), C.byref(nf), None, None),
               "icosphere")
    nodes = np.empty((nn.value, 3), np.float64)
    faces = np.empty((nf.value, 3), np.int32)
    _check_gen(L.hmvgen_icosphere(subdiv, radius, *center, C.byref(nn), C.byref(nf),
                                  _dptr(nodes), faces.ctypes.data_as(_I)), "icosphere")
    return nodes, faces


def merge(*meshes):
    """Concatenates (nodes, faces) meshes into one."""
    nodes, faces, off = [], [], 0
    for n, f in meshes:
        nodes.append(n)
        faces.append(f + off)
        off += n.shape[0]
    return np.ascontiguousarray(np.vstack(nodes)), np.ascontiguousarray(np.vstack(faces).astype(np.int32))


def read_pbf(path):
    """ppohBEM mesh file -> (nodes, faces, first double parameter per face)."""
    L = gen_lib()
    nn, nf = C.c_int(), C.c_int()
    _check_gen(L.hmvgen_read_pbf(path.encode(), C.byref(nn), C.byref(nf), None, None, None), "read_pbf")
    nodes = np.empty((nn.value, 3), np.float64)
    faces = np.empty((nf.value, 3), np.int32)
    para = np.empty(nf.value, np.float64)
    _check_gen(L.hmvgen_read_pbf(path.encode(), C.byref(nn), C.byref(nf), _dptr(nodes),
                                 faces.ctypes.data_as(_I), _dptr(para)), "read_pbf")
    return nodes, faces, para


def write_pbf(path, nodes, faces, dble_para=None):
    """Mesh file in the reference's .pbf layout (hmvgen_write_pbf); dble_para:
    None, one value per face, or an (nface, ndble) array."""
    nodes = np.ascontiguousarray(nodes, np.float64).reshape(-1, 3)
    faces = np.ascontiguousarray(faces, np.int32).reshape(-1, 3)
    if dble_para is None:
        nd, dp = 0, None
    else:
        dp = np.ascontiguousarray(dble_para, np.float64).reshape(len(faces), -1)
        nd = dp.shape[1]
    _check_gen(gen_lib().hmvgen_write_pbf(os.fsencode(path), len(nodes), _dptr(nodes), len(faces),
                                          faces.ctypes.data_as(_I), nd,
                                          _dptr(dp) if dp is not None else None), "write_pbf")


def write_vtk(path, nodes, faces, sol):
    """Solution per face as the reference's VTK output (hmvgen_write_vtk)."""
    nodes = np.ascontiguousarray(nodes, np.float64).reshape(-1, 3)
    faces = np.ascontiguousarray(faces, np.int32).reshape(-1, 3)
    sol = np.ascontiguousarray(sol, np.float64)
    assert sol.shape == (len(faces),)
    _check_gen(gen_lib().hmvgen_write_vtk(os.fsencode(path), len(nodes), _dptr(nodes), len(faces),
                                          faces.ctypes.data_as(_I), _dptr(sol)), "write_vtk")


def uniform(seed, n):
    """Deterministic i.i.d. uniform[-1,1) vector (splitmix64 stream)."""
    out = np.empty(int(n), np.float64)
    gen_lib().hmvgen_uniform(C.c_uint64(seed), int(n), _dptr(out))
    return out


# ---------------------------------------------------------------------------
# H-matrix construction
# ---------------------------------------------------------------------------
class HParams:
    """Construction parameters; names follow the reference's param() slots
    (defaults m_HACApK_base.f90:294-313)."""

    def __init__(self, **kw):
        p = _GenParams()
        gen_lib().hmvgen_default_params(C.byref(p))
        self._p = p
        for k, v in kw.items():
            self[k] = v

    def __setitem__(self, k, v):
        if k == "kernel" and isinstance(v, str):
            v = KERNEL[v]
        if not hasattr(self._p, k):
            raise KeyError(k)
        setattr(self._p, k, v)

    def __getitem__(self, k):
        return getattr(self._p, k)

    def as_dict(self):
        return {f[0]: getattr(self._p, f[0]) for f in _GenParams._fields_}


def baseline_params(aca_eps=1e-4, leaf_size=20, eta=2.0, kernel="centroid", nthreads=0):
    """BEM configuration of the H-matvec benchmarks: 1/(4 pi r) kernel with
    centroid quadrature off the diagonal, analytic self terms."""
    return HParams(leaf_size=leaf_size, eta=eta, aca_eps=aca_eps, kernel=kernel, nthreads=nthreads)


class _ConstructParams(C.Structure):
    _fields_ = [("kernel", C.c_int), ("kmax", C.c_int), ("kmin", C.c_int), ("pad", C.c_int),
                ("aca_eps", C.c_double), ("aca_eps_factor", C.c_double)]


class DeviceHMatrix:
    """An H-matrix whose leaves are filled on the device (hmv_construct): the
    host builds the cluster tree and the leaf list (hmvgen_build_blocks), the
    device evaluates the dense leaves and runs ACA+ on the low-rank ones.
    The factors stay in device memory; arena() repacks them without a host
    copy, download() gives host leaves in the reference's layout."""

    def __init__(self, nodes, faces, params=None):
        L = gen_lib()
        nodes = np.ascontiguousarray(nodes, np.float64)
        faces = np.ascontiguousarray(faces, np.int32)
        params = params or HParams()
        self.params = params
        t0 = time.perf_counter()
        h = C.c_void_p()
        _check_gen(L.hmvgen_build_blocks(nodes.shape[0], _dptr(nodes), faces.shape[0],
                                         faces.ctypes.data_as(_I), C.byref(params._p), C.byref(h)),
                   "hmvgen_build_blocks")
        try:
            st = _GenStats()
            _check_gen(L.hmvgen_get_stats(h, C.byref(st)), "stats")
            self.nd = st.nd
            nlf, stride = C.c_int(), C.c_int()
            specs = L.hmvgen_leaves(h, C.byref(nlf), C.byref(stride))
            self.nlf = nlf.value
            self.lod = np.ctypeslib.as_array(L.hmvgen_lod(h), shape=(self.nd,)).copy()
            self.centroids = np.ctypeslib.as_array(L.hmvgen_centroids(h), shape=(self.nd, 3)).copy()
            self.areas = np.ctypeslib.as_array(L.hmvgen_areas(h), shape=(self.nd,)).copy()
            self.block_seconds = time.perf_counter() - t0
            cp = _ConstructParams(kernel=params["kernel"], kmax=params["kmax"], kmin=params["kmin"],
                                  aca_eps=params["aca_eps"], aca_eps_factor=params["aca_eps_factor"])
            d = C.c_void_p()
            lod = np.ascontiguousarray(self.lod, np.int32)
            _check(hmv_lib().hmv_construct(nodes.shape[0], _dptr(nodes), faces.shape[0],
                                           faces.ctypes.data_as(_I), _dptr(self.centroids),
                                           _dptr(self.areas), lod.ctypes.data_as(_I), self.nlf,
                                           specs, stride.value, C.byref(cp), C.byref(d)),
                   "hmv_construct")
            self._d = d
            self.table = _leaf_ints(specs, self.nlf, stride.value)
            self.table[:, 1] = 0
        finally:
            L.hmvgen_free(h)
        kt = np.zeros(self.nlf, np.int32)
        sec = C.c_double()
        _check(hmv_lib().hmv_dleaves_info(self._d, None, kt.ctypes.data_as(_I), C.byref(sec)),
               "hmv_dleaves_info")
        self.table[:, 1] = kt
        self.kt = kt
        self.fill_seconds = sec.value

    def arena(self, precision="f64", tuning=None):
        """Arena repacked from the device factors (hmv_arena_build_dleaves)."""
        a = C.c_void_p()
        t = Tuning.from_dict(tuning) if tuning else None
        _check(hmv_lib().hmv_arena_build_dleaves(self._d, PREC[precision],
                                                 C.byref(t) if t is not None else None, C.byref(a)),
               "hmv_arena_build_dleaves")
        ar = Arena.__new__(Arena)
        ar.nd = self.nd
        ar.precision = precision
        ar._finish(a)
        return ar

    def download(self):
        """Host copy of every leaf: an object with nd, nlf, leaves_ptr, stride
        and leaf_table(), usable wherever an HMatrix's leaves are."""
        return _HostLeaves(self)

    def free(self):
        if getattr(self, "_d", None):
            hmv_lib().hmv_dleaves_free(self._d)
            self._d = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _HostLeaves:
    def __init__(self, dh):
        self.nd, self.nlf = dh.nd, dh.nlf
        self._rec = (LeafRec * self.nlf)()
        self._bufs = []
        for i, (lt, kt, l0, ndl, t0, ndt) in enumerate(dh.table.tolist()):
            n1 = ndt * kt if lt == 1 else ndt * ndl
            n2 = ndl * kt if lt == 1 else 0
            buf = np.zeros(max(1, n1 + n2))
            if n1 + n2 > 0:
                _check(hmv_lib().hmv_dleaves_download(dh._d, i, _dptr(buf)), "hmv_dleaves_download")
            r = self._rec[i]
            r.ltmtx, r.kt, r.nstrtl, r.ndl, r.nstrtt, r.ndt = lt, kt, l0, ndl, t0, ndt
            r.a1size = 8 * n1
            r.a1 = buf.ctypes.data_as(C.POINTER(C.c_double))
            r.a2 = C.cast(buf.ctypes.data + 8 * n1, C.POINTER(C.c_double)) if lt == 1 else None
            self._bufs.append(buf)
        self.leaves_ptr = C.addressof(self._rec)
        self.stride = C.sizeof(LeafRec)
        self._table = dh.table.copy()

    def leaves(self):
        return self._rec

    def leaf_table(self):
        return self._table


class HMatrix:
    """An H-matrix built on the host: leaves in the reference's record layout."""

    def __init__(self, nodes, faces, params=None):
        L = gen_lib()
        nodes = np.ascontiguousarray(nodes, np.float64)
        faces = np.ascontiguousarray(faces, np.int32)
        params = params or HParams()
        h = C.c_void_p()
        _check_gen(L.hmvgen_build(nodes.shape[0], _dptr(nodes), faces.shape[0],
                                  faces.ctypes.data_as(_I), C.byref(params._p), C.byref(h)),
                   "hmvgen_build")
        self._h = h
        self.params = params
        st = _GenStats()
        _check_gen(L.hmvgen_get_stats(h, C.byref(st)), "stats")
        self.stats = {f[0]: getattr(st, f[0]) for f in _GenStats._fields_ if not f[0].startswith("pad")}
        self.nd = st.nd
        nlf, stride = C.c_int(), C.c_int()
        self.leaves_ptr = L.hmvgen_leaves(h, C.byref(nlf), C.byref(stride))
        self.nlf, self.stride = nlf.value, stride.value
        lod = L.hmvgen_lod(h)
        self.lod = np.ctypeslib.as_array(lod, shape=(self.nd,)).copy()
        cen = L.hmvgen_centroids(h)
        self.centroids = np.ctypeslib.as_array(cen, shape=(self.nd, 3)).copy()

    def close(self):
        if getattr(self, "_h", None):
            gen_lib().hmvgen_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def entry(self, i, j):
        """A(i,j) for original 1-based indices (HACApK_entry_ij)."""
        return gen_lib().hmvgen_entry(self._h, i, j)

    def leaves(self):
        """Leaf records as a ctypes array view (no copy)."""
        return (LeafRec * self.nlf).from_address(self.leaves_ptr)

    def leaf_table(self):
        """numpy table of (ltmtx, kt, nstrtl, ndl, nstrtt, ndt) per leaf."""
        return _leaf_ints(self.leaves_ptr, self.nlf, self.stride)

    def to_cluster_order(self, v):
        """Original order -> cluster order (b(:nd)=rhs(lod(:nd)), m_HACApK_use.f90:285)."""
        return np.ascontiguousarray(np.asarray(v)[self.lod - 1])

    def to_original_order(self, v):
        """Cluster order -> original order (sol(lod(:nd))=u(:nd), m_HACApK_use.f90:469)."""
        out = np.empty_like(v)
        out[self.lod - 1] = v
        return out


# ---------------------------------------------------------------------------
# device buffers
# ---------------------------------------------------------------------------
class DeviceBuffer:
    def __init__(self, nbytes):
        L = hmv_lib()
        p = C.c_void_p()
        _check(L.hmv_device_malloc(C.byref(p), int(nbytes)), "device malloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a, np.float64)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        _check(hmv_lib().hmv_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes), "htod")

    def download(self, n=None):
        n = self.nbytes // 8 if n is None else n
        out = np.empty(n, np.float64)
        _check(hmv_lib().hmv_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes), "dtoh")
        return out

    def free(self):
        if self.ptr:
            hmv_lib().hmv_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def stream_test(kind, nbytes, reps=20):
    """Average ms of a plain read (kind 0) or copy (kind 1) kernel over nbytes."""
    src = DeviceBuffer(nbytes)
    dst = DeviceBuffer(nbytes) if kind == 1 else DeviceBuffer(64)
    _check(hmv_lib().hmv_memcpy_htod(src.ptr, np.zeros(1).ctypes.data, 8), "init")
    ms = C.c_double()
    _check(hmv_lib().hmv_stream_test(kind, src.ptr, dst.ptr, int(nbytes), int(reps), C.byref(ms)),
           "hmv_stream_test")
    src.free()
    dst.free()
    return ms.value


def mfma_rate(mfmas_per_wave=65536, reps=5):
    """Measured FP64 MFMA rate in TFLOP/s (hmv_stream_test kind 2: 16x16x4
    FP64 MFMAs on register operands, 4 waves per SIMD)."""
    dst = DeviceBuffer(64)
    ms = C.c_double()
    _check(hmv_lib().hmv_stream_test(2, None, dst.ptr, int(mfmas_per_wave), int(reps),
                                     C.byref(ms)), "hmv_stream_test")
    dst.free()
    return 2048.0 * mfmas_per_wave * 4096 / (ms.value * 1e-3) / 1e12


def device_info():
    L = hmv_lib()
    buf = C.create_string_buffer(256)
    cu = C.c_int()
    _check(L.hmv_device_info(buf, 256, C.byref(cu)), "device info")
    return buf.value.decode(), cu.value


# ---------------------------------------------------------------------------
# engine
# ---------------------------------------------------------------------------
class Arena:
    """Device-resident arena of an H-matrix (hmv_arena_build)."""

    def __init__(self, hm, precision="f64", nd=None, nlf=None, leaves_ptr=None, stride=None,
                 tuning=None):
        L = hmv_lib()
        if hm is not None:
            nd, nlf, leaves_ptr, stride = hm.nd, hm.nlf, hm.leaves_ptr, hm.stride
        self.nd = nd
        self.precision = precision
        a = C.c_void_p()
        if tuning:
            t = Tuning.from_dict(tuning)
            _check(L.hmv_arena_build_tuned(nd, nlf, leaves_ptr, stride, PREC[precision], C.byref(t),
                                           C.byref(a)), "hmv_arena_build_tuned")
        else:
            _check(L.hmv_arena_build(nd, nlf, leaves_ptr, stride, PREC[precision], C.byref(a)),
                   "hmv_arena_build")
        self._finish(a)

    @classmethod
    def load(cls, path):
        """Arena from an arena file (hmv_arena_load): no leaves, no repack."""
        a = C.c_void_p()
        _check(hmv_lib().hmv_arena_load(os.fsencode(path), C.byref(a)), "hmv_arena_load")
        self = cls.__new__(cls)
        st = _ArenaStats()
        _check(hmv_lib().hmv_arena_stats_get(a, C.byref(st)), "arena stats")
        self.nd = st.nd
        self.precision = {v: k for k, v in PREC.items()}[st.precision]
        self._finish(a)
        return self

    def save(self, path):
        """Writes the arena file (hmv_arena_save)."""
        _check(hmv_lib().hmv_arena_save(self._a, os.fsencode(path)), "hmv_arena_save")

    def _finish(self, a):
        self._a = a
        st = _ArenaStats()
        _check(hmv_lib().hmv_arena_stats_get(a, C.byref(st)), "arena stats")
        self.stats = {f[0]: getattr(st, f[0]) for f in _ArenaStats._fields_ if f[0] != "pad"}

    @property
    def handle(self):
        return self._a

    @property
    def tuning(self):
        """The tuning this arena was built with (defaults resolved)."""
        t = Tuning()
        _check(hmv_lib().hmv_arena_tuning_get(self._a, C.byref(t)), "hmv_arena_tuning_get")
        return t.as_dict()

    def free(self):
        if getattr(self, "_a", None):
            hmv_lib().hmv_arena_free(self._a)
            self._a = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def apply(self, x_ptr, y_ptr, beta=0.0, stream=None):
        """y = H x + beta y on device pointers (asynchronous)."""
        _check(hmv_lib().hmv_apply(self._a, x_ptr, y_ptr, float(beta), stream), "hmv_apply")

    def matvec(self, x):
        """Host convenience: returns H x (x, result in cluster order)."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        _check(hmv_lib().hmv_apply_host(self._a, _dptr(x), _dptr(y)), "hmv_apply_host")
        return y

    def apply_host(self, zu, zau):
        """zau += H zu on host arrays, in place (hmv_apply_host, the reference's
        zau += H*zu); zu and zau contiguous float64 arrays of length nd."""
        for v in (zu, zau):
            if v.dtype != np.float64 or v.shape != (self.nd,) or not v.flags.c_contiguous:
                raise ValueError("zu and zau must be contiguous float64 arrays of length %d" % self.nd)
        _check(hmv_lib().hmv_apply_host(self._a, _dptr(zu), _dptr(zau)), "hmv_apply_host")

    def apply_block(self, x_ptr, y_ptr, k, beta=0.0, stream=None):
        """Y = H X + beta Y for k row-interleaved right-hand sides on device (asynchronous)."""
        _check(hmv_lib().hmv_apply_block(self._a, int(k), x_ptr, y_ptr, float(beta), stream),
               "hmv_apply_block")

    def matmat(self, X):
        """Host convenience: H X for an (N, k) array (cluster order); returns (N, k)."""
        X = np.ascontiguousarray(X, np.float64)
        assert X.ndim == 2 and X.shape[0] == self.nd
        k = X.shape[1]
        dx, dy = DeviceBuffer.from_array(X.reshape(-1)), DeviceBuffer.from_array(np.zeros(X.size))
        self.apply_block(dx.ptr, dy.ptr, k)
        return dy.download().reshape(self.nd, k)

    def block_tuning(self, k):
        """The H-matmat path for k right-hand sides (hmv_block_tuning_get)."""
        t = BlockTuning()
        _check(hmv_lib().hmv_block_tuning_get(self._a, int(k), C.byref(t)), "hmv_block_tuning_get")
        return t.as_dict()

    def set_block_tuning(self, k, vector, tiles_per_unit=0):
        """Sets the H-matmat path of k's classes (hmv_block_tuning_set)."""
        t = BlockTuning(int(vector), int(tiles_per_unit))
        _check(hmv_lib().hmv_block_tuning_set(self._a, int(k), C.byref(t)), "hmv_block_tuning_set")

    def autotune_block(self, k, reps=5):
        """hmv_autotune_block: sets and returns the fastest H-matmat path for
        k right-hand sides, with its ms per product and the default's."""
        t = BlockTuning()
        best, dflt = C.c_double(), C.c_double()
        _check(hmv_lib().hmv_autotune_block(self._a, int(k), int(reps), C.byref(t), C.byref(best),
                                            C.byref(dflt)), "hmv_autotune_block")
        return t.as_dict(), best.value, dflt.value

    def apply_block_timed(self, x_ptr, y_ptr, k, reps):
        ms = C.c_double()
        _check(hmv_lib().hmv_apply_block_timed(self._a, int(k), x_ptr, y_ptr, int(reps), C.byref(ms)),
               "hmv_apply_block_timed")
        return ms.value

    def block_bicgstab(self, B, U0=None, eps=1e-8, maxit=1000, check_every=8):
        """k independent BiCGSTAB solves sharing one H-matmat per step; B is (N, k)
        (k a power of two <= 128).  Returns (U, per-column info list, info)."""
        B = np.ascontiguousarray(B, np.float64)
        k = B.shape[1]
        U0 = np.zeros_like(B) if U0 is None else np.ascontiguousarray(U0, np.float64)
        db, du = DeviceBuffer.from_array(B.reshape(-1)), DeviceBuffer.from_array(U0.reshape(-1))
        cols = (ColumnInfo * k)()
        info = SolveInfo()
        _check(hmv_lib().hmv_block_bicgstab(self._a, k, db.ptr, du.ptr, float(eps), int(maxit),
                                            int(check_every), cols, C.byref(info)),
               "hmv_block_bicgstab")
        return du.download().reshape(self.nd, k), [c.as_dict() for c in cols], info.as_dict()

    def apply_timed(self, x_ptr, y_ptr, reps):
        m1, m2, ma = C.c_double(), C.c_double(), C.c_double()
        _check(hmv_lib().hmv_apply_timed(self._a, x_ptr, y_ptr, int(reps), C.byref(m1), C.byref(m2),
                                         C.byref(ma)), "hmv_apply_timed")
        return m1.value, m2.value, ma.value

    def bicgstab(self, b, u0=None, eps=1e-8, maxit=500, check_every=1, history=True):
        """BiCGSTAB on device; host b/u0 in cluster order.  Returns (u, info, hist)."""
        b = np.ascontiguousarray(b, np.float64)
        u0 = np.zeros_like(b) if u0 is None else np.ascontiguousarray(u0, np.float64)
        db, du = DeviceBuffer.from_array(b), DeviceBuffer.from_array(u0)
        hist = np.full(maxit + 1, np.nan) if history else None
        info = SolveInfo()
        _check(hmv_lib().hmv_bicgstab(self._a, db.ptr, du.ptr, eps, maxit, check_every,
                                      _dptr(hist) if history else None, C.byref(info)), "hmv_bicgstab")
        u = du.download()
        db.free()
        du.free()
        if history:
            hist = hist[: info.nstp + 1]
        return u, info.as_dict(), hist

    def bicgstab_pipe(self, b, u0=None, eps=1e-8, maxit=500, check_every=1, history=True):
        """Pipelined BiCGSTAB on device (HACApK_SOLVER_PIPE.c:144-637); host b/u0
        in cluster order.  Returns (u, info, hist)."""
        b = np.ascontiguousarray(b, np.float64)
        u0 = np.zeros_like(b) if u0 is None else np.ascontiguousarray(u0, np.float64)
        db, du = DeviceBuffer.from_array(b), DeviceBuffer.from_array(u0)
        hist = np.full(maxit + 1, np.nan) if history else None
        info = SolveInfo()
        try:
            _check(hmv_lib().hmv_bicgstab_pipe(self._a, db.ptr, du.ptr, eps, maxit, check_every,
                                               _dptr(hist) if history else None, C.byref(info)),
                   "hmv_bicgstab_pipe")
            u = du.download()
        finally:
            db.free()
            du.free()
        if history:
            hist = hist[: info.nstp + 1]
        return u, info.as_dict(), hist

    def gcrm(self, b, u0=None, eps=1e-8, maxit=500, mreset=8, check_every=1, history=True):
        """GCR(mreset) on device (HACApK_gcrm_lfmtx, m_HACApK_solve.f90:506-571);
        host b/u0 in cluster order.  Returns (u, info, hist)."""
        b = np.ascontiguousarray(b, np.float64)
        u0 = np.zeros_like(b) if u0 is None else np.ascontiguousarray(u0, np.float64)
        db, du = DeviceBuffer.from_array(b), DeviceBuffer.from_array(u0)
        hist = np.full(maxit + 1, np.nan) if history else None
        info = SolveInfo()
        try:
            _check(hmv_lib().hmv_gcrm(self._a, db.ptr, du.ptr, eps, maxit, mreset, check_every,
                                      _dptr(hist) if history else None, C.byref(info)), "hmv_gcrm")
            u = du.download()
        finally:
            db.free()
            du.free()
        if history:
            hist = hist[: info.nstp + 1]
        return u, info.as_dict(), hist


def _plan_create(hm, precision, tuning):
    p = C.c_void_p()
    t = Tuning.from_dict(tuning) if tuning else None
    _check(hmv_lib().hmv_plan_create_tuned(hm.nd, hm.nlf, hm.leaves_ptr, hm.stride, PREC[precision],
                                           C.byref(t) if t is not None else None, C.byref(p)),
           "hmv_plan_create")
    return p


def plan_tables(hm, precision="f64", tuning=None):
    """Host-side repack of an HMatrix (hmv_plan_create_tuned): numpy copies of
    the arena image and the engine's segment/work tables.  Needs no device."""
    return _plan_view(_plan_create(hm, precision, tuning))


def save_plan(hm, path, precision="f64", tuning=None):
    """Repacks an HMatrix on the host and writes the arena file (hmv_plan_save).
    Needs no device."""
    p = _plan_create(hm, precision, tuning)
    try:
        _check(hmv_lib().hmv_plan_save(p, os.fsencode(path)), "hmv_plan_save")
    finally:
        hmv_lib().hmv_plan_free(p)


def load_plan_tables(path):
    """Reads and checks an arena file (hmv_plan_load); returns the tables as
    plan_tables does.  Needs no device."""
    p = C.c_void_p()
    _check(hmv_lib().hmv_plan_load(os.fsencode(path), C.byref(p)), "hmv_plan_load")
    return _plan_view(p)


def _plan_view(p):
    """Tables of plan handle p as numpy copies; frees the plan."""
    L = hmv_lib()
    try:
        t = _PlanTables()
        _check(L.hmv_plan_view(p, C.byref(t)), "hmv_plan_view")
        assert t.work_bytes == WORK_DTYPE.itemsize and t.zseg_bytes == ZSEG_DTYPE.itemsize
        assert t.tfin_bytes == TFIN_DTYPE.itemsize

        def arr(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dt)
            return np.frombuffer((C.c_char * (n * np.dtype(dt).itemsize)).from_address(ptr), dt).copy()

        return {
            "arena": arr(t.arena, t.arena_bytes, np.uint8),
            "work1": arr(t.work1, t.nwork1, WORK_DTYPE), "work2": arr(t.work2, t.nwork2, WORK_DTYPE),
            "zseg": arr(t.zseg, t.nzseg, ZSEG_DTYPE), "outidx": arr(t.outidx, t.noutidx, np.int32),
            "codes": arr(t.codes, t.ncodes, np.int32), "tfin": arr(t.tfin, t.ntfin, TFIN_DTYPE),
            "tsize": t.tsize, "tpart_size": t.tpart_size,
            "nwork1_small": t.nwork1_small, "nt1_small": t.nt1_small,
            "nwork2_small": t.nwork2_small, "nt2_small": t.nt2_small,
        }
    finally:
        L.hmv_plan_free(p)


class MultiArena(Arena):
    """Many independent H-matrices in one arena (hmv_arena_build_multi): one
    launch per phase applies all of them; vectors are concatenated."""

    def __init__(self, hms, precision="f64", tuning=None):
        L = hmv_lib()
        n = len(hms)
        strides = {hm.stride for hm in hms}
        assert len(strides) == 1, "leaf records of one layout"
        nd = (C.c_int * n)(*[hm.nd for hm in hms])
        nlf = (C.c_int * n)(*[hm.nlf for hm in hms])
        lv = (C.c_void_p * n)(*[hm.leaves_ptr for hm in hms])
        t = Tuning.from_dict(tuning) if tuning else None
        a = C.c_void_p()
        _check(L.hmv_arena_build_multi(n, nd, nlf, lv, strides.pop(), PREC[precision],
                                       C.byref(t) if t is not None else None, C.byref(a)),
               "hmv_arena_build_multi")
        self.nd = sum(hm.nd for hm in hms)
        self.precision = precision
        self._finish(a)
        off = (C.c_int * (n + 1))()
        assert L.hmv_arena_segments(a, off, n + 1) == n
        self.offsets = np.array(off[:], dtype=np.int64)

    def bicgstab_multi(self, b, u0=None, eps=1e-8, maxit=1000, check_every=8):
        """One BiCGSTAB per instance, all in lockstep.  Returns (u, per-instance
        info list, per-instance seconds, info)."""
        b = np.ascontiguousarray(b, np.float64)
        u0 = np.zeros_like(b) if u0 is None else np.ascontiguousarray(u0, np.float64)
        n = len(self.offsets) - 1
        db, du = DeviceBuffer.from_array(b), DeviceBuffer.from_array(u0)
        inst = (ColumnInfo * n)()
        secs = np.zeros(n)
        info = SolveInfo()
        _check(hmv_lib().hmv_multi_bicgstab(self._a, db.ptr, du.ptr, float(eps), int(maxit),
                                            int(check_every), inst, _dptr(secs), C.byref(info)),
               "hmv_multi_bicgstab")
        return du.download(), [c.as_dict() for c in inst], secs, info.as_dict()


def block_bicgstab_ir(a64, alow, B, U0=None, eps=1e-8, inner_eps=1e-4, max_outer=30,
                      max_inner=2000, check_every=8):
    """Mixed-precision refinement of a block solve (hmv_block_bicgstab_ir): B
    is (N, k), k a power of two <= 128.  Returns (U, per-column info list,
    info)."""
    B = np.ascontiguousarray(B, np.float64)
    k = B.shape[1]
    U0 = np.zeros_like(B) if U0 is None else np.ascontiguousarray(U0, np.float64)
    db, du = DeviceBuffer.from_array(B.reshape(-1)), DeviceBuffer.from_array(U0.reshape(-1))
    cols = (ColumnInfo * k)()
    info = SolveInfo()
    try:
        _check(hmv_lib().hmv_block_bicgstab_ir(a64.handle, alow.handle, k, db.ptr, du.ptr,
                                               float(eps), float(inner_eps), int(max_outer),
                                               int(max_inner), int(check_every), cols,
                                               C.byref(info)), "hmv_block_bicgstab_ir")
        U = du.download().reshape(a64.nd, k)
    finally:
        db.free()
        du.free()
    return U, [c.as_dict() for c in cols], info.as_dict()


def multi_bicgstab_ir(m64, mlow, b, u0=None, eps=1e-8, inner_eps=1e-4, max_outer=30,
                      max_inner=2000, check_every=4):
    """Mixed-precision refinement per instance over two MultiArenas of the
    same instances (hmv_multi_bicgstab_ir).  Returns (u, per-instance info
    list, per-instance seconds, info)."""
    b = np.ascontiguousarray(b, np.float64)
    u0 = np.zeros_like(b) if u0 is None else np.ascontiguousarray(u0, np.float64)
    n = len(m64.offsets) - 1
    db, du = DeviceBuffer.from_array(b), DeviceBuffer.from_array(u0)
    inst = (ColumnInfo * n)()
    secs = np.zeros(n)
    info = SolveInfo()
    try:
        _check(hmv_lib().hmv_multi_bicgstab_ir(m64.handle, mlow.handle, db.ptr, du.ptr, float(eps),
                                               float(inner_eps), int(max_outer), int(max_inner),
                                               int(check_every), inst, _dptr(secs), C.byref(info)),
               "hmv_multi_bicgstab_ir")
        u = du.download()
    finally:
        db.free()
        du.free()
    return u, [c.as_dict() for c in inst], secs, info.as_dict()


def multi_gcr_ir(m64, mlow, b, u0=None, eps=1e-8, inner_eps=5e-3, m=32, max_outer=30,
                 max_inner=2000, check_every=4):
    """Mixed-precision refinement per instance with GCR(m) inner solves over
    two MultiArenas of the same instances (hmv_multi_gcr_ir).  Returns (u,
    per-instance info list, per-instance seconds, info)."""
    b = np.ascontiguousarray(b, np.float64)
    u0 = np.zeros_like(b) if u0 is None else np.ascontiguousarray(u0, np.float64)
    n = len(m64.offsets) - 1
    db, du = DeviceBuffer.from_array(b), DeviceBuffer.from_array(u0)
    inst = (ColumnInfo * n)()
    secs = np.zeros(n)
    info = SolveInfo()
    try:
        _check(hmv_lib().hmv_multi_gcr_ir(m64.handle, mlow.handle, db.ptr, du.ptr, float(eps),
                                          float(inner_eps), int(m), int(max_outer), int(max_inner),
                                          int(check_every), inst, _dptr(secs), C.byref(info)),
               "hmv_multi_gcr_ir")
        u = du.download()
    finally:
        db.free()
        du.free()
    return u, [c.as_dict() for c in inst], secs, info.as_dict()


def autotune_multi(hms, precision="f64", reps=10):
    """hmv_autotune_multi over the instances hms: returns (best tuning dict,
    ms per apply of all instances, ms with the defaults)."""
    n = len(hms)
    strides = {hm.stride for hm in hms}
    assert len(strides) == 1, "leaf records of one layout"
    nd = (C.c_int * n)(*[hm.nd for hm in hms])
    nlf = (C.c_int * n)(*[hm.nlf for hm in hms])
    lv = (C.c_void_p * n)(*[hm.leaves_ptr for hm in hms])
    t = Tuning()
    best, dflt = C.c_double(), C.c_double()
    _check(hmv_lib().hmv_autotune_multi(n, nd, nlf, lv, strides.pop(), PREC[precision], int(reps),
                                        C.byref(t), C.byref(best), C.byref(dflt)),
           "hmv_autotune_multi")
    return t.as_dict(), best.value, dflt.value


def default_tuning(precision="f64"):
    """The repack tunables hmv_arena_build uses for a precision."""
    t = Tuning()
    _check(hmv_lib().hmv_default_tuning(PREC[precision], C.byref(t)), "hmv_default_tuning")
    return t.as_dict()


def autotune(hm, precision="f64", reps=20):
    """hmv_autotune: returns (best tuning dict, ms per apply, ms per apply with defaults)."""
    t = Tuning()
    best, dflt = C.c_double(), C.c_double()
    _check(hmv_lib().hmv_autotune(hm.nd, hm.nlf, hm.leaves_ptr, hm.stride, PREC[precision], int(reps),
                                  C.byref(t), C.byref(best), C.byref(dflt)), "hmv_autotune")
    return t.as_dict(), best.value, dflt.value


def bicgstab_ir(a64, alow, b, u0=None, eps=1e-8, inner_eps=1e-4, max_outer=30, max_inner=1000,
                check_every=8):
    """Mixed-precision refinement (FP64 residual, low-precision inner BiCGSTAB;
    the inner solver reads its residual every `check_every` iterations)."""
    b = np.ascontiguousarray(b, np.float64)
    u0 = np.zeros_like(b) if u0 is None else np.ascontiguousarray(u0, np.float64)
    db, du = DeviceBuffer.from_array(b), DeviceBuffer.from_array(u0)
    hist = np.full(max_outer + 1, np.nan)
    info = SolveInfo()
    _check(hmv_lib().hmv_bicgstab_ir(a64.handle, alow.handle, db.ptr, du.ptr, eps, inner_eps,
                                     max_outer, max_inner, int(check_every), _dptr(hist),
                                     C.byref(info)),
           "hmv_bicgstab_ir")
    u = du.download()
    return u, info.as_dict(), hist[: info.outer + 1]


def gcr_ir(a64, alow, b, u0=None, eps=1e-8, inner_eps=5e-3, m=32, max_outer=30, max_inner=1000,
           check_every=8):
    """Mixed-precision refinement with GCR(m) inner solves on the
    low-precision arena (hmv_gcr_ir).  Defaults measured best at 327,680
    unknowns with BF16 arenas (inner 5e-3, m = 32).  Returns (u, info, outer
    residuals)."""
    b = np.ascontiguousarray(b, np.float64)
    u0 = np.zeros_like(b) if u0 is None else np.ascontiguousarray(u0, np.float64)
    db, du = DeviceBuffer.from_array(b), DeviceBuffer.from_array(u0)
    hist = np.full(max_outer + 1, np.nan)
    info = SolveInfo()
    try:
        _check(hmv_lib().hmv_gcr_ir(a64.handle, alow.handle, db.ptr, du.ptr, float(eps),
                                    float(inner_eps), int(m), int(max_outer), int(max_inner),
                                    int(check_every), _dptr(hist), C.byref(info)), "hmv_gcr_ir")
        u = du.download()
    finally:
        db.free()
        du.free()
    return u, info.as_dict(), hist[: info.outer + 1]


def bicgstab_ru(a64, alow, b, u0=None, eps=1e-8, maxit=3000, check_every=4, delta=0.1):
    """Mixed-precision BiCGSTAB with reliable updates (hmv_bicgstab_ru).
    Returns (u, info); info["outer"] = residual replacements."""
    b = np.ascontiguousarray(b, np.float64)
    u0 = np.zeros_like(b) if u0 is None else np.ascontiguousarray(u0, np.float64)
    db, du = DeviceBuffer.from_array(b), DeviceBuffer.from_array(u0)
    info = SolveInfo()
    _check(hmv_lib().hmv_bicgstab_ru(a64.handle, alow.handle, db.ptr, du.ptr, float(eps), int(maxit),
                                     int(check_every), float(delta), C.byref(info)), "hmv_bicgstab_ru")
    return du.download(), info.as_dict()


# ---------------------------------------------------------------------------
# reference-named entry points
# ---------------------------------------------------------------------------
def make_leafmtxp(hm, layout="magma", fill=None):
    """st_HACApK_leafmtxp view of an HMatrix for the reference-named calls, in
    the layout of a HAVE_MAGMA build ("magma") or of a build without it
    ("plain"); fill (magma layout): a byte value written over the whole MAGMA
    block first, standing for whatever the caller's build keeps there."""
    t = hm.leaf_table()
    lp = RefLeafmtxp() if layout == "magma" else RefLeafmtxpPlain()
    if fill is not None:
        assert layout == "magma"
        C.memset(C.addressof(lp), fill, C.sizeof(lp))
    lp.nd = hm.nd
    lp.nlf = hm.nlf
    lp.nlfkt = int((t[:, 0] == 1).sum()) if len(t) else 0
    lp.ktmax = int(t[t[:, 0] == 1, 1].max()) if lp.nlfkt else 0
    lp.st_lf_stride = hm.stride
    lp.st_lf = hm.leaves_ptr
    lp._owner = hm
    return lp


def set_ref_layout(layout):
    """hmv_ref_set_layout: the st_HACApK_leafmtxp layout the reference-named
    entry points read ("magma" or "plain"); returns the previous one."""
    prev = hmv_lib().hmv_ref_set_layout(REF_LAYOUT[layout])
    if prev < 0:
        _check(prev, "hmv_ref_set_layout")
    return {v: k for k, v in REF_LAYOUT.items()}[prev]


def adot_body_lfcpy_batch_sorted(nd, leafmtxp):
    """c_hacapk_adot_body_lfcpy_batch_sorted_ (HACApK_BATCH_SORT.c:86)."""
    hmv_lib().c_hacapk_adot_body_lfcpy_batch_sorted_(C.byref(C.c_int(nd)), C.byref(leafmtxp))


def adot_body_lfmtx_batch(zau, leafmtxp, zu):
    """c_hacapk_adot_body_lfmtx_batch_ (HACApK_BATCH.c:257): zau += H zu (in place)."""
    assert zau.dtype == np.float64 and zau.flags.c_contiguous
    zu = np.ascontiguousarray(zu, np.float64)
    tb, ts, tc = C.c_double(0), C.c_double(0), C.c_double(0)
    hmv_lib().c_hacapk_adot_body_lfmtx_batch_(_dptr(zau), C.byref(leafmtxp), _dptr(zu), None,
                                              C.byref(tb), C.byref(ts), C.byref(tc))
    return tb.value


def adot_body_lfmtx_host(zau, leafmtxp, zu, variant="lfmtx"):
    """c_hacapk_adot_body_lfmtx_ (HACApK_MAGMA.c:5) or, with variant
    "lfmtx_gpu", c_hacapk_adot_body_lfmtx_gpu_ (HACApK_MAGMA.c:79): zau += H zu
    (in place) through the arena bound to leafmtxp (built on first use)."""
    assert variant in ("lfmtx", "lfmtx_gpu")
    assert zau.dtype == np.float64 and zau.flags.c_contiguous
    zu = np.ascontiguousarray(zu, np.float64)
    getattr(hmv_lib(), "c_hacapk_adot_body_%s_" % variant)(_dptr(zau), C.byref(leafmtxp),
                                                           _dptr(zu), None)


def adot_body_lfdel_batch(leafmtxp):
    """c_hacapk_adot_body_lfdel_batch_ (HACApK_BATCH.c:812)."""
    hmv_lib().c_hacapk_adot_body_lfdel_batch_(C.byref(leafmtxp))


def _ref_solver(fn, leafmtxp, u, b, param, nrank=1):
    """Calls a reference-named solver with an lcontrol whose lpmd array (the
    communicator, ranks, this rank, ...: m_HACApK_use.f90 reads lpmd(1..3))
    follows the param array, lpmd_offset bytes past it."""
    assert u.dtype == np.float64 and u.flags.c_contiguous
    b = np.ascontiguousarray(b, np.float64)
    param = np.asarray(param, np.float64)
    buf = np.zeros(len(param) + 8)
    buf[:len(param)] = param
    lpmd = buf[len(param):].view(np.int32)
    lpmd[1] = nrank
    ctl = RefLcontrol()
    ctl.param = _dptr(buf)
    ctl.lpmd_offset = 8 * len(param)
    nd, nstp, lrtrn = C.c_int(len(u)), C.c_int(0), C.c_int(0)
    fn(C.byref(leafmtxp), C.byref(ctl), _dptr(u), _dptr(b), _dptr(buf), C.byref(nd), C.byref(nstp),
       C.byref(lrtrn))
    return nstp.value, lrtrn.value


def bicgstab_cax_lfmtx_gpu(leafmtxp, u, b, param):
    """c_hacapk_bicgstab_cax_lfmtx_gpu_ (HACApK_SOLVER.c:421): u updated in place.
    param is the 100-slot Fortran param array seen 0-based (param[82] max
    steps, param[90] tolerance).  Returns (nstp, lrtrn)."""
    return _ref_solver(hmv_lib().c_hacapk_bicgstab_cax_lfmtx_gpu_, leafmtxp, u, b, param)


def bicgstab_cax_lfmtx_flat(leafmtxp, u, b, param):
    """c_hacapk_bicgstab_cax_lfmtx_flat_ (HACApK_SOLVER.c:254): as
    bicgstab_cax_lfmtx_gpu."""
    return _ref_solver(hmv_lib().c_hacapk_bicgstab_cax_lfmtx_flat_, leafmtxp, u, b, param)


def bicgstab_cax_lfmtx_mgpu1(leafmtxp, u, b, param, nrank=1):
    """c_hacapk_bicgstab_cax_lfmtx_mgpu1_ (HACApK_SOLVER.c:815) with lpmd(2) =
    nrank; more than one rank returns lrtrn = HMV_ERR_ARG."""
    return _ref_solver(hmv_lib().c_hacapk_bicgstab_cax_lfmtx_mgpu1_, leafmtxp, u, b, param, nrank)


def adot_body_lfcpy_gpu(nd, leafmtxp):
    """c_hacapk_adot_body_lfcpy_gpu_ (HACApK_MAGMA.c:231): as lfcpy_batch_sorted."""
    hmv_lib().c_hacapk_adot_body_lfcpy_gpu_(C.byref(C.c_int(nd)), C.byref(leafmtxp))


def adot_body_lfdel_gpu(leafmtxp):
    """c_hacapk_adot_body_lfdel_gpu_ (HACApK_MAGMA.c:369): as lfdel_batch."""
    hmv_lib().c_hacapk_adot_body_lfdel_gpu_(C.byref(leafmtxp))


def bicgstab_cax_lfmtx_pipe(leafmtxp, u, b, param, variant=1):
    """c_hacapk_bicgstab_cax_lfmtx_pipe_ (HACApK_SOLVER_PIPE.c:144) or, with
    variant=2, c_hacapk_bicgstab_cax_lfmtx_pipe2_ (:640, what the PIPE_BICG
    driver calls, m_HACApK_use.f90:418): the pipelined BiCGSTAB; u receives
    the solution.  Returns (nstp, lrtrn)."""
    L = hmv_lib()
    fn = L.c_hacapk_bicgstab_cax_lfmtx_pipe2_ if variant == 2 else L.c_hacapk_bicgstab_cax_lfmtx_pipe_
    return _ref_solver(fn, leafmtxp, u, b, param)


def solve(hm, rhs, sol0=None, ztol=1e-8, mstep=500, arena=None, solver="bicgstab", mreset=8):
    """HACApK_solve (m_HACApK_use.f90:190-262) on the engine: permute with lod,
    solve on the device, un-permute.  solver selects as param(85) does
    (m_HACApK_use.f90:235-242): "bicgstab" (1) or "gcrm" (2), the latter with
    mreset = param(87) directions.  Returns (sol, info, hist)."""
    if solver not in ("bicgstab", "gcrm"):
        raise ValueError("solver must be 'bicgstab' or 'gcrm'")
    own = arena is None
    a = arena or Arena(hm, "f64")
    b = hm.to_cluster_order(rhs)
    u0 = None if sol0 is None else hm.to_cluster_order(sol0)
    if solver == "gcrm":
        u, info, hist = a.gcrm(b, u0, eps=ztol, maxit=mstep, mreset=mreset)
    else:
        u, info, hist = a.bicgstab(b, u0, eps=ztol, maxit=mstep)
    if own:
        a.free()
    return hm.to_original_order(u), info, hist
