"""ctypes loader of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, as the checker or as the timed CPU baseline; the engine never
does.  See hmv_oracle.c for the reference code each function restates and
for what pins it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_D = C.POINTER(C.c_double)
_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: %s" % path)
        L = C.CDLL(path)
        vp = C.c_void_p
        L.oracle_adot_body_lfmtx.argtypes = [_D, vp, C.c_int, C.c_int, _D, _D]
        L.oracle_adot_lfmtx_hyp.argtypes = [_D, vp, C.c_int, C.c_int, C.c_int, _D, C.c_int]
        L.oracle_setcutthread.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.oracle_bicgstab.argtypes = [vp, C.c_int, C.c_int, C.c_int, _D, _D, C.c_int, C.c_double,
                                      C.c_int, _D]
        L.oracle_gcrm.argtypes = [vp, C.c_int, C.c_int, C.c_int, _D, _D, C.c_int, C.c_int,
                                  C.c_double, C.c_int, _D]
        L.oracle_bicgstab_pipe.argtypes = [vp, C.c_int, C.c_int, C.c_int, _D, _D, C.c_int,
                                           C.c_double, C.c_int, _D]
        L.oracle_dotp.argtypes = [C.c_int, _D, _D]
        L.oracle_dotp.restype = C.c_double
        L.oracle_dense.argtypes = [vp, C.c_int, C.c_int, C.c_int, _D]
        L.oracle_max_threads.restype = C.c_int
        _lib = L
    return _lib


def ref_lib():
    """oracle/_ref/libhacapk_ref.so (reference Fortran sources), or None."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libhacapk_ref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_unrm_d.argtypes = [C.c_int, _D]
        L.ref_unrm_d.restype = C.c_double
        L.ref_maxabsvalloc_d.argtypes = [C.c_int, _D, _D, C.POINTER(C.c_int)]
        L.ref_minabsvalloc_d.argtypes = [C.c_int, _D, _D, C.POINTER(C.c_int)]
        L.ref_face_integral.argtypes = [_D, _D, _D, C.c_double, C.c_double, C.c_double]
        L.ref_face_integral.restype = C.c_double
        if hasattr(L, "ref_acaplus"):  # HACApK_acaplus (ref_aca_bind.f90)
            _I = C.POINTER(C.c_int)
            L.ref_acaplus.argtypes = [C.c_int, _D, C.c_int, _I, _I, C.c_int, C.c_int, C.c_int,
                                      C.c_int, _D, _D, _D]
            L.ref_acaplus.restype = C.c_int
        _ref = L
    return _ref


def _d(a):
    return a.ctypes.data_as(_D)


def adot_body_lfmtx(hm, zu, zau=None):
    """HACApK_MAGMA.c:5-57: zau += H zu (zau zero if not given)."""
    zu = np.ascontiguousarray(zu, np.float64)
    zau = np.zeros(hm.nd) if zau is None else np.ascontiguousarray(zau, np.float64).copy()
    zbu = np.zeros(max(1, _ktmax(hm)))
    lib().oracle_adot_body_lfmtx(_d(zau), hm.leaves_ptr, hm.nlf, hm.stride, _d(zu), _d(zbu))
    return zau


def adot_lfmtx_hyp(hm, zu, nthr=1):
    """m_HACApK_solve.f90:127-161,203-253: returns H zu (threaded, fixed order)."""
    zu = np.ascontiguousarray(zu, np.float64)
    zau = np.empty(hm.nd)
    rc = lib().oracle_adot_lfmtx_hyp(_d(zau), hm.leaves_ptr, hm.nlf, hm.stride, hm.nd, _d(zu), nthr)
    if rc:
        raise RuntimeError("oracle_adot_lfmtx_hyp failed")
    return zau


def setcutthread(hm, nthr):
    out = (C.c_int * (nthr + 1))()
    lib().oracle_setcutthread(hm.leaves_ptr, hm.nlf, hm.stride, nthr, out)
    return np.array(out[:], np.int64)


def bicgstab(hm, b, u0=None, mstep=500, eps=1e-8, nthr=1):
    """m_HACApK_solve.f90:342-449.  Returns (u, nstp, hist)."""
    b = np.ascontiguousarray(b, np.float64)
    u = np.zeros(hm.nd) if u0 is None else np.ascontiguousarray(u0, np.float64).copy()
    hist = np.full(mstep + 1, np.nan)
    nstp = lib().oracle_bicgstab(hm.leaves_ptr, hm.nlf, hm.stride, hm.nd, _d(u), _d(b), mstep, eps,
                                 nthr, _d(hist))
    return u, nstp, hist[: nstp + 1]


def gcrm(hm, b, u0=None, mstep=500, mreset=8, eps=1e-8, nthr=1):
    """m_HACApK_solve.f90:506-571.  Returns (u, nstp, hist)."""
    b = np.ascontiguousarray(b, np.float64)
    u = np.zeros(hm.nd) if u0 is None else np.ascontiguousarray(u0, np.float64).copy()
    hist = np.full(mstep + 1, np.nan)
    nstp = lib().oracle_gcrm(hm.leaves_ptr, hm.nlf, hm.stride, hm.nd, _d(u), _d(b), mstep, mreset,
                             eps, nthr, _d(hist))
    return u, nstp, hist[: nstp + 1]


def bicgstab_pipe(hm, b, u0=None, mstep=500, eps=1e-8, nthr=1):
    """HACApK_SOLVER_PIPE.c:144-637 (iterate returned).  Returns (u, nstp, hist)."""
    b = np.ascontiguousarray(b, np.float64)
    u = np.zeros(hm.nd) if u0 is None else np.ascontiguousarray(u0, np.float64).copy()
    hist = np.full(mstep + 1, np.nan)
    nstp = lib().oracle_bicgstab_pipe(hm.leaves_ptr, hm.nlf, hm.stride, hm.nd, _d(u), _d(b), mstep,
                                      eps, nthr, _d(hist))
    return u, nstp, hist[: nstp + 1]


def dense(hm):
    a = np.empty((hm.nd, hm.nd))
    lib().oracle_dense(hm.leaves_ptr, hm.nlf, hm.stride, hm.nd, _d(a))
    return a


def max_threads():
    return lib().oracle_max_threads()


def _ktmax(hm):
    t = hm.leaf_table()
    lr = t[t[:, 0] == 1]
    return int(lr[:, 1].max()) if len(lr) else 1
