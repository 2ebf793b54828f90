"""The C-ABI libraries load and export every function their headers declare;
without a GPU every compute entry point fails loudly (no CPU fallback)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([a-z_][a-z0-9_]*)\s*\(", src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


@pytest.mark.parametrize("header,lib", [("hmv.h", "libhmv.so"), ("hmvgen.h", "libhmvgen.so")])
def test_every_declared_symbol_is_exported(H, header, lib):
    names = _declared(header)
    assert len(names) > 8
    L = C.CDLL(os.path.join(ROOT, "hacapk-magma_amd", "lib", lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_named_symbols_present(H):
    L = H.hmv_lib()
    for n in ("c_hacapk_adot_body_lfcpy_batch_sorted_", "c_hacapk_adot_body_lfmtx_batch_",
              "c_hacapk_adot_body_lfdel_batch_", "c_hacapk_bicgstab_cax_lfmtx_gpu_"):
        assert hasattr(L, n)


def _have_gpu(H):
    n = C.c_int(0)
    return os.path.exists("/dev/kfd") and H.hmv_lib().hmv_init(0) == 0


def test_no_cpu_fallback(H, sphere3):
    if _have_gpu(H):
        pytest.skip("a GPU is present")
    with pytest.raises(H.HmvError, match="no HIP device"):
        H.Arena(sphere3, "f64")
    with pytest.raises(H.HmvError):
        H.DeviceBuffer(64)
    # the reference-named host applies compute nothing on the CPU: zau is
    # left as it was and the error names the missing device
    lp = H.make_leafmtxp(sphere3)
    for variant in ("lfmtx", "lfmtx_gpu"):
        zau = np.full(sphere3.nd, 7.0)
        H.adot_body_lfmtx_host(zau, lp, np.ones(sphere3.nd), variant)
        assert np.all(zau == 7.0)
        assert b"no HIP device" in H.hmv_lib().hmv_last_error()


def test_new_reference_named_symbols_present(H):
    L = H.hmv_lib()
    for n in ("c_hacapk_bicgstab_cax_lfmtx_pipe_", "c_hacapk_bicgstab_cax_lfmtx_pipe2_",
              "c_hacapk_adot_body_lfmtx_", "c_hacapk_adot_body_lfmtx_gpu_",
              "c_hacapk_adot_body_lfcpy_gpu_", "c_hacapk_adot_body_lfdel_gpu_",
              "c_hacapk_bicgstab_cax_lfmtx_flat_", "c_hacapk_bicgstab_cax_lfmtx_mgpu1_"):
        assert hasattr(L, n)


def test_argument_checks_before_any_device_call(H):
    """Invalid arguments of the solver, construction and file entry points
    fail with HMV_ERR_ARG and a message, before touching a device."""
    L = H.hmv_lib()
    info = H.SolveInfo()
    assert L.hmv_gcrm(None, None, None, 1e-8, 10, 8, 1, None, C.byref(info)) == -1
    assert b"hmv_gcrm" in L.hmv_last_error()
    assert L.hmv_bicgstab_pipe(None, None, None, 1e-8, 10, 1, None, C.byref(info)) == -1
    assert L.hmv_gcr_ir(None, None, None, None, 1e-8, 1e-2, 8, 10, 10, 1, None, C.byref(info)) == -1
    d = C.c_void_p()
    cp = H._ConstructParams(kernel=7, kmax=10, kmin=1, aca_eps=1e-4, aca_eps_factor=1.0)
    nodes = (C.c_double * 9)()
    faces = (C.c_int * 3)(0, 1, 2)
    lod = (C.c_int * 1)(1)
    rec = (H.LeafRec * 1)()
    assert L.hmv_construct(3, nodes, 1, faces, nodes, nodes, lod, 1, C.addressof(rec),
                           C.sizeof(H.LeafRec), C.byref(cp), C.byref(d)) == -1
    cp.kernel = 1
    faces[2] = 5  # missing node
    assert L.hmv_construct(3, nodes, 1, faces, nodes, nodes, lod, 1, C.addressof(rec),
                           C.sizeof(H.LeafRec), C.byref(cp), C.byref(d)) == -1
    assert b"missing node" in L.hmv_last_error()
    faces[2] = 2
    rec[0].ltmtx, rec[0].nstrtl, rec[0].ndl, rec[0].nstrtt, rec[0].ndt = 2, 1, 2, 1, 1  # 2 rows of 1
    assert L.hmv_construct(3, nodes, 1, faces, nodes, nodes, lod, 1, C.addressof(rec),
                           C.sizeof(H.LeafRec), C.byref(cp), C.byref(d)) == -1
    assert b"outside the matrix" in L.hmv_last_error()
    p = C.c_void_p()
    assert L.hmv_plan_load(b"/nonexistent/arena.hmv", C.byref(p)) == -1
    assert b"cannot open" in L.hmv_last_error()
    assert L.hmv_arena_load(b"/nonexistent/arena.hmv", C.byref(p)) == -1


def test_calibration_argument_checks(H):
    """hmv_stream_test rejects unknown kinds and malformed MFMA-rate sizes
    before any device call."""
    L = H.hmv_lib()
    ms = C.c_double()
    buf = (C.c_char * 64)()
    assert L.hmv_stream_test(3, buf, buf, 64, 1, C.byref(ms)) == -1
    assert L.hmv_stream_test(2, None, buf, 20, 1, C.byref(ms)) == -1  # not a multiple of 8
    assert L.hmv_stream_test(2, None, buf, (1 << 27) + 8, 1, C.byref(ms)) == -1
    assert L.hmv_stream_test(2, None, None, 64, 1, C.byref(ms)) == -1
    assert b"hmv_stream_test" in L.hmv_last_error()
