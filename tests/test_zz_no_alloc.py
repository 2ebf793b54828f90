"""hmv_apply_block_static, the H-matmat that allocates nothing: its t
buffer is the library's static device workspace.  It must give
hmv_apply_block's result bit for bit (same kernels, same summation order)
and the oracle's columns to the FP64 bar (1e-12 relative), on two calls in
a row with the second one smaller (the workspace keeps the first call's
contents), and refuse a product whose t buffer exceeds the workspace."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def _static_matmat(H, a, X):
    X = np.ascontiguousarray(X, np.float64)
    k = X.shape[1]
    dx = H.DeviceBuffer.from_array(X.reshape(-1))
    dy = H.DeviceBuffer.from_array(np.full(X.size, np.nan))
    rc = H.hmv_lib().hmv_apply_block_static(a.handle, k, dx.ptr, dy.ptr, 0.0, None)
    H._check(rc, "hmv_apply_block_static")
    H.hmv_lib().hmv_device_synchronize()
    Y = dy.download().reshape(a.nd, k)
    dx.free()
    dy.free()
    return Y


def test_static_matmat_matches_block_apply(gpu, O, sphere4, sphere3):
    for hm, k, seed in ((sphere4, 16, 40), (sphere3, 3, 41)):
        a = gpu.Arena(hm, "f64")
        X = np.stack([gpu.uniform(seed + j, hm.nd) for j in range(k)], axis=1)
        print("static H-matmat nd", hm.nd, "k", k, "tparts", a.stats["tparts"], flush=True)
        Y = _static_matmat(gpu, a, X)
        Yb = a.matmat(X)
        assert np.array_equal(Y, Yb)
        for j in range(k):
            ref = O.adot_body_lfmtx(hm, X[:, j])
            assert np.linalg.norm(Y[:, j] - ref) / np.linalg.norm(ref) <= TOL64
        a.free()


def test_static_matmat_rejects_oversize(gpu):
    """A 81,920-unknown sphere at k = 128 needs more t buffer than the
    256 MiB workspace: HMV_ERR_ARG with the sizes, before any launch."""
    hm = gpu.HMatrix(*gpu.icosphere(6), gpu.baseline_params(aca_eps=1e-4))
    a = gpu.Arena(hm, "f64")
    k = 128
    need = max(1, a.stats["tparts"]) * k * 8
    assert need > 256 << 20
    dx = gpu.DeviceBuffer(8 * hm.nd * k)
    dy = gpu.DeviceBuffer(8 * hm.nd * k)
    rc = gpu.hmv_lib().hmv_apply_block_static(a.handle, k, dx.ptr, dy.ptr, 0.0, None)
    assert rc == -1
    assert b"static scratch holds" in gpu.hmv_lib().hmv_last_error()
    rc = gpu.hmv_lib().hmv_apply_block_static(a.handle, 0, dx.ptr, dy.ptr, 0.0, None)
    assert rc == -1
    dx.free()
    dy.free()
    a.free()
