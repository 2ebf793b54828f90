"""Multi-right-hand-side H-matmat (vector units for few right-hand sides,
MFMA for more) and block BiCGSTAB through the C ABI.

Bar: every column of the H-matmat equals the single-vector apply of that
column (FP64: <= 1e-13 relative, FP32 / BF16 arenas: <= 1e-5 relative, both
accumulate in FP32 but in different orders) and the CPU oracle (FP64 <= 1e-12);
block BiCGSTAB columns reach the true residual bar and agree with the
single-vector solver's solutions.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("k", [1, 3, 8, 16, 32, 40, 64, 72, 128])
def test_matmat_f64_columns(gpu, O, two_spheres4, k):
    hm = two_spheres4
    a = gpu.Arena(hm, "f64")
    X = np.stack([gpu.uniform(100 + j, hm.nd) for j in range(k)], axis=1)
    Y = a.matmat(X)
    worst = 0.0
    for j in sorted({0, k // 2, k - 1}):
        y1 = a.matvec(X[:, j])
        worst = max(worst, _rel(Y[:, j], y1))
        assert _rel(Y[:, j], O.adot_body_lfmtx(hm, X[:, j])) <= 1e-12
    print("k", k, "max rel diff vs single apply %.2e" % worst, flush=True)
    assert worst <= 1e-13
    a.free()


@pytest.mark.parametrize("prec,k", [("f32", 8), ("f32", 32), ("bf16", 16), ("bf16_all", 8),
                                    ("f32", 4), ("bf16", 2), ("bf16_all", 3), ("f32", 128),
                                    ("bf16_all", 72)])
def test_matmat_low_precision(gpu, O, two_spheres4, prec, k):
    """Narrow arenas: k <= 4 runs on the vector units (hmv_kstream.hip),
    larger k on FP32 MFMAs (two 16-row tiles per wave above k = 64)."""
    hm = two_spheres4
    a = gpu.Arena(hm, prec)
    X = np.stack([gpu.uniform(200 + j, hm.nd) for j in range(k)], axis=1)
    Y = a.matmat(X)
    for j in (0, k - 1):
        assert _rel(Y[:, j], a.matvec(X[:, j])) <= 1e-5
        assert _rel(Y[:, j], O.adot_body_lfmtx(hm, X[:, j])) <= {"f32": 1e-5, "bf16": 2e-2, "bf16_all": 2e-2}[prec]
    a.free()


TOL = {"f64": 1e-12, "f32": 1e-5, "bf16": 2e-2, "bf16_all": 2e-2}


def _paths(prec, k):
    """Every H-matmat path choice of k's classes: (vector, tiles per unit)."""
    vec_max = 8 if prec == "f64" else 4
    kt = (k + 15) // 16
    tiles = [4, 2] if kt == 1 else [2, 1] if kt <= 4 else ([1, 2] if prec == "f64" else [2, 1])
    out = [(1, tiles[0])] if k <= vec_max else []
    return out + [(0, t) for t in tiles]


@pytest.mark.parametrize("prec,k", [("f64", 3), ("f64", 8), ("f64", 16), ("f64", 40), ("f64", 128),
                                    ("f32", 4), ("f32", 72), ("bf16_all", 128)])
def test_matmat_every_path(gpu, O, two_spheres4, prec, k):
    """Each path the block tuning can select (vector units, MFMA with each
    instantiated tiles-per-unit width) gives the oracle's product; the
    arena reports the path it was set to."""
    hm = two_spheres4
    a = gpu.Arena(hm, prec)
    X = np.stack([gpu.uniform(300 + j, hm.nd) for j in range(k)], axis=1)
    for vec, tiles in _paths(prec, k):
        a.set_block_tuning(k, vec, tiles)
        assert a.block_tuning(k) == {"vector": vec, "tiles_per_unit": tiles}
        Y = a.matmat(X)
        for j in (0, k - 1):
            assert _rel(Y[:, j], O.adot_body_lfmtx(hm, X[:, j])) <= TOL[prec], (vec, tiles, j)
    a.free()


@pytest.mark.parametrize("prec,k", [("f64", 7), ("f64", 8), ("f64", 64), ("f64", 128),
                                    ("f32", 128)])
def test_matmat_inputs_at_any_alignment(gpu, sphere4, prec, k):
    """Inputs that start 8 bytes past a 16-byte boundary (a caller's x may
    start anywhere) take the paths that read inputs one value at a time
    instead of in 16-byte pairs; the product is bit-equal to the aligned
    one (the same values reach the same FMAs), and odd k uses the unpaired
    staging too."""
    a = gpu.Arena(sphere4, prec)
    X = np.stack([gpu.uniform(700 + j, sphere4.nd) for j in range(k)], axis=1)
    Y_al = a.matmat(X)
    buf = gpu.DeviceBuffer.from_array(np.concatenate([[0.0], X.reshape(-1)]))
    dy = gpu.DeviceBuffer.from_array(np.zeros(X.size))
    a.apply_block(buf.ptr + 8, dy.ptr, k)
    assert np.array_equal(dy.download().reshape(sphere4.nd, k), Y_al)
    buf.free()
    dy.free()
    a.free()


@pytest.mark.parametrize("prec,k", [("f64", 8), ("f64", 32), ("f32", 4)])
def test_block_autotune(gpu, O, sphere4, prec, k):
    """hmv_autotune_block sets the fastest path of k's classes: no slower
    than the defaults, reported back by the arena, and its product matches
    the oracle."""
    a = gpu.Arena(sphere4, prec)
    dflt = a.block_tuning(k)
    best, best_ms, dflt_ms = a.autotune_block(k, reps=3)
    print(prec, k, "default", dflt, "best", best, "%.4f ms (default %.4f ms)" % (best_ms, dflt_ms),
          flush=True)
    assert 0 < best_ms <= dflt_ms
    assert a.block_tuning(k) == best
    X = np.stack([gpu.uniform(400 + j, sphere4.nd) for j in range(k)], axis=1)
    Y = a.matmat(X)
    assert _rel(Y[:, k - 1], O.adot_body_lfmtx(sphere4, X[:, k - 1])) <= TOL[prec]
    a.free()


def test_block_tuning_rejects(gpu, sphere3):
    a, n = gpu.Arena(sphere3, "f64"), gpu.Arena(sphere3, "f32")
    with pytest.raises(gpu.HmvError, match="vector path"):
        a.set_block_tuning(16, 1)
    with pytest.raises(gpu.HmvError, match="vector path"):
        n.set_block_tuning(8, 1)
    with pytest.raises(gpu.HmvError, match="tiles_per_unit"):
        a.set_block_tuning(32, 0, 4)
    with pytest.raises(gpu.HmvError):
        a.block_tuning(129)
    a.free()
    n.free()


def test_matmat_beta_accumulates(gpu, sphere4):
    a = gpu.Arena(sphere4, "f64")
    k = 8
    X = np.stack([gpu.uniform(300 + j, sphere4.nd) for j in range(k)], axis=1)
    Y0 = np.stack([gpu.uniform(400 + j, sphere4.nd) for j in range(k)], axis=1)
    dx, dy = gpu.DeviceBuffer.from_array(X.reshape(-1)), gpu.DeviceBuffer.from_array(Y0.reshape(-1))
    a.apply_block(dx.ptr, dy.ptr, k, beta=1.0)
    gpu.hmv_lib().hmv_device_synchronize()
    Y = dy.download().reshape(sphere4.nd, k)
    assert _rel(Y, Y0 + a.matmat(X)) <= 1e-13
    a.free()


def test_block_bicgstab_matches_single(gpu, sphere4):
    hm = sphere4
    a = gpu.Arena(hm, "f64")
    k = 8
    B = np.stack([gpu.uniform(500 + j, hm.nd) for j in range(k)], axis=1)
    B[:, 0] = 1.0
    U, cols, info = a.block_bicgstab(B, eps=1e-8, maxit=500, check_every=4)
    print("block", info, [c["nstp"] for c in cols], flush=True)
    assert info["status_name"] == "converged"
    for j in range(k):
        assert cols[j]["status_name"] == "converged" and cols[j]["true_relres"] < 1e-8 * 1.5
        u1, i1, _ = a.bicgstab(B[:, j], eps=1e-8, maxit=500)
        assert abs(cols[j]["nstp"] - i1["nstp"]) <= 3
        assert _rel(U[:, j], u1) < 1e-6
    a.free()


def test_block_bicgstab_stops_columns_independently(gpu, sphere4):
    """A zero right-hand side stops its column at setup (non-finite relative
    residual); the other columns still converge."""
    hm = sphere4
    a = gpu.Arena(hm, "f64")
    B = np.stack([gpu.uniform(600 + j, hm.nd) for j in range(4)], axis=1)
    B[:, 2] = 0.0
    U, cols, info = a.block_bicgstab(B, eps=1e-8, maxit=500)
    assert cols[2]["status_name"] == "nonfinite" and cols[2]["nstp"] == 0
    assert np.all(U[:, 2] == 0.0)
    for j in (0, 1, 3):
        assert cols[j]["status_name"] == "converged"
    a.free()


def test_block_rejects_bad_k(gpu, sphere3):
    a = gpu.Arena(sphere3, "f64")
    with pytest.raises(gpu.HmvError):
        a.block_bicgstab(np.ones((sphere3.nd, 3)))
    with pytest.raises(gpu.HmvError):
        a.matmat(np.ones((sphere3.nd, 129)))
    a.free()


@pytest.mark.parametrize("prec,inner", [("f32", 1e-4), ("bf16", 1e-2)])
def test_block_refinement(gpu, O, sphere4, prec, inner):
    """hmv_block_bicgstab_ir: FP64 block residuals on the FP64 arena, inner
    block BiCGSTAB on the low-precision arena; every column reaches a true
    1e-8 (recomputed by the oracle) and the single FP64 solve's solution
    within cond * 2e-8 (cond = 2 sqrt(N) + 1, the sphere single-layer
    operator's eigenvalue spread); a zero column stops at once."""
    hm = sphere4
    a64, alow = gpu.Arena(hm, "f64"), gpu.Arena(hm, prec)
    k = 8
    B = np.stack([gpu.uniform(700 + j, hm.nd) for j in range(k)], axis=1)
    B[:, 1] = 1.0
    B[:, 5] = 0.0
    U, cols, info = gpu.block_bicgstab_ir(a64, alow, B, eps=1e-8, inner_eps=inner, max_outer=30,
                                          max_inner=500)
    print(prec, info, [(c["nstp"], "%.1e" % c["true_relres"]) for c in cols], flush=True)
    assert info["outer"] >= 2
    assert cols[5]["status_name"] == "nonfinite" and cols[5]["nstp"] == 0
    assert np.all(U[:, 5] == 0.0)
    bound = (2.0 * np.sqrt(hm.nd) + 1.0) * 2e-8
    for j in range(k):
        if j == 5:
            continue
        tr = _rel(O.adot_body_lfmtx(hm, U[:, j]), B[:, j])
        assert cols[j]["status_name"] == "converged" and tr < 1e-8
        u1, i1, _ = a64.bicgstab(B[:, j], eps=1e-8, maxit=500)
        assert _rel(U[:, j], u1) < bound
    a64.free()
    alow.free()
