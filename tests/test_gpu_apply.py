"""GPU parity of the H-matvec (libhmv.so through the C ABI) against the CPU oracle.

Bar: FP64 relative error <= 1e-12 in the 2-norm against the oracle's C-loop
apply (HACApK_MAGMA.c:5-57) on the same leaves and inputs; FP32 storage
<= 1e-5 and BF16 low-rank storage <= 2e-2 (storage rounding 2^-24 / 2^-9 per
entry, summed over long rows).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("case", ["sphere3", "sphere4", "two_spheres4", "ref_mesh_sample"])
def test_f64_apply_matches_oracle(gpu, O, case, request):
    hm = request.getfixturevalue(case)
    a = gpu.Arena(hm, "f64")
    for seed in (0, 1):
        x = gpu.uniform(seed, hm.nd)
        y = a.matvec(x)
        y_ref = O.adot_body_lfmtx(hm, x)
        print(case, "nd", hm.nd, "seed", seed, "rel err %.3e" % _rel(y, y_ref), flush=True)
        assert _rel(y, y_ref) <= TOL64
    a.free()


def test_f64_device_apply_overwrite_and_accumulate(gpu, O, sphere4):
    hm = sphere4
    a = gpu.Arena(hm, "f64")
    x = gpu.uniform(3, hm.nd)
    y0 = gpu.uniform(4, hm.nd)
    dx, dy = gpu.DeviceBuffer.from_array(x), gpu.DeviceBuffer.from_array(np.full(hm.nd, np.nan))
    a.apply(dx.ptr, dy.ptr, 0.0)          # beta=0 ignores NaNs in y
    gpu.hmv_lib().hmv_device_synchronize()
    y_ref = O.adot_body_lfmtx(hm, x)
    assert _rel(dy.download(), y_ref) <= TOL64
    dy.upload(y0)
    a.apply(dx.ptr, dy.ptr, 1.0)          # y += H x
    gpu.hmv_lib().hmv_device_synchronize()
    assert _rel(dy.download(), y0 + y_ref) <= TOL64
    a.free()


def test_apply_is_deterministic(gpu, sphere4):
    a = gpu.Arena(sphere4, "f64")
    x = gpu.uniform(5, sphere4.nd)
    y1 = a.matvec(x)
    y2 = a.matvec(x)
    assert np.array_equal(y1, y2)
    # in place into a caller-owned array, zau += H zu as the reference's entry point
    zau = np.full(sphere4.nd, 7.0)
    a.apply_host(x, zau)
    assert np.array_equal(zau, 7.0 + y1)
    with pytest.raises(ValueError):
        a.apply_host(x, np.zeros(sphere4.nd + 1))
    a.free()


@pytest.mark.parametrize("prec,tol", [("f32", 1e-5), ("bf16", 2e-2), ("bf16_all", 2e-2)])
def test_low_precision_apply(gpu, O, sphere4, prec, tol):
    hm = sphere4
    a = gpu.Arena(hm, prec)
    x = gpu.uniform(6, hm.nd)
    y = a.matvec(x)
    y_ref = O.adot_body_lfmtx(hm, x)
    err = _rel(y, y_ref)
    print(prec, "rel err vs FP64 oracle %.3e" % err, flush=True)
    assert err <= tol
    a.free()


@pytest.mark.parametrize("prec,tuning", [
    ("f64", {"p1_small_threads": 64, "p2_small_threads": 64}),
    ("f64", {"p1_small_threads": 128, "p2_small_threads": 128, "p1_target_bytes": 16384,
             "p1_row_tile": 32, "p2_max_rows": 16}),
    ("f32", {"p1_small_threads": -1, "p2_small_threads": 128}),
    ("bf16", {"p1_small_threads": 64, "p2_small_threads": -1}),
])
def test_tuned_arenas_match_oracle(gpu, O, two_spheres4, prec, tuning):
    """Arenas built with other tunings (small workgroups for short blocks,
    other tile sizes) give the same product to the precision's bar."""
    hm = two_spheres4
    a = gpu.Arena(hm, prec, tuning=tuning)
    got = a.tuning
    for k, v in tuning.items():
        assert got[k] == v
    x = gpu.uniform(8, hm.nd)
    err = _rel(a.matvec(x), O.adot_body_lfmtx(hm, x))
    print(prec, tuning, "rel err %.3e" % err, flush=True)
    assert err <= {"f64": TOL64, "f32": 1e-5, "bf16": 2e-2}[prec]
    a.free()


def test_autotune(gpu, O, sphere4):
    """The autotuner returns a valid tuning no slower than the defaults (it
    keeps the defaults unless a candidate is faster), and an arena built
    with it matches the oracle."""
    best, best_ms, dflt_ms = gpu.autotune(sphere4, "f64", reps=5)
    print("autotune", best, "%.4f ms (defaults %.4f ms)" % (best_ms, dflt_ms), flush=True)
    assert 0 < best_ms <= dflt_ms
    a = gpu.Arena(sphere4, "f64", tuning=best)
    x = gpu.uniform(9, sphere4.nd)
    assert _rel(a.matvec(x), O.adot_body_lfmtx(sphere4, x)) <= TOL64
    a.free()


def test_linearity(gpu, sphere4):
    a = gpu.Arena(sphere4, "f64")
    x1, x2 = gpu.uniform(10, sphere4.nd), gpu.uniform(11, sphere4.nd)
    y = a.matvec(2.0 * x1 - 3.0 * x2)
    y12 = 2.0 * a.matvec(x1) - 3.0 * a.matvec(x2)
    assert _rel(y, y12) <= 1e-13
    a.free()


def test_reference_named_batch_apply(gpu, O, sphere3):
    """c_hacapk_adot_body_lfcpy_batch_sorted_ / lfmtx_batch_ / lfdel_batch_:
    zau += H zu on host vectors, the Fortran driver's calling sequence
    (m_HACApK_solve.f90:711-712)."""
    hm = sphere3
    lp = gpu.make_leafmtxp(hm)
    gpu.adot_body_lfcpy_batch_sorted(hm.nd, lp)
    zu = np.ones(hm.nd)
    zau = np.ones(hm.nd)               # the reference's v(:)=1 before the call
    gpu.adot_body_lfmtx_batch(zau, lp, zu)
    gpu.adot_body_lfdel_batch(lp)
    ref = O.adot_body_lfmtx(hm, zu, np.ones(hm.nd))
    assert _rel(zau, ref) <= TOL64


@pytest.mark.parametrize("variant", ["lfmtx", "lfmtx_gpu"])
def test_reference_named_host_apply(gpu, O, sphere3, variant):
    """c_hacapk_adot_body_lfmtx_ / lfmtx_gpu_ without a prior lfcpy: the
    first call binds an arena, a second call reuses it (the timing loop of
    m_HACApK_solve.f90:655-664 calls it repeatedly with u(:)=1, b(:)=1), and
    the reference-named solver then finds the bound arena."""
    hm = sphere3
    lp = gpu.make_leafmtxp(hm)
    b = np.ones(hm.nd)
    ref = O.adot_body_lfmtx(hm, b, np.ones(hm.nd))
    for _ in range(2):
        u = np.ones(hm.nd)
        gpu.adot_body_lfmtx_host(u, lp, b, variant)
        assert _rel(u, ref) <= TOL64
    param = np.zeros(100)
    param[82], param[90] = 200, 1e-8
    x = np.zeros(hm.nd)
    nstp, lrtrn = gpu.bicgstab_cax_lfmtx_gpu(lp, x, b, param)
    assert lrtrn == 0 and nstp > 0
    gpu.adot_body_lfdel_batch(lp)


def test_edge_leaves(gpu, O, H):
    """Hand-made leaf lists: a rank-0 low-rank leaf, a 1x1 dense leaf, rows
    covered by no leaf (y = 0 there), a long low-rank block (> 2048 columns,
    several LDS passes) and blocks whose row stride needs padding."""
    rng = np.random.default_rng(0)
    nd = 5000
    specs = [(1, 0, 1, 10, 11, 10), (2, 0, 21, 1, 22, 1), (1, 3, 1, 2500, 2501, 2499),
             (2, 0, 2501, 7, 1, 13), (1, 5, 3001, 1000, 1, 2500), (2, 0, 4001, 999, 4001, 999)]
    bufs = []
    recs = (H.LeafRec * len(specs))()
    for i, (lt, kt, r0, ndl, c0, ndt) in enumerate(specs):
        if lt == 1:
            a1 = rng.standard_normal(ndt * kt + ndl * kt)
            recs[i].a1size = ndt * kt * 8
        else:
            a1 = rng.standard_normal(ndt * ndl)
            recs[i].a1size = ndt * ndl * 8
        a1 = np.ascontiguousarray(np.concatenate([a1, [0.0]]))
        bufs.append(a1)
        recs[i].ltmtx, recs[i].kt = lt, kt
        recs[i].nstrtl, recs[i].ndl, recs[i].nstrtt, recs[i].ndt = r0, ndl, c0, ndt
        recs[i].a1 = a1.ctypes.data_as(C.POINTER(C.c_double))
        recs[i].a2 = C.cast(C.addressof(recs[i].a1.contents) + recs[i].a1size, C.POINTER(C.c_double)) \
            if lt == 1 else None

    ptr, stride = C.addressof(recs), C.sizeof(H.LeafRec)
    a = H.Arena(None, "f64", nd=nd, nlf=len(specs), leaves_ptr=ptr, stride=stride)
    x = rng.uniform(-1, 1, nd)
    y = a.matvec(x)
    zau = np.zeros(nd)
    zbu = np.zeros(16)
    O.lib().oracle_adot_body_lfmtx(zau.ctypes.data_as(O._D), ptr, len(specs), stride,
                                   x.ctypes.data_as(O._D), zbu.ctypes.data_as(O._D))
    assert _rel(y, zau) <= TOL64
    covered = np.zeros(nd, bool)
    for lt, kt, r0, ndl, c0, ndt in specs:
        if lt == 2 or kt > 0:
            covered[r0 - 1:r0 - 1 + ndl] = True
    assert np.all(y[~covered] == 0.0)
    a.free()


@pytest.mark.parametrize("prec,tol", [("f64", TOL64), ("f32", 1e-5), ("bf16", 2e-2),
                                      ("bf16_all", 2e-2)])
def test_input_runs_and_code_fallback(gpu, O, prec, tol):
    """Blocks whose inputs come from a tabled run list (31 runs), a code table
    (101 runs, beyond the table) and single runs, through the H-matvec and the
    H-matmat (k = 3)."""
    import synthetic_leaves as SL
    hm = SL.many_runs(gpu)
    a = gpu.Arena(hm, prec)
    x = gpu.uniform(6, hm.nd)
    y = a.matvec(x)
    y_ref = O.adot_body_lfmtx(hm, x)
    print(prec, "apply rel err %.3e" % _rel(y, y_ref), flush=True)
    assert _rel(y, y_ref) <= tol
    X = np.stack([gpu.uniform(7 + j, hm.nd) for j in range(3)], axis=1)
    Y = a.matmat(X)
    for j in range(3):
        assert _rel(Y[:, j], O.adot_body_lfmtx(hm, X[:, j])) <= tol
    a.free()


def test_invalid_leaves_rejected(gpu, H):
    recs = (H.LeafRec * 1)()
    recs[0].ltmtx, recs[0].kt, recs[0].nstrtl, recs[0].ndl = 2, 0, 10, 5
    recs[0].nstrtt, recs[0].ndt = 1, 5
    buf = np.zeros(25)
    recs[0].a1 = buf.ctypes.data_as(C.POINTER(C.c_double))
    with pytest.raises(H.HmvError, match="outside"):
        H.Arena(None, "f64", nd=12, nlf=1, leaves_ptr=C.addressof(recs), stride=C.sizeof(H.LeafRec))
    recs[0].ltmtx = 3
    with pytest.raises(H.HmvError, match="ltmtx"):
        H.Arena(None, "f64", nd=20, nlf=1, leaves_ptr=C.addressof(recs), stride=C.sizeof(H.LeafRec))


@pytest.mark.parametrize("prec", ["f64", "bf16"])
def test_arena_file_round_trip(gpu, O, sphere4, tmp_path, prec):
    """An arena written by hmv_arena_save and one written on the host from
    the plan both load (hmv_arena_load) into arenas whose apply and H-matmat
    are bit-equal to the original's."""
    a = gpu.Arena(sphere4, prec)
    x = gpu.uniform(12, sphere4.nd)
    y = a.matvec(x)
    X = np.stack([gpu.uniform(20 + j, sphere4.nd) for j in range(8)], axis=1)
    Y = a.matmat(X)
    a.save(tmp_path / "dev.hmv")
    gpu.save_plan(sphere4, tmp_path / "host.hmv", prec)
    for name in ("dev.hmv", "host.hmv"):
        b = gpu.Arena.load(tmp_path / name)
        assert b.nd == sphere4.nd and b.precision == prec
        assert b.stats["arena_bytes"] == a.stats["arena_bytes"]
        assert np.array_equal(b.matvec(x), y)
        assert np.array_equal(b.matmat(X), Y)
        b.free()
    a.free()
    assert (tmp_path / "dev.hmv").read_bytes() == (tmp_path / "host.hmv").read_bytes()


def test_multi_arena_file_keeps_instances(gpu, O, sphere3, sphere4, tmp_path):
    m = gpu.MultiArena([sphere3, sphere4], "f64")
    m.save(tmp_path / "multi.hmv")
    b = gpu.Arena.load(tmp_path / "multi.hmv")
    offs = (C.c_int * 3)()
    assert gpu.hmv_lib().hmv_arena_segments(b.handle, offs, 3) == 2
    assert list(offs) == [0, sphere3.nd, sphere3.nd + sphere4.nd]
    x = gpu.uniform(13, b.nd)
    assert np.array_equal(b.matvec(x), m.matvec(x))
    b.free()
    m.free()


def test_calibration_kernels(gpu):
    """The plain read/copy and FP64 MFMA-rate calibrations run and give
    plausible rates (bandwidth below 8 TB/s, MFMA rate below the published
    78.6 TFLOP/s FP64 peak)."""
    nbytes = 256 << 20
    ms_read = gpu.stream_test(0, nbytes, reps=3)
    ms_copy = gpu.stream_test(1, nbytes, reps=3)
    tf = gpu.mfma_rate(mfmas_per_wave=8192, reps=2)
    print("read %.0f GB/s, copy %.0f GB/s, FP64 MFMA %.1f TFLOP/s"
          % (nbytes / ms_read / 1e6, 2 * nbytes / ms_copy / 1e6, tf), flush=True)
    assert 0 < nbytes / ms_read / 1e6 < 8000 and 0 < 2 * nbytes / ms_copy / 1e6 < 8000
    assert 1.0 < tf < 78.6

