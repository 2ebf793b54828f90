"""GPU parity at the benchmark configurations' full sizes, against the
oracle (threaded restatement of the reference's CPU path,
m_HACApK_solve.f90:203-253 with the per-leaf GEMVs of HACApK_MAGMA.c:114-199,
and its BiCGSTAB, m_HACApK_solve.f90:342-449).  The reference checks its
GPU apply against the CPU one at production size the same way
(m_HACApK_solve.f90:684-702,717-735).

What only these sizes exercise: FP64 arenas past 2^31 bytes (4.8 GB at
327,680 unknowns, 20.8 GB for the 48-instance batch), column-chunked
phase-1 groups (long low-rank blocks of the largest clusters), tens of
thousands of workgroups per phase.

Configurations:
  two spheres   two icospheres s=6 (r=1, centres 3 apart), 163,840 unknowns,
                ACA+ 1e-5: FP64 apply and BiCGSTAB (RHS U[-1,1] seed 1)
  sphere s=7    327,680 unknowns, ACA+ 1e-5: FP32 / BF16 / all-BF16 applies,
                mixed-precision refinement to a true 1e-8, H-matmat and block
                BiCGSTAB with 8, 32 and 128 right-hand sides (U[-1,1] seed 3)
  48 instances  icospheres s in {4,5,6}, radii U[0.5,2] (seed 4), ACA+ 1e-4,
                2,042,880 unknowns: per-instance applies and solves

Bars: FP64 products 1e-12 relative (the FP64 bar of every GPU test); FP32
storage 1e-5, BF16 2e-2 (storage rounding, as tests/test_gpu_apply.py);
solves: true residual ||b - A u|| / ||b|| recomputed here with the oracle's
FP64 apply below 1e-8 (x 1.5 where the solver stops on its recursive
residual and the rounding of the final update shows); solutions against the
oracle's FP64 solution within cond * 2e-8, cond = 2 sqrt(N) + 1: on a
sphere the single-layer operator's eigenvalues are R/(2l+1), l = 0..L with
(L+1)^2 ~ N modes, so two solutions whose residuals differ by 2e-8 of ||b||
can differ by that much.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _nthr():
    return max(1, min(16, os.cpu_count() or 1))


def _true_relres(O, hm, u, b):
    return _rel(O.adot_lfmtx_hyp(hm, u, _nthr()), b)


def _cond_bound(nd):
    return (2.0 * np.sqrt(nd) + 1.0) * 2e-8


@pytest.fixture(scope="module")
def two_spheres6(gpu):
    H = gpu
    m = H.merge(H.icosphere(6), H.icosphere(6, 1.0, (3.0, 0.0, 0.0)))
    return H.HMatrix(*m, H.baseline_params(aca_eps=1e-5))


@pytest.fixture(scope="module")
def sphere7(gpu):
    return gpu.HMatrix(*gpu.icosphere(7), gpu.baseline_params(aca_eps=1e-5))


@pytest.fixture(scope="module")
def sphere7_f64(gpu, sphere7):
    a = gpu.Arena(sphere7, "f64")
    print("s=7 FP64 arena %.2f GB" % (a.stats["arena_bytes"] / 1e9), flush=True)
    assert a.stats["arena_bytes"] > 2 ** 31
    yield a
    a.free()


@pytest.fixture(scope="module")
def sphere7_oracle_solution(gpu, O, sphere7):
    """The oracle's FP64 BiCGSTAB solutions (RHS = 1 and U[-1,1] seed 2) to 1e-10."""
    H = gpu
    out = {}
    for name, b in (("ones", np.ones(sphere7.nd)), ("uniform2", H.uniform(2, sphere7.nd))):
        u, nstp, hist = O.bicgstab(sphere7, b, None, mstep=3000, eps=1e-10, nthr=_nthr())
        print("oracle s=7 %s: %d steps to %.1e" % (name, nstp, hist[-1]), flush=True)
        out[name] = (b, u)
    return out


# ---- two spheres, 163,840 unknowns -------------------------------------------
def test_two_spheres_f64_apply(gpu, O, two_spheres6):
    hm = two_spheres6
    a = gpu.Arena(hm, "f64")
    print("two spheres: nd %d, arena %.2f GB, %d + %d workgroups, %d t parts (%d slots)"
          % (hm.nd, a.stats["arena_bytes"] / 1e9, a.stats["nwork_phase1"], a.stats["nwork_phase2"],
             a.stats["tparts"], a.stats["tslots"]), flush=True)
    assert a.stats["tparts"] > a.stats["tslots"]  # column-chunked groups present
    for seed in (1, 2):
        x = gpu.uniform(seed, hm.nd)
        err = _rel(a.matvec(x), O.adot_lfmtx_hyp(hm, x, _nthr()))
        print("seed %d: rel err vs oracle %.3e" % (seed, err), flush=True)
        assert err <= 1e-12
    a.free()


def test_two_spheres_bicgstab(gpu, O, two_spheres6):
    """BiCGSTAB to 1e-8 on the bench RHS: the early residual history equals
    the oracle's, the true residual is below the bar, the solution within
    the conditioning bound of the oracle's.  The GPU sums its dot products
    in a tree, the oracle in sequence, and BiCGSTAB amplifies that rounding
    difference quickly at this size: measured log10 differences 1e-8 at
    step 7, 6e-7 at step 8, 6e-5 at step 9.  The histories are compared to
    1e-6 while the residual is above 0.25 (steps 0-7)."""
    hm = two_spheres6
    b = gpu.uniform(1, hm.nd)
    a = gpu.Arena(hm, "f64")
    u, info, hist = a.bicgstab(b, eps=1e-8, maxit=2000, check_every=1)
    a.free()
    u_ref, nstp_ref, hist_ref = O.bicgstab(hm, b, None, mstep=2000, eps=1e-8, nthr=_nthr())
    tr = _true_relres(O, hm, u, b)
    print("two spheres BiCGSTAB: gpu %d steps, oracle %d; true relres %.2e (gpu's own %.2e); "
          "|u-u_ref|/|u_ref| %.2e" % (info["nstp"], nstp_ref, tr, info["true_relres"],
                                      _rel(u, u_ref)), flush=True)
    assert info["status_name"] == "converged"
    n = min(len(hist), len(hist_ref))
    early = (hist_ref[:n] > 0.25) & (hist[:n] > 0.25)
    assert early.sum() >= 6
    assert np.allclose(np.log10(hist[:n][early]), np.log10(hist_ref[:n][early]), atol=1e-6)
    assert abs(info["nstp"] - nstp_ref) <= 0.1 * nstp_ref
    assert tr < 1.5e-8
    assert _rel(u, u_ref) < _cond_bound(hm.nd)


# ---- one sphere, 327,680 unknowns --------------------------------------------
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5), ("bf16", 2e-2),
                                      ("bf16_all", 2e-2)])
def test_sphere7_apply(gpu, O, sphere7, sphere7_f64, prec, tol):
    hm = sphere7
    a = sphere7_f64 if prec == "f64" else gpu.Arena(hm, prec)
    x = gpu.uniform(5, hm.nd)
    err = _rel(a.matvec(x), O.adot_lfmtx_hyp(hm, x, _nthr()))
    print("s=7 %s: arena %.2f GB, rel err vs FP64 oracle %.3e"
          % (prec, a.stats["arena_bytes"] / 1e9, err), flush=True)
    assert err <= tol
    if prec != "f64":
        a.free()


@pytest.mark.parametrize("method,prec", [("bicgstab_ir", "f32"), ("gcr_ir", "bf16"),
                                         ("gcr_ir", "bf16_all")])
@pytest.mark.parametrize("rhs", ["ones", "uniform2"])
def test_sphere7_mixed_refinement(gpu, O, sphere7, sphere7_f64, sphere7_oracle_solution, method,
                                  prec, rhs):
    """FP64 iterative refinement over low-precision inner solves reaches a
    true 1e-8 (FP64 residual recomputed by the oracle) and the oracle's FP64
    solution within the conditioning bound.  bicgstab_ir: inner BiCGSTAB
    to 1e-4 (the FP32 configuration); gcr_ir: inner GCR(32) to 5e-3."""
    hm = sphere7
    b, u_ref = sphere7_oracle_solution[rhs]
    alow = gpu.Arena(hm, prec)
    if method == "bicgstab_ir":
        u, info, hist = gpu.bicgstab_ir(sphere7_f64, alow, b, eps=1e-8, inner_eps=1e-4,
                                        max_outer=40, max_inner=3000)
    else:
        u, info, hist = gpu.gcr_ir(sphere7_f64, alow, b, eps=1e-8, inner_eps=5e-3, m=32)
    alow.free()
    tr = _true_relres(O, hm, u, b)
    err = _rel(u, u_ref)
    print("s=7 %s %s %s: outer %d, inner %d, outer residuals %s; true relres %.2e, err vs oracle "
          "solution %.2e (bound %.1e)" % (method, prec, rhs, info["outer"], info["inner_total"],
                                          ["%.1e" % v for v in hist], tr, err,
                                          _cond_bound(hm.nd)), flush=True)
    assert info["status_name"] == "converged"
    assert tr < 1e-8
    assert err < _cond_bound(hm.nd)


@pytest.mark.parametrize("k", [8, 32, 128])
def test_sphere7_matmat(gpu, O, sphere7, sphere7_f64, k):
    """H-matmat with k right-hand sides: sampled columns against the oracle."""
    hm = sphere7
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (hm.nd, k))
    Y = sphere7_f64.matmat(X)
    for j in sorted({0, k // 3, k - 1}):
        err = _rel(Y[:, j], O.adot_lfmtx_hyp(hm, X[:, j], _nthr()))
        print("s=7 H-matmat k=%d column %d rel err %.3e" % (k, j, err), flush=True)
        assert err <= 1e-12


def test_sphere7_matmat_f32_factors(gpu, O, sphere7):
    hm = sphere7
    a = gpu.Arena(hm, "f32")
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (hm.nd, 32))
    Y = a.matmat(X)
    a.free()
    for j in (0, 31):
        err = _rel(Y[:, j], O.adot_lfmtx_hyp(hm, X[:, j], _nthr()))
        print("s=7 FP32 H-matmat k=32 column %d rel err %.3e" % (j, err), flush=True)
        assert err <= 1e-5


@pytest.mark.parametrize("k", [8, 32, 128])
def test_sphere7_block_bicgstab(gpu, O, sphere7, sphere7_f64, k):
    """Block BiCGSTAB: every column converges, the solver's own true
    residuals are below the bar, and sampled columns' true residuals
    recomputed by the oracle agree."""
    hm = sphere7
    rng = np.random.default_rng(3)
    B = rng.uniform(-1, 1, (hm.nd, k))
    U, cols, info = sphere7_f64.block_bicgstab(B, eps=1e-8, maxit=3000, check_every=8)
    worst = max(c["true_relres"] for c in cols)
    print("s=7 block BiCGSTAB k=%d: %d iterations, %.3f s, worst true relres %.2e"
          % (k, info["nstp"], info["seconds"], worst), flush=True)
    assert info["status_name"] == "converged"
    assert all(c["status_name"] == "converged" for c in cols)
    assert worst < 1.5e-8
    for j in sorted({0, k - 1}):
        tr = _true_relres(O, hm, U[:, j], B[:, j])
        assert tr < 1.5e-8
        assert abs(tr - cols[j]["true_relres"]) <= 1e-3 * tr


@pytest.mark.parametrize("k", [8, 32, 128])
def test_sphere7_block_refinement_f32(gpu, O, sphere7, sphere7_f64, k):
    """Block BiCGSTAB with FP32 factors, refined in FP64 to 1e-8 per column
    (hmv_block_bicgstab_ir, inner 1e-4): every column converges, sampled
    columns' true residuals recomputed by the oracle."""
    hm = sphere7
    alow = gpu.Arena(hm, "f32")
    rng = np.random.default_rng(3)
    B = rng.uniform(-1, 1, (hm.nd, k))
    U, cols, info = gpu.block_bicgstab_ir(sphere7_f64, alow, B, eps=1e-8, inner_eps=1e-4,
                                          max_outer=40, max_inner=3000)
    alow.free()
    print("s=7 FP32 block refinement k=%d: %.3f s, %d outer, %d inner iterations, worst relres %.2e"
          % (k, info["seconds"], info["outer"], info["inner_total"], info["true_relres"]),
          flush=True)
    assert info["status_name"] == "converged"
    assert all(c["status_name"] == "converged" and c["true_relres"] < 1e-8 for c in cols)
    for j in sorted({0, k - 1}):
        assert _true_relres(O, hm, U[:, j], B[:, j]) < 1e-8


# ---- 48 instances, 2,042,880 unknowns -----------------------------------------
@pytest.fixture(scope="module")
def batch48(gpu):
    H = gpu
    rng = np.random.default_rng(4)
    subs = rng.choice([4, 5, 6], size=48)
    radii = rng.uniform(0.5, 2.0, size=48)
    hms = [H.HMatrix(*H.icosphere(int(s), float(r)), H.baseline_params(aca_eps=1e-4))
           for s, r in zip(subs, radii)]
    return hms


@pytest.fixture(scope="module")
def batch48_f64(gpu, batch48):
    m = gpu.MultiArena(batch48, "f64")
    print("48 instances: %d unknowns, FP64 arena %.1f GB" % (m.nd, m.stats["arena_bytes"] / 1e9),
          flush=True)
    assert m.nd == 2042880 and m.stats["arena_bytes"] > 2 ** 34
    yield m
    m.free()


def _sample(batch):
    """One instance of each size: the first, plus the last instance."""
    first = {}
    for i, hm in enumerate(batch):
        first.setdefault(hm.nd, i)
    return sorted(set(first.values()) | {len(batch) - 1})


def test_batch48_apply(gpu, O, batch48, batch48_f64):
    m = batch48_f64
    x = gpu.uniform(7, m.nd)
    y = m.matvec(x)
    for i in _sample(batch48):
        lo, hi = m.offsets[i], m.offsets[i + 1]
        err = _rel(y[lo:hi], O.adot_lfmtx_hyp(batch48[i], x[lo:hi], _nthr()))
        print("instance %d (nd %d, rows %d..%d) rel err %.3e" % (i, hi - lo, lo, hi, err), flush=True)
        assert err <= 1e-12


def test_batch48_bicgstab(gpu, O, batch48, batch48_f64):
    """One BiCGSTAB per instance in lockstep: every instance converges to the
    bar; sampled instances' true residuals recomputed by the oracle, and
    their solutions against the oracle's."""
    m = batch48_f64
    b = np.ones(m.nd)
    u, inst, secs, info = m.bicgstab_multi(b, eps=1e-8, maxit=3000, check_every=4)
    print("48 instances: %.3f s, steps %d..%d, worst true relres %.2e, p50 %.3f s p99 %.3f s"
          % (info["seconds"], min(d["nstp"] for d in inst), max(d["nstp"] for d in inst),
             info["true_relres"], np.percentile(secs, 50), np.percentile(secs, 99)), flush=True)
    assert info["status_name"] == "converged"
    assert all(d["status_name"] == "converged" and d["true_relres"] < 1.5e-8 for d in inst)
    for i in _sample(batch48):
        hm = batch48[i]
        lo, hi = m.offsets[i], m.offsets[i + 1]
        tr = _true_relres(O, hm, u[lo:hi], b[lo:hi])
        u_ref, nstp_ref, _ = O.bicgstab(hm, b[lo:hi], None, mstep=3000, eps=1e-8, nthr=_nthr())
        err = _rel(u[lo:hi], u_ref)
        print("instance %d: steps %d (oracle %d), true relres %.2e, err vs oracle %.2e"
              % (i, inst[i]["nstp"], nstp_ref, tr, err), flush=True)
        assert tr < 1.5e-8
        assert err < _cond_bound(hm.nd)


@pytest.mark.parametrize("prec,method,inner", [("bf16", "bicgstab", 1e-2), ("f32", "bicgstab", 1e-4),
                                               ("bf16_all", "gcr", 5e-3)])
def test_batch48_mixed_refinement(gpu, O, batch48, batch48_f64, prec, method, inner):
    """The batch with mixed-precision factors: per-instance FP64 refinement
    with BiCGSTAB (hmv_multi_bicgstab_ir) or GCR(32) (hmv_multi_gcr_ir) inner
    solves over one low-precision multi-instance arena; every instance
    reaches a true 1e-8, sampled instances checked by the oracle's apply and
    against the oracle's FP64 solution."""
    mlow = gpu.MultiArena(batch48, prec)
    b = np.ones(batch48_f64.nd)
    if method == "gcr":
        u, inst, secs, info = gpu.multi_gcr_ir(batch48_f64, mlow, b, eps=1e-8, inner_eps=inner, m=32,
                                               max_outer=40, max_inner=3000)
    else:
        u, inst, secs, info = gpu.multi_bicgstab_ir(batch48_f64, mlow, b, eps=1e-8, inner_eps=inner,
                                                    max_outer=40, max_inner=3000)
    mlow.free()
    print("48 instances %s %s refinement: %.3f s, %d outer steps, worst relres %.2e, p50 %.3f s "
          "p99 %.3f s" % (prec, method, info["seconds"], info["outer"], info["true_relres"],
                          np.percentile(secs, 50), np.percentile(secs, 99)), flush=True)
    assert info["status_name"] == "converged"
    assert all(d["status_name"] == "converged" and d["true_relres"] < 1e-8 for d in inst)
    m = batch48_f64
    for i in _sample(batch48):
        hm = batch48[i]
        lo, hi = m.offsets[i], m.offsets[i + 1]
        tr = _true_relres(O, hm, u[lo:hi], b[lo:hi])
        u_ref, _, _ = O.bicgstab(hm, b[lo:hi], None, mstep=3000, eps=1e-10, nthr=_nthr())
        err = _rel(u[lo:hi], u_ref)
        print("instance %d: true relres %.2e, err vs oracle %.2e" % (i, tr, err), flush=True)
        assert tr < 1e-8
        assert err < _cond_bound(hm.nd)
