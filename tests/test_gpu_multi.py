"""Many independent instances in one arena (one launch per phase) and their
per-instance BiCGSTAB and GCR(m) solves, through the C ABI.

Bar: the concatenated apply and H-matmat equal each instance's own apply
(FP64 <= 1e-13 relative per instance) and the oracle; each instance's solve
reaches the true-residual bar and agrees with its single-instance solve.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def instances(gpu):
    H = gpu
    out = []
    for i, (s, r) in enumerate([(2, 0.5), (3, 1.7), (2, 1.2), (3, 0.8), (4, 1.0)]):
        nodes, faces = H.icosphere(s, r)
        out.append(H.HMatrix(nodes, faces, H.baseline_params(aca_eps=1e-4)))
    return out


def test_multi_apply_matches_instances(gpu, O, instances):
    m = gpu.MultiArena(instances, "f64")
    assert m.offsets[-1] == sum(hm.nd for hm in instances) == m.nd
    x = gpu.uniform(7, m.nd)
    y = m.matvec(x)
    for i, hm in enumerate(instances):
        lo, hi = m.offsets[i], m.offsets[i + 1]
        a = gpu.Arena(hm, "f64")
        assert _rel(y[lo:hi], a.matvec(x[lo:hi])) <= 1e-13
        assert _rel(y[lo:hi], O.adot_body_lfmtx(hm, x[lo:hi])) <= 1e-12
        a.free()
    m.free()


@pytest.mark.parametrize("k", [8, 32])
def test_multi_matmat_matches_instances(gpu, O, instances, k):
    """The H-matmat over a multi-instance arena is block diagonal: each
    instance's rows of every column equal that instance's own matvec."""
    m = gpu.MultiArena(instances, "f64")
    X = np.stack([gpu.uniform(50 + j, m.nd) for j in range(k)], axis=1)
    Y = m.matmat(X)
    for i, hm in enumerate(instances):
        lo, hi = m.offsets[i], m.offsets[i + 1]
        a = gpu.Arena(hm, "f64")
        for j in (0, k - 1):
            assert _rel(Y[lo:hi, j], a.matvec(X[lo:hi, j])) <= 1e-13
        assert _rel(Y[lo:hi, k // 2], O.adot_body_lfmtx(hm, X[lo:hi, k // 2])) <= 1e-12
        a.free()
    m.free()


def test_multi_bicgstab_per_instance(gpu, instances):
    m = gpu.MultiArena(instances, "f64")
    b = np.concatenate([np.ones(hm.nd) if i % 2 else gpu.uniform(20 + i, hm.nd)
                        for i, hm in enumerate(instances)])
    u, inst, secs, info = m.bicgstab_multi(b, eps=1e-8, maxit=500, check_every=4)
    print("multi", info, [d["nstp"] for d in inst], secs, flush=True)
    assert info["status_name"] == "converged"
    assert np.all(secs > 0) and np.all(secs <= info["seconds"] + 1e-9)
    for i, hm in enumerate(instances):
        lo, hi = m.offsets[i], m.offsets[i + 1]
        assert inst[i]["status_name"] == "converged" and inst[i]["true_relres"] < 1.5e-8
        a = gpu.Arena(hm, "f64")
        u1, i1, _ = a.bicgstab(b[lo:hi], eps=1e-8, maxit=500)
        assert abs(inst[i]["nstp"] - i1["nstp"]) <= 3
        assert _rel(u[lo:hi], u1) < 1e-6
        a.free()
    m.free()


def test_multi_zero_instance_frozen(gpu, instances):
    m = gpu.MultiArena(instances[:3], "f64")
    b = gpu.uniform(30, m.nd)
    b[m.offsets[1]:m.offsets[2]] = 0.0
    u, inst, secs, info = m.bicgstab_multi(b, eps=1e-8, maxit=500)
    assert inst[1]["status_name"] == "nonfinite" and inst[1]["nstp"] == 0
    assert np.all(u[m.offsets[1]:m.offsets[2]] == 0.0)
    assert inst[0]["status_name"] == "converged" and inst[2]["status_name"] == "converged"
    m.free()


def test_multi_low_precision_apply(gpu, O, instances):
    m = gpu.MultiArena(instances, "bf16")
    x = gpu.uniform(8, m.nd)
    y = m.matvec(x)
    for i, hm in enumerate(instances):
        lo, hi = m.offsets[i], m.offsets[i + 1]
        assert _rel(y[lo:hi], O.adot_body_lfmtx(hm, x[lo:hi])) <= 2e-2
    m.free()


def _cond_bound(nd):
    """cond * 2e-8 with cond = 2 sqrt(N) + 1, the spread of the sphere
    single-layer operator's eigenvalues R/(2l+1) over (L+1)^2 ~ N modes."""
    return (2.0 * np.sqrt(nd) + 1.0) * 2e-8


@pytest.mark.parametrize("prec,inner", [("f32", 1e-4), ("bf16", 1e-2), ("bf16_all", 1e-2)])
def test_multi_refinement_per_instance(gpu, O, instances, prec, inner):
    """hmv_multi_bicgstab_ir: FP64 residuals on the FP64 arena, inner
    per-instance BiCGSTAB on the low-precision arena.  Every instance reaches
    a true 1e-8 (recomputed by the oracle) and its FP64 solution within the
    conditioning bound; a zero right-hand side stops its instance at once
    (non-finite relative residual) and leaves its iterate at zero."""
    m64, mlow = gpu.MultiArena(instances, "f64"), gpu.MultiArena(instances, prec)
    b = np.concatenate([np.ones(hm.nd) if i % 2 else gpu.uniform(40 + i, hm.nd)
                        for i, hm in enumerate(instances)])
    u, inst, secs, info = gpu.multi_bicgstab_ir(m64, mlow, b, eps=1e-8, inner_eps=inner,
                                                max_outer=30, max_inner=500)
    print(prec, info, [(d["nstp"], "%.1e" % d["true_relres"]) for d in inst], flush=True)
    assert info["status_name"] == "converged" and info["outer"] >= 2
    assert np.all(secs > 0) and np.all(secs <= info["seconds"] + 1e-9)
    for i, hm in enumerate(instances):
        lo, hi = m64.offsets[i], m64.offsets[i + 1]
        tr = _rel(O.adot_body_lfmtx(hm, u[lo:hi]), b[lo:hi])
        assert inst[i]["status_name"] == "converged" and tr < 1e-8
        assert abs(tr - inst[i]["true_relres"]) <= 1e-3 * tr
        u_ref, _, _ = O.bicgstab(hm, b[lo:hi], None, mstep=500, eps=1e-10)
        assert _rel(u[lo:hi], u_ref) < _cond_bound(hm.nd)
    b[m64.offsets[1]:m64.offsets[2]] = 0.0
    u, inst, secs, info = gpu.multi_bicgstab_ir(m64, mlow, b, eps=1e-8, inner_eps=inner)
    assert inst[1]["status_name"] == "nonfinite" and inst[1]["nstp"] == 0
    assert np.all(u[m64.offsets[1]:m64.offsets[2]] == 0.0)
    assert all(inst[i]["status_name"] == "converged" for i in (0, 2, 3, 4))
    m64.free()
    mlow.free()


@pytest.mark.parametrize("prec,m", [("bf16_all", 32), ("f32", 32), ("bf16", 4)])
def test_multi_gcr_refinement_per_instance(gpu, O, instances, prec, m):
    """hmv_multi_gcr_ir: the same refinement with one GCR(m) per instance.
    Every instance reaches a true 1e-8 (recomputed by the oracle) and its
    FP64 solution within the conditioning bound; each instance takes the
    inner steps of its own single-instance hmv_gcr_ir (the per-instance
    dots sum other partial groupings, so up to 3 steps apart over all outer
    steps); m = 4 reuses direction slots (the restart step).  A zero
    right-hand side stops its instance at once and leaves it at zero."""
    m64, mlow = gpu.MultiArena(instances, "f64"), gpu.MultiArena(instances, prec)
    b = np.concatenate([np.ones(hm.nd) if i % 2 else gpu.uniform(40 + i, hm.nd)
                        for i, hm in enumerate(instances)])
    u, inst, secs, info = gpu.multi_gcr_ir(m64, mlow, b, eps=1e-8, inner_eps=5e-3, m=m,
                                           max_outer=30, max_inner=500)
    print(prec, m, info, [(d["nstp"], "%.1e" % d["true_relres"]) for d in inst], flush=True)
    assert info["status_name"] == "converged" and info["outer"] >= 2
    assert np.all(secs > 0) and np.all(secs <= info["seconds"] + 1e-9)
    for i, hm in enumerate(instances):
        lo, hi = m64.offsets[i], m64.offsets[i + 1]
        tr = _rel(O.adot_body_lfmtx(hm, u[lo:hi]), b[lo:hi])
        assert inst[i]["status_name"] == "converged" and tr < 1e-8
        assert abs(tr - inst[i]["true_relres"]) <= 1e-3 * tr
        u_ref, _, _ = O.bicgstab(hm, b[lo:hi], None, mstep=500, eps=1e-10)
        assert _rel(u[lo:hi], u_ref) < _cond_bound(hm.nd)
        a64, alow = gpu.Arena(hm, "f64"), gpu.Arena(hm, prec)
        _, inf1, _ = gpu.gcr_ir(a64, alow, b[lo:hi], eps=1e-8, inner_eps=5e-3, m=m)
        assert abs(inf1["inner_total"] - inst[i]["nstp"]) <= 3, (i, inf1, inst[i])
        a64.free()
        alow.free()
    b[m64.offsets[1]:m64.offsets[2]] = 0.0
    u, inst, secs, info = gpu.multi_gcr_ir(m64, mlow, b, eps=1e-8, inner_eps=5e-3, m=m)
    assert inst[1]["status_name"] == "nonfinite" and inst[1]["nstp"] == 0
    assert np.all(u[m64.offsets[1]:m64.offsets[2]] == 0.0)
    assert all(inst[i]["status_name"] == "converged" for i in (0, 2, 3, 4))
    m64.free()
    mlow.free()


def test_multi_refinement_rejects_other_instances(gpu, instances):
    m64 = gpu.MultiArena(instances, "f64")
    other = gpu.MultiArena(instances[:3], "bf16")
    with pytest.raises(gpu.HmvError, match="same instances"):
        gpu.multi_bicgstab_ir(m64, other, np.ones(m64.nd))
    with pytest.raises(gpu.HmvError, match="same instances"):
        gpu.multi_gcr_ir(m64, other, np.ones(m64.nd))
    m64.free()
    other.free()


def test_multi_autotune(gpu, O, instances):
    """hmv_autotune_multi returns a tuning no slower than the defaults on the
    whole block-diagonal apply; a multi arena built with it matches the
    oracle per instance."""
    best, best_ms, dflt_ms = gpu.autotune_multi(instances, "bf16", reps=5)
    print("multi autotune", best, "%.4f ms (defaults %.4f ms)" % (best_ms, dflt_ms), flush=True)
    assert 0 < best_ms <= dflt_ms
    m = gpu.MultiArena(instances, "bf16", tuning=best)
    assert m.tuning == best
    x = gpu.uniform(9, m.nd)
    y = m.matvec(x)
    for i, hm in enumerate(instances):
        lo, hi = m.offsets[i], m.offsets[i + 1]
        assert _rel(y[lo:hi], O.adot_body_lfmtx(hm, x[lo:hi])) <= 2e-2
    m.free()
