"""Many independent instances in one arena (one launch per phase) and their
per-instance BiCGSTAB solves, through the C ABI.

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
