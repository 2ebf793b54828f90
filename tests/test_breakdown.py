"""Breakdown and non-finite systems: hand-made 2 x 2 and 4 x 4 dense
systems (tests/synthetic_leaves.py).  The reference's BiCGSTAB has no guard
(m_HACApK_solve.f90:404-415 divides by (r^, A p) and (A t, A t) as they
come); the engine's solvers stop and report HMV_SOLVE_BREAKDOWN or
HMV_SOLVE_NONFINITE, and take BiCGSTAB's half step when A t = 0 solves the
system.

CPU part: the repacks of these tiny systems pass the kernel-walk emulation
(every access in bounds) before any GPU run; GPU part: every solver exits
normally with the status on the host.
"""
import numpy as np
import pytest

import plan_emulator as PE
import synthetic_leaves as SL


@pytest.mark.parametrize("make", [SL.skew_system, SL.skew_and_scaled_identity,
                                  SL.scaled_identity, lambda H: SL.with_nan(H, 5.0)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_tiny_plans_emulate(H, make, prec):
    hm = make(H)
    T = H.plan_tables(hm, prec)
    x = np.arange(1.0, hm.nd + 1.0)
    t, y = PE.run(T, x, hm.nd)
    assert np.allclose(y, hm.dense @ x, rtol=1e-6, atol=0)


gpu = pytest.mark.gpu


def _solve_all(H, hm, b):
    """(status, nstp, u) of hmv_bicgstab, hmv_bicgstab_pipe and hmv_gcrm."""
    a = H.Arena(hm, "f64")
    out = {}
    try:
        u, info, _ = a.bicgstab(b, eps=1e-8, maxit=20)
        out["bicgstab"] = (info["status_name"], info["nstp"], u)
        u, info, _ = a.bicgstab_pipe(b, eps=1e-8, maxit=20)
        out["pipe"] = (info["status_name"], info["nstp"], u)
        u, info, _ = a.gcrm(b, eps=1e-8, maxit=20, mreset=4)
        out["gcrm"] = (info["status_name"], info["nstp"], u)
    finally:
        a.free()
    for k, v in out.items():
        print(k, v[0], "nstp", v[1], "u", v[2], flush=True)
    return out


@gpu
def test_skew_system_breaks_down(gpu):
    """(r^, A p) = 0 at step 1: BiCGSTAB and the pipelined BiCGSTAB stop with
    BREAKDOWN after 0 steps (the pipelined one at its setup, where alpha =
    (r0,r0)/(r0,A r0) is infinite).  GCR's first step has alpha = (r, A r) /
    (A r, A r) = 0, so r does not change; the next direction, A r
    orthogonalised against A p = A r, is zero, and GCR stops with BREAKDOWN
    after that one step (the reference's GCR would divide by zero,
    m_HACApK_solve.f90:546-548).  Every iterate stays finite."""
    hm = SL.skew_system(gpu)
    b = np.array([1.0, 0.0])
    out = _solve_all(gpu, hm, b)
    assert out["bicgstab"][:2] == ("breakdown", 0)
    assert out["pipe"][:2] == ("breakdown", 0)
    assert out["gcrm"][:2] == ("breakdown", 1)
    for name in ("bicgstab", "pipe", "gcrm"):
        assert np.array_equal(out[name][2], [0.0, 0.0]), name


@gpu
def test_half_step_solves(gpu):
    """t = r - alpha A p = 0 (A = 2 I): BiCGSTAB's (A t, A t) is 0; the half
    step u = alpha p is the solution, reported as converged after one step
    (before, the solver returned the previous iterate as a breakdown)."""
    hm = SL.scaled_identity(gpu)
    b = np.array([1.0, 1.0])
    out = _solve_all(gpu, hm, b)
    for name in ("bicgstab", "pipe", "gcrm"):
        status, nstp, u = out[name]
        assert status == "converged" and nstp == 1, name
        assert np.array_equal(u, b / 2), name


@gpu
def test_nonfinite_leaf(gpu):
    """A leaf holding a NaN: the first product is non-finite, and every
    solver reports NONFINITE after 0 steps."""
    hm = SL.with_nan(gpu)
    b = np.array([1.0, 1.0])
    a = gpu.Arena(hm, "f64")
    y = a.matvec(b)
    assert np.isnan(y[0]) and y[1] == 3.0
    a.free()
    out = _solve_all(gpu, hm, b)
    for name in ("bicgstab", "pipe", "gcrm"):
        assert out[name][0] == "nonfinite", name
        assert out[name][1] == 0, name


@gpu
def test_block_bicgstab_columns_stop_on_their_own(gpu):
    """Block BiCGSTAB on [SKEW, 2 I] with k = 2: column 0 (b on the skew
    block) breaks down at step 1, column 1 (b on the 2 I block) is solved by
    its first half step; each column reports its own status and the solve
    returns normally.  With a NaN leaf both columns are non-finite."""
    hm = SL.skew_and_scaled_identity(gpu)
    a = gpu.Arena(hm, "f64")
    B = np.array([[1.0, 0.0], [0.0, 0.0], [0.0, 1.0], [0.0, 1.0]])
    U, cols, info = a.block_bicgstab(B, eps=1e-8, maxit=20, check_every=1)
    print(cols, info, flush=True)
    assert cols[0]["status_name"] == "breakdown" and cols[0]["nstp"] == 0
    assert cols[1]["status_name"] == "converged" and cols[1]["nstp"] == 1
    assert np.array_equal(U[:, 1], [0.0, 0.0, 0.5, 0.5])
    assert np.all(np.isfinite(U))
    a.free()
    hm = SL.with_nan(gpu)
    a = gpu.Arena(hm, "f64")
    U, cols, info = a.block_bicgstab(np.ones((2, 2)), eps=1e-8, maxit=20, check_every=1)
    assert [c["status_name"] for c in cols] == ["nonfinite", "nonfinite"]
    a.free()
