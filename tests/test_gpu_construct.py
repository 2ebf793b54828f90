"""Device construction (hmv_construct): dense entries and ACA+ on the GPU
against the host fill (the restatement of HACApK_generate's fill,
m_HACApK_base.f90:783-1051), and arenas repacked from device factors against
arenas from host factors."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _leaf_values(hm, lod):
    """Per leaf: (ltmtx, factors, diagonal mask or None)."""
    out = []
    for r in hm.leaves():
        if r.ltmtx == 1:
            n = (r.ndt + r.ndl) * r.kt
            out.append((1, np.ctypeslib.as_array(r.a1, shape=(n,)).copy() if n else np.zeros(0), None))
        else:
            a = np.ctypeslib.as_array(r.a1, shape=(r.ndl, r.ndt)).copy()
            rows = lod[r.nstrtl - 1:r.nstrtl - 1 + r.ndl]
            cols = lod[r.nstrtt - 1:r.nstrtt - 1 + r.ndt]
            out.append((2, a, rows[:, None] == cols[None, :]))
    return out


def test_centroid_kernel_fill_matches_host(gpu, O, H):
    """With the centroid kernel an off-diagonal entry is one sqrt and one
    division, correctly rounded on either side, and ACA+ runs the same
    arithmetic in the same order: ranks, every low-rank factor and every
    off-diagonal dense entry equal the host fill's bit for bit.  The diagonal
    (the analytic self term, through atan2 and log) may differ by an ulp."""
    nodes, faces = H.icosphere(4)
    p = H.baseline_params(aca_eps=1e-5)
    host = H.HMatrix(nodes, faces, p)
    dev = H.DeviceHMatrix(nodes, faces, p)
    assert np.array_equal(dev.table, host.leaf_table())
    got = dev.download()
    ndiff = 0
    for (kind, x, diag), (_, y, _) in zip(_leaf_values(host, host.lod), _leaf_values(got, host.lod)):
        if kind == 1:
            assert np.array_equal(x, y)
        else:
            assert np.array_equal(x[~diag], y[~diag])
            assert np.all(np.abs(x[diag] - y[diag]) <= 4e-16 * np.abs(x[diag]))
            ndiff += int((x[diag] != y[diag]).sum())
    print("leaves", host.nlf, "diagonal entries differing by an ulp:", ndiff,
          "device fill %.4f s" % dev.fill_seconds, flush=True)
    x = H.uniform(3, host.nd)
    y, yr = O.adot_body_lfmtx(got, x), O.adot_body_lfmtx(host, x)
    assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < 1e-15
    dev.free()


def test_analytic_kernel_fill(gpu, O, H):
    """The reference's own kernel (face_integral for every pair) through the
    device's atan2/log: entries differ from the host's in the last bits, which
    can change a near-tied ACA+ pivot; both fills approximate the same matrix
    to the ACA tolerance."""
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "input_1ts_mesh.npz"))
    p = H.HParams(leaf_size=15, aca_eps=1e-6, kernel="ppohbem")
    host = H.HMatrix(d["nodes"], d["faces"], p)
    dev = H.DeviceHMatrix(d["nodes"], d["faces"], p)
    ht = host.leaf_table()
    assert np.array_equal(dev.table[:, [0, 2, 3, 4, 5]], ht[:, [0, 2, 3, 4, 5]])
    assert np.mean(dev.table[:, 1] != ht[:, 1]) < 0.02  # ranks
    got = dev.download()
    A = O.dense(host)
    x = H.uniform(4, host.nd)
    ya = A @ x
    for hm in (host, got):
        assert np.linalg.norm(O.adot_body_lfmtx(hm, x) - ya) / np.linalg.norm(ya) < 2e-5
    dense_dev = O.dense(got)
    assert np.linalg.norm(dense_dev - A) / np.linalg.norm(A) < 2e-5
    dev.free()


def test_high_rank_blocks_take_the_second_pass(gpu, O, H):
    """ACA+ to 1e-14 needs more columns than the first pass reserves (64) on
    the larger blocks; those blocks run again with their full bound and
    still equal the host fill."""
    nodes, faces = H.icosphere(4)
    p = H.baseline_params(aca_eps=1e-14)
    host = H.HMatrix(nodes, faces, p)
    dev = H.DeviceHMatrix(nodes, faces, p)
    assert dev.kt.max() > 64
    assert np.array_equal(dev.table, host.leaf_table())
    got = dev.download()
    for (kind, x, diag), (_, y, _) in zip(_leaf_values(host, host.lod), _leaf_values(got, host.lod)):
        if kind == 1:
            assert np.array_equal(x, y)
    dev.free()


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16", "bf16_all"])
def test_arena_from_device_factors(gpu, O, H, tmp_path, prec):
    """The arena repacked on the device from device factors is byte-equal to
    the host repack of the same factors (downloaded), and applies bit-equal."""
    nodes, faces = H.icosphere(4)
    p = H.baseline_params(aca_eps=1e-5)
    dev = H.DeviceHMatrix(nodes, faces, p)
    got = dev.download()
    a = dev.arena(prec)
    b = H.Arena(got, prec)
    x = H.uniform(5, dev.nd)
    assert np.array_equal(a.matvec(x), b.matvec(x))
    a.save(tmp_path / "dev.hmv")
    b.save(tmp_path / "host.hmv")
    assert (tmp_path / "dev.hmv").read_bytes() == (tmp_path / "host.hmv").read_bytes()
    y = a.matvec(x)
    y_ref = O.adot_body_lfmtx(got, x)
    tol = {"f64": 1e-12, "f32": 1e-6, "bf16": 2e-2, "bf16_all": 2e-2}[prec]
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < tol
    a.free()
    b.free()
    dev.free()
