"""The random block structures of test_plan_random.py (ragged sizes, dense and
low-rank leaves of every rank from 0, rows covered by no leaf) through the
GPU kernels: the single-vector apply in every storage precision, with the
default and a tight repack tuning, and the H-matmat on every path its
tuning can select, against the oracle's apply (HACApK_MAGMA.c:5-57)."""
import numpy as np
import pytest

from test_gpu_block import _paths
from test_plan_random import TUNINGS, _leaf_set

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5, "bf16": 2e-2, "bf16_all": 2e-2}


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("seed", range(0, 24, 2))
def test_random_structures_on_gpu(gpu, O, seed):
    hm = _leaf_set(gpu, seed)
    x = gpu.uniform(70 + seed, hm.nd)
    y_ref = O.adot_body_lfmtx(hm, x)
    for prec in TOL:
        for tune in (None, TUNINGS[1]):
            a = gpu.Arena(hm, prec, tuning=tune)
            assert _rel(a.matvec(x), y_ref) <= TOL[prec], (seed, prec, tune)
            if tune is None and prec in ("f64", "f32"):
                for k in (3, 20):
                    X = np.stack([gpu.uniform(400 + 10 * seed + j, hm.nd) for j in range(k)], axis=1)
                    refs = {j: O.adot_body_lfmtx(hm, X[:, j]) for j in (0, k - 1)}
                    for vec, tiles in _paths(prec, k):
                        a.set_block_tuning(k, vec, tiles)
                        Y = a.matmat(X)
                        for j, r in refs.items():
                            assert _rel(Y[:, j], r) <= TOL[prec], (seed, prec, k, vec, tiles, j)
            a.free()


@pytest.mark.parametrize("seed", range(1, 24, 4))
def test_random_structures_through_reference_entry_points(gpu, O, seed):
    """The same structures through the reference-named entry points on the
    HAVE_MAGMA container (its MAGMA block filled with sentinel bytes):
    lfcpy_batch_sorted_ + lfmtx_batch_ (zau += H zu), lfmtx_ and lfmtx_gpu_
    (HACApK_MAGMA.c:5-57, 79-227), then lfdel_batch_; each product within
    1e-12 of the oracle's."""
    hm = _leaf_set(gpu, seed)
    zu = gpu.uniform(90 + seed, hm.nd)
    z0 = gpu.uniform(95 + seed, hm.nd)
    ref = O.adot_body_lfmtx(hm, zu, z0.copy())
    prev = gpu.set_ref_layout("magma")
    try:
        lp = gpu.make_leafmtxp(hm, "magma", fill=0x5A)
        gpu.adot_body_lfcpy_batch_sorted(hm.nd, lp)
        zau = z0.copy()
        gpu.adot_body_lfmtx_batch(zau, lp, zu)
        assert _rel(zau, ref) <= TOL["f64"], seed
        for variant in ("lfmtx", "lfmtx_gpu"):
            zau = z0.copy()
            gpu.adot_body_lfmtx_host(zau, lp, zu, variant)
            assert _rel(zau, ref) <= TOL["f64"], (seed, variant)
        gpu.adot_body_lfdel_batch(lp)
    finally:
        gpu.set_ref_layout(prev)


@pytest.mark.parametrize("prec", ["f64", "bf16_all"])
def test_random_structures_as_instances_of_one_arena(gpu, O, prec):
    """Five random structures (ragged sizes, rank-0 leaves, a row with no
    leaf) as the instances of one multi-instance arena: every instance's rows
    of the apply and of the H-matmat (k = 3 on the vector units, k = 20 on
    MFMA) equal the oracle's apply of that instance alone."""
    hms = [_leaf_set(gpu, seed) for seed in (3, 5, 8, 13, 21)]
    m = gpu.MultiArena(hms, prec)
    assert m.offsets[-1] == sum(hm.nd for hm in hms)
    x = gpu.uniform(123, m.nd)
    y = m.matvec(x)
    for i, hm in enumerate(hms):
        lo, hi = m.offsets[i], m.offsets[i + 1]
        assert _rel(y[lo:hi], O.adot_body_lfmtx(hm, x[lo:hi])) <= TOL[prec], (i, prec)
    for k in (3, 20):
        X = np.stack([gpu.uniform(600 + j, m.nd) for j in range(k)], axis=1)
        Y = m.matmat(X)
        for i, hm in enumerate(hms):
            lo, hi = m.offsets[i], m.offsets[i + 1]
            for j in (0, k - 1):
                assert _rel(Y[lo:hi, j], O.adot_body_lfmtx(hm, X[lo:hi, j])) <= TOL[prec], (i, k, j)
    m.free()


@pytest.mark.parametrize("seed", [2, 9, 15])
def test_random_structures_arena_file_round_trip(gpu, O, seed, tmp_path):
    """An arena of a random structure saved to an arena file and loaded back
    (no leaves, no repack) applies bit for bit as the original does, in every
    storage precision, and within the precision's bar of the oracle."""
    hm = _leaf_set(gpu, seed)
    x = gpu.uniform(140 + seed, hm.nd)
    y_ref = O.adot_body_lfmtx(hm, x)
    for prec in TOL:
        a = gpu.Arena(hm, prec, tuning=TUNINGS[2] if seed % 2 else None)
        path = str(tmp_path / ("%s_%d.hmv" % (prec, seed)))
        a.save(path)
        b = gpu.Arena.load(path)
        assert b.nd == hm.nd and b.precision == prec
        ya, yb = a.matvec(x), b.matvec(x)
        assert np.array_equal(ya, yb), (seed, prec)
        assert _rel(yb, y_ref) <= TOL[prec], (seed, prec)
        a.free()
        b.free()
