"""Random block structures through the arena repack, checked on the CPU.

Seeded random tilings of the N x N matrix (ragged block sizes from 1, dense
and low-rank leaves, ranks from 0, rows and columns covered by no leaf) go
through hmv_plan_create and the numpy emulation of the two kernel phases
(tests/plan_emulator.py, which asserts every access is in bounds and every
t / y entry is written once); the emulated apply must equal the oracle's
(oracle_adot_body_lfmtx, HACApK_MAGMA.c:5-57), to rounding in FP64 and to
storage precision otherwise.  Tight tunings force column chunks, short row
tiles and both workgroup sizes.
"""
import numpy as np
import pytest

import plan_emulator as PE
import synthetic_leaves as SL

TOL = {"f64": 1e-14, "f32": 1e-6, "bf16": 1e-2, "bf16_all": 1e-2}

TUNINGS = [
    None,
    {"p1_target_bytes": 4096, "p1_row_tile": 16, "p2_max_rows": 16,
     "p1_small_threads": 64, "p2_small_threads": 64},
    {"p1_target_bytes": 2048, "p1_row_tile": 8, "p2_max_rows": 32,
     "p1_small_threads": -1, "p2_small_threads": 128},
]


def _tiling(rng, r0, nr, c0, nc, depth, out):
    """Quadtree split at random points; a leaf is dense, low rank (rank 0 to
    12, capped by the block) or absent (5%)."""
    if depth > 0 and nr > 4 and nc > 4 and rng.random() < 0.75:
        rs, cs = int(rng.integers(1, nr)), int(rng.integers(1, nc))
        for a, b in ((r0, rs), (r0 + rs, nr - rs)):
            for c, d in ((c0, cs), (c0 + cs, nc - cs)):
                _tiling(rng, a, b, c, d, depth - 1, out)
        return
    u = rng.random()
    if u < 0.05:
        return
    if u < 0.5 or min(nr, nc) < 2:
        out.append(("dense", r0, nr, c0, nc, 0))
    else:
        out.append(("lowrank", r0, nr, c0, nc, int(rng.integers(0, min(nr, nc, 12) + 1))))


def _leaf_set(H, seed):
    rng = np.random.default_rng(1000 + seed)
    nd = int(rng.integers(40, 900))
    leaves = []
    _tiling(rng, 0, nd, 0, nd, int(rng.integers(2, 7)), leaves)
    if seed % 3 == 0:  # leave one row with no leaf at all (its y entry is 0)
        r = int(rng.integers(0, nd))
        leaves = [l for l in leaves if not l[1] <= r < l[1] + l[2]]
    return SL.LeafSet(H, nd, leaves, seed)


@pytest.mark.parametrize("seed", range(24))
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16", "bf16_all"])
def test_random_tilings(H, O, seed, prec):
    hm = _leaf_set(H, seed)
    tab = hm.leaf_table()
    x = H.uniform(50 + seed, hm.nd)
    y_ref = O.adot_body_lfmtx(hm, x)
    scale = max(np.linalg.norm(y_ref), 1e-300)
    for tune in TUNINGS:
        T = H.plan_tables(hm, prec, tune)
        _, y = PE.run(T, x, hm.nd)
        err = np.linalg.norm(y - y_ref) / scale
        assert err < TOL[prec], (seed, prec, tune, err, hm.nd, len(tab))
    # the tiling has what the test is for
    if seed == 0:
        assert (tab[:, 0] == 1).any() and (tab[:, 0] == 2).any()


def test_random_tilings_cover_edge_cases(H):
    """Across the seeds: rank-0 and rank-1 leaves, 1-row and 1-column
    blocks, and rows that no leaf covers all occur."""
    seen = set()
    for seed in range(24):
        hm = _leaf_set(H, seed)
        t = hm.leaf_table()
        if ((t[:, 0] == 1) & (t[:, 1] == 0)).any():
            seen.add("rank0")
        if ((t[:, 0] == 1) & (t[:, 1] == 1)).any():
            seen.add("rank1")
        if (t[:, 3] == 1).any():
            seen.add("one row")
        if (t[:, 5] == 1).any():
            seen.add("one column")
        covered = np.zeros(hm.nd, bool)
        for lt, kt, r0, ndl, c0, ndt in t:
            covered[r0 - 1:r0 - 1 + ndl] = True
        if not covered.all():
            seen.add("uncovered rows")
    assert seen == {"rank0", "rank1", "one row", "one column", "uncovered rows"}, seen
