"""The arena repack (hmv_plan_create, host only) checked on the CPU.

A numpy emulation of the two kernel phases (tests/plan_emulator.py) walks the
same tables the HIP kernels walk, asserts every access is in bounds and
every t / y entry is written exactly once, and must reproduce the oracle's
apply: to rounding for FP64, to storage precision for FP32 / BF16.
"""
import numpy as np
import pytest

import plan_emulator as PE

TOL = {"f64": 1e-14, "f32": 1e-6, "bf16": 1e-2, "bf16_all": 1e-2}


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16", "bf16_all"])
@pytest.mark.parametrize("case", ["sphere3", "sphere4", "two_spheres4", "ref_mesh_sample"])
def test_plan_emulation_matches_oracle(H, O, request, case, prec):
    hm = request.getfixturevalue(case)
    T = H.plan_tables(hm, prec)
    x = H.uniform(9, hm.nd)
    t, y = PE.run(T, x, hm.nd)
    y_ref = O.adot_body_lfmtx(hm, x)
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < TOL[prec]


def test_plan_alignment_and_order(H, sphere4):
    for prec in ("f64", "bf16"):
        T = H.plan_tables(sphere4, prec)
        for w in (T["work1"], T["work2"]):
            assert np.all(w["off0"][w["ncol0"] > 0] % 256 == 0)
            assert np.all(w["off1"][w["ncol1"] > 0] % 256 == 0)
        # the dispatcher sends workgroup b to XCD b % 8: within each XCD's
        # stream workgroups come largest first, and the XCDs' phase-2 row
        # ranges are disjoint (each XCD owns a contiguous eighth of the rows)
        for w in (T["work1"], T["work2"]):
            es0 = np.where(w["dt0"] == 0, 8, np.where(w["dt0"] == 1, 4, 2))
            es1 = np.where(w["dt1"] == 0, 8, np.where(w["dt1"] == 1, 4, 2))
            size = w["rows_p"] * (w["ncol0"] * es0 + w["ncol1"] * es1)
            for k in range(8):
                assert np.all(np.diff(size[k::8]) <= 0)
        rows = T["work2"]["out_base"]
        spans = sorted((rows[k::8].min(), rows[k::8].max()) for k in range(8))
        assert all(a[1] < b[0] for a, b in zip(spans, spans[1:]))


def test_long_column_groups_are_chunked(H, O):
    """A wide low-rank block is cut into column chunks whose partial sums sit
    past the t slots and are summed by the finalize step; the emulated apply
    still equals the oracle."""
    m = H.merge(H.icosphere(4), H.icosphere(4, 1.0, (3.0, 0.0, 0.0)))
    hm = H.HMatrix(*m, H.baseline_params(aca_eps=1e-5))
    tune = H.default_tuning("f64")
    tune["p1_target_bytes"] = 96 * 1024  # chunks this mesh's widest groups
    T = H.plan_tables(hm, "f64", tuning=tune)
    fin = T["tfin"]
    assert len(fin) > 0 and np.all(fin["np"] >= 2)
    assert T["tpart_size"] == T["tsize"] + fin["np"].sum()
    # partial ranges tile [tsize, tpart_size) and each chunked slot appears once
    order = np.argsort(fin["base"])
    assert fin["base"][order][0] == T["tsize"]
    assert np.all(np.diff(fin["base"][order]) == fin["np"][order][:-1])
    assert len(np.unique(fin["slot"])) == len(fin)
    # inputs address final slots only
    zs = T["zseg"]
    trun = zs[zs["kind"] == 2]
    assert len(trun) > 0 and np.all(trun["src"] + trun["ncols"] <= T["tsize"])
    assert T["codes"].size == 0 or T["codes"].min() >= -T["tsize"]
    x = H.uniform(2, hm.nd)
    _, y = PE.run(T, x, hm.nd)
    y_ref = O.adot_body_lfmtx(hm, x)
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < 1e-14


@pytest.mark.parametrize("prec,tuning", [
    ("f64", {"p1_small_threads": 64, "p2_small_threads": 128}),
    ("f64", {"p1_target_bytes": 16384, "p1_row_tile": 32, "p2_max_rows": 16}),
    ("f32", {"p1_small_threads": 64, "p2_small_threads": 64, "p2_max_rows": 64}),
    ("bf16", {"p1_small_threads": -1, "p2_small_threads": 128, "p1_target_bytes": 200000}),
])
def test_tuned_plans(H, O, two_spheres4, prec, tuning):
    """Every tuning the autotuner can pick gives a repack whose emulated apply
    equals the oracle, with short blocks first in each list."""
    T = H.plan_tables(two_spheres4, prec, tuning)
    for k in (1, 2):
        key = "p%d_small_threads" % k
        if tuning.get(key, 0) in (64, 128):
            assert T["nt%d_small" % k] == tuning[key] and T["nwork%d_small" % k] > 0
            w = T["work%d" % k]
            vr = {"f64": 1, "f32": 2, "bf16": 4 if k == 1 else 2}[prec]
            assert np.all(w["rows_p"][:T["nwork%d_small" % k]] // vr <= tuning[key])
        elif tuning.get(key, 0) == -1:
            assert T["nwork%d_small" % k] == 0
    x = H.uniform(4, two_spheres4.nd)
    _, y = PE.run(T, x, two_spheres4.nd)
    y_ref = O.adot_body_lfmtx(two_spheres4, x)
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < TOL[prec]


@pytest.mark.parametrize("prec", ["f64", "bf16"])
def test_input_runs_and_code_fallback(H, O, prec):
    """Phase-2 blocks take their inputs as runs (x ranges of dense leaves, t
    ranges of row groups) up to the run table's size, and a per-column code
    table beyond it; both give the oracle's apply."""
    import synthetic_leaves as SL
    hm = SL.many_runs(H)
    T = H.plan_tables(hm, prec)
    zs, w2 = T["zseg"], T["work2"]
    nz = w2["zseg_end"] - w2["zseg_begin"]
    kinds = [set(zs["kind"][b:e]) for b, e in zip(w2["zseg_begin"], w2["zseg_end"])]
    assert any(k == {1} for k in kinds)                 # code table (101 runs)
    assert any(n > 16 and 1 not in k for n, k in zip(nz, kinds))  # 31 runs tabled
    assert any(2 in k for k in kinds)                   # t runs
    assert T["codes"].size == 100 + 5  # rows 0-15 only: 100 x columns, 5 t slots
    x = H.uniform(5, hm.nd)
    _, y = PE.run(T, x, hm.nd)
    y_ref = O.adot_body_lfmtx(hm, x)
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < TOL[prec]


def test_default_tuning(H):
    assert H.default_tuning("f64")["p1_small_threads"] == -1
    assert H.default_tuning("bf16")["p2_small_threads"] == 64


def test_plan_rejects_bad_input(H, sphere3):
    import ctypes as C
    p = C.c_void_p()
    rc = H.hmv_lib().hmv_plan_create(sphere3.nd - 5, sphere3.nlf, sphere3.leaves_ptr, sphere3.stride, 0,
                                     C.byref(p))
    assert rc != 0 and "outside" in H.hmv_lib().hmv_last_error().decode()
    rc = H.hmv_lib().hmv_plan_create(sphere3.nd, sphere3.nlf, sphere3.leaves_ptr, sphere3.stride, 7,
                                     C.byref(p))
    assert rc != 0 and "precision" in H.hmv_lib().hmv_last_error().decode()
    with pytest.raises(H.HmvError, match="small-workgroup"):
        H.plan_tables(sphere3, "f64", {"p1_small_threads": 96})
    with pytest.raises(H.HmvError, match="tuning"):
        H.plan_tables(sphere3, "f64", {"p2_max_rows": 1000})
