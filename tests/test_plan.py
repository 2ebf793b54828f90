"""The arena repack (hmv_plan_create, host only) checked on the CPU.

A numpy emulation of the two kernel phases (tests/plan_emulator.py) walks the
same tables the HIP kernels walk, asserts every access is in bounds and
every t / y entry is written exactly once, and must reproduce the oracle's
apply: to rounding for FP64, to storage precision for FP32 / BF16.
"""
import numpy as np
import pytest

import plan_emulator as PE

TOL = {"f64": 1e-14, "f32": 1e-6, "bf16": 1e-2}


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("case", ["sphere3", "sphere4", "two_spheres4", "ref_mesh_sample"])
def test_plan_emulation_matches_oracle(H, O, request, case, prec):
    hm = request.getfixturevalue(case)
    T = H.plan_tables(hm, prec)
    x = H.uniform(9, hm.nd)
    t, y = PE.run(T, x, hm.nd)
    y_ref = O.adot_body_lfmtx(hm, x)
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < TOL[prec]


def test_plan_alignment_and_order(H, sphere4):
    T = H.plan_tables(sphere4, "f64")
    for segs in (T["seg1"], T["seg2"]):
        assert np.all(segs["data_off"] % 16 == 0)
        assert np.all((segs["ld"] * 8) % 16 == 0)
    # workgroups are dispatched largest first
    for w, segs in ((T["work1"], T["seg1"]), (T["work2"], T["seg2"])):
        sizes = [int(sum(segs[i]["nrows"] * segs[i]["ld"] for i in range(a, b)))
                 for a, b in zip(w["seg_begin"], w["seg_end"])]
        assert sizes == sorted(sizes, reverse=True)


def test_plan_rejects_bad_input(H, sphere3):
    import ctypes as C
    p = C.c_void_p()
    rc = H.hmv_lib().hmv_plan_create(sphere3.nd - 5, sphere3.nlf, sphere3.leaves_ptr, sphere3.stride, 0,
                                     C.byref(p))
    assert rc != 0 and "outside" in H.hmv_lib().hmv_last_error().decode()
    rc = H.hmv_lib().hmv_plan_create(sphere3.nd, sphere3.nlf, sphere3.leaves_ptr, sphere3.stride, 7,
                                     C.byref(p))
    assert rc != 0 and "precision" in H.hmv_lib().hmv_last_error().decode()
