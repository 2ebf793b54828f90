"""Pins the host input generator (libhmvgen) against the reference.

* entry function, ACA helpers and ACA+ itself: bit-for-bit against the
  reference's own Fortran sources compiled with flang (oracle/_ref), plus
  closed forms;
* cluster tree / block tree: the counts of the reference's recorded run on
  input_10ts.pbs (ivb/job_cpu.sh.o78632, ivb/log0000.txt), and invariants;
* ACA+: error against the dense block within the tolerance.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
D = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(D)


@pytest.fixture(scope="module")
def ref(O):
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return R


def test_face_integral_bitexact_vs_reference(H, ref):
    G = H.gen_lib()
    rng = np.random.default_rng(1)
    for k in range(3000):
        xs, ys, zs = rng.normal(size=3), rng.normal(size=3), rng.normal(size=3)
        if k % 4 == 0:
            p = np.array([xs.mean(), ys.mean(), zs.mean()])   # self term
        else:
            p = rng.normal(size=3) * (0.05 if k % 3 == 0 else 5.0)
        a = ref.ref_face_integral(_p(xs), _p(ys), _p(zs), *p)
        b = G.hmvgen_face_integral(0, _p(xs), _p(ys), _p(zs), *p)
        assert a == b, (k, a, b)


def test_aca_helpers_bitexact_vs_reference(H, ref):
    G = H.gen_lib()
    rng = np.random.default_rng(2)
    for k in range(300):
        n = int(rng.integers(1, 60))
        za = rng.standard_normal(n) * 10.0 ** rng.integers(-40, 40)
        if k % 5 == 0:
            za[rng.integers(0, n)] = 0.0
        assert ref.ref_unrm_d(n, _p(za)) == G.hmvgen_util_unrm(n, _p(za))
        zr, ir = C.c_double(), C.c_int()
        zg = C.c_double()
        ref.ref_maxabsvalloc_d(n, _p(za), C.byref(zr), C.byref(ir))
        assert G.hmvgen_util_maxabs(n, _p(za), C.byref(zg)) == ir.value and zg.value == zr.value
        ref.ref_minabsvalloc_d(n, _p(za), C.byref(zr), C.byref(ir))
        assert G.hmvgen_util_minabs(n, _p(za), C.byref(zg)) == ir.value and zg.value == zr.value


def test_unrm_is_the_two_norm(H):
    rng = np.random.default_rng(3)
    G = H.gen_lib()
    for _ in range(50):
        za = rng.standard_normal(int(rng.integers(1, 200)))
        assert math.isclose(G.hmvgen_util_unrm(len(za), _p(za)), np.linalg.norm(za), rel_tol=1e-14)


def test_face_integral_closed_forms(H):
    """Self potential of an equilateral triangle at its centroid is
    sqrt(3) a ln(2+sqrt(3)); far away the potential tends to area/r."""
    G = H.gen_lib()
    a = 0.7
    xs = np.array([0.0, a, a / 2]); ys = np.array([0.0, 0.0, a * math.sqrt(3) / 2]); zs = np.zeros(3)
    c = (xs.mean(), ys.mean(), 0.0)
    v = G.hmvgen_face_integral(1, _p(xs), _p(ys), _p(zs), *c) * 4 * math.pi
    assert math.isclose(v, math.sqrt(3) * a * math.log(2 + math.sqrt(3)), rel_tol=1e-13)
    area = a * a * math.sqrt(3) / 4
    r = 1000.0
    v = G.hmvgen_face_integral(1, _p(xs), _p(ys), _p(zs), c[0], c[1], r) * 4 * math.pi
    assert math.isclose(v, area / r, rel_tol=1e-6)
    # ppohBEM mode scales by 1/(4 pi eps0) with the single-precision literals
    w = G.hmvgen_face_integral(0, _p(xs), _p(ys), _p(zs), c[0], c[1], r)
    pi_f = float(np.float32(3.141592653589793))
    eps_f = float(np.float32(8.854187818)) * 1e-12
    assert math.isclose(w, area / r / (4 * pi_f * eps_f), rel_tol=1e-6)


def test_recorded_run_structure(H):
    """Cluster and block counts of the reference's recorded run (leaf 15, eta 2,
    ACA+ 1e-6) on input_10ts.pbs; these do not depend on ACA."""
    rec = json.load(open(os.path.join(HERE, "golden", "recorded_run_10ts.json")))
    d = np.load(os.path.join(HERE, "golden", "input_10ts_mesh.npz"))
    hm = H.HMatrix(d["nodes"], d["faces"], H.HParams(leaf_size=15, aca_eps=1e-6, kernel="ppohbem"))
    s = hm.stats
    assert s["nclst"] == rec["nclst"]
    assert (s["nlfkt"], s["nlfdense"], s["nhblocks"]) == (rec["nlfkt"], rec["nlfdense"], rec["nhblocks"])
    assert s["ncpc"] == rec["ncpc"]
    assert s["ktmax"] == rec["ktmax"]


def test_read_pbf_matches_fixture(H):
    path = "/root/reference/bem-bb_inputs/input_1ts.pbs"
    if not os.path.exists(path):
        pytest.skip("reference mesh files not present")
    nodes, faces, para = H.read_pbf(path)
    d = np.load(os.path.join(HERE, "golden", "input_1ts_mesh.npz"))
    assert np.array_equal(nodes, d["nodes"]) and np.array_equal(faces, d["faces"])
    assert np.array_equal(para, d["dble_para"])


def test_icosphere(H):
    nodes, faces = H.icosphere(4, 2.0, (1.0, 0.0, 0.0))
    assert faces.shape == (5120, 3) and nodes.shape == (2562, 3)
    assert np.allclose(np.linalg.norm(nodes - [1, 0, 0], axis=1), 2.0)
    assert nodes.shape[0] - 3 * faces.shape[0] // 2 + faces.shape[0] == 2   # Euler characteristic


def test_leaves_tile_matrix_and_admissibility(H, sphere3):
    hm = sphere3
    nd = hm.nd
    t = hm.leaf_table()
    cover = np.zeros((nd, nd), np.int8)
    for lt, kt, r0, ndl, c0, ndt in t:
        cover[r0 - 1:r0 - 1 + ndl, c0 - 1:c0 - 1 + ndt] += 1
    assert cover.min() == 1 and cover.max() == 1
    order = np.lexsort((t[:, 4], t[:, 2]))
    assert np.array_equal(order, np.arange(len(t)))          # sorted by (row, column)
    nleaf = hm.params["leaf_size"] + 1
    lr = t[t[:, 0] == 1]
    assert np.all(lr[:, 3] >= nleaf) and np.all(lr[:, 5] >= nleaf)
    assert sorted(hm.lod) == list(range(1, nd + 1))


def test_aca_error_within_tolerance(H):
    nodes, faces = H.icosphere(3)
    for eps in (1e-4, 1e-6):
        hm = H.HMatrix(nodes, faces, H.baseline_params(aca_eps=eps))
        errs = []
        for l in hm.leaves():
            if l.ltmtx != 1:
                continue
            rows = hm.lod[l.nstrtl - 1:l.nstrtl - 1 + l.ndl]
            cols = hm.lod[l.nstrtt - 1:l.nstrtt - 1 + l.ndt]
            A = np.array([[hm.entry(int(i), int(j)) for j in cols] for i in rows])
            V = np.ctypeslib.as_array(l.a1, shape=(l.kt * l.ndt,)).reshape(l.kt, l.ndt)
            U = np.ctypeslib.as_array(l.a2, shape=(l.kt * l.ndl,)).reshape(l.kt, l.ndl)
            errs.append(np.linalg.norm(A - U.T @ V) / np.linalg.norm(A))
            if len(errs) >= 60:
                break
        assert np.median(errs) < eps and max(errs) < 20 * eps, (eps, np.median(errs), max(errs))


def _ref_param(hm):
    """The reference's param() slots that HACApK_acaplus reads (1-based)."""
    p = hm.params.as_dict()
    param = np.zeros(100)
    param[60 - 1] = p["aca_kind"]
    param[61 - 1] = p["aca_norm"]
    param[63 - 1] = p["kmax"]
    param[64 - 1] = p["kmin"]
    param[71 - 1] = p["aca_eps"]
    param[72 - 1] = p["aca_eps_factor"]
    return param


@pytest.mark.parametrize("mesh,eps", [("input_10ts", 1e-6), ("icosphere3", 1e-4),
                                      ("input_1ts", 1e-5), ("icosphere4", 1e-8)])
def test_acaplus_bitexact_vs_reference(H, ref, mesh, eps):
    """Every low-rank leaf of the host fill equals the reference's own
    HACApK_acaplus (m_HACApK_base.f90:783-975, compiled from its source by
    oracle/ref_build.sh) on the same block: the same rank and the same U and
    V factors bit for bit.  The input_10ts mesh with leaf 15 / eta 2 / ACA+
    1e-6 is the configuration of the reference's recorded run."""
    if not hasattr(ref, "ref_acaplus"):
        pytest.skip("oracle/_ref built without the ACA+ routine")
    if mesh.startswith("input_"):
        d = np.load(os.path.join(HERE, "golden", mesh + "_mesh.npz"))
        nodes, faces = d["nodes"], d["faces"]
        hm = H.HMatrix(nodes, faces, H.HParams(leaf_size=15, aca_eps=eps, kernel="ppohbem"))
    else:
        nodes, faces = H.icosphere(int(mesh[-1]))
        hm = H.HMatrix(nodes, faces, H.baseline_params(aca_eps=eps, kernel="ppohbem"))
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    faces = np.ascontiguousarray(faces, dtype=np.int32)
    lod = np.ascontiguousarray(hm.lod, dtype=np.int32)
    param = _ref_param(hm)
    kmax = int(param[62])
    I = C.POINTER(C.c_int)
    nlr = 0
    kts = []
    for l in hm.leaves():
        if l.ltmtx != 1:
            continue
        zaa = np.zeros((kmax, l.ndl))
        zab = np.zeros((kmax, l.ndt))
        kt = ref.ref_acaplus(nodes.shape[0], _p(nodes), faces.shape[0], faces.ctypes.data_as(I),
                             lod.ctypes.data_as(I), l.ndl, l.ndt, l.nstrtl, l.nstrtt, _p(param),
                             _p(zaa), _p(zab))
        assert kt == l.kt, (l.nstrtl, l.nstrtt, kt, l.kt)
        V = np.ctypeslib.as_array(l.a1, shape=(l.kt * l.ndt,)).reshape(l.kt, l.ndt)
        U = np.ctypeslib.as_array(l.a2, shape=(l.kt * l.ndl,)).reshape(l.kt, l.ndl)
        assert np.array_equal(zab[:kt], V) and np.array_equal(zaa[:kt], U), (l.nstrtl, l.nstrtt)
        nlr += 1
        kts.append(kt)
    assert nlr == hm.stats["nlfkt"] and nlr > 0
    print(mesh, "low-rank leaves", nlr, "ranks", min(kts), "-", max(kts), flush=True)


def test_fill_independent_of_thread_count(H):
    nodes, faces = H.icosphere(3)
    a = H.HMatrix(nodes, faces, H.baseline_params(aca_eps=1e-5, nthreads=1))
    b = H.HMatrix(nodes, faces, H.baseline_params(aca_eps=1e-5, nthreads=5))
    assert np.array_equal(a.leaf_table(), b.leaf_table())
    for la, lb in zip(a.leaves(), b.leaves()):
        n = la.a1size // 8 + (la.kt * la.ndl if la.ltmtx == 1 else 0)
        assert np.array_equal(np.ctypeslib.as_array(la.a1, shape=(n,)), np.ctypeslib.as_array(lb.a1, shape=(n,)))


def test_uniform_stream(H):
    a, b = H.uniform(7, 1000), H.uniform(7, 1000)
    assert np.array_equal(a, b) and a.min() >= -1 and a.max() < 1 and abs(a.mean()) < 0.1
