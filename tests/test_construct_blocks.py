"""The host part of the device construction: hmvgen_build_blocks gives the
same tree and leaf list as the full host build, without factors."""
import ctypes as C

import numpy as np


def test_blocks_match_full_build(H):
    nodes, faces = H.icosphere(3)
    p = H.baseline_params(aca_eps=1e-4)
    full = H.HMatrix(nodes, faces, p)
    L = H.gen_lib()
    h = C.c_void_p()
    nodes = np.ascontiguousarray(nodes, np.float64)
    faces = np.ascontiguousarray(faces, np.int32)
    assert L.hmvgen_build_blocks(len(nodes), nodes.ctypes.data_as(C.POINTER(C.c_double)), len(faces),
                                 faces.ctypes.data_as(C.POINTER(C.c_int)), C.byref(p._p),
                                 C.byref(h)) == 0
    try:
        nlf, stride = C.c_int(), C.c_int()
        ptr = L.hmvgen_leaves(h, C.byref(nlf), C.byref(stride))
        assert nlf.value == full.nlf and stride.value == full.stride
        recs = (H.LeafRec * nlf.value).from_address(ptr)
        blocks = np.array([(r.ltmtx, r.nstrtl, r.ndl, r.nstrtt, r.ndt) for r in recs])
        ref = full.leaf_table()[:, [0, 2, 3, 4, 5]]
        assert np.array_equal(blocks, ref)
        assert all(r.kt == 0 and not r.a1 for r in recs)
        lod = np.ctypeslib.as_array(L.hmvgen_lod(h), shape=(full.nd,))
        assert np.array_equal(lod, full.lod)
        area = np.ctypeslib.as_array(L.hmvgen_areas(h), shape=(full.nd,))
        tri = nodes[faces]
        want = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
        assert np.allclose(area, want, rtol=1e-14)
    finally:
        L.hmvgen_free(h)
