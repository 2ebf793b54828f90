"""The struct the reference-named entry points read is the one the
reference's Fortran passes.

flang compiles the reference's own declarations of st_HACApK_leafmtx and
st_HACApK_leafmtxp (m_HACApK_base.f90:54-235, preprocessed with and
without -DHAVE_MAGMA -DHAVE_MAGMA_BATCH -DISO_C_BINDING by
oracle/ref_build.sh) and prints their byte offsets; these tests compare
them with include/hmv.h (compiled here with gcc) and with the ctypes mirror
in the package.  The MAGMA build's layout is the one every GPU call site
of the reference passes (m_HACApK_use.f90:317-455, m_HACApK_solve.f90:652-741);
the plain one is what c_hacapk_adot_body_lfmtx_ receives from a build without
the MAGMA flags (m_HACApK_solve.f90:102).
"""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_OUT = os.path.join(ROOT, "oracle", "_ref")

# name in the Fortran declaration -> C expression (offsetof) in include/hmv.h
MAGMA_FIELDS = {
    "nd": "nd", "nlf": "nlf", "nlfkt": "nlfkt", "ktmax": "ktmax", "st_lf_stride": "st_lf_stride",
    "m": "m", "max_block": "max_block", "gflops": "gflops",
    "mtx1_gpu": "gpu_vectors[0]", "zbu_gpu": "gpu_vectors[5]",
    "num_batch": "num_batch", "transA": "transA",
    "d_A_array": "batch_tables[0]", "batch_size": "batch_tables[19]",
    "num_streamed": "num_streamed", "num_streamed_t": "num_streamed_t",
    "h_A_array_streamed": "streamed_tables[0]", "h_lda_streamed": "streamed_tables[5]",
    "d_A_mgpu": "mgpu_tables[0]", "iwork": "mgpu_tables[24]",
    "mpi_rank": "mpi_rank", "st_lf": "st_lf",
}
PLAIN_FIELDS = {"nd": "nd", "nlf": "nlf", "nlfkt": "nlfkt", "ktmax": "ktmax",
                "st_lf_stride": "st_lf_stride", "st_lf": "st_lf"}
LEAF_FIELDS = {"leaf_ltmtx": "ltmtx", "leaf_kt": "kt", "leaf_nstrtl": "nstrtl", "leaf_ndl": "ndl",
               "leaf_nstrtt": "nstrtt", "leaf_ndt": "ndt", "leaf_a1size": "a1size", "leaf_a1": "a1"}


def _fortran_layout(kind):
    path = os.path.join(REF_OUT, "ref_layout_%s.txt" % kind)
    if not os.path.exists(path):
        pytest.skip("%s not built (oracle/ref_build.sh needs /root/reference and flang)" % path)
    out = {}
    for line in open(path):
        name, val = line.split()
        out[name] = int(val)
    return out


@pytest.fixture(scope="module")
def c_offsets(tmp_path_factory):
    """offsetof/sizeof of the header's structs, from a gcc-compiled probe."""
    d = tmp_path_factory.mktemp("cprobe")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "hmv.h"', "int main(void) {"]
    for name, expr in MAGMA_FIELDS.items():
        lines.append('printf("magma %s %%zu\\n", offsetof(hmv_ref_leafmtxp, %s));' % (name, expr))
    for name, expr in PLAIN_FIELDS.items():
        lines.append('printf("plain %s %%zu\\n", offsetof(hmv_ref_leafmtxp_plain, %s));'
                     % (name, expr))
    for name, expr in LEAF_FIELDS.items():
        lines.append('printf("leaf %s %%zu\\n", offsetof(hmv_leaf, %s));' % (name, expr))
    lines.append('printf("magma size %zu\\n", sizeof(hmv_ref_leafmtxp));')
    lines.append('printf("plain size %zu\\n", sizeof(hmv_ref_leafmtxp_plain));')
    lines.append('printf("leaf size %zu\\n", sizeof(hmv_leaf));')
    lines.append("return 0; }")
    src = d / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True,
                               text=True).stdout.splitlines():
        kind, name, val = line.split()
        out[(kind, name)] = int(val)
    return out


def test_magma_build_layout_matches_reference_declaration(c_offsets):
    f = _fortran_layout("magma")
    for name in MAGMA_FIELDS:
        assert c_offsets[("magma", name)] == f[name], name
    # C reads only the first word of st_lf's descriptor (the base address)
    assert c_offsets[("magma", "size")] == f["st_lf"] + 8
    assert f["size"] >= f["st_lf"] + 8


def test_plain_build_layout_matches_reference_declaration(c_offsets):
    f = _fortran_layout("plain")
    for name in PLAIN_FIELDS:
        assert c_offsets[("plain", name)] == f[name], name


def test_leaf_record_matches_reference_declaration(c_offsets):
    """C's view of one st_HACApK_leafmtx: six ints, a1size, and a1's base
    address in the first word of its pointer descriptor; records sit
    st_lf_stride bytes apart (the Fortran record is larger than the C one,
    so the engine always steps by the caller's stride)."""
    for kind in ("magma", "plain"):
        f = _fortran_layout(kind)
        for name in LEAF_FIELDS:
            assert c_offsets[("leaf", name)] == f[name], name
        assert f["leaf_stride"] >= c_offsets[("leaf", "size")]


def test_ctypes_mirror_matches_header(H, c_offsets):
    for name, expr in MAGMA_FIELDS.items():
        base, _, idx = expr.partition("[")
        off = getattr(H.RefLeafmtxp, base).offset + (8 * int(idx[:-1]) if idx else 0)
        assert off == c_offsets[("magma", name)], name
    assert C.sizeof(H.RefLeafmtxp) == c_offsets[("magma", "size")]
    for name in PLAIN_FIELDS:
        assert getattr(H.RefLeafmtxpPlain, name).offset == c_offsets[("plain", name)]
    assert C.sizeof(H.RefLeafmtxpPlain) == c_offsets[("plain", "size")]


def test_layout_selection(H):
    """hmv_ref_set_layout switches the struct the entry points read and
    rejects unknown values; the default is the MAGMA build's."""
    L = H.hmv_lib()
    assert L.hmv_ref_get_layout() == 0
    assert H.set_ref_layout("plain") == "magma"
    assert L.hmv_ref_get_layout() == 1
    assert H.set_ref_layout("magma") == "plain"
    assert L.hmv_ref_set_layout(7) == -1
    assert b"unknown layout" in L.hmv_last_error()
    assert L.hmv_ref_get_layout() == 0


def test_magma_init_and_finalize_exported(H):
    L = H.hmv_lib()
    assert hasattr(L, "magma_init_") and hasattr(L, "magma_finalize_")
    L.magma_finalize_()  # does nothing, needs no device
