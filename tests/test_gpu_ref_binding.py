"""The reference-named entry points on the struct the reference's GPU
builds pass (st_HACApK_leafmtxp with its MAGMA block,
m_HACApK_base.f90:64-235), from ctypes and from a Fortran program.

Bar: FP64 products within 1e-12 of the oracle's C-loop apply
(HACApK_MAGMA.c:5-57); the BiCGSTAB shim within one step of the oracle's
BiCGSTAB (m_HACApK_solve.f90:342-449) and its solution within 1e-6
relative; the MAGMA block unchanged except mpi_rank, which the engine sets
to 0 as the reference's lfcpy sets the MPI rank (HACApK_BATCH_SORT.c:98).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
TOL64 = 1e-12


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _magma_block(lp):
    """Bytes of the container between st_lf_stride and st_lf, with mpi_rank
    cut out."""
    raw = C.string_at(C.addressof(lp), C.sizeof(lp))
    R = type(lp)
    return raw[20:R.mpi_rank.offset] + raw[R.mpi_rank.offset + 4:R.st_lf.offset]


def test_ctypes_magma_layout_with_sentinels(gpu, O, sphere3):
    """lfcpy_batch_sorted_, lfmtx_batch_, bicgstab_cax_lfmtx_gpu_ and
    lfdel_batch_ on a 544-byte container whose MAGMA block holds 0xA5 bytes
    (the caller's own MAGMA state)."""
    hm = sphere3
    lp = gpu.make_leafmtxp(hm, "magma", fill=0xA5)
    before = _magma_block(lp)
    assert lp.mpi_rank != 0
    gpu.adot_body_lfcpy_batch_sorted(hm.nd, lp)
    assert lp.mpi_rank == 0
    zu = gpu.uniform(21, hm.nd)
    zau = np.full(hm.nd, 0.5)
    gpu.adot_body_lfmtx_batch(zau, lp, zu)
    ref = O.adot_body_lfmtx(hm, zu, np.full(hm.nd, 0.5))
    print("lfmtx_batch_ rel err %.3e" % _rel(zau, ref), flush=True)
    assert _rel(zau, ref) <= TOL64
    b = np.ones(hm.nd)
    param = np.zeros(100)
    param[82], param[90] = 300, 1e-8
    u = np.zeros(hm.nd)
    nstp, lrtrn = gpu.bicgstab_cax_lfmtx_gpu(lp, u, b, param)
    u_ref, nstp_ref, _ = O.bicgstab(hm, b, None, mstep=300, eps=1e-8)
    print("bicgstab_cax_lfmtx_gpu_ steps %d (oracle %d), |u-u_ref|/|u_ref| %.2e"
          % (nstp, nstp_ref, _rel(u, u_ref)), flush=True)
    assert lrtrn == 0 and abs(nstp - nstp_ref) <= 1 and _rel(u, u_ref) <= 1e-6
    gpu.adot_body_lfdel_batch(lp)
    assert _magma_block(lp) == before


def test_ctypes_plain_layout(gpu, O, sphere3):
    """A build without HAVE_MAGMA reaches c_hacapk_adot_body_lfmtx_ only
    (m_HACApK_solve.f90:102), with the 32-byte container; the layout is
    selected for the process and restored afterwards."""
    hm = sphere3
    lp = gpu.make_leafmtxp(hm, "plain")
    prev = gpu.set_ref_layout("plain")
    try:
        zu = gpu.uniform(22, hm.nd)
        zau = np.zeros(hm.nd)
        gpu.adot_body_lfmtx_host(zau, lp, zu, "lfmtx")
        assert _rel(zau, O.adot_body_lfmtx(hm, zu)) <= TOL64
        gpu.adot_body_lfdel_batch(lp)
    finally:
        gpu.set_ref_layout(prev)


def _write_case(path, hm, zu, b, mstep, eps, zau0):
    """The driver's input file (see oracle/ref_driver.f90)."""
    recs = hm.leaves()
    with open(path, "wb") as f:
        f.write(np.array([hm.nd, hm.nlf, mstep], np.int32).tobytes())
        f.write(np.array([eps, zau0], np.float64).tobytes())
        for r in recs:
            f.write(np.array([r.ltmtx, r.kt, r.nstrtl, r.ndl, r.nstrtt, r.ndt], np.int32).tobytes())
            n1 = r.ndt * r.kt if r.ltmtx == 1 else r.ndt * r.ndl
            a1 = np.ctypeslib.as_array(r.a1, shape=(n1,)) if n1 else np.zeros(0)
            f.write(np.ascontiguousarray(a1, np.float64).tobytes())
            if r.ltmtx == 1 and r.kt > 0:
                addr = C.cast(r.a1, C.c_void_p).value + r.a1size
                a2 = np.ctypeslib.as_array((C.c_double * (r.ndl * r.kt)).from_address(addr))
                f.write(np.ascontiguousarray(a2, np.float64).tobytes())
        f.write(np.ascontiguousarray(zu, np.float64).tobytes())
        f.write(np.ascontiguousarray(b, np.float64).tobytes())


@pytest.mark.parametrize("case", ["sphere3", "sphere4"])
def test_fortran_driver_calling_sequence(gpu, O, case, request, tmp_path):
    """oracle/_ref/ref_driver: the reference's own type declarations (a
    HAVE_MAGMA + ISO_C_BINDING build), leaves allocated one by one in
    Fortran, st_lf_stride and a1size measured as m_HACApK_solve.f90:632-650
    does, then magma_init, lfcpy_batch_sorted, lfmtx_batch,
    bicgstab_cax_lfmtx_gpu, lfdel_batch, lfmtx, lfdel_batch, magma_finalize
    through implicit interfaces."""
    if not os.path.exists(DRIVER):
        pytest.skip("oracle/_ref/ref_driver not built (needs /root/reference and flang)")
    hm = request.getfixturevalue(case)
    zu = gpu.uniform(23, hm.nd)
    b = np.ones(hm.nd)
    mstep, eps, zau0 = 400, 1e-8, 0.25
    fin, fout = tmp_path / "case.bin", tmp_path / "out.bin"
    _write_case(fin, hm, zu, b, mstep, eps, zau0)
    print("running", DRIVER, case, "nd", hm.nd, "nlf", hm.nlf, flush=True)
    p = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    raw = fout.read_bytes()
    nd = hm.nd
    v = np.frombuffer(raw[:24 * nd], np.float64)
    zau, zau2, u = v[:nd], v[nd:2 * nd], v[2 * nd:]
    nstp, lrtrn, mpi_rank, kept, stride, _ = np.frombuffer(raw[24 * nd:], np.int32)
    ref = O.adot_body_lfmtx(hm, zu, np.full(nd, zau0))
    print("lfmtx_batch rel err %.3e, lfmtx rel err %.3e" % (_rel(zau, ref), _rel(zau2, ref)),
          flush=True)
    assert _rel(zau, ref) <= TOL64 and _rel(zau2, ref) <= TOL64
    u_ref, nstp_ref, _ = O.bicgstab(hm, b, None, mstep=mstep, eps=eps)
    print("Fortran-driven BiCGSTAB steps %d (oracle %d), |u-u_ref|/|u_ref| %.2e, stride %d"
          % (nstp, nstp_ref, _rel(u, u_ref), stride), flush=True)
    assert lrtrn == 0 and abs(nstp - nstp_ref) <= 1 and _rel(u, u_ref) <= 1e-6
    assert mpi_rank == 0 and kept == 1
    assert stride > 48  # Fortran records carry pointer descriptors
