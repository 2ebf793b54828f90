"""The C-ABI libraries load and export every function their headers declare;
without a GPU every compute entry point fails loudly (no CPU fallback)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([a-z_][a-z0-9_]*)\s*\(", src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


@pytest.mark.parametrize("header,lib", [("hmv.h", "libhmv.so"), ("hmvgen.h", "libhmvgen.so")])
def test_every_declared_symbol_is_exported(H, header, lib):
    names = _declared(header)
    assert len(names) > 8
    L = C.CDLL(os.path.join(ROOT, "hacapk-magma_amd", "lib", lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_named_symbols_present(H):
    L = H.hmv_lib()
    for n in ("c_hacapk_adot_body_lfcpy_batch_sorted_", "c_hacapk_adot_body_lfmtx_batch_",
              "c_hacapk_adot_body_lfdel_batch_", "c_hacapk_bicgstab_cax_lfmtx_gpu_"):
        assert hasattr(L, n)


def _have_gpu(H):
    n = C.c_int(0)
    return os.path.exists("/dev/kfd") and H.hmv_lib().hmv_init(0) == 0


def test_no_cpu_fallback(H, sphere3):
    if _have_gpu(H):
        pytest.skip("a GPU is present")
    with pytest.raises(H.HmvError, match="no HIP device"):
        H.Arena(sphere3, "f64")
    with pytest.raises(H.HmvError):
        H.DeviceBuffer(64)
