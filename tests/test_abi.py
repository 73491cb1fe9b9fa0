"""The C-ABI library loads and exports every function include/zkhip.h
declares, and the ctypes binding covers all of them (no device calls)."""
import ctypes as ct
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "zkhip.h")
LIB = os.path.join(ROOT, "pyzkp_amd", "libzkhip.so")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(zk_\w+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("zk_init", "zk_ntt", "zk_msm_g1", "zk_last_error", "zk_fixed_bases_create",
                 "zk_msm_g1_fixed_batch"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzkhip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ct.CDLL(LIB)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from pyzkp_amd._native import SIGNATURES
    assert sorted(SIGNATURES) == _declared()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzkhip.so not built")
def test_no_device_errors_cleanly():
    from pyzkp_amd._native import load_library
    lib = load_library()
    n = ct.c_int(-1)
    assert lib.zk_device_count(ct.byref(n)) == 0
    if n.value == 0:
        h = ct.c_void_p()
        assert lib.zk_init(0, ct.byref(h)) == 3   # ZK_ERR_NO_DEVICE
        assert b"no HIP device" in lib.zk_last_error(None)
    assert lib.zk_version() >= 100
