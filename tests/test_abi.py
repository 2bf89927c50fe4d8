"""The C-ABI library loads and exports every entry point include/sbp.h
declares; argument validation that happens before any device work."""
import ctypes as C
import os
import re

import pytest

from sc_2022_amd import _capi as capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sbp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sbp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(capi.SIGNATURES), set(syms) ^ set(capi.SIGNATURES)


def test_version_and_status_strings():
    lib = capi.load()
    assert b"gfx950" in lib.sbp_version()
    for s in range(0, 11):
        assert lib.sbp_status_string(s)


def test_default_params():
    p = capi.default_params()
    assert (p.nu_pre, p.nu_post, p.coarse_n) == (2, 2, 17)
    assert abs(p.omega - 2 / 3) < 1e-15 and p.restrict_scale == 4.0
    with pytest.raises(KeyError):
        capi.default_params(nonsense=1)


def test_create_rejects_bad_order_and_size_before_device_work():
    lib = capi.load()
    h = C.c_void_p()
    prm = capi.default_params()
    assert lib.sbp_create(17, 3, C.byref(prm), C.byref(h)) == capi.SBP_ERR_INVALID_ORDER
    assert lib.sbp_create(17, 12, C.byref(prm), C.byref(h)) == capi.SBP_ERR_INVALID_ORDER
    # p = 6 / 8 / 10 closures need N >= 2 C = 18 / 24 / 32
    assert lib.sbp_create(17, 6, C.byref(prm), C.byref(h)) == capi.SBP_ERR_INVALID_SIZE
    assert lib.sbp_create(23, 8, C.byref(prm), C.byref(h)) == capi.SBP_ERR_INVALID_SIZE
    assert lib.sbp_create(31, 10, C.byref(prm), C.byref(h)) == capi.SBP_ERR_INVALID_SIZE
    assert lib.sbp_create(7, 2, C.byref(prm), C.byref(h)) == capi.SBP_ERR_INVALID_SIZE
    assert lib.sbp_create(11, 4, C.byref(prm), C.byref(h)) == capi.SBP_ERR_INVALID_SIZE
    bad = capi.default_params(omega=-1.0)
    assert lib.sbp_create(17, 2, C.byref(bad), C.byref(h)) == capi.SBP_ERR_INVALID_ARG
    bad = capi.default_params(smoother=5)
    assert lib.sbp_create(17, 2, C.byref(bad), C.byref(h)) == capi.SBP_ERR_INVALID_ARG
    assert lib.sbp_destroy(None) == capi.SBP_ERR_INVALID_ARG
