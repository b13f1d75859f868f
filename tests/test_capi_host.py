"""CPU checks of the C ABI: the library loads, exports every function that
include/pointint_hip.h declares, and its host utilities agree with the oracle.
No device calls."""
import ctypes
import os
import re

import numpy as np
import pytest

from pointint_amd import _native as N
from pointint_amd import core

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "pointint_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pii_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(N.LIB_PATH)
    names = header_functions()
    assert len(names) >= 20
    for nm in names:
        assert hasattr(lib, nm), nm
    # the Python binding covers exactly the declared surface
    assert sorted(N.SIGNATURES) == names


def test_gauss_legendre_and_kgrid_match_oracle(oracle):
    for n in (1, 7, 32, 64, 96):
        x, w = core.gauss_legendre(n, 0.0, 0.5 * np.pi)
        xo, wo = oracle.gauss_legendre(n, 0.0, 0.5 * np.pi)
        assert np.allclose(x, xo, rtol=0, atol=2e-15)
        assert np.allclose(w, wo, rtol=1e-13, atol=0)
    d, w = core.load_lebedev(core.lebedev_path(29))
    do, wo = oracle.load_lebedev(core.lebedev_path(29))
    assert np.array_equal(d, do) and np.array_equal(w, wo)
    dirs, kmag, wq = core.build_kgrid_arrays(d, w, 32)
    od, ok, ow = oracle.build_kgrid(do, wo, 32)
    assert np.array_equal(dirs, od)
    assert np.allclose(kmag, ok, rtol=1e-14, atol=0)
    assert np.allclose(wq, ow, rtol=1e-12, atol=0)


def test_errors_surface_as_status_codes():
    lib = N.load()
    out = ctypes.c_int()
    rc = lib.pii_lebedev_load(b"/nonexistent/rule.txt", None, None, 0, ctypes.byref(out))
    assert rc == N.PII_ERR_IO
    assert b"nonexistent" in lib.pii_last_error()
    with pytest.raises(N.PiiError):
        core.load_lebedev("/nonexistent/rule.txt")
    assert lib.pii_gauss_legendre(0, 0.0, 1.0, None, None) == N.PII_ERR_INVALID


def test_matrix_entries_reject_null_arguments():
    """Argument checks run before any device call: NULL context / grid or no
    requested output give PII_ERR_INVALID with a message naming the entry."""
    lib = N.load()
    one = (ctypes.c_int * 1)(12)
    rc = lib.pii_volume_matrix_derivatives(None, 1, None, one, None, one, None, 0, None, None,
                                           None)
    assert rc == N.PII_ERR_INVALID
    assert b"pii_volume_matrix_derivatives" in lib.pii_last_error()
    rc = lib.pii_volume_matrix_hessian(None, 1, None, one, None, one, None, 0, None, None)
    assert rc == N.PII_ERR_INVALID
    assert b"pii_volume_matrix_hessian" in lib.pii_last_error()
