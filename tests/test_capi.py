"""CPU checks of the boundary: libidrs_hip.so loads without a GPU and exports
every function include/idrs_capi.h declares; the ctypes binding covers them;
the library reports errors instead of crashing when no device is present."""
import ctypes
import os
import re

import numpy as np

import idrs_amd
from idrs_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "idrs_capi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(idrs_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_binding_covers_header():
    assert set(declared_functions()) == set(_lib.exported_symbols())


def test_opts_struct_layout():
    # the Fortran bind(C) type and the ctypes structure mirror the C layout
    assert ctypes.sizeof(_lib.idrs_opts_t) == 104
    assert _lib.idrs_opts_t.p_seed.offset == 88
    assert ctypes.sizeof(_lib.idrs_stats_t) == 40


def test_no_device_reports_error():
    lib = idrs_amd.load_library()
    if lib.idrs_device_count() > 0:
        return
    rp = np.array([1, 2], dtype=np.int32)
    ci = np.array([1], dtype=np.int32)
    try:
        idrs_amd.crs_format(1, 1, rp, ci, rval=np.ones(1))
    except idrs_amd.IdrsError as e:
        assert e.status == _lib.IDRS_ERR_HIP
    else:
        raise AssertionError("expected a HIP error without a device")


def test_argument_errors_before_device_use():
    lib = idrs_amd.load_library()
    h = ctypes.c_void_p()
    # row_ptr(1) must be 1
    rp = np.array([0, 1], dtype=np.int32)
    ci = np.array([1], dtype=np.int32)
    st = lib.idrs_crs_format_f64(1, 1, rp.ctypes.data, ci.ctypes.data, None, ctypes.byref(h))
    assert st == _lib.IDRS_ERR_ARG
    assert b"row_ptr" in lib.idrs_last_error()


def test_coefficients_without_device():
    """The polynomial coefficient entry points are host code."""
    lib = idrs_amd.load_library()
    foci = np.array([0.15, 0.0, 1.85, 0.0])
    out = np.zeros(2 * 9)
    assert lib.idrs_chebyshev_coeffs(8, foci.ctypes.data, 0, out.ctypes.data) == 0
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    assert np.array_equal(out[0::2], orc.chebyshev(8, np.array([0.15, 1.85])))
