"""CPU checks of the boundary: libidrs_hip.so loads without a GPU and exports
every function include/idrs_capi.h declares; the ctypes binding covers them;
the library reports errors instead of crashing when no device is present."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

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


def _c_layout(struct, fields, tmp_path):
    """sizeof and offsetof of every field as the C compiler lays out the
    header's struct."""
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "idrs_capi.h"', "int main(void){",
             f'printf("%zu\\n", sizeof({struct}));']
    lines += [f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    return int(out[0]), [int(v) for v in out[1:]]


@pytest.mark.parametrize("struct", ["idrs_opts_t", "idrs_krylov_opts_t", "idrs_stats_t"])
def test_struct_layout_matches_header(struct, tmp_path):
    # the ctypes structures (and the Fortran bind(C) types, which list the
    # same fields in the same order) mirror the C layout field by field
    cls = getattr(_lib, struct)
    names = [f[0] for f in cls._fields_]
    size, offs = _c_layout(struct, names, tmp_path)
    assert ctypes.sizeof(cls) == size
    assert [getattr(cls, f).offset for f in names] == offs


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
