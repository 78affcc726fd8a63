"""The C ABI library loads and exports every symbol include/gpurt.h declares; argument
errors are reported as status codes before any device work (no GPU needed)."""
import ctypes
import os
import re

import pytest

from gpurt import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gpurt.h")).read()
    return sorted(set(re.findall(r"GPURT_API\s+[\w\s\*]+?\b(gpurt_\w+)\s*\(", text)))


def test_header_declares_the_binding_table():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_struct_layouts_match_header():
    assert ctypes.sizeof(_lib.BuildOpts) == 16
    assert ctypes.sizeof(_lib.QueryOpts) == 16
    assert ctypes.sizeof(_lib.BvhInfo) == 56


def test_null_arguments_return_invalid_without_device_work():
    lib = _lib.load()
    out = ctypes.c_void_p()
    rc = lib.gpurt_bvh_build(None, None, 12, 0, None, 0, None, None, ctypes.byref(out))
    assert rc == _lib.GPURT_E_INVALID
    assert b"null" in lib.gpurt_last_error()
    assert lib.gpurt_closest_point(None, None, 1, None, None, None, None, None, None) == _lib.GPURT_E_INVALID
    assert lib.gpurt_trace_closest(None, None, 1, None, None, None, None, None) == _lib.GPURT_E_INVALID
    assert lib.gpurt_trace_any(None, None, 1, None, None, None) == _lib.GPURT_E_INVALID
    assert lib.gpurt_closest_point_bwd(None, None, 1, None, None, None, None, None, None, None, None) == _lib.GPURT_E_INVALID
    assert lib.gpurt_trace_bwd(None, None, 1, None, None, None, None, None, None, None, None) == _lib.GPURT_E_INVALID
    assert lib.gpurt_bvh_get_info(None, None) == _lib.GPURT_E_INVALID
    assert lib.gpurt_ctx_create(0, None) == _lib.GPURT_E_INVALID
    lib.gpurt_bvh_destroy(None)
    lib.gpurt_ctx_destroy(None)


def test_check_raises_with_message():
    lib = _lib.load()
    lib.gpurt_bvh_get_info(None, None)
    with pytest.raises(_lib.GpurtError) as e:
        _lib.check(_lib.GPURT_E_INVALID)
    assert "GPURT_E_INVALID" in str(e.value)
