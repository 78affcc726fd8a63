"""The C ABI library loads and exports every symbol include/gpurt.h declares; argument
errors are reported as status codes before any device work (no GPU needed)."""
import ctypes
import os
import re
import subprocess
import tempfile

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


def _c_layout(struct, fields):
    """sizeof and member offsets of a struct of include/gpurt.h, from the C compiler."""
    body = "\n".join(f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields)
    src = (f'#include <stdio.h>\n#include "gpurt.h"\nint main(void){{{body}'
           f'printf("%zu\\n", sizeof({struct})); return 0;}}')
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "probe.c"), os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        vals = [int(x) for x in subprocess.run([exe], check=True, capture_output=True,
                                               text=True).stdout.split()]
    return vals[-1], vals[:-1]


@pytest.mark.parametrize("cls,struct", [(_lib.BuildOpts, "gpurt_build_opts"),
                                        (_lib.QueryOpts, "gpurt_query_opts"),
                                        (_lib.BvhInfo, "gpurt_bvh_info"),
                                        (_lib.BvhArrays, "gpurt_bvh_arrays"),
                                        (_lib.BatchMesh, "gpurt_batch_mesh"),
                                        (_lib.BatchGrad, "gpurt_batch_grad"),
                                        (_lib.Instance, "gpurt_instance"),
                                        (_lib.Object, "gpurt_object"),
                                        (_lib.RenderParams, "gpurt_render_params")])
def test_struct_layouts_match_header(cls, struct):
    fields = [f for f, _ in cls._fields_]
    size, offs = _c_layout(struct, fields)
    assert ctypes.sizeof(cls) == size
    assert [getattr(cls, f).offset for f in fields] == offs


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
    assert lib.gpurt_bvh_export(None, None, None) == _lib.GPURT_E_INVALID
    assert lib.gpurt_trace_closest_noalloc(None, None, 1, None, None, None, None, None) == _lib.GPURT_E_INVALID
    assert lib.gpurt_ctx_create(0, None) == _lib.GPURT_E_INVALID
    assert lib.gpurt_ctx_status(None, None, 0) == _lib.GPURT_E_INVALID
    assert lib.gpurt_ctx_set_vertex_sums(None, 1) == _lib.GPURT_E_INVALID
    assert lib.gpurt_tune_cache_load(None, None, None) == _lib.GPURT_E_INVALID
    assert lib.gpurt_tune_batch(None, None, 0, None, None, None) == _lib.GPURT_E_INVALID
    assert lib.gpurt_scene_create(None, None, 1, None, ctypes.byref(out)) == _lib.GPURT_E_INVALID
    assert lib.gpurt_render(None, None, None, None, None, None, None) == _lib.GPURT_E_INVALID
    assert lib.gpurt_scene_light_pdf(None, None, None, 1, None, None) == _lib.GPURT_E_INVALID
    lib.gpurt_scene_destroy(None)
    lib.gpurt_bvh_destroy(None)
    lib.gpurt_ctx_destroy(None)


def test_check_raises_with_message():
    lib = _lib.load()
    lib.gpurt_bvh_get_info(None, None)
    with pytest.raises(_lib.GpurtError) as e:
        _lib.check(_lib.GPURT_E_INVALID)
    assert "GPURT_E_INVALID" in str(e.value)
