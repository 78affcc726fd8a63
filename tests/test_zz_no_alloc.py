"""gpurt_trace_closest_noalloc (traversal scratch in a fixed per-device array, no device
allocation) against gpurt_trace_closest on the same BVH and rays, through the C ABI.

The fixed array is shared by every call on a device and keeps the previous call's
contents, so two calls run back to back, the second with fewer rays; each must equal the
allocating entry point bit for bit (prim, t, uv).  The full 2.31 M-triangle scene has a
16-level tree whose worst-case stack spill does not fit the array at the default grid, so
the call also exercises the fitted (smaller) grid.
"""
import ctypes

import numpy as np
import pytest
import torch

import gpurt
from gpurt import _lib
from gpurt import workloads as W

pytestmark = pytest.mark.gpu


def _trace(fn, bvh, rays, opts=None):
    dev = bvh.device
    n = rays.shape[0]
    prim = torch.full((n,), -7, dtype=torch.int32, device=dev)
    t = torch.full((n,), -7.0, dtype=torch.float32, device=dev)
    uv = torch.full((n, 2), -7.0, dtype=torch.float32, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(fn(bvh.handle, ctypes.c_void_p(rays.data_ptr()), n, ctypes.c_void_p(prim.data_ptr()),
                  ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(uv.data_ptr()),
                  ctypes.byref(opts) if opts is not None else None, stream))
    torch.cuda.synchronize(dev)
    return prim.cpu().numpy(), t.cpu().numpy(), uv.cpu().numpy()


def _same(a, b, what):
    for x, y, name in zip(a, b, ("prim", "t", "uv")):
        assert np.array_equal(x, y), f"{what}: {name} differs in {(x != y).sum()} places"


@pytest.mark.parametrize("scene", ["full", "small"])
def test_noalloc_equals_allocating_entry_point_twice(ctx, dev, scene):
    lib = _lib.load()
    if scene == "full":
        w = W.rays_scene()  # 2,310,418 triangles, 1,048,576 rays
        n2 = 70_001
    else:
        w = W.rays_scene(n_rays=20_000, grid=101, n_spheres=6)
        n2 = 3_001
    print("noalloc", scene, w["meta"], flush=True)
    V = torch.from_numpy(w["verts"]).to(dev)
    F = torch.from_numpy(w["tris"]).to(dev)
    bvh = gpurt.Bvh(ctx, V, F)
    rays1 = torch.from_numpy(w["rays"]).to(dev)
    # second call: fewer rays, from another part of the array, reversed
    rays2 = torch.flip(rays1[-n2:], [0]).contiguous()
    for k, rays in enumerate((rays1, rays2)):
        ref = _trace(lib.gpurt_trace_closest, bvh, rays)
        got = _trace(lib.gpurt_trace_closest_noalloc, bvh, rays)
        print("noalloc", scene, "call", k, "n", rays.shape[0], "hits", int((ref[0] >= 0).sum()), flush=True)
        _same(got, ref, f"{scene} call {k}")
        assert (ref[0] >= 0).any() and (ref[0] < 0).any()


def test_noalloc_options_and_errors(ctx, dev):
    lib = _lib.load()
    w = W.rays_scene(n_rays=8192, grid=101, n_spheres=6)
    bvh = gpurt.Bvh(ctx, torch.from_numpy(w["verts"]).to(dev), torch.from_numpy(w["tris"]).to(dev),
                    node_width=8)
    rays = torch.from_numpy(w["rays"]).to(dev)
    for opts in (_lib.QueryOpts(16, 4, None, 1, 64), _lib.QueryOpts(64, 0, None, 0, 0)):
        _same(_trace(lib.gpurt_trace_closest_noalloc, bvh, rays, opts),
              _trace(lib.gpurt_trace_closest, bvh, rays, opts), "options")
    stats = torch.zeros(4, dtype=torch.int64, device=dev)
    opts = _lib.QueryOpts(0, 0, ctypes.c_void_p(stats.data_ptr()), 0, 0)
    with pytest.raises(_lib.GpurtError) as e:
        _trace(lib.gpurt_trace_closest_noalloc, bvh, rays, opts)
    assert e.value.code == _lib.GPURT_E_UNSUPPORTED
