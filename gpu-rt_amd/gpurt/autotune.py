"""Autotuner for the traversal and build parameters (Python face of gpurt_tune_queries /
gpurt_tune_batch, include/gpurt.h).

The library sweeps node width, node encoding (FP32 or 8-bit child boxes), LDS stack
depth, resident waves per CU and treelet size over the caller's own workload, times every
candidate with HIP events on the current stream and keeps the fastest; the library
defaults are timed beside it.  Answers are independent of every parameter (the traversal
is exact for any tree), so only speed is compared.  Winners are kept in a text cache
keyed by the device and log2 buckets of the workload's sizes, the same file C++ callers
read with gpurt::load_tuned, so a second call with a similar workload costs nothing.

    best = autotune.tune_queries(ctx, verts, tris, queries=q, rays=r)
    bvh = gpurt.Bvh(ctx, verts, tris, **best["build"])
    bvh.closest_point(q, **best["query"])

    best = autotune.tune_batch(ctx, meshes)
    batch = gpurt.Batch(ctx, meshes, **best["build"], **best["query"])
"""
import ctypes
import os

import torch

from . import _lib

DEFAULT_SPACE = {
    "node_width": (2, 4, 8),
    "stack_depth": (8, 16, 32, 64),
    "waves_per_cu": (0, 16, 24),
    "treelet_size": (0, 6),
    "node_format": (1, 2),
}


def cache_path():
    return os.environ.get("GPURT_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "gpurt", "autotune.txt"))


def _space(space, path, reps, builds_per_run=0.0):
    sp = dict(DEFAULT_SPACE, **(space or {}))
    keep = []
    s = _lib.TuneSpace()
    for k in DEFAULT_SPACE:
        arr = (ctypes.c_int32 * len(sp[k]))(*sp[k])
        keep.append(arr)
        setattr(s, k, ctypes.cast(arr, ctypes.c_void_p))
        setattr(s, "n_" + k, len(sp[k]))
    s.reps = int(reps)
    s.builds_per_run = float(builds_per_run)
    if path:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        s.cache_path = os.path.abspath(path).encode()
    return s, keep


def _result(r):
    return {"build": {"node_width": r.build.node_width, "treelet_size": r.build.treelet_size,
                      "node_format": r.build.node_format},
            "query": {"stack_depth": r.query.stack_depth, "waves_per_cu": r.query.waves_per_cu},
            "objective_ms": r.best_ms, "default_ms": r.default_ms, "candidates": r.candidates,
            "key": r.key.decode(), "cached": bool(r.from_cache)}


def load(key, path=None):
    """The cached result for `key`, or None (gpurt_tune_cache_load)."""
    r = _lib.TuneResult()
    rc = _lib.load().gpurt_tune_cache_load((path or cache_path()).encode(), key.encode(), ctypes.byref(r))
    return _result(r) if rc == _lib.GPURT_OK else None


def store(result, path=None):
    """Append a result (a dict as returned by the tuners) to the cache."""
    r = _lib.TuneResult()
    b, q = result["build"], result["query"]
    r.build.node_width, r.build.treelet_size, r.build.node_format = (b["node_width"], b["treelet_size"],
                                                                      b["node_format"])
    r.query.stack_depth, r.query.waves_per_cu = q["stack_depth"], q["waves_per_cu"]
    r.best_ms, r.default_ms = result["objective_ms"], result.get("default_ms", 0.0)
    r.key = result["key"].encode()
    _lib.check(_lib.load().gpurt_tune_cache_store((path or cache_path()).encode(), ctypes.byref(r)))


def _sample(t, n):
    if t is None or t.shape[0] <= n:
        return t
    idx = torch.linspace(0, t.shape[0] - 1, n, device=t.device).long()
    return t[idx].contiguous()


def tune_queries(ctx, verts, tris, queries=None, rays=None, space=None, sample=262144,
                 builds_per_call=0.0, use_cache=True, path=None, reps=3):
    """Pick build and query parameters for one mesh (gpurt_tune_queries).

    The objective is query time on `sample` of the queries and rays, plus the build time
    times `builds_per_call` (0 = the BVH is built once and queried many times; 1 = a
    rebuild before every query batch)."""
    q = _sample(queries, sample)
    r = _sample(rays, sample)
    sp, keep = _space(space, (path or cache_path()) if use_cache else None, reps, builds_per_call)
    out = _lib.TuneResult()
    stream = ctypes.c_void_p(torch.cuda.current_stream(ctx.device).cuda_stream)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    _lib.check(_lib.load().gpurt_tune_queries(
        ctx.handle, ptr(verts), verts.shape[1] * 4, verts.shape[0], ptr(tris), tris.shape[0],
        ptr(q), 0 if q is None else q.shape[0], ptr(r), 0 if r is None else r.shape[0],
        ctypes.byref(sp), stream, ctypes.byref(out)))
    del keep
    return _result(out)


def tune_batch(ctx, meshes, space=None, use_cache=True, path=None, reps=3):
    """Pick parameters for a batch (gpurt_tune_batch): every candidate is a full replay
    (all builds and all queries) of these meshes, which is what the batch runs.  `meshes`
    are the dicts gpurt.Batch takes; their outputs are written by the trial replays."""
    import gpurt
    probe = gpurt.Batch(ctx, meshes)  # validates the meshes and lays out the C array
    sp, keep = _space(space, (path or cache_path()) if use_cache else None, reps)
    out = _lib.TuneResult()
    stream = ctypes.c_void_p(torch.cuda.current_stream(ctx.device).cuda_stream)
    _lib.check(_lib.load().gpurt_tune_batch(ctx.handle, probe._arr, len(meshes), ctypes.byref(sp), stream,
                                            ctypes.byref(out)))
    torch.cuda.synchronize(ctx.device)
    probe.close()
    del keep
    return _result(out)
