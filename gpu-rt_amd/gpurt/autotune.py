"""Autotuner for the traversal and build parameters.

Sweeps node width {2, 4, 8}, node encoding (FP32 or 8-bit child boxes), LDS stack depth
{8, 16, 32, 64}, resident waves per CU and treelet size over a sample of the caller's own queries, times every candidate with HIP
events on the current stream, and keeps the fastest.  Answers are independent of every
parameter (the traversal is exact for any tree), so only speed is compared.  Winners
are cached in a JSON file keyed by the device name and a coarse signature of the
workload (log2 buckets of triangle and query counts, query kinds), so a second call with
a similar workload costs nothing.

    best = autotune.tune_queries(ctx, verts, tris, queries=q, rays=r)
    bvh = gpurt.Bvh(ctx, verts, tris, **best["build"])
    bvh.closest_point(q, **best["query"])

    best = autotune.tune_batch(ctx, meshes)
    batch = gpurt.Batch(ctx, meshes, **best["build"], **best["query"])
"""
import itertools
import json
import math
import os

import torch

DEFAULT_SPACE = {
    "node_width": (2, 4, 8),
    "stack_depth": (8, 16, 32, 64),
    "waves_per_cu": (0, 16, 24),
    "treelet_size": (0, 6),
    "node_format": (1, 2),
}


def cache_path():
    return os.environ.get("GPURT_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "gpurt", "autotune.json"))


def _bucket(n):
    return int(math.floor(math.log2(max(1, int(n)))))


def signature(device_name, n_tris, n_queries, n_rays, kind, objective):
    """Cache key: device, log2 size buckets, query kinds and objective."""
    return "|".join([device_name, kind, f"t{_bucket(n_tris)}", f"q{_bucket(n_queries)}",
                     f"r{_bucket(n_rays)}", objective])


def load_cache(path=None):
    path = path or cache_path()
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def store_cache(key, value, path=None):
    path = path or cache_path()
    data = load_cache(path)
    data[key] = value
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _time_ms(fn, reps=5):
    fn()  # warm-up (first launch also allocates stream-ordered scratch)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        torch.cuda.synchronize()
        times.append(s.elapsed_time(e))
    times.sort()
    return times[len(times) // 2]


def _sample(t, n):
    if t is None or t.shape[0] <= n:
        return t
    idx = torch.linspace(0, t.shape[0] - 1, n, device=t.device).long()
    return t[idx].contiguous()


def tune_queries(ctx, verts, tris, queries=None, rays=None, space=None, sample=262144,
                 builds_per_call=0.0, use_cache=True, path=None):
    """Pick build and query parameters for one mesh.

    The objective is query time on `sample` of the queries and rays, plus the build time
    times `builds_per_call` (0 = the BVH is built once and queried many times; 1 = a
    rebuild before every query batch)."""
    import gpurt
    space = dict(DEFAULT_SPACE, **(space or {}))
    dev = ctx.device
    name = torch.cuda.get_device_name(dev)
    nq = 0 if queries is None else queries.shape[0]
    nr = 0 if rays is None else rays.shape[0]
    kind = ("cpq" if nq else "") + ("+" if nq and nr else "") + ("ray" if nr else "")
    key = signature(name, tris.shape[0], nq, nr, kind, f"b{builds_per_call:g}")
    if use_cache:
        hit = load_cache(path).get(key)
        if hit:
            return dict(hit, cached=True)
    q = _sample(queries, sample)
    r = _sample(rays, sample)
    table = []
    for W, T, NF in itertools.product(space["node_width"], space["treelet_size"], space["node_format"]):
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        bvh = gpurt.Bvh(ctx, verts, tris, node_width=W, treelet_size=T, node_format=NF)
        ev1.record()
        torch.cuda.synchronize()
        build_ms = ev0.elapsed_time(ev1)
        for S, WV in itertools.product(space["stack_depth"], space["waves_per_cu"]):
            def run():
                if q is not None:
                    bvh.closest_point(q, stack_depth=S, waves_per_cu=WV)
                if r is not None:
                    bvh.trace_closest(r, stack_depth=S, waves_per_cu=WV)
            ms = _time_ms(run)
            table.append({"node_width": W, "treelet_size": T, "node_format": NF, "stack_depth": S,
                          "waves_per_cu": WV, "query_ms": ms, "build_ms": build_ms,
                          "objective_ms": ms + builds_per_call * build_ms})
        bvh.close()
    best = min(table, key=lambda e: e["objective_ms"])
    result = {"build": {"node_width": best["node_width"], "treelet_size": best["treelet_size"],
                        "node_format": best["node_format"]},
              "query": {"stack_depth": best["stack_depth"], "waves_per_cu": best["waves_per_cu"]},
              "objective_ms": best["objective_ms"], "key": key, "table": table}
    if use_cache:
        store_cache(key, {k: v for k, v in result.items() if k != "table"}, path)
    return dict(result, cached=False)


def tune_batch(ctx, meshes, space=None, use_cache=True, path=None, reps=3):
    """Pick parameters for a batch (gpurt.Batch): every candidate is a full replay (all
    builds and all queries), which is what the batch runs."""
    import gpurt
    space = dict(DEFAULT_SPACE, **(space or {}))
    name = torch.cuda.get_device_name(ctx.device)
    n_tris = sum(m["tris"].shape[0] for m in meshes)
    nq = sum(0 if m.get("queries") is None else m["queries"].shape[0] for m in meshes)
    nr = sum(0 if m.get("rays") is None else m["rays"].shape[0] for m in meshes)
    key = signature(name, n_tris, nq, nr, f"batch{_bucket(len(meshes))}", "replay")
    if use_cache:
        hit = load_cache(path).get(key)
        if hit:
            return dict(hit, cached=True)
    table = []
    for W, T, NF, S, WV in itertools.product(space["node_width"], space["treelet_size"],
                                             space["node_format"], space["stack_depth"],
                                             space["waves_per_cu"]):
        b = gpurt.Batch(ctx, meshes, node_width=W, treelet_size=T, stack_depth=S, waves_per_cu=WV,
                        node_format=NF)
        ms = _time_ms(b.run, reps)
        b.close()
        table.append({"node_width": W, "treelet_size": T, "node_format": NF, "stack_depth": S,
                      "waves_per_cu": WV, "replay_ms": ms})
    best = min(table, key=lambda e: e["replay_ms"])
    result = {"build": {"node_width": best["node_width"], "treelet_size": best["treelet_size"],
                        "node_format": best["node_format"]},
              "query": {"stack_depth": best["stack_depth"], "waves_per_cu": best["waves_per_cu"]},
              "objective_ms": best["replay_ms"], "key": key, "table": table}
    if use_cache:
        store_cache(key, {k: v for k, v in result.items() if k != "table"}, path)
    return dict(result, cached=False)
