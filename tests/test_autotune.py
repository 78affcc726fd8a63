"""Autotuner: the tuned-result cache (CPU: the library's text cache, also read by C++
callers), and tuned configurations that stay exact (GPU)."""
import numpy as np
import pytest
import torch

from gpurt import autotune
from gpurt import workloads as W
from oracle import oracle as O


def _res(key, w, s, ms):
    return {"build": {"node_width": w, "treelet_size": 0, "node_format": 2},
            "query": {"stack_depth": s, "waves_per_cu": 16}, "objective_ms": ms, "default_ms": 2 * ms,
            "key": key}


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.txt")
    assert autotune.load("k1", path) is None
    autotune.store(_res("gfx950x256|cpq|t20|q20|r-1|b0", 4, 8, 1.5), path)
    autotune.store(_res("k2", 8, 16, 2.5), path)
    autotune.store(_res("gfx950x256|cpq|t20|q20|r-1|b0", 2, 32, 1.25), path)  # a later line wins
    a = autotune.load("gfx950x256|cpq|t20|q20|r-1|b0", path)
    assert a["build"]["node_width"] == 2 and a["query"]["stack_depth"] == 32 and a["cached"]
    assert a["objective_ms"] == 1.25 and a["default_ms"] == 2.5
    b = autotune.load("k2", path)
    assert b["build"] == {"node_width": 8, "treelet_size": 0, "node_format": 2}
    assert autotune.load("k", path) is None  # prefixes do not match
    with open(path) as f:
        assert len(f.read().splitlines()) == 3


def test_cache_rejects_bad_keys(tmp_path):
    from gpurt import _lib
    with pytest.raises(_lib.GpurtError):
        autotune.store(_res("a\tb", 4, 8, 1.0), str(tmp_path / "t.txt"))


@pytest.mark.gpu
def test_tuned_queries_stay_exact(ctx, dev, tmp_path):
    import gpurt
    w = W.cpq_small(n_queries=8192)
    V, F, Q = w["verts"], w["tris"], w["queries"]
    tv, tf, tq = (torch.from_numpy(x).to(dev) for x in (V, F, Q))
    rng = np.random.default_rng(3)
    R = W.pack_rays(rng.uniform(-1.5, 1.5, (8192, 3)), W.random_dirs(rng, 8192), 1e-5, 1e7)
    tr = torch.from_numpy(R).to(dev)
    path = str(tmp_path / "tune.json")
    space = {"node_width": (2, 4), "stack_depth": (8, 16), "waves_per_cu": (0,), "treelet_size": (0, 5),
             "node_format": (1, 2)}
    best = autotune.tune_queries(ctx, tv, tf, queries=tq, rays=tr, space=space, path=path)
    assert not best["cached"] and best["candidates"] == 16
    assert 0 < best["objective_ms"] and 0 < best["default_ms"]
    again = autotune.tune_queries(ctx, tv, tf, queries=tq, rays=tr, space=space, path=path)
    assert again["cached"] and again["build"] == best["build"] and again["query"] == best["query"]
    bvh = gpurt.Bvh(ctx, tv, tf, **best["build"])
    prim, _, dist, _ = bvh.closest_point(tq, **best["query"])
    torch.cuda.synchronize()
    ep, _, ed, _ = O.cpq_brute_f32(V, F, Q)
    assert np.array_equal(prim.cpu().numpy(), ep) and np.array_equal(dist.cpu().numpy(), ed)


@pytest.mark.gpu
def test_tuned_batch(ctx, dev, tmp_path):
    import gpurt
    meshes = []
    for k in range(5):
        v, f = W.displaced_icosphere(2 + k % 3, np.random.default_rng(k))
        q = np.random.default_rng(10 + k).uniform(-1.2, 1.2, (500, 3)).astype(np.float32)
        meshes.append({"verts": torch.from_numpy(v.astype(np.float32)).to(dev),
                       "tris": torch.from_numpy(f.astype(np.int32)).to(dev),
                       "queries": torch.from_numpy(q).to(dev)})
    space = {"node_width": (4, 8), "stack_depth": (8,), "waves_per_cu": (0,), "treelet_size": (0,),
             "node_format": (2,)}
    best = autotune.tune_batch(ctx, meshes, space=space, path=str(tmp_path / "t.txt"))
    assert best["candidates"] == 2 and best["build"]["node_width"] in (4, 8)
    res = gpurt.Batch(ctx, meshes, **best["build"], **best["query"]).run()
    torch.cuda.synchronize()
    for m, r in zip(meshes, res):
        ep, _, ed, _ = O.cpq_brute_f32(m["verts"].cpu().numpy(), m["tris"].cpu().numpy(),
                                       m["queries"].cpu().numpy())
        assert np.array_equal(r["q_prim"].cpu().numpy(), ep)
