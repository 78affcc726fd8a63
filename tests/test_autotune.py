"""Autotuner: cache keys and storage (CPU), and a tuned configuration that stays exact
(GPU)."""
import numpy as np
import pytest
import torch

from gpurt import autotune
from gpurt import workloads as W
from oracle import oracle as O


def test_signature_buckets_sizes():
    a = autotune.signature("dev", 1_000_000, 1 << 20, 0, "cpq", "b0")
    b = autotune.signature("dev", 1_040_000, (1 << 20) + 5, 0, "cpq", "b0")
    c = autotune.signature("dev", 4_000_000, 1 << 20, 0, "cpq", "b0")
    assert a == b and a != c


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    assert autotune.load_cache(path) == {}
    autotune.store_cache("k1", {"build": {"node_width": 4}}, path)
    autotune.store_cache("k2", {"build": {"node_width": 8}}, path)
    c = autotune.load_cache(path)
    assert c["k1"]["build"]["node_width"] == 4 and c["k2"]["build"]["node_width"] == 8


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
    assert not best["cached"] and len(best["table"]) == 16
    again = autotune.tune_queries(ctx, tv, tf, queries=tq, rays=tr, space=space, path=path)
    assert again["cached"] and again["build"] == best["build"]
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
    best = autotune.tune_batch(ctx, meshes, space=space, path=str(tmp_path / "t.json"))
    res = gpurt.Batch(ctx, meshes, **best["build"], **best["query"]).run()
    torch.cuda.synchronize()
    for m, r in zip(meshes, res):
        ep, _, ed, _ = O.cpq_brute_f32(m["verts"].cpu().numpy(), m["tris"].cpu().numpy(),
                                       m["queries"].cpu().numpy())
        assert np.array_equal(r["q_prim"].cpu().numpy(), ep)
