"""Batched path: many meshes built and queried in one recorded graph must give exactly
the plain per-mesh answers (and the oracle's)."""
import numpy as np
import pytest
import torch

import gpurt
from gpurt import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _small_meshes(seed=0):
    rng = np.random.default_rng(seed)
    out = []
    kinds = [("ico", 1), ("ico", 2), ("terrain", 9), ("ico", 3), ("tri", 1), ("tri", 2),
             ("terrain", 20), ("ico", 0), ("terrain", 3)]
    for k, (kind, p) in enumerate(kinds):
        if kind == "ico":
            v, f = W.displaced_icosphere(p, np.random.default_rng([seed, k]))
        elif kind == "terrain":
            v, f = W.terrain(p, np.random.default_rng([seed, k]), octaves=3)
            v = v / 50.0
        else:
            v = rng.normal(size=(3 * p, 3))
            f = np.arange(3 * p).reshape(-1, 3)
        V, F = v.astype(np.float32), f.astype(np.int32)
        nq = int(rng.integers(1, 700))
        nr = int(rng.integers(1, 700))
        q = rng.uniform(-1.5, 1.5, (nq, 3)).astype(np.float32)
        o = rng.uniform(-2, 2, (nr, 3))
        rays = W.pack_rays(o, W.random_dirs(rng, nr), 1e-5, 1e7)
        out.append({"verts": V, "tris": F, "queries": q, "rays": rays})
    return out


def _dev(meshes, dev):
    return [{k: torch.from_numpy(v).to(dev) for k, v in m.items()} for m in meshes]


@pytest.mark.parametrize("fmt,width,sched", [(1, 2, {}), (1, 4, {}), (1, 8, {}), (2, 2, {}), (2, 4, {}),
                                             (2, 8, {}), (2, 4, dict(leaf_trigger=1, refill_trigger=64)),
                                             (1, 4, dict(leaf_trigger=64, refill_trigger=1))])
def test_batch_matches_oracle_and_plain(ctx, dev, width, fmt, sched):
    host = _small_meshes()
    dm = _dev(host, dev)
    print("batch width", width, "format", fmt, "meshes", len(host), sched, flush=True)
    batch = gpurt.Batch(ctx, dm, node_width=width, node_format=fmt, **sched)
    res = batch.run()
    torch.cuda.synchronize()
    for i, (h, r) in enumerate(zip(host, res)):
        ep, ept, ed, eb = O.cpq_brute_f32(h["verts"], h["tris"], h["queries"])
        assert np.array_equal(r["q_prim"].cpu().numpy(), ep), i
        assert np.array_equal(r["q_dist"].cpu().numpy(), ed), i
        assert np.array_equal(r["q_pt"].cpu().numpy(), ept), i
        assert np.array_equal(r["q_bary"].cpu().numpy(), eb), i
        rp, rt, ruv = O.ray_closest_brute_f32(h["verts"], h["tris"], h["rays"])
        assert np.array_equal(r["r_prim"].cpu().numpy(), rp), i
        assert np.array_equal(r["r_t"].cpu().numpy(), rt), i
        assert np.array_equal(r["r_uv"].cpu().numpy(), ruv), i
        # the plain per-mesh path gives the same answers
        bvh = gpurt.Bvh(ctx, dm[i]["verts"], dm[i]["tris"], node_width=width, node_format=fmt)
        p, _, d, _ = bvh.closest_point(dm[i]["queries"])
        assert torch.equal(p, r["q_prim"]) and torch.equal(d, r["q_dist"])


def test_batch_replay_rebuilds_from_current_vertices(ctx, dev):
    host = _small_meshes(seed=3)
    dm = _dev(host, dev)
    batch = gpurt.Batch(ctx, dm)
    first = [{k: (v.clone() if v is not None else None) for k, v in r.items()} for r in batch.run()]
    again = batch.run()
    torch.cuda.synchronize()
    for a, b in zip(first, again):
        assert torch.equal(a["q_prim"], b["q_prim"]) and torch.equal(a["r_t"], b["r_t"])
    # move the vertices in place: the next replay must rebuild every BVH from them
    for m in dm:
        m["verts"].mul_(1.25).add_(0.1)
    res = batch.run()
    torch.cuda.synchronize()
    for m, r in zip(dm, res):
        V = m["verts"].cpu().numpy()
        F = m["tris"].cpu().numpy()
        ep, _, ed, _ = O.cpq_brute_f32(V, F, m["queries"].cpu().numpy())
        assert np.array_equal(r["q_prim"].cpu().numpy(), ep)
        assert np.array_equal(r["q_dist"].cpu().numpy(), ed)
        rp, rt, _ = O.ray_closest_brute_f32(V, F, m["rays"].cpu().numpy())
        assert np.array_equal(r["r_prim"].cpu().numpy(), rp)
        assert np.array_equal(r["r_t"].cpu().numpy(), rt)


def test_batch_meshes_without_queries_or_rays(ctx, dev):
    host = _small_meshes(seed=5)[:4]
    host[1] = {"verts": host[1]["verts"], "tris": host[1]["tris"]}
    host[2] = {"verts": host[2]["verts"], "tris": host[2]["tris"], "rays": host[2]["rays"]}
    dm = _dev(host, dev)
    res = gpurt.Batch(ctx, dm).run()
    torch.cuda.synchronize()
    assert res[1]["q_prim"].numel() == 0 and res[1]["r_prim"].numel() == 0
    rp, rt, _ = O.ray_closest_brute_f32(host[2]["verts"], host[2]["tris"], host[2]["rays"])
    assert np.array_equal(res[2]["r_prim"].cpu().numpy(), rp)


def test_batch_rejects_bad_index(ctx, dev):
    host = _small_meshes()[:3]
    host[2]["tris"] = host[2]["tris"].copy()
    host[2]["tris"][0, 0] = host[2]["verts"].shape[0] + 5
    with pytest.raises(gpurt.GpurtError) as e:
        gpurt.Batch(ctx, _dev(host, dev))
    assert e.value.code == -1


def test_batch_with_treelets_matches_oracle(ctx, dev):
    host = _small_meshes(seed=7)
    res = gpurt.Batch(ctx, _dev(host, dev), treelet_size=6).run()
    torch.cuda.synchronize()
    for h, r in zip(host, res):
        ep, _, ed, _ = O.cpq_brute_f32(h["verts"], h["tris"], h["queries"])
        assert np.array_equal(r["q_prim"].cpu().numpy(), ep)
        assert np.array_equal(r["q_dist"].cpu().numpy(), ed)
        rp, rt, _ = O.ray_closest_brute_f32(h["verts"], h["tris"], h["rays"])
        assert np.array_equal(r["r_prim"].cpu().numpy(), rp)
        assert np.array_equal(r["r_t"].cpu().numpy(), rt)


@pytest.mark.parametrize("fmt,det", [(1, False), (2, False), (2, True)])
def test_batch_backward_matches_plain_backward(dev, fmt, det):
    """The recorded batched backward gives the plain per-mesh backward's gradients:
    query/origin/direction gradients bit for bit (same per-query arithmetic), vertex
    gradients up to the order of the FP32 sums; a second replay clears and recomputes the
    vertex gradients, with the same bits under fixed-order sums."""
    ctx = gpurt.Context(0, vertex_sums="fixed" if det else "atomic")
    host = _small_meshes(seed=11)
    dm = _dev(host, dev)
    batch = gpurt.Batch(ctx, dm, node_format=fmt)
    res = batch.run()
    grads = batch.enable_backward()
    gen = torch.Generator(device=dev).manual_seed(3)
    for g in grads:
        g["g_dist"].uniform_(0.5, 2.0, generator=gen)
        g["g_t"].uniform_(0.5, 2.0, generator=gen)
    first = None
    for rep in range(2):
        batch.run()
        batch.backward()
        torch.cuda.synchronize()
        if det:
            now = [g["g_verts"].clone() for g in grads]
            if first is not None:
                assert all(torch.equal(a, b) for a, b in zip(first, now))
            first = now
        for i, (m, r, g) in enumerate(zip(dm, res, grads)):
            bvh = gpurt.Bvh(ctx, m["verts"], m["tris"], node_format=fmt)
            prim, pt, dist, bary = bvh.closest_point(m["queries"])
            rp, rt, ruv = bvh.trace_closest(m["rays"])
            assert torch.equal(prim, r["q_prim"]) and torch.equal(rp, r["r_prim"])
            gq = torch.empty_like(m["queries"])
            go = torch.empty((rp.shape[0], 3), device=dev)
            gd = torch.empty((rp.shape[0], 3), device=dev)
            gv = torch.zeros((m["verts"].shape[0], 3), device=dev)
            bvh.closest_point_bwd(m["queries"], prim, bary, pt, dist, g["g_dist"], gq, gv)
            bvh.trace_bwd(m["rays"], rp, rt, ruv, g["g_t"], go, gd, gv)
            torch.cuda.synchronize()
            assert torch.equal(gq, g["g_q"]), (rep, i)
            assert torch.equal(go, g["g_origin"]) and torch.equal(gd, g["g_dir"]), (rep, i)
            scale = float(gv.abs().max()) + 1e-6
            assert float((gv - g["g_verts"]).abs().max()) <= 1e-5 * scale, (rep, i)
            # FP64 closed form of the distance term at the batch's own argmin
            V = m["verts"].cpu().numpy().astype(np.float64)
            F = m["tris"].cpu().numpy()
            Q = m["queries"].cpu().numpy().astype(np.float64)
            eq, _ = O.cpq_grad_f64(V, F, Q, r["q_prim"].cpu().numpy(), r["q_pt"].cpu().numpy().astype(np.float64),
                                   r["q_bary"].cpu().numpy().astype(np.float64),
                                   g["g_dist"].cpu().numpy().astype(np.float64))
            ok = r["q_dist"].cpu().numpy() > 1e-3
            np.testing.assert_allclose(g["g_q"].cpu().numpy()[ok], eq[ok], atol=1e-3, rtol=1e-3)
            bvh.close()
    batch.close()
    ctx.close()


def test_batch_backward_saved_outputs(ctx, dev):
    """Distance gradients recompute the saved triangle's closest point, so they need no
    saved barycentrics; vertex gradients of hit t need the saved (u,v)."""
    host = _small_meshes(seed=2)[:3]
    with_b = gpurt.Batch(ctx, _dev(host, dev))
    no_b = gpurt.Batch(ctx, _dev(host, dev), with_bary=False)
    ga = with_b.enable_backward(hit_t=False)
    gb = no_b.enable_backward(hit_t=False)
    for b in (with_b, no_b):
        b.run()
        b.backward()
    torch.cuda.synchronize()
    for a, b in zip(ga, gb):
        assert torch.equal(a["g_q"], b["g_q"])
        assert torch.allclose(a["g_verts"], b["g_verts"], atol=1e-5)  # atomic order
    no_uv = gpurt.Batch(ctx, _dev(host, dev), with_uv=False)
    with pytest.raises(ValueError):
        no_uv.enable_backward()
    g = no_uv.enable_backward(vertex_grads=False)
    no_uv.run()
    no_uv.backward()
    torch.cuda.synchronize()
    assert g[0]["g_verts"] is None and g[0]["g_q"].shape[1] == 3
