"""Bad input on the GPU path: non-finite vertex positions are rejected by build, refit and
batch creation (GPURT_E_INVALID; a rejected refit leaves the tree as it was), and rays or
query points that cannot be answered get the defined answer -- a miss, or no triangle --
equal to the oracle's bit for bit.  Also the shape checks of the Python layer that keep
the backward kernels' scatter inside the caller's buffers."""
import numpy as np
import pytest
import torch

import gpurt
from gpurt import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _np(*ts):
    torch.cuda.synchronize()
    return [t.cpu().numpy() if t is not None else None for t in ts]


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_non_finite_vertex_is_rejected_by_build(ctx, dev, bad):
    w = W.cpq_small(n_queries=16)
    V = w["verts"].copy()
    V[17, 1] = bad
    with pytest.raises(gpurt.GpurtError) as e:
        gpurt.Bvh(ctx, _t(V, dev), _t(w["tris"], dev))
    assert e.value.code == -1 and "non-finite" in str(e.value)
    # a vertex no triangle uses is never read
    V2 = np.concatenate([w["verts"], np.full((1, 3), bad, np.float32)])
    bvh = gpurt.Bvh(ctx, _t(V2, dev), _t(w["tris"], dev))
    prim, = _np(bvh.closest_point(_t(w["queries"], dev))[0])
    assert np.array_equal(prim, O.cpq_brute_f32(w["verts"], w["tris"], w["queries"])[0])


def test_non_finite_refit_is_rejected_and_keeps_the_tree(ctx, dev):
    w = W.cpq_small(n_queries=2000)
    V, F, Q = w["verts"], w["tris"], w["queries"]
    tv = _t(V, dev)
    bvh = gpurt.Bvh(ctx, tv, _t(F, dev))
    V2 = V.copy()
    V2[5, 0] = np.nan
    with pytest.raises(gpurt.GpurtError) as e:
        bvh.refit(_t(V2, dev))
    assert e.value.code == -1
    prim, pt, dist, bary = _np(*bvh.closest_point(_t(Q, dev)))
    ep, ept, ed, eb = O.cpq_brute_f32(V, F, Q)
    assert np.array_equal(prim, ep) and np.array_equal(dist, ed) and np.array_equal(pt, ept)


def test_non_finite_vertex_is_rejected_by_batch(ctx, dev):
    meshes = []
    for k in range(3):
        w = W.cpq_small(n_queries=100)
        V = w["verts"].copy()
        if k == 1:
            V[3, 2] = np.inf
        meshes.append({"verts": _t(V, dev), "tris": _t(w["tris"], dev), "queries": _t(w["queries"], dev)})
    with pytest.raises(gpurt.GpurtError) as e:
        gpurt.Batch(ctx, meshes)
    assert e.value.code == -1 and "non-finite" in str(e.value)


def _bad_rays(R):
    """Each ray of R followed by its variants that cannot be traced, and a variant with
    tmax = +inf that can."""
    out = []
    for r in R:
        out.append(r)
        for k, v in ((0, np.nan), (1, np.inf), (2, -np.inf), (4, np.nan), (5, np.inf), (3, np.nan), (7, np.nan)):
            x = r.copy()
            x[k] = v
            out.append(x)
        z = r.copy()
        z[4:7] = 0.0
        out.append(z)
        f = r.copy()
        f[7] = np.inf
        out.append(f)
    return np.array(out, np.float32)


@pytest.mark.parametrize("fmt", [1, 2])
def test_rays_that_cannot_be_traced_are_misses(ctx, dev, fmt):
    w = W.rays_scene(n_rays=512, grid=61, n_spheres=3)
    V, F = w["verts"], w["tris"]
    R = _bad_rays(w["rays"][:256])
    RA = _bad_rays(w["rays_any"][:256])
    bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev), node_format=fmt)
    prim, t, uv = _np(*bvh.trace_closest(_t(R, dev)))
    ep, et, euv = O.ray_closest_brute_f32(V, F, R)
    assert np.array_equal(prim, ep) and np.array_equal(t, et) and np.array_equal(uv, euv)
    occ, = _np(bvh.trace_any(_t(RA, dev)))
    assert np.array_equal(occ, O.ray_any_brute_f32(V, F, RA))
    # every variant but the last (tmax = +inf) is a miss; the originals do hit
    per = 10
    assert (prim.reshape(-1, per)[:, 1:per - 1] == -1).all()
    assert (t.reshape(-1, per)[:, 1:per - 1] == np.inf).all()
    assert (prim.reshape(-1, per)[:, 0] >= 0).any()
    assert (occ.reshape(-1, per)[:, 1:per - 1] == 0).all()


def test_rays_that_cannot_be_traced_are_misses_through_a_tlas(ctx, dev):
    sc = W.instanced_scene(n_rays=256, grid=61, n_spheres=3)
    blas = [gpurt.Bvh(ctx, _t(V, dev), _t(F, dev)) for V, F in sc["meshes"]]
    tlas = gpurt.Tlas(ctx, [(blas[m], sc["xf"][i].reshape(3, 4)) for i, m in enumerate(sc["mesh_of"])])
    R = _bad_rays(sc["rays"][:128])
    obj, prim, t, uv = _np(*tlas.trace_closest(_t(R, dev)))
    eo, ep, et, euv = O.inst_ray_brute_f32(sc["xf"], sc["mesh_of"], sc["meshes"], R)
    assert np.array_equal(obj, eo) and np.array_equal(prim, ep) and np.array_equal(t, et)
    occ, = _np(tlas.trace_any(_t(R, dev)))
    assert np.array_equal(occ, O.inst_ray_brute_f32(sc["xf"], sc["mesh_of"], sc["meshes"], R, any_hit=True))
    assert (prim.reshape(-1, 10)[:, 1:9] == -1).all()


def test_query_points_that_cannot_be_answered(ctx, dev):
    w = W.cpq_small(n_queries=1000)
    V, F = w["verts"], w["tris"]
    Q = w["queries"].copy()
    Q[::7, 0] = np.nan
    Q[1::7, 2] = np.inf
    Q[2::7, 1] = -np.inf
    bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev))
    prim, pt, dist, bary = _np(*bvh.closest_point(_t(Q, dev)))
    ep, ept, ed, eb = O.cpq_brute_f32(V, F, Q)
    assert np.array_equal(prim, ep) and np.array_equal(dist, ed)
    assert np.array_equal(pt, ept) and np.array_equal(bary, eb)
    bad = ~np.isfinite(Q).all(1)
    assert (prim[bad] == -1).all() and (dist[bad] == np.inf).all() and (pt[bad] == 0).all()
    assert (prim[~bad] >= 0).all()
    # a launch of nothing but such points ends at once (each visits only the root)
    st = torch.zeros(4, dtype=torch.int64, device=dev)
    bvh.closest_point(_t(np.full((4096, 3), np.nan, np.float32), dev), stats=st)
    nodes, tris = st[:2].tolist()
    assert nodes <= 4096 and tris == 0
    # their backward contributes nothing
    tq = _t(Q, dev).requires_grad_(True)
    tv = _t(V, dev).requires_grad_(True)
    d, _ = gpurt.closest_point_distance(bvh, tq, tv)
    d[torch.isfinite(d)].sum().backward()
    g = tq.grad.cpu().numpy()
    assert (g[bad] == 0).all() and np.isfinite(tv.grad.cpu().numpy()).all()


@pytest.mark.parametrize("k", [4, 12])
def test_wide_vertex_records_get_a_gradient_of_their_shape(ctx, dev, k):
    w = W.cpq_small(n_queries=500)
    V = np.zeros((w["verts"].shape[0], k), np.float32)
    V[:, :3] = w["verts"]
    V[:, 3:] = 5.0
    tv = _t(V, dev).requires_grad_(True)
    bvh = gpurt.Bvh(ctx, tv.detach(), _t(w["tris"], dev))
    q = _t(w["queries"], dev)
    d, prim = gpurt.closest_point_distance(bvh, q, tv)
    d.sum().backward()
    g = tv.grad.cpu().numpy()
    assert g.shape == V.shape and (g[:, 3:] == 0).all() and np.abs(g[:, :3]).sum() > 0
    # the same positions as a (n, 3) array give the same position gradient
    t3 = _t(w["verts"], dev).requires_grad_(True)
    d3, _ = gpurt.closest_point_distance(bvh, q, t3)
    d3.sum().backward()
    np.testing.assert_allclose(g[:, :3], t3.grad.cpu().numpy(), rtol=0, atol=1e-5)


def test_mismatched_shapes_raise_before_any_kernel(ctx, dev):
    w = W.cpq_small(n_queries=64)
    V = _t(w["verts"], dev)
    bvh = gpurt.Bvh(ctx, V, _t(w["tris"], dev))
    q = _t(w["queries"], dev)
    with pytest.raises(ValueError):
        gpurt.closest_point_distance(bvh, q, V[:-1])
    o = torch.zeros((8, 3), device=dev)
    d = torch.ones((8, 3), device=dev)
    with pytest.raises(ValueError):
        gpurt.ray_hit_t(bvh, o, d, V[:10])
    prim, pt, dist, bary = bvh.closest_point(q)
    with pytest.raises(ValueError):
        bvh.closest_point_bwd(q, prim, bary, pt, dist, torch.ones(64, device=dev),
                              g_verts=torch.zeros((V.shape[0] - 1, 3), device=dev))
    with pytest.raises(ValueError):
        bvh.closest_point_bwd(q, prim[:10], bary, pt, dist, torch.ones(64, device=dev))
    with pytest.raises(ValueError):
        bvh.closest_point(q, out=(prim[:5], pt, dist, bary))
    rays = torch.zeros((8, 8), device=dev)
    with pytest.raises(ValueError):
        bvh.trace_any(rays, out=torch.empty(4, dtype=torch.uint8, device=dev))
