"""Two-level instancing (gpurt_tlas_*) against the oracle's instanced brute force
(orc_inst_ray_brute_f32): the same FP32 object-space rays and the same triangle test, so
object ids, primitive ids, t, (u,v) and occlusion are compared bit for bit."""
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


def _check(ctx, dev, sc, rays, rays_any, custom=None, mask=None, qkw=None, **bkw):
    blas = [gpurt.Bvh(ctx, _t(V, dev), _t(F, dev), **bkw) for V, F in sc["meshes"]]
    n = sc["xf"].shape[0]
    custom = list(range(100, 100 + n)) if custom is None else custom
    mask = [0xFF] * n if mask is None else mask
    inst = [(blas[m], sc["xf"][i].reshape(3, 4), custom[i], mask[i]) for i, m in enumerate(sc["mesh_of"])]
    tlas = gpurt.Tlas(ctx, inst)
    qkw = qkw or {}
    obj, prim, t, uv = _np(*tlas.trace_closest(_t(rays, dev), **qkw))
    eo, ep, et, euv = O.inst_ray_brute_f32(sc["xf"], sc["mesh_of"], sc["meshes"], rays, custom=custom, mask=mask)
    assert np.array_equal(prim, ep), f"{(prim != ep).sum()} prim ids differ"
    assert np.array_equal(obj, eo), f"{(obj != eo).sum()} object ids differ"
    assert np.array_equal(t, et)
    assert np.array_equal(uv, euv)
    occ, = _np(tlas.trace_any(_t(rays_any, dev), **qkw))
    eocc = O.inst_ray_brute_f32(sc["xf"], sc["mesh_of"], sc["meshes"], rays_any, custom=custom, mask=mask,
                                any_hit=True)
    assert np.array_equal(occ, eocc)
    return obj, tlas


@pytest.mark.parametrize("fmt", [1, 2, 3])
@pytest.mark.parametrize("width", [2, 4, 8])
def test_tlas_matches_instanced_brute_force(ctx, dev, width, fmt):
    sc = W.instanced_scene(n_rays=6000, grid=41, n_spheres=12, sphere_level=2, seed=5)
    print("tlas width", width, "format", fmt, sc["meta"], flush=True)
    obj, _ = _check(ctx, dev, sc, sc["rays"], sc["rays_any"], node_width=width, node_format=fmt)
    assert len(set(obj.tolist())) > 5  # rays hit many different instances


@pytest.mark.parametrize("refill", [1, 13, 64])
@pytest.mark.parametrize("leaf", [1, 64])
def test_tlas_wave_scheduling_options_keep_answers(ctx, dev, leaf, refill):
    sc = W.instanced_scene(n_rays=6000, grid=41, n_spheres=12, sphere_level=2, seed=5)
    print("tlas leaf_trigger", leaf, "refill_trigger", refill, flush=True)
    _check(ctx, dev, sc, sc["rays"], sc["rays_any"],
           qkw=dict(leaf_trigger=leaf, refill_trigger=refill, waves_per_cu=4))


@pytest.mark.parametrize("n", [1, 3, 70])
def test_tlas_few_rays_shared_by_the_whole_wave(ctx, dev, n):
    """A handful of rays leave most lanes idle from the start, so subtrees of both levels
    are handed to helper lanes (work donation) and merged by (t, instance, id): still the
    instanced brute force's answers."""
    sc = W.instanced_scene(n_rays=n, grid=41, n_spheres=12, sphere_level=2, seed=9)
    print("tlas few rays", n, flush=True)
    _check(ctx, dev, sc, sc["rays"], sc["rays_any"])
    # coincident instances: equal t in several instances, the lowest instance must win
    V, F = sc["meshes"][1]
    blas = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev))
    xf = np.eye(4, dtype=np.float32)[:3]
    tlas = gpurt.Tlas(ctx, [(blas, xf, 10 + i, 0xFF) for i in range(6)])
    R = W.pack_rays(np.tile([[0.05, 0.1, 5.0]], (n, 1)), np.tile([[0.0, 0.0, -1.0]], (n, 1)), 0.0, 100.0)
    obj, prim, t, _ = _np(*tlas.trace_closest(_t(R, dev)))
    eo, ep, et, _ = O.inst_ray_brute_f32(np.tile(xf.reshape(1, 12), (6, 1)), np.zeros(6, np.int32), [(V, F)], R,
                                          custom=list(range(10, 16)))
    assert np.array_equal(obj, eo) and np.array_equal(prim, ep) and np.array_equal(t, et)
    assert (obj == 10).all()


def test_tlas_masks_and_custom_indices(ctx, dev):
    sc = W.instanced_scene(n_rays=4000, grid=33, n_spheres=8, sphere_level=2, seed=6)
    n = sc["xf"].shape[0]
    mask = [0xFF] * n
    mask[0] = 0     # the terrain is invisible
    mask[3] = 0x100  # no bit of the 0xFF ray mask: invisible
    custom = [7 * i + 3 for i in range(n)]
    obj, _ = _check(ctx, dev, sc, sc["rays"], sc["rays_any"], custom=custom, mask=mask)
    assert 3 not in obj.tolist() and custom[3] not in obj.tolist()


def test_tlas_ties_between_coincident_instances(ctx, dev):
    """Two instances with the same transform of the same mesh: every hit is an exact tie,
    resolved to the lower instance."""
    sc = W.instanced_scene(n_rays=3000, grid=25, n_spheres=3, sphere_level=2, seed=7)
    sc["xf"] = np.concatenate([sc["xf"], sc["xf"][2:3]])
    sc["mesh_of"] = np.concatenate([sc["mesh_of"], sc["mesh_of"][2:3]])
    _check(ctx, dev, sc, sc["rays"], sc["rays_any"])


def test_identity_instance_equals_plain_bvh(ctx, dev):
    w = W.rays_scene(n_rays=4096, grid=61, n_spheres=3)
    V, F = _t(w["verts"], dev), _t(w["tris"], dev)
    bvh = gpurt.Bvh(ctx, V, F)
    tlas = gpurt.Tlas(ctx, [(bvh, np.eye(4), 42)])
    R = _t(w["rays"], dev)
    o, p, t, uv = tlas.trace_closest(R)
    pp, pt, puv = bvh.trace_closest(R)
    torch.cuda.synchronize()
    assert torch.equal(p, pp) and torch.equal(t, pt) and torch.equal(uv, puv)
    assert torch.equal(o, torch.where(pp >= 0, 42, -1).to(torch.int32))


def test_tlas_rejects_bad_instances(ctx, dev):
    sc = W.instanced_scene(n_rays=16, grid=9, n_spheres=2, sphere_level=1, seed=8)
    (V0, F0), (V1, F1) = sc["meshes"]
    b4 = gpurt.Bvh(ctx, _t(V0, dev), _t(F0, dev), node_width=4)
    b8 = gpurt.Bvh(ctx, _t(V1, dev), _t(F1, dev), node_width=8)
    with pytest.raises(gpurt.GpurtError):
        gpurt.Tlas(ctx, [(b4, np.eye(4)), (b8, np.eye(4))])  # mixed node widths
    singular = np.eye(4)
    singular[2, 2] = 0.0
    with pytest.raises(gpurt.GpurtError):
        gpurt.Tlas(ctx, [(b4, singular)])


def test_rays_outside_the_exact_domain_are_reported(ctx, dev):
    """gpurt_ctx_status reports a TLAS ray whose origin lies beyond 2^8 x the scene's
    largest coordinate (answered, but outside the domain where exactness is argued);
    rays inside it leave the status clear."""
    sc = W.instanced_scene(n_rays=512, grid=41, n_spheres=2)
    blas = [gpurt.Bvh(ctx, torch.from_numpy(V).to(dev), torch.from_numpy(F).to(dev)) for V, F in sc["meshes"]]
    tlas = gpurt.Tlas(ctx, [(blas[m], sc["xf"][i].reshape(3, 4)) for i, m in enumerate(sc["mesh_of"])])
    ctx.status(clear=True)
    tlas.trace_closest(torch.from_numpy(sc["rays"]).to(dev))
    assert ctx.status() == set()
    far = sc["rays"][:4].copy()
    far[:, 0] = 1e9  # far beyond 2^8 scene extents
    tlas.trace_any(torch.from_numpy(far).to(dev))
    assert ctx.status(clear=True) == {"tlas_outside_domain"}
    assert ctx.status() == set()
