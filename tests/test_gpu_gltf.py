"""The reference's own scenes (media/cube, mis_test, cbox; committed as fixtures in
tests/golden by make_gltf_golden.py) through the GPU path: the flattened scene through a
BVH and the per-object scene through a TLAS, against the oracle's FP32 answers stored
with the fixture, bit for bit."""
import os

import numpy as np
import pytest
import torch

import gpurt

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("name", ["cube", "mis_test", "cbox"])
def test_reference_scene_flattened(ctx, dev, name):
    fx = np.load(os.path.join(GOLDEN, f"gltf_{name}.npz"))
    bvh = gpurt.Bvh(ctx, _t(fx["world_verts"], dev), _t(fx["world_tris"].astype(np.int32), dev))
    p, t, uv = bvh.trace_closest(_t(fx["rays"], dev))
    occ = bvh.trace_any(_t(fx["rays_any"], dev))
    qp, qpt, qd, qb = bvh.closest_point(_t(fx["queries"], dev))
    torch.cuda.synchronize()
    assert np.array_equal(p.cpu().numpy(), fx["ray_prim"]) and np.array_equal(t.cpu().numpy(), fx["ray_t"])
    assert np.array_equal(uv.cpu().numpy(), fx["ray_uv"])
    assert np.array_equal(occ.cpu().numpy(), fx["occluded"])
    assert np.array_equal(qp.cpu().numpy(), fx["cpq_prim"]) and np.array_equal(qd.cpu().numpy(), fx["cpq_dist"])
    assert np.array_equal(qpt.cpu().numpy(), fx["cpq_pt"]) and np.array_equal(qb.cpu().numpy(), fx["cpq_bary"])


@pytest.mark.parametrize("name", ["cube", "mis_test", "cbox"])
def test_reference_scene_instanced(ctx, dev, name):
    """One BLAS per glTF object (48-byte Mesh::Vertex records, as the loader packs them)
    and a TLAS over the objects' world matrices: GPURT::build_accel's structure."""
    fx = np.load(os.path.join(GOLDEN, f"gltf_{name}.npz"))
    n_obj = fx["xf"].shape[0]
    blas = []
    for k in range(n_obj):
        pos = fx[f"obj{k}_pos"]
        v = np.zeros((pos.shape[0], 12), np.float32)
        v[:, :3] = pos
        v[:, 3:] = 9.0  # normals / tangents / uv in the other columns are ignored
        blas.append(gpurt.Bvh(ctx, _t(v, dev), _t(fx[f"obj{k}_tris"].astype(np.int32), dev)))
    tlas = gpurt.Tlas(ctx, [(b, fx["xf"][k].reshape(3, 4), k) for k, b in enumerate(blas)])
    o, p, t, uv = tlas.trace_closest(_t(fx["rays"], dev))
    occ = tlas.trace_any(_t(fx["rays_any"], dev))
    torch.cuda.synchronize()
    assert np.array_equal(o.cpu().numpy(), fx["inst_obj"]) and np.array_equal(p.cpu().numpy(), fx["inst_prim"])
    assert np.array_equal(t.cpu().numpy(), fx["inst_t"]) and np.array_equal(uv.cpu().numpy(), fx["inst_uv"])
    assert np.array_equal(occ.cpu().numpy(), fx["inst_occluded"])
