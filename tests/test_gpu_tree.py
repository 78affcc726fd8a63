"""The GPU builder's tree, read back with gpurt_bvh_export and checked against the
invariants that define a valid BVH over the mesh (the reference's tree is opaque, so the
build is pinned by these and by the query parity tests):

  * the nodes form a tree: every node reachable from the root exactly once, stored level
    by level, as many levels as the handle reports;
  * leaves are disjoint ranges of packed triangles that cover all of them, each of at
    most leaf_size triangles; the packed order is a permutation and every packed triangle
    holds its mesh triangle's vertices (bit for bit) and id;
  * every child box contains its subtree: a child node's boxes, or a leaf's vertices
    (FP32 comparisons); each child's count equals the triangles under it;
  * every 8-bit box, decoded with the traversal's own arithmetic, contains its FP32 box;
  * without treelets the packed order is the oracle's Morton order (30-bit codes of the
    triangle box centres over the centroid bounds, stable sort), and with treelets it is
    still a permutation while the SAH cost the handle reports matches the exported boxes.
"""
import numpy as np
import pytest
import torch

import gpurt
from gpurt import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EMPTY = 0xFFFFFFFF
QBYTES = {2: 48, 4: 64, 8: 96}


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _export(bvh):
    ex = bvh.export()
    torch.cuda.synchronize()
    return {k: (v.cpu().numpy() if v is not None else None) for k, v in ex.items()}


def _split(nodes, W):
    f = nodes.view(np.float32)
    u = nodes.view(np.uint32)
    lo = np.stack([f[:, 0:W], f[:, W:2 * W], f[:, 2 * W:3 * W]], -1)   # (n, W, 3)
    hi = np.stack([f[:, 3 * W:4 * W], f[:, 4 * W:5 * W], f[:, 5 * W:6 * W]], -1)
    return lo, hi, u[:, 6 * W:7 * W], u[:, 7 * W:8 * W]


def _check_tree(bvh, V, F, leaf_size):
    info = bvh.info()
    W, nn, nt = info["node_width"], info["n_nodes"], info["n_tris"]
    ex = _export(bvh)
    lo, hi, child, count = _split(ex["nodes"], W)
    # --- a tree, level by level
    seen = np.zeros(nn, np.int32)
    seen[0] = 1
    level = [0]
    depth = 0
    leaves = []
    under = {}
    while level:
        depth += 1
        nxt = []
        for i in level:
            for j in range(W):
                c = int(child[i, j])
                if c == EMPTY:
                    continue
                if c & 0x80000000:
                    first, cnt = c & 0x0FFFFFFF, ((c >> 28) & 7) + 1
                    assert cnt <= leaf_size
                    leaves.append((first, cnt, i, j))
                else:
                    assert 0 < c < nn and seen[c] == 0, (i, j, c)
                    seen[c] = 1
                    assert c > max(level), "nodes are stored level by level"
                    nxt.append(c)
        level = nxt
    assert seen.all(), f"{(seen == 0).sum()} nodes unreachable"
    assert depth == info["depth"]
    # --- leaves cover the packed triangles exactly once
    cover = np.zeros(nt, np.int32)
    for first, cnt, _, _ in leaves:
        cover[first:first + cnt] += 1
    assert (cover == 1).all(), f"{(cover == 0).sum()} uncovered, {(cover > 1).sum()} covered twice"
    order = ex["leaf_order"]
    assert np.array_equal(np.sort(order), np.arange(nt))
    tri = ex["tris"].reshape(nt, 3, 4)
    assert np.array_equal(tri[:, 0, 3].view(np.uint32), order.astype(np.uint32))
    assert np.array_equal(tri[:, :, :3], V[F[order]])
    # --- boxes contain their subtrees, counts are right
    def subtree_count(i, j):
        c = int(child[i, j])
        if c & 0x80000000:
            return ((c >> 28) & 7) + 1
        return int(sum(subtree_count(c, k) for k in range(W) if child[c, k] != EMPTY))
    for i in range(nn):
        for j in range(W):
            c = int(child[i, j])
            if c == EMPTY:
                continue
            if c & 0x80000000:
                first, cnt = c & 0x0FFFFFFF, ((c >> 28) & 7) + 1
                p = tri[first:first + cnt, :, :3].reshape(-1, 3)
                assert (p >= lo[i, j]).all() and (p <= hi[i, j]).all(), (i, j)
                assert count[i, j] == cnt
            else:
                ne = child[c] != EMPTY
                assert (lo[c][ne] >= lo[i, j]).all() and (hi[c][ne] <= hi[i, j]).all(), (i, j, c)
    # counts of internal children (checked on the top levels; the leaves above cover the rest)
    for j in range(W):
        if child[0, j] != EMPTY:
            assert count[0, j] == subtree_count(0, j)
    assert count[0][child[0] != EMPTY].sum() == nt
    return info, ex, lo, hi, child


def _check_q8(ex, W, lo, hi, child):
    q = ex["qnodes"]
    assert q.shape[1] == QBYTES[W]
    origin = q[:, 0:12].copy().view(np.float32)                      # (n, 3)
    if W <= 4:  # float scales, then the child words
        scale = q[:, 12:24].copy().view(np.float32)                   # (n, 3)
        c0 = 24
    else:       # biased exponent bytes
        exps = q[:, 12:15].astype(np.int32)                           # (n, 3)
        scale = np.ldexp(np.float32(1), exps - 127).astype(np.float32)
        c0 = 16
    # every scale is a power of two
    m, _ = np.frexp(scale.astype(np.float64))
    assert (m == 0.5).all()
    qchild = q[:, c0:c0 + 4 * W].copy().view(np.uint32)
    assert np.array_equal(qchild, child)
    planes = q[:, c0 + 4 * W:c0 + 10 * W].reshape(-1, 6, W).astype(np.float32)
    # empty slots hold the inverted box lo = 255, hi = 0
    em = child == EMPTY
    assert (planes[:, 0:3].transpose(0, 2, 1)[em] == 255).all()
    assert (planes[:, 3:6].transpose(0, 2, 1)[em] == 0).all()
    # fmaf(q, scale, origin): q * scale is exact in FP32, so this is one rounding as fmaf
    dec_lo = (planes[:, 0:3].transpose(0, 2, 1) * scale[:, None, :]).astype(np.float32) + origin[:, None, :]
    dec_hi = (planes[:, 3:6].transpose(0, 2, 1) * scale[:, None, :]).astype(np.float32) + origin[:, None, :]
    ne = child != EMPTY
    assert (dec_lo[ne] <= lo[ne]).all() and (dec_hi[ne] >= hi[ne]).all()


@pytest.mark.parametrize("fmt", [1, 2, 3])
@pytest.mark.parametrize("width", [2, 4, 8])
def test_tree_invariants(ctx, dev, width, fmt):
    w = W.cpq_small()
    V, F = w["verts"], w["tris"]
    bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev), node_width=width, node_format=fmt)
    info, ex, lo, hi, child = _check_tree(bvh, V, F, 4)
    if fmt in (2, 3):
        _check_q8(ex, width, lo, hi, child)
    else:
        assert ex["qnodes"] is None
    # no treelets: the packed order is the oracle's Morton order
    assert np.array_equal(ex["leaf_order"], O.CpuBvh(V, F, 4).order())


@pytest.mark.parametrize("tsize", [5, 7])
@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_tree_invariants_with_treelets_and_leaf_sizes(ctx, dev, tsize, leaf):
    w = W.rays_scene(n_rays=16, grid=81, n_spheres=4)
    V, F = w["verts"], w["tris"]
    bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev), treelet_size=tsize, leaf_size=leaf)
    info, ex, lo, hi, child = _check_tree(bvh, V, F, leaf)
    _check_q8(ex, info["node_width"], lo, hi, child)
    # the reported SAH cost is the one of the exported boxes: 1.2 x area per internal
    # child plus area x triangles per leaf, over the root area (here the union of the
    # root's inflated child boxes, a hair larger than the binary root box it is
    # reported against)
    e = (hi - lo).astype(np.float64)
    area = 2 * (e[..., 0] * e[..., 1] + e[..., 1] * e[..., 2] + e[..., 0] * e[..., 2])
    leafc = (child & 0x80000000) != 0
    ne = child != EMPTY
    cnt = ((child >> 28) & 7) + 1
    cost = (1.2 * area[ne & ~leafc]).sum() + (area * cnt)[ne & leafc].sum()
    rl, rh = lo[0][ne[0]].min(0).astype(np.float64), hi[0][ne[0]].max(0).astype(np.float64)
    ext = rh - rl
    root = 2 * (ext[0] * ext[1] + ext[1] * ext[2] + ext[0] * ext[2])
    assert abs(cost / root - info["sah_cost"]) <= 0.01 * info["sah_cost"], (cost / root, info["sah_cost"])


def test_tree_after_refit(ctx, dev):
    w = W.cpq_small()
    V, F = w["verts"], w["tris"]
    tv = _t(V, dev)
    bvh = gpurt.Bvh(ctx, tv, _t(F, dev))
    V2 = (V * np.float32(1.3) + np.random.default_rng(1).normal(0, 0.05, V.shape)).astype(np.float32)
    tv.copy_(_t(V2, dev))
    bvh.refit(tv)
    info, ex, lo, hi, child = _check_tree(bvh, V2, F, 4)
    _check_q8(ex, 4, lo, hi, child)
