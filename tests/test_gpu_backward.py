"""Backward passes (gpurt_closest_point_bwd / gpurt_trace_bwd and the autograd
Functions) against the oracle: FP64 closed forms evaluated at the GPU's argmin, FP64
closed forms at the FP64 argmin (brute force), and FP64 central differences.

Vertex gradients are FP32 sums of per-query terms, compared with tolerances (stated per
test) against the FP64 forms; they are added with FP32 atomics (the order, and so the
last bits, vary) or sorted by vertex and added in a fixed order, the same bits on every
run (the context's vertex_sums mode; "auto" picks by size)."""
import numpy as np
import pytest
import torch

import gpurt
from gpurt import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _sphere(level, seed):
    v, f = W.displaced_icosphere(level, np.random.default_rng(seed))
    return v.astype(np.float32), f.astype(np.int32)


def test_distance_gradient_matches_closed_form(ctx, dev):
    V, F = _sphere(5, 8)
    Q = W.surface_samples(V.astype(np.float64), F, np.random.default_rng(9), 20000, 0.01).astype(np.float32)
    tv, tq = _t(V, dev), _t(Q, dev).requires_grad_(True)
    tv.requires_grad_(True)
    bvh = gpurt.Bvh(ctx, tv.detach(), _t(F, dev))
    g = torch.from_numpy(np.random.default_rng(1).uniform(0.5, 2.0, len(Q)).astype(np.float32)).to(dev)
    # the unsigned distance has a kink on the surface; the direction is recomputed in
    # FP64, so only queries within 1e-6 of it are left out (both sides)
    with torch.no_grad():
        d0 = gpurt.closest_point_distance(bvh, tq.detach(), tv.detach())[0]
    g = g * (d0 > 1e-6).float()
    dist, prim = gpurt.closest_point_distance(bvh, tq, tv)
    (dist * g).sum().backward()
    torch.cuda.synchronize()
    gq, gv = tq.grad.cpu().numpy(), tv.grad.cpu().numpy()
    # FP64 closed form at the GPU's own argmin (same triangle, exact arithmetic)
    p = prim.cpu().numpy()
    pt, bary = np.empty((len(Q), 3)), np.empty((len(Q), 3))
    V64, Q64 = V.astype(np.float64), Q.astype(np.float64)
    for i in range(len(Q)):
        tri = F[p[i]]
        _, pt[i], bary[i] = O.cp_tri_f64(Q64[i], V64[tri[0]], V64[tri[1]], V64[tri[2]])
    eq, ev = O.cpq_grad_f64(V64, F, Q64, p, pt, bary, g.cpu().numpy().astype(np.float64))
    # |dd/dq| = 1: the direction (q - p)/d comes from FP64, rounded once to FP32
    d = dist.detach().cpu().numpy()
    ok = d > 1e-6
    assert (d[ok] < 1e-4).sum() > 100  # queries a hair off the surface are in the test
    np.testing.assert_allclose(gq[ok], eq[ok], atol=1e-6, rtol=1e-6)
    nrm = np.linalg.norm(gq[ok] / g.cpu().numpy()[ok, None], axis=1)
    assert np.all(np.abs(nrm - 1.0) <= 1e-6)
    # vertex gradients: sums of ~10 FP32 atomic contributions per vertex
    np.testing.assert_allclose(gv, ev, atol=2e-5, rtol=1e-5)


def test_distance_gradient_matches_fp64_brute_force_and_central_differences(ctx, dev):
    V, F = _sphere(3, 2)  # 1,280 triangles: FP64 brute force and FD stay cheap
    rng = np.random.default_rng(4)
    Q = (W.random_dirs(rng, 1024) * rng.uniform(0.6, 1.6, 1024)[:, None]).astype(np.float32)
    tv = _t(V, dev).requires_grad_(True)
    tq = _t(Q, dev).requires_grad_(True)
    bvh = gpurt.Bvh(ctx, tv.detach(), _t(F, dev))
    dist, prim = gpurt.closest_point_distance(bvh, tq, tv)
    dist.sum().backward()
    torch.cuda.synchronize()
    gq, gv = tq.grad.cpu().numpy(), tv.grad.cpu().numpy()
    V64, Q64 = V.astype(np.float64), Q.astype(np.float64)
    bp, bpt, bd, bb = O.cpq_brute_f64(V64, F, Q64)
    eq, ev = O.cpq_grad_f64(V64, F, Q64, bp, bpt, bb)
    same = prim.cpu().numpy() == bp
    assert same.mean() > 0.97
    np.testing.assert_allclose(gq[same], eq[same], atol=1e-4)
    # central differences in FP64 (h = 1e-4), skipping stencils across a kink
    h = 1e-4
    fd_ok = 0
    for k in range(3):
        e = np.zeros(3); e[k] = h
        rp, rm = O.cpq_brute_f64(V64, F, Q64 + e), O.cpq_brute_f64(V64, F, Q64 - e)
        gp = O.cpq_grad_f64(V64, F, Q64 + e, rp[0], rp[1], rp[3])[0]
        gm = O.cpq_grad_f64(V64, F, Q64 - e, rm[0], rm[1], rm[3])[0]
        smooth = np.max(np.abs(gp - gm), axis=1) < 1e-2
        fd = (rp[2] - rm[2]) / (2 * h)
        np.testing.assert_allclose(gq[smooth, k], fd[smooth], atol=2e-4)
        fd_ok += smooth.sum()
    assert fd_ok > 0.95 * 3 * len(Q)
    # vertex gradient of sum(dist) by central differences on sampled vertices
    touched = np.unique(F[bp].ravel())
    for j in rng.choice(touched, 16, replace=False):
        for k in range(3):
            Vp = V64.copy(); Vp[j, k] += h
            Vm = V64.copy(); Vm[j, k] -= h
            fd = (O.cpq_brute_f64(Vp, F, Q64)[2].sum() - O.cpq_brute_f64(Vm, F, Q64)[2].sum()) / (2 * h)
            assert gv[j, k] == pytest.approx(fd, abs=5e-3)


def test_hit_distance_gradient(ctx, dev):
    V, F = _sphere(4, 3)
    rng = np.random.default_rng(10)
    n = 4096
    o = W.random_dirs(rng, n) * 3.0
    tgt = W.random_dirs(rng, n) * 0.4
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    to = _t(o.astype(np.float32), dev).requires_grad_(True)
    td = _t(d.astype(np.float32), dev).requires_grad_(True)
    tv = _t(V, dev).requires_grad_(True)
    bvh = gpurt.Bvh(ctx, tv.detach(), _t(F, dev))
    t, prim = gpurt.ray_hit_t(bvh, to, td, tv)
    hit = torch.isfinite(t)
    assert bool(hit.all())
    t.sum().backward()
    torch.cuda.synchronize()
    R = np.concatenate([to.detach().cpu().numpy(), np.full((n, 1), gpurt.EPS, np.float32),
                        td.detach().cpu().numpy(), np.full((n, 1), gpurt.LARGE_DIST, np.float32)], 1)
    V64, R64 = V.astype(np.float64), R.astype(np.float64)
    p = prim.cpu().numpy()
    # FP64 closed form at the GPU's hit triangle, FP64 hit (t, u, v) recomputed
    tt, uv = np.empty(n), np.empty((n, 2))
    for i in range(n):
        tri = F[p[i]]
        ok, tt[i], uv[i, 0], uv[i, 1] = O.ray_tri_f64(R64[i, :3], R64[i, 4:7], V64[tri[0]], V64[tri[1]], V64[tri[2]])
    ego, egd, egv = O.ray_grad_f64(V64, F, R64, p, tt, uv)
    np.testing.assert_allclose(to.grad.cpu().numpy(), ego, rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(td.grad.cpu().numpy(), egd, rtol=1e-3, atol=3e-3)
    np.testing.assert_allclose(tv.grad.cpu().numpy(), egv, atol=1e-2, rtol=1e-3)
    # central differences on the origins of 256 rays (FP64 brute force)
    idx = np.arange(0, n, 16)
    h = 1e-5
    for k in range(3):
        Rp, Rm = R64[idx].copy(), R64[idx].copy()
        Rp[:, k] += h
        Rm[:, k] -= h
        fd = (O.ray_closest_brute_f64(V64, F, Rp)[1] - O.ray_closest_brute_f64(V64, F, Rm)[1]) / (2 * h)
        smooth = np.abs(fd) < 50
        np.testing.assert_allclose(to.grad.cpu().numpy()[idx][smooth, k], fd[smooth], atol=2e-3, rtol=1e-3)


def test_backward_reuses_saved_ids_no_retraversal(ctx, dev):
    """The backward call takes the saved ids and barycentrics: feeding it the forward
    outputs reproduces the autograd gradient exactly (same kernel, same inputs)."""
    V, F = _sphere(3, 5)
    Q = np.random.default_rng(2).uniform(-1.3, 1.3, (3000, 3)).astype(np.float32)
    bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev))
    tq = _t(Q, dev)
    prim, pt, dist, bary = bvh.closest_point(tq)
    g = torch.ones(len(Q), device=dev)
    gq = torch.empty_like(tq)
    bvh.closest_point_bwd(tq, prim, bary, pt, dist, g, gq, None)
    tq2 = tq.clone().requires_grad_(True)
    d2, _ = gpurt.closest_point_distance(bvh, tq2, _t(V, dev))
    d2.sum().backward()
    torch.cuda.synchronize()
    assert torch.equal(gq, tq2.grad)


def test_misses_and_zero_distance_give_zero_gradient(ctx, dev):
    V, F = _sphere(2, 1)
    # rays pointing away from the sphere miss; a query on a vertex has distance 0
    o = torch.tensor([[0, 0, 5.0], [0, 0, -5.0]], device=dev, requires_grad=True)
    d = torch.tensor([[0, 0, 1.0], [0, 0, -1.0]], device=dev)
    tv = _t(V, dev).requires_grad_(True)
    bvh = gpurt.Bvh(ctx, tv.detach(), _t(F, dev))
    t, prim = gpurt.ray_hit_t(bvh, o, d, tv)
    assert prim.tolist() == [-1, -1]
    torch.where(torch.isfinite(t), t, torch.zeros_like(t)).sum().backward()
    assert torch.count_nonzero(o.grad) == 0 and torch.count_nonzero(tv.grad) == 0
    q = _t(V[:5], dev).requires_grad_(True)
    dist, _ = gpurt.closest_point_distance(bvh, q, tv)
    assert torch.count_nonzero(dist) == 0
    dist.sum().backward()
    assert torch.count_nonzero(q.grad) == 0


@pytest.mark.parametrize("mode", ["atomic", "fixed", "auto"])
def test_plain_vertex_gradient_sums(dev, mode):
    """The same backward, run five times on the differentiable workload's kind of input
    (many queries per vertex): the sum over every term matches the FP64 closed form in
    every mode, and fixed-order sums give identical bits every run."""
    det = mode == "fixed"
    ctx = gpurt.Context(0, vertex_sums=mode)
    V, F = _sphere(6, 3)
    Q = W.surface_samples(V.astype(np.float64), F, np.random.default_rng(4), 200000, 0.005).astype(np.float32)
    rng = np.random.default_rng(5)
    o = W.random_dirs(rng, 50000) * 3.0
    d = -o / np.linalg.norm(o, axis=1, keepdims=True)
    R = W.pack_rays(o, d, 1e-5, 1e7)
    tv, tq, tr = _t(V, dev), _t(Q, dev), _t(R, dev)
    bvh = gpurt.Bvh(ctx, tv, _t(F, dev))
    prim, pt, dist, bary = bvh.closest_point(tq)
    rp, rt, ruv = bvh.trace_closest(tr)
    g1 = torch.ones(len(Q), device=dev)
    g2 = torch.ones(len(R), device=dev)
    runs = []
    for _ in range(5):
        gv = torch.zeros((V.shape[0], 3), device=dev)
        bvh.closest_point_bwd(tq, prim, bary, pt, dist, g1, None, gv)
        bvh.trace_bwd(tr, rp, rt, ruv, g2, None, None, gv)
        torch.cuda.synchronize()
        runs.append(gv.cpu().numpy())
    if det:
        for r in runs[1:]:
            assert np.array_equal(r, runs[0])
    p = prim.cpu().numpy()
    _, ev = O.cpq_grad_f64(V.astype(np.float64), F, Q.astype(np.float64), p)
    _, _, erv = O.ray_grad_f64(V.astype(np.float64), F, R.astype(np.float64), rp.cpu().numpy(),
                            rt.cpu().numpy().astype(np.float64), ruv.cpu().numpy().astype(np.float64))
    scale = np.abs(ev + erv).max()
    for r in runs:
        np.testing.assert_allclose(r, ev + erv, atol=1e-5 * scale, rtol=1e-4)
    bvh.close()
    ctx.close()


def test_vertex_sum_modes_are_checked(dev):
    ctx = gpurt.Context(0)
    assert ctx.vertex_sums == "auto"
    with pytest.raises(ValueError):
        ctx.vertex_sums = "sorted"
    from gpurt import _lib
    assert _lib.load().gpurt_ctx_set_vertex_sums(ctx.handle, 7) == _lib.GPURT_E_INVALID
    ctx.close()


def test_auto_vertex_sums_pick_the_fixed_order_for_large_calls(dev):
    """From 2^21 terms (here 1,048,576 queries = 3,145,728 terms) the default mode uses
    the fixed-order sums: its gradients equal the fixed mode's bit for bit, and match
    the atomic mode's within rounding."""
    V, F = _sphere(6, 11)
    Q = W.surface_samples(V.astype(np.float64), F, np.random.default_rng(12), 1 << 20, 0.01).astype(np.float32)
    out = {}
    for mode in ("auto", "fixed", "atomic"):
        ctx = gpurt.Context(0, vertex_sums=mode)
        bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev))
        tq = _t(Q, dev)
        prim, pt, dist, bary = bvh.closest_point(tq)
        gv = torch.zeros((V.shape[0], 3), device=dev)
        bvh.closest_point_bwd(tq, prim, bary, pt, dist, torch.ones(len(Q), device=dev), None, gv)
        torch.cuda.synchronize()
        out[mode] = gv.cpu().numpy()
        bvh.close()
        ctx.close()
    assert np.array_equal(out["auto"], out["fixed"])
    scale = np.abs(out["fixed"]).max()
    np.testing.assert_allclose(out["atomic"], out["fixed"], atol=1e-5 * scale, rtol=1e-4)
