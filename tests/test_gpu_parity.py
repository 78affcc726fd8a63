"""GPU parity: the HIP kernels (through the C ABI) against the oracle on the same inputs.

Integer outputs (primitive ids, occlusion) and every FP32 output of the forward queries
are compared BIT-EXACTLY with the oracle's FP32 brute force: the kernels and the oracle
perform the same rounded operations, and the traversal never culls a subtree that may
hold an equal-or-better primitive, so the answer cannot depend on the tree.

Both traversed node encodings are covered: FP32 child boxes (node_format 1) and 8-bit
child boxes relative to the node (node_format 2), whose decoded boxes contain the FP32
ones, so they change the work done but never the answers.
"""
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


def _cpq_check(ctx, dev, V, F, Q, **kw):
    bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev), **{k: v for k, v in kw.items() if k in
                                                      ("node_width", "leaf_size", "node_format")})
    qkw = {k: v for k, v in kw.items() if k in ("stack_depth", "waves_per_cu")}
    prim, pt, dist, bary = _np(*bvh.closest_point(_t(Q, dev), **qkw))
    ep, ept, ed, eb = O.cpq_brute_f32(V, F, Q)
    assert np.array_equal(prim, ep), f"{(prim != ep).sum()} ids differ"
    assert np.array_equal(dist, ed)
    assert np.array_equal(pt, ept)
    assert np.array_equal(bary, eb)
    return bvh


def test_cpq_small_full_workload_bitexact(ctx, dev):
    w = W.cpq_small()
    print("cpq_small", w["meta"], flush=True)
    bvh = _cpq_check(ctx, dev, w["verts"], w["tris"], w["queries"])
    info = bvh.info()
    assert info["n_tris"] == 5120 and info["leaf_size"] == 4 and info["node_width"] == 4


FORMATS = [1, 2]


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("width", [2, 4, 8])
@pytest.mark.parametrize("stack", [8, 16, 32, 64])
def test_cpq_widths_and_stacks(ctx, dev, width, stack, fmt):
    w = W.cpq_small(n_queries=4096)
    print("cpq width", width, "stack", stack, "format", fmt, flush=True)
    _cpq_check(ctx, dev, w["verts"], w["tris"], w["queries"], node_width=width, stack_depth=stack,
               node_format=fmt)


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("leaf", [1, 2, 8])
def test_cpq_leaf_sizes(ctx, dev, leaf, fmt):
    w = W.cpq_small(n_queries=4096)
    _cpq_check(ctx, dev, w["verts"], w["tris"], w["queries"], leaf_size=leaf, node_format=fmt)


def _ray_check(ctx, dev, V, F, R, RA=None, **kw):
    bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev), **{k: v for k, v in kw.items() if k in
                                                      ("node_width", "leaf_size", "node_format")})
    qkw = {k: v for k, v in kw.items() if k in ("stack_depth", "waves_per_cu", "leaf_trigger", "refill_trigger")}
    prim, t, uv = _np(*bvh.trace_closest(_t(R, dev), **qkw))
    ep, et, euv = O.ray_closest_brute_f32(V, F, R)
    assert np.array_equal(prim, ep), f"{(prim != ep).sum()} ids differ"
    assert np.array_equal(t, et)
    assert np.array_equal(uv, euv)
    RA = R if RA is None else RA
    occ, = _np(bvh.trace_any(_t(RA, dev), **qkw))
    assert np.array_equal(occ, O.ray_any_brute_f32(V, F, RA))
    return prim


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("width", [2, 4, 8])
def test_rays_small_scene_bitexact(ctx, dev, width, fmt):
    w = W.rays_scene(n_rays=8192, grid=101, n_spheres=6)
    print("rays", w["meta"], "width", width, "format", fmt, flush=True)
    prim = _ray_check(ctx, dev, w["verts"], w["tris"], w["rays"], w["rays_any"], node_width=width,
                      node_format=fmt)
    assert 0.2 < (prim >= 0).mean() < 0.9


@pytest.mark.parametrize("kind", ["rays", "cpq"])
def test_stack_spills_keep_answers_exact(ctx, dev, kind):
    """8-wide nodes with the smallest LDS stack (8 entries): a node visit pushes up to 7
    children, so pushes often do not fit below the last LDS slot; they then take the
    per-push path, which sends the entries beyond the LDS slots to the device spill buffer.
    The counting kernel (same tree, same options) confirms that entries spilled, and the
    answers still equal the brute force."""
    st = torch.zeros(4, dtype=torch.int64, device=dev)
    if kind == "rays":
        w = W.rays_scene(n_rays=8192, grid=101, n_spheres=6)
        _ray_check(ctx, dev, w["verts"], w["tris"], w["rays"], w["rays_any"], node_width=8, stack_depth=8)
        bvh = gpurt.Bvh(ctx, _t(w["verts"], dev), _t(w["tris"], dev), node_width=8)
        bvh.trace_closest(_t(w["rays"], dev), stack_depth=8, stats=st)
    else:
        w = W.cpq_small(n_queries=4096)
        bvh = _cpq_check(ctx, dev, w["verts"], w["tris"], w["queries"], node_width=8, stack_depth=8)
        bvh.closest_point(_t(w["queries"], dev), stack_depth=8, stats=st)
    nodes, tris, spills, _ = st.cpu().tolist()
    print(kind, "nodes", nodes, "tris", tris, "spills", spills, flush=True)
    assert spills > 0


@pytest.mark.parametrize("refill", [1, 13, 64])
@pytest.mark.parametrize("leaf", [1, 64])
def test_wave_scheduling_options_keep_answers(ctx, dev, leaf, refill):
    """leaf_trigger (parked leaves before a leaf step) and refill_trigger (idle lanes before
    finished rays are written and new ones fetched) reorder work inside a wave only."""
    w = W.rays_scene(n_rays=8192, grid=101, n_spheres=6)
    print("rays leaf_trigger", leaf, "refill_trigger", refill, flush=True)
    _ray_check(ctx, dev, w["verts"], w["tris"], w["rays"], w["rays_any"], leaf_trigger=leaf,
               refill_trigger=refill, waves_per_cu=4)


def test_repeated_launches_are_identical(ctx, dev):
    """Work donation groups lanes by timing; the answers must not depend on it: repeated
    launches (full grid, few queries, few waves) all equal the brute force."""
    w = W.rays_scene(n_rays=20000, grid=101, n_spheres=6)
    V, F, R = w["verts"], w["tris"], w["rays"]
    bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev))
    ep, et, euv = O.ray_closest_brute_f32(V, F, R)
    eo = O.ray_any_brute_f32(V, F, w["rays_any"])
    c = W.cpq_small(n_queries=3000)
    cb = gpurt.Bvh(ctx, _t(c["verts"], dev), _t(c["tris"], dev))
    cp = O.cpq_brute_f32(c["verts"], c["tris"], c["queries"])[0]
    tr, tra, tq = _t(R, dev), _t(w["rays_any"], dev), _t(c["queries"], dev)
    for it in range(6):
        for kw in ({}, dict(waves_per_cu=4), dict(waves_per_cu=16, refill_trigger=8)):
            p, t, uv = _np(*bvh.trace_closest(tr, **kw))
            assert np.array_equal(p, ep) and np.array_equal(t, et) and np.array_equal(uv, euv), (it, kw)
            occ, = _np(bvh.trace_any(tra, **kw))
            assert np.array_equal(occ, eo), (it, kw)
            prim, = _np(cb.closest_point(tq, **{k: v for k, v in kw.items() if k != "refill_trigger"})[0])
            assert np.array_equal(prim, cp), (it, kw)


@pytest.mark.parametrize("stride", [3, 4, 12])
def test_vertex_strides(ctx, dev, stride):
    """12-, 16- and 48-byte vertex records (48 = the reference's Mesh::Vertex)."""
    w = W.cpq_small(n_queries=2048)
    V, F, Q = w["verts"], w["tris"], w["queries"]
    Vs = np.zeros((V.shape[0], stride), np.float32)
    Vs[:, :3] = V
    Vs[:, 3:] = 7.0  # normals/tangents/uv in the other columns must be ignored
    bvh = gpurt.Bvh(ctx, _t(Vs, dev), _t(F, dev))
    prim, dist = _np(*bvh.closest_point(_t(Q, dev))[:3:2])
    ep, _, ed, _ = O.cpq_brute_f32(V, F, Q)
    assert np.array_equal(prim, ep) and np.array_equal(dist, ed)


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("n_tris", [1, 2, 3, 5, 17])
def test_tiny_meshes(ctx, dev, n_tris, fmt):
    rng = np.random.default_rng(n_tris)
    V = rng.normal(size=(3 * n_tris, 3)).astype(np.float32)
    F = np.arange(3 * n_tris, dtype=np.int32).reshape(-1, 3)
    Q = rng.normal(size=(777, 3)).astype(np.float32)
    _cpq_check(ctx, dev, V, F, Q, node_format=fmt)
    o = rng.normal(size=(777, 3)) * 3
    d = W.random_dirs(rng, 777)
    _ray_check(ctx, dev, V, F, W.pack_rays(o, d, 0.0, 1e7), node_format=fmt)


@pytest.mark.parametrize("width", [2, 4, 8])
def test_quantized_nodes_are_smaller_and_exact(ctx, dev, width):
    """8-bit child boxes: smaller nodes, boxes that can only be looser (so never fewer
    node visits than FP32 boxes on the same tree), and bit-identical answers."""
    w = W.rays_scene(n_rays=8192, grid=101, n_spheres=6)
    V, F, R = _t(w["verts"], dev), _t(w["tris"], dev), _t(w["rays"], dev)
    f32 = gpurt.Bvh(ctx, V, F, node_width=width, node_format=1)
    q8 = gpurt.Bvh(ctx, V, F, node_width=width, node_format=2)
    i32, i8 = f32.info(), q8.info()
    assert i32["node_format"] == 1 and i8["node_format"] == 2
    assert i32["node_bytes"] == 32 * width and i8["node_bytes"] == {2: 48, 4: 64, 8: 96}[width]
    assert i32["n_nodes"] == i8["n_nodes"]
    s32 = torch.zeros(4, dtype=torch.int64, device=dev)
    s8 = torch.zeros(4, dtype=torch.int64, device=dev)
    a = _np(*f32.trace_closest(R, stats=s32))
    b = _np(*q8.trace_closest(R, stats=s8))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    n32, t32 = s32[:2].tolist()
    n8, t8 = s8[:2].tolist()
    assert n8 >= n32 and t8 >= t32
    # the looser boxes cost little: at most a few percent more visits here
    assert n8 <= 1.2 * n32, (n8, n32)


@pytest.mark.parametrize("fmt", FORMATS)
def test_exact_ties_go_to_lowest_index(ctx, dev, fmt):
    a, b, c = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    V = np.array([a, b, c], np.float32)
    F = np.array([[2, 0, 1], [0, 1, 2], [1, 2, 0]] * 5, np.int32)  # 15 copies of one triangle
    Q = np.array([[0.25, 0.25, 1], [-1, -1, 0], [0.5, 0.5, 0], [3, 3, 3]], np.float32)
    bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev), leaf_size=2, node_format=fmt)
    prim, = _np(bvh.closest_point(_t(Q, dev))[0])
    assert prim.tolist() == [0, 0, 0, 0]
    R = W.pack_rays(np.array([[0.2, 0.3, 1.0], [0.2, 0.3, -1.0]]),
                    np.array([[0, 0, -1.0], [0, 0, 1.0]]), 0.0, 10.0)
    p, t, _ = _np(*bvh.trace_closest(_t(R, dev)))
    assert p.tolist() == [0, 0] and t.tolist() == [1.0, 1.0]


@pytest.mark.parametrize("n", [1, 2, 5, 70])
def test_few_queries_shared_by_the_whole_wave(ctx, dev, n):
    """A handful of queries leave most lanes of the wave idle from the start, so nearly
    every subtree is handed to a helper lane (work donation) and the answers are merged
    across lanes: still the brute-force answers, ties included."""
    w = W.rays_scene(n_rays=n, grid=101, n_spheres=6)
    print("few queries", n, flush=True)
    _ray_check(ctx, dev, w["verts"], w["tris"], w["rays"], w["rays_any"])
    c = W.cpq_small(n_queries=n)
    _cpq_check(ctx, dev, c["verts"], c["tris"], c["queries"])
    # queries at the centre of the sphere: every triangle is nearly as close as the best
    _cpq_check(ctx, dev, c["verts"], c["tris"], np.zeros((n, 3), np.float32) + np.float32(0.01))
    # 15 copies of one triangle per leaf pair: the tie must resolve to index 0 across helpers
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    F = np.array([[2, 0, 1], [0, 1, 2], [1, 2, 0]] * 40, np.int32)
    bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev), leaf_size=1)
    Q = np.tile(np.array([[0.25, 0.25, 1]], np.float32), (n, 1))
    prim, = _np(bvh.closest_point(_t(Q, dev))[0])
    assert (prim == 0).all()
    R = W.pack_rays(np.tile([[0.2, 0.3, 1.0]], (n, 1)), np.tile([[0, 0, -1.0]], (n, 1)), 0.0, 10.0)
    p, t, _ = _np(*bvh.trace_closest(_t(R, dev)))
    assert (p == 0).all() and (t == 1.0).all()


@pytest.mark.parametrize("fmt", FORMATS)
def test_shared_vertices_and_edges_of_a_closed_mesh(ctx, dev, fmt):
    """Queries exactly at vertices and on edges of an icosphere: exact FP32 ties
    between neighbouring triangles resolve to the lowest index."""
    v, f = W.icosphere(3)
    V, F = v.astype(np.float32), f.astype(np.int32)
    mid = ((V[F[:, 0]] + V[F[:, 1]]) * np.float32(0.5)).astype(np.float32)
    Q = np.concatenate([V, mid, V * np.float32(1.5), mid * np.float32(0.5)]).astype(np.float32)
    _cpq_check(ctx, dev, V, F, Q, node_format=fmt)
    # rays aimed at vertices and edge midpoints from outside
    o = np.concatenate([V, mid]) * 3.0
    d = -o / np.linalg.norm(o, axis=1, keepdims=True)
    _ray_check(ctx, dev, V, F, W.pack_rays(o, d, 0.0, 1e7), node_format=fmt)


@pytest.mark.parametrize("fmt", FORMATS)
def test_degenerate_triangles(ctx, dev, fmt):
    rng = np.random.default_rng(5)
    w = W.cpq_small(n_queries=3000)
    V, F = w["verts"], w["tris"].copy()
    k = rng.choice(F.shape[0], 300, replace=False)
    F[k[:100], 1] = F[k[:100], 0]            # two equal vertices
    F[k[100:200], 2] = F[k[100:200], 0]
    F[k[200:], 1] = F[k[200:], 0]
    F[k[200:], 2] = F[k[200:], 0]            # a point triangle
    _cpq_check(ctx, dev, V, F, w["queries"], node_format=fmt)
    o = rng.uniform(-1.5, 1.5, (3000, 3))
    _ray_check(ctx, dev, V, F, W.pack_rays(o, W.random_dirs(rng, 3000), 1e-5, 1e7), node_format=fmt)


def test_ray_interval_bounds(ctx, dev):
    """tmin/tmax are inclusive (rt.rgen:257-291): hits at exactly tmax count."""
    w = W.rays_scene(n_rays=4096, grid=61, n_spheres=3)
    V, F, R = w["verts"], w["tris"], w["rays"]
    ep, et, _ = O.ray_closest_brute_f32(V, F, R)
    hit = ep >= 0
    R2 = R[hit].copy()
    R2[:, 7] = et[hit]                   # tmax == t: still a hit
    R3 = R[hit].copy()
    R3[:, 7] = np.nextafter(et[hit], np.float32(-np.inf))  # just below: the hit is excluded
    for RR in (R2, R3):
        _ray_check(ctx, dev, V, F, RR)


def test_out_of_range_index_is_rejected(ctx, dev):
    V = np.zeros((3, 3), np.float32)
    F = np.array([[0, 1, 3]], np.int32)
    with pytest.raises(gpurt.GpurtError) as e:
        gpurt.Bvh(ctx, _t(V, dev), _t(F, dev))
    assert e.value.code == -1


def test_bad_options_are_rejected(ctx, dev):
    w = W.cpq_small(n_queries=16)
    with pytest.raises(gpurt.GpurtError):
        gpurt.Bvh(ctx, _t(w["verts"], dev), _t(w["tris"], dev), node_width=3)
    with pytest.raises(gpurt.GpurtError):
        gpurt.Bvh(ctx, _t(w["verts"], dev), _t(w["tris"], dev), node_format=7)
    bvh = gpurt.Bvh(ctx, _t(w["verts"], dev), _t(w["tris"], dev))
    with pytest.raises(gpurt.GpurtError):
        bvh.closest_point(_t(w["queries"], dev), stack_depth=24)
    rays = torch.zeros((4, 8), device=dev)
    for bad in (dict(refill_trigger=65), dict(refill_trigger=-1), dict(leaf_trigger=65)):
        with pytest.raises(gpurt.GpurtError):
            bvh.trace_closest(rays, **bad)


def test_empty_query_sets(ctx, dev):
    w = W.cpq_small(n_queries=16)
    bvh = gpurt.Bvh(ctx, _t(w["verts"], dev), _t(w["tris"], dev))
    prim, pt, dist, bary = bvh.closest_point(torch.empty((0, 3), device=dev))
    assert prim.numel() == 0
    p, t, uv = bvh.trace_closest(torch.empty((0, 8), device=dev))
    assert p.numel() == 0


def test_stats_count_visits(ctx, dev):
    w = W.cpq_small(n_queries=4096)
    bvh = gpurt.Bvh(ctx, _t(w["verts"], dev), _t(w["tris"], dev))
    st = torch.zeros(4, dtype=torch.int64, device=dev)
    prim, = _np(bvh.closest_point(_t(w["queries"], dev), stats=st)[0])
    nodes, tris, spills, _ = st.cpu().tolist()
    assert nodes >= 4096 and tris >= 4096
    ep, _, _, _ = O.cpq_brute_f32(w["verts"], w["tris"], w["queries"])
    assert np.array_equal(prim, ep)
    # the BVH prunes: far fewer triangle tests than brute force
    assert tris < 0.05 * 4096 * 5120


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("width", [2, 4, 8])
def test_refit_after_moving_vertices(ctx, dev, width, fmt):
    """Refit (same triangles, moved vertices) answers exactly like brute force on the
    moved mesh, for closest points and rays, and twice in a row (quantized nodes are
    re-encoded from the refitted boxes)."""
    w = W.cpq_small(n_queries=3000)
    V, F, Q = w["verts"], w["tris"], w["queries"]
    tv = _t(V, dev)
    bvh = gpurt.Bvh(ctx, tv, _t(F, dev), node_width=width, node_format=fmt)
    rng = np.random.default_rng(9)
    o = rng.uniform(-1.5, 1.5, (3000, 3))
    R = W.pack_rays(o, W.random_dirs(rng, 3000), 1e-5, 1e7)
    for step in range(2):
        V = (V * np.float32(1.0 + 0.2 * (step + 1)) +
             rng.normal(0, 0.02, V.shape).astype(np.float32)).astype(np.float32)
        tv.copy_(_t(V, dev))
        bvh.refit(tv)
        prim, pt, dist, bary = _np(*bvh.closest_point(_t(Q, dev)))
        ep, ept, ed, eb = O.cpq_brute_f32(V, F, Q)
        assert np.array_equal(prim, ep) and np.array_equal(dist, ed) and np.array_equal(pt, ept)
        p, t, uv = _np(*bvh.trace_closest(_t(R, dev)))
        rp, rt, ruv = O.ray_closest_brute_f32(V, F, R)
        assert np.array_equal(p, rp) and np.array_equal(t, rt)
        # backward uses the refitted positions
        g = torch.ones(len(Q), device=dev)
        gv = torch.zeros((V.shape[0], 3), device=dev)
        tq = _t(Q, dev)
        prim_t, pt_t, dist_t, bary_t = bvh.closest_point(tq)
        bvh.closest_point_bwd(tq, prim_t, bary_t, pt_t, dist_t, g, None, gv)
        _, egv = O.cpq_grad_f64(V.astype(np.float64), F, Q.astype(np.float64), ep)
        np.testing.assert_allclose(gv.cpu().numpy(), egv, atol=5e-5)


@pytest.mark.parametrize("tsize", [5, 6, 7])
def test_treelet_builds_are_bitexact(ctx, dev, tsize):
    """SAH treelet restructuring changes the tree, never the answers."""
    w = W.cpq_small(n_queries=4096)
    V, F, Q = w["verts"], w["tris"], w["queries"]
    bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev), treelet_size=tsize)
    prim, pt, dist, bary = _np(*bvh.closest_point(_t(Q, dev)))
    ep, ept, ed, eb = O.cpq_brute_f32(V, F, Q)
    assert np.array_equal(prim, ep) and np.array_equal(dist, ed) and np.array_equal(bary, eb)
    ws = W.rays_scene(n_rays=4096, grid=81, n_spheres=4)
    _ray_check(ctx, dev, ws["verts"], ws["tris"], ws["rays"], ws["rays_any"])
    bvh2 = gpurt.Bvh(ctx, _t(ws["verts"], dev), _t(ws["tris"], dev), treelet_size=tsize)
    p, t, uv = _np(*bvh2.trace_closest(_t(ws["rays"], dev)))
    rp, rt, ruv = O.ray_closest_brute_f32(ws["verts"], ws["tris"], ws["rays"])
    assert np.array_equal(p, rp) and np.array_equal(t, rt) and np.array_equal(uv, ruv)


def test_treelets_lower_the_sah_cost(ctx, dev):
    for V, F in [W.displaced_icosphere(6, np.random.default_rng(1)),
                 W.terrain(201, np.random.default_rng(2))]:
        tv, tf = _t(V.astype(np.float32), dev), _t(F.astype(np.int32), dev)
        base = gpurt.Bvh(ctx, tv, tf).info()["sah_cost"]
        opt = gpurt.Bvh(ctx, tv, tf, treelet_size=7).info()["sah_cost"]
        print("sah", base, "->", opt)
        assert 0 < opt < base * 0.99
