"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the test oracle.

The reference has no tests, fixtures or CPU BVH (its BVH lives in the Vulkan driver),
so the oracle is pinned by: known answers a plausible mistake breaks (wrong region,
swapped barycentric, sign error, dropped term), an independent FP64 formulation of
each query, brute force over all triangles, sortedness/permutation invariants of the
LBVH, and FP64 central differences for the gradients.
"""
import numpy as np
import pytest

from gpurt import workloads as W
from oracle import oracle as O

TRI = (np.array([0, 0, 0.]), np.array([1, 0, 0.]), np.array([0, 1, 0.]))

# (query, closest point, squared distance, barycentrics) on the unit right triangle
CP_KAT = [
    ((0.25, 0.25, 1.0), (0.25, 0.25, 0.0), 1.0, (0.5, 0.25, 0.25)),   # face interior
    ((-1.0, -1.0, 0.0), (0.0, 0.0, 0.0), 2.0, (1.0, 0.0, 0.0)),       # vertex a
    ((2.0, -1.0, 0.0), (1.0, 0.0, 0.0), 2.0, (0.0, 1.0, 0.0)),        # vertex b
    ((-1.0, 2.0, 0.0), (0.0, 1.0, 0.0), 2.0, (0.0, 0.0, 1.0)),        # vertex c
    ((0.5, -1.0, 0.0), (0.5, 0.0, 0.0), 1.0, (0.5, 0.5, 0.0)),        # edge ab
    ((-1.0, 0.5, 0.0), (0.0, 0.5, 0.0), 1.0, (0.5, 0.0, 0.5)),        # edge ac
    ((1.0, 1.0, 0.0), (0.5, 0.5, 0.0), 0.5, (0.0, 0.5, 0.5)),         # edge bc
    ((0.25, 0.25, -2.0), (0.25, 0.25, 0.0), 4.0, (0.5, 0.25, 0.25)),  # other side
    ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0, (1.0, 0.0, 0.0)),         # on a vertex
]


@pytest.mark.parametrize("p,cp,d2,bary", CP_KAT)
def test_closest_point_known_answers(p, cp, d2, bary):
    a, b, c = TRI
    for fn in (O.cp_tri_f32, O.cp_tri_f64):
        got_d2, got_cp, got_b = fn(np.array(p), a, b, c)
        assert got_d2 == pytest.approx(d2, abs=1e-7)
        np.testing.assert_allclose(got_cp, cp, atol=1e-7)
        np.testing.assert_allclose(got_b, bary, atol=1e-7)
        # barycentrics reproduce the closest point
        np.testing.assert_allclose(got_b[0] * a + got_b[1] * b + got_b[2] * c, cp, atol=1e-6)


def test_closest_point_f32_vs_f64_and_dense_sampling():
    rng = np.random.default_rng(123)
    for _ in range(400):
        a, b, c = rng.normal(size=(3, 3))
        p = rng.normal(size=3) * 2
        d2f, cpf, bf = O.cp_tri_f32(p, a, b, c)
        a32, b32, c32, p32 = (x.astype(np.float32).astype(np.float64) for x in (a, b, c, p))
        d2d, cpd, bd = O.cp_tri_f64(p32, a32, b32, c32)
        assert abs(np.sqrt(d2f) - np.sqrt(d2d)) <= 2e-6 * (1 + np.abs(p32).max())
        # brute force over a dense barycentric grid never beats the closed form
        s = np.linspace(0, 1, 61)
        u, v = np.meshgrid(s, s)
        m = u + v <= 1
        pts = (1 - u[m] - v[m])[:, None] * a32 + u[m][:, None] * b32 + v[m][:, None] * c32
        dmin = np.min(np.sum((pts - p32) ** 2, axis=1))
        assert d2d <= dmin + 1e-12
        assert dmin - d2d <= 0.05 * (1 + d2d)


MT_KAT = [
    # origin, direction, hit, t, u, v   (u weights pb, v weights pc: rt.rchit:13)
    ((0.2, 0.3, 1.0), (0, 0, -1.0), True, 1.0, 0.2, 0.3),
    ((0.2, 0.3, -1.0), (0, 0, 1.0), True, 1.0, 0.2, 0.3),   # back face: no culling
    ((0.8, 0.8, 1.0), (0, 0, -1.0), False, 0, 0, 0),        # u + v > 1
    ((-0.1, 0.3, 1.0), (0, 0, -1.0), False, 0, 0, 0),       # u < 0
    ((0.0, 0.0, 1.0), (1.0, 0, 0), False, 0, 0, 0),         # parallel: det == 0
    ((0.5, 0.0, 1.0), (0, 0, -1.0), True, 1.0, 0.5, 0.0),   # on edge ab (closed)
    ((0.2, 0.3, 3.0), (0, 0, -2.0), True, 1.5, 0.2, 0.3),   # t in units of |d|
    ((0.2, 0.3, 1.0), (0, 0, 1.0), True, -1.0, 0.2, 0.3),   # behind: t < 0 reported
]


@pytest.mark.parametrize("o,d,hit,t,u,v", MT_KAT)
def test_ray_triangle_known_answers(o, d, hit, t, u, v):
    a, b, c = TRI
    for fn in (O.mt_f32, O.ray_tri_f64):
        ok, gt, gu, gv = fn(np.array(o), np.array(d), a, b, c)
        assert ok == hit
        if hit:
            assert gt == pytest.approx(t, abs=1e-6)
            assert gu == pytest.approx(u, abs=1e-6)
            assert gv == pytest.approx(v, abs=1e-6)


def test_ray_interval_is_closed():
    a, b, c = TRI
    V = np.stack([a, b, c]).astype(np.float32)
    F = np.array([[0, 1, 2]], np.uint32)
    o, d = np.array([0.2, 0.3, 1.0]), np.array([0, 0, -1.0])
    for tmin, tmax, expect in [(0.0, 1.0, 1), (1.0, 2.0, 1), (0.0, np.nextafter(np.float32(1), np.float32(0)), 0),
                               (np.nextafter(np.float32(1), np.float32(2)), 5.0, 0)]:
        R = W.pack_rays(o[None], d[None], tmin, tmax)
        assert O.ray_any_brute_f32(V, F, R)[0] == expect
        prim, t, _ = O.ray_closest_brute_f32(V, F, R)
        assert (prim[0] == 0) == bool(expect)
        if expect:
            assert t[0] == 1.0


def test_lowest_index_wins_exact_ties():
    # the same triangle three times: every query and ray must report index 0
    a, b, c = TRI
    V = np.stack([a, b, c]).astype(np.float32)
    F = np.array([[0, 1, 2], [1, 2, 0], [2, 0, 1]], np.uint32)
    Q = np.array([[0.25, 0.25, 1.0], [-1, -1, 0]], np.float32)
    prim, _, _, _ = O.cpq_brute_f32(V, F, Q)
    assert prim.tolist() == [0, 0]
    bvh = O.CpuBvh(V, F, 1)
    assert bvh.closest_point(Q)[0].tolist() == [0, 0]
    R = W.pack_rays(np.array([[0.2, 0.3, 1.0]]), np.array([[0, 0, -1.0]]), 0.0, 10.0)
    assert O.ray_closest_brute_f32(V, F, R)[0].tolist() == [0]
    assert bvh.trace_closest(R)[0].tolist() == [0]


def _scale(V):
    return float(np.abs(V).max())


def test_cpq_brute_f32_vs_f64():
    w = W.cpq_small(n_queries=2048)
    V, F, Q = w["verts"], w["tris"], w["queries"]
    p32, pt32, d32, b32 = O.cpq_brute_f32(V, F, Q)
    p64, pt64, d64, b64 = O.cpq_brute_f64(V, F, Q)
    s = _scale(V)
    assert np.max(np.abs(d32 - d64)) <= 1e-6 * s
    same = p32 == p64
    assert np.max(np.abs(pt32[same] - pt64[same])) <= 1e-6 * s
    # where the ids differ the two triangles are an FP32 near-tie in FP64 too, and the
    # FP32 point is the FP64 closest point of the triangle FP32 chose
    diff = np.nonzero(~same)[0]
    assert diff.size < 0.1 * len(Q)
    Fd = F.astype(np.int64)
    V64 = V.astype(np.float64)
    for i in diff:
        tri = Fd[p32[i]]
        d2, cp, _ = O.cp_tri_f64(Q[i].astype(np.float64), V64[tri[0]], V64[tri[1]], V64[tri[2]])
        assert np.sqrt(d2) - d64[i] <= 1e-6 * s
        assert np.max(np.abs(cp - pt32[i])) <= 1e-6 * s
    # barycentrics reproduce the closest point
    rec = np.einsum("ij,ijk->ik", b64, V64[Fd[p64]])
    np.testing.assert_allclose(rec, pt64, atol=1e-12)


def test_reported_distance_is_the_fp64_distance_rounded():
    """The reported point / distance / barycentrics of the chosen triangle come from the
    FP64 refinement: against the independent FP64 formulation (plane projection +
    clamped edges) the distance is within one FP32 ulp (rounding of two FP64 values that
    agree to ~1e-15) and relative error <= 1e-7 over every query -- including queries a
    hair off the surface, where the FP32 formula alone loses digits -- and where the ids
    agree the point matches to 1e-7 of the scene scale."""
    w = W.cpq_small(n_queries=4096)
    V, F, Q = w["verts"], w["tris"], w["queries"]
    rng = np.random.default_rng(9)
    # queries within 1e-4 .. 1e-6 of the surface
    tri = F[rng.integers(0, F.shape[0], 2048)].astype(np.int64)
    b = rng.dirichlet([1, 1, 1], 2048)
    base = np.einsum("ij,ijk->ik", b, V.astype(np.float64)[tri])
    nrm = base / np.linalg.norm(base, axis=1, keepdims=True)
    near = (base + nrm * (10.0 ** rng.uniform(-6, -4, 2048))[:, None] * rng.choice([-1, 1], 2048)[:, None])
    Q = np.concatenate([Q, near.astype(np.float32)])
    p32, pt32, d32, b32 = O.cpq_brute_f32(V, F, Q)
    p64, pt64, d64, b64 = O.cpq_brute_f64(V, F, Q)
    Fd, V64 = F.astype(np.int64), V.astype(np.float64)
    for i in range(0, len(Q), 7):  # the refinement against the FP64 formulation, same triangle
        t = Fd[p32[i]]
        d2, cp, _ = O.cp_tri_f64(Q[i].astype(np.float64), V64[t[0]], V64[t[1]], V64[t[2]])
        ref = np.float32(np.sqrt(d2))
        assert abs(int(d32[i].view(np.int32)) - int(ref.view(np.int32))) <= 1, (i, d32[i], ref)
    rel = np.abs(d32.astype(np.float64) - d64) / np.maximum(d64, 1e-30)
    assert rel.max() <= 1e-7, rel.max()
    same = p32 == p64
    assert np.max(np.abs(pt32[same] - pt64[same])) <= 1e-7 * _scale(V)
    assert d64.min() < 1e-5  # the sample does reach a hair off the surface


def test_lbvh_matches_brute_force_cpq():
    w = W.cpq_small(n_queries=8192)
    V, F, Q = w["verts"], w["tris"], w["queries"]
    bvh = O.CpuBvh(V, F, 4)
    info = bvh.info()
    assert info["bad"] == 0 and info["max_leaf"] <= 4
    got = bvh.closest_point(Q)
    exp = O.cpq_brute_f32(V, F, Q)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_lbvh_matches_brute_force_rays():
    w = W.rays_scene(n_rays=4096, grid=61, n_spheres=4)
    V, F = w["verts"], w["tris"]
    bvh = O.CpuBvh(V, F, 4)
    for g, e in zip(bvh.trace_closest(w["rays"]), O.ray_closest_brute_f32(V, F, w["rays"])):
        assert np.array_equal(g, e)
    assert np.array_equal(bvh.trace_any(w["rays_any"]), O.ray_any_brute_f32(V, F, w["rays_any"]))


def test_inputs_that_cannot_be_answered():
    """Known answers of the bad-input rule: a ray straight down onto the unit right
    triangle hits it; the same ray with a NaN or infinite origin / direction component,
    a zero direction or a NaN bound is a miss in every ray routine (brute force, LBVH,
    instanced), while an infinite tmax is an ordinary ray.  A query point with a
    non-finite coordinate finds no triangle (-1, +inf, zeros), in brute force and LBVH."""
    V = np.array(TRI, np.float32)
    F = np.array([[0, 1, 2]], np.uint32)
    good = np.array([0.25, 0.25, 1.0, 0.0, 0.0, 0.0, -1.0, 10.0], np.float32)
    rays = [good]
    for k, v in ((0, np.nan), (1, np.inf), (2, -np.inf), (4, np.nan), (6, -np.inf), (3, np.nan), (7, np.nan)):
        r = good.copy()
        r[k] = v
        rays.append(r)
    z = good.copy()
    z[4:7] = 0.0
    rays.append(z)
    f = good.copy()
    f[7] = np.inf
    rays.append(f)
    R = np.array(rays, np.float32)
    miss = np.array([False] + [True] * 8 + [False])
    prim, t, uv = O.ray_closest_brute_f32(V, F, R)
    assert (prim[miss] == -1).all() and (t[miss] == np.inf).all() and (uv[miss] == 0).all()
    assert (prim[~miss] == 0).all() and (t[~miss] == 1.0).all()
    assert np.array_equal(O.ray_any_brute_f32(V, F, R), (~miss).astype(np.uint8))
    bvh = O.CpuBvh(V, F, 4)
    for g, e in zip(bvh.trace_closest(R), (prim, t, uv)):
        assert np.array_equal(g, e)
    assert np.array_equal(bvh.trace_any(R), (~miss).astype(np.uint8))
    obj, ip, it, iuv = O.inst_ray_brute_f32(np.eye(3, 4), [0], [(V, F)], R)
    assert np.array_equal(ip, prim) and np.array_equal(it, t) and np.array_equal(obj[miss], prim[miss])
    assert np.array_equal(O.inst_ray_brute_f32(np.eye(3, 4), [0], [(V, F)], R, any_hit=True),
                          (~miss).astype(np.uint8))
    Q = np.array([[0.25, 0.25, 1.0], [np.nan, 0, 0], [0, np.inf, 0], [0, 0, -np.inf]], np.float32)
    for p, pt, d, b in (O.cpq_brute_f32(V, F, Q), bvh.closest_point(Q)):
        assert p.tolist() == [0, -1, -1, -1] and d[0] == 1.0 and (d[1:] == np.inf).all()
        assert (pt[1:] == 0).all() and (b[1:] == 0).all()


def test_ray_brute_f32_vs_f64():
    w = W.rays_scene(n_rays=2048, grid=41, n_spheres=3)
    V, F, R = w["verts"], w["tris"], w["rays"]
    p32, t32, uv32 = O.ray_closest_brute_f32(V, F, R)
    p64, t64, uv64 = O.ray_closest_brute_f64(V, F, R)
    hit = p64 >= 0
    assert (p32 != p64).mean() < 0.002
    same = hit & (p32 == p64)
    rel = np.abs(t32[same] - t64[same]) / t64[same]
    assert rel.max() <= 1e-5
    np.testing.assert_allclose(uv32[same], uv64[same], atol=1e-4)


def test_morton_known_answers_and_lbvh_order():
    lo = np.zeros(3, np.float32)
    hi = np.ones(3, np.float32)
    c = np.array([[0, 0, 0], [1, 1, 1], [0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5],
                  [1 / 1024, 0, 0]], np.float32)
    codes = O.morton_codes(c, lo, hi)
    # x occupies bit 3k+2, y bit 3k+1, z bit 3k (Karras 2012 interleave)
    assert codes.tolist() == [0, 0x3FFFFFFF, 1 << 29, 1 << 28, 1 << 27, 1 << 2]
    w = W.cpq_small()
    V, F = w["verts"], w["tris"]
    bvh = O.CpuBvh(V, F, 4)
    order = bvh.order()
    assert np.array_equal(np.sort(order), np.arange(F.shape[0]))  # a permutation
    tri = V[F.astype(np.int64)]
    cen = ((tri.min(1) + tri.max(1)) * np.float32(0.5)).astype(np.float32)
    k = O.morton_codes(cen, cen.min(0), cen.max(0))[order]
    assert np.all((k[1:] > k[:-1]) | ((k[1:] == k[:-1]) & (order[1:] > order[:-1])))


def test_icosphere_closed_forms():
    v, f = W.icosphere(3)
    assert f.shape == (1280, 3) and v.shape == (642, 3)
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-12)
    V, F = v.astype(np.float32), f.astype(np.uint32)
    # a query on a vertex has distance 0 and closest point = the vertex
    prim, pt, dist, _ = O.cpq_brute_f32(V, F, V[:50])
    assert np.all(dist == 0) and np.array_equal(pt, V[:50])
    # queries outside the sphere: r - 1 <= d <= distance to the nearest vertex
    rng = np.random.default_rng(4)
    dirs = W.random_dirs(rng, 500)
    r = rng.uniform(1.1, 2.0, 500)
    Q = (dirs * r[:, None]).astype(np.float32)
    _, _, d, _ = O.cpq_brute_f64(V, F, Q)
    assert np.all(d >= r - 1 - 1e-6)
    nearest_v = np.min(np.linalg.norm(Q[:, None, :].astype(np.float64) - V[None], axis=2), axis=1)
    assert np.all(d <= nearest_v + 1e-9)
    # a ray from the centre hits at t in [inradius, 1]
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    nrm = np.cross(b - a, c - a)
    inr = np.min(np.abs(np.sum(nrm * a, 1)) / np.linalg.norm(nrm, axis=1))
    R = W.pack_rays(np.zeros((500, 3)), dirs, 0.0, 10.0)
    p, t, _ = O.ray_closest_brute_f32(V, F, R)
    assert np.all(p >= 0) and np.all(t >= inr - 1e-6) and np.all(t <= 1 + 1e-6)


def test_cpq_gradients_match_fp64_central_differences():
    w = W.cpq_small(n_queries=256)
    V = w["verts"].astype(np.float64)
    F = w["tris"]
    Q = w["queries"].astype(np.float64)
    prim, pt, dist, bary = O.cpq_brute_f64(V, F, Q)
    gq, gv = O.cpq_grad_f64(V, F, Q, prim, pt, bary)
    h = 1e-4
    # w.r.t. queries; central differences are only meaningful where d is smooth on the
    # stencil, so points whose analytic gradient changes within +-h (medial axis) skip
    for k in range(3):
        e = np.zeros(3); e[k] = h
        rp = O.cpq_brute_f64(V, F, Q + e)
        rm = O.cpq_brute_f64(V, F, Q - e)
        gp = O.cpq_grad_f64(V, F, Q + e, rp[0], rp[1], rp[3])[0]
        gm = O.cpq_grad_f64(V, F, Q - e, rm[0], rm[1], rm[3])[0]
        smooth = np.max(np.abs(gp - gm), axis=1) < 1e-2
        assert smooth.mean() > 0.95
        fd = (rp[2] - rm[2]) / (2 * h)
        np.testing.assert_allclose(gq[smooth, k], fd[smooth], atol=2e-5)
    # w.r.t. vertices: loss = sum of distances; check the vertices with gradient mass
    rng = np.random.default_rng(0)
    touched = np.unique(F[prim].ravel())
    for j in rng.choice(touched, 12, replace=False):
        for k in range(3):
            Vp = V.copy(); Vp[j, k] += h
            Vm = V.copy(); Vm[j, k] -= h
            fd = (O.cpq_brute_f64(Vp, F, Q)[2].sum() - O.cpq_brute_f64(Vm, F, Q)[2].sum()) / (2 * h)
            assert gv[j, k] == pytest.approx(fd, abs=5e-5)


def test_ray_gradients_match_fp64_central_differences():
    v, f = W.icosphere(3)
    V, F = v, f.astype(np.uint32)
    rng = np.random.default_rng(2)
    o = W.random_dirs(rng, 128) * 3
    d = -o / np.linalg.norm(o, axis=1, keepdims=True) + rng.normal(0, 0.1, (128, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    R = np.concatenate([o, np.zeros((128, 1)), d, np.full((128, 1), 10.0)], 1)
    prim, t, uv = O.ray_closest_brute_f64(V, F, R)
    assert np.all(prim >= 0)
    go, gd, gv = O.ray_grad_f64(V, F, R, prim, t, uv)
    h = 1e-5
    for k in range(3):
        Rp = R.copy(); Rp[:, k] += h
        Rm = R.copy(); Rm[:, k] -= h
        fd = (O.ray_closest_brute_f64(V, F, Rp)[1] - O.ray_closest_brute_f64(V, F, Rm)[1]) / (2 * h)
        ok = np.abs(fd) < 100  # skip rays whose hit switches triangle inside +-h
        np.testing.assert_allclose(go[ok, k], fd[ok], atol=1e-4)
        Rp = R.copy(); Rp[:, 4 + k] += h
        Rm = R.copy(); Rm[:, 4 + k] -= h
        fd = (O.ray_closest_brute_f64(V, F, Rp)[1] - O.ray_closest_brute_f64(V, F, Rm)[1]) / (2 * h)
        np.testing.assert_allclose(gd[:, k], fd, atol=1e-3)
    for j in np.unique(F[prim].ravel())[:10]:
        for k in range(3):
            Vp = V.copy(); Vp[j, k] += h
            Vm = V.copy(); Vm[j, k] -= h
            fd = (O.ray_closest_brute_f64(Vp, F, R)[1].sum() -
                  O.ray_closest_brute_f64(Vm, F, R)[1].sum()) / (2 * h)
            assert gv[j, k] == pytest.approx(fd, abs=1e-3)


# ---- two-level instancing -------------------------------------------------------------

def _xf(M=None, c=(0.0, 0.0, 0.0)):
    M = np.eye(3) if M is None else np.asarray(M, np.float64)
    X = np.zeros((3, 4), np.float32)
    X[:, :3] = M
    X[:, 3] = c
    return X.reshape(12)


def test_invert3_matches_numpy_and_rejects_singular():
    rng = np.random.default_rng(21)
    for _ in range(50):
        M = rng.normal(size=(3, 3)) + 3 * np.eye(3)
        inv = O.invert3(_xf(M))
        np.testing.assert_allclose(inv, np.linalg.inv(M.astype(np.float32).astype(np.float64)), rtol=2e-6,
                                   atol=2e-7)
    with pytest.raises(ValueError):
        O.invert3(_xf(np.array([[1, 2, 3], [2, 4, 6], [0, 0, 1.0]])))


def _dyadic_scene(rng, n_tris=40, n_rays=600):
    # coordinates on a 1/16 grid in [-4, 4]: every transformed vertex and ray origin below
    # is exactly representable, so only the MT arithmetic rounds
    V = rng.integers(-64, 65, size=(3 * n_tris, 3)).astype(np.float32) / 16
    F = np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3)
    o = rng.integers(-128, 129, size=(n_rays, 3)).astype(np.float64) / 16
    d = rng.integers(-16, 17, size=(n_rays, 3)).astype(np.float64) / 16
    d[np.all(d == 0, axis=1)] = [0, 0, 1]
    R = W.pack_rays(o, d, 0.0, 1e7)
    return V, F, R


def test_instance_identity_equals_plain_brute_force():
    V, F, R = _dyadic_scene(np.random.default_rng(22))
    obj, prim, t, uv = O.inst_ray_brute_f32([_xf()], [0], [(V, F)], R, custom=[7])
    ep, et, euv = O.ray_closest_brute_f32(V, F, R)
    assert np.array_equal(prim, ep) and np.array_equal(t, et) and np.array_equal(uv, euv)
    assert np.array_equal(obj, np.where(ep >= 0, 7, -1))
    assert np.array_equal(O.inst_ray_brute_f32([_xf()], [0], [(V, F)], R, any_hit=True),
                          O.ray_any_brute_f32(V, F, R))


@pytest.mark.parametrize("s", [0.5, 1.0, 2.0, 4.0])
def test_scaled_translated_instance_equals_transformed_mesh(s):
    """A power-of-two scale and a dyadic translation transform the scene exactly, and IEEE
    rounding commutes with power-of-two scaling: the instance answers exactly like the
    plain brute force on the mesh moved into world space."""
    rng = np.random.default_rng(23)
    V, F, R = _dyadic_scene(rng)
    c = np.array([0.75, -1.5, 2.25], np.float32)
    Vw = (V * np.float32(s) + c).astype(np.float32)
    # aim two thirds of the rays at dyadic interior points (2a + b + c) / 4 of triangles
    k = rng.integers(0, F.shape[0], 400)
    tgt = (2 * Vw[F[k, 0]] + Vw[F[k, 1]] + Vw[F[k, 2]]) / 4
    R[:400, 4:7] = tgt - R[:400, :3]
    R = R[np.any(R[:, 4:7] != 0, axis=1)]
    obj, prim, t, uv = O.inst_ray_brute_f32([_xf(np.eye(3) * s, c)], [0], [(V, F)], R)
    ep, et, euv = O.ray_closest_brute_f32(Vw, F, R)
    assert (ep >= 0).mean() > 0.2
    assert np.array_equal(prim, ep) and np.array_equal(t, et) and np.array_equal(uv, euv)


def test_instances_masks_ties_and_custom_indices():
    V, F, R = _dyadic_scene(np.random.default_rng(24))
    # three copies of the same placement: exact ties go to the lowest visible instance
    xfs = [_xf()] * 3
    obj, prim, t, _ = O.inst_ray_brute_f32(xfs, [0, 0, 0], [(V, F)], R, custom=[10, 11, 12], mask=[0, 0xFF, 0xFF])
    ep, et, _ = O.ray_closest_brute_f32(V, F, R)
    assert np.array_equal(obj, np.where(ep >= 0, 11, -1)) and np.array_equal(prim, ep)
    # a mask with no bit of the ray's 0xFF cull mask hides an instance
    occ = O.inst_ray_brute_f32(xfs, [0, 0, 0], [(V, F)], R, mask=[0x100, 0, 0x200], any_hit=True)
    assert not occ.any()


def test_instanced_rays_agree_with_fp64_world_space():
    """Rotated/scaled instances of two meshes against the FP64 brute force over the
    meshes transformed to world space in FP64: ids agree except at near-ties, t to the
    rounding of the transform."""
    rng = np.random.default_rng(25)
    v1, f1 = W.icosphere(2)
    v2, f2 = W.terrain(9, rng)
    v2 = v2 / 10.0
    meshes = [(v1.astype(np.float32), f1), (v2.astype(np.float32), f2)]
    xfs, mesh_of, Vw, Fw, base = [], [], [], [], 0
    for k in range(6):
        a = rng.normal(size=3)
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        M = Q * rng.uniform(0.5, 2.0)
        c = rng.uniform(-3, 3, 3)
        X = _xf(M, c)
        xfs.append(X)
        m = k % 2
        mesh_of.append(m)
        X64 = X.astype(np.float64).reshape(3, 4)
        Vw.append(meshes[m][0].astype(np.float64) @ X64[:, :3].T + X64[:, 3])
        Fw.append(meshes[m][1].astype(np.uint32) + base)
        base += meshes[m][0].shape[0]
    o = rng.uniform(-5, 5, (2000, 3))
    R = W.pack_rays(o, W.random_dirs(rng, 2000), 1e-5, 1e7)
    obj, prim, t, _ = O.inst_ray_brute_f32(xfs, mesh_of, meshes, R)
    ep, et, _ = O.ray_closest_brute_f64(np.concatenate(Vw), np.concatenate(Fw), R.astype(np.float64))
    # world triangle index -> (instance, local prim)
    counts = [meshes[m][1].shape[0] for m in mesh_of]
    starts = np.concatenate([[0], np.cumsum(counts)])
    inst64 = np.searchsorted(starts, ep, side="right") - 1
    local64 = ep - starts[np.clip(inst64, 0, len(counts) - 1)]
    hit = ep >= 0
    same = (obj == np.where(hit, inst64, -1)) & (prim == np.where(hit, local64, -1))
    assert same.mean() > 0.995
    both = same & hit
    # the object-space ray carries the FP32 rounding of (o - c) and M^-1: an absolute
    # position error of a few 2^-24 of the scene scale (~5), plus FP32 MT rounding
    np.testing.assert_allclose(t[both], et[both], rtol=1e-5, atol=4e-6)
