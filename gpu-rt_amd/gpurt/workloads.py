"""Synthetic workloads for the geometric-query path (numpy, deterministic by seed).

All generators use numpy.random.default_rng(seed) (PCG64).  Choices the workload
descriptions leave open are pinned here and recorded in each workload's `meta`:

  icosphere(L)     20 * 4^L triangles: icosahedron subdivided L times by edge midpoints
                   projected to the unit sphere; vertices in creation order (12 base, then
                   each level's midpoints in ascending (min,max) vertex-pair order), faces of
                   a level in four blocks (corner a, corner b, corner c, centre).
  value noise      lattice values U[-1,1] at a 256-entry permutation hash, quintic fade,
                   fbm with lacunarity 2 and gain 0.5, normalised by the sum of amplitudes.
  terrain          G x G vertex grid over x,z in [-50, 50]; height 6 * fbm2(x/20, z/20) with
                   `octaves` octaves; cell (i,j) -> triangles (v00, v10, v01), (v01, v10, v11).
  rays             origin U(box), direction = normalised N(0, I_3).

Workloads (sizes):
  cpq_small       icosphere L4 (5,120 tris) jittered radially by N(0,0.01) (seed 0);
                  65,536 queries U[-1.5,1.5]^3 (seed 1)
  rays_scene      terrain 708x708 (999,698 tris, 6 octaves, seed 2) + 64 icospheres L5
                  (20,480 tris each) radius U[0.5,3], centre x,z U[-45,45], y U[hmin, hmax+8]
                  (seed 2); 1,048,576 rays, origins U(scene box), tmin 1e-5, tmax 1e7
                  (closest hit) and tmax U[1,50] (any hit) (seed 3)
  cpq_large       icosphere L9 (5,242,880 tris), radius 1 + 0.05 fbm3(2 n) 4 octaves (seed 4);
                  1,048,576 queries: 80% surface samples (area-weighted triangle, uniform
                  barycentric) + N(0,0.005) along the face normal, 20% U[-2,2]^3, shuffled
                  (seed 5)
  diff            icosphere L8 (1,310,720 tris) displaced as cpq_large (seed 8); 1,048,576
                  near-surface queries (seed 9); 262,144 rays from U(sphere r=3) towards
                  U(ball r=0.5) (seed 10)
  batched         256 meshes, triangle targets log-uniform [1024, 524288] (seed 6): even
                  index -> displaced icosphere at the level nearest the target (L3..L7),
                  odd index -> terrain grid with 2(G-1)^2 ~ target scaled to x,z in [-1,1];
                  query and ray counts log-uniform [256, 262144] (seed 7); queries 80%
                  near-surface + 20% uniform in the mesh's bounding box grown 20% about
                  its centre; rays from the mesh's bounding sphere inwards
"""
import numpy as np

_ICO_CACHE = {}


def icosphere(level):
    """Unit icosphere: (verts float64 (V,3), faces int64 (F,3))."""
    if level in _ICO_CACHE:
        v, f = _ICO_CACHE[level]
        return v.copy(), f.copy()
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
                  (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)],
                 dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9),
                  (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2),
                  (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
                  (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    for _ in range(level):
        nv, nf = v.shape[0], f.shape[0]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        e = np.stack([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)], 0)
        es = np.sort(e, axis=2).reshape(-1, 2)
        key = es[:, 0] * nv + es[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        mid = nv + inv.reshape(3, nf)
        m = v[uniq // nv] + v[uniq % nv]
        m /= np.linalg.norm(m, axis=1, keepdims=True)
        v = np.concatenate([v, m], 0)
        ab, bc, ca = mid[0], mid[1], mid[2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)], 0)
    _ICO_CACHE[level] = (v.copy(), f.copy())
    return v, f


class ValueNoise:
    """Lattice value noise with a seeded 256-entry permutation (2D and 3D)."""

    def __init__(self, rng):
        self.perm = np.concatenate([rng.permutation(256)] * 2).astype(np.int64)
        self.vals = rng.uniform(-1.0, 1.0, 256)

    @staticmethod
    def _fade(t):
        return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)

    def _h3(self, x, y, z):
        p = self.perm
        return self.vals[p[p[p[x & 255] + (y & 255)] + (z & 255)]]

    def noise3(self, p):
        i = np.floor(p).astype(np.int64)
        f = p - i
        u = self._fade(f)
        x, y, z = i[:, 0], i[:, 1], i[:, 2]
        ux, uy, uz = u[:, 0], u[:, 1], u[:, 2]
        def lerp(a, b, t):
            return a + (b - a) * t
        c000 = self._h3(x, y, z); c100 = self._h3(x + 1, y, z)
        c010 = self._h3(x, y + 1, z); c110 = self._h3(x + 1, y + 1, z)
        c001 = self._h3(x, y, z + 1); c101 = self._h3(x + 1, y, z + 1)
        c011 = self._h3(x, y + 1, z + 1); c111 = self._h3(x + 1, y + 1, z + 1)
        x00 = lerp(c000, c100, ux); x10 = lerp(c010, c110, ux)
        x01 = lerp(c001, c101, ux); x11 = lerp(c011, c111, ux)
        return lerp(lerp(x00, x10, uy), lerp(x01, x11, uy), uz)

    def fbm3(self, p, octaves):
        s = np.zeros(p.shape[0]); amp = 1.0; freq = 1.0; norm = 0.0
        for _ in range(octaves):
            s += amp * self.noise3(p * freq)
            norm += amp; amp *= 0.5; freq *= 2.0
        return s / norm

    def fbm2(self, x, z, octaves):
        p = np.stack([x, np.zeros_like(x), z], 1)
        return self.fbm3(p, octaves)


def _as_f32(v, f):
    return np.ascontiguousarray(v, dtype=np.float32), np.ascontiguousarray(f, dtype=np.int32)


def displaced_icosphere(level, rng, amp=0.05, octaves=4, freq=2.0):
    v, f = icosphere(level)
    nz = ValueNoise(rng)
    r = 1.0 + amp * nz.fbm3(v * freq, octaves)
    return v * r[:, None], f


def terrain(G, rng, octaves=6, extent=50.0, height=6.0, scale=20.0):
    nz = ValueNoise(rng)
    xs = np.linspace(-extent, extent, G)
    X, Z = np.meshgrid(xs, xs)  # X varies along columns (j), Z along rows (i)
    Y = height * nz.fbm2((X / scale).ravel(), (Z / scale).ravel(), octaves)
    v = np.stack([X.ravel(), Y, Z.ravel()], 1)
    i, j = np.meshgrid(np.arange(G - 1), np.arange(G - 1), indexing="ij")
    v00 = (i * G + j).ravel()
    v01 = v00 + 1
    v10 = v00 + G
    v11 = v10 + 1
    t1 = np.stack([v00, v10, v01], 1)
    t2 = np.stack([v01, v10, v11], 1)
    f = np.stack([t1, t2], 1).reshape(-1, 3)
    return v, f


def random_dirs(rng, n):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d


def surface_samples(v, f, rng, n, sigma):
    """Area-weighted surface points offset by N(0, sigma) along the face normal."""
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    cr = np.cross(b - a, c - a)
    area = np.linalg.norm(cr, axis=1)
    cdf = np.cumsum(area)
    cdf /= cdf[-1]
    tri = np.searchsorted(cdf, rng.uniform(size=n), side="right")
    tri = np.minimum(tri, f.shape[0] - 1)
    r1 = np.sqrt(rng.uniform(size=n))
    r2 = rng.uniform(size=n)
    w0, w1, w2 = 1 - r1, r1 * (1 - r2), r1 * r2
    p = w0[:, None] * a[tri] + w1[:, None] * b[tri] + w2[:, None] * c[tri]
    nrm = cr[tri] / np.maximum(area[tri], 1e-300)[:, None]
    return p + nrm * rng.normal(0.0, sigma, n)[:, None]


def pack_rays(o, d, tmin, tmax):
    n = o.shape[0]
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


# --------------------------------------------------------------------------------------
# the five workloads
# --------------------------------------------------------------------------------------

def cpq_small(n_queries=65536):
    rng = np.random.default_rng(0)
    v, f = icosphere(4)
    v = v * (1.0 + rng.normal(0.0, 0.01, v.shape[0]))[:, None]
    q = np.random.default_rng(1).uniform(-1.5, 1.5, (n_queries, 3))
    V, F = _as_f32(v, f)
    return {"verts": V, "tris": F, "queries": q.astype(np.float32),
            "meta": {"workload": "cpq_small", "n_tris": int(F.shape[0]),
                     "n_queries": int(n_queries), "leaf_size": 4}}


def rays_scene(n_rays=1 << 20, grid=708, n_spheres=64, sphere_level=5):
    rng = np.random.default_rng(2)
    tv, tf = terrain(grid, rng)
    hmin, hmax = tv[:, 1].min(), tv[:, 1].max()
    sv, sf = icosphere(sphere_level)
    radii = rng.uniform(0.5, 3.0, n_spheres)
    cx = rng.uniform(-45.0, 45.0, n_spheres)
    cz = rng.uniform(-45.0, 45.0, n_spheres)
    cy = rng.uniform(hmin, hmax + 8.0, n_spheres)
    vs, fs = [tv], [tf]
    off = tv.shape[0]
    for k in range(n_spheres):
        vs.append(sv * radii[k] + np.array([cx[k], cy[k], cz[k]]))
        fs.append(sf + off)
        off += sv.shape[0]
    V, F = _as_f32(np.concatenate(vs), np.concatenate(fs))
    lo, hi = V.min(0).astype(np.float64), V.max(0).astype(np.float64)
    r = np.random.default_rng(3)
    o = r.uniform(lo, hi, (n_rays, 3))
    d = random_dirs(r, n_rays)
    tmax_any = r.uniform(1.0, 50.0, n_rays)
    rays = pack_rays(o, d, 1e-5, 1e7)
    rays_any = pack_rays(o, d, 1e-5, 0.0)
    rays_any[:, 7] = tmax_any
    return {"verts": V, "tris": F, "rays": rays, "rays_any": rays_any,
            "meta": {"workload": "rays_scene", "n_tris": int(F.shape[0]),
                     "n_terrain_tris": int(tf.shape[0]), "n_rays": int(n_rays),
                     "grid": grid, "n_spheres": n_spheres}}


def random_rotation(rng):
    """Uniform random rotation (QR of a Gaussian matrix with the sign fixed)."""
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def instanced_scene(n_rays=1 << 20, grid=708, n_spheres=64, sphere_level=5, seed=2):
    """rays_scene as the reference lays it out: one terrain mesh and ONE icosphere mesh,
    placed by 1 + n_spheres instances (object-to-world 3x4 transforms: identity for the
    terrain; rotation * radius + position for each sphere, the way GPURT::build_accel
    places objects by scale * pose), with the same incoherent rays (origins uniform in the
    scene box, uniform directions; any-hit tmax U[1, 50])."""
    rng = np.random.default_rng(seed)
    tv, tf = terrain(grid, rng)
    hmin, hmax = tv[:, 1].min(), tv[:, 1].max()
    sv, sf = icosphere(sphere_level)
    radii = rng.uniform(0.5, 3.0, n_spheres)
    cx = rng.uniform(-45.0, 45.0, n_spheres)
    cz = rng.uniform(-45.0, 45.0, n_spheres)
    cy = rng.uniform(hmin, hmax + 8.0, n_spheres)
    xfs = [np.concatenate([np.eye(3), np.zeros((3, 1))], 1)]
    for k in range(n_spheres):
        M = random_rotation(rng) * radii[k]
        xfs.append(np.concatenate([M, np.array([[cx[k]], [cy[k]], [cz[k]]])], 1))
    xfs = np.asarray(xfs, np.float32)
    lo = np.minimum(tv.min(0), (np.array([cx, cy, cz]).T - radii[:, None]).min(0))
    hi = np.maximum(tv.max(0), (np.array([cx, cy, cz]).T + radii[:, None]).max(0))
    r = np.random.default_rng(seed + 1)
    o = r.uniform(lo, hi, (n_rays, 3))
    d = random_dirs(r, n_rays)
    rays = pack_rays(o, d, 1e-5, 1e7)
    rays_any = pack_rays(o, d, 1e-5, 0.0)
    rays_any[:, 7] = r.uniform(1.0, 50.0, n_rays)
    meshes = [_as_f32(tv, tf), _as_f32(sv, sf)]
    mesh_of = np.array([0] + [1] * n_spheres, np.int32)
    n_tris = int(tf.shape[0] + n_spheres * sf.shape[0])
    return {"meshes": meshes, "xf": xfs.reshape(-1, 12), "mesh_of": mesh_of, "rays": rays,
            "rays_any": rays_any,
            "meta": {"workload": "instanced_scene", "n_instances": int(n_spheres + 1),
                     "n_blas_tris": int(tf.shape[0] + sf.shape[0]), "n_tris_instanced": n_tris,
                     "n_rays": int(n_rays), "grid": grid, "n_spheres": n_spheres}}


def near_surface_queries(v, f, rng, n, frac_surface=0.8, sigma=0.005, box=2.0):
    """80% surface samples + N(0, sigma) normal offsets, 20% uniform in a box: the cube
    [-box, box]^3 for a number, or an explicit (lo, hi) pair."""
    n_s = int(round(n * frac_surface))
    qs = surface_samples(v, f, rng, n_s, sigma)
    if isinstance(box, tuple):
        qu = rng.uniform(box[0], box[1], (n - n_s, 3))
    else:
        qu = rng.uniform(-box, box, (n - n_s, 3))
    q = np.concatenate([qs, qu], 0)
    return q[rng.permutation(n)]


def cpq_large(n_queries=1 << 20, level=9):
    v, f = displaced_icosphere(level, np.random.default_rng(4))
    V, F = _as_f32(v, f)
    q = near_surface_queries(V.astype(np.float64), F, np.random.default_rng(5), n_queries)
    return {"verts": V, "tris": F, "queries": q.astype(np.float32),
            "meta": {"workload": "cpq_large", "n_tris": int(F.shape[0]),
                     "n_queries": int(n_queries), "level": level}}


def diff(n_queries=1 << 20, n_rays=1 << 18, level=8):
    v, f = displaced_icosphere(level, np.random.default_rng(8))
    V, F = _as_f32(v, f)
    q = surface_samples(V.astype(np.float64), F, np.random.default_rng(9), n_queries, 0.005)
    r = np.random.default_rng(10)
    o = random_dirs(r, n_rays) * 3.0
    tgt = random_dirs(r, n_rays) * (0.5 * r.uniform(size=n_rays) ** (1.0 / 3.0))[:, None]
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return {"verts": V, "tris": F, "queries": q.astype(np.float32),
            "ray_o": o.astype(np.float32), "ray_d": d.astype(np.float32),
            "meta": {"workload": "diff", "n_tris": int(F.shape[0]), "n_queries": int(n_queries),
                     "n_rays": int(n_rays), "level": level}}


def batched(n_meshes=256, tri_lo=1024, tri_hi=524288, q_lo=256, q_hi=262144):
    rng = np.random.default_rng(6)
    targets = np.exp(rng.uniform(np.log(tri_lo), np.log(tri_hi), n_meshes))
    qr = np.random.default_rng(7)
    counts = np.exp(qr.uniform(np.log(q_lo), np.log(q_hi), n_meshes)).astype(np.int64)
    meshes = []
    for k in range(n_meshes):
        mr = np.random.default_rng([6, k])
        if k % 2 == 0:
            lvl = int(np.clip(np.round(np.log(targets[k] / 20.0) / np.log(4.0)), 3, 7))
            v, f = displaced_icosphere(lvl, mr)
        else:
            G = int(round((targets[k] / 2.0) ** 0.5)) + 1
            v, f = terrain(G, mr, octaves=5)
            v = v / 50.0  # unit-scale terrain
        V, F = _as_f32(v, f)
        n = int(counts[k])
        lo, hi = V.min(0).astype(np.float64), V.max(0).astype(np.float64)
        mid, half = (lo + hi) / 2, (hi - lo) / 2 * 1.2
        q = near_surface_queries(V.astype(np.float64), F, np.random.default_rng([7, k]), n,
                                 box=(mid - half, mid + half))
        rr = np.random.default_rng([7, k, 1])
        c = V.mean(0).astype(np.float64)
        rad = float(np.linalg.norm(V - c, axis=1).max()) * 1.5 + 1e-3
        o = c + random_dirs(rr, n) * rad
        tgt = c + random_dirs(rr, n) * (0.5 * rad) * rr.uniform(size=n)[:, None]
        d = tgt - o
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        meshes.append({"verts": V, "tris": F, "queries": q.astype(np.float32),
                       "rays": pack_rays(o, d, 1e-5, 1e7)})
    meta = {"workload": "batched", "n_meshes": n_meshes,
            "n_tris_total": int(sum(m["tris"].shape[0] for m in meshes)),
            "n_queries_total": int(sum(m["queries"].shape[0] for m in meshes))}
    return {"meshes": meshes, "meta": meta}


# ---- path-tracer scenes ----------------------------------------------------------------

def _mesh_record(pos, nrm, tris):
    """Mesh::Vertex records (n, 12): pos.xyz, u=0, normal.xyz, v=0, tangent 0."""
    v = np.zeros((pos.shape[0], 12), np.float32)
    v[:, 0:3] = pos
    v[:, 4:7] = nrm
    return v, np.ascontiguousarray(tris, np.uint32)


def _quad():
    """Unit quad in the xz plane (y = 0, x, z in [-1, 1]), normal +y, two triangles."""
    p = np.array([(-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)], np.float64)
    return _mesh_record(p, np.tile([0.0, 1.0, 0.0], (4, 1)), [(0, 2, 1), (0, 3, 2)])


def _cube():
    """Unit cube [-1, 1]^3 with flat per-face normals (24 vertices, 12 triangles)."""
    pos, nrm, tri = [], [], []
    for axis in range(3):
        for s in (-1.0, 1.0):
            n = np.zeros(3)
            n[axis] = s
            u, w = (axis + 1) % 3, (axis + 2) % 3
            base = len(pos)
            for a, b in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
                p = np.zeros(3)
                p[axis], p[u], p[w] = s, a, b
                pos.append(p)
                nrm.append(n)
            tri += [(base, base + 1, base + 2), (base, base + 2, base + 3)]
    return _mesh_record(np.array(pos), np.array(nrm), tri)


def _sphere(level):
    v, f = icosphere(level)
    return _mesh_record(v, v, f)


def _affine(scale=(1, 1, 1), rot_deg=(0, 0, 0), trans=(0, 0, 0)):
    """Row-major 4x4 T * Rz * Ry * Rx * S."""
    ax, ay, az = np.radians(rot_deg)
    rx = np.array([[1, 0, 0], [0, np.cos(ax), -np.sin(ax)], [0, np.sin(ax), np.cos(ax)]])
    ry = np.array([[np.cos(ay), 0, np.sin(ay)], [0, 1, 0], [-np.sin(ay), 0, np.cos(ay)]])
    rz = np.array([[np.cos(az), -np.sin(az), 0], [np.sin(az), np.cos(az), 0], [0, 0, 1]])
    m = np.eye(4)
    m[:3, :3] = rz @ ry @ rx @ np.diag(scale)
    m[:3, 3] = trans
    return m


def _obj(mesh, model, albedo, emissive=(0, 0, 0), metal_rough=(0.0, 1.0), name=""):
    v, f = mesh
    return {"verts": v, "indices": f, "model": model, "albedo": np.asarray(albedo, np.float32),
            "emissive": np.asarray(emissive, np.float32), "metal_rough": np.asarray(metal_rough, np.float32),
            "name": name}


def tracer_box(sphere_level=3):
    """A closed Cornell-style box for the path tracer: five walls (quads placed by
    rotations and non-uniform scales), a two-triangle light under the ceiling (emission
    10), a glossy sphere (roughness 0.2), a rotated and sheared-scale block (roughness 1),
    a mirror sphere (roughness 0) and metallic factors for use_metalness.  Camera at
    (0, 0, 14) looking at the origin, 45 degree vertical field of view."""
    q, c, s = _quad(), _cube(), _sphere(sphere_level)
    objs = [
        _obj(q, _affine((5, 1, 5), (0, 0, 0), (0, -5, 0)), (0.75, 0.75, 0.75), name="floor"),
        _obj(q, _affine((5, 1, 5), (180, 0, 0), (0, 5, 0)), (0.75, 0.75, 0.75), name="ceiling"),
        _obj(q, _affine((5, 1, 5), (90, 0, 0), (0, 0, -5)), (0.7, 0.7, 0.7), name="back"),
        _obj(q, _affine((5, 1, 5), (0, 0, -90), (-5, 0, 0)), (0.75, 0.1, 0.1), name="left"),
        _obj(q, _affine((5, 1, 5), (0, 0, 90), (5, 0, 0)), (0.1, 0.75, 0.1), name="right"),
        _obj(q, _affine((1.5, 1, 1.2), (180, 0, 0), (0, 4.95, 0)), (0.0, 0.0, 0.0), emissive=(10, 10, 10),
             metal_rough=(0.0, 0.0), name="light"),
        _obj(s, _affine((1.6, 1.6, 1.6), (0, 30, 0), (-2.2, -3.4, -1.0)), (0.9, 0.8, 0.6), metal_rough=(0.5, 0.2),
             name="glossy sphere"),
        _obj(c, _affine((1.2, 2.5, 1.2), (0, 25, 0), (2.3, -2.5, -1.5)), (0.6, 0.6, 0.7), metal_rough=(0.0, 1.0),
             name="block"),
        _obj(s, _affine((1.0, 1.0, 1.0), (0, 0, 0), (1.5, -4.0, 2.5)), (0.95, 0.95, 0.95), metal_rough=(1.0, 0.0),
             name="mirror sphere"),
    ]
    cam = {"eye": (0.0, 0.0, 14.0), "target": (0.0, 0.0, 0.0), "yfov_deg": 45.0}
    return {"objects": objs, "camera": cam}


def tracer_floor(occluder=True, roughness=0.5):
    """An open scene whose every light path has at most one non-black bounce: a floor
    (Blinn-Phong / GGX, `roughness`), an emitting quad above it and, optionally, a black
    (albedo 0) block that casts a shadow.  Direct light sampling, material sampling and
    MIS therefore estimate the same image (with env_light = 0)."""
    q, c = _quad(), _cube()
    objs = [
        _obj(q, _affine((6, 1, 6), (0, 0, 0), (0, 0, 0)), (0.8, 0.7, 0.6), metal_rough=(0.0, roughness),
             name="floor"),
        _obj(q, _affine((1.5, 1, 1.0), (180, 20, 0), (0.5, 4.0, 0.0)), (0.0, 0.0, 0.0), emissive=(8, 8, 8),
             metal_rough=(0.0, 0.0), name="light"),
    ]
    if occluder:
        objs.append(_obj(c, _affine((0.6, 0.8, 0.6), (0, 30, 0), (0.3, 1.6, 0.2)), (0.0, 0.0, 0.0),
                         metal_rough=(0.0, 1.0), name="black block"))
    cam = {"eye": (0.0, 7.0, 9.0), "target": (0.0, 0.0, 0.0), "yfov_deg": 50.0}
    return {"objects": objs, "camera": cam}
