"""numpy/ctypes wrapper of liboracle.so (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the test oracle or the timed CPU port -- never by the product
package.  See oracle.c for what each function restates and which reference lines it
follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    I64, I, P = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p
    opt = lambda t: ctypes.c_void_p  # nullable arrays are passed as raw pointers
    lib.orc_set_threads.argtypes = [I]
    lib.orc_num_threads.restype = I
    lib.orc_cp_tri_f32.argtypes = [_f32p] * 4 + [_f32p, _f32p]
    lib.orc_cp_tri_f32.restype = ctypes.c_float
    lib.orc_cp_tri_refine.argtypes = [_f32p] * 4 + [_f32p, _f32p]
    lib.orc_cp_tri_refine.restype = ctypes.c_float
    lib.orc_mt_f32.argtypes = [_f32p] * 5 + [ctypes.POINTER(ctypes.c_float)] * 3
    lib.orc_mt_f32.restype = I
    lib.orc_cp_tri_f64.argtypes = [_f64p] * 4 + [_f64p, _f64p]
    lib.orc_cp_tri_f64.restype = ctypes.c_double
    lib.orc_ray_tri_f64.argtypes = [_f64p] * 5 + [ctypes.POINTER(ctypes.c_double)] * 3
    lib.orc_ray_tri_f64.restype = I
    lib.orc_cpq_brute_f32.argtypes = [_f32p, _u32p, I64, _f32p, I64, _i32p, P, P, P]
    lib.orc_cpq_brute_f64.argtypes = [_f64p, _u32p, I64, _f64p, I64, _i32p, P, P, P]
    lib.orc_ray_closest_brute_f32.argtypes = [_f32p, _u32p, I64, _f32p, I64, _i32p, _f32p, P]
    lib.orc_ray_any_brute_f32.argtypes = [_f32p, _u32p, I64, _f32p, I64, _u8p]
    lib.orc_ray_closest_brute_f64.argtypes = [_f64p, _u32p, I64, _f64p, I64, _i32p, _f64p, P]
    lib.orc_cpq_grad_f64.argtypes = [_f64p, _u32p, _f64p, I64, _i32p, P, P, P, P, P]
    lib.orc_ray_grad_f64.argtypes = [_f64p, _u32p, _f64p, I64, _i32p, _f64p, _f64p, P, P, P, P]
    lib.orc_morton_codes.argtypes = [_f32p, I64, _f32p, _f32p, _u32p]
    lib.orc_invert3.argtypes = [_f32p, _f32p]
    lib.orc_invert3.restype = I
    lib.orc_inst_ray_brute_f32.argtypes = [I64, _f32p, _i32p, _u32p, _u32p, _f32p, _u32p, _i64p, _f32p, I64,
                                           P, P, P, P, P]
    lib.orc_inst_ray_brute_f32.restype = I
    lib.orc_lbvh_build.argtypes = [_f32p, I64, _u32p, I64, I]
    lib.orc_lbvh_build.restype = P
    lib.orc_lbvh_free.argtypes = [P]
    lib.orc_lbvh_info.argtypes = [P, _i64p]
    lib.orc_lbvh_order.argtypes = [P, _i32p]
    lib.orc_lbvh_cpq.argtypes = [P, _f32p, I64, _i32p, P, P, P, P]
    lib.orc_lbvh_ray_closest.argtypes = [P, _f32p, I64, _i32p, _f32p, P, P]
    lib.orc_lbvh_ray_any.argtypes = [P, _f32p, I64, _u8p, P]
    U32 = ctypes.c_uint32
    lib.orc_tea_rounds.argtypes = [U32, U32, _u32p, I]
    lib.orc_tea_rounds.restype = U32
    lib.orc_tea.argtypes = [U32, U32]
    lib.orc_tea.restype = U32
    lib.orc_randf_seq.argtypes = [U32, I64, _f32p]
    lib.orc_mat_samples.argtypes = [I, _f32p, ctypes.c_float, _f32p, _f32p, U32, I64, _f32p]
    lib.orc_mat_eval.argtypes = [I, _f32p, ctypes.c_float, _f32p, _f32p, _f32p, I64, _f32p]
    lib.orc_render.argtypes = [I64, _f32p, _i64p, _f32p, _f32p, _i64p, _u32p, _f32p, _f32p, _i32p, _f32p, P, P, P]
    lib.orc_render.restype = I
    lib.orc_light_pdf.argtypes = [I64, _f32p, _i64p, _f32p, _i64p, _u32p, _f32p, _f32p, _f32p, I64, _f32p]
    lib.orc_light_pdf.restype = I
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def set_threads(n):
    load().orc_set_threads(int(n))


def num_threads():
    return load().orc_num_threads()


# ---- per-primitive -------------------------------------------------------------------

def cp_tri_f32(p, a, b, c):
    cp = np.zeros(3, np.float32); br = np.zeros(3, np.float32)
    d2 = load().orc_cp_tri_f32(_c(p, np.float32), _c(a, np.float32), _c(b, np.float32),
                               _c(c, np.float32), cp, br)
    return np.float32(d2), cp, br


def cp_tri_f64(p, a, b, c):
    cp = np.zeros(3); br = np.zeros(3)
    d2 = load().orc_cp_tri_f64(_c(p, np.float64), _c(a, np.float64), _c(b, np.float64),
                               _c(c, np.float64), cp, br)
    return d2, cp, br


def mt_f32(o, d, a, b, c):
    t, u, v = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    ok = load().orc_mt_f32(_c(o, np.float32), _c(d, np.float32), _c(a, np.float32),
                           _c(b, np.float32), _c(c, np.float32), ctypes.byref(t),
                           ctypes.byref(u), ctypes.byref(v))
    return bool(ok), np.float32(t.value), np.float32(u.value), np.float32(v.value)


def ray_tri_f64(o, d, a, b, c):
    t, u, v = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    ok = load().orc_ray_tri_f64(_c(o, np.float64), _c(d, np.float64), _c(a, np.float64),
                                _c(b, np.float64), _c(c, np.float64), ctypes.byref(t),
                                ctypes.byref(u), ctypes.byref(v))
    return bool(ok), t.value, u.value, v.value


# ---- brute force -----------------------------------------------------------------------

def cpq_brute_f32(V, F, Q):
    V, F, Q = _c(V, np.float32), _c(F, np.uint32), _c(Q, np.float32)
    n = Q.shape[0]
    prim = np.empty(n, np.int32); pt = np.empty((n, 3), np.float32)
    dist = np.empty(n, np.float32); bary = np.empty((n, 3), np.float32)
    load().orc_cpq_brute_f32(V, F, F.shape[0], Q, n, prim, _p(pt), _p(dist), _p(bary))
    return prim, pt, dist, bary


def cpq_brute_f64(V, F, Q):
    V, F, Q = _c(V, np.float64), _c(F, np.uint32), _c(Q, np.float64)
    n = Q.shape[0]
    prim = np.empty(n, np.int32); pt = np.empty((n, 3)); dist = np.empty(n)
    bary = np.empty((n, 3))
    load().orc_cpq_brute_f64(V, F, F.shape[0], Q, n, prim, _p(pt), _p(dist), _p(bary))
    return prim, pt, dist, bary


def ray_closest_brute_f32(V, F, R):
    V, F, R = _c(V, np.float32), _c(F, np.uint32), _c(R, np.float32)
    n = R.shape[0]
    prim = np.empty(n, np.int32); t = np.empty(n, np.float32); uv = np.empty((n, 2), np.float32)
    load().orc_ray_closest_brute_f32(V, F, F.shape[0], R, n, prim, t, _p(uv))
    return prim, t, uv


def ray_any_brute_f32(V, F, R):
    V, F, R = _c(V, np.float32), _c(F, np.uint32), _c(R, np.float32)
    occ = np.empty(R.shape[0], np.uint8)
    load().orc_ray_any_brute_f32(V, F, F.shape[0], R, R.shape[0], occ)
    return occ


def invert3(xf):
    """FP32 inverse of the 3x3 part of a row-major 3x4 transform (FP64 cofactors)."""
    xf = _c(np.asarray(xf, np.float32).reshape(12), np.float32)
    out = np.empty(9, np.float32)
    if not load().orc_invert3(xf, out):
        raise ValueError("singular transform")
    return out.reshape(3, 3)


def _flatten_meshes(meshes):
    Vs, Fs, off, base = [], [], [0], 0
    for V, F in meshes:
        V = np.asarray(V, np.float32).reshape(-1, 3)
        Fs.append(np.asarray(F, np.uint32).reshape(-1, 3) + np.uint32(base))
        Vs.append(V)
        base += V.shape[0]
        off.append(off[-1] + Fs[-1].shape[0])
    return (_c(np.concatenate(Vs), np.float32), _c(np.concatenate(Fs), np.uint32),
            _c(np.asarray(off, np.int64), np.int64))


def inst_ray_brute_f32(xf, mesh_of, meshes, R, custom=None, mask=None, any_hit=False):
    """Two-level brute force: instance n places meshes[mesh_of[n]] = (V, F) by the
    row-major 3x4 transform xf[n].  Closest hit: (obj, prim, t, uv); any hit: occ."""
    xf = _c(np.asarray(xf, np.float32).reshape(-1, 12), np.float32)
    n_inst = xf.shape[0]
    mesh_of = _c(np.asarray(mesh_of, np.int32), np.int32)
    custom = _c(np.arange(n_inst, dtype=np.uint32) if custom is None else np.asarray(custom, np.uint32), np.uint32)
    mask = _c(np.full(n_inst, 0xFF, np.uint32) if mask is None else np.asarray(mask, np.uint32), np.uint32)
    V, F, off = _flatten_meshes(meshes)
    R = _c(R, np.float32)
    n = R.shape[0]
    if any_hit:
        occ = np.empty(n, np.uint8)
        ok = load().orc_inst_ray_brute_f32(n_inst, xf, mesh_of, custom, mask, V, F, off, R, n,
                                           None, None, None, None, _p(occ))
        if not ok:
            raise ValueError("singular transform")
        return occ
    obj = np.empty(n, np.int32); prim = np.empty(n, np.int32)
    t = np.empty(n, np.float32); uv = np.empty((n, 2), np.float32)
    ok = load().orc_inst_ray_brute_f32(n_inst, xf, mesh_of, custom, mask, V, F, off, R, n,
                                       _p(obj), _p(prim), _p(t), _p(uv), None)
    if not ok:
        raise ValueError("singular transform")
    return obj, prim, t, uv


def ray_closest_brute_f64(V, F, R):
    V, F, R = _c(V, np.float64), _c(F, np.uint32), _c(R, np.float64)
    n = R.shape[0]
    prim = np.empty(n, np.int32); t = np.empty(n); uv = np.empty((n, 2))
    load().orc_ray_closest_brute_f64(V, F, F.shape[0], R, n, prim, t, _p(uv))
    return prim, t, uv


# ---- gradients (FP64 closed forms) ---------------------------------------------------

def cpq_grad_f64(V, F, Q, prim, pt=None, bary=None, g=None):
    """FP64 closed-form distance gradients at the given argmin; pt / bary None: the
    closest point of each prim is computed in FP64 here (independent formulation)."""
    V, F, Q = _c(V, np.float64), _c(F, np.uint32), _c(Q, np.float64)
    prim = _c(prim, np.int32)
    pt = _c(pt, np.float64) if pt is not None else None
    bary = _c(bary, np.float64) if bary is not None else None
    g = _c(g, np.float64) if g is not None else None
    n = Q.shape[0]
    gq = np.zeros((n, 3)); gv = np.zeros((V.shape[0], 3))
    load().orc_cpq_grad_f64(V, F, Q, n, prim, _p(pt), _p(bary), _p(g), _p(gq), _p(gv))
    return gq, gv


def ray_grad_f64(V, F, R, prim, t, uv, g=None):
    V, F, R = _c(V, np.float64), _c(F, np.uint32), _c(R, np.float64)
    prim = _c(prim, np.int32); t = _c(t, np.float64); uv = _c(uv, np.float64)
    g = _c(g, np.float64) if g is not None else None
    n = R.shape[0]
    go = np.zeros((n, 3)); gd = np.zeros((n, 3)); gv = np.zeros((V.shape[0], 3))
    load().orc_ray_grad_f64(V, F, R, n, prim, t, uv, _p(g), _p(go), _p(gd), _p(gv))
    return go, gd, gv


# ---- CPU LBVH ----------------------------------------------------------------------------

def morton_codes(centroids, lo, hi):
    c = _c(centroids, np.float32)
    out = np.empty(c.shape[0], np.uint32)
    load().orc_morton_codes(c, c.shape[0], _c(lo, np.float32), _c(hi, np.float32), out)
    return out


class CpuBvh:
    """The oracle's CPU LBVH (Karras 2012, leaves <= leaf_size) -- the CPU baseline."""

    def __init__(self, V, F, leaf_size=4):
        self.V = _c(V, np.float32)
        self.F = _c(F, np.uint32)
        self.h = load().orc_lbvh_build(self.V, self.V.shape[0], self.F, self.F.shape[0],
                                       int(leaf_size))
        if not self.h:
            raise ValueError("oracle LBVH build failed (bad input)")

    def info(self):
        out = np.zeros(5, np.int64)
        load().orc_lbvh_info(self.h, out)
        return {"n_nodes": int(out[0]), "depth": int(out[1]), "max_leaf": int(out[2]),
                "n_leaves": int(out[3]), "bad": int(out[4])}

    def order(self):
        out = np.empty(self.F.shape[0], np.int32)
        load().orc_lbvh_order(self.h, out)
        return out

    def closest_point(self, Q, counters=None):
        Q = _c(Q, np.float32)
        n = Q.shape[0]
        prim = np.empty(n, np.int32); pt = np.empty((n, 3), np.float32)
        dist = np.empty(n, np.float32); bary = np.empty((n, 3), np.float32)
        load().orc_lbvh_cpq(self.h, Q, n, prim, _p(pt), _p(dist), _p(bary), _p(counters))
        return prim, pt, dist, bary

    def trace_closest(self, R, counters=None):
        R = _c(R, np.float32)
        n = R.shape[0]
        prim = np.empty(n, np.int32); t = np.empty(n, np.float32); uv = np.empty((n, 2), np.float32)
        load().orc_lbvh_ray_closest(self.h, R, n, prim, t, _p(uv), _p(counters))
        return prim, t, uv

    def trace_any(self, R, counters=None):
        R = _c(R, np.float32)
        occ = np.empty(R.shape[0], np.uint8)
        load().orc_lbvh_ray_any(self.h, R, R.shape[0], occ, _p(counters))
        return occ

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_lbvh_free(self.h)
            self.h = None


# ---- path tracer (pathtrace.c) ---------------------------------------------------------

def tea(v0, v1, key=None, rounds=16):
    """tea of rtcommon.glsl:99-111; key / rounds default to the reference's."""
    if key is None:
        return int(load().orc_tea(v0 & 0xFFFFFFFF, v1 & 0xFFFFFFFF))
    return int(load().orc_tea_rounds(v0 & 0xFFFFFFFF, v1 & 0xFFFFFFFF, _c(key, np.uint32), rounds))


def randf_seq(seed, n):
    out = np.empty(n, np.float32)
    load().orc_randf_seq(seed & 0xFFFFFFFF, n, out)
    return out


def mat_samples(brdf, albedo, roughness, wo, N, seed, n):
    """n draws of MAT_sample at the frame of N: (ok, wi (n,3), pdf(wi), eval(wi) (n,3))."""
    out = np.empty((n, 8), np.float32)
    load().orc_mat_samples(brdf, _c(albedo, np.float32), roughness, _c(wo, np.float32), _c(N, np.float32),
                           seed & 0xFFFFFFFF, n, out)
    return out[:, 0] > 0, out[:, 1:4], out[:, 4], out[:, 5:8]


def mat_eval(brdf, albedo, roughness, wo, N, wi):
    wi = _c(wi, np.float32).reshape(-1, 3)
    out = np.empty((wi.shape[0], 4), np.float32)
    load().orc_mat_eval(brdf, _c(albedo, np.float32), roughness, _c(wo, np.float32), _c(N, np.float32), wi,
                        wi.shape[0], out)
    return out[:, 0], out[:, 1:4]


def _tracer_scene(objects):
    """objects: dicts with verts (n, 12) Mesh::Vertex floats, indices (m, 3), model (4, 4)
    row-major object-to-world, albedo, emissive, metal_rough."""
    model = _c(np.stack([np.asarray(o["model"], np.float32).reshape(4, 4)[:3].reshape(12) for o in objects]),
               np.float32)
    vs = [np.asarray(o["verts"], np.float32).reshape(-1, 12) for o in objects]
    fs = [np.asarray(o["indices"], np.uint32).reshape(-1, 3) for o in objects]
    voff = _c(np.concatenate([[0], np.cumsum([v.shape[0] for v in vs])]), np.int64)
    foff = _c(np.concatenate([[0], np.cumsum([f.shape[0] for f in fs])]), np.int64)
    V = _c(np.concatenate([v[:, 0:3] for v in vs]), np.float32)
    N = _c(np.concatenate([v[:, 4:7] for v in vs]), np.float32)
    F = _c(np.concatenate(fs), np.uint32)
    mat = _c(np.stack([np.concatenate([np.asarray(o["albedo"], np.float32).reshape(3),
                                       np.asarray(o["emissive"], np.float32).reshape(3),
                                       np.asarray(o["metal_rough"], np.float32).reshape(2)]) for o in objects]),
             np.float32)
    return len(objects), model, voff, V, N, foff, F, mat


_RENDER_INT = ("width", "height", "frame", "samples", "max_frame", "qmc", "max_depth", "use_metalness",
               "integrator", "brdf", "use_rr", "seed")


def render(objects, params, image=None, gbuffer=False):
    """One frame of the reference's path tracer on the CPU (pathtrace.c).  params: dict with
    inv_view, inv_proj (4x4 row-major), clear_col, env_light and the integer fields of
    _RENDER_INT.  Returns image (h, w, 4) [, (pos, norm, albedo)]."""
    n, model, voff, V, N, foff, F, mat = _tracer_scene(objects)
    fp = _c(np.concatenate([np.asarray(params["inv_view"], np.float32).reshape(16),
                            np.asarray(params["inv_proj"], np.float32).reshape(16),
                            np.asarray(params["clear_col"], np.float32).reshape(3),
                            np.asarray(params["env_light"], np.float32).reshape(3)]), np.float32)
    ip = _c(np.array([int(params[k]) & 0xFFFFFFFF if k == "seed" else int(params[k]) for k in _RENDER_INT],
                     np.int64).astype(np.uint32).view(np.int32), np.int32)
    h, w = int(params["height"]), int(params["width"])
    img = np.zeros((h, w, 4), np.float32) if image is None else _c(image, np.float32).copy()
    g = [np.zeros((h, w, 4), np.float32) for _ in range(3)] if gbuffer else [None] * 3
    if not load().orc_render(n, model, voff, V, N, foff, F, mat, fp, ip, img, _p(g[0]), _p(g[1]), _p(g[2])):
        raise ValueError("singular model matrix")
    return (img, tuple(g)) if gbuffer else img


def light_pdf(objects, P, D):
    """light_pdf (rt.rgen:222-255) for (point, direction) pairs, looping over every light
    triangle as the reference does."""
    n, model, voff, V, N, foff, F, mat = _tracer_scene(objects)
    P, D = _c(P, np.float32).reshape(-1, 3), _c(D, np.float32).reshape(-1, 3)
    out = np.empty(P.shape[0], np.float32)
    if not load().orc_light_pdf(n, model, voff, V, foff, F, mat, P, D, P.shape[0], out):
        raise ValueError("singular model matrix")
    return out
