"""Wavefront path tracer over the library's TLAS (gpurt_scene_create / gpurt_render).

The host side of the reference's ray-tracing pipeline (RTPipe, src/vk/rt.cpp:26-62 and
346-394) as a Python object: a scene is a list of objects, each a mesh of Mesh::Vertex
records with a model matrix and material factors; Scene builds one BLAS per object and a
TLAS whose custom index i names object i (GPURT::build_accel, src/gpurt.cpp:220-241), then
renders frames with the reference's integrators (direct light sampling, material sampling,
MIS), BRDFs (Blinn-Phong, GGX) and RNG.

    objs = tracer.objects_from_gltf(gltf.load("cbox.gltf"))
    scene = tracer.Scene(ctx, objs)
    iv, ip = tracer.camera(eye=(0, 0, 20), target=(0, 0, 0), yfov_deg=41.1, aspect=4 / 3)
    img = scene.render(640, 480, iv, ip, samples=8, max_depth=8, integrator="mis")

Object dicts: verts (n, 12) float32 (pos.xyz, u, normal.xyz, v, tangent.xyzw), indices
(m, 3) uint32/int32, model (4, 4) row-major object-to-world, albedo (3), emissive (3),
metal_rough (2).  Images are (height, width, 4) float32 device tensors; row y is launch
row y (camera-space y grows with the row index).
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from . import Bvh, Tlas, _need, _ptr, _stream, _track

INTEGRATORS = {"direct": 0, "mats": 1, "mis": 2}
BRDFS = {"blinn-phong": 0, "ggx": 1}


def camera(eye, target, up=(0.0, 1.0, 0.0), yfov_deg=53.13, aspect=1.0, near=0.1):
    """(inv_view, inv_proj) as float64 (4, 4) row-major matrices for a right-handed look-at
    camera (looking down -z in view space) and an infinite reversed-depth perspective
    (depth n/z: 1 at the near plane, 0 at infinity), the conventions of the reference's
    Mat4::look_at / Mat4::project (src/lib/mat4.h:39-49).  The tracer only uses
    inv_proj * (ndc.x, ndc.y, 0, 1) as a view-space direction, so any projection with the
    same x / y scaling gives the same rays."""
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f = f / np.linalg.norm(f)
    r = np.cross(f, up)
    r = r / np.linalg.norm(r)
    u = np.cross(r, f)
    iv = np.eye(4)
    iv[:3, 0], iv[:3, 1], iv[:3, 2], iv[:3, 3] = r, u, -f, eye
    t = 1.0 / math.tan(math.radians(yfov_deg) / 2.0)
    ip = np.zeros((4, 4))
    ip[0, 0] = aspect / t
    ip[1, 1] = 1.0 / t
    ip[2, 3] = -1.0
    ip[3, 2] = 1.0 / near
    return iv, ip


def objects_from_gltf(scene):
    """Tracer objects from gltf.load()'s result (factors only, as the reference reads
    them at src/scene/scene.cpp:286-303)."""
    out = []
    for o in scene["objects"]:
        m = o["material"]
        out.append({"verts": o["verts"], "indices": o["indices"], "model": np.asarray(o["transform"], np.float64),
                    "albedo": m["albedo"], "emissive": m["emissive"], "metal_rough": m["metal_rough"],
                    "name": o.get("name", "")})
    return out


def _vec(v, n):
    a = [float(x) for x in np.asarray(v, np.float64).reshape(-1)]
    if len(a) == 1:
        a = a * n
    if len(a) != n:
        raise ValueError(f"expected {n} values")
    return a


class Scene:
    """A renderable scene (gpurt_scene_create): BLAS per object, TLAS over them, the light
    list and the light BVH.  Keeps every device buffer it hands to the library alive."""

    def __init__(self, ctx, objects, node_width=4, leaf_size=4, node_format=0):
        lib = _lib.load()
        self.ctx = ctx
        dev = ctx.device
        self.objects = objects
        self._verts, self._idx, self._blas = [], [], []
        arr = (_lib.Object * len(objects))()
        inst = []
        for i, o in enumerate(objects):
            v = torch.as_tensor(np.ascontiguousarray(np.asarray(o["verts"], np.float32).reshape(-1, 12)), device=dev)
            f = torch.as_tensor(np.ascontiguousarray(np.asarray(o["indices"]).astype(np.int32).reshape(-1, 3)),
                                device=dev)
            model = np.asarray(o["model"], np.float64).reshape(4, 4)
            b = Bvh(ctx, v, f, node_width=node_width, leaf_size=leaf_size, node_format=node_format)
            self._verts.append(v)
            self._idx.append(f)
            self._blas.append(b)
            inst.append((b, model[:3], i))
            a = arr[i]
            a.verts, a.indices = _ptr(v), _ptr(f)
            a.n_verts, a.n_tris = v.shape[0], f.shape[0]
            a.model[:] = [float(x) for x in model.reshape(-1)]
            a.albedo[:] = _vec(o.get("albedo", 1.0), 3)
            a.emissive[:] = _vec(o.get("emissive", 0.0), 3)
            a.metal_rough[:] = _vec(o.get("metal_rough", (1.0, 1.0)), 2)
        self.tlas = Tlas(ctx, inst)
        h = ctypes.c_void_p()
        _lib.check(lib.gpurt_scene_create(self.tlas._h, arr, len(objects), _stream(dev), ctypes.byref(h)))
        self._h = h
        _track("scene", self)

    @property
    def device(self):
        return self.ctx.device

    def render(self, width, height, inv_view, inv_proj, samples=8, max_depth=8, integrator=0, brdf=0,
               use_rr=True, clear_col=0.3, env_light=0.0, frame=0, seed=0, qmc=False, max_frame=256,
               use_metalness=False, image=None, gbuffer=False, ray_counts=None, concurrency=0):
        """One frame (gpurt_render).  Defaults are RTPipe's (src/vk/rt.h:36-51); env_light is
        env * env_scale.  With frame > 0, `image` (the previous frame) is blended in place.
        ray_counts: optional numpy uint64 array of 2 that accumulates the rays traced
        (closest-hit bounce rays, shadow / light rays).
        Returns the image, or (image, (pos, normal, albedo)) with gbuffer=True."""
        dev = self.device
        integrator = INTEGRATORS.get(integrator, integrator)
        brdf = BRDFS.get(brdf, brdf)
        if image is None:
            image = torch.zeros((height, width, 4), dtype=torch.float32, device=dev)
        else:
            _need(image, torch.float32, (width, 4), "image", dev)
            if image.shape[0] != height:
                raise ValueError("image must be (height, width, 4)")
        g = [torch.zeros((height, width, 4), dtype=torch.float32, device=dev) for _ in range(3)] if gbuffer else None
        p = _lib.RenderParams()
        p.inv_view[:] = _vec(inv_view, 16)
        p.inv_proj[:] = _vec(inv_proj, 16)
        p.clear_col[:] = _vec(clear_col, 3)
        p.env_light[:] = _vec(env_light, 3)
        p.width, p.height, p.frame, p.samples = int(width), int(height), int(frame), int(samples)
        p.max_frame, p.qmc, p.max_depth = int(max_frame), int(bool(qmc)), int(max_depth)
        p.use_metalness, p.integrator, p.brdf = int(bool(use_metalness)), int(integrator), int(brdf)
        p.use_rr, p.seed = int(bool(use_rr)), int(seed) & 0xFFFFFFFF
        p.concurrency = int(concurrency)
        if ray_counts is not None:
            if not (isinstance(ray_counts, np.ndarray) and ray_counts.dtype == np.uint64 and ray_counts.size >= 2
                    and ray_counts.flags.c_contiguous):
                raise ValueError("ray_counts must be a contiguous numpy uint64 array of >= 2 elements")
            p.ray_counts = ray_counts.ctypes.data
        _lib.check(_lib.load().gpurt_render(self._h, ctypes.byref(p), _ptr(image),
                                            _ptr(g[0]) if g else None, _ptr(g[1]) if g else None,
                                            _ptr(g[2]) if g else None, _stream(dev)))
        return (image, tuple(g)) if gbuffer else image

    def light_pdf(self, p, d):
        """light_pdf (rt.rgen:222-255) for (n, 3) points and unit directions."""
        dev = self.device
        _need(p, torch.float32, (3,), "p", dev)
        _need(d, torch.float32, (3,), "d", dev)
        out = torch.empty(p.shape[0], dtype=torch.float32, device=dev)
        _lib.check(_lib.load().gpurt_scene_light_pdf(self._h, _ptr(p), _ptr(d), p.shape[0], _ptr(out),
                                                     _stream(dev)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().gpurt_scene_destroy(self._h)
            self._h = None
        if getattr(self, "tlas", None) is not None:
            self.tlas.close()
            self.tlas = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
