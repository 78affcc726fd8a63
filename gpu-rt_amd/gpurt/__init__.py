"""gpurt -- MI355X geometric queries (BVH build, closest point, closest/any-hit rays).

Python host layer over libgpurt.so (include/gpurt.h).  Torch tensors supply device
memory and the current HIP stream; every computation runs in the library's HIP kernels.

    ctx = gpurt.Context(0)
    bvh = gpurt.Bvh(ctx, verts, tris)             # VK::Accel(const Mesh&) equivalent
    prim, pt, dist, bary = bvh.closest_point(q)
    prim, t, uv = bvh.trace_closest(rays)          # rays: (n, 8) = o.xyz, tmin, d.xyz, tmax
    occ = bvh.trace_any(rays)
    d = gpurt.closest_point_distance(bvh, q, verts)   # differentiable
    t = gpurt.ray_hit_t(bvh, rays_o, rays_d, verts)   # differentiable
"""
import atexit
import ctypes
import weakref

import torch

from . import _lib
from ._lib import BuildOpts, BvhInfo, GpurtError, QueryOpts

__all__ = ["Context", "Bvh", "Batch", "GpurtError", "closest_point_distance", "ray_hit_t",
           "make_rays", "default_context"]

EPS = 1e-5          # rtcommon.glsl:7 -- ray tmin of trace_ray / visibility
LARGE_DIST = 1e7    # rtcommon.glsl:6 -- ray tmax of trace_ray


# Every open library object, by kind.  At interpreter exit they are closed in dependency
# order (scenes, batches, TLASes, BVHs, then contexts) before torch and the HIP runtime
# shut down, instead of in whatever order garbage collection reaches them.
_OPEN = {k: weakref.WeakSet() for k in ("scene", "batch", "tlas", "bvh", "ctx")}


def _track(kind, obj):
    _OPEN[kind].add(obj)


@atexit.register
def _close_all():
    for kind in ("scene", "batch", "tlas", "bvh", "ctx"):
        for obj in list(_OPEN[kind]):
            try:
                obj.close()
            except Exception:
                pass


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _need(t, dtype, shape_tail, name, device):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if t.device != device:
        raise ValueError(f"{name} must be on {device}, got {t.device}")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if tuple(t.shape[1:]) != tuple(shape_tail):
        raise ValueError(f"{name} must have shape (n, {', '.join(map(str, shape_tail))}), "
                         f"got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t


def _check_out(n, device, **arrays):
    """Caller-supplied output tensors: n rows of the right dtype and width, contiguous, on
    the device (None skips an optional output)."""
    for name, (t, dtype, tail) in arrays.items():
        if t is None:
            continue
        _need(t, dtype, tail, name, device)
        if t.shape[0] != n:
            raise ValueError(f"{name} has {t.shape[0]} rows for {n} queries")


class Context:
    """Library context on one HIP device (gpurt_ctx_create)."""

    def __init__(self, device=0, vertex_sums="auto"):
        lib = _lib.load()
        if isinstance(device, torch.device):
            device = device.index or 0
        self.device = torch.device("cuda", int(device))
        h = ctypes.c_void_p()
        _lib.check(lib.gpurt_ctx_create(int(device), ctypes.byref(h)))
        self._h = h
        _track("ctx", self)
        self.vertex_sums = vertex_sums

    _SUMS = {"auto": _lib.SUMS_AUTO, "fixed": _lib.SUMS_FIXED_ORDER, "atomic": _lib.SUMS_ATOMIC}

    @property
    def vertex_sums(self):
        """How backward passes add vertex-gradient terms (gpurt_ctx_set_vertex_sums):
        "fixed" (sorted by vertex, a fixed order: the same bits every run), "atomic"
        (FP32 atomics) or "auto" (the faster of the two for each call's size)."""
        return self._sums

    @vertex_sums.setter
    def vertex_sums(self, mode):
        if mode not in self._SUMS:
            raise ValueError(f"vertex_sums must be one of {sorted(self._SUMS)}, got {mode!r}")
        _lib.check(_lib.load().gpurt_ctx_set_vertex_sums(self._h, self._SUMS[mode]))
        self._sums = mode

    @property
    def handle(self):
        return self._h

    def status(self, clear=True):
        """Sticky conditions the kernels reported (gpurt_ctx_status): a set of
        "stack_overflow", "tlas_outside_domain"; waits for the device."""
        f = ctypes.c_uint32()
        _lib.check(_lib.load().gpurt_ctx_status(self._h, ctypes.byref(f), int(bool(clear))))
        names = {_lib.STATUS_STACK_OVERFLOW: "stack_overflow",
                 _lib.STATUS_TLAS_OUTSIDE_DOMAIN: "tlas_outside_domain"}
        return {n for bit, n in names.items() if f.value & bit}

    def close(self):
        if getattr(self, "_h", None):
            torch.cuda.synchronize(self.device)
            _lib.load().gpurt_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = {}


def default_context(device=0):
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


def make_rays(o, d, tmin=EPS, tmax=LARGE_DIST):
    """Pack origins/directions (n,3) and scalar or (n,) bounds into the (n,8) ray layout."""
    n = o.shape[0]
    r = torch.empty((n, 8), dtype=torch.float32, device=o.device)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


class Bvh:
    """A BVH over one triangle mesh (gpurt_bvh_build); mirrors VK::Accel(const Mesh&).

    verts: (n_verts, k) float32 device tensor with positions in columns 0..2 (k = 3, 4 or
    12 -- the reference's 48-byte Mesh::Vertex); tris: (n_tris, 3) int32/uint32 indices.
    """

    def __init__(self, ctx, verts, tris, node_width=4, leaf_size=4, treelet_size=0,
                 treelet_rounds=0, node_format=0):
        lib = _lib.load()
        self.ctx = ctx
        dev = ctx.device
        if verts.dim() != 2 or verts.shape[1] < 3:
            raise ValueError("verts must be (n, k>=3)")
        _need(verts, torch.float32, verts.shape[1:], "verts", dev)
        if tris.dtype not in (torch.int32, torch.uint32):
            raise TypeError("tris must be int32 or uint32")
        _need(tris, tris.dtype, (3,), "tris", dev)
        self.n_verts = verts.shape[0]
        self.n_tris = tris.shape[0]
        self._tris = tris  # indices are needed by no kernel after the build; kept for users
        opts = BuildOpts(node_width, leaf_size, treelet_size, treelet_rounds, node_format)
        h = ctypes.c_void_p()
        _lib.check(lib.gpurt_bvh_build(ctx.handle, _ptr(verts), verts.shape[1] * 4,
                                       verts.shape[0], _ptr(tris), tris.shape[0],
                                       ctypes.byref(opts), _stream(dev), ctypes.byref(h)))
        self._h = h
        _track("bvh", self)

    @property
    def handle(self):
        return self._h

    @property
    def device(self):
        return self.ctx.device

    def info(self):
        inf = BvhInfo()
        _lib.check(_lib.load().gpurt_bvh_get_info(self._h, ctypes.byref(inf)))
        return {f: getattr(inf, f) for f, _ in BvhInfo._fields_}

    def export(self):
        """The built tree (gpurt_bvh_export) as device tensors: "nodes" (n_nodes, 8*W)
        float32 FP32 wide nodes (child words and counts are the bit patterns of the last
        2*W columns), "qnodes" (n_nodes, bytes) uint8 8-bit nodes or None, "tris"
        (n_tris, 12) packed triangles, "leaf_order" (n_tris,) int32."""
        inf = self.info()
        W, nn, nt = inf["node_width"], inf["n_nodes"], inf["n_tris"]
        dev = self.device
        out = {"nodes": torch.empty((nn, 8 * W), dtype=torch.float32, device=dev),
               "qnodes": (torch.empty((nn, {2: 48, 4: 64, 8: 96}[W]), dtype=torch.uint8, device=dev)
                          if inf["node_format"] in (2, 3) else None),
               "tris": torch.empty((nt, 12), dtype=torch.float32, device=dev),
               "leaf_order": torch.empty(nt, dtype=torch.int32, device=dev)}
        arr = _lib.BvhArrays(*(_ptr(out[k]) for k in ("nodes", "qnodes", "tris", "leaf_order")))
        _lib.check(_lib.load().gpurt_bvh_export(self._h, ctypes.byref(arr), _stream(dev)))
        return out

    def refit(self, verts):
        _need(verts, torch.float32, verts.shape[1:], "verts", self.device)
        if verts.shape[0] != self.n_verts:
            raise ValueError("refit needs the same vertex count")
        _lib.check(_lib.load().gpurt_bvh_refit(self._h, _ptr(verts), verts.shape[1] * 4,
                                               _stream(self.device)))

    def _qopts(self, stack_depth, waves_per_cu, stats, leaf_trigger=0, refill_trigger=0):
        """stats: optional int64 device tensor of 4 counters (wide nodes visited,
        triangles tested, stack spills, reserved), accumulated by the counting kernel
        variant."""
        if stats is not None:
            if (stats.dtype != torch.int64 or stats.numel() < 4 or stats.device != self.device
                    or not stats.is_contiguous()):
                raise ValueError("stats must be a contiguous int64 tensor of >= 4 elements "
                                 "on the BVH's device")
        return QueryOpts(stack_depth, waves_per_cu,
                         ctypes.c_void_p(stats.data_ptr()) if stats is not None else None,
                         leaf_trigger, refill_trigger)

    def closest_point(self, q, with_bary=True, stack_depth=0, waves_per_cu=0, stats=None,
                      out=None, leaf_trigger=0):
        dev = self.device
        _need(q, torch.float32, (3,), "q", dev)
        n = q.shape[0]
        if out is None:
            prim = torch.empty(n, dtype=torch.int32, device=dev)
            pt = torch.empty((n, 3), dtype=torch.float32, device=dev)
            dist = torch.empty(n, dtype=torch.float32, device=dev)
            bary = torch.empty((n, 3), dtype=torch.float32, device=dev) if with_bary else None
        else:
            prim, pt, dist, bary = out
            _check_out(n, dev, prim=(prim, torch.int32, ()), pt=(pt, torch.float32, (3,)),
                       dist=(dist, torch.float32, ()), bary=(bary, torch.float32, (3,)))
        opts = self._qopts(stack_depth, waves_per_cu, stats, leaf_trigger)
        _lib.check(_lib.load().gpurt_closest_point(self._h, _ptr(q), n, _ptr(prim), _ptr(pt),
                                                   _ptr(dist), _ptr(bary), ctypes.byref(opts),
                                                   _stream(dev)))
        return prim, pt, dist, bary

    def trace_closest(self, rays, with_uv=True, stack_depth=0, waves_per_cu=0, stats=None,
                      out=None, leaf_trigger=0, refill_trigger=0):
        dev = self.device
        _need(rays, torch.float32, (8,), "rays", dev)
        n = rays.shape[0]
        if out is None:
            prim = torch.empty(n, dtype=torch.int32, device=dev)
            t = torch.empty(n, dtype=torch.float32, device=dev)
            uv = torch.empty((n, 2), dtype=torch.float32, device=dev) if with_uv else None
        else:
            prim, t, uv = out
            _check_out(n, dev, prim=(prim, torch.int32, ()), t=(t, torch.float32, ()),
                       uv=(uv, torch.float32, (2,)))
        opts = self._qopts(stack_depth, waves_per_cu, stats, leaf_trigger, refill_trigger)
        _lib.check(_lib.load().gpurt_trace_closest(self._h, _ptr(rays), n, _ptr(prim), _ptr(t),
                                                   _ptr(uv), ctypes.byref(opts), _stream(dev)))
        return prim, t, uv

    def trace_any(self, rays, stack_depth=0, waves_per_cu=0, stats=None, out=None, leaf_trigger=0,
                  refill_trigger=0):
        dev = self.device
        _need(rays, torch.float32, (8,), "rays", dev)
        n = rays.shape[0]
        occ = out if out is not None else torch.empty(n, dtype=torch.uint8, device=dev)
        _check_out(n, dev, occluded=(occ, torch.uint8, ()))
        opts = self._qopts(stack_depth, waves_per_cu, stats, leaf_trigger, refill_trigger)
        _lib.check(_lib.load().gpurt_trace_any(self._h, _ptr(rays), n, _ptr(occ),
                                               ctypes.byref(opts), _stream(dev)))
        return occ

    def _check_bwd(self, n, per_query, g_verts):
        """Shapes, dtypes and devices of a backward call's tensors: the kernels index
        g_verts by this BVH's vertex indices and the per-query arrays by query."""
        dev = self.device
        _check_out(n, dev, **per_query)
        if g_verts is not None:
            _need(g_verts, torch.float32, (3,), "g_verts", dev)
            if g_verts.shape[0] != self.n_verts:
                raise ValueError(f"g_verts must be ({self.n_verts}, 3) for this BVH, "
                                 f"got {tuple(g_verts.shape)}")

    def closest_point_bwd(self, q, prim, bary, pt, dist, g_dist, g_q=None, g_verts=None):
        """Accumulates g_verts (n_verts, 3) and writes g_q (n, 3) from the forward pass's
        saved ids (gpurt_closest_point_bwd)."""
        _need(q, torch.float32, (3,), "q", self.device)
        n = q.shape[0]
        self._check_bwd(n, {"prim": (prim, torch.int32, ()), "g_dist": (g_dist, torch.float32, ()),
                            "g_q": (g_q, torch.float32, (3,)), "bary": (bary, torch.float32, (3,)),
                            "pt": (pt, torch.float32, (3,)), "dist": (dist, torch.float32, ())}, g_verts)
        _lib.check(_lib.load().gpurt_closest_point_bwd(
            self._h, _ptr(q), n, _ptr(prim), _ptr(bary), _ptr(pt), _ptr(dist),
            _ptr(g_dist), _ptr(g_q), _ptr(g_verts), _stream(self.device)))

    def trace_bwd(self, rays, prim, t, uv, g_t, g_origin=None, g_dir=None, g_verts=None):
        """Accumulates g_verts (n_verts, 3) and writes g_origin / g_dir (n, 3) from the
        forward pass's saved ids, t and (u, v) (gpurt_trace_bwd)."""
        _need(rays, torch.float32, (8,), "rays", self.device)
        n = rays.shape[0]
        self._check_bwd(n, {"prim": (prim, torch.int32, ()), "t": (t, torch.float32, ()),
                            "uv": (uv, torch.float32, (2,)), "g_t": (g_t, torch.float32, ()),
                            "g_origin": (g_origin, torch.float32, (3,)),
                            "g_dir": (g_dir, torch.float32, (3,))}, g_verts)
        _lib.check(_lib.load().gpurt_trace_bwd(
            self._h, _ptr(rays), n, _ptr(prim), _ptr(t), _ptr(uv), _ptr(g_t),
            _ptr(g_origin), _ptr(g_dir), _ptr(g_verts), _stream(self.device)))

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().gpurt_bvh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many meshes built and queried together (gpurt_batch_create / gpurt_batch_run).

    meshes: list of dicts with device tensors "verts" (n,k>=3) float32, "tris" (m,3)
    int32, and optionally "queries" (nq,3) float32 and "rays" (nr,8) float32.  Output
    tensors are allocated once; every run() rebuilds each BVH from the current vertex
    values and answers every query, as one recorded HIP graph.  results[i] holds mesh
    i's outputs: q_prim, q_pt, q_dist, q_bary, r_prim, r_t, r_uv.
    """

    def __init__(self, ctx, meshes, node_width=4, leaf_size=4, treelet_size=0, stack_depth=0,
                 waves_per_cu=0, with_bary=True, with_uv=True, node_format=0, leaf_trigger=0,
                 refill_trigger=0):
        lib = _lib.load()
        self.ctx = ctx
        dev = ctx.device
        self._keep = []
        self.results = []
        arr = (_lib.BatchMesh * len(meshes))()
        for i, m in enumerate(meshes):
            v, f = m["verts"], m["tris"]
            _need(v, torch.float32, v.shape[1:], "verts", dev)
            _need(f, f.dtype, (3,), "tris", dev)
            q = m.get("queries")
            r = m.get("rays")
            nq = 0 if q is None else q.shape[0]
            nr = 0 if r is None else r.shape[0]
            if q is not None:
                _need(q, torch.float32, (3,), "queries", dev)
            if r is not None:
                _need(r, torch.float32, (8,), "rays", dev)
            out = {
                "q_prim": torch.empty(nq, dtype=torch.int32, device=dev),
                "q_pt": torch.empty((nq, 3), dtype=torch.float32, device=dev),
                "q_dist": torch.empty(nq, dtype=torch.float32, device=dev),
                "q_bary": torch.empty((nq, 3), dtype=torch.float32, device=dev) if with_bary else None,
                "r_prim": torch.empty(nr, dtype=torch.int32, device=dev),
                "r_t": torch.empty(nr, dtype=torch.float32, device=dev),
                "r_uv": torch.empty((nr, 2), dtype=torch.float32, device=dev) if with_uv else None,
            }
            self.results.append(out)
            self._keep += [v, f, q, r]
            bm = arr[i]
            bm.verts, bm.vstride_bytes, bm.n_verts = v.data_ptr(), v.shape[1] * 4, v.shape[0]
            bm.tri_idx, bm.n_tris = f.data_ptr(), f.shape[0]
            bm.q, bm.n_q = (q.data_ptr() if q is not None else None), nq
            bm.q_prim, bm.q_pt, bm.q_dist = (out["q_prim"].data_ptr(), out["q_pt"].data_ptr(),
                                             out["q_dist"].data_ptr())
            bm.q_bary = out["q_bary"].data_ptr() if with_bary else None
            bm.rays, bm.n_r = (r.data_ptr() if r is not None else None), nr
            bm.r_prim, bm.r_t = out["r_prim"].data_ptr(), out["r_t"].data_ptr()
            bm.r_uv = out["r_uv"].data_ptr() if with_uv else None
        self._arr = arr
        self._meshes = meshes
        bo = BuildOpts(node_width, leaf_size, treelet_size, 0, node_format)
        qo = QueryOpts(stack_depth, waves_per_cu, None, leaf_trigger, refill_trigger)
        torch.cuda.synchronize(dev)  # inputs must be resident before the first (eager) run
        h = ctypes.c_void_p()
        _lib.check(lib.gpurt_batch_create(ctx.handle, arr, len(meshes), ctypes.byref(bo),
                                          ctypes.byref(qo), ctypes.byref(h)))
        self._h = h
        _track("batch", self)

    def run(self):
        _lib.check(_lib.load().gpurt_batch_run(self._h, _stream(self.ctx.device)))
        return self.results

    def enable_backward(self, query_grads=True, vertex_grads=True, distance=True, hit_t=True):
        """Allocate gradient buffers and record the batched backward (gpurt_batch_set_grads).

        grads[i] holds mesh i's inputs g_dist (n_q), g_t (n_r) -- upstream gradients of
        q_dist and r_t, ones by default -- and outputs g_q (n_q,3), g_origin, g_dir
        (n_r,3), g_verts (n_verts,3) (None where not requested).  backward() fills the
        outputs from the last run()."""
        dev = self.ctx.device
        arr = (_lib.BatchGrad * len(self.results))()
        self.grads = []
        for i, (r, m) in enumerate(zip(self.results, self._meshes)):
            nq, nr = r["q_prim"].shape[0], r["r_prim"].shape[0]
            if vertex_grads and hit_t and nr and r["r_uv"] is None:
                raise ValueError("vertex gradients of hit t need with_uv=True")
            g = {
                "g_dist": torch.ones(nq, device=dev) if distance else None,
                "g_q": torch.empty((nq, 3), device=dev) if (distance and query_grads) else None,
                "g_t": torch.ones(nr, device=dev) if hit_t else None,
                "g_origin": torch.empty((nr, 3), device=dev) if (hit_t and query_grads) else None,
                "g_dir": torch.empty((nr, 3), device=dev) if (hit_t and query_grads) else None,
                "g_verts": torch.empty((m["verts"].shape[0], 3), device=dev) if vertex_grads else None,
            }
            self.grads.append(g)
            for k, v in g.items():
                setattr(arr[i], k, v.data_ptr() if v is not None else None)
        self._garr = arr
        torch.cuda.synchronize(dev)
        _lib.check(_lib.load().gpurt_batch_set_grads(self._h, arr))
        return self.grads

    def backward(self):
        """Run the recorded backward graph (after run()); returns the gradient dicts."""
        _lib.check(_lib.load().gpurt_batch_run_backward(self._h, _stream(self.ctx.device)))
        return self.grads

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().gpurt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tlas:
    """Two-level acceleration structure over instances of BLASes (gpurt_tlas_build);
    mirrors VK::Accel(blas, T) (src/vk/vulkan.cpp:777-856).

    instances: list of (bvh, transform) or (bvh, transform, custom_index[, mask]) with
    transform a row-major 3x4 (or 4x4, last row ignored) object-to-world matrix; the
    custom index defaults to the instance's position (as Accel::recreate sets it) and
    the mask to 0xFF.  All BVHs need the same node width and ray node format and must
    stay alive (this object keeps references)."""

    def __init__(self, ctx, instances, treelet_size=0):
        lib = _lib.load()
        self.ctx = ctx
        arr = (_lib.Instance * len(instances))()
        self._blas = []
        for i, inst in enumerate(instances):
            bvh, xf = inst[0], inst[1]
            custom = inst[2] if len(inst) > 2 and inst[2] is not None else i
            mask = inst[3] if len(inst) > 3 else 0xFF
            m = torch.as_tensor(xf, dtype=torch.float64).reshape(-1)
            if m.numel() == 16:
                m = m.reshape(4, 4)[:3].reshape(-1)
            if m.numel() != 12:
                raise ValueError("transform must be 3x4 or 4x4")
            arr[i].transform[:] = [float(x) for x in m.tolist()]
            arr[i].custom_index = int(custom)
            arr[i].mask = int(mask)
            arr[i].blas = bvh.handle
            self._blas.append(bvh)
        self.n_instances = len(instances)
        opts = BuildOpts(0, 0, treelet_size, 0, 0)
        h = ctypes.c_void_p()
        _lib.check(lib.gpurt_tlas_build(ctx.handle, arr, len(instances), ctypes.byref(opts),
                                        _stream(ctx.device), ctypes.byref(h)))
        self._h = h
        _track("tlas", self)

    @property
    def device(self):
        return self.ctx.device

    def _qopts(self, stack_depth, waves_per_cu, stats, leaf_trigger, refill_trigger=0):
        if stats is not None and (stats.dtype != torch.int64 or stats.numel() < 4):
            raise ValueError("stats must be an int64 tensor of >= 4 elements")
        return QueryOpts(stack_depth, waves_per_cu,
                         ctypes.c_void_p(stats.data_ptr()) if stats is not None else None,
                         leaf_trigger, refill_trigger)

    def trace_closest(self, rays, with_uv=True, stack_depth=0, waves_per_cu=0, stats=None,
                      out=None, leaf_trigger=0, refill_trigger=0):
        """Returns (obj, prim, t, uv): obj = custom index of the hit instance (-1: miss)."""
        dev = self.device
        _need(rays, torch.float32, (8,), "rays", dev)
        n = rays.shape[0]
        if out is None:
            obj = torch.empty(n, dtype=torch.int32, device=dev)
            prim = torch.empty(n, dtype=torch.int32, device=dev)
            t = torch.empty(n, dtype=torch.float32, device=dev)
            uv = torch.empty((n, 2), dtype=torch.float32, device=dev) if with_uv else None
        else:
            obj, prim, t, uv = out
            _check_out(n, dev, obj=(obj, torch.int32, ()), prim=(prim, torch.int32, ()),
                       t=(t, torch.float32, ()), uv=(uv, torch.float32, (2,)))
        opts = self._qopts(stack_depth, waves_per_cu, stats, leaf_trigger, refill_trigger)
        _lib.check(_lib.load().gpurt_tlas_trace_closest(self._h, _ptr(rays), n, _ptr(obj), _ptr(prim),
                                                        _ptr(t), _ptr(uv), ctypes.byref(opts),
                                                        _stream(dev)))
        return obj, prim, t, uv

    def trace_any(self, rays, stack_depth=0, waves_per_cu=0, stats=None, out=None, leaf_trigger=0,
                  refill_trigger=0):
        dev = self.device
        _need(rays, torch.float32, (8,), "rays", dev)
        n = rays.shape[0]
        occ = out if out is not None else torch.empty(n, dtype=torch.uint8, device=dev)
        _check_out(n, dev, occluded=(occ, torch.uint8, ()))
        opts = self._qopts(stack_depth, waves_per_cu, stats, leaf_trigger, refill_trigger)
        _lib.check(_lib.load().gpurt_tlas_trace_any(self._h, _ptr(rays), n, _ptr(occ),
                                                    ctypes.byref(opts), _stream(dev)))
        return occ

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().gpurt_tlas_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _check_verts(bvh, verts):
    """The differentiable wrappers return a gradient for `verts`, which must be the vertex
    array the BVH was built on (or one of the same shape, e.g. after a refit)."""
    if verts.dim() != 2 or verts.shape[1] < 3 or verts.shape[0] != bvh.n_verts:
        raise ValueError(f"verts must be ({bvh.n_verts}, k>=3) like the BVH's vertex array, "
                         f"got {tuple(verts.shape)}")


def _vertex_grad(g_v, shape):
    """The (n_verts, 3) position gradient laid out like verts: columns 0..2 of a zero
    (n_verts, k) tensor for the 16- and 48-byte vertex records."""
    if g_v is None or shape[1] == 3:
        return g_v
    out = torch.zeros(tuple(shape), dtype=g_v.dtype, device=g_v.device)
    out[:, :3] = g_v
    return out


class _ClosestPointDistance(torch.autograd.Function):
    """d_i = min over triangles |q_i - p| ; backward via gpurt_closest_point_bwd."""

    @staticmethod
    def forward(ctx, bvh, q, verts):
        _check_verts(bvh, verts)
        q = q.contiguous()
        prim, pt, dist, bary = bvh.closest_point(q)
        ctx.bvh = bvh
        ctx.verts_shape = verts.shape
        ctx.save_for_backward(q, prim, pt, dist, bary)
        ctx.mark_non_differentiable(prim)
        return dist, prim

    @staticmethod
    def backward(ctx, g_dist, _g_prim):
        q, prim, pt, dist, bary = ctx.saved_tensors
        bvh = ctx.bvh
        need_q, need_v = ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        g_dist = g_dist.contiguous().float()
        g_q = torch.empty_like(q) if need_q else None
        g_v = torch.zeros((bvh.n_verts, 3), dtype=torch.float32, device=q.device) if need_v else None
        bvh.closest_point_bwd(q, prim, bary, pt, dist, g_dist, g_q, g_v)
        return None, g_q, _vertex_grad(g_v, ctx.verts_shape)


class _RayHitT(torch.autograd.Function):
    """t_i = closest-hit distance ; backward via gpurt_trace_bwd (misses give zero)."""

    @staticmethod
    def forward(ctx, bvh, o, d, verts, tmin, tmax):
        _check_verts(bvh, verts)
        rays = make_rays(o, d, tmin, tmax)
        prim, t, uv = bvh.trace_closest(rays)
        ctx.bvh = bvh
        ctx.verts_shape = verts.shape
        ctx.save_for_backward(rays, prim, t, uv)
        ctx.mark_non_differentiable(prim)
        return t, prim

    @staticmethod
    def backward(ctx, g_t, _g_prim):
        rays, prim, t, uv = ctx.saved_tensors
        bvh = ctx.bvh
        need_o, need_d, need_v = ctx.needs_input_grad[1:4]
        n = rays.shape[0]
        g_t = g_t.contiguous().float()
        g_o = torch.empty((n, 3), dtype=torch.float32, device=rays.device) if need_o else None
        g_d = torch.empty((n, 3), dtype=torch.float32, device=rays.device) if need_d else None
        g_v = torch.zeros((bvh.n_verts, 3), dtype=torch.float32, device=rays.device) if need_v else None
        bvh.trace_bwd(rays, prim, t, uv, g_t, g_o, g_d, g_v)
        return None, g_o, g_d, _vertex_grad(g_v, ctx.verts_shape), None, None


def closest_point_distance(bvh, q, verts):
    """Unsigned distance from each q to the mesh of `bvh` (built on `verts`), with
    gradients w.r.t. q and verts.  Returns (dist, prim)."""
    return _ClosestPointDistance.apply(bvh, q, verts)


def ray_hit_t(bvh, o, d, verts, tmin=EPS, tmax=LARGE_DIST):
    """Closest-hit distance of rays (o, d) against the mesh of `bvh` (built on `verts`),
    with gradients w.r.t. o, d and verts.  Returns (t, prim); t = +inf on a miss."""
    return _RayHitT.apply(bvh, o, d, verts, tmin, tmax)
