"""ctypes binding of libgpurt.so (the C ABI declared in include/gpurt.h).

The library is loaded from this directory.  There is no fallback: if the HIP library
is missing or cannot be loaded, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpurt.so")

GPURT_OK = 0
GPURT_E_INVALID = -1
GPURT_E_OOM = -2
GPURT_E_STACK_OVERFLOW = -3
GPURT_E_DEVICE = -4
GPURT_E_UNSUPPORTED = -5
STATUS_STACK_OVERFLOW = 1
STATUS_TLAS_OUTSIDE_DOMAIN = 2
SUMS_AUTO, SUMS_FIXED_ORDER, SUMS_ATOMIC = 0, 1, 2

_ERR_NAMES = {
    GPURT_E_INVALID: "GPURT_E_INVALID",
    GPURT_E_OOM: "GPURT_E_OOM",
    GPURT_E_STACK_OVERFLOW: "GPURT_E_STACK_OVERFLOW",
    GPURT_E_DEVICE: "GPURT_E_DEVICE",
    GPURT_E_UNSUPPORTED: "GPURT_E_UNSUPPORTED",
}

# Every symbol include/gpurt.h declares.
EXPORTS = (
    "gpurt_ctx_create", "gpurt_ctx_destroy", "gpurt_ctx_status", "gpurt_ctx_set_vertex_sums", "gpurt_bvh_build", "gpurt_bvh_refit",
    "gpurt_bvh_destroy", "gpurt_bvh_get_info", "gpurt_bvh_export", "gpurt_closest_point",
    "gpurt_trace_closest",
    "gpurt_trace_closest_noalloc", "gpurt_trace_any", "gpurt_closest_point_bwd", "gpurt_trace_bwd",
    "gpurt_batch_create",
    "gpurt_batch_run", "gpurt_batch_set_grads", "gpurt_batch_run_backward", "gpurt_batch_destroy",
    "gpurt_tune_batch", "gpurt_tune_queries", "gpurt_tune_cache_load", "gpurt_tune_cache_store",
    "gpurt_tlas_build", "gpurt_tlas_destroy", "gpurt_tlas_trace_closest", "gpurt_tlas_trace_any",
    "gpurt_scene_create", "gpurt_scene_destroy", "gpurt_render", "gpurt_scene_light_pdf",
    "gpurt_last_error",
)


class GpurtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class BuildOpts(ctypes.Structure):
    _fields_ = [("node_width", ctypes.c_int32), ("leaf_size", ctypes.c_int32),
                ("treelet_size", ctypes.c_int32), ("treelet_rounds", ctypes.c_int32),
                ("node_format", ctypes.c_int32)]


class QueryOpts(ctypes.Structure):
    _fields_ = [("stack_depth", ctypes.c_int32), ("waves_per_cu", ctypes.c_int32),
                ("stats", ctypes.c_void_p), ("leaf_trigger", ctypes.c_int32),
                ("refill_trigger", ctypes.c_int32)]


class BvhInfo(ctypes.Structure):
    _fields_ = [("n_tris", ctypes.c_int64), ("n_verts", ctypes.c_int64),
                ("n_nodes", ctypes.c_int64), ("node_bytes", ctypes.c_int64),
                ("tri_bytes", ctypes.c_int64), ("depth", ctypes.c_int32),
                ("node_width", ctypes.c_int32), ("leaf_size", ctypes.c_int32),
                ("sah_cost", ctypes.c_float), ("node_format", ctypes.c_int32),
                ("cpq_node_bytes", ctypes.c_int32)]


class BvhArrays(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_void_p), ("qnodes", ctypes.c_void_p), ("tris", ctypes.c_void_p),
                ("leaf_order", ctypes.c_void_p)]


class BatchMesh(ctypes.Structure):
    _fields_ = [("verts", ctypes.c_void_p), ("vstride_bytes", ctypes.c_size_t),
                ("n_verts", ctypes.c_size_t), ("tri_idx", ctypes.c_void_p),
                ("n_tris", ctypes.c_size_t), ("q", ctypes.c_void_p), ("n_q", ctypes.c_size_t),
                ("q_prim", ctypes.c_void_p), ("q_pt", ctypes.c_void_p),
                ("q_dist", ctypes.c_void_p), ("q_bary", ctypes.c_void_p),
                ("rays", ctypes.c_void_p), ("n_r", ctypes.c_size_t),
                ("r_prim", ctypes.c_void_p), ("r_t", ctypes.c_void_p),
                ("r_uv", ctypes.c_void_p)]


class TuneSpace(ctypes.Structure):
    _fields_ = [("node_width", ctypes.c_void_p), ("n_node_width", ctypes.c_int32),
                ("stack_depth", ctypes.c_void_p), ("n_stack_depth", ctypes.c_int32),
                ("waves_per_cu", ctypes.c_void_p), ("n_waves_per_cu", ctypes.c_int32),
                ("treelet_size", ctypes.c_void_p), ("n_treelet_size", ctypes.c_int32),
                ("node_format", ctypes.c_void_p), ("n_node_format", ctypes.c_int32),
                ("reps", ctypes.c_int32), ("builds_per_run", ctypes.c_float),
                ("cache_path", ctypes.c_char_p)]


class TuneResult(ctypes.Structure):
    _fields_ = [("build", BuildOpts), ("query", QueryOpts), ("best_ms", ctypes.c_float),
                ("default_ms", ctypes.c_float), ("candidates", ctypes.c_int32),
                ("from_cache", ctypes.c_int32), ("key", ctypes.c_char * 128)]


class Instance(ctypes.Structure):
    _fields_ = [("transform", ctypes.c_float * 12), ("custom_index", ctypes.c_uint32),
                ("mask", ctypes.c_uint32), ("blas", ctypes.c_void_p)]


class BatchGrad(ctypes.Structure):
    _fields_ = [("g_dist", ctypes.c_void_p), ("g_q", ctypes.c_void_p), ("g_t", ctypes.c_void_p),
                ("g_origin", ctypes.c_void_p), ("g_dir", ctypes.c_void_p),
                ("g_verts", ctypes.c_void_p)]


class Object(ctypes.Structure):
    _fields_ = [("verts", ctypes.c_void_p), ("indices", ctypes.c_void_p), ("n_verts", ctypes.c_uint32),
                ("n_tris", ctypes.c_uint32), ("model", ctypes.c_float * 16), ("albedo", ctypes.c_float * 3),
                ("emissive", ctypes.c_float * 3), ("metal_rough", ctypes.c_float * 2)]


class RenderParams(ctypes.Structure):
    _fields_ = [("inv_view", ctypes.c_float * 16), ("inv_proj", ctypes.c_float * 16),
                ("clear_col", ctypes.c_float * 3), ("env_light", ctypes.c_float * 3),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("frame", ctypes.c_int32),
                ("samples", ctypes.c_int32), ("max_frame", ctypes.c_int32), ("qmc", ctypes.c_int32),
                ("max_depth", ctypes.c_int32), ("use_metalness", ctypes.c_int32), ("integrator", ctypes.c_int32),
                ("brdf", ctypes.c_int32), ("use_rr", ctypes.c_int32), ("seed", ctypes.c_uint32),
                ("ray_counts", ctypes.c_void_p), ("concurrency", ctypes.c_int32), ("reserved", ctypes.c_int32)]


_lib = None


def load():
    """Load libgpurt.so once; raises if it is absent (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C gpu-rt_amd` "
                          "(or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    P, I, S, V = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, None
    sig = {
        "gpurt_ctx_create": (I, [I, ctypes.POINTER(P)]),
        "gpurt_ctx_destroy": (V, [P]),
        "gpurt_ctx_status": (I, [P, ctypes.POINTER(ctypes.c_uint32), I]),
        "gpurt_ctx_set_vertex_sums": (I, [P, I]),
        "gpurt_bvh_build": (I, [P, P, S, S, P, S, ctypes.POINTER(BuildOpts), P, ctypes.POINTER(P)]),
        "gpurt_bvh_refit": (I, [P, P, S, P]),
        "gpurt_bvh_destroy": (V, [P]),
        "gpurt_bvh_get_info": (I, [P, ctypes.POINTER(BvhInfo)]),
        "gpurt_bvh_export": (I, [P, ctypes.POINTER(BvhArrays), P]),
        "gpurt_closest_point": (I, [P, P, S, P, P, P, P, ctypes.POINTER(QueryOpts), P]),
        "gpurt_trace_closest": (I, [P, P, S, P, P, P, ctypes.POINTER(QueryOpts), P]),
        "gpurt_trace_closest_noalloc": (I, [P, P, S, P, P, P, ctypes.POINTER(QueryOpts), P]),
        "gpurt_trace_any": (I, [P, P, S, P, ctypes.POINTER(QueryOpts), P]),
        "gpurt_closest_point_bwd": (I, [P, P, S, P, P, P, P, P, P, P, P]),
        "gpurt_trace_bwd": (I, [P, P, S, P, P, P, P, P, P, P, P]),
        "gpurt_batch_create": (I, [P, ctypes.POINTER(BatchMesh), S, ctypes.POINTER(BuildOpts),
                                   ctypes.POINTER(QueryOpts), ctypes.POINTER(P)]),
        "gpurt_batch_run": (I, [P, P]),
        "gpurt_batch_set_grads": (I, [P, ctypes.POINTER(BatchGrad)]),
        "gpurt_batch_run_backward": (I, [P, P]),
        "gpurt_tune_batch": (I, [P, ctypes.POINTER(BatchMesh), S, ctypes.POINTER(TuneSpace), P,
                                 ctypes.POINTER(TuneResult)]),
        "gpurt_tune_queries": (I, [P, P, S, S, P, S, P, S, P, S, ctypes.POINTER(TuneSpace), P,
                                   ctypes.POINTER(TuneResult)]),
        "gpurt_tune_cache_load": (I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(TuneResult)]),
        "gpurt_tune_cache_store": (I, [ctypes.c_char_p, ctypes.POINTER(TuneResult)]),
        "gpurt_tlas_build": (I, [P, ctypes.POINTER(Instance), S, ctypes.POINTER(BuildOpts), P,
                                 ctypes.POINTER(P)]),
        "gpurt_tlas_destroy": (V, [P]),
        "gpurt_tlas_trace_closest": (I, [P, P, S, P, P, P, P, ctypes.POINTER(QueryOpts), P]),
        "gpurt_tlas_trace_any": (I, [P, P, S, P, ctypes.POINTER(QueryOpts), P]),
        "gpurt_batch_destroy": (V, [P]),
        "gpurt_scene_create": (I, [P, ctypes.POINTER(Object), S, P, ctypes.POINTER(P)]),
        "gpurt_scene_destroy": (V, [P]),
        "gpurt_render": (I, [P, ctypes.POINTER(RenderParams), P, P, P, P, P]),
        "gpurt_scene_light_pdf": (I, [P, P, P, S, P, P]),
        "gpurt_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != GPURT_OK:
        msg = load().gpurt_last_error()
        raise GpurtError(rc, msg.decode() if msg else "")
    return rc
