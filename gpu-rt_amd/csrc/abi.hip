// abi.hip -- the extern "C" entry points of include/gpurt.h: argument checks, device
// selection, ownership, and dispatch to the builder / traversal / backward launchers.
#include <cstring>
#include <string>

#include "gpurt_internal.h"

namespace gpurt_impl {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_error(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) return GPURT_E_OOM;
    return GPURT_E_DEVICE;
}

// Makes `dev` current for the scope and restores the caller's device afterwards.
struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        ok = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

__global__ void k_compact_verts(uint32_t n, const float* __restrict__ v, uint64_t stride,
                                float* __restrict__ out) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float* p = v + (size_t)i * stride;
    out[3 * (size_t)i] = p[0];
    out[3 * (size_t)i + 1] = p[1];
    out[3 * (size_t)i + 2] = p[2];
}

void note_use(gpurt_bvh_s* b, hipStream_t s) {
    std::lock_guard<std::mutex> lk(b->use_mu);
    // work recorded into a graph runs when the graph is launched: no event can mark it
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
        (void)hipGetLastError();
        b->uses_lost = true;
        return;
    }
    for (auto& u : b->uses)
        if (u.first == s) {
            if (hipEventRecord(u.second, s) != hipSuccess) {
                (void)hipGetLastError();
                b->uses_lost = true;
            }
            return;
        }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess || hipEventRecord(e, s) != hipSuccess) {
        (void)hipGetLastError();
        if (e) (void)hipEventDestroy(e);
        b->uses_lost = true;
        return;
    }
    b->uses.emplace_back(s, e);
}

// note_use without creating an event (the no-allocation entry point): a stream not seen
// before makes destroy wait for the whole device instead
static void note_use_existing(gpurt_bvh_s* b, hipStream_t s) {
    std::lock_guard<std::mutex> lk(b->use_mu);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
        (void)hipGetLastError();
        b->uses_lost = true;
        return;
    }
    for (auto& u : b->uses)
        if (u.first == s) {
            if (hipEventRecord(u.second, s) != hipSuccess) {
                (void)hipGetLastError();
                b->uses_lost = true;
            }
            return;
        }
    b->uses_lost = true;
}

void wait_uses(gpurt_bvh_s* b, hipStream_t s) {
    std::lock_guard<std::mutex> lk(b->use_mu);
    for (auto& u : b->uses)
        if (hipStreamWaitEvent(s, u.second, 0) != hipSuccess) {
            (void)hipGetLastError();
            b->uses_lost = true;
        }
}

// frees a BVH's buffers in order on stream s (all come from the stream-ordered pool)
static void free_bvh_buffers(gpurt_bvh_s* b, hipStream_t s) {
    for (void* p : {b->nodes, b->qnodes, (void*)b->tris, (void*)b->tri_idx, (void*)b->verts, (void*)b->leaf_order})
        if (p) (void)hipFreeAsync(p, s);
    b->nodes = b->qnodes = nullptr;
    b->tris = nullptr;
    b->tri_idx = b->leaf_order = nullptr;
    b->verts = nullptr;
}

int compact_verts(const float* verts, size_t stride_floats, size_t n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_compact_verts, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, (uint32_t)n, verts,
                       (uint64_t)stride_floats, out);
    GPURT_HIP(hipGetLastError());
    return GPURT_OK;
}

int ctx_scratch(gpurt_ctx_s* c, hipStream_t s, size_t bytes, CallScratch& out) {
    out = CallScratch{};
    out.s = s;
    hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
    const bool capturing = hipStreamIsCapturing(s, &cst) != hipSuccess || cst != hipStreamCaptureStatusNone;
    if (capturing) (void)hipGetLastError();
    if (!capturing) {
        std::lock_guard<std::mutex> lk(c->scratch_mu);
        int k = -1;
        for (size_t i = 0; i < c->scratch.size(); i++)
            if (c->scratch[i].s == s) {
                k = (int)i;
                break;
            }
        if (k < 0) {
            gpurt_ctx_s::StreamBuf nb;
            nb.s = s;
            if (hipEventCreateWithFlags(&nb.last, hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                nb.last = nullptr;
            }
            if (nb.last) {
                c->scratch.push_back(nb);
                k = (int)c->scratch.size() - 1;
            }
        }
        if (k >= 0 && !c->scratch[k].busy) {
            gpurt_ctx_s::StreamBuf& e = c->scratch[k];
            // a recycled stream handle: the buffer's previous user may be another stream
            if (e.mem && hipStreamWaitEvent(s, e.last, 0) != hipSuccess)
                return hip_error(hipGetLastError(), "hipStreamWaitEvent");
            if (e.bytes < bytes) {
                // Outside the stream-ordered pool, so the builds' own stream-ordered
                // allocations keep reusing the pool's memory.  Growing waits for the
                // device (hipFree), once per new largest call on the stream.
                if (e.mem) (void)hipFree(e.mem);
                e.mem = nullptr;
                e.bytes = 0;
                const size_t want = bytes + bytes / 4;  // room for somewhat larger calls
                if (hipMalloc(&e.mem, want) != hipSuccess) {
                    (void)hipGetLastError();
                    e.mem = nullptr;
                    return set_error(GPURT_E_OOM, "query scratch allocation failed");
                }
                e.bytes = want;
            }
            e.busy = true;
            out.mem = e.mem;
            out.slot = k;
            return GPURT_OK;
        }
    }
    // a call recorded into a graph, or the stream's buffer is held by another call
    if (hipMallocAsync(&out.mem, bytes, s) != hipSuccess) {
        (void)hipGetLastError();
        out.mem = nullptr;
        return set_error(GPURT_E_OOM, "query scratch allocation failed");
    }
    return GPURT_OK;
}

void ctx_scratch_done(gpurt_ctx_s* c, CallScratch& cs) {
    if (!cs.mem) return;
    if (cs.slot < 0) {
        (void)hipFreeAsync(cs.mem, cs.s);
    } else {
        std::lock_guard<std::mutex> lk(c->scratch_mu);
        gpurt_ctx_s::StreamBuf& e = c->scratch[cs.slot];
        if (hipEventRecord(e.last, cs.s) != hipSuccess) (void)hipGetLastError();
        e.busy = false;
    }
    cs = CallScratch{};
}

void ctx_scratch_release(gpurt_ctx_s* c) {
    std::lock_guard<std::mutex> lk(c->scratch_mu);
    for (auto& e : c->scratch) {
        if (e.mem) (void)hipFree(e.mem);
        if (e.last) (void)hipEventDestroy(e.last);
    }
    c->scratch.clear();
}

// node format used when gpurt_build_opts::node_format is 0
constexpr int kDefaultFormat = kFormatMixed;

int parse_build_opts(const gpurt_build_opts* o, BuildParams& p) {
    p = BuildParams{};
    if (o) {
        if (o->node_width) p.width = o->node_width;
        if (o->leaf_size) p.leaf_size = o->leaf_size;
        p.treelet_size = o->treelet_size;
        p.treelet_rounds = o->treelet_rounds;
        p.format = o->node_format;
    }
    if (p.width != 2 && p.width != 4 && p.width != 8)
        return set_error(GPURT_E_UNSUPPORTED, "node_width must be 2, 4 or 8");
    if (p.leaf_size < 1 || p.leaf_size > 8) return set_error(GPURT_E_UNSUPPORTED, "leaf_size must be 1..8");
    if (p.treelet_size != 0 && (p.treelet_size < 5 || p.treelet_size > 7))
        return set_error(GPURT_E_UNSUPPORTED, "treelet_size must be 0 or 5..7");
    if (p.treelet_rounds < 0 || p.treelet_rounds > 1)
        return set_error(GPURT_E_UNSUPPORTED, "treelet_rounds must be 0 or 1");
    if (p.treelet_size && p.treelet_rounds == 0) p.treelet_rounds = 1;
    if (p.format < 0 || p.format > 3) return set_error(GPURT_E_UNSUPPORTED, "node_format must be 0..3");
    if (p.format == 0) p.format = kDefaultFormat;
    return GPURT_OK;
}

}  // namespace gpurt_impl

using namespace gpurt_impl;

extern "C" {

const char* gpurt_last_error(void) { return g_last_error.c_str(); }

int gpurt_ctx_create(int device, gpurt_ctx* out) {
    if (!out) return set_error(GPURT_E_INVALID, "out is null");
    *out = nullptr;
    int n_dev = 0;
    GPURT_HIP(hipGetDeviceCount(&n_dev));
    if (device < 0 || device >= n_dev) return set_error(GPURT_E_INVALID, "device out of range");
    DeviceGuard g(device);
    if (!g.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    auto* c = new gpurt_ctx_s();
    c->device = device;
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && v > 0)
        c->n_cu = v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) ==
            hipSuccess &&
        v > 0)
        c->lds_per_cu = v;
    if (hipMalloc((void**)&c->err, 256) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return set_error(GPURT_E_OOM, "context allocation failed");
    }
    if (hipMemset(c->err, 0, 256) != hipSuccess ||
        hipStreamCreateWithFlags(&c->release, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(c->err);
        delete c;
        return set_error(GPURT_E_DEVICE, "context init failed");
    }
    // the fixed traversal scratch of gpurt_trace_closest_noalloc: its device address,
    // once per device (this device is current here)
    {
        const int rc = noalloc_scratch_init(device);
        if (rc) {
            (void)hipStreamDestroy(c->release);
            (void)hipFree(c->err);
            delete c;
            return rc;
        }
    }
    // keep freed stream-ordered allocations (traversal spill stacks) in the pool
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
        uint64_t thr = ~0ull;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    (void)hipGetLastError();
    *out = c;
    return GPURT_OK;
}

void gpurt_ctx_destroy(gpurt_ctx ctx) {
    if (!ctx) return;
    DeviceGuard g(ctx->device);
    (void)hipDeviceSynchronize();
    {
        // handles still open: their device memory goes with the context
        std::lock_guard<std::mutex> lk(ctx->live_mu);
        for (gpurt_tlas_s* t : ctx->tlases) tlas_orphan(t);
        ctx->tlases.clear();
        for (gpurt_bvh_s* b : ctx->bvhs) {
            free_bvh_buffers(b, ctx->release);
            for (auto& u : b->uses) (void)hipEventDestroy(u.second);
            b->uses.clear();
            b->ctx = nullptr;
        }
        ctx->bvhs.clear();
    }
    ctx_scratch_release(ctx);
    if (ctx->release) (void)hipStreamSynchronize(ctx->release);
    (void)hipGetLastError();
    if (ctx->release) (void)hipStreamDestroy(ctx->release);
    if (ctx->err) (void)hipFree(ctx->err);
    delete ctx;
}

int gpurt_ctx_status(gpurt_ctx ctx, uint32_t* flags, int clear) {
    if (!ctx || !flags) return set_error(GPURT_E_INVALID, "null argument");
    DeviceGuard g(ctx->device);
    if (!g.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    GPURT_HIP(hipDeviceSynchronize());
    uint32_t f = 0;
    GPURT_HIP(hipMemcpy(&f, ctx->err, sizeof(f), hipMemcpyDeviceToHost));
    if (clear) GPURT_HIP(hipMemset(ctx->err, 0, sizeof(f)));
    *flags = f;
    return GPURT_OK;
}

int gpurt_ctx_set_vertex_sums(gpurt_ctx ctx, int mode) {
    if (!ctx) return set_error(GPURT_E_INVALID, "null context");
    if (mode != GPURT_SUMS_AUTO && mode != GPURT_SUMS_FIXED_ORDER && mode != GPURT_SUMS_ATOMIC)
        return set_error(GPURT_E_INVALID, "unknown vertex-sum mode");
    ctx->vertex_sums = mode;
    return GPURT_OK;
}

static int check_mesh(const float* verts, size_t vstride_bytes, size_t n_verts,
                      const uint32_t* tri_idx, size_t n_tris) {
    if (!verts || !tri_idx) return set_error(GPURT_E_INVALID, "null vertex or index pointer");
    if (vstride_bytes < 12 || vstride_bytes % 4 != 0)
        return set_error(GPURT_E_INVALID, "vstride_bytes must be a multiple of 4 and >= 12");
    if (n_verts == 0 || n_verts > 0xFFFFFFFFull)
        return set_error(GPURT_E_INVALID, "n_verts out of range");
    if (n_tris == 0 || n_tris > (size_t)kMaxLeafFirst - 8)
        return set_error(GPURT_E_INVALID, "n_tris out of range (1 .. 2^28-9)");
    return GPURT_OK;
}

int gpurt_bvh_build(gpurt_ctx ctx, const float* verts, size_t vstride_bytes, size_t n_verts,
                    const uint32_t* tri_idx, size_t n_tris, const gpurt_build_opts* opts,
                    void* stream, gpurt_bvh* out) {
    if (!ctx || !out) return set_error(GPURT_E_INVALID, "null context or output handle");
    *out = nullptr;
    int rc = check_mesh(verts, vstride_bytes, n_verts, tri_idx, n_tris);
    if (rc) return rc;
    BuildParams bp;
    if ((rc = parse_build_opts(opts, bp))) return rc;
    DeviceGuard g(ctx->device);
    if (!g.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    hipStream_t s = (hipStream_t)stream;

    SegDesc sd{};
    sd.verts = verts;
    sd.vstride = vstride_bytes / 4;
    sd.n_verts = (uint32_t)n_verts;
    sd.n_tris = (uint32_t)n_tris;
    sd.idx = tri_idx;
    sd.tri_offset = 0;
    auto* b = new gpurt_bvh_s();
    b->ctx = ctx;
    b->n_verts = (int64_t)n_verts;
    // on failure the buffers made so far are freed in order on the build stream
    auto fail = [&](int code) {
        free_bvh_buffers(b, s);
        delete b;
        return code;
    };
    rc = build_segments(ctx, &sd, 1, bp.width, bp.leaf_size, bp.treelet_size, bp.treelet_rounds, bp.format, s, b);
    if (rc) return fail(rc);
    // keep the index buffer and compact positions for the backward passes and refit
    if (hipMallocAsync((void**)&b->tri_idx, sizeof(uint32_t) * 3 * n_tris, s) != hipSuccess ||
        hipMallocAsync((void**)&b->verts, sizeof(float) * 3 * n_verts, s) != hipSuccess) {
        (void)hipGetLastError();
        return fail(set_error(GPURT_E_OOM, "allocation of the mesh copy failed"));
    }
    if (hipMemcpyAsync(b->tri_idx, tri_idx, sizeof(uint32_t) * 3 * n_tris, hipMemcpyDeviceToDevice, s) !=
        hipSuccess)
        return fail(hip_error(hipGetLastError(), "copying indices"));
    if ((rc = compact_verts(verts, vstride_bytes / 4, n_verts, b->verts, s)) != GPURT_OK) return fail(rc);
    // the handle is complete when this returns, so queries on any stream may follow
    if (hipStreamSynchronize(s) != hipSuccess) return fail(hip_error(hipGetLastError(), "build"));
    {
        std::lock_guard<std::mutex> lk(ctx->live_mu);
        ctx->bvhs.insert(b);
    }
    *out = b;
    return GPURT_OK;
}

int gpurt_bvh_refit(gpurt_bvh bvh, const float* verts, size_t vstride_bytes, void* stream) {
    if (!bvh || !verts) return set_error(GPURT_E_INVALID, "null argument");
    if (vstride_bytes < 12 || vstride_bytes % 4 != 0)
        return set_error(GPURT_E_INVALID, "vstride_bytes must be a multiple of 4 and >= 12");
    DeviceGuard g(bvh->ctx->device);
    if (!g.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    // a refit rewrites the tree: it waits for earlier queries on other streams
    wait_uses(bvh, (hipStream_t)stream);
    int rc = launch_refit(bvh, verts, vstride_bytes / 4, (hipStream_t)stream);
    if (rc) return rc;
    GPURT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return GPURT_OK;
}

void gpurt_bvh_destroy(gpurt_bvh b) {
    if (!b) return;
    if (!b->ctx) {  // its context was destroyed first and released its device memory
        delete b;
        return;
    }
    {
        std::lock_guard<std::mutex> lk(b->ctx->live_mu);
        b->ctx->bvhs.erase(b);
    }
    DeviceGuard g(b->ctx->device);
    hipStream_t rs = b->ctx->release;
    // uses that no event marks (graph capture, a failed record): wait for the device
    if (b->uses_lost) (void)hipDeviceSynchronize();
    wait_uses(b, rs);
    free_bvh_buffers(b, rs);
    for (auto& u : b->uses) (void)hipEventDestroy(u.second);
    (void)hipGetLastError();
    delete b;
}

int gpurt_bvh_get_info(gpurt_bvh b, gpurt_bvh_info* out) {
    if (!b || !out) return set_error(GPURT_E_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->n_tris = b->n_tris;
    out->n_verts = b->n_verts;
    out->n_nodes = b->n_nodes;
    out->node_bytes = (int64_t)(rays_use_q8(b->format) ? qnode_bytes(b->width) : node_bytes(b->width));
    out->cpq_node_bytes = (int32_t)(cpq_use_q8(b->format) ? qnode_bytes(b->width) : node_bytes(b->width));
    out->tri_bytes = 48;
    out->depth = b->depth;
    out->node_width = b->width;
    out->leaf_size = b->leaf_size;
    out->sah_cost = b->sah_cost;
    out->node_format = b->format;
    return GPURT_OK;
}

int gpurt_bvh_export(gpurt_bvh b, const gpurt_bvh_arrays* out, void* stream) {
    if (!b || !out) return set_error(GPURT_E_INVALID, "null argument");
    if (out->qnodes && !b->qnodes) return set_error(GPURT_E_INVALID, "this BVH keeps no 8-bit nodes (node format 1)");
    DeviceGuard g(b->ctx->device);
    if (!g.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    hipStream_t s = (hipStream_t)stream;
    const size_t nn = (size_t)b->n_nodes, nt = (size_t)b->n_tris;
    if (out->nodes) GPURT_HIP(hipMemcpyAsync(out->nodes, b->nodes, node_bytes(b->width) * nn, hipMemcpyDeviceToDevice, s));
    if (out->qnodes)
        GPURT_HIP(hipMemcpyAsync(out->qnodes, b->qnodes, qnode_bytes(b->width) * nn, hipMemcpyDeviceToDevice, s));
    if (out->tris) GPURT_HIP(hipMemcpyAsync(out->tris, b->tris, sizeof(float4) * 3 * nt, hipMemcpyDeviceToDevice, s));
    if (out->leaf_order)
        GPURT_HIP(hipMemcpyAsync(out->leaf_order, b->leaf_order, sizeof(uint32_t) * nt, hipMemcpyDeviceToDevice, s));
    note_use(b, s);
    return GPURT_OK;
}

static int check_opts(const gpurt_query_opts* o) {
    if (!o) return GPURT_OK;
    if (o->stack_depth != 0 && o->stack_depth != 8 && o->stack_depth != 16 && o->stack_depth != 32 &&
        o->stack_depth != 64)
        return set_error(GPURT_E_UNSUPPORTED, "stack_depth must be 0, 8, 16, 32 or 64");
    if (o->waves_per_cu < 0 || o->waves_per_cu > 64)
        return set_error(GPURT_E_UNSUPPORTED, "waves_per_cu must be 0..64");
    if (o->leaf_trigger < 0 || o->leaf_trigger > 64)
        return set_error(GPURT_E_UNSUPPORTED, "leaf_trigger must be 0..64");
    if (o->refill_trigger < 0 || o->refill_trigger > 64)
        return set_error(GPURT_E_UNSUPPORTED, "refill_trigger must be 0..64");
    return GPURT_OK;
}

int gpurt_closest_point(gpurt_bvh b, const float* q, size_t n, int32_t* prim, float* pt,
                        float* dist, float* bary, const gpurt_query_opts* opts, void* stream) {
    if (!b) return set_error(GPURT_E_INVALID, "null bvh");
    if (n == 0) return GPURT_OK;
    if (!q || !prim || !pt || !dist) return set_error(GPURT_E_INVALID, "null query/output pointer");
    int rc = check_opts(opts);
    if (rc) return rc;
    DeviceGuard g(b->ctx->device);
    if (!g.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    const int rc2 = launch_cpq(b, q, n, prim, pt, dist, bary, opts, (hipStream_t)stream);
    note_use(b, (hipStream_t)stream);
    return rc2;
}

int gpurt_trace_closest(gpurt_bvh b, const gpurt_ray* rays, size_t n, int32_t* prim, float* t,
                        float* uv, const gpurt_query_opts* opts, void* stream) {
    if (!b) return set_error(GPURT_E_INVALID, "null bvh");
    if (n == 0) return GPURT_OK;
    if (!rays || !prim || !t) return set_error(GPURT_E_INVALID, "null ray/output pointer");
    int rc = check_opts(opts);
    if (rc) return rc;
    DeviceGuard g(b->ctx->device);
    if (!g.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    const int rc2 = launch_ray(b, rays, n, prim, t, uv, nullptr, false, opts, (hipStream_t)stream);
    note_use(b, (hipStream_t)stream);
    return rc2;
}

int gpurt_trace_closest_noalloc(gpurt_bvh b, const gpurt_ray* rays, size_t n, int32_t* prim, float* t,
                                float* uv, const gpurt_query_opts* opts, void* stream) {
    if (!b) return set_error(GPURT_E_INVALID, "null bvh");
    if (n == 0) return GPURT_OK;
    if (!rays || !prim || !t) return set_error(GPURT_E_INVALID, "null ray/output pointer");
    int rc = check_opts(opts);
    if (rc) return rc;
    if (opts && opts->stats) return set_error(GPURT_E_UNSUPPORTED, "stats are not counted by the no-allocation variant");
    DeviceGuard g(b->ctx->device);
    if (!g.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    const int rc2 = launch_ray_noalloc(b, rays, n, prim, t, uv, opts, (hipStream_t)stream);
    note_use_existing(b, (hipStream_t)stream);
    return rc2;
}

int gpurt_trace_any(gpurt_bvh b, const gpurt_ray* rays, size_t n, uint8_t* occluded,
                    const gpurt_query_opts* opts, void* stream) {
    if (!b) return set_error(GPURT_E_INVALID, "null bvh");
    if (n == 0) return GPURT_OK;
    if (!rays || !occluded) return set_error(GPURT_E_INVALID, "null ray/output pointer");
    int rc = check_opts(opts);
    if (rc) return rc;
    DeviceGuard g(b->ctx->device);
    if (!g.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    const int rc2 = launch_ray(b, rays, n, nullptr, nullptr, nullptr, occluded, true, opts, (hipStream_t)stream);
    note_use(b, (hipStream_t)stream);
    return rc2;
}

int gpurt_closest_point_bwd(gpurt_bvh b, const float* q, size_t n, const int32_t* prim,
                            const float* bary, const float* pt, const float* dist,
                            const float* g_dist, float* g_q, float* g_verts, void* stream) {
    if (!b) return set_error(GPURT_E_INVALID, "null bvh");
    if (n == 0) return GPURT_OK;
    (void)bary, (void)pt, (void)dist;  // recomputed in FP64 from the saved triangle
    if (!q || !prim || !g_dist) return set_error(GPURT_E_INVALID, "null input pointer");
    if (n > 0xFFFFFFFFull) return set_error(GPURT_E_INVALID, "too many queries in one call");
    DeviceGuard g(b->ctx->device);
    if (!g.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    const int rc2 = launch_cpq_bwd(b, q, n, prim, g_dist, g_q, g_verts, (hipStream_t)stream);
    note_use(b, (hipStream_t)stream);
    return rc2;
}

int gpurt_trace_bwd(gpurt_bvh b, const gpurt_ray* rays, size_t n, const int32_t* prim,
                    const float* t, const float* uv, const float* g_t, float* g_origin,
                    float* g_dir, float* g_verts, void* stream) {
    if (!b) return set_error(GPURT_E_INVALID, "null bvh");
    if (n == 0) return GPURT_OK;
    if (!rays || !prim || !t || !g_t) return set_error(GPURT_E_INVALID, "null input pointer");
    if (g_verts && !uv) return set_error(GPURT_E_INVALID, "g_verts needs the saved (u,v)");
    if (n > 0xFFFFFFFFull) return set_error(GPURT_E_INVALID, "too many rays in one call");
    DeviceGuard g(b->ctx->device);
    if (!g.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    const int rc2 = launch_ray_bwd(b, rays, n, prim, t, uv, g_t, g_origin, g_dir, g_verts, (hipStream_t)stream);
    note_use(b, (hipStream_t)stream);
    return rc2;
}

}  // extern "C"
