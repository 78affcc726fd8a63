// batch.hip -- many meshes built and queried together.
//
// The builder is segmented by construction (the mesh id sits in the high bits of every
// Morton key, so each mesh is its own subtree of one radix tree), so the BVHs of all
// meshes are built by the same ~30 launches as one mesh.  Queries then run in two
// persistent launches (closest points, then rays) of the streaming traversal: a wave
// takes 64-query chunks from device work queues and hands their queries to its lanes
// as lanes finish, until the chunk list is exhausted; a query starts at its mesh's root
// (wide node m of the shared node array).  Chunks are ordered by decreasing mesh size so
// the tail of a launch is made of cheap chunks.  The whole sequence -- build of every
// mesh plus both query launches -- is recorded once into a HIP graph and replayed by
// gpurt_batch_run.
#include <algorithm>
#include <numeric>
#include <vector>

#include "backward.h"
#include "traverse.h"

namespace gpurt_impl {

struct BatchMeshDev {
    const float* verts;  // the mesh (for the closest-point finishing pass)
    uint64_t vstride;    // floats
    const uint32_t* tri_idx;
    uint32_t n_tris, n_verts;
    const float* q;
    int32_t* q_prim;
    float* q_pt;
    float* q_dist;
    float* q_bary;
    const float4* rays;
    int32_t* r_prim;
    float* r_t;
    float* r_uv;
    uint32_t n_q;
    uint32_t n_r;
};

// Clears the queue counters of both query kinds (2 * kQueues, kQueueStride apart).
__global__ void k_reset_counters(uint32_t* c) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * kQueues) c[i * kQueueStride] = 0u;
}

// Query source over a chunk list (traverse.h): the chunks are dealt round-robin to
// nq = min(kQueues, blocks) queues with a device counter each (as the plain kernels'
// QueueSource), and block b's waves take the next chunk (mesh m, first query) of queue
// b % nq when their pool is empty, handing the chunk's queries to the lanes that want
// one, in lane order.  Called by all lanes together, so the pool stays wave-uniform.
struct ChunkSource {
    struct Handle {
        uint32_t m, i;
    };
    const BatchMeshDev* meshes;
    const uint2* chunks;
    uint32_t n_chunks;
    uint32_t* counters;
    bool rays;
    uint32_t pm = 0, pnext = 0, pend = 0;  // current chunk: mesh and [pnext, pend)
    bool more = true;
    __device__ __forceinline__ bool refill(bool has, Handle& h) {
        const uint32_t lane = threadIdx.x & 63u;
        const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
        const uint32_t nq = min(kQueues, gridDim.x);  // every queue has blocks serving it
        const uint32_t q = blockIdx.x % nq;
        bool got = false;
        uint64_t want = __ballot(!has);
        while (want && more) {
            if (pnext >= pend) {
                uint32_t c = 0;
                if (lane == 0) c = gatomic_add(counters + q * kQueueStride, 1u);
                c = __builtin_amdgcn_readfirstlane(c);
                const uint64_t j = (uint64_t)c * nq + q;
                if (j >= n_chunks) {
                    more = false;
                    break;
                }
                const uint2 ch = chunks[j];
                pm = ch.x;
                pnext = ch.y;
                const uint32_t nq = rays ? meshes[pm].n_r : meshes[pm].n_q;
                pend = min(ch.y + 64u, nq);
            }
            const uint32_t k = min(pend - pnext, (uint32_t)__popcll(want));
            const uint32_t rank = (uint32_t)__popcll(want & below);
            if (!has && !got && rank < k) {
                h.m = pm;
                h.i = pnext + rank;
                got = true;
            }
            pnext += k;
            want = __ballot(!has && !got);
        }
        return got;
    }
    // wave-uniform: no lane of this wave will get another query
    __device__ __forceinline__ bool exhausted() const { return !more; }
    __device__ __forceinline__ static Handle shfl(const Handle& h, uint32_t from) {
        return Handle{(uint32_t)__shfl(h.m, (int)from), (uint32_t)__shfl(h.i, (int)from)};
    }
    __device__ __forceinline__ uint32_t root(const Handle& h) const { return h.m; }
    __device__ __forceinline__ void ray(const Handle& h, float4& r0, float4& r1) const {
        const float4* r = meshes[h.m].rays;
        r0 = r[2 * (size_t)h.i];
        r1 = r[2 * (size_t)h.i + 1];
    }
    __device__ __forceinline__ V3 point(const Handle& h) const {
        const float* q = meshes[h.m].q;
        return v3(q[3 * (size_t)h.i], q[3 * (size_t)h.i + 1], q[3 * (size_t)h.i + 2]);
    }
};

struct BatchArgs {
    const void* nodes;
    const float4* tris;
    const BatchMeshDev* meshes;
    const uint2* chunks;
    uint32_t n_chunks;
    uint32_t* counter;
    uint2* spill;
    uint32_t spill_slots;
    uint32_t* err;
    uint32_t leaf_trigger;
};

template <int W, bool Q, int S, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WavesHint<W, Q>::value))) void k_batch_cpq(
    BatchArgs a) {
    const NodeOf<W, Q>* __restrict__ nodes = reinterpret_cast<const NodeOf<W, Q>*>(a.nodes);
    LaneStack<S, BS> st;
    st.init(a.spill, a.spill_slots, blockIdx.x * BS + threadIdx.x, gridDim.x * BS);
    ChunkSource src{a.meshes, a.chunks, a.n_chunks, a.counter, false};
    auto out = [&](const ChunkSource::Handle& h, uint32_t id) {
        a.meshes[h.m].q_prim[h.i] = id == 0xFFFFFFFFu ? -1 : (int32_t)id;
    };
    unsigned long long dn = 0, dt = 0;
    if (!cpq_stream<NodeOf<W, Q>, S, BS, false>(nodes, a.tris, src, a.leaf_trigger, st, out, dn, dt)) gatomic_or(a.err, 1u);
}

template <int W, bool Q, int S, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WavesHint<W, Q>::value))) void k_batch_ray(
    BatchArgs a) {
    const NodeOf<W, Q>* __restrict__ nodes = reinterpret_cast<const NodeOf<W, Q>*>(a.nodes);
    LaneStack<S, BS> st;
    st.init(a.spill, a.spill_slots, blockIdx.x * BS + threadIdx.x, gridDim.x * BS);
    ChunkSource src{a.meshes, a.chunks, a.n_chunks, a.counter, true};
    auto out = [&](const ChunkSource::Handle& h, uint32_t id, float t, float u, float v, bool) {
        const BatchMeshDev& md = a.meshes[h.m];
        const size_t i = h.i;
        const bool hit = id != 0xFFFFFFFFu;
        md.r_prim[i] = hit ? (int32_t)id : -1;
        md.r_t[i] = hit ? t : INFINITY;
        if (md.r_uv) {
            md.r_uv[2 * i] = hit ? u : 0.0f;
            md.r_uv[2 * i + 1] = hit ? v : 0.0f;
        }
    };
    unsigned long long dn = 0, dt = 0;
    if (!ray_stream<NodeOf<W, Q>, S, BS, false, false>(nodes, a.tris, src, a.leaf_trigger, st, out, dn, dt)) gatomic_or(a.err, 1u);
}

template <int W, bool Q, int S, int BS>
static void launch_batch_kind(bool rays, unsigned grid, hipStream_t s, const BatchArgs& a) {
    if (rays)
        hipLaunchKernelGGL((k_batch_ray<W, Q, S, BS>), dim3(grid), dim3(BS), (S * sizeof(uint2) + kGroupKeyBytes) * BS, s, a);
    else
        hipLaunchKernelGGL((k_batch_cpq<W, Q, S, BS>), dim3(grid), dim3(BS), (S * sizeof(uint2) + kGroupKeyBytes) * BS, s, a);
}

template <int W, bool Q>
static int launch_batch_w(int S, bool rays, unsigned grid, hipStream_t s, const BatchArgs& a) {
    switch (S) {
        case 8: launch_batch_kind<W, Q, 8, 256>(rays, grid, s, a); return 256;
        case 16: launch_batch_kind<W, Q, 16, 256>(rays, grid, s, a); return 256;
        case 32: launch_batch_kind<W, Q, 32, 128>(rays, grid, s, a); return 128;
        case 64: launch_batch_kind<W, Q, 64, 64>(rays, grid, s, a); return 64;
    }
    return 0;
}

static int launch_batch(int W, bool q8, int S, bool rays, unsigned grid, hipStream_t s, const BatchArgs& a) {
    switch (W * 2 + (q8 ? 1 : 0)) {
        case 4: return launch_batch_w<2, false>(S, rays, grid, s, a);
        case 5: return launch_batch_w<2, true>(S, rays, grid, s, a);
        case 8: return launch_batch_w<4, false>(S, rays, grid, s, a);
        case 9: return launch_batch_w<4, true>(S, rays, grid, s, a);
        case 16: return launch_batch_w<8, false>(S, rays, grid, s, a);
        case 17: return launch_batch_w<8, true>(S, rays, grid, s, a);
    }
    return 0;
}

static int block_for_stack_b(int S) { return S <= 16 ? 256 : (S == 32 ? 128 : 64); }

// ------------------------------------------------------------------------------------
// batched backward: gradients of every mesh's distances and hit distances, from the
// forward outputs the batch wrote (ids, barycentrics, points, distances, hit t, (u,v))
// ------------------------------------------------------------------------------------

struct BatchGradDev {
    const float* verts;
    uint64_t vstride;  // floats
    const uint32_t* tri_idx;
    uint32_t n_tris, n_verts;
    const float* g_dist;
    float* g_q;
    const float* g_t;
    float* g_origin;
    float* g_dir;
    float* g_verts;
    uint64_t voff;  // the mesh's first vertex in the batch-wide numbering of the fixed-order sums
};

constexpr uint32_t kZeroSeg = 4096;  // floats cleared per block of the zeroing kernel

// Clears every requested vertex gradient: segment k = (mesh, first float) covers up to
// kZeroSeg floats of that mesh's n_verts x 3 array.
__global__ __launch_bounds__(256) void k_batch_zero_vgrads(const BatchGradDev* __restrict__ grads,
                                                           const uint2* __restrict__ segs) {
    const uint2 sg = segs[blockIdx.x];
    const BatchGradDev& gd = grads[sg.x];
    const uint32_t end = min(sg.y + kZeroSeg, 3u * gd.n_verts);
    for (uint32_t i = sg.y + threadIdx.x; i < end; i += 256) gd.g_verts[i] = 0.0f;
}

// Closest-point finishing pass over the forward chunk list (one wave per 64-query
// chunk): the winners' point, distance and barycentrics in FP64.
__global__ __launch_bounds__(256) void k_batch_cpq_finish(const BatchMeshDev* __restrict__ meshes,
                                                          const uint2* __restrict__ chunks, uint32_t n_chunks) {
    const uint32_t c = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (c >= n_chunks) return;
    const uint2 ch = chunks[c];
    const BatchMeshDev& md = meshes[ch.x];
    const uint32_t i = ch.y + (threadIdx.x & 63u);
    if (i >= md.n_q) return;
    cpq_finish_one(i, md.q, md.q_prim, md.tri_idx, md.n_tris, md.verts, md.vstride, md.n_verts, md.q_pt, md.q_dist,
                   md.q_bary);
}

// Batched backward, one wave per 64-item chunk of the forward chunk list (closest-point
// chunks, then ray chunks).  Query and ray gradients are written here.  Vertex terms are
// scatter-added with FP32 atomics when `keys` is null; otherwise item (chunk c, lane l)
// writes them to term slots 3 (64 c + l) .. + 2, keyed by the vertex's batch-wide index
// (the mesh's voff + its vertex), or `none` when it adds nothing (no such item, no vertex
// gradient requested), for the fixed-order sum.
template <bool RAYS>
__global__ __launch_bounds__(256) void k_batch_bwd(const BatchMeshDev* __restrict__ meshes,
                                                   const BatchGradDev* __restrict__ grads,
                                                   const uint2* __restrict__ chunks, uint32_t n_chunks,
                                                   uint32_t first_chunk, uint64_t none,
                                                   uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                   float* __restrict__ terms) {
    const uint32_t c = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (c >= n_chunks) return;
    const uint2 ch = chunks[c];
    const BatchMeshDev& md = meshes[ch.x];
    const BatchGradDev& gd = grads[ch.x];
    const uint32_t i = ch.y + (threadIdx.x & 63u);
    CornerTerms t{{kNoVertex, kNoVertex, kNoVertex}, {}};
    if (RAYS) {
        if (i < md.n_r && gd.g_t)
            t = ray_bwd_terms(i, md.rays, md.r_prim, md.r_t, md.r_uv, gd.g_t, gd.tri_idx, gd.n_tris, gd.verts,
                              gd.vstride, gd.g_origin, gd.g_dir);
    } else {
        if (i < md.n_q && gd.g_dist)
            t = cpq_bwd_terms(i, md.q, md.q_prim, gd.g_dist, gd.tri_idx, gd.n_tris, gd.verts, gd.vstride, gd.g_q);
    }
    if (!keys) {
        scatter_terms(t, gd.g_verts);
        return;
    }
    const size_t k0 = 3 * (64 * ((size_t)first_chunk + c) + (threadIdx.x & 63u));
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const size_t k = k0 + j;
        keys[k] = (t.v[j] == kNoVertex || !gd.g_verts) ? none : gd.voff + t.v[j];
        vals[k] = (uint32_t)k;
        terms[3 * k] = t.c[j][0];
        terms[3 * k + 1] = t.c[j][1];
        terms[3 * k + 2] = t.c[j][2];
    }
}

// adds a batch-wide vertex's total to its mesh's vertex gradient (cleared by
// k_batch_zero_vgrads; each key has one run).  voff[0..n_meshes] are the meshes' first
// batch-wide vertices (a mesh without a vertex gradient adds none), so the mesh is the
// last m with voff[m] <= key.
struct MeshVertexSink {
    const BatchGradDev* grads;
    const uint64_t* voff;
    uint32_t n_meshes;
    __device__ void operator()(uint64_t key, float x, float y, float z) const {
        uint32_t lo = 0, hi = n_meshes;  // voff[lo] <= key < voff[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (voff[mid] <= key) lo = mid;
            else hi = mid;
        }
        float* g = grads[lo].g_verts;
        const size_t v = key - voff[lo];
        g[3 * v] += x;
        g[3 * v + 1] += y;
        g[3 * v + 2] += z;
    }
};

}  // namespace gpurt_impl

using namespace gpurt_impl;

struct gpurt_batch_s {
    gpurt_ctx ctx = nullptr;
    int device = 0;  // the context's device (destroy does not read the context)
    int width = 4, leaf_size = 4, stack = 8, block = 256, format = kFormatF32;
    uint32_t leaf_trigger = 32;
    BuildBufs build;
    void* qnodes = nullptr;  // QNode<W>[build.capacity] when has_qnodes(format)
    BatchMeshDev* meshes_d = nullptr;
    uint2* chunks_d = nullptr;    // closest-point chunks, then ray chunks
    uint32_t n_cpq_chunks = 0, n_ray_chunks = 0;
    uint32_t* counter_d = nullptr;  // kQueues closest-point queue counters, then kQueues for rays
    uint2* spill = nullptr;
    uint32_t spill_slots = 0;
    unsigned grid = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // backward (gpurt_batch_set_grads): its own recorded graph
    std::vector<gpurt_batch_mesh> host_meshes;
    BatchGradDev* grads_d = nullptr;
    uint2* zero_segs_d = nullptr;
    void* terms_mem = nullptr;  // vertex terms of the backward and their sort buffers
    hipGraph_t bgraph = nullptr;
    hipGraphExec_t bexec = nullptr;
    void drop_backward() {
        if (bexec) (void)hipGraphExecDestroy(bexec);
        if (bgraph) (void)hipGraphDestroy(bgraph);
        if (grads_d) (void)hipFree(grads_d);
        if (zero_segs_d) (void)hipFree(zero_segs_d);
        if (terms_mem) (void)hipFree(terms_mem);
        terms_mem = nullptr;
        bexec = nullptr;
        bgraph = nullptr;
        grads_d = nullptr;
        zero_segs_d = nullptr;
    }
    ~gpurt_batch_s() {
        drop_backward();
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        if (meshes_d) (void)hipFree(meshes_d);
        if (chunks_d) (void)hipFree(chunks_d);
        if (counter_d) (void)hipFree(counter_d);
        if (spill) (void)hipFree(spill);
        if (qnodes) (void)hipFree(qnodes);
    }
};

namespace {

struct DevGuard {
    int prev = -1;
    bool ok = false;
    explicit DevGuard(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        ok = hipSetDevice(d) == hipSuccess;
    }
    ~DevGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

int enqueue_batch(gpurt_batch_s* b, int treelet_size, int treelet_rounds, hipStream_t s) {
    int rc;
    if ((rc = build_enqueue_binary(b->build, treelet_size, treelet_rounds, s))) return rc;
    // all levels: empty levels exit at once, and the tree cannot be deeper (kMaxLevels)
    if ((rc = build_enqueue_levels(b->build, b->leaf_size, 0, kMaxLevels, s))) return rc;
    if (has_qnodes(b->format)) {
        // counters[0] holds the wide-node count the levels produced
        launch_quantize(b->width, b->build.nodes, b->qnodes, 0, b->build.capacity, b->build.counters, s);
    }
    hipLaunchKernelGGL(k_reset_counters, dim3(cdiv(2 * kQueues, 256)), dim3(256), 0, s, b->counter_d);
    BatchArgs a{};
    a.tris = b->build.final_tris();
    a.meshes = b->meshes_d;
    a.spill = b->spill;
    a.spill_slots = b->spill_slots;
    a.err = b->ctx->err;
    a.leaf_trigger = b->leaf_trigger;
    for (int rays = 0; rays < 2; rays++) {
        const uint32_t n = rays ? b->n_ray_chunks : b->n_cpq_chunks;
        if (n == 0) continue;
        const bool q8 = rays ? rays_use_q8(b->format) : cpq_use_q8(b->format);
        a.nodes = q8 ? b->qnodes : b->build.nodes;
        a.chunks = b->chunks_d + (rays ? b->n_cpq_chunks : 0u);
        a.n_chunks = n;
        a.counter = b->counter_d + (size_t)rays * kQueues * kQueueStride;
        const unsigned grid = std::max<size_t>(1, std::min<size_t>(b->grid, (size_t)n * 64 / b->block + 1));
        if (!launch_batch(b->width, q8, b->stack, rays != 0, grid, s, a))
            return set_error(GPURT_E_UNSUPPORTED, "batch stack/width");
        if (!rays)
            hipLaunchKernelGGL(k_batch_cpq_finish, dim3(cdiv(n, 4)), dim3(256), 0, s, b->meshes_d, b->chunks_d, n);
    }
    GPURT_HIP(hipGetLastError());
    return GPURT_OK;
}

}  // namespace

extern "C" {

int gpurt_batch_create(gpurt_ctx ctx, const gpurt_batch_mesh* meshes, size_t n_meshes,
                       const gpurt_build_opts* bopts, const gpurt_query_opts* qopts, gpurt_batch* out) {
    if (!out) return set_error(GPURT_E_INVALID, "out is null");
    *out = nullptr;
    if (!ctx || !meshes) return set_error(GPURT_E_INVALID, "null context or mesh array");
    if (n_meshes == 0 || n_meshes > (1u << 20)) return set_error(GPURT_E_INVALID, "n_meshes must be 1..2^20");
    BuildParams bp;
    int rc0 = parse_build_opts(bopts, bp);
    if (rc0) return rc0;
    const int W = bp.width, leaf = bp.leaf_size, tsize = bp.treelet_size, trounds = bp.treelet_rounds;
    // 16 stack entries by default: the closest-point searches that dominate a batch spill
    // with 8, and 16 waves per CU (the default below) fit 16-entry stacks in LDS (256-mesh
    // batch: 26.8 -> 25.5 ms per replay; the autotuner's pick on that batch)
    int S = (qopts && qopts->stack_depth) ? qopts->stack_depth : 16;
    if (S != 8 && S != 16 && S != 32 && S != 64)
        return set_error(GPURT_E_UNSUPPORTED, "stack_depth must be 0, 8, 16, 32 or 64");
    if (qopts && (qopts->waves_per_cu < 0 || qopts->waves_per_cu > 64))
        return set_error(GPURT_E_UNSUPPORTED, "waves_per_cu must be 0..64");
    if (qopts && (qopts->leaf_trigger < 0 || qopts->leaf_trigger > 64))
        return set_error(GPURT_E_UNSUPPORTED, "leaf_trigger must be 0..64");
    if (qopts && (qopts->refill_trigger < 0 || qopts->refill_trigger > 64))
        return set_error(GPURT_E_UNSUPPORTED, "refill_trigger must be 0..64");

    std::vector<SegDesc> segs(n_meshes);
    std::vector<BatchMeshDev> md(n_meshes);
    uint64_t off = 0;
    for (size_t m = 0; m < n_meshes; m++) {
        const gpurt_batch_mesh& g = meshes[m];
        const std::string at = " (mesh " + std::to_string(m) + ")";
        if (!g.verts || !g.tri_idx) return set_error(GPURT_E_INVALID, "null vertex or index pointer" + at);
        if (g.vstride_bytes < 12 || g.vstride_bytes % 4)
            return set_error(GPURT_E_INVALID, "vstride_bytes must be a multiple of 4 and >= 12" + at);
        if (g.n_verts == 0 || g.n_verts > 0xFFFFFFFFull || g.n_tris == 0)
            return set_error(GPURT_E_INVALID, "empty or oversized mesh" + at);
        if (g.n_q && (!g.q || !g.q_prim || !g.q_pt || !g.q_dist))
            return set_error(GPURT_E_INVALID, "null closest-point query/output pointer" + at);
        if (g.n_r && (!g.rays || !g.r_prim || !g.r_t))
            return set_error(GPURT_E_INVALID, "null ray/output pointer" + at);
        if (g.n_q > 0xFFFFFF00ull || g.n_r > 0xFFFFFF00ull)
            return set_error(GPURT_E_INVALID, "too many queries for one mesh" + at);
        segs[m].verts = g.verts;
        segs[m].vstride = g.vstride_bytes / 4;
        segs[m].n_verts = (uint32_t)g.n_verts;
        segs[m].n_tris = (uint32_t)g.n_tris;
        segs[m].idx = g.tri_idx;
        segs[m].tri_offset = (uint32_t)std::min<uint64_t>(off, 0xFFFFFFFFull);
        off += g.n_tris;
        md[m] = BatchMeshDev{g.verts, g.vstride_bytes / 4, g.tri_idx, (uint32_t)g.n_tris, (uint32_t)g.n_verts,
                             g.q, g.q_prim, g.q_pt, g.q_dist, g.q_bary,
                             reinterpret_cast<const float4*>(g.rays), g.r_prim, g.r_t, g.r_uv,
                             (uint32_t)g.n_q, (uint32_t)g.n_r};
    }
    if (off > kMaxLeafFirst - 8) return set_error(GPURT_E_INVALID, "total triangle count exceeds 2^28-9");

    DevGuard dg(ctx->device);
    if (!dg.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    auto* b = new gpurt_batch_s();
    b->ctx = ctx;
    b->device = ctx->device;
    b->width = W;
    b->leaf_size = leaf;
    b->stack = S;
    b->format = bp.format;
    if (qopts && qopts->leaf_trigger) b->leaf_trigger = (uint32_t)qopts->leaf_trigger;
    // bits 8-15: the ray kernel's refill trigger (16 lanes by default, as for single meshes)
    b->leaf_trigger |= (qopts && qopts->refill_trigger ? (uint32_t)qopts->refill_trigger : 16u) << 8;
    b->host_meshes.assign(meshes, meshes + n_meshes);
    int rc = build_alloc(segs.data(), (int)n_meshes, W, leaf, tsize, b->build);
    if (rc) {
        delete b;
        return rc;
    }
    // work list: 64-query chunks, meshes with the most triangles first
    std::vector<uint32_t> order(n_meshes);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return segs[x].n_tris > segs[y].n_tris; });
    std::vector<uint2> chunks;
    for (uint32_t m : order)
        for (uint32_t st = 0; st < md[m].n_q; st += 64) chunks.push_back(make_uint2(m, st));
    b->n_cpq_chunks = (uint32_t)chunks.size();
    for (uint32_t m : order)
        for (uint32_t st = 0; st < md[m].n_r; st += 64) chunks.push_back(make_uint2(m, st));
    b->n_ray_chunks = (uint32_t)chunks.size() - b->n_cpq_chunks;
    const int BS = block_for_stack_b(S);
    b->block = BS;
    const int waves_per_block = BS / 64;
    // register-limited waves: the larger of the two kernels' needs (FP32 8-wide nodes
    // need the most registers)
    const bool any_f32 = !rays_use_q8(bp.format) || !cpq_use_q8(bp.format);
    const int hint = W == 8 ? (any_f32 ? WavesHint<8, false>::value : WavesHint<8, true>::value)
                            : WavesHint<4, false>::value;
    int blocks_per_cu = std::min(ctx->lds_per_cu / ((S * 8 + kGroupKeyBytes) * BS), std::min(32, 4 * hint) / waves_per_block);
    // default 4 waves/SIMD, as the single-mesh closest-point kernels, which dominate a
    // batch (256-mesh batch: 36.1 -> 34.9 ms per replay against every slot registers allow)
    constexpr int kBatchDefaultWaves = 16;
    if (qopts && qopts->waves_per_cu > 0)
        blocks_per_cu = std::max(1, qopts->waves_per_cu / waves_per_block);
    else
        blocks_per_cu = std::min(blocks_per_cu, std::max(1, kBatchDefaultWaves / waves_per_block));
    blocks_per_cu = std::max(1, blocks_per_cu);
    b->grid = (unsigned)std::max<size_t>(1, (size_t)ctx->n_cu * blocks_per_cu);
    // deepest possible wide tree (kMaxLevels levels) bounds the stack: no overflow
    const int need = (W - 1) * kMaxLevels + 1;
    b->spill_slots = need > S ? (uint32_t)(need - S) : 0u;
    bool fail = hipMalloc((void**)&b->meshes_d, sizeof(BatchMeshDev) * n_meshes) != hipSuccess ||
                hipMalloc((void**)&b->chunks_d, sizeof(uint2) * std::max<size_t>(1, chunks.size())) != hipSuccess ||
                hipMalloc((void**)&b->counter_d, sizeof(uint32_t) * 2 * kQueues * kQueueStride) != hipSuccess ||
                (has_qnodes(bp.format) && hipMalloc(&b->qnodes, qnode_bytes(W) * std::max<uint32_t>(1u, b->build.capacity)) != hipSuccess) ||
                (b->spill_slots &&
                 hipMalloc((void**)&b->spill, (size_t)b->spill_slots * b->grid * BS * sizeof(uint2)) != hipSuccess);
    if (fail) {
        (void)hipGetLastError();
        delete b;
        return set_error(GPURT_E_OOM, "batch allocation failed");
    }
    if (hipMemcpy(b->meshes_d, md.data(), sizeof(BatchMeshDev) * n_meshes, hipMemcpyHostToDevice) != hipSuccess ||
        (!chunks.empty() &&
         hipMemcpy(b->chunks_d, chunks.data(), sizeof(uint2) * chunks.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess) {
        rc = hip_error(hipGetLastError(), "batch setup");
        delete b;
        return rc;
    }
    // record build + queries into a graph
    if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        rc = hip_error(hipGetLastError(), "hipStreamBeginCapture");
        delete b;
        return rc;
    }
    rc = enqueue_batch(b, tsize, trounds, b->cap);
    hipGraph_t g = nullptr;
    hipError_t ec = hipStreamEndCapture(b->cap, &g);
    b->graph = g;
    if (rc == GPURT_OK && ec != hipSuccess) rc = hip_error(ec, "hipStreamEndCapture");
    if (rc == GPURT_OK && hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
        rc = hip_error(hipGetLastError(), "hipGraphInstantiate");
    if (rc) {
        delete b;
        return rc;
    }
    // one eager run validates the inputs (vertex indices, finite positions) before the
    // handle is returned
    uint32_t h_counters[4] = {0, 0, 0, 0};
    uint32_t h_last = 0;
    if (hipGraphLaunch(b->exec, b->cap) != hipSuccess ||
        hipMemcpyAsync(h_counters, b->build.counters, sizeof(h_counters), hipMemcpyDeviceToHost, b->cap) != hipSuccess ||
        hipMemcpyAsync(&h_last, b->build.lvl + kMaxLevels, sizeof(uint32_t), hipMemcpyDeviceToHost, b->cap) != hipSuccess ||
        hipStreamSynchronize(b->cap) != hipSuccess) {
        rc = hip_error(hipGetLastError(), "first batch run");
        delete b;
        return rc;
    }
    if (h_counters[1] & 1u) rc = set_error(GPURT_E_INVALID, "triangle index >= n_verts");
    else if (h_counters[1] & 4u) rc = set_error(GPURT_E_INVALID, "non-finite vertex position");
    else if (h_counters[1] & 2u) rc = set_error(GPURT_E_DEVICE, "wide-node capacity exceeded");
    else if (h_last) rc = set_error(GPURT_E_UNSUPPORTED, "BVH deeper than the supported level count");
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return GPURT_OK;
}

int gpurt_batch_run(gpurt_batch b, void* stream) {
    if (!b) return set_error(GPURT_E_INVALID, "null batch");
    DevGuard dg(b->ctx->device);
    if (!dg.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    GPURT_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return GPURT_OK;
}

int gpurt_batch_set_grads(gpurt_batch b, const gpurt_batch_grad* grads) {
    if (!b || !grads) return set_error(GPURT_E_INVALID, "null batch or gradient array");
    const size_t n_meshes = b->host_meshes.size();
    std::vector<BatchGradDev> gd(n_meshes);
    std::vector<uint2> segs;
    std::vector<uint64_t> voff(n_meshes + 1, 0);  // batch-wide vertex numbering
    for (size_t m = 0; m < n_meshes; m++) {
        const gpurt_batch_mesh& hm = b->host_meshes[m];
        const gpurt_batch_grad& g = grads[m];
        const std::string at = " (mesh " + std::to_string(m) + ")";
        if (g.g_t && hm.n_r && g.g_verts && !hm.r_uv)
            return set_error(GPURT_E_INVALID, "vertex gradients of hit t need the saved (u,v)" + at);
        if (g.g_verts && 3 * (uint64_t)hm.n_verts >= 0xFFFFFFFFull)
            return set_error(GPURT_E_INVALID, "too many vertices for a batched vertex gradient" + at);
        gd[m] = BatchGradDev{hm.verts, hm.vstride_bytes / 4, hm.tri_idx, (uint32_t)hm.n_tris, (uint32_t)hm.n_verts,
                             hm.n_q ? g.g_dist : nullptr, g.g_q, hm.n_r ? g.g_t : nullptr, g.g_origin, g.g_dir,
                             g.g_verts, voff[m]};
        voff[m + 1] = voff[m] + (g.g_verts ? (uint64_t)hm.n_verts : 0);
        if (g.g_verts)
            for (uint64_t f = 0; f < 3 * (uint64_t)hm.n_verts; f += kZeroSeg) segs.push_back(make_uint2((uint32_t)m, (uint32_t)f));
    }
    // vertex terms: 3 per item slot of every chunk, keyed by (mesh, vertex)
    // (for the fixed-order sums; otherwise they are added with atomics)
    const uint32_t n_chunks = b->n_cpq_chunks + b->n_ray_chunks;
    const bool fixed = !segs.empty() && fixed_order_sums(b->ctx, 3ull * 64 * n_chunks);
    const uint64_t m64 = fixed ? 3ull * 64 * n_chunks : 0;
    if (m64 > 0xFFFFFFFFull - kBlock) return set_error(GPURT_E_INVALID, "too many queries for a batched backward");
    const uint32_t m = (uint32_t)m64;
    // sort keys: batch-wide vertex indices, and their count itself for "no vertex"
    const uint64_t none = voff[n_meshes];
    int kbits = 1;
    while (kbits < 64 && (none >> kbits)) kbits++;
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t tbytes = 2 * up(8 * (size_t)m) + 2 * up(4 * (size_t)m) + 2 * up(12 * (size_t)m) +
                          up(4 * radix_sort_scratch_words(std::max<uint32_t>(m, 1))) + up(8 * (n_meshes + 1));
    DevGuard dg(b->ctx->device);
    if (!dg.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    b->drop_backward();
    if (hipMalloc((void**)&b->grads_d, sizeof(BatchGradDev) * n_meshes) != hipSuccess ||
        (!segs.empty() && hipMalloc((void**)&b->zero_segs_d, sizeof(uint2) * segs.size()) != hipSuccess) ||
        (m && hipMalloc(&b->terms_mem, tbytes) != hipSuccess)) {
        (void)hipGetLastError();
        b->drop_backward();
        return set_error(GPURT_E_OOM, "batch gradient allocation failed");
    }
    uint64_t *k0 = nullptr, *k1 = nullptr;
    uint32_t *v0 = nullptr, *v1 = nullptr, *sort_scratch = nullptr;
    float *terms = nullptr, *partial = nullptr;
    uint64_t* voff_d = nullptr;
    if (m) {
        char* tp = static_cast<char*>(b->terms_mem);
        const size_t m8 = up(8 * (size_t)m), m4 = up(4 * (size_t)m), m12 = up(12 * (size_t)m);
        k0 = reinterpret_cast<uint64_t*>(tp);
        k1 = reinterpret_cast<uint64_t*>(tp + m8);
        v0 = reinterpret_cast<uint32_t*>(tp + 2 * m8);
        v1 = reinterpret_cast<uint32_t*>(tp + 2 * m8 + m4);
        terms = reinterpret_cast<float*>(tp + 2 * m8 + 2 * m4);
        partial = reinterpret_cast<float*>(tp + 2 * m8 + 2 * m4 + m12);
        sort_scratch = reinterpret_cast<uint32_t*>(tp + 2 * m8 + 2 * m4 + 2 * m12);
        voff_d = reinterpret_cast<uint64_t*>(tp + 2 * m8 + 2 * m4 + 2 * m12 +
                                             up(4 * radix_sort_scratch_words(std::max<uint32_t>(m, 1))));
    }
    if (hipMemcpy(b->grads_d, gd.data(), sizeof(BatchGradDev) * n_meshes, hipMemcpyHostToDevice) != hipSuccess ||
        (!segs.empty() &&
         hipMemcpy(b->zero_segs_d, segs.data(), sizeof(uint2) * segs.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        (voff_d && hipMemcpy(voff_d, voff.data(), sizeof(uint64_t) * (n_meshes + 1), hipMemcpyHostToDevice) != hipSuccess)) {
        int rc = hip_error(hipGetLastError(), "batch gradient setup");
        b->drop_backward();
        return rc;
    }
    if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        int rc = hip_error(hipGetLastError(), "hipStreamBeginCapture");
        b->drop_backward();
        return rc;
    }
    if (!segs.empty())
        hipLaunchKernelGGL(k_batch_zero_vgrads, dim3((unsigned)segs.size()), dim3(256), 0, b->cap, b->grads_d,
                           b->zero_segs_d);
    if (b->n_cpq_chunks)
        hipLaunchKernelGGL(k_batch_bwd<false>, dim3(cdiv(b->n_cpq_chunks, 4)), dim3(256), 0, b->cap, b->meshes_d,
                           b->grads_d, b->chunks_d, b->n_cpq_chunks, 0u, none, k0, v0, terms);
    if (b->n_ray_chunks)
        hipLaunchKernelGGL(k_batch_bwd<true>, dim3(cdiv(b->n_ray_chunks, 4)), dim3(256), 0, b->cap, b->meshes_d,
                           b->grads_d, b->chunks_d + b->n_cpq_chunks, b->n_ray_chunks, b->n_cpq_chunks, none,
                           k0, v0, terms);
    if (m) {
        // each (mesh, vertex)'s terms in chunk order, summed in a fixed order
        const bool in1 = radix_sort_pairs(k0, v0, k1, v1, m, kbits, sort_scratch, b->cap);
        hipLaunchKernelGGL(k_wave_segsum<kBlock>, dim3(cdiv(m, kBlock)), dim3(kBlock), 0, b->cap, m, in1 ? k1 : k0,
                           in1 ? v1 : v0, terms, partial);
        hipLaunchKernelGGL((k_run_sum<kBlock, MeshVertexSink>), dim3(cdiv(m, kBlock)), dim3(kBlock), 0, b->cap, m,
                           in1 ? k1 : k0, partial, none, MeshVertexSink{b->grads_d, voff_d, (uint32_t)n_meshes});
    }
    hipError_t el = hipGetLastError();
    hipGraph_t g = nullptr;
    hipError_t ec = hipStreamEndCapture(b->cap, &g);
    b->bgraph = g;
    int rc = GPURT_OK;
    if (el != hipSuccess) rc = hip_error(el, "batch backward launch");
    else if (ec != hipSuccess) rc = hip_error(ec, "hipStreamEndCapture");
    else if (hipGraphInstantiate(&b->bexec, b->bgraph, nullptr, nullptr, 0) != hipSuccess)
        rc = hip_error(hipGetLastError(), "hipGraphInstantiate");
    if (rc) b->drop_backward();
    return rc;
}

int gpurt_batch_run_backward(gpurt_batch b, void* stream) {
    if (!b) return set_error(GPURT_E_INVALID, "null batch");
    if (!b->bexec) return set_error(GPURT_E_INVALID, "gpurt_batch_set_grads was not called");
    DevGuard dg(b->ctx->device);
    if (!dg.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    GPURT_HIP(hipGraphLaunch(b->bexec, (hipStream_t)stream));
    return GPURT_OK;
}

void gpurt_batch_destroy(gpurt_batch b) {
    if (!b) return;
    DevGuard dg(b->device);
    (void)hipDeviceSynchronize();
    delete b;
}

}  // extern "C"
