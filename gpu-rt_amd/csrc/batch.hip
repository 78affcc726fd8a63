// batch.hip -- many meshes built and queried together.
//
// The builder is segmented by construction (the mesh id sits in the high bits of every
// Morton key, so each mesh is its own subtree of one radix tree), so the BVHs of all
// meshes are built by the same ~30 launches as one mesh.  Queries then run in ONE
// persistent launch: every wave pulls a 64-query chunk (closest-point or closest-hit,
// so the branch is wave-uniform) from a device work counter until the chunk list is
// exhausted, and starts at its mesh's root (wide node m of the shared node array).
// Chunks are ordered by decreasing mesh size so the tail of the launch is made of
// cheap chunks.  The whole sequence -- build of every mesh plus the query launch --
// is recorded once into a HIP graph and replayed by gpurt_batch_run.
#include <algorithm>
#include <numeric>
#include <vector>

#include "traverse.h"

namespace gpurt_impl {

struct BatchMeshDev {
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

constexpr uint32_t kRayChunk = 0x80000000u;  // chunk.x bit 31: ray chunk (else closest point)

__global__ void k_reset_counter(uint32_t* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0u;
}

template <int W, int S, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WavesHint<W>::value))) void k_batch_query(
    const void* nodes_v, const float4* __restrict__ tris, const BatchMeshDev* __restrict__ meshes,
    const uint2* __restrict__ chunks, uint32_t n_chunks, uint32_t* counter, uint2* spill,
    uint32_t spill_slots, uint32_t* err) {
    const WNode<W>* __restrict__ nodes = reinterpret_cast<const WNode<W>*>(nodes_v);
    LaneStack<S, BS> st;
    st.init(spill, spill_slots, blockIdx.x * BS + threadIdx.x, gridDim.x * BS);
    const uint32_t lane = threadIdx.x & 63u;
    unsigned long long dn = 0, dt = 0;
    for (;;) {
        uint32_t c = 0;
        if (lane == 0) c = gatomic_add(counter, 1u);
        // wave-uniform from here on: the chunk and mesh descriptors live in SGPRs
        c = __builtin_amdgcn_readfirstlane(c);
        if (c >= n_chunks) break;
        const uint2 ch = chunks[c];
        const uint32_t m = ch.x & ~kRayChunk;
        const uint32_t i = ch.y + lane;
        const BatchMeshDev md = meshes[m];
        if (ch.x & kRayChunk) {
            if (i < md.n_r) {
                const float4 r0 = md.rays[2 * (size_t)i];
                const float4 r1 = md.rays[2 * (size_t)i + 1];
                RayHit h;
                if (!ray_one<W, S, BS, false, false>(nodes, tris, m, v3(r0.x, r0.y, r0.z),
                                                     v3(r1.x, r1.y, r1.z), r0.w, r1.w, st, h, dn, dt))
                    gatomic_or(err, 1u);
                const bool hit = h.id != 0xFFFFFFFFu;
                md.r_prim[i] = hit ? (int32_t)h.id : -1;
                md.r_t[i] = hit ? h.t : INFINITY;
                if (md.r_uv) {
                    md.r_uv[2 * (size_t)i] = hit ? h.u : 0.0f;
                    md.r_uv[2 * (size_t)i + 1] = hit ? h.v : 0.0f;
                }
            }
        } else {
            if (i < md.n_q) {
                const V3 q = v3(md.q[3 * (size_t)i], md.q[3 * (size_t)i + 1], md.q[3 * (size_t)i + 2]);
                CpqHit h;
                if (!cpq_one<W, S, BS, false>(nodes, tris, m, q, st, h, dn, dt)) gatomic_or(err, 1u);
                md.q_prim[i] = (int32_t)h.id;
                md.q_pt[3 * (size_t)i] = h.cp.x;
                md.q_pt[3 * (size_t)i + 1] = h.cp.y;
                md.q_pt[3 * (size_t)i + 2] = h.cp.z;
                md.q_dist[i] = (h.id != 0xFFFFFFFFu) ? sqrtf(h.d2) : INFINITY;
                if (md.q_bary) {
                    md.q_bary[3 * (size_t)i] = h.bary.x;
                    md.q_bary[3 * (size_t)i + 1] = h.bary.y;
                    md.q_bary[3 * (size_t)i + 2] = h.bary.z;
                }
            }
        }
    }
}

template <int W, int S, int BS>
static void launch_batch_query(unsigned grid, hipStream_t s, const void* nodes, const float4* tris,
                               const BatchMeshDev* meshes, const uint2* chunks, uint32_t n_chunks,
                               uint32_t* counter, uint2* spill, uint32_t spill_slots, uint32_t* err) {
    hipLaunchKernelGGL((k_batch_query<W, S, BS>), dim3(grid), dim3(BS), S * BS * sizeof(uint2), s, nodes, tris, meshes, chunks,
                       n_chunks, counter, spill, spill_slots, err);
}

template <int W>
static int launch_batch_w(int S, unsigned grid, hipStream_t s, const void* nodes, const float4* tris,
                          const BatchMeshDev* meshes, const uint2* chunks, uint32_t n_chunks,
                          uint32_t* counter, uint2* spill, uint32_t spill_slots, uint32_t* err) {
    switch (S) {
        case 8: launch_batch_query<W, 8, 256>(grid, s, nodes, tris, meshes, chunks, n_chunks, counter, spill, spill_slots, err); return 256;
        case 16: launch_batch_query<W, 16, 256>(grid, s, nodes, tris, meshes, chunks, n_chunks, counter, spill, spill_slots, err); return 256;
        case 32: launch_batch_query<W, 32, 128>(grid, s, nodes, tris, meshes, chunks, n_chunks, counter, spill, spill_slots, err); return 128;
        case 64: launch_batch_query<W, 64, 64>(grid, s, nodes, tris, meshes, chunks, n_chunks, counter, spill, spill_slots, err); return 64;
    }
    return 0;
}

static int block_for_stack_b(int S) { return S <= 16 ? 256 : (S == 32 ? 128 : 64); }

}  // namespace gpurt_impl

using namespace gpurt_impl;

struct gpurt_batch_s {
    gpurt_ctx ctx = nullptr;
    int width = 4, leaf_size = 4, stack = 8;
    BuildBufs build;
    BatchMeshDev* meshes_d = nullptr;
    uint2* chunks_d = nullptr;
    uint32_t n_chunks = 0;
    uint32_t* counter_d = nullptr;
    uint2* spill = nullptr;
    uint32_t spill_slots = 0;
    unsigned grid = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~gpurt_batch_s() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        if (meshes_d) (void)hipFree(meshes_d);
        if (chunks_d) (void)hipFree(chunks_d);
        if (counter_d) (void)hipFree(counter_d);
        if (spill) (void)hipFree(spill);
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
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, s, b->counter_d);
    int bs = 0;
    switch (b->width) {
        case 2: bs = launch_batch_w<2>(b->stack, b->grid, s, b->build.nodes, b->build.final_tris(), b->meshes_d, b->chunks_d, b->n_chunks, b->counter_d, b->spill, b->spill_slots, b->ctx->err); break;
        case 4: bs = launch_batch_w<4>(b->stack, b->grid, s, b->build.nodes, b->build.final_tris(), b->meshes_d, b->chunks_d, b->n_chunks, b->counter_d, b->spill, b->spill_slots, b->ctx->err); break;
        case 8: bs = launch_batch_w<8>(b->stack, b->grid, s, b->build.nodes, b->build.final_tris(), b->meshes_d, b->chunks_d, b->n_chunks, b->counter_d, b->spill, b->spill_slots, b->ctx->err); break;
    }
    if (!bs) return set_error(GPURT_E_UNSUPPORTED, "batch stack/width");
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
    int W = 4, leaf = 4, tsize = 0, trounds = 0;
    if (bopts) {
        if (bopts->node_width) W = bopts->node_width;
        if (bopts->leaf_size) leaf = bopts->leaf_size;
        tsize = bopts->treelet_size;
        trounds = bopts->treelet_rounds;
    }
    if (W != 2 && W != 4 && W != 8) return set_error(GPURT_E_UNSUPPORTED, "node_width must be 2, 4 or 8");
    if (leaf < 1 || leaf > 8) return set_error(GPURT_E_UNSUPPORTED, "leaf_size must be 1..8");
    if (tsize != 0 && (tsize < 5 || tsize > 7))
        return set_error(GPURT_E_UNSUPPORTED, "treelet_size must be 0 or 5..7");
    if (trounds < 0 || trounds > 1) return set_error(GPURT_E_UNSUPPORTED, "treelet_rounds must be 0 or 1");
    if (tsize && !trounds) trounds = 1;
    int S = (qopts && qopts->stack_depth) ? qopts->stack_depth : (W == 8 ? 16 : 8);
    if (S != 8 && S != 16 && S != 32 && S != 64)
        return set_error(GPURT_E_UNSUPPORTED, "stack_depth must be 0, 8, 16, 32 or 64");
    if (qopts && (qopts->waves_per_cu < 0 || qopts->waves_per_cu > 64))
        return set_error(GPURT_E_UNSUPPORTED, "waves_per_cu must be 0..64");

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
        md[m] = BatchMeshDev{g.q, g.q_prim, g.q_pt, g.q_dist, g.q_bary,
                             reinterpret_cast<const float4*>(g.rays), g.r_prim, g.r_t, g.r_uv,
                             (uint32_t)g.n_q, (uint32_t)g.n_r};
    }
    if (off > kMaxLeafFirst - 8) return set_error(GPURT_E_INVALID, "total triangle count exceeds 2^28-9");

    DevGuard dg(ctx->device);
    if (!dg.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    auto* b = new gpurt_batch_s();
    b->ctx = ctx;
    b->width = W;
    b->leaf_size = leaf;
    b->stack = S;
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
    for (uint32_t m : order) {
        for (uint32_t st = 0; st < md[m].n_q; st += 64) chunks.push_back(make_uint2(m, st));
        for (uint32_t st = 0; st < md[m].n_r; st += 64) chunks.push_back(make_uint2(m | kRayChunk, st));
    }
    b->n_chunks = (uint32_t)chunks.size();
    const int BS = block_for_stack_b(S);
    const int waves_per_block = BS / 64;
    const int reg_waves = 4 * WavesHint<4>::value;
    int blocks_per_cu = std::min(ctx->lds_per_cu / (S * BS * 8),
                                 std::min(32, W == 8 ? 4 * WavesHint<8>::value : reg_waves) / waves_per_block);
    if (qopts && qopts->waves_per_cu > 0) blocks_per_cu = std::max(1, qopts->waves_per_cu / waves_per_block);
    blocks_per_cu = std::max(1, blocks_per_cu);
    b->grid = (unsigned)std::max<size_t>(1, std::min<size_t>((size_t)ctx->n_cu * blocks_per_cu,
                                                             (size_t)b->n_chunks * 64 / BS + 1));
    // deepest possible wide tree (kMaxLevels levels) bounds the stack: no overflow
    const int need = (W - 1) * kMaxLevels + 1;
    b->spill_slots = need > S ? (uint32_t)(need - S) : 0u;
    bool fail = hipMalloc((void**)&b->meshes_d, sizeof(BatchMeshDev) * n_meshes) != hipSuccess ||
                hipMalloc((void**)&b->chunks_d, sizeof(uint2) * std::max<size_t>(1, chunks.size())) != hipSuccess ||
                hipMalloc((void**)&b->counter_d, 256) != hipSuccess ||
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
    // one eager run validates the inputs (vertex indices) before the handle is returned
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

void gpurt_batch_destroy(gpurt_batch b) {
    if (!b) return;
    DevGuard dg(b->ctx->device);
    (void)hipDeviceSynchronize();
    delete b;
}

}  // extern "C"
