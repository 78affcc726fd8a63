// query.hip -- plain (per-mesh) closest-point and ray traversal launches for gfx950.
//
// Replaces the hardware traversal behind traceRayEXT (src/shaders/rt/rt.rgen:257-291)
// and adds the closest-point query on the same BVH.  Each kernel runs the streaming
// (lane-refill) traversal of traverse.h over the queries with a grid sized to the waves
// that fit on the chip at once: lane g takes queries g, g + total, g + 2 total, ...
#include <algorithm>
#include <mutex>

#include "traverse.h"

namespace gpurt_impl {

struct TravArgs {
    const void* nodes;
    const float4* tris;
    uint32_t root;
    uint32_t n;
    // closest point (the traversal writes the winner's id; k_cpq_finish the rest)
    const float* q;
    // rays
    const float4* rays;
    float* t;
    float* uv;
    uint8_t* occ;
    // shared
    int32_t* prim;
    uint2* spill;
    uint32_t spill_slots;
    uint32_t* counter;  // kQueues work-queue positions (64-query chunks), zero at launch
    unsigned long long* stats;
    uint32_t* err;
    uint32_t leaf_trigger;  // 1..64, see leaf_step_now
};

template <int W, bool Q, int S, int BS, bool STATS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WavesHint<W, Q>::value))) void k_cpq(
    TravArgs a) {
    const NodeOf<W, Q>* __restrict__ nodes = reinterpret_cast<const NodeOf<W, Q>*>(a.nodes);
    const uint32_t total = gridDim.x * BS;
    const uint32_t gid = blockIdx.x * BS + threadIdx.x;
    LaneStack<S, BS> st;
    st.init(a.spill, a.spill_slots, gid, total);
    unsigned long long n_nodes = 0, n_tris = 0;
    QueueSource src{nullptr, a.q, a.n, first_query(gid, total), total, a.counter};
    auto out = [&](uint32_t qi, uint32_t id) { a.prim[qi] = id == 0xFFFFFFFFu ? -1 : (int32_t)id; };
    if (!cpq_stream<NodeOf<W, Q>, S, BS, STATS>(nodes, a.tris, src, a.leaf_trigger, st, out, n_nodes, n_tris))
        gatomic_or(a.err, 1u);
    if (STATS) {
        atomicAdd(&a.stats[0], n_nodes);
        atomicAdd(&a.stats[1], n_tris);
        atomicAdd(&a.stats[2], (unsigned long long)st.spills);
    }
}

template <int W, bool Q, int S, int BS, bool ANY, bool STATS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WavesHint<W, Q>::value))) void k_ray(
    TravArgs a) {
    const NodeOf<W, Q>* __restrict__ nodes = reinterpret_cast<const NodeOf<W, Q>*>(a.nodes);
    const uint32_t total = gridDim.x * BS;
    const uint32_t gid = blockIdx.x * BS + threadIdx.x;
    LaneStack<S, BS> st;
    st.init(a.spill, a.spill_slots, gid, total);
    unsigned long long n_nodes = 0, n_tris = 0;
    QueueSource src{a.rays, nullptr, a.n, first_query(gid, total), total, a.counter};
    auto out = [&](uint32_t ri, uint32_t id, float t, float u, float v, bool any) {
        if (ANY) {
            a.occ[ri] = any ? 1 : 0;
        } else {
            const bool hit = id != 0xFFFFFFFFu;
            a.prim[ri] = hit ? (int32_t)id : -1;
            a.t[ri] = hit ? t : INFINITY;
            if (a.uv) {
                a.uv[2 * (size_t)ri] = hit ? u : 0.0f;
                a.uv[2 * (size_t)ri + 1] = hit ? v : 0.0f;
            }
        }
    };
    if (!ray_stream<NodeOf<W, Q>, S, BS, ANY, STATS>(nodes, a.tris, src, a.leaf_trigger, st, out, n_nodes, n_tris))
        gatomic_or(a.err, 1u);
    if (STATS) {
        atomicAdd(&a.stats[0], n_nodes);
        atomicAdd(&a.stats[1], n_tris);
        atomicAdd(&a.stats[2], (unsigned long long)st.spills);
    }
}

__global__ void k_clear_counters(uint32_t* c) {
    if (threadIdx.x < kQueues) c[threadIdx.x * kQueueStride] = 0u;
}

// ------------------------------------------------------------------------------------
// launch plumbing
// ------------------------------------------------------------------------------------

template <int W, bool Q, int S, int BS, bool STATS>
static void launch_kind(QKind kind, const TravArgs& a, unsigned grid, hipStream_t s) {
    switch (kind) {
        case QKind::Cpq:
            hipLaunchKernelGGL((k_cpq<W, Q, S, BS, STATS>), dim3(grid), dim3(BS), (S * sizeof(uint2) + kGroupKeyBytes) * BS, s, a);
            break;
        case QKind::Closest:
            hipLaunchKernelGGL((k_ray<W, Q, S, BS, false, STATS>), dim3(grid), dim3(BS), (S * sizeof(uint2) + kGroupKeyBytes) * BS, s, a);
            break;
        case QKind::Any:
            hipLaunchKernelGGL((k_ray<W, Q, S, BS, true, STATS>), dim3(grid), dim3(BS), (S * sizeof(uint2) + kGroupKeyBytes) * BS, s, a);
            break;
    }
}

// LDS per block is S * BS * 8 bytes: 16 KB for S = 8, 32 KB for the others.
template <int W, bool Q, bool STATS>
static int dispatch_stack(int S, QKind kind, const TravArgs& a, unsigned grid, hipStream_t s) {
    switch (S) {
        case 8: launch_kind<W, Q, 8, 256, STATS>(kind, a, grid, s); return 256;
        case 16: launch_kind<W, Q, 16, 256, STATS>(kind, a, grid, s); return 256;
        case 32: launch_kind<W, Q, 32, 128, STATS>(kind, a, grid, s); return 128;
        case 64: launch_kind<W, Q, 64, 64, STATS>(kind, a, grid, s); return 64;
    }
    return 0;
}

// Defaults per query kind, from sweeps on the rays_scene and closest-point workloads:
// the parked-leaf share that switches a wave to leaf steps, the idle-lane share at which
// a ray wave writes its finished rays and fetches new ones (a refill costs the whole wave
// one pass through the ray setup, so it pays to gather several finished lanes), and
// resident waves per CU (fewer waves leave the L2 less crowded; with work donation the
// end of a launch no longer needs extra waves to hide its idle lanes).  rays_scene, 1M +
// 1M rays: closest 3.29 -> 4.05, any hit 4.51 -> 5.70 G rays/s against a refill on every
// finished lane at 8 waves/SIMD; closest-point queries near a 1.3M-triangle surface: 0.88
// -> 1.06 G queries/s at 4 waves/SIMD.
// Rays re-swept on rays_scene after the two-visit node step: closest hit at 28 waves per
// CU and a refill at 24 idle lanes 218 us (24 waves, 16 lanes: 230 us), any hit at 28
// waves and 16 lanes 148-151 us (158 us); 32 waves crowd the L2 (+5%).  (The TLAS
// passes its own launch options.)
static uint32_t default_leaf_trigger(QKind) { return 32u; }
static uint32_t default_refill_trigger(QKind kind) { return kind == QKind::Closest ? 24u : 16u; }
static int default_waves_per_cu(QKind kind) { return kind == QKind::Cpq ? 16 : 28; }

static int block_for_stack(int S) { return S <= 16 ? 256 : (S == 32 ? 128 : 64); }

int plan_traversal(gpurt_ctx ctx, int width, bool q8, QKind kind, int stack_need, const gpurt_query_opts* opts,
                   size_t n, TravPlan& p, int waves_per_simd, int lds_extra) {
    if (n > 0xFFFFFFFFull - 1024u) return set_error(GPURT_E_INVALID, "too many queries in one call");
    // default: the smallest LDS stack (most resident waves) for rays; 8-wide nodes push
    // more entries per visit, so they default to 16, and so do closest-point queries,
    // whose 16 waves per CU fit 16-entry stacks in LDS and whose wide searches otherwise
    // spill (diff 943 -> 867 us, cpq_large 13.97 -> 13.60 ms, cpq_small -3%)
    p.S = opts && opts->stack_depth ? opts->stack_depth : (width == 8 || kind == QKind::Cpq ? 16 : 8);
    if (p.S != 8 && p.S != 16 && p.S != 32 && p.S != 64)
        return set_error(GPURT_E_UNSUPPORTED, "stack_depth must be 8, 16, 32 or 64");
    p.BS = block_for_stack(p.S);
    // resident blocks per CU: limited by LDS (the stacks), by the 32 wave slots of a CU
    // and by registers (WavesHint waves per SIMD), unless the caller asks for waves/CU
    const int waves_per_block = p.BS / 64;
    const int hint = waves_per_simd > 0 ? waves_per_simd
                     : width == 8 ? (q8 ? WavesHint<8, true>::value : WavesHint<8, false>::value)
                                  : (q8 ? WavesHint<4, true>::value : WavesHint<4, false>::value);
    int blocks_per_cu = std::min(ctx->lds_per_cu / ((p.S * 8 + lds_extra) * p.BS), std::min(32, 4 * hint) / waves_per_block);
    if (opts && opts->waves_per_cu > 0)
        blocks_per_cu = std::max(1, opts->waves_per_cu / waves_per_block);
    else if (default_waves_per_cu(kind) > 0)
        blocks_per_cu = std::min(blocks_per_cu, std::max(1, default_waves_per_cu(kind) / waves_per_block));
    blocks_per_cu = std::max(blocks_per_cu, 1);
    // up to kSpread lanes per query: a small set is spread over more waves (first_query),
    // whose idle lanes take work from the busy ones once the queues are empty
    constexpr size_t kSpread = 4;
    p.grid = (unsigned)std::min<size_t>(cdiv(n * kSpread, p.BS), (size_t)ctx->n_cu * blocks_per_cu);
    if (p.grid == 0) p.grid = 1;
    p.total = p.grid * p.BS;
    // the part of the worst-case stack beyond the LDS slots spills to device memory
    p.spill_slots = stack_need > p.S ? (uint32_t)(stack_need - p.S) : 0u;
    p.leaf_trigger = opts && opts->leaf_trigger ? (uint32_t)opts->leaf_trigger : default_leaf_trigger(kind);
    const uint32_t refill = opts && opts->refill_trigger ? (uint32_t)opts->refill_trigger : default_refill_trigger(kind);
    p.leaf_trigger |= refill << 8;  // bits 8-15 of the trigger word
    return GPURT_OK;
}

static constexpr size_t kCounterBytes = (size_t)kQueues * kQueueStride * sizeof(uint32_t);

// the queue counters and the stack entries beyond the LDS slots
size_t traversal_scratch_bytes(const TravPlan& p) {
    return kCounterBytes + (size_t)p.spill_slots * p.total * sizeof(uint2);
}

void carve_traversal_scratch(const TravPlan& p, void* mem, TravScratch& t) {
    t.mem = mem;
    t.counters = (uint32_t*)mem;
    t.spill = p.spill_slots ? (uint2*)((char*)mem + kCounterBytes) : nullptr;
}

void clear_traversal_counters(const TravScratch& t, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(kQueues), 0, s, t.counters);
}

int alloc_traversal_scratch(gpurt_ctx_s* c, const TravPlan& p, hipStream_t s, TravScratch& t) {
    CallScratch cs;
    const int rc = ctx_scratch(c, s, traversal_scratch_bytes(p), cs);
    if (rc) {
        t.mem = nullptr;
        return rc;
    }
    carve_traversal_scratch(p, cs.mem, t);
    t.cs = cs;
    clear_traversal_counters(t, s);
    return GPURT_OK;
}

void free_traversal_scratch(gpurt_ctx_s* c, TravScratch& t) {
    ctx_scratch_done(c, t.cs);
    t.mem = nullptr;
}

// ------------------------------------------------------------------------------------
// fixed scratch of gpurt_trace_closest_noalloc
// ------------------------------------------------------------------------------------

// The queue counters and stack spill area of the no-allocation entry point.  Each device
// holds its own copy from the first context created on it until the process ends; its
// address is taken once per device by noalloc_scratch_init (hipGetSymbolAddress), and
// only that address is handed to kernels (the array's name is not a device address in
// host code).
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return set_error(GPURT_E_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return GPURT_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    }
    g_scratch_ptr[device] = ptr;
    return GPURT_OK;
}

static void* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

// Fits a plan into `bytes` of scratch: the spill area grows with the resident lanes, so
// a deep tree is traversed with fewer resident blocks (whole blocks, at least one).
static int fit_plan_to_scratch(TravPlan& p, size_t bytes) {
    if (traversal_scratch_bytes(p) <= bytes) return GPURT_OK;
    const size_t per_block = (size_t)p.spill_slots * p.BS * sizeof(uint2);
    const size_t room = bytes - kCounterBytes;
    const size_t blocks = room / per_block;
    if (blocks == 0)
        return set_error(GPURT_E_INVALID, "tree too deep for the fixed traversal scratch of "
                                          "gpurt_trace_closest_noalloc (64 MiB)");
    p.grid = (unsigned)std::min<size_t>(p.grid, blocks);
    p.total = p.grid * p.BS;
    return GPURT_OK;
}

// fixed_scratch: memory of `fixed_bytes` the caller holds for the queue counters and
// stack spill (the grid is fitted into it); nullptr: a stream-ordered allocation per call
static int run_query(gpurt_bvh_s* b, QKind kind, TravArgs a, size_t n, const gpurt_query_opts* opts,
                     hipStream_t s, void* fixed_scratch = nullptr, size_t fixed_bytes = 0) {
    if (n == 0) return GPURT_OK;
    const bool q8 = kind == QKind::Cpq ? cpq_use_q8(b->format) : rays_use_q8(b->format);
    // worst case stack: (W-1) entries per level above the last plus W for the last
    TravPlan p;
    int rc = plan_traversal(b->ctx, b->width, q8, kind, (b->width - 1) * b->depth + 1, opts, n, p);
    if (rc) return rc;
    TravScratch t;
    if (fixed_scratch) {
        if ((rc = fit_plan_to_scratch(p, fixed_bytes))) return rc;
        carve_traversal_scratch(p, fixed_scratch, t);
        clear_traversal_counters(t, s);
        t.mem = nullptr;  // not freed below
    } else if ((rc = alloc_traversal_scratch(b->ctx, p, s, t))) {
        return rc;
    }
    a.tris = b->tris;
    a.root = 0;
    a.n = (uint32_t)n;
    a.spill = t.spill;
    a.spill_slots = p.spill_slots;
    a.counter = t.counters;
    a.stats = opts ? (unsigned long long*)opts->stats : nullptr;
    a.err = b->ctx->err;
    a.leaf_trigger = p.leaf_trigger;
    const bool stats = a.stats != nullptr;
    int bs = 0;
    a.nodes = q8 ? b->qnodes : b->nodes;
    switch (b->width * 4 + (q8 ? 2 : 0) + (stats ? 1 : 0)) {
        case 8: bs = dispatch_stack<2, false, false>(p.S, kind, a, p.grid, s); break;
        case 9: bs = dispatch_stack<2, false, true>(p.S, kind, a, p.grid, s); break;
        case 10: bs = dispatch_stack<2, true, false>(p.S, kind, a, p.grid, s); break;
        case 11: bs = dispatch_stack<2, true, true>(p.S, kind, a, p.grid, s); break;
        case 16: bs = dispatch_stack<4, false, false>(p.S, kind, a, p.grid, s); break;
        case 17: bs = dispatch_stack<4, false, true>(p.S, kind, a, p.grid, s); break;
        case 18: bs = dispatch_stack<4, true, false>(p.S, kind, a, p.grid, s); break;
        case 19: bs = dispatch_stack<4, true, true>(p.S, kind, a, p.grid, s); break;
        case 32: bs = dispatch_stack<8, false, false>(p.S, kind, a, p.grid, s); break;
        case 33: bs = dispatch_stack<8, false, true>(p.S, kind, a, p.grid, s); break;
        case 34: bs = dispatch_stack<8, true, false>(p.S, kind, a, p.grid, s); break;
        case 35: bs = dispatch_stack<8, true, true>(p.S, kind, a, p.grid, s); break;
        default: break;
    }
    hipError_t le = hipGetLastError();
    free_traversal_scratch(b->ctx, t);
    if (bs == 0) return set_error(GPURT_E_UNSUPPORTED, "node width");
    if (le != hipSuccess) return hip_error(le, "traversal launch");
    return GPURT_OK;
}

__global__ __launch_bounds__(kBlock) void k_cpq_finish(uint32_t n, const float* __restrict__ q,
                                                       const int32_t* __restrict__ prim,
                                                       const uint32_t* __restrict__ tri_idx, uint32_t n_tris,
                                                       const float* __restrict__ verts, uint32_t n_verts,
                                                       float* __restrict__ pt, float* __restrict__ dist,
                                                       float* __restrict__ bary) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    cpq_finish_one(i, q, prim, tri_idx, n_tris, verts, 3, n_verts, pt, dist, bary);
}

int launch_cpq(gpurt_bvh_s* b, const float* q, size_t n, int32_t* prim, float* pt, float* dist,
               float* bary, const gpurt_query_opts* opts, hipStream_t s) {
    TravArgs a{};
    a.q = q;
    a.prim = prim;
    int rc = run_query(b, QKind::Cpq, a, n, opts, s);
    if (rc || n == 0) return rc;
    // the winners' point, distance and barycentrics in FP64 (one gather per query)
    hipLaunchKernelGGL(k_cpq_finish, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, (uint32_t)n, q, prim, b->tri_idx,
                       (uint32_t)b->n_tris, b->verts, (uint32_t)b->n_verts, pt, dist, bary);
    GPURT_HIP(hipGetLastError());
    return GPURT_OK;
}

int launch_ray(gpurt_bvh_s* b, const gpurt_ray* rays, size_t n, int32_t* prim, float* t, float* uv,
               uint8_t* occ, bool any, const gpurt_query_opts* opts, hipStream_t s) {
    TravArgs a{};
    a.rays = reinterpret_cast<const float4*>(rays);
    a.prim = prim;
    a.t = t;
    a.uv = uv;
    a.occ = occ;
    return run_query(b, any ? QKind::Any : QKind::Closest, a, n, opts, s);
}

int launch_ray_noalloc(gpurt_bvh_s* b, const gpurt_ray* rays, size_t n, int32_t* prim, float* t, float* uv,
                       const gpurt_query_opts* opts, hipStream_t s) {
    void* mem = noalloc_scratch(b->ctx->device);
    if (!mem) return set_error(GPURT_E_DEVICE, "fixed traversal scratch not set up on this device");
    TravArgs a{};
    a.rays = reinterpret_cast<const float4*>(rays);
    a.prim = prim;
    a.t = t;
    a.uv = uv;
    return run_query(b, QKind::Closest, a, n, opts, s, mem, kNoAllocScratchBytes);
}

}  // namespace gpurt_impl
