// query.hip -- closest-point and ray traversal kernels for gfx950.
//
// Replaces the hardware traversal behind traceRayEXT (src/shaders/rt/rt.rgen:257-291)
// and adds the closest-point query on the same BVH.
//
// Traversal: one lane per query, grid-stride over the queries.  A lane keeps a stack of
// (node-or-leaf reference, lower bound) pairs: the first S entries live in LDS laid out
// [slot][lane] (conflict-free: consecutive lanes hit consecutive banks), deeper entries
// spill to a device buffer sized from the tree depth, so no traversal can overflow.
// A wide node is fetched as 2*W float4 loads (W=4: one 128-byte line); children are
// bounded, sorted by the bound with a sorting network, and pushed far-to-near so the
// nearest is taken next without a stack round trip.  Culls compare against the best
// value with a small relative slack and boxes are inflated at build time, so a child
// that can hold an equal-or-better primitive is never culled: results do not depend
// on the tree's shape, and ties resolve to the lowest primitive id.
#include <algorithm>

#include "gpurt_internal.h"
#include "tri_math.h"

namespace gpurt_impl {

struct TravArgs {
    const void* nodes;
    const float4* tris;
    uint32_t root;
    uint32_t n;
    // closest point
    const float* q;
    float* pt;
    float* dist;
    float* bary;
    // rays
    const float4* rays;
    float* t;
    float* uv;
    uint8_t* occ;
    // shared
    int32_t* prim;
    uint2* spill;
    uint32_t spill_slots;
    unsigned long long* stats;
    uint32_t* err;
};

template <int S, int BS>
struct LaneStack {
    uint2* lds;
    uint2* spill;
    uint32_t gid, total, cap;
    int sp;
    __device__ __forceinline__ bool push(uint32_t ref, float key) {
        if ((uint32_t)sp >= cap) return false;
        uint2 e = make_uint2(ref, __float_as_uint(key));
        if (sp < S) lds[sp * BS] = e;
        else spill[(size_t)(sp - S) * total + gid] = e;
        sp++;
        return true;
    }
    __device__ __forceinline__ uint2 pop() {
        sp--;
        if (sp < S) return lds[sp * BS];
        return spill[(size_t)(sp - S) * total + gid];
    }
};

template <int W>
__device__ __forceinline__ void sort_children(float (&d)[W], uint32_t (&c)[W]) {
    auto cx = [&](int i, int j) {
        if (d[j] < d[i]) {
            float tf = d[i]; d[i] = d[j]; d[j] = tf;
            uint32_t tc = c[i]; c[i] = c[j]; c[j] = tc;
        }
    };
    if constexpr (W == 2) {
        cx(0, 1);
    } else if constexpr (W == 4) {
        cx(0, 1); cx(2, 3); cx(0, 2); cx(1, 3); cx(1, 2);
    } else {
        // Batcher odd-even merge sort network for 8 keys (19 comparators)
        cx(0, 1); cx(2, 3); cx(4, 5); cx(6, 7);
        cx(0, 2); cx(1, 3); cx(4, 6); cx(5, 7);
        cx(1, 2); cx(5, 6);
        cx(0, 4); cx(1, 5); cx(2, 6); cx(3, 7);
        cx(2, 4); cx(3, 5);
        cx(1, 2); cx(3, 4); cx(5, 6);
    }
}

// ------------------------------------------------------------------------------------
// closest point
// ------------------------------------------------------------------------------------

// Resident waves per SIMD the register allocator targets: 8 (<= 64 VGPRs) for 2- and
// 4-wide nodes; 8-wide nodes need more live registers for the child boxes.
template <int W>
struct WavesHint {
    static constexpr int value = W == 8 ? 4 : 8;
};

constexpr float kCullSlack = 1.0f + 0x1p-16f;  // relative slack on squared distances

template <int W, int S, int BS, bool STATS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WavesHint<W>::value))) void k_cpq(TravArgs a) {
    __shared__ uint2 stk[S * BS];
    const WNode<W>* __restrict__ nodes = reinterpret_cast<const WNode<W>*>(a.nodes);
    const float4* __restrict__ tris = a.tris;
    const uint32_t total = gridDim.x * BS;
    const uint32_t gid = blockIdx.x * BS + threadIdx.x;
    LaneStack<S, BS> st;
    st.lds = stk + threadIdx.x;
    st.spill = a.spill;
    st.gid = gid;
    st.total = total;
    st.cap = S + a.spill_slots;
    unsigned long long n_nodes = 0, n_tris = 0;
    for (uint32_t qi = gid; qi < a.n; qi += total) {
        const V3 q = v3(a.q[3 * (size_t)qi], a.q[3 * (size_t)qi + 1], a.q[3 * (size_t)qi + 2]);
        float best = INFINITY, thr = INFINITY;
        uint32_t bid = 0xFFFFFFFFu;
        V3 bcp = v3(0.f, 0.f, 0.f), bb = v3(0.f, 0.f, 0.f);
        st.sp = 0;
        uint32_t cur = a.root;
        bool have = true;
        bool overflow = false;
        for (;;) {
            if (!have) {
                if (st.sp == 0) break;
                uint2 e = st.pop();
                if (__uint_as_float(e.y) > thr) continue;
                cur = e.x;
            }
            have = false;
            if (child_is_leaf(cur)) {
                const uint32_t f = leaf_first(cur), cnt = leaf_count(cur);
                for (uint32_t k = 0; k < cnt; k++) {
                    const float4 A = tris[3 * (size_t)(f + k)];
                    const float4 B = tris[3 * (size_t)(f + k) + 1];
                    const float4 C = tris[3 * (size_t)(f + k) + 2];
                    V3 cp, br;
                    float d2 = closest_pt_tri(q, v3(A.x, A.y, A.z), v3(B.x, B.y, B.z),
                                              v3(C.x, C.y, C.z), cp, br);
                    uint32_t id = __float_as_uint(A.w);
                    if (STATS) n_tris++;
                    if (d2 < best || (d2 == best && id < bid)) {
                        best = d2;
                        bid = id;
                        bcp = cp;
                        bb = br;
                        thr = best * kCullSlack;
                    }
                }
                continue;
            }
            if (STATS) n_nodes++;
            const WNode<W> nd = nodes[cur];
            float d[W];
            uint32_t c[W];
#pragma unroll
            for (int j = 0; j < W; j++) {
                float dx = fmaxf(fmaxf(nd.lox[j] - q.x, q.x - nd.hix[j]), 0.0f);
                float dy = fmaxf(fmaxf(nd.loy[j] - q.y, q.y - nd.hiy[j]), 0.0f);
                float dz = fmaxf(fmaxf(nd.loz[j] - q.z, q.z - nd.hiz[j]), 0.0f);
                float dd = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                c[j] = nd.child[j];
                d[j] = (c[j] == kEmpty) ? INFINITY : dd;
            }
            sort_children<W>(d, c);
#pragma unroll
            for (int j = W - 1; j >= 1; j--) {
                if (c[j] != kEmpty && d[j] <= thr) {
                    if (!st.push(c[j], d[j])) overflow = true;
                }
            }
            if (c[0] != kEmpty && d[0] <= thr) {
                cur = c[0];
                have = true;
            }
        }
        if (overflow) atomicOr(a.err, 1u);
        a.prim[qi] = (int32_t)bid;
        a.pt[3 * (size_t)qi] = bcp.x;
        a.pt[3 * (size_t)qi + 1] = bcp.y;
        a.pt[3 * (size_t)qi + 2] = bcp.z;
        a.dist[qi] = (bid != 0xFFFFFFFFu) ? sqrtf(best) : INFINITY;
        if (a.bary) {
            a.bary[3 * (size_t)qi] = bb.x;
            a.bary[3 * (size_t)qi + 1] = bb.y;
            a.bary[3 * (size_t)qi + 2] = bb.z;
        }
    }
    if (STATS) {
        atomicAdd(&a.stats[0], n_nodes);
        atomicAdd(&a.stats[1], n_tris);
    }
}

// ------------------------------------------------------------------------------------
// rays: closest hit and any hit
// ------------------------------------------------------------------------------------

constexpr float kTSlack = 0x1p-20f;

template <int W, int S, int BS, bool ANY, bool STATS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WavesHint<W>::value))) void k_ray(TravArgs a) {
    __shared__ uint2 stk[S * BS];
    const WNode<W>* __restrict__ nodes = reinterpret_cast<const WNode<W>*>(a.nodes);
    const float4* __restrict__ tris = a.tris;
    const uint32_t total = gridDim.x * BS;
    const uint32_t gid = blockIdx.x * BS + threadIdx.x;
    LaneStack<S, BS> st;
    st.lds = stk + threadIdx.x;
    st.spill = a.spill;
    st.gid = gid;
    st.total = total;
    st.cap = S + a.spill_slots;
    unsigned long long n_nodes = 0, n_tris = 0;
    for (uint32_t ri = gid; ri < a.n; ri += total) {
        const float4 r0 = a.rays[2 * (size_t)ri];
        const float4 r1 = a.rays[2 * (size_t)ri + 1];
        const V3 o = v3(r0.x, r0.y, r0.z), d = v3(r1.x, r1.y, r1.z);
        const float tmin = r0.w;
        float best = r1.w;  // tmax, shrinks to the best hit
        uint32_t bid = 0xFFFFFFFFu;
        float bu = 0.f, bv = 0.f;
        bool hit_any = false;
        V3 inv;
        {
            float x = fabsf(d.x) < 1e-30f ? copysignf(1e-30f, d.x) : d.x;
            float y = fabsf(d.y) < 1e-30f ? copysignf(1e-30f, d.y) : d.y;
            float z = fabsf(d.z) < 1e-30f ? copysignf(1e-30f, d.z) : d.z;
            inv = v3(1.0f / x, 1.0f / y, 1.0f / z);
        }
        st.sp = 0;
        uint32_t cur = a.root;
        bool have = true;
        bool overflow = false;
        for (;;) {
            if (!have) {
                if (st.sp == 0) break;
                uint2 e = st.pop();
                if (__uint_as_float(e.y) > best) continue;
                cur = e.x;
            }
            have = false;
            if (child_is_leaf(cur)) {
                const uint32_t f = leaf_first(cur), cnt = leaf_count(cur);
                for (uint32_t k = 0; k < cnt; k++) {
                    const float4 A = tris[3 * (size_t)(f + k)];
                    const float4 B = tris[3 * (size_t)(f + k) + 1];
                    const float4 C = tris[3 * (size_t)(f + k) + 2];
                    if (STATS) n_tris++;
                    float t, u, v;
                    if (!ray_tri(o, d, v3(A.x, A.y, A.z), v3(B.x, B.y, B.z), v3(C.x, C.y, C.z), t,
                                 u, v))
                        continue;
                    if (!(t >= tmin && t <= best)) continue;
                    if (ANY) {
                        hit_any = true;
                        break;
                    }
                    uint32_t id = __float_as_uint(A.w);
                    if (t < best || id < bid) {
                        best = t;
                        bid = id;
                        bu = u;
                        bv = v;
                    }
                }
                if (ANY && hit_any) break;
                continue;
            }
            if (STATS) n_nodes++;
            const WNode<W> nd = nodes[cur];
            float tn[W];
            uint32_t c[W];
#pragma unroll
            for (int j = 0; j < W; j++) {
                float t0x = (nd.lox[j] - o.x) * inv.x, t1x = (nd.hix[j] - o.x) * inv.x;
                float t0y = (nd.loy[j] - o.y) * inv.y, t1y = (nd.hiy[j] - o.y) * inv.y;
                float t0z = (nd.loz[j] - o.z) * inv.z, t1z = (nd.hiz[j] - o.z) * inv.z;
                float lo = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
                float hi = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), best));
                c[j] = nd.child[j];
                bool h = (c[j] != kEmpty) && (lo <= hi + fabsf(hi) * kTSlack);
                tn[j] = h ? lo : INFINITY;
                if (!h) c[j] = kEmpty;
            }
            if (!ANY) sort_children<W>(tn, c);
            // push all but the nearest hit child (far to near), then continue with it;
            // the pending entry is carried in scalars so every index stays static
            uint32_t pend = kEmpty;
            float pend_t = 0.0f;
#pragma unroll
            for (int j = W - 1; j >= 0; j--) {
                if (c[j] == kEmpty) continue;
                if (pend != kEmpty) {
                    if (!st.push(pend, pend_t)) overflow = true;
                }
                pend = c[j];
                pend_t = tn[j];
            }
            if (pend != kEmpty) {
                cur = pend;
                have = true;
            }
        }
        if (overflow) atomicOr(a.err, 1u);
        if (ANY) {
            a.occ[ri] = hit_any ? 1 : 0;
        } else {
            bool hit = bid != 0xFFFFFFFFu;
            a.prim[ri] = hit ? (int32_t)bid : -1;
            a.t[ri] = hit ? best : INFINITY;
            if (a.uv) {
                a.uv[2 * (size_t)ri] = hit ? bu : 0.0f;
                a.uv[2 * (size_t)ri + 1] = hit ? bv : 0.0f;
            }
        }
    }
    if (STATS) {
        atomicAdd(&a.stats[0], n_nodes);
        atomicAdd(&a.stats[1], n_tris);
    }
}

// ------------------------------------------------------------------------------------
// launch plumbing
// ------------------------------------------------------------------------------------

enum class QKind { Cpq, Closest, Any };

template <int W, int S, int BS, bool STATS>
static void launch_kind(QKind kind, const TravArgs& a, unsigned grid, hipStream_t s) {
    switch (kind) {
        case QKind::Cpq:
            hipLaunchKernelGGL((k_cpq<W, S, BS, STATS>), dim3(grid), dim3(BS), 0, s, a);
            break;
        case QKind::Closest:
            hipLaunchKernelGGL((k_ray<W, S, BS, false, STATS>), dim3(grid), dim3(BS), 0, s, a);
            break;
        case QKind::Any:
            hipLaunchKernelGGL((k_ray<W, S, BS, true, STATS>), dim3(grid), dim3(BS), 0, s, a);
            break;
    }
}

// LDS per block is S * BS * 8 bytes: 16 KB for S = 8, 32 KB for the others.
template <int W, bool STATS>
static int dispatch_stack(int S, QKind kind, const TravArgs& a, unsigned grid, hipStream_t s) {
    switch (S) {
        case 8: launch_kind<W, 8, 256, STATS>(kind, a, grid, s); return 256;
        case 16: launch_kind<W, 16, 256, STATS>(kind, a, grid, s); return 256;
        case 32: launch_kind<W, 32, 128, STATS>(kind, a, grid, s); return 128;
        case 64: launch_kind<W, 64, 64, STATS>(kind, a, grid, s); return 64;
    }
    return 0;
}

static int block_for_stack(int S) { return S <= 16 ? 256 : (S == 32 ? 128 : 64); }

static int run_query(gpurt_bvh_s* b, QKind kind, TravArgs a, size_t n, const gpurt_query_opts* opts,
                     hipStream_t s) {
    if (n == 0) return GPURT_OK;
    if (n > 0xFFFFFFFFull - 1024u) return set_error(GPURT_E_INVALID, "too many queries in one call");
    // default: the smallest LDS stack (most resident waves); 8-wide nodes push more
    // entries per visit, so they default to 16
    int S = opts && opts->stack_depth ? opts->stack_depth : (b->width == 8 ? 16 : 8);
    if (S != 8 && S != 16 && S != 32 && S != 64)
        return set_error(GPURT_E_UNSUPPORTED, "stack_depth must be 8, 16, 32 or 64");
    const int BS = block_for_stack(S);
    gpurt_ctx ctx = b->ctx;
    // resident blocks per CU: limited by LDS (the stacks), by the 32 wave slots of a CU
    // and by registers (WavesHint waves per SIMD), unless the caller asks for waves/CU
    const int waves_per_block = BS / 64;
    const int reg_waves = 4 * (b->width == 8 ? WavesHint<8>::value : WavesHint<4>::value);
    int blocks_per_cu = std::min(ctx->lds_per_cu / (S * BS * 8), std::min(32, reg_waves) / waves_per_block);
    if (opts && opts->waves_per_cu > 0) blocks_per_cu = std::max(1, opts->waves_per_cu / waves_per_block);
    blocks_per_cu = std::max(blocks_per_cu, 1);
    unsigned grid = (unsigned)std::min<size_t>(cdiv(n, BS), (size_t)ctx->n_cu * blocks_per_cu);
    if (grid == 0) grid = 1;
    const uint32_t total = grid * BS;
    // worst case stack: (W-1) entries per level above the last plus W for the last; the
    // part beyond the LDS slots spills to a stream-ordered allocation owned by this call
    const int need = (b->width - 1) * b->depth + 1;
    uint32_t spill_slots = need > S ? (uint32_t)(need - S) : 0u;
    void* spill = nullptr;
    if (spill_slots) {
        size_t bytes = (size_t)spill_slots * total * sizeof(uint2);
        if (hipMallocAsync(&spill, bytes, s) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(GPURT_E_OOM, "traversal spill buffer allocation failed");
        }
    }
    a.nodes = b->nodes;
    a.tris = b->tris;
    a.root = 0;
    a.n = (uint32_t)n;
    a.spill = (uint2*)spill;
    a.spill_slots = spill_slots;
    a.stats = opts ? (unsigned long long*)opts->stats : nullptr;
    a.err = ctx->err;
    const bool stats = a.stats != nullptr;
    int bs = 0;
    switch (b->width) {
        case 2: bs = stats ? dispatch_stack<2, true>(S, kind, a, grid, s) : dispatch_stack<2, false>(S, kind, a, grid, s); break;
        case 4: bs = stats ? dispatch_stack<4, true>(S, kind, a, grid, s) : dispatch_stack<4, false>(S, kind, a, grid, s); break;
        case 8: bs = stats ? dispatch_stack<8, true>(S, kind, a, grid, s) : dispatch_stack<8, false>(S, kind, a, grid, s); break;
        default: return set_error(GPURT_E_UNSUPPORTED, "node width");
    }
    (void)bs;
    hipError_t le = hipGetLastError();
    if (spill) (void)hipFreeAsync(spill, s);
    if (le != hipSuccess) return hip_error(le, "traversal launch");
    return GPURT_OK;
}

int launch_cpq(gpurt_bvh_s* b, const float* q, size_t n, int32_t* prim, float* pt, float* dist,
               float* bary, const gpurt_query_opts* opts, hipStream_t s) {
    TravArgs a{};
    a.q = q;
    a.prim = prim;
    a.pt = pt;
    a.dist = dist;
    a.bary = bary;
    return run_query(b, QKind::Cpq, a, n, opts, s);
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

}  // namespace gpurt_impl
