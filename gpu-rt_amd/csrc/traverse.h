// traverse.h -- per-query BVH traversal shared by the plain kernels (query.hip) and the
// persistent batch kernel (batch.hip).
//
// One lane per query.  A lane keeps a stack of (node-or-leaf reference, lower bound)
// pairs: the first S entries live in LDS laid out [slot][lane] (consecutive lanes hit
// consecutive banks at every depth), deeper entries spill to a device buffer that the
// launcher sizes from the tree depth, so no traversal can overflow.  A wide node is
// fetched as 2*W float4 loads (W=4: one 128-byte line); children are bounded, sorted
// by the bound with a sorting network, and pushed far-to-near so the nearest is taken
// next without a stack round trip.  Culls compare against the best value with a small
// relative slack and boxes are inflated at build time, so a child that can hold an
// equal-or-better primitive is never culled: results do not depend on the tree's
// shape, and ties resolve to the lowest primitive id.
#pragma once

#include "gpurt_internal.h"
#include "tri_math.h"

namespace gpurt_impl {

// Per-lane traversal stack: slots [0, S) in LDS at lds_stack[slot * BS + lane]
// (dynamic shared memory of S * BS * 8 bytes, indexed directly so the compiler keeps
// LDS addressing); deeper slots in the device spill buffer, lane-contiguous (a lane's
// slot S + k at byte (gid * spill_slots + k) * 8), accessed with cached buffer
// instructions: a lane's consecutive spills share cache lines, the hardware range check
// of the buffer descriptor bounds every access, and the buffer intrinsics cannot be
// merged with the LDS side into one flat access.
extern __shared__ uint2 lds_stack[];

constexpr int kBufferDword3 = 0x00020000;  // gfx9 raw buffer descriptor word 3

__device__ __forceinline__ __amdgpu_buffer_rsrc_t spill_rsrc(void* spill, uint32_t slots, uint32_t total) {
    const uint64_t bytes = (uint64_t)slots * total * sizeof(uint2);
    return __builtin_amdgcn_make_buffer_rsrc(spill, 0, (int)(bytes > 0xFFFFFFF0ull ? 0xFFFFFFF0ull : bytes),
                                             kBufferDword3);
}

template <int S, int BS>
struct LaneStack {
    __amdgpu_buffer_rsrc_t spill;
    uint32_t lane, base, cap;  // base: byte offset of this lane's spill slots
    int sp;
    uint32_t spills;  // pushes that went to the spill buffer (reported by counting kernels)
    __device__ __forceinline__ void init(void* spill_buf, uint32_t spill_slots, uint32_t gid, uint32_t total) {
        spill = spill_rsrc(spill_buf, spill_slots, total);
        lane = threadIdx.x;
        base = gid * spill_slots * (uint32_t)sizeof(uint2);
        cap = S + spill_slots;
        sp = 0;
        spills = 0;
    }
    __device__ __forceinline__ bool push(uint32_t ref, float key) {
        if ((uint32_t)sp >= cap) return false;
        if (sp < S) {
            lds_stack[sp * BS + lane] = make_uint2(ref, __float_as_uint(key));
        } else {
            const uint32_t off = base + (uint32_t)(sp - S) * (uint32_t)sizeof(uint2);
            spills++;
            __builtin_amdgcn_raw_buffer_store_b32(ref, spill, off, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(key), spill, off + 4, 0, 0);
        }
        sp++;
        return true;
    }
    __device__ __forceinline__ uint2 pop() {
        sp--;
        if (sp < S) return lds_stack[sp * BS + lane];
        const uint32_t off = base + (uint32_t)(sp - S) * (uint32_t)sizeof(uint2);
        return make_uint2(__builtin_amdgcn_raw_buffer_load_b32(spill, off, 0, 0),
                          __builtin_amdgcn_raw_buffer_load_b32(spill, off + 4, 0, 0));
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
constexpr float kTSlack = 0x1p-20f;             // relative slack on ray box exits

struct CpqHit {
    uint32_t id;  // 0xFFFFFFFF: none
    float d2;
    V3 cp, bary;
};

// Closest point of q on the subtree at `root`.  Returns false on stack overflow.
template <int W, int S, int BS, bool STATS>
__device__ __forceinline__ bool cpq_one(const WNode<W>* __restrict__ nodes,
                                        const float4* __restrict__ tris, uint32_t root, V3 q,
                                        LaneStack<S, BS>& st, CpqHit& h,
                                        unsigned long long& n_nodes, unsigned long long& n_tris) {
    float best = INFINITY, thr = INFINITY;
    uint32_t bid = 0xFFFFFFFFu;
    V3 bcp = v3(0.f, 0.f, 0.f), bb = v3(0.f, 0.f, 0.f);
    st.sp = 0;
    uint32_t cur = root;
    bool have = true;
    bool ok = true;
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
                float d2 = closest_pt_tri(q, v3(A.x, A.y, A.z), v3(B.x, B.y, B.z), v3(C.x, C.y, C.z),
                                          cp, br);
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
                if (!st.push(c[j], d[j])) ok = false;
            }
        }
        if (c[0] != kEmpty && d[0] <= thr) {
            cur = c[0];
            have = true;
        }
    }
    h.id = bid;
    h.d2 = best;
    h.cp = bcp;
    h.bary = bb;
    return ok;
}

struct RayHit {
    uint32_t id;  // 0xFFFFFFFF: none
    float t, u, v;
    bool any;
};

// Closest (ANY=false) or any (ANY=true) hit of the ray on the subtree at `root`.
// Returns false on stack overflow.
template <int W, int S, int BS, bool ANY, bool STATS>
__device__ __forceinline__ bool ray_one(const WNode<W>* __restrict__ nodes,
                                        const float4* __restrict__ tris, uint32_t root, V3 o, V3 d,
                                        float tmin, float tmax, LaneStack<S, BS>& st, RayHit& h,
                                        unsigned long long& n_nodes, unsigned long long& n_tris) {
    float best = tmax;  // shrinks to the best hit
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
    const float oix = -o.x * inv.x, oiy = -o.y * inv.y, oiz = -o.z * inv.z;
    st.sp = 0;
    uint32_t cur = root;
    bool have = true;
    bool ok = true;
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
                if (!ray_tri(o, d, v3(A.x, A.y, A.z), v3(B.x, B.y, B.z), v3(C.x, C.y, C.z), t, u, v))
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
            const float t0x = fmaf(nd.lox[j], inv.x, oix), t1x = fmaf(nd.hix[j], inv.x, oix);
            const float t0y = fmaf(nd.loy[j], inv.y, oiy), t1y = fmaf(nd.hiy[j], inv.y, oiy);
            const float t0z = fmaf(nd.loz[j], inv.z, oiz), t1z = fmaf(nd.hiz[j], inv.z, oiz);
            float lo = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
            float hi = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), best));
            c[j] = nd.child[j];
            bool hb = (c[j] != kEmpty) && (lo <= hi + fabsf(hi) * kTSlack);
            tn[j] = hb ? lo : INFINITY;
            if (!hb) c[j] = kEmpty;
        }
        if (!ANY) sort_children<W>(tn, c);
        // push all but the nearest hit child (far to near), then continue with it; the
        // pending entry is carried in scalars so every index stays static
        uint32_t pend = kEmpty;
        float pend_t = 0.0f;
#pragma unroll
        for (int j = W - 1; j >= 0; j--) {
            if (c[j] == kEmpty) continue;
            if (pend != kEmpty) {
                if (!st.push(pend, pend_t)) ok = false;
            }
            pend = c[j];
            pend_t = tn[j];
        }
        if (pend != kEmpty) {
            cur = pend;
            have = true;
        }
    }
    h.id = bid;
    h.t = best;
    h.u = bu;
    h.v = bv;
    h.any = hit_any;
    return ok;
}

}  // namespace gpurt_impl

namespace gpurt_impl {

// ------------------------------------------------------------------------------------
// Streaming traversal: lane-level refill
//
// A lane runs one traversal step per loop iteration; when its query finishes it writes
// the result and loads its next query (index += stride) in the same iteration, so lanes
// of a wave do not wait for the wave's slowest query before taking new work (the
// per-query loop above keeps a finished lane idle until all 64 finish).  Every query
// ends (each node is visited at most once), a lane leaves when its indices run out,
// and the loop ends when no lane of the wave is active.
// ------------------------------------------------------------------------------------

template <int W, int S, int BS, bool ANY, bool STATS, typename Out>
__device__ __forceinline__ bool ray_stream(const WNode<W>* __restrict__ nodes,
                                           const float4* __restrict__ tris, uint32_t root,
                                           const float4* __restrict__ rays, uint32_t n, uint32_t first,
                                           uint32_t stride, LaneStack<S, BS>& st, Out out,
                                           unsigned long long& n_nodes, unsigned long long& n_tris) {
    bool ok = true;
    uint32_t ri = first;
    bool active = ri < n;
    V3 o = v3(0.f, 0.f, 0.f), d = o, inv = o;
    float oix = 0.f, oiy = 0.f, oiz = 0.f;
    float tmin = 0.f, best = 0.f, bu = 0.f, bv = 0.f;
    uint32_t bid = 0xFFFFFFFFu, cur = root;
    bool hit_any = false, have = false;
    auto start = [&](uint32_t i) {
        const float4 r0 = rays[2 * (size_t)i];
        const float4 r1 = rays[2 * (size_t)i + 1];
        o = v3(r0.x, r0.y, r0.z);
        d = v3(r1.x, r1.y, r1.z);
        tmin = r0.w;
        best = r1.w;
        const float x = fabsf(d.x) < 1e-30f ? copysignf(1e-30f, d.x) : d.x;
        const float y = fabsf(d.y) < 1e-30f ? copysignf(1e-30f, d.y) : d.y;
        const float z = fabsf(d.z) < 1e-30f ? copysignf(1e-30f, d.z) : d.z;
        inv = v3(1.0f / x, 1.0f / y, 1.0f / z);
        oix = -o.x * inv.x;
        oiy = -o.y * inv.y;
        oiz = -o.z * inv.z;
        bid = 0xFFFFFFFFu;
        bu = bv = 0.f;
        hit_any = false;
        st.sp = 0;
        cur = root;
        have = true;
    };
    if (active) start(ri);
    while (__any(active)) {
        if (!active) continue;
        if (!have) {
            if (st.sp == 0 || (ANY && hit_any)) {
                out(ri, bid, best, bu, bv, hit_any);
                ri += stride;
                if (ri < n) start(ri);
                else active = false;
                continue;
            }
            const uint2 e = st.pop();
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
                if (!ray_tri(o, d, v3(A.x, A.y, A.z), v3(B.x, B.y, B.z), v3(C.x, C.y, C.z), t, u, v)) continue;
                if (!(t >= tmin && t <= best)) continue;
                if (ANY) {
                    hit_any = true;
                    break;
                }
                const uint32_t id = __float_as_uint(A.w);
                if (t < best || id < bid) {
                    best = t;
                    bid = id;
                    bu = u;
                    bv = v;
                }
            }
            continue;
        }
        if (STATS) n_nodes++;
        const WNode<W> nd = nodes[cur];
        float tn[W];
        uint32_t c[W];
#pragma unroll
        for (int j = 0; j < W; j++) {
            // t = (b - o) / d as one FMA with o/d precomputed; its rounding is far below
            // the build-time box inflation, so the test stays conservative
            const float t0x = fmaf(nd.lox[j], inv.x, oix), t1x = fmaf(nd.hix[j], inv.x, oix);
            const float t0y = fmaf(nd.loy[j], inv.y, oiy), t1y = fmaf(nd.hiy[j], inv.y, oiy);
            const float t0z = fmaf(nd.loz[j], inv.z, oiz), t1z = fmaf(nd.hiz[j], inv.z, oiz);
            const float lo = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
            const float hi = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), best));
            c[j] = nd.child[j];
            const bool hb = (c[j] != kEmpty) && (lo <= hi + fabsf(hi) * kTSlack);
            tn[j] = hb ? lo : INFINITY;
            if (!hb) c[j] = kEmpty;
        }
        if (!ANY) sort_children<W>(tn, c);
        uint32_t pend = kEmpty;
        float pend_t = 0.0f;
#pragma unroll
        for (int j = W - 1; j >= 0; j--) {
            if (c[j] == kEmpty) continue;
            if (pend != kEmpty) {
                if (!st.push(pend, pend_t)) ok = false;
            }
            pend = c[j];
            pend_t = tn[j];
        }
        if (pend != kEmpty) {
            cur = pend;
            have = true;
        }
    }
    return ok;
}

template <int W, int S, int BS, bool STATS, typename Out>
__device__ __forceinline__ bool cpq_stream(const WNode<W>* __restrict__ nodes,
                                           const float4* __restrict__ tris, uint32_t root,
                                           const float* __restrict__ qs, uint32_t n, uint32_t first,
                                           uint32_t stride, LaneStack<S, BS>& st, Out out,
                                           unsigned long long& n_nodes, unsigned long long& n_tris) {
    bool ok = true;
    uint32_t qi = first;
    bool active = qi < n;
    V3 q = v3(0.f, 0.f, 0.f), bcp = q, bb = q;
    float best = INFINITY, thr = INFINITY;
    uint32_t bid = 0xFFFFFFFFu, cur = root;
    bool have = false;
    auto start = [&](uint32_t i) {
        q = v3(qs[3 * (size_t)i], qs[3 * (size_t)i + 1], qs[3 * (size_t)i + 2]);
        best = thr = INFINITY;
        bid = 0xFFFFFFFFu;
        bcp = bb = v3(0.f, 0.f, 0.f);
        st.sp = 0;
        cur = root;
        have = true;
    };
    if (active) start(qi);
    while (__any(active)) {
        if (!active) continue;
        if (!have) {
            if (st.sp == 0) {
                out(qi, bid, best, bcp, bb);
                qi += stride;
                if (qi < n) start(qi);
                else active = false;
                continue;
            }
            const uint2 e = st.pop();
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
                const float d2 = closest_pt_tri(q, v3(A.x, A.y, A.z), v3(B.x, B.y, B.z), v3(C.x, C.y, C.z), cp, br);
                const uint32_t id = __float_as_uint(A.w);
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
        float dd[W];
        uint32_t c[W];
#pragma unroll
        for (int j = 0; j < W; j++) {
            const float dx = fmaxf(fmaxf(nd.lox[j] - q.x, q.x - nd.hix[j]), 0.0f);
            const float dy = fmaxf(fmaxf(nd.loy[j] - q.y, q.y - nd.hiy[j]), 0.0f);
            const float dz = fmaxf(fmaxf(nd.loz[j] - q.z, q.z - nd.hiz[j]), 0.0f);
            c[j] = nd.child[j];
            dd[j] = (c[j] == kEmpty) ? INFINITY : fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        }
        sort_children<W>(dd, c);
#pragma unroll
        for (int j = W - 1; j >= 1; j--) {
            if (c[j] != kEmpty && dd[j] <= thr) {
                if (!st.push(c[j], dd[j])) ok = false;
            }
        }
        if (c[0] != kEmpty && dd[0] <= thr) {
            cur = c[0];
            have = true;
        }
    }
    return ok;
}

}  // namespace gpurt_impl
