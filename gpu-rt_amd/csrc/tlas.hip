// tlas.hip -- two-level instancing: a BVH over instances (a BLAS each, placed by an
// object-to-world transform), traversed in world space, with each reached instance's
// BLAS traversed in object space.
//
// Build.  An instance's world box is its BLAS's object box (the union of the root
// node's padded child boxes) with the 8 corners mapped through the transform in FP64,
// rounded outward to FP32 and padded by 2^-12 * scene scale * max(1, condition number)
// (the bound on how far the FP32 object-space ray can stray from the exact transform
// of the world ray, for origins within 2^8 scene extents).  The instance BVH is built by
// the triangle builder over one degenerate "triangle" (lo, hi, lo) per instance (whose
// box is exactly [lo, hi]) with one instance per leaf; its leaves carry the instance
// index where a triangle carries its id.
//
// Traversal (one lane per ray, the streaming loop of traverse.h extended by a level):
// at the top level the lane tests the instance BVH with the world ray; reaching an
// instance leaf it switches to the instance's object-space ray, remembers the stack
// height, and continues in the BLAS; when the BLAS part of the stack is exhausted (and
// no parked leaf of this instance is pending) it reloads the world ray and continues at
// the top level.  t is the same parameter in both spaces, so a single best-t culls both
// levels.  Leaves of the BLASes are postponed exactly as in single-level traversal.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "traverse.h"

namespace gpurt_impl {

struct alignas(16) InstDev {
    float minv[9];      // rows of the FP32 inverse of the 3x3 part
    float c[3];         // translation of the object-to-world transform
    const void* nodes;  // the BLAS's ray nodes (FP32 or 8-bit, as the TLAS)
    const float4* tris;
    uint32_t custom;
    uint32_t mask;
    uint32_t pad_[2];
};
static_assert(sizeof(InstDev) == 80, "instance descriptor size");

constexpr float kTSlackTop = 0x1p-16f;  // relative slack on the world-space ray box exits

// o' = Minv (o - c), d' = Minv d; each row summed left to right, products rounded.
__device__ __forceinline__ void to_object(const InstDev& I, V3 o, V3 d, V3& oo, V3& od) {
    const V3 x = v3(o.x - I.c[0], o.y - I.c[1], o.z - I.c[2]);
    oo = v3((I.minv[0] * x.x + I.minv[1] * x.y) + I.minv[2] * x.z,
            (I.minv[3] * x.x + I.minv[4] * x.y) + I.minv[5] * x.z,
            (I.minv[6] * x.x + I.minv[7] * x.y) + I.minv[8] * x.z);
    od = v3((I.minv[0] * d.x + I.minv[1] * d.y) + I.minv[2] * d.z,
            (I.minv[3] * d.x + I.minv[4] * d.y) + I.minv[5] * d.z,
            (I.minv[6] * d.x + I.minv[7] * d.y) + I.minv[8] * d.z);
}

struct RaySetup {
    V3 o, d, inv;
    float oix, oiy, oiz;
    __device__ __forceinline__ void set(V3 oo, V3 dd) {
        o = oo;
        d = dd;
        const float x = fabsf(d.x) < 1e-30f ? copysignf(1e-30f, d.x) : d.x;
        const float y = fabsf(d.y) < 1e-30f ? copysignf(1e-30f, d.y) : d.y;
        const float z = fabsf(d.z) < 1e-30f ? copysignf(1e-30f, d.z) : d.z;
        inv = v3(1.0f / x, 1.0f / y, 1.0f / z);
        oix = -o.x * inv.x;
        oiy = -o.y * inv.y;
        oiz = -o.z * inv.z;
    }
};

// Work donation state of the two-level kernels, after the block's S * BS stack slots:
// BS (instance, id) keys (8 bytes), BS best-t words and BS member counts (4 bytes each).
constexpr int kTlasGroupBytes = 16;
template <int S, int BS>
__device__ __forceinline__ unsigned long long* tlas_group_k() {
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x) & ~63u;
    return reinterpret_cast<unsigned long long*>(lds_stack + S * BS) + w;
}
template <int S, int BS>
__device__ __forceinline__ uint32_t* tlas_group_t() {
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x) & ~63u;
    return reinterpret_cast<uint32_t*>(reinterpret_cast<unsigned long long*>(lds_stack + S * BS) + BS) + w;
}
template <int S, int BS>
__device__ __forceinline__ uint32_t* tlas_group_count() {
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x) & ~63u;
    return reinterpret_cast<uint32_t*>(reinterpret_cast<unsigned long long*>(lds_stack + S * BS) + BS) + BS + w;
}
// t as an integer that orders like t (the high word of result_key), and back
__device__ __forceinline__ uint32_t t_order(float t) { return (uint32_t)(result_key(t, 0u) >> 32); }
__device__ __forceinline__ float t_from_order(uint32_t o) { return key_value((unsigned long long)o << 32); }

// `outside` collects (as a wave ballot) the lanes whose ray origin lies beyond `obound`
// in some coordinate: outside the domain where the instance boxes' padding keeps the
// answers exact (the kernel reports it through the context's status word)
template <typename NodeT, int S, int BS, bool ANY, bool STATS, typename Src, typename Out>
__device__ __forceinline__ bool tlas_ray_stream(const NodeT* __restrict__ top_nodes,
                                                const float4* __restrict__ top_tris,
                                                const InstDev* __restrict__ inst, Src& src, uint32_t trig,
                                                LaneStack<S, BS>& st, Out out, unsigned long long& n_nodes,
                                                unsigned long long& n_tris, float obound, uint64_t& outside) {
    constexpr int W = NodeT::kWidth;
    bool ok = true;
    typename Src::Handle h{};
    bool has = false;
    RaySetup r;
    r.set(v3(0.f, 0.f, 1.f), v3(0.f, 0.f, 1.f));
    float tmin = 0.f, best = 0.f;
    uint32_t bid = 0xFFFFFFFFu, binst = 0xFFFFFFFFu, bpos = 0, cur = 0, pleaf = kEmpty;
    bool hit_any = false, have = false, in_inst = false;
    uint32_t cinst = 0;
    int sp_base = 0;
    const NodeT* bnodes = top_nodes;
    const float4* btris = top_tris;
    // work donation (traverse.h) with the two-level tie rule (t, instance, id): members
    // share the group's best t through one LDS word (order-preserving bits; any hit: 0
    // once a member hit), and when the group is complete the members holding that t
    // merge (instance << 32 | id) through a second LDS key, whose minimum is the winner
    const uint32_t lane = __lane_id();
    uint32_t owner = lane;
    bool grouped = false, wait = false, drained = false;
    bool mine = true;  // (closest hit) binst/bid/bpos were found by this lane
    uint32_t* const gt = tlas_group_t<S, BS>();
    unsigned long long* const gk = tlas_group_k<S, BS>();
    uint32_t* const gcnt = tlas_group_count<S, BS>();
    // the trigger word: bits 0-7 the leaf trigger, bits 8-15 the refill trigger
    const uint32_t leaf_trig = trig & 0xFFu;
    const uint32_t refill_trig = max((trig >> 8) & 0xFFu, 1u);
    for (;;) {
        // leave an instance whose BLAS part of the stack is done (its parked leaf, if
        // any, must be tested first: it needs the object-space ray)
        if (has && in_inst && !have && pleaf == kEmpty && st.sp == sp_base) {
            in_inst = false;
            float4 r0, r1;
            src.ray(h, r0, r1);
            r.set(v3(r0.x, r0.y, r0.z), v3(r1.x, r1.y, r1.z));
        }
        bool done;
        if (!drained) {
            done = has && !in_inst && !have && pleaf == kEmpty && (st.sp == 0 || (ANY && hit_any));
        } else {
            // groups exist only once the source has run dry (wave-uniform branch)
            if (grouped && has && !wait) {
                const uint32_t g = gt[owner];
                if (ANY) {
                    if (g == 0u) {  // a member hit: drop both levels
                        have = false;
                        pleaf = kEmpty;
                        st.sp = 0;
                        in_inst = false;
                    }
                } else if (g < t_order(best)) {  // a member's hit is closer: cull with it
                    best = t_from_order(g);
                    bid = binst = 0xFFFFFFFFu;
                    mine = false;
                }
            }
            done = has && !wait && !in_inst && !have && pleaf == kEmpty && (st.sp == 0 || (ANY && hit_any));
            if (grouped && done) {
                if (ANY ? hit_any : bid != 0xFFFFFFFFu) atomicMin(gt + owner, ANY ? 0u : t_order(best));
                atomicSub(gcnt + owner, 1u);
                wait = true;
                done = false;
            }
        }
        if (wait && gcnt[owner] == 0u) {  // the last member has finished
            bool write, any_hit = false;
            if (ANY) {
                any_hit = gt[owner] == 0u;
                write = lane == owner;
            } else {
                const bool cand = mine && bid != 0xFFFFFFFFu && t_order(best) == gt[owner];
                const unsigned long long mk = (unsigned long long)binst << 32 | bid;
                if (cand) atomicMin(gk + owner, mk);
                const unsigned long long kf = gk[owner];
                write = cand ? mk == kf : (kf == ~0ull && lane == owner);
            }
            if (write) {
                float bu = 0.f, bv = 0.f;
                uint32_t obj = 0xFFFFFFFFu;
                if (!ANY && bid != 0xFFFFFFFFu) {
                    const InstDev& I = inst[binst];
                    V3 oo, od;
                    to_object(I, r.o, r.d, oo, od);
                    const float4 A = I.tris[3 * (size_t)bpos];
                    const float4 B = I.tris[3 * (size_t)bpos + 1];
                    const float4 C = I.tris[3 * (size_t)bpos + 2];
                    float tt;
                    (void)ray_tri(oo, od, v3(A.x, A.y, A.z), v3(B.x, B.y, B.z), v3(C.x, C.y, C.z), tt, bu, bv);
                    obj = I.custom;
                }
                out(h, obj, bid, best, bu, bv, ANY ? any_hit : hit_any);
            }
            has = wait = grouped = false;
            owner = lane;
        }
        // a finished lane keeps its slot until refill_trig lanes of the wave are idle (or
        // all are, or the source has run dry); results are then written and new rays
        // fetched for all of them at once
        const uint32_t n_idle = (uint32_t)__popcll(__ballot(!has || done));
        if (drained || n_idle >= refill_trig || n_idle == (uint32_t)__popcll(__ballot(true))) {
            if (done) {
                float bu = 0.f, bv = 0.f;
                uint32_t obj = 0xFFFFFFFFu;
                if (!ANY && bid != 0xFFFFFFFFu) {
                    // (u,v) of the winner, recomputed once with the same arithmetic
                    const InstDev& I = inst[binst];
                    V3 oo, od;
                    to_object(I, r.o, r.d, oo, od);
                    const float4 A = I.tris[3 * (size_t)bpos];
                    const float4 B = I.tris[3 * (size_t)bpos + 1];
                    const float4 C = I.tris[3 * (size_t)bpos + 2];
                    float tt;
                    (void)ray_tri(oo, od, v3(A.x, A.y, A.z), v3(B.x, B.y, B.z), v3(C.x, C.y, C.z), tt, bu, bv);
                    obj = I.custom;
                }
                out(h, obj, bid, best, bu, bv, hit_any);
                has = false;
                done = false;
            }
            if (!drained) {
                if (src.refill(has, h)) {
                    float4 r0, r1;
                    src.ray(h, r0, r1);
                    r.set(v3(r0.x, r0.y, r0.z), v3(r1.x, r1.y, r1.z));
                    tmin = r0.w;
                    best = r1.w;
                    empty_if_invalid_ray(r0, r1, tmin, best);
                    outside |= __ballot(fmaxf(fmaxf(fabsf(r0.x), fabsf(r0.y)), fabsf(r0.z)) > obound);
                    bid = binst = 0xFFFFFFFFu;
                    hit_any = false;
                    st.sp = 0;
                    sp_base = 0;
                    in_inst = false;
                    cur = 0;
                    pleaf = kEmpty;
                    have = true;
                    has = true;
                    mine = true;
                    gt[lane] = ~0u;
                    gk[lane] = ~0ull;
                    gcnt[lane] = 1u;
                }
                drained = !STATS && src.exhausted();  // counting kernels count one lane's visits
            }
            if (!__any(has)) break;
        }
        if (drained) {
            // donation: an idle lane takes the top stack entry of a lane with other work
            // left -- a BLAS entry when the donor is inside an instance (the recipient then
            // works in the same object space), else a top-level entry
            const bool idle = !has;
            const bool donor = has && !wait && !done && (!in_inst || st.sp > sp_base) &&
                               st.sp >= ((have || pleaf != kEmpty) ? 1 : 2);
            if ((uint32_t)__popcll(__ballot(idle)) >= kRayDonateIdle && __ballot(donor) != 0) {
                bool take, give;
                const uint32_t from = pair_donors(idle, donor, take, give);
                uint2 e = make_uint2(0u, 0u);
                if (give) {
                    e = st.pop();
                    if (!ANY && !grouped && bid != 0xFFFFFFFFu) atomicMin(gt + owner, t_order(best));
                    grouped = true;
                }
                e.x = __shfl(e.x, (int)from);
                e.y = __shfl(e.y, (int)from);
                owner = __shfl(owner, (int)from);
                h = Src::shfl(h, from);
                r.o = v3(__shfl(r.o.x, (int)from), __shfl(r.o.y, (int)from), __shfl(r.o.z, (int)from));
                r.d = v3(__shfl(r.d.x, (int)from), __shfl(r.d.y, (int)from), __shfl(r.d.z, (int)from));
                r.inv = v3(__shfl(r.inv.x, (int)from), __shfl(r.inv.y, (int)from), __shfl(r.inv.z, (int)from));
                r.oix = __shfl(r.oix, (int)from);
                r.oiy = __shfl(r.oiy, (int)from);
                r.oiz = __shfl(r.oiz, (int)from);
                tmin = __shfl(tmin, (int)from);
                best = __shfl(best, (int)from);
                const bool d_in = __shfl((int)in_inst, (int)from) != 0;
                const uint32_t d_inst = __shfl(cinst, (int)from);
                if (take) {
                    atomicAdd(gcnt + owner, 1u);
                    has = true;
                    grouped = true;
                    mine = true;
                    bid = binst = 0xFFFFFFFFu;
                    hit_any = false;
                    st.sp = 0;
                    sp_base = 0;
                    pleaf = kEmpty;
                    in_inst = d_in;
                    cinst = d_inst;
                    if (d_in) {
                        const InstDev& I = inst[d_inst];
                        bnodes = reinterpret_cast<const NodeT*>(I.nodes);
                        btris = I.tris;
                    }
                    cur = e.x;
                    have = __uint_as_float(e.y) <= best;
                }
            }
        }
        if (!has || done || wait) continue;
        if (!have && st.sp > (in_inst ? sp_base : 0)) {
            const uint2 e = st.pop();
            if (__uint_as_float(e.y) <= best) {
                cur = e.x;
                have = true;
            }
        }
        // an instance leaf of the top level: switch to the instance's object space
        if (have && !in_inst && child_is_leaf(cur)) {
            const uint32_t id = __float_as_uint(top_tris[3 * (size_t)leaf_first(cur)].w);
            const InstDev& I = inst[id];
            have = false;
            if (I.mask & 0xFFu) {
                V3 oo, od;
                to_object(I, r.o, r.d, oo, od);
                r.set(oo, od);
                in_inst = true;
                cinst = id;
                sp_base = st.sp;
                bnodes = reinterpret_cast<const NodeT*>(I.nodes);
                btris = I.tris;
                cur = 0;
                have = true;
            }
        }
        if (have && in_inst && pleaf == kEmpty && child_is_leaf(cur)) {
            pleaf = cur;
            have = false;
        }
        const bool node_ready = have && !child_is_leaf(cur);
        if (leaf_step_now(pleaf != kEmpty, node_ready, leaf_trig)) {
            if (pleaf == kEmpty) continue;
            const uint32_t f = leaf_first(pleaf), cnt = leaf_count(pleaf);
            bool better = false;
            for (uint32_t k = 0; k < cnt; k++) {
                const float4 A = btris[3 * (size_t)(f + k)];
                const float4 B = btris[3 * (size_t)(f + k) + 1];
                const float4 C = btris[3 * (size_t)(f + k) + 2];
                if (STATS) n_tris++;
                float t, u, v;
                if (!ray_tri(r.o, r.d, v3(A.x, A.y, A.z), v3(B.x, B.y, B.z), v3(C.x, C.y, C.z), t, u, v)) continue;
                if (!(t >= tmin && t <= best)) continue;
                if (ANY) {
                    hit_any = true;
                    break;
                }
                const uint32_t id = __float_as_uint(A.w);
                if (t < best || cinst < binst || (cinst == binst && id < bid)) {
                    best = t;
                    bid = id;
                    binst = cinst;
                    bpos = f + k;
                    mine = true;
                    better = true;
                }
            }
            pleaf = kEmpty;
            // share the lane's best with the rest of its group
            if (grouped && (ANY ? hit_any : better)) atomicMin(gt + owner, ANY ? 0u : t_order(best));
            if (ANY && hit_any) {
                // done: drop both levels; the world ray is restored before the result
                have = false;
                st.sp = sp_base;
            } else if (have && child_is_leaf(cur)) {
                pleaf = cur;  // the leaf this lane was waiting with
                have = false;
            }
            continue;
        }
        if (!node_ready) continue;
        // Node step: up to kRayNodeVisits visits in a row within one level, as in the
        // one-level kernel; the stack pop stays above the instance's stack base, and a
        // leaf child (a triangle leaf, or an instance at the top level) ends the step.
        const float slack = in_inst ? kTSlack : kTSlackTop;
        const NodeT* const lnodes = in_inst ? bnodes : top_nodes;
        const int floor_sp = in_inst ? sp_base : 0;
#pragma unroll 1
        for (int visit = 0; visit < kRayNodeVisits; visit++) {
            if (STATS) n_nodes++;
            float tn[W];
            uint32_t c[W];
            ray_children<W>(lnodes, cur, r.inv, r.oix, r.oiy, r.oiz, tmin, best, slack, tn, c);
            if (!ANY) sort_children<W>(tn, c);
            have = false;
            bool hit[W];
#pragma unroll
            for (int j = 0; j < W; j++) hit[j] = c[j] != kEmpty;
            const uint32_t pend = push_children<!ANY>(st, c, tn, hit, ok);
            if (pend != kEmpty) {
                cur = pend;
                have = true;
            } else if (st.sp > floor_sp) {
                const uint2 e = st.pop();
                if (__uint_as_float(e.y) <= best) {
                    cur = e.x;
                    have = true;
                }
            }
            if (!have || child_is_leaf(cur)) break;
        }
    }
    return ok;
}

// Resident waves per SIMD the two-level kernels are compiled for: the second level's
// state (instance, stack base, BLAS pointers) and the work-donation state do not fit 64
// VGPRs without spilling.  4-wide nodes get up to 85 (6 waves): at 72 the closest-hit
// kernel spilled the LDS stack address to scratch and every pop waited for it, and with
// two node visits per step the any-hit kernel spilled 8 VGPRs at 72 and none at 6 waves
// (65-instance rays workload: any hit 240 -> 236 us).  8-wide nodes take what they need.
template <int W>
constexpr int kTlasWaves = W == 8 ? 4 : 6;

struct TlasArgs {
    const void* top_nodes;
    const float4* top_tris;
    const InstDev* inst;
    const float4* rays;
    uint32_t n;             // rays (or, with n_dev, the capacity of the ray array)
    const uint32_t* n_dev;  // nullable: the ray count, read from device memory at launch
    int32_t* obj;
    int32_t* prim;
    float* t;
    float* uv;
    uint8_t* occ;
    uint2* spill;
    uint32_t spill_slots;
    uint32_t* counter;
    unsigned long long* stats;
    uint32_t* err;
    uint32_t leaf_trigger;
    float origin_bound;  // |origin| coordinates beyond this are outside the exact domain
};

template <int W, bool Q, int S, int BS, bool ANY, bool STATS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(kTlasWaves<W>))) void k_tlas_ray(
    TlasArgs a) {
    const uint32_t total = gridDim.x * BS;
    const uint32_t gid = blockIdx.x * BS + threadIdx.x;
    LaneStack<S, BS> st;
    st.init(a.spill, a.spill_slots, gid, total);
    const uint32_t n = a.n_dev ? min(*a.n_dev, a.n) : a.n;
    QueueSource src{a.rays, nullptr, n, first_query(gid, total), total, a.counter};
    auto out = [&](uint32_t ri, uint32_t obj, uint32_t id, float t, float u, float v, bool any) {
        if (ANY) {
            a.occ[ri] = any ? 1 : 0;
        } else {
            const bool hit = id != 0xFFFFFFFFu;
            a.obj[ri] = hit ? (int32_t)obj : -1;
            a.prim[ri] = hit ? (int32_t)id : -1;
            a.t[ri] = hit ? t : INFINITY;
            if (a.uv) {
                a.uv[2 * (size_t)ri] = hit ? u : 0.0f;
                a.uv[2 * (size_t)ri + 1] = hit ? v : 0.0f;
            }
        }
    };
    unsigned long long n_nodes = 0, n_tris = 0;
    uint64_t outside = 0;
    if (!tlas_ray_stream<NodeOf<W, Q>, S, BS, ANY, STATS>(reinterpret_cast<const NodeOf<W, Q>*>(a.top_nodes),
                                                           a.top_tris, a.inst, src, a.leaf_trigger, st, out,
                                                           n_nodes, n_tris, a.origin_bound, outside))
        gatomic_or(a.err, 1u);
    if (outside && __lane_id() == 0) gatomic_or(a.err, 2u);
    if (STATS) {
        atomicAdd(&a.stats[0], n_nodes);
        atomicAdd(&a.stats[1], n_tris);
        atomicAdd(&a.stats[2], (unsigned long long)st.spills);
    }
}

template <int W, bool Q, int S, int BS, bool STATS>
static void launch_tlas_kind(bool any, const TlasArgs& a, unsigned grid, hipStream_t s) {
    if (any)
        hipLaunchKernelGGL((k_tlas_ray<W, Q, S, BS, true, STATS>), dim3(grid), dim3(BS), (S * sizeof(uint2) + kTlasGroupBytes) * BS, s, a);
    else
        hipLaunchKernelGGL((k_tlas_ray<W, Q, S, BS, false, STATS>), dim3(grid), dim3(BS), (S * sizeof(uint2) + kTlasGroupBytes) * BS, s, a);
}

template <int W, bool Q, bool STATS>
static bool launch_tlas_stack(int S, bool any, const TlasArgs& a, unsigned grid, hipStream_t s) {
    switch (S) {
        case 8: launch_tlas_kind<W, Q, 8, 256, STATS>(any, a, grid, s); return true;
        case 16: launch_tlas_kind<W, Q, 16, 256, STATS>(any, a, grid, s); return true;
        case 32: launch_tlas_kind<W, Q, 32, 128, STATS>(any, a, grid, s); return true;
        case 64: launch_tlas_kind<W, Q, 64, 64, STATS>(any, a, grid, s); return true;
    }
    return false;
}

// FP64 inverse of the 3x3 part (cofactors over the determinant), rounded to FP32;
// false if the matrix is singular or not finite.
static bool invert3(const float* t, float* minv, double* cond) {
    const double m00 = t[0], m01 = t[1], m02 = t[2];
    const double m10 = t[4], m11 = t[5], m12 = t[6];
    const double m20 = t[8], m21 = t[9], m22 = t[10];
    const double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
    if (!(det != 0.0) || !std::isfinite(det)) return false;
    const double inv[9] = {(m11 * m22 - m12 * m21) / det, (m02 * m21 - m01 * m22) / det, (m01 * m12 - m02 * m11) / det,
                           (m12 * m20 - m10 * m22) / det, (m00 * m22 - m02 * m20) / det, (m02 * m10 - m00 * m12) / det,
                           (m10 * m21 - m11 * m20) / det, (m01 * m20 - m00 * m21) / det, (m00 * m11 - m01 * m10) / det};
    double fm = 0.0, fi = 0.0;
    const double mm[9] = {m00, m01, m02, m10, m11, m12, m20, m21, m22};
    for (int k = 0; k < 9; k++) {
        if (!std::isfinite(inv[k])) return false;
        minv[k] = (float)inv[k];
        fm += mm[k] * mm[k];
        fi += inv[k] * inv[k];
    }
    *cond = std::sqrt(fm) * std::sqrt(fi);
    return true;
}

}  // namespace gpurt_impl

using namespace gpurt_impl;

struct gpurt_tlas_s {
    gpurt_ctx ctx = nullptr;
    gpurt_bvh_s* top = nullptr;  // the instance BVH (owned)
    InstDev* inst_d = nullptr;
    size_t n_inst = 0;
    int width = 4;
    bool q8 = false;
    int max_blas_depth = 0;
    float origin_bound = INFINITY;      // 2^8 x the largest |world box coordinate|
    std::vector<uint32_t> custom;       // custom index of every instance (host copy)
    std::vector<gpurt_bvh_s*> blas;     // BLAS of every instance
    // frees in stream order on the context's release stream (see gpurt_tlas_destroy);
    // a TLAS whose context is gone (tlas_orphan) holds no device memory
    ~gpurt_tlas_s() {
        if (inst_d && ctx) {
            if (top) wait_uses(top, ctx->release);
            (void)hipFreeAsync(inst_d, ctx->release);
        }
        if (top) gpurt_bvh_destroy(top);
    }
};

namespace {

struct TlasDevGuard {
    int prev = -1;
    bool ok = false;
    explicit TlasDevGuard(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        ok = hipSetDevice(d) == hipSuccess;
    }
    ~TlasDevGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Object box of a BLAS: the union of its root node's (padded) child boxes.
int blas_object_box(gpurt_bvh_s* b, double lo[3], double hi[3]) {
    const int W = b->width;
    std::vector<float> node(8 * W);
    GPURT_HIP(hipMemcpy(node.data(), b->nodes, node_bytes(W), hipMemcpyDeviceToHost));
    for (int a = 0; a < 3; a++) {
        lo[a] = INFINITY;
        hi[a] = -INFINITY;
    }
    bool any = false;
    for (int j = 0; j < W; j++) {
        uint32_t c;
        std::memcpy(&c, &node[6 * W + j], 4);
        if (c == kEmpty) continue;
        any = true;
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], (double)node[a * W + j]);
            hi[a] = std::max(hi[a], (double)node[(3 + a) * W + j]);
        }
    }
    if (!any) return set_error(GPURT_E_INVALID, "empty BLAS");
    return GPURT_OK;
}

}  // namespace

namespace gpurt_impl {

int tlas_host_info(gpurt_tlas t, gpurt_ctx* ctx, const std::vector<uint32_t>** custom,
                   const std::vector<gpurt_bvh_s*>** blas) {
    if (!t) return set_error(GPURT_E_INVALID, "null tlas");
    *ctx = t->ctx;
    *custom = &t->custom;
    *blas = &t->blas;
    return GPURT_OK;
}

}  // namespace gpurt_impl

extern "C" {

int gpurt_tlas_build(gpurt_ctx ctx, const gpurt_instance* instances, size_t n, const gpurt_build_opts* opts,
                     void* stream, gpurt_tlas* out) {
    if (!out) return set_error(GPURT_E_INVALID, "out is null");
    *out = nullptr;
    if (!ctx || !instances) return set_error(GPURT_E_INVALID, "null context or instance array");
    if (n == 0 || n > ((size_t)1 << 24)) return set_error(GPURT_E_INVALID, "instance count must be 1..2^24");
    BuildParams bp;
    int rc = parse_build_opts(opts, bp);
    if (rc) return rc;
    const gpurt_bvh_s* b0 = instances[0].blas;
    if (!b0) return set_error(GPURT_E_INVALID, "null BLAS (instance 0)");
    const int W = b0->width;
    const bool q8 = rays_use_q8(b0->format);
    TlasDevGuard dg(ctx->device);
    if (!dg.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    hipStream_t s = (hipStream_t)stream;
    GPURT_HIP(hipStreamSynchronize(s));
    // per instance: inverse, world box (FP64 corners), condition number
    std::vector<InstDev> dev(n);
    std::vector<double> box(6 * n);
    std::map<gpurt_bvh_s*, std::vector<double>> obj_boxes;
    double scale = 0.0, cond_max = 1.0;
    int max_depth = 0;
    for (size_t i = 0; i < n; i++) {
        const gpurt_instance& in = instances[i];
        const std::string at = " (instance " + std::to_string(i) + ")";
        gpurt_bvh_s* b = in.blas;
        if (!b) return set_error(GPURT_E_INVALID, "null BLAS" + at);
        if (b->ctx != ctx) return set_error(GPURT_E_INVALID, "BLAS of another context" + at);
        if (b->width != W || rays_use_q8(b->format) != q8)
            return set_error(GPURT_E_UNSUPPORTED, "all BLASes need the same node width and ray node format" + at);
        for (int k = 0; k < 12; k++)
            if (!std::isfinite(in.transform[k])) return set_error(GPURT_E_INVALID, "non-finite transform" + at);
        InstDev& d = dev[i];
        double cond = 1.0;
        if (!invert3(in.transform, d.minv, &cond)) return set_error(GPURT_E_INVALID, "singular transform" + at);
        cond_max = std::max(cond_max, cond);
        d.c[0] = in.transform[3];
        d.c[1] = in.transform[7];
        d.c[2] = in.transform[11];
        d.nodes = q8 ? b->qnodes : b->nodes;
        d.tris = b->tris;
        d.custom = in.custom_index;
        d.mask = in.mask;
        d.pad_[0] = d.pad_[1] = 0;
        max_depth = std::max(max_depth, b->depth);
        auto it = obj_boxes.find(b);
        if (it == obj_boxes.end()) {
            std::vector<double> ob(6);
            if ((rc = blas_object_box(b, ob.data(), ob.data() + 3))) return rc;
            it = obj_boxes.emplace(b, ob).first;
        }
        const std::vector<double>& ob = it->second;
        double wlo[3] = {INFINITY, INFINITY, INFINITY}, whi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 8; k++) {
            const double p[3] = {(k & 1) ? ob[3] : ob[0], (k & 2) ? ob[4] : ob[1], (k & 4) ? ob[5] : ob[2]};
            for (int a = 0; a < 3; a++) {
                const double w = (double)in.transform[4 * a] * p[0] + (double)in.transform[4 * a + 1] * p[1] +
                                 (double)in.transform[4 * a + 2] * p[2] + (double)in.transform[4 * a + 3];
                wlo[a] = std::min(wlo[a], w);
                whi[a] = std::max(whi[a], w);
            }
        }
        for (int a = 0; a < 3; a++) {
            box[6 * i + a] = wlo[a];
            box[6 * i + 3 + a] = whi[a];
            scale = std::max(scale, std::max(std::fabs(wlo[a]), std::fabs(whi[a])));
        }
    }
    // pad, round outward, and lay out one (lo, hi, lo) "triangle" per instance
    const double pad = std::ldexp(std::max(scale, 1e-30), -12) * cond_max;
    std::vector<float> verts(9 * n);
    std::vector<uint32_t> idx(3 * n);
    for (size_t i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) {
            const float lo = std::nextafter((float)(box[6 * i + a] - pad), -INFINITY);
            const float hi = std::nextafter((float)(box[6 * i + 3 + a] + pad), INFINITY);
            verts[9 * i + a] = lo;
            verts[9 * i + 3 + a] = hi;
            verts[9 * i + 6 + a] = lo;
        }
        for (int k = 0; k < 3; k++) idx[3 * i + k] = (uint32_t)(3 * i + k);
    }
    auto* T = new gpurt_tlas_s();
    T->ctx = ctx;
    T->n_inst = n;
    T->width = W;
    T->q8 = q8;
    T->max_blas_depth = max_depth;
    T->origin_bound = (float)std::ldexp(std::max(scale, 1e-30), 8);
    T->custom.resize(n);
    T->blas.resize(n);
    for (size_t i = 0; i < n; i++) {
        T->custom[i] = instances[i].custom_index;
        T->blas[i] = instances[i].blas;
    }
    // stream-ordered allocations on the build stream; the instance table lives with the
    // TLAS and is freed in stream order by gpurt_tlas_destroy
    float* vd = nullptr;
    uint32_t* id = nullptr;
    rc = GPURT_OK;
    if (hipMallocAsync((void**)&vd, sizeof(float) * verts.size(), s) != hipSuccess ||
        hipMallocAsync((void**)&id, sizeof(uint32_t) * idx.size(), s) != hipSuccess ||
        hipMallocAsync((void**)&T->inst_d, sizeof(InstDev) * n, s) != hipSuccess) {
        (void)hipGetLastError();
        rc = set_error(GPURT_E_OOM, "TLAS allocation failed");
    } else if (hipMemcpyAsync(vd, verts.data(), sizeof(float) * verts.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
               hipMemcpyAsync(id, idx.data(), sizeof(uint32_t) * idx.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
               hipMemcpyAsync(T->inst_d, dev.data(), sizeof(InstDev) * n, hipMemcpyHostToDevice, s) != hipSuccess) {
        rc = hip_error(hipGetLastError(), "TLAS upload");
    }
    if (rc == GPURT_OK) {
        gpurt_build_opts to{};
        to.node_width = W;
        to.leaf_size = 1;
        to.treelet_size = bp.treelet_size;
        to.treelet_rounds = bp.treelet_rounds;
        to.node_format = q8 ? kFormatQ8 : kFormatF32;
        rc = gpurt_bvh_build(ctx, vd, 12, 3 * n, id, n, &to, stream, &T->top);
    }
    if (vd) (void)hipFreeAsync(vd, s);
    if (id) (void)hipFreeAsync(id, s);
    // the host arrays above feed asynchronous copies: wait for them on every path
    if (hipStreamSynchronize(s) != hipSuccess && rc == GPURT_OK) rc = hip_error(hipGetLastError(), "TLAS build");
    if (rc) {
        delete T;
        return rc;
    }
    {
        std::lock_guard<std::mutex> lk(ctx->live_mu);
        ctx->tlases.insert(T);
    }
    *out = T;
    return GPURT_OK;
}

void gpurt_tlas_destroy(gpurt_tlas t) {
    if (!t) return;
    if (!t->ctx) {  // its context was destroyed first and released its device memory
        delete t;
        return;
    }
    {
        std::lock_guard<std::mutex> lk(t->ctx->live_mu);
        t->ctx->tlases.erase(t);
    }
    TlasDevGuard dg(t->ctx->device);
    // The TLAS's queries read its BLASes: the release stream waits for them and then
    // counts as a use of every BLAS, so a BLAS destroyed afterwards is freed after them.
    if (t->top) {
        hipStream_t rs = t->ctx->release;
        wait_uses(t->top, rs);
        bool lost;
        {
            std::lock_guard<std::mutex> lk(t->top->use_mu);
            lost = t->top->uses_lost;
        }
        std::vector<gpurt_bvh_s*> seen;
        for (gpurt_bvh_s* b : t->blas) {
            if (std::find(seen.begin(), seen.end(), b) != seen.end()) continue;
            seen.push_back(b);
            note_use(b, rs);
            if (lost) {
                std::lock_guard<std::mutex> lk(b->use_mu);
                b->uses_lost = true;
            }
        }
    }
    delete t;
}

}  // extern "C"

namespace gpurt_impl {

// Defaults of the two-level kernels (opts fields left 0): every wave slot the register
// budget allows (kTlasWaves), a leaf trigger of 32 lanes, and a refill when 24 lanes of a
// closest-hit wave or 8 of an any-hit wave are idle (swept on the 65-instance rays
// workload with work donation: closest 2.58 -> 2.70, any hit 4.17 -> 4.29 G rays/s
// against 16 and 1).
constexpr int kTlasDefaultTrigger = 32;

static int tlas_waves(int W) { return W == 8 ? kTlasWaves<8> : kTlasWaves<4>; }

static gpurt_query_opts tlas_opts(const gpurt_query_opts* opts, int W, bool any) {
    gpurt_query_opts o = opts ? *opts : gpurt_query_opts{};
    if (o.waves_per_cu == 0) o.waves_per_cu = 4 * tlas_waves(W);
    if (o.leaf_trigger == 0) o.leaf_trigger = kTlasDefaultTrigger;
    if (o.refill_trigger == 0) o.refill_trigger = any ? 8 : 24;
    return o;
}

static int tlas_plan(gpurt_tlas T, size_t n, bool any, const gpurt_query_opts* opts_in, TravPlan& p) {
    const gpurt_query_opts o = tlas_opts(opts_in, T->width, any);
    const gpurt_query_opts* opts = &o;
    const int W = T->width;
    // worst-case stack: the top levels' entries plus one BLAS's
    const int need = (W - 1) * (T->top->depth + T->max_blas_depth) + 2;
    return plan_traversal(T->ctx, W, T->q8, any ? QKind::Any : QKind::Closest, need, opts, n, p,
                          tlas_waves(W), kTlasGroupBytes);
}

size_t tlas_scratch_bytes(gpurt_tlas T, size_t n) {
    size_t bytes = 0;
    for (bool any : {false, true}) {
        TravPlan p;
        if (tlas_plan(T, n, any, nullptr, p)) return 0;
        bytes = std::max(bytes, traversal_scratch_bytes(p));
    }
    return bytes;
}

void tlas_note_use(gpurt_tlas t, hipStream_t s) {
    if (t && t->top) note_use(t->top, s);
}

void tlas_orphan(gpurt_tlas t) {
    if (t->inst_d) (void)hipFreeAsync(t->inst_d, t->ctx->release);
    t->inst_d = nullptr;
    t->ctx = nullptr;
}

int tlas_trace_ex(gpurt_tlas T, const gpurt_ray* rays, size_t n, const uint32_t* n_dev, int32_t* obj,
                  int32_t* prim, float* t, float* uv, uint8_t* occ, bool any, const gpurt_query_opts* opts,
                  hipStream_t stream, void* scratch, size_t scratch_bytes) {
    if (!T) return set_error(GPURT_E_INVALID, "null tlas");
    if (n == 0) return GPURT_OK;
    if (!rays || (any ? !occ : (!obj || !prim || !t))) return set_error(GPURT_E_INVALID, "null ray/output pointer");
    if (opts && (opts->stack_depth < 0 || opts->waves_per_cu < 0 || opts->waves_per_cu > 64 ||
                 opts->leaf_trigger < 0 || opts->leaf_trigger > 64 || opts->refill_trigger < 0 ||
                 opts->refill_trigger > 64))
        return set_error(GPURT_E_UNSUPPORTED, "query options out of range");
    TlasDevGuard dg(T->ctx->device);
    if (!dg.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    hipStream_t s = stream;
    const int W = T->width;
    TravPlan p;
    int rc = tlas_plan(T, n, any, opts, p);
    if (rc) return rc;
    TravScratch sc;
    if (scratch) {  // the caller's buffer: clear the queue counters, keep the memory
        if (scratch_bytes < traversal_scratch_bytes(p))
            return set_error(GPURT_E_INVALID, "traversal scratch buffer too small");
        carve_traversal_scratch(p, scratch, sc);
        clear_traversal_counters(sc, s);
    } else if ((rc = alloc_traversal_scratch(T->ctx, p, s, sc))) {
        return rc;
    }
    TlasArgs a{};
    a.top_nodes = T->q8 ? T->top->qnodes : T->top->nodes;
    a.top_tris = T->top->tris;
    a.inst = T->inst_d;
    a.rays = reinterpret_cast<const float4*>(rays);
    a.n = (uint32_t)n;
    a.n_dev = n_dev;
    a.obj = obj;
    a.prim = prim;
    a.t = t;
    a.uv = uv;
    a.occ = occ;
    a.spill = sc.spill;
    a.spill_slots = p.spill_slots;
    a.counter = sc.counters;
    a.stats = opts ? (unsigned long long*)opts->stats : nullptr;
    a.err = T->ctx->err;
    a.leaf_trigger = p.leaf_trigger;
    a.origin_bound = T->origin_bound;
    const bool stats = a.stats != nullptr;
    bool launched = false;
    switch (W * 4 + (T->q8 ? 2 : 0) + (stats ? 1 : 0)) {
        case 8: launched = launch_tlas_stack<2, false, false>(p.S, any, a, p.grid, s); break;
        case 9: launched = launch_tlas_stack<2, false, true>(p.S, any, a, p.grid, s); break;
        case 10: launched = launch_tlas_stack<2, true, false>(p.S, any, a, p.grid, s); break;
        case 11: launched = launch_tlas_stack<2, true, true>(p.S, any, a, p.grid, s); break;
        case 16: launched = launch_tlas_stack<4, false, false>(p.S, any, a, p.grid, s); break;
        case 17: launched = launch_tlas_stack<4, false, true>(p.S, any, a, p.grid, s); break;
        case 18: launched = launch_tlas_stack<4, true, false>(p.S, any, a, p.grid, s); break;
        case 19: launched = launch_tlas_stack<4, true, true>(p.S, any, a, p.grid, s); break;
        case 32: launched = launch_tlas_stack<8, false, false>(p.S, any, a, p.grid, s); break;
        case 33: launched = launch_tlas_stack<8, false, true>(p.S, any, a, p.grid, s); break;
        case 34: launched = launch_tlas_stack<8, true, false>(p.S, any, a, p.grid, s); break;
        case 35: launched = launch_tlas_stack<8, true, true>(p.S, any, a, p.grid, s); break;
        default: break;
    }
    hipError_t le = hipGetLastError();
    if (!scratch) free_traversal_scratch(T->ctx, sc);
    if (!launched) return set_error(GPURT_E_UNSUPPORTED, "node width / stack depth");
    if (le != hipSuccess) return hip_error(le, "TLAS traversal launch");
    return GPURT_OK;
}

}  // namespace gpurt_impl

extern "C" {

static int tlas_trace(gpurt_tlas T, const gpurt_ray* rays, size_t n, int32_t* obj, int32_t* prim, float* t,
                      float* uv, uint8_t* occ, bool any, const gpurt_query_opts* opts, void* stream) {
    const int rc = tlas_trace_ex(T, rays, n, nullptr, obj, prim, t, uv, occ, any, opts, (hipStream_t)stream, nullptr, 0);
    if (T && n) {
        TlasDevGuard dg(T->ctx->device);
        note_use(T->top, (hipStream_t)stream);
    }
    return rc;
}

int gpurt_tlas_trace_closest(gpurt_tlas T, const gpurt_ray* rays, size_t n, int32_t* obj, int32_t* prim, float* t,
                             float* uv, const gpurt_query_opts* opts, void* stream) {
    return tlas_trace(T, rays, n, obj, prim, t, uv, nullptr, false, opts, stream);
}

int gpurt_tlas_trace_any(gpurt_tlas T, const gpurt_ray* rays, size_t n, uint8_t* occluded,
                         const gpurt_query_opts* opts, void* stream) {
    return tlas_trace(T, rays, n, nullptr, nullptr, nullptr, nullptr, occluded, true, opts, stream);
}

}  // extern "C"
