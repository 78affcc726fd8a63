// traverse.h -- BVH traversal shared by the plain kernels (query.hip) and the persistent
// batch kernels (batch.hip).
//
// One lane per query while the work queues last; at the end of a launch the lanes left
// without a query help the busy ones (work donation, below).  A lane keeps a stack of
// (node-or-leaf reference, lower bound) pairs: the first S entries live in LDS laid out
// [slot][lane] (consecutive lanes hit consecutive banks at every depth), deeper entries
// spill to a device buffer that the launcher sizes from the tree depth, so no traversal
// can overflow.  A wide node is fetched with unconditional dwordx4 loads (FP32 boxes: 7
// for W=4; 8-bit boxes: 4); children are bounded, sorted by the bound with a sorting
// network, and pushed far-to-near so the nearest is taken next without a stack round
// trip.  Culls compare against the best value with a small relative slack and boxes are
// inflated at build time, so a child that can hold an equal-or-better primitive is never
// culled: results do not depend on the tree's shape, and ties resolve to the lowest
// primitive id.
#pragma once

#include "gpurt_internal.h"
#include "tri_math.h"

#include <type_traits>

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
    // Only sp (and the spill count of counting kernels) is per lane: the wave's LDS
    // column and spill base are wave-uniform (scalar registers), and the lane's own
    // offsets are rebuilt from its lane id when needed, so no vector register holds an
    // address for the whole traversal.
    __amdgpu_buffer_rsrc_t spill;
    uint32_t wave_lds;    // first LDS column of this wave (wave index in block * 64)
    uint32_t wave_gid;    // global lane index of this wave's lane 0
    uint32_t slots, cap;  // spill slots per lane; LDS + spill slots per lane
    int sp;
    uint32_t spills;  // pushes that went to the spill buffer (reported by counting kernels)
    __device__ __forceinline__ void init(void* spill_buf, uint32_t spill_slots, uint32_t gid, uint32_t total) {
        (void)gid;
        spill = spill_rsrc(spill_buf, spill_slots, total);
        const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x) & ~63u;
        wave_lds = w;
        wave_gid = blockIdx.x * BS + w;
        slots = spill_slots;
        cap = S + spill_slots;
        sp = 0;
        spills = 0;
    }
    __device__ __forceinline__ uint32_t lds_index(int slot) const {
        return (uint32_t)slot * BS + wave_lds + __lane_id();
    }
    __device__ __forceinline__ uint32_t spill_off(int slot) const {
        return ((wave_gid + __lane_id()) * slots + (uint32_t)(slot - S)) * (uint32_t)sizeof(uint2);
    }
    __device__ __forceinline__ bool push(uint32_t ref, float key) {
        if ((uint32_t)sp >= cap) return false;
        if (sp < S) {
            lds_stack[lds_index(sp)] = make_uint2(ref, __float_as_uint(key));
        } else {
            const uint32_t off = spill_off(sp);
            spills++;
            __builtin_amdgcn_raw_buffer_store_b32(ref, spill, off, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(key), spill, off + 4, 0, 0);
        }
        sp++;
        return true;
    }
    __device__ __forceinline__ uint2 pop() {
        sp--;
        if (sp < S) return lds_stack[lds_index(sp)];
        const uint32_t off = spill_off(sp);
        return make_uint2(__builtin_amdgcn_raw_buffer_load_b32(spill, off, 0, 0),
                          __builtin_amdgcn_raw_buffer_load_b32(spill, off + 4, 0, 0));
    }
};

// Pushes the children of a node step that the lane visits later and returns the one it
// continues with (kEmpty when no child is hit).  hit[j] marks the children to visit.
// SORTED: the children are ordered near to far, the first hit one is next and the others
// are pushed far to near; otherwise the first hit child in slot order is next and the
// others are pushed in descending slot order.  When every push fits below the last LDS
// slot, each child is written to the next free slot unconditionally and the slot
// advances only past a pushed child, so a skipped child's entry is overwritten by the
// next write or lies above the new stack top: a select per child instead of nested
// branches per push (rays_scene closest hit 212-223 -> 198-205 us).
template <bool SORTED, int W, int S, int BS>
__device__ __forceinline__ uint32_t push_children(LaneStack<S, BS>& st, const uint32_t (&c)[W], const float (&key)[W],
                                                  const bool (&hit)[W], bool& ok) {
    uint32_t next = kEmpty;
    if constexpr (SORTED) {
        next = hit[0] ? c[0] : kEmpty;
    } else {
#pragma unroll
        for (int j = W - 1; j >= 0; j--)
            if (hit[j]) next = c[j];
    }
    constexpr int J0 = SORTED ? 1 : 0;  // slots that may be pushed: [J0, W)
    bool push[W];
    uint32_t n = 0;
#pragma unroll
    for (int j = J0; j < W; j++) {
        push[j] = hit[j] && (SORTED || c[j] != next);  // child words of hit children differ
        n += push[j] ? 1u : 0u;
    }
    if ((uint32_t)st.sp + n < (uint32_t)S) {
        // the writes reach slot sp + n at most, inside the lane's LDS column
        uint32_t slot = (uint32_t)st.sp;
#pragma unroll
        for (int j = W - 1; j >= J0; j--) {
            lds_stack[st.lds_index((int)slot)] = make_uint2(c[j], __float_as_uint(key[j]));
            slot += push[j] ? 1u : 0u;
        }
        st.sp = (int)slot;
    } else {
#pragma unroll
        for (int j = W - 1; j >= J0; j--)
            if (push[j] && !st.push(c[j], key[j])) ok = false;
    }
    return next;
}

template <int W>
__device__ __forceinline__ void sort_children(float (&d)[W], uint32_t (&c)[W]) {
    // keys by min / max, children by selects: no branch (a branch per comparator cost
    // the exchange's moves under divergent exec masks)
    auto cx = [&](int i, int j) {
        const bool s = d[j] < d[i];
        const float lo = fminf(d[i], d[j]), hi = fmaxf(d[i], d[j]);
        const uint32_t ci = s ? c[j] : c[i], cj = s ? c[i] : c[j];
        d[i] = lo;
        d[j] = hi;
        c[i] = ci;
        c[j] = cj;
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

// Child boxes and child words of one wide node, in registers.
template <int W>
struct ChildBoxes {
    float lox[W], loy[W], loz[W];
    float hix[W], hiy[W], hiz[W];
    uint32_t child[W];
};

// FP32 node: the box and child words (7 * W dwords; all 16 dwords of a 2-wide node) as
// unconditional dwordx4 loads issued together, before any use.
template <int W>
__device__ __forceinline__ void fetch_node(const WNode<W>* __restrict__ nodes, uint32_t i, ChildBoxes<W>& b) {
    constexpr int NV = W == 2 ? 4 : 7 * W / 4;  // float4 loads
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    float f[4 * NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const float4 v = p[k];
        f[4 * k] = v.x;
        f[4 * k + 1] = v.y;
        f[4 * k + 2] = v.z;
        f[4 * k + 3] = v.w;
    }
#pragma unroll
    for (int j = 0; j < W; j++) {
        b.lox[j] = f[j];
        b.loy[j] = f[W + j];
        b.loz[j] = f[2 * W + j];
        b.hix[j] = f[3 * W + j];
        b.hiy[j] = f[4 * W + j];
        b.hiz[j] = f[5 * W + j];
        b.child[j] = __float_as_uint(f[6 * W + j]);
    }
}

// Quantized node: sizeof(QNode<W>) / 16 dwordx4 loads, then every box plane decodes as
// fmaf(q, 2^(e-127), origin) -- the expression the builder checked it with.
// the per-axis plane scales of a quantized node (its words w)
template <int W>
__device__ __forceinline__ V3 qnode_scales(const uint32_t* w) {
    if constexpr (QNode<W>::kFloatScale)
        return v3(__uint_as_float(w[3]), __uint_as_float(w[4]), __uint_as_float(w[5]));
    else
        return v3(__uint_as_float((w[3] & 0xFFu) << 23), __uint_as_float(((w[3] >> 8) & 0xFFu) << 23),
                  __uint_as_float(((w[3] >> 16) & 0xFFu) << 23));
}

template <int W>
__device__ __forceinline__ void fetch_node(const QNode<W>* __restrict__ nodes, uint32_t i, ChildBoxes<W>& b) {
    using QN = QNode<W>;
    constexpr int NW = QN::kWords;
    constexpr int Q0 = QN::kQ0;  // first word of the 8-bit planes
    const uint4* p = reinterpret_cast<const uint4*>(nodes + i);
    uint32_t w[NW];
#pragma unroll
    for (int k = 0; k < NW / 4; k++) {
        const uint4 v = p[k];
        w[4 * k] = v.x;
        w[4 * k + 1] = v.y;
        w[4 * k + 2] = v.z;
        w[4 * k + 3] = v.w;
    }
    const float ox = __uint_as_float(w[0]), oy = __uint_as_float(w[1]), oz = __uint_as_float(w[2]);
    const V3 sc = qnode_scales<W>(w);
    const float sx = sc.x, sy = sc.y, sz = sc.z;
    auto qb = [&](int row, int j) -> float {
        const int byte = row * W + j;
        return (float)((w[Q0 + byte / 4] >> (8 * (byte % 4))) & 0xFFu);
    };
#pragma unroll
    for (int j = 0; j < W; j++) {
        b.child[j] = w[QN::kChild0 + j];
        b.lox[j] = fmaf(qb(0, j), sx, ox);
        b.loy[j] = fmaf(qb(1, j), sy, oy);
        b.loz[j] = fmaf(qb(2, j), sz, oz);
        b.hix[j] = fmaf(qb(3, j), sx, ox);
        b.hiy[j] = fmaf(qb(4, j), sy, oy);
        b.hiz[j] = fmaf(qb(5, j), sz, oz);
    }
}

template <int W, bool Q>
using NodeOf = typename std::conditional<Q, QNode<W>, WNode<W>>::type;

// Ray / child-box slab tests of one wide node: tn[j] = entry distance of child j and c[j]
// its child word, or INFINITY / kEmpty when the ray's [tmin, best] misses the box.  The
// plane distances are t = (plane - o) / d formed as one FMA with inv = 1/d and
// oi = -o * inv precomputed; `slack` is the relative slack on box exits.
template <int W>
__device__ __forceinline__ void ray_children(const WNode<W>* __restrict__ nodes, uint32_t i, V3 inv, float oix,
                                             float oiy, float oiz, float tmin, float best, float slack, float* tn,
                                             uint32_t* c) {
    ChildBoxes<W> nd;
    fetch_node(nodes, i, nd);
#pragma unroll
    for (int j = 0; j < W; j++) {
        const float t0x = fmaf(nd.lox[j], inv.x, oix), t1x = fmaf(nd.hix[j], inv.x, oix);
        const float t0y = fmaf(nd.loy[j], inv.y, oiy), t1y = fmaf(nd.hiy[j], inv.y, oiy);
        const float t0z = fmaf(nd.loz[j], inv.z, oiz), t1z = fmaf(nd.hiz[j], inv.z, oiz);
        const float lo = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
        const float hi = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), best));
        const bool hb = (nd.child[j] != kEmpty) & (lo <= fmaf(fabsf(hi), slack, hi));
        tn[j] = hb ? lo : INFINITY;
        c[j] = hb ? nd.child[j] : kEmpty;
    }
}

// 8-bit node: the plane distances come straight from the byte planes.  With
// a = 2^(e-127) * inv and b = origin * inv + oi per axis (once per node), a plane of
// byte q is at t = q * a + b: the distance of the decoded plane q * 2^(e-127) + origin,
// with two FMAs per child and axis fewer than decoding first.  Its rounding differs from
// the decode-then-slab form by a few FP32 ulps of the coordinates, far inside the
// builder's box inflation (2^-18 of the scene's extent), so the cull stays conservative.
template <int W>
__device__ __forceinline__ void ray_children(const QNode<W>* __restrict__ nodes, uint32_t i, V3 inv, float oix,
                                             float oiy, float oiz, float tmin, float best, float slack, float* tn,
                                             uint32_t* c) {
    using QN = QNode<W>;
    constexpr int NW = QN::kWords;
    constexpr int Q0 = QN::kQ0;  // first word of the 8-bit planes
    const uint4* p = reinterpret_cast<const uint4*>(nodes + i);
    uint32_t w[NW];
#pragma unroll
    for (int k = 0; k < NW / 4; k++) {
        const uint4 v = p[k];
        w[4 * k] = v.x;
        w[4 * k + 1] = v.y;
        w[4 * k + 2] = v.z;
        w[4 * k + 3] = v.w;
    }
    const V3 sc = qnode_scales<W>(w);
    const float ax = sc.x * inv.x;
    const float ay = sc.y * inv.y;
    const float az = sc.z * inv.z;
    const float bx = fmaf(__uint_as_float(w[0]), inv.x, oix);
    const float by = fmaf(__uint_as_float(w[1]), inv.y, oiy);
    const float bz = fmaf(__uint_as_float(w[2]), inv.z, oiz);
    auto qb = [&](int row, int j) -> float {
        const int byte = row * W + j;
        return (float)((w[Q0 + byte / 4] >> (8 * (byte % 4))) & 0xFFu);
    };
    if constexpr (W % 4 == 0) {
        // Near and far planes chosen once per node by the direction's signs.  The
        // builder gives every non-empty child lo byte <= hi byte, and q -> fmaf(q, a, b)
        // is non-decreasing for a >= 0 and non-increasing for a <= 0 (a has inv's sign),
        // so the row picked is exactly the min / max of the two plane distances.  With
        // W % 4 == 0 each row is whole words, so the choice is a word select per row word
        // instead of a min and a max per child and axis.  An empty slot holds the
        // inverted box lo = 255, hi = 0 on every axis: for either sign its near plane
        // lies 255 steps beyond its far one, so the test rejects it without a compare of
        // the child word; a degenerate ray that still passes it (a scale step below 2^-20
        // of the plane distances) gets its child word, kEmpty, and never visits it.
        constexpr int RW = W / 4;  // words per byte row
        const bool nx = inv.x < 0.0f, ny = inv.y < 0.0f, nz = inv.z < 0.0f;
        uint32_t nw[3][RW], fw[3][RW];
#pragma unroll
        for (int k = 0; k < RW; k++) {
            nw[0][k] = nx ? w[Q0 + 3 * RW + k] : w[Q0 + k];
            fw[0][k] = nx ? w[Q0 + k] : w[Q0 + 3 * RW + k];
            nw[1][k] = ny ? w[Q0 + 4 * RW + k] : w[Q0 + RW + k];
            fw[1][k] = ny ? w[Q0 + RW + k] : w[Q0 + 4 * RW + k];
            nw[2][k] = nz ? w[Q0 + 5 * RW + k] : w[Q0 + 2 * RW + k];
            fw[2][k] = nz ? w[Q0 + 2 * RW + k] : w[Q0 + 5 * RW + k];
        }
        auto bt = [](uint32_t v, int j) -> float { return (float)((v >> (8 * (j % 4))) & 0xFFu); };
#pragma unroll
        for (int j = 0; j < W; j++) {
            const uint32_t ch = w[QN::kChild0 + j];
            const float lo = fmaxf(fmaxf(fmaf(bt(nw[0][j / 4], j), ax, bx), fmaf(bt(nw[1][j / 4], j), ay, by)),
                                   fmaxf(fmaf(bt(nw[2][j / 4], j), az, bz), tmin));
            const float hi = fminf(fminf(fmaf(bt(fw[0][j / 4], j), ax, bx), fmaf(bt(fw[1][j / 4], j), ay, by)),
                                   fminf(fmaf(bt(fw[2][j / 4], j), az, bz), best));
            const bool hb = lo <= fmaf(fabsf(hi), slack, hi);
            tn[j] = hb ? lo : INFINITY;
            c[j] = hb ? ch : kEmpty;  // an empty slot stays kEmpty either way
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < W; j++) {
        const uint32_t ch = w[QN::kChild0 + j];
        const float t0x = fmaf(qb(0, j), ax, bx), t1x = fmaf(qb(3, j), ax, bx);
        const float t0y = fmaf(qb(1, j), ay, by), t1y = fmaf(qb(4, j), ay, by);
        const float t0z = fmaf(qb(2, j), az, bz), t1z = fmaf(qb(5, j), az, bz);
        const float lo = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
        const float hi = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), best));
        // (the per-axis min / max turns the inverted box of an empty slot into a proper
        // one, so this form tests the child word)
        const bool hb = (ch != kEmpty) & (lo <= fmaf(fabsf(hi), slack, hi));
        tn[j] = hb ? lo : INFINITY;
        c[j] = hb ? ch : kEmpty;
    }
}

// ------------------------------------------------------------------------------------
// closest point
// ------------------------------------------------------------------------------------

// Resident waves per SIMD the register allocator targets: 7 (<= 72 VGPRs) for 2- and
// 4-wide nodes.  At 8 (64 VGPRs) the 4-wide closest-hit kernel spilled 5 VGPRs to
// scratch and 20 SGPRs to VGPR lanes, reloading them with v_readlane in the loop; at 7
// it spills nothing (70 VGPRs), and every 4-wide kernel got faster with identical
// answers: rays_scene closest hit 219 -> 213 us, diff closest point 875 -> 861 us,
// cpq_large 13.7 -> 13.4 ms, batched replay 21.55 -> 21.33 ms.  Two such kernels at 16
// waves per CU each (the two-stream ray step) still run side by side as fast as before.
// 8-wide nodes need more live registers for the child boxes and sort keys, so they get
// what they need (5-7 waves).
template <int W, bool Q = false>
struct WavesHint {
    static constexpr int value = W == 8 ? 4 : 7;
};

constexpr float kCullSlack = 1.0f + 0x1p-16f;  // relative slack on squared distances
constexpr float kTSlack = 0x1p-20f;             // relative slack on ray box exits

// ------------------------------------------------------------------------------------
// Streaming traversal with postponed leaves
//
// A lane runs one step per loop iteration, and in each iteration a wave runs either node
// steps or leaf steps, never both (the two paths are long and would otherwise both be
// issued whenever lanes disagree).  A lane that reaches a leaf parks it -- one parked
// leaf per lane -- and keeps visiting nodes from its stack meanwhile; a lane that
// reaches a second leaf waits.  The wave runs leaf steps when a share (trig/64) of the
// lanes with a query have parked a leaf, or when no lane has a node to visit
// (postponed leaves with speculative traversal: Aila & Laine, "Understanding the
// Efficiency of Ray Traversal on GPUs", HPG 2009).  Extra visits cannot change the
// answer: culls only compare against the best value found so far.
//
// Queries come from a source (Src) that every lane of the wave calls together at the
// top of each iteration: a lane whose query is done writes its result and takes the
// next one in the same iteration, so lanes do not wait for the wave's slowest query.
// Every query ends (each node is visited at most once), and each iteration visits a
// node, tests a parked leaf, pops an entry or ends a query.
//
// A source provides
//   typename Handle;                         which query (index, or mesh and index)
//   bool refill(bool has, Handle& h);        called by all lanes together; gives lanes
//                                            with has == false their next query
//   bool exhausted();                        (wave-uniform) no further query for the wave
//   static Handle shfl(Handle, uint32_t l);  lane l's handle (a wave shuffle)
//   uint32_t root(const Handle&);            root node of the query's tree
//   ray(const Handle&, float4&, float4&)     (rays) origin, tmin / direction, tmax
//   V3 point(const Handle&)                  (closest point) the query point
// ------------------------------------------------------------------------------------

// Wave-uniform choice of the step kind: leaf steps when some lane has a parked leaf and
// either no lane has a node to visit or at least trig/64 of the lanes with a query have
// parked a leaf (trig = 1..64).
__device__ __forceinline__ bool leaf_step_now(bool parked, bool node_ready, uint32_t trig) {
    const uint64_t pk = __ballot(parked);
    if (pk == 0) return false;
    if (__ballot(node_ready) == 0) return true;
    return (uint32_t)__popcll(pk) * 64u >= (uint32_t)__popcll(__ballot(true)) * trig;
}

// The first query of global lane g (of total lanes, a multiple of 64): lane k of wave w
// takes query k * (total / 64) + w, a permutation of [0, total).  A query set smaller
// than the grid is thereby spread over every wave, a few queries per wave with the other
// lanes free to help through work donation, instead of filling the first waves; for a
// set that fills the grid it measured faster than lane order as well (1 M closest-hit
// rays on rays_scene: 4.20 vs 3.76 G rays/s on one box, any hit unchanged).
__device__ __forceinline__ uint32_t first_query(uint32_t g, uint32_t total) {
    return (g & 63u) * (total >> 6) + (g >> 6);
}

// Plain source with load balancing: lane g takes query first_query(g) first; the queries after the
// first `total` are cut into 64-query chunks dealt round-robin to nq = min(kQueues,
// blocks) queues (chunk j belongs to queue j % nq), each with its own device counter
// (zero at launch, 256 bytes apart).  Block b's waves take chunks of queue b % nq, so
// every queue is served by at least one block: a wave takes a
// chunk when lanes want work and its pool is empty, and gives the chunk's queries to
// those lanes in lane order.  Lanes that get cheap queries take more of them, so no
// lane idles while others still hold several; many counters keep the device-wide
// atomics from serializing on one address (one shared counter costs ~0.1 ms per
// million queries).  refill is called by all lanes together, so the pool stays
// wave-uniform.
constexpr uint32_t kQueues = 256;
constexpr uint32_t kQueueStride = 64;  // uint32 words between counters

struct QueueSource {
    using Handle = uint32_t;
    const float4* rays;
    const float* pts;
    uint32_t n, first, total;
    uint32_t* counters;
    bool started = false;
    uint32_t pnext = 0, pend = 0;  // wave's pool: queries [pnext, pend) not handed out
    bool more = true;
    __device__ __forceinline__ bool refill(bool has, Handle& h) {
        if (!started) {  // first call (all lanes together): the static first query only
            started = true;
            more = total < n;
            h = first;
            return first < n;
        }
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
                const uint64_t beg = (uint64_t)total + ((uint64_t)c * nq + q) * 64u;
                if (beg >= n) {
                    more = false;
                    break;
                }
                pnext = (uint32_t)beg;
                pend = beg + 64u < n ? (uint32_t)(beg + 64u) : n;
            }
            const uint32_t k = min(pend - pnext, (uint32_t)__popcll(want));
            const uint32_t rank = (uint32_t)__popcll(want & below);
            if (!has && !got && rank < k) {
                h = pnext + rank;
                got = true;
            }
            pnext += k;
            want = __ballot(!has && !got);
        }
        return got;
    }
    // wave-uniform: no lane of this wave will get another query
    __device__ __forceinline__ bool exhausted() const { return started && !more; }
    __device__ __forceinline__ static Handle shfl(Handle h, uint32_t from) { return __shfl(h, (int)from); }
    __device__ __forceinline__ uint32_t root(Handle) const { return 0u; }
    __device__ __forceinline__ void ray(Handle h, float4& r0, float4& r1) const {
        r0 = rays[2 * (size_t)h];
        r1 = rays[2 * (size_t)h + 1];
    }
    __device__ __forceinline__ V3 point(Handle h) const {
        return v3(pts[3 * (size_t)h], pts[3 * (size_t)h + 1], pts[3 * (size_t)h + 2]);
    }
};

// ------------------------------------------------------------------------------------
// Work donation at the end of a launch
//
// Once a wave's source has no more queries, its lanes without a query would idle while
// the last queries finish one node at a time (measured: a launch of 262,144 rays takes
// 82% as long as one of 1,048,576).  Instead an idle lane takes the top stack entry of
// a lane that still traverses and keeps other work: both then work on the same query,
// each on its own subtrees, the helper starting from the best value the donor had.
// The lanes working on one query form its group -- the query's own lane (the owner) and
// its helpers; helpers donate further in the same way.  A member merges each better
// result into the owner's 64-bit key in LDS with an atomic min and adopts the key when
// it is better than its own (so every member culls with the group's best).  The key orders
// (value, primitive id) lexicographically (result_key), so the merge keeps the
// lowest-id tie rule.  The owner's LDS count holds the members still traversing (one at
// the start, one more per helper, one less per member that finishes); when it reaches
// zero, the member that found the winning key writes the result (the owner when nothing
// was found) and all members go idle.  For closest points only the id leaves the
// traversal and it is the key's low word, so the owner writes it and a helper is free to
// take more work as soon as its own subtrees are done.  Every subtree stays with exactly one lane, and
// every lane culls only against a value that some primitive (or the query's own limit)
// reached, so the answer is the one a single lane finds.
// ------------------------------------------------------------------------------------

// position of the n-th (from 0) set bit of m; n < popcount(m)
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t n) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t w = 32; w >= 1; w >>= 1) {
        const uint32_t c = (uint32_t)__popcll((m >> pos) & ((1ull << w) - 1ull));
        if (n >= c) {
            n -= c;
            pos += w;
        }
    }
    return pos;
}

__device__ __forceinline__ uint64_t lanes_below() {
    const uint32_t lane = __lane_id();
    return lane ? (~0ull >> (64u - lane)) : 0ull;
}

// this wave's 64 group keys and member counts: after the block's S * BS stack slots,
// BS keys (8 bytes) then BS counts (4 bytes)
template <int S, int BS>
__device__ __forceinline__ unsigned long long* group_keys() {
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x) & ~63u;
    return reinterpret_cast<unsigned long long*>(lds_stack + S * BS) + w;
}
template <int S, int BS>
__device__ __forceinline__ uint32_t* group_counts() {
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x) & ~63u;
    return reinterpret_cast<uint32_t*>(reinterpret_cast<unsigned long long*>(lds_stack + S * BS) + BS) + w;
}

// Pairs idle lanes with donors (called by all lanes together): the r-th idle lane, in
// lane order, takes from the r-th donor.  Returns the donor lane for a recipient (take),
// else this lane; give marks a donor that has been paired.
__device__ __forceinline__ uint32_t pair_donors(bool idle, bool donor, bool& take, bool& give) {
    const uint64_t I = __ballot(idle), D = __ballot(donor);
    const uint32_t k = min((uint32_t)__popcll(I), (uint32_t)__popcll(D));
    const uint64_t below = lanes_below();
    const uint32_t ri = (uint32_t)__popcll(I & below), rd = (uint32_t)__popcll(D & below);
    give = donor && rd < k;
    take = idle && ri < k;
    return take ? nth_set_bit(D, ri) : __lane_id();
}

// Inputs that cannot be answered.  A ray with a non-finite origin or direction component,
// a zero direction or a NaN bound is a miss: its interval is emptied (tmin = +inf, tmax =
// -inf), so no box and no triangle passes the interval tests and the traversal ends
// after the root.  A query point with a non-finite coordinate finds no triangle (its
// cull bound is -inf), like a query on an empty mesh.  Without this, NaN entries would
// pass the fmaxf / fminf box tests (they drop NaN operands) and visit every node.
__device__ __forceinline__ void empty_if_invalid_ray(float4 r0, float4 r1, float& tmin, float& tmax) {
    const bool ok = isfinite(r0.x) && isfinite(r0.y) && isfinite(r0.z) && isfinite(r1.x) && isfinite(r1.y) &&
                    isfinite(r1.z) && (r1.x != 0.0f || r1.y != 0.0f || r1.z != 0.0f) && !isnan(r0.w) &&
                    !isnan(r1.w);
    if (!ok) {
        tmin = INFINITY;
        tmax = -INFINITY;
    }
}
__device__ __forceinline__ bool finite_point(V3 q) { return isfinite(q.x) && isfinite(q.y) && isfinite(q.z); }

// (value, id) as one integer that orders like the pair: the float's bits mapped to an
// order-preserving integer (negative values have their bits inverted, the rest get the
// sign bit set), with -0 first turned into +0 since the traversal treats them as equal
__device__ __forceinline__ unsigned long long result_key(float value, uint32_t id) {
    const uint32_t b = __float_as_uint(value + 0.0f);
    const uint32_t ord = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return (unsigned long long)ord << 32 | id;
}
__device__ __forceinline__ float key_value(unsigned long long key) {
    const uint32_t ord = (uint32_t)(key >> 32);
    return __uint_as_float((ord & 0x80000000u) ? (ord & 0x7FFFFFFFu) : ~ord);
}

// Ray traversal tuning, from sweeps on rays_scene (2.3 M triangles, 1 M + 1 M incoherent
// rays): two node visits per node step (closest hit 242 -> 230 us, any hit 168 -> 158 us,
// both passes on two streams 6.4 -> 6.7 G rays/s, all at 24 waves per CU), and donation
// only once 16 lanes of a wave are idle (a donation round costs the whole wave ~20
// shuffles: 235-236 us with a threshold of 12 or 20 lanes vs 242 us with 1; 32 lanes
// gives the gain back).  A third visit per step, or parking a leaf child inside the step
// to pop the next entry, measured no better.
constexpr int kRayNodeVisits = 2;
constexpr uint32_t kRayDonateIdle = 16;

template <typename NodeT, int S, int BS, bool ANY, bool STATS, typename Src, typename Out>
__device__ __forceinline__ bool ray_stream(const NodeT* __restrict__ nodes, const float4* __restrict__ tris,
                                           Src& src, uint32_t trig, LaneStack<S, BS>& st, Out out,
                                           unsigned long long& n_nodes, unsigned long long& n_tris) {
    constexpr int W = NodeT::kWidth;
    bool ok = true;
    typename Src::Handle h{};
    bool has = false;
    V3 o = v3(0.f, 0.f, 0.f), d = o, inv = o;
    float oix = 0.f, oiy = 0.f, oiz = 0.f;
    float tmin = 0.f, best = 0.f;
    uint32_t bid = 0xFFFFFFFFu, bpos = 0, cur = 0, pleaf = kEmpty;
    bool hit_any = false, have = false;
    // work donation: the lane whose ray this lane works on, whether the ray is shared,
    // whether this lane's part is finished and waits for the rest of the group, and
    // (wave-uniform) whether the source has run dry
    const uint32_t lane = __lane_id();
    uint32_t owner = lane;
    bool grouped = false, wait = false, drained = false;
    bool mine = true;  // (closest hit) bid was found by this lane, so bpos locates it
    unsigned long long* const gkey = group_keys<S, BS>();
    uint32_t* const gcnt = group_counts<S, BS>();
    // the trigger word: bits 0-7 the leaf trigger, bits 8-15 the refill trigger
    const uint32_t leaf_trig = trig & 0xFFu;
    const uint32_t refill_trig = max((trig >> 8) & 0xFFu, 1u);
    for (;;) {
        bool done;
        if (!drained) {
            done = has && !have && pleaf == kEmpty && (st.sp == 0 || (ANY && hit_any));
        } else {
            // groups exist only once the source has run dry (wave-uniform branch)
            if (grouped && has && !wait) {
                // the best hit of the whole group so far, every iteration: any hit ends the
                // search; a closer closest hit tightens this lane's culls
                const unsigned long long g = gkey[owner];
                if (ANY) {
                    if (g == 0ull) {
                        have = false;
                        pleaf = kEmpty;
                        st.sp = 0;
                    }
                } else if (g < result_key(best, bid)) {
                    best = key_value(g);
                    bid = (uint32_t)g;
                    mine = false;
                }
            }
            done = has && !wait && !have && pleaf == kEmpty && (st.sp == 0 || (ANY && hit_any));
            if (grouped && done) {
                const unsigned long long k = ANY ? (hit_any ? 0ull : ~0ull) : result_key(best, bid);
                atomicMin(gkey + owner, k);
                atomicSub(gcnt + owner, 1u);
                done = false;
                // A helper that cannot write the result leaves at once to take more work:
                // any-hit results are written by the owner, and a closest-hit helper is the
                // writer only if it found a hit whose key is still the group's (keys only
                // fall, and each primitive lies in one subtree).
                if (lane != owner && (ANY || !mine || bid == 0xFFFFFFFFu || gkey[owner] < k)) {
                    has = grouped = false;
                    owner = lane;
                } else {
                    wait = true;
                }
            }
        }
        if (wait && gcnt[owner] == 0u) {  // the last member has finished
            const unsigned long long fin = gkey[owner];
            const bool write = ANY || (uint32_t)fin == 0xFFFFFFFFu
                                   ? lane == owner
                                   : mine && bid != 0xFFFFFFFFu && result_key(best, bid) == fin;
            if (write) {
                float bu = 0.f, bv = 0.f;
                if (!ANY && bid != 0xFFFFFFFFu) {
                    const float4 A = tris[3 * (size_t)bpos];
                    const float4 B = tris[3 * (size_t)bpos + 1];
                    const float4 C = tris[3 * (size_t)bpos + 2];
                    float tt;
                    (void)ray_tri(o, d, v3(A.x, A.y, A.z), v3(B.x, B.y, B.z), v3(C.x, C.y, C.z), tt, bu, bv);
                }
                out(h, bid, best, bu, bv, ANY ? fin == 0ull : hit_any);
            }
            has = wait = grouped = false;
            owner = lane;
        }
        // A lane whose ray is finished keeps its slot until refill_trig lanes of the wave are
        // idle (or all are, or the source has run dry): the results are then written and
        // new rays fetched for all of them at once, so the wave pays for the refill code
        // once per batch rather than every time some lane finishes.
        const uint32_t n_idle = (uint32_t)__popcll(__ballot(!has || done));
        if (drained || n_idle >= refill_trig || n_idle == (uint32_t)__popcll(__ballot(true))) {
            if (done) {
                // barycentrics of the winner, recomputed once with the same arithmetic
                float bu = 0.f, bv = 0.f;
                if (!ANY && bid != 0xFFFFFFFFu) {
                    const float4 A = tris[3 * (size_t)bpos];
                    const float4 B = tris[3 * (size_t)bpos + 1];
                    const float4 C = tris[3 * (size_t)bpos + 2];
                    float tt;
                    (void)ray_tri(o, d, v3(A.x, A.y, A.z), v3(B.x, B.y, B.z), v3(C.x, C.y, C.z), tt, bu, bv);
                }
                out(h, bid, best, bu, bv, hit_any);
                has = false;
                done = false;
            }
            if (!drained) {
                if (src.refill(has, h)) {
                    float4 r0, r1;
                    src.ray(h, r0, r1);
                    o = v3(r0.x, r0.y, r0.z);
                    d = v3(r1.x, r1.y, r1.z);
                    tmin = r0.w;
                    best = r1.w;
                    empty_if_invalid_ray(r0, r1, tmin, best);
                    const float x = fabsf(d.x) < 1e-30f ? copysignf(1e-30f, d.x) : d.x;
                    const float y = fabsf(d.y) < 1e-30f ? copysignf(1e-30f, d.y) : d.y;
                    const float z = fabsf(d.z) < 1e-30f ? copysignf(1e-30f, d.z) : d.z;
                    inv = v3(1.0f / x, 1.0f / y, 1.0f / z);
                    oix = -o.x * inv.x;
                    oiy = -o.y * inv.y;
                    oiz = -o.z * inv.z;
                    bid = 0xFFFFFFFFu;
                    hit_any = false;
                    st.sp = 0;
                    cur = src.root(h);
                    pleaf = kEmpty;
                    have = true;
                    has = true;
                    mine = true;
                    gkey[lane] = ~0ull;
                    gcnt[lane] = 1u;
                }
                drained = !STATS && src.exhausted();  // counting kernels count one lane's visits
            }
            if (!__any(has)) break;
        }
        if (drained) {
            // donation: an idle lane takes the top stack entry of a lane with other work left
            const bool idle = !has;
            const bool donor = has && !wait && !done && st.sp >= ((have || pleaf != kEmpty) ? 1 : 2);
            if ((uint32_t)__popcll(__ballot(idle)) >= kRayDonateIdle && __ballot(donor) != 0) {
                bool take, give;
                const uint32_t from = pair_donors(idle, donor, take, give);
                uint2 e = make_uint2(0u, 0u);
                if (give) {
                    e = st.pop();
                    if (!ANY && !grouped && bid != 0xFFFFFFFFu) atomicMin(gkey + owner, result_key(best, bid));
                    grouped = true;
                }
                e.x = __shfl(e.x, (int)from);
                e.y = __shfl(e.y, (int)from);
                owner = __shfl(owner, (int)from);
                h = Src::shfl(h, from);
                o = v3(__shfl(o.x, (int)from), __shfl(o.y, (int)from), __shfl(o.z, (int)from));
                d = v3(__shfl(d.x, (int)from), __shfl(d.y, (int)from), __shfl(d.z, (int)from));
                inv = v3(__shfl(inv.x, (int)from), __shfl(inv.y, (int)from), __shfl(inv.z, (int)from));
                oix = __shfl(oix, (int)from);
                oiy = __shfl(oiy, (int)from);
                oiz = __shfl(oiz, (int)from);
                tmin = __shfl(tmin, (int)from);
                best = __shfl(best, (int)from);
                if (take) {
                    atomicAdd(gcnt + owner, 1u);
                    has = true;
                    grouped = true;
                    bid = 0xFFFFFFFFu;
                    mine = true;
                    hit_any = false;
                    st.sp = 0;
                    pleaf = kEmpty;
                    cur = e.x;
                    have = __uint_as_float(e.y) <= best;
                }
            }
        }
        if (!has || done || wait) continue;
        if (!have && st.sp > 0) {
            const uint2 e = st.pop();
            if (__uint_as_float(e.y) <= best) {
                cur = e.x;
                have = true;
            }
        }
        if (have && pleaf == kEmpty && child_is_leaf(cur)) {
            pleaf = cur;
            have = false;
        }
        const bool node_ready = have && !child_is_leaf(cur);
        if (leaf_step_now(pleaf != kEmpty, node_ready, leaf_trig)) {
            if (pleaf == kEmpty) continue;
            const uint32_t f = leaf_first(pleaf), cnt = leaf_count(pleaf);
            bool better = false;
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
                    bpos = f + k;
                    mine = true;
                    better = true;
                }
            }
            pleaf = kEmpty;
            // share the lane's best with the rest of its group
            if (grouped && (ANY ? hit_any : better)) atomicMin(gkey + owner, ANY ? 0ull : result_key(best, bid));
            if (ANY && hit_any) {
                have = false;  // done: the result is written in the next iteration
            } else if (have && child_is_leaf(cur)) {
                pleaf = cur;  // the leaf this lane was waiting with
                have = false;
            }
            continue;
        }
        if (!node_ready) continue;
        // Node step: up to kRayNodeVisits visits in a row, the next one when the lane's
        // nearest child (or the entry it pops when every child was culled) is another
        // node, so the loop's bookkeeping above is paid once per several visits.
#pragma unroll 1
        for (int visit = 0; visit < kRayNodeVisits; visit++) {
            if (STATS) n_nodes++;
            // t = (b - o) / d as one FMA with o/d precomputed; its rounding is far below
            // the build-time box inflation, so the test stays conservative
            float tn[W];
            uint32_t c[W];
            ray_children<W>(nodes, cur, inv, oix, oiy, oiz, tmin, best, kTSlack, tn, c);
            if (!ANY) sort_children<W>(tn, c);
            have = false;
            // closest hit: continue with the nearest hit child and push the others far
            // to near; any hit: continue with the first in slot order
            bool hit[W];
#pragma unroll
            for (int j = 0; j < W; j++) hit[j] = c[j] != kEmpty;
            const uint32_t pend = push_children<!ANY>(st, c, tn, hit, ok);
            if (pend != kEmpty) {
                cur = pend;
                have = true;
            } else if (st.sp > 0) {
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

template <typename NodeT, int S, int BS, bool STATS, typename Src, typename Out>
__device__ __forceinline__ bool cpq_stream(const NodeT* __restrict__ nodes, const float4* __restrict__ tris,
                                           Src& src, uint32_t trig, LaneStack<S, BS>& st, Out out,
                                           unsigned long long& n_nodes, unsigned long long& n_tris) {
    constexpr int W = NodeT::kWidth;
    bool ok = true;
    typename Src::Handle h{};
    bool has = false;
    V3 q = v3(0.f, 0.f, 0.f);
    float best = INFINITY, thr = INFINITY;
    uint32_t bid = 0xFFFFFFFFu, cur = 0, pleaf = kEmpty;
    bool have = false;
    // work donation state, as in ray_stream
    const uint32_t lane = __lane_id();
    uint32_t owner = lane;
    bool grouped = false, wait = false, drained = false;
    unsigned long long* const gkey = group_keys<S, BS>();
    uint32_t* const gcnt = group_counts<S, BS>();
    for (;;) {
        bool done;
        if (!drained) {
            done = has && !have && pleaf == kEmpty && st.sp == 0;
        } else {
            // groups exist only once the source has run dry (wave-uniform branch)
            done = has && !wait && !have && pleaf == kEmpty && st.sp == 0;
            if (grouped && done) {
                atomicMin(gkey + owner, result_key(best, bid));
                atomicSub(gcnt + owner, 1u);
                done = false;
                if (lane == owner) {
                    wait = true;  // the owner holds the query and writes the group's result
                } else {
                    // a helper's best is in the group key, so it is free at once to take
                    // more of any group's work (one query near the centre of a closed
                    // surface can otherwise keep its wave busy many times longer than
                    // every other query of the launch together)
                    has = grouped = false;
                    owner = lane;
                }
            }
        }
        if (wait && gcnt[owner] == 0u) {  // the last member has finished
            // the owner writes the group's winner (its id is the key's low word)
            if (lane == owner) out(h, (uint32_t)gkey[owner]);
            has = wait = grouped = false;
            owner = lane;
        }
        if (done) {
            // only the winner's id leaves the traversal: its point, distance and
            // barycentrics are computed by a separate FP64 pass (cpq_finish_one)
            out(h, bid);
            has = false;
        }
        if (!drained) {
            if (src.refill(has, h)) {
                q = src.point(h);
                best = thr = finite_point(q) ? INFINITY : -INFINITY;
                bid = 0xFFFFFFFFu;
                st.sp = 0;
                cur = src.root(h);
                pleaf = kEmpty;
                have = true;
                has = true;
                gkey[lane] = ~0ull;
                gcnt[lane] = 1u;
            }
            drained = !STATS && src.exhausted();  // counting kernels count one lane's visits
        }
        if (!__any(has)) break;
        if (drained) {
            // donation: an idle lane takes the top stack entry of a lane with other work left
            const bool idle = !has;
            const bool donor = has && !wait && st.sp >= ((have || pleaf != kEmpty) ? 1 : 2);
            if (__ballot(idle) != 0 && __ballot(donor) != 0) {
                bool take, give;
                const uint32_t from = pair_donors(idle, donor, take, give);
                uint2 e = make_uint2(0u, 0u);
                if (give) {
                    e = st.pop();
                    if (!grouped && bid != 0xFFFFFFFFu) atomicMin(gkey + owner, result_key(best, bid));
                    grouped = true;
                }
                e.x = __shfl(e.x, (int)from);
                e.y = __shfl(e.y, (int)from);
                owner = __shfl(owner, (int)from);
                h = Src::shfl(h, from);
                q = v3(__shfl(q.x, (int)from), __shfl(q.y, (int)from), __shfl(q.z, (int)from));
                best = __shfl(best, (int)from);
                thr = __shfl(thr, (int)from);
                if (take) {
                    atomicAdd(gcnt + owner, 1u);
                    has = true;
                    grouped = true;
                    bid = 0xFFFFFFFFu;
                    st.sp = 0;
                    pleaf = kEmpty;
                    cur = e.x;
                    have = __uint_as_float(e.y) <= thr;
                }
            }
        }
        if (!has || wait) continue;
        if (!have && st.sp > 0) {
            const uint2 e = st.pop();
            if (grouped) {
                // the group's best so far, read beside the stack entry (a read of its own
                // every iteration costs the long closest-point searches more than the
                // fresher bound saves) tightens the culls
                const unsigned long long g = gkey[owner];
                if (g < result_key(best, bid)) {
                    best = key_value(g);
                    bid = (uint32_t)g;
                    thr = best * kCullSlack;
                }
            }
            if (__uint_as_float(e.y) <= thr) {
                cur = e.x;
                have = true;
            }
        }
        if (have && pleaf == kEmpty && child_is_leaf(cur)) {
            pleaf = cur;
            have = false;
        }
        const bool node_ready = have && !child_is_leaf(cur);
        if (leaf_step_now(pleaf != kEmpty, node_ready, trig & 0xFFu)) {
            if (pleaf == kEmpty) continue;
            const uint32_t f = leaf_first(pleaf), cnt = leaf_count(pleaf);
            bool better = false;
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
                    thr = best * kCullSlack;
                    better = true;
                }
            }
            pleaf = kEmpty;
            // share the lane's best with the rest of its group
            if (grouped && better) atomicMin(gkey + owner, result_key(best, bid));
            if (have && child_is_leaf(cur)) {
                pleaf = cur;
                have = false;
            }
            continue;
        }
        if (!node_ready) continue;
        if (STATS) n_nodes++;
        ChildBoxes<W> nd;
        fetch_node(nodes, cur, nd);
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
        have = false;
        bool hit[W];
#pragma unroll
        for (int j = 0; j < W; j++) hit[j] = c[j] != kEmpty && dd[j] <= thr;
        const uint32_t next = push_children<true>(st, c, dd, hit, ok);
        if (next != kEmpty) {
            cur = next;
            have = true;
        }
    }
    return ok;
}

}  // namespace gpurt_impl
