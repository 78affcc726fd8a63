// build.hip -- GPU BVH builder for gfx950.
//
// Replaces the driver build behind VK::Accel::recreate(const Mesh&)
// (src/vk/vulkan.cpp:881-936 -> vkCmdBuildAccelerationStructuresKHR at :877).
//
// Pipeline (every stage is one launch over ALL meshes of the build; a plain build is a
// batch of one mesh):
//   k_prep         triangle AABBs, centroid bounds and coordinate scale per mesh
//                  (wave-reduced, ordered-uint atomics)
//   k_morton       key = mesh_id << 30 | 30-bit Morton code of the AABB centre over the
//                  mesh's centroid bounds; value = triangle index
//   radix sort     LSD, 8-bit digits, LDS histograms + LDS ranking by wave ballots
//                  (stable, so equal codes stay in triangle order)
//   k_pack         triangles and AABBs gathered into sorted (leaf) order
//   k_karras       Karras (HPG 2012) binary radix tree over the sorted keys; because the
//                  mesh id is in the high key bits every mesh is one subtree, whose root
//                  is recorded
//   k_bottom_up    node AABBs by atomic-counter bottom-up climb (second arrival merges)
//   k_treelet      optional SAH treelet restructuring (treelet.hip)
//   k_collapse     one launch per wide level: a node opens the child with the largest
//                  surface area until it has W children; subtrees of <= leaf_size
//                  triangles become leaves (contiguous ranges of packed triangles)
#include <algorithm>
#include <cstring>
#include <vector>

#include "gpurt_internal.h"

namespace gpurt_impl {

// ------------------------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------------------------

__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t u) {
    u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return __uint_as_float(u);
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__device__ __forceinline__ uint32_t quant10(float c, float lo, float inv) {
    float f = ((c - lo) * inv) * 1024.0f;
    int v = (int)floorf(f);
    if (!(f >= 0.0f)) v = 0;
    if (v > 1023) v = 1023;
    return (uint32_t)v;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}

__device__ __forceinline__ int find_seg(const SegDesc* segs, int n_seg, uint32_t g) {
    int lo = 0, hi = n_seg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[mid].tri_offset <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// per-mesh statistics: [cmin xyz, cmax xyz, max|coord|, unused] as ordered uints
constexpr int kStat = 8;

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_init_build(const SegDesc* segs, int n_seg, uint32_t* segstat, uint32_t* seg_root,
                             uint32_t* lvl_cnt, uint32_t* counters) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_seg) {
        for (int k = 0; k < 3; k++) {
            segstat[i * kStat + k] = 0xFFFFFFFFu;
            segstat[i * kStat + 3 + k] = 0u;
        }
        segstat[i * kStat + 6] = 0u;
        segstat[i * kStat + 7] = 0u;
        seg_root[i] = kLeafBit | segs[i].tri_offset;  // single-triangle mesh: leaf root
    }
    if (i <= kMaxLevels) lvl_cnt[i] = (i == 0) ? (uint32_t)n_seg : 0u;
    if (i < 4) counters[i] = (i == 0) ? (uint32_t)n_seg : 0u;  // [0] wide nodes, [1] error
}

// ------------------------------------------------------------------------------------
// k_prep: AABBs + per-mesh centroid bounds
// ------------------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_prep(const SegDesc* __restrict__ segs, int n_seg,
                                                 uint32_t N, float4* __restrict__ aabb_lo,
                                                 float4* __restrict__ aabb_hi,
                                                 float4* __restrict__ tri_v,
                                                 uint32_t* __restrict__ segstat,
                                                 uint32_t* __restrict__ counters) {
    uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    bool valid = g < N;
    int s = valid ? find_seg(segs, n_seg, g) : 0;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float cen[3] = {INFINITY, INFINITY, INFINITY}, cenx[3] = {-INFINITY, -INFINITY, -INFINITY};
    float amax = 0.0f;
    if (valid) {
        const SegDesc sd = segs[s];
        uint32_t t = g - sd.tri_offset;
        uint32_t vi[3];
        bool bad = false;
        for (int k = 0; k < 3; k++) {
            vi[k] = sd.idx[3 * (size_t)t + k];
            if (vi[k] >= sd.n_verts) { bad = true; vi[k] = 0; }
        }
        if (bad) atomicOr(&counters[1], 1u);
        float v[3][3];
        for (int k = 0; k < 3; k++) {
            const float* p = sd.verts + (size_t)vi[k] * sd.vstride;
            v[k][0] = p[0]; v[k][1] = p[1]; v[k][2] = p[2];
        }
        for (int a = 0; a < 3; a++) {
            for (int k = 0; k < 3; k++) {
                lo[a] = fminf(lo[a], v[k][a]);
                hi[a] = fmaxf(hi[a], v[k][a]);
                amax = fmaxf(amax, fabsf(v[k][a]));
            }
            cen[a] = (lo[a] + hi[a]) * 0.5f;
            cenx[a] = cen[a];
        }
        aabb_lo[g] = make_float4(lo[0], lo[1], lo[2], __uint_as_float((uint32_t)s));
        aabb_hi[g] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        tri_v[3 * (size_t)g + 0] = make_float4(v[0][0], v[0][1], v[0][2], __uint_as_float(t));
        tri_v[3 * (size_t)g + 1] = make_float4(v[1][0], v[1][1], v[1][2], 0.0f);
        tri_v[3 * (size_t)g + 2] = make_float4(v[2][0], v[2][1], v[2][2], 0.0f);
    }
    // reduce over the wave when it covers one mesh, else per lane
    int s0 = __shfl(s, 0);
    bool uniform = __all(s == s0 || !valid);
    uint32_t o[7];
    for (int a = 0; a < 3; a++) {
        o[a] = valid ? f2o(cen[a]) : 0xFFFFFFFFu;
        o[3 + a] = valid ? f2o(cenx[a]) : 0u;
    }
    o[6] = valid ? f2o(amax) : 0u;
    if (uniform) {
        for (int a = 0; a < 3; a++) {
            o[a] = wave_min_u32(o[a]);
            o[3 + a] = wave_max_u32(o[3 + a]);
        }
        o[6] = wave_max_u32(o[6]);
        if ((threadIdx.x & 63) == 0 && __any(valid)) {
            for (int a = 0; a < 3; a++) {
                atomicMin(&segstat[s0 * kStat + a], o[a]);
                atomicMax(&segstat[s0 * kStat + 3 + a], o[3 + a]);
            }
            atomicMax(&segstat[s0 * kStat + 6], o[6]);
        }
    } else if (valid) {
        for (int a = 0; a < 3; a++) {
            atomicMin(&segstat[s * kStat + a], o[a]);
            atomicMax(&segstat[s * kStat + 3 + a], o[3 + a]);
        }
        atomicMax(&segstat[s * kStat + 6], o[6]);
    }
}

__global__ __launch_bounds__(kBlock) void k_morton(uint32_t N, const float4* __restrict__ aabb_lo,
                                                   const float4* __restrict__ aabb_hi,
                                                   const uint32_t* __restrict__ segstat,
                                                   uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
    uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= N) return;
    float4 lo = aabb_lo[g], hi = aabb_hi[g];
    uint32_t s = __float_as_uint(lo.w);
    const uint32_t* st = segstat + s * kStat;
    float c[3] = {(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f};
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        float l = o2f(st[a]), h = o2f(st[3 + a]);
        float inv = h > l ? 1.0f / (h - l) : 0.0f;
        q[a] = quant10(c[a], l, inv);
    }
    uint32_t code = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
    keys[g] = ((uint64_t)s << 30) | code;
    vals[g] = g;
}

// ------------------------------------------------------------------------------------
// LSD radix sort of (uint64 key, uint32 value), 8-bit digits, stable
// ------------------------------------------------------------------------------------

constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;

__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint64_t* __restrict__ keys, uint32_t n,
                                                       int shift, uint32_t* __restrict__ hist,
                                                       uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kSortTile;
#pragma unroll 4
    for (int r = 0; r < kSortItems; r++) {
        uint32_t i = base + r * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of `total` counts in one block of 1024 threads
__global__ __launch_bounds__(1024) void k_scan_excl(uint32_t* __restrict__ a, uint32_t total) {
    __shared__ uint32_t part[1024];
    uint32_t chunk = (total + 1023) / 1024;
    uint32_t beg = threadIdx.x * chunk;
    uint32_t end = min(beg + chunk, total);
    uint32_t s = 0;
    for (uint32_t i = beg; i < end; i++) s += a[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (uint32_t i = beg; i < end; i++) {
        uint32_t x = a[i];
        a[i] = run;
        run += x;
    }
}

__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint64_t* __restrict__ kin,
                                                          const uint32_t* __restrict__ vin,
                                                          uint64_t* __restrict__ kout,
                                                          uint32_t* __restrict__ vout, uint32_t n,
                                                          int shift,
                                                          const uint32_t* __restrict__ offs,
                                                          uint32_t nblocks) {
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    base[tid] = offs[(size_t)tid * nblocks + blockIdx.x];
    for (int w = 0; w < 4; w++) wcnt[w][tid] = 0;
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    uint32_t tile = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; r++) {
        uint32_t i = tile + r * kBlock + tid;
        bool valid = i < n;
        uint64_t key = valid ? kin[i] : 0ull;
        uint32_t val = valid ? vin[i] : 0u;
        uint32_t d = (uint32_t)(key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        uint32_t rank = __popcll(peers & lt_mask);
        bool leader = valid && rank == 0;
        if (leader) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = base[d] + rank;
            for (int w = 0; w < wave; w++) off += wcnt[w][d];
            kout[off] = key;
            vout[off] = val;
        }
        __syncthreads();
        uint32_t tot = wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        base[tid] += tot;
        for (int w = 0; w < 4; w++) wcnt[w][tid] = 0;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------
// k_pack: gather triangles and AABBs into sorted order
// ------------------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_pack(uint32_t N, const uint32_t* __restrict__ vals,
                                                 const float4* __restrict__ aabb_lo,
                                                 const float4* __restrict__ aabb_hi,
                                                 const float4* __restrict__ tri_v,
                                                 float4* __restrict__ saabb,
                                                 float4* __restrict__ tris,
                                                 uint32_t* __restrict__ leaf_order) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= N) return;
    uint32_t g = vals[p];
    saabb[2 * (size_t)p] = aabb_lo[g];
    saabb[2 * (size_t)p + 1] = aabb_hi[g];
    tris[3 * (size_t)p + 0] = tri_v[3 * (size_t)g + 0];
    tris[3 * (size_t)p + 1] = tri_v[3 * (size_t)g + 1];
    tris[3 * (size_t)p + 2] = tri_v[3 * (size_t)g + 2];
    if (leaf_order) leaf_order[p] = g;
}

// ------------------------------------------------------------------------------------
// k_karras: binary radix tree (Karras 2012, Algorithm 2) over 64-bit keys
// ------------------------------------------------------------------------------------

__device__ __forceinline__ int kdelta(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

// binary references: internal node i -> i; leaf (sorted primitive p) -> kLeafBit | p
__global__ __launch_bounds__(kBlock) void k_karras(int N, const uint64_t* __restrict__ keys,
                                                   const SegDesc* __restrict__ segs,
                                                   uint32_t* __restrict__ left,
                                                   uint32_t* __restrict__ right,
                                                   uint2* __restrict__ range,
                                                   uint32_t* __restrict__ node_parent,
                                                   uint32_t* __restrict__ leaf_parent,
                                                   uint32_t* __restrict__ seg_root,
                                                   uint8_t* __restrict__ is_root) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N - 1) return;
    int d = (kdelta(keys, N, i, i + 1) - kdelta(keys, N, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = kdelta(keys, N, i, i - d);
    int lmax = 2;
    while (kdelta(keys, N, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (kdelta(keys, N, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = kdelta(keys, N, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) / 2;
        if (kdelta(keys, N, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int split = i + s * d + min(d, 0);
    int first = min(i, j), last = max(i, j);
    uint32_t L = (first == split) ? (kLeafBit | (uint32_t)split) : (uint32_t)split;
    uint32_t R = (last == split + 1) ? (kLeafBit | (uint32_t)(split + 1)) : (uint32_t)(split + 1);
    left[i] = L;
    right[i] = R;
    range[i] = make_uint2((uint32_t)first, (uint32_t)last);
    if (L & kLeafBit) leaf_parent[split] = (uint32_t)i;
    else node_parent[split] = (uint32_t)i;
    if (R & kLeafBit) leaf_parent[split + 1] = (uint32_t)i;
    else node_parent[split + 1] = (uint32_t)i;
    uint32_t sf = (uint32_t)(keys[first] >> 30), sl = (uint32_t)(keys[last] >> 30);
    bool root = false;
    if (sf == sl) {
        const SegDesc sd = segs[sf];
        root = (uint32_t)first == sd.tri_offset && (uint32_t)last == sd.tri_offset + sd.n_tris - 1u;
    }
    is_root[i] = root ? 1 : 0;
    if (root) seg_root[sf] = (uint32_t)i;
}

// ------------------------------------------------------------------------------------
// k_bottom_up: internal node boxes
// ------------------------------------------------------------------------------------

__device__ __forceinline__ void ref_box(uint32_t r, const float4* __restrict__ saabb,
                                        const float4* __restrict__ nbox, float4& lo, float4& hi) {
    if (r & kLeafBit) {
        size_t p = r & ~kLeafBit;
        lo = saabb[2 * p];
        hi = saabb[2 * p + 1];
    } else {
        lo = nbox[2 * (size_t)r];
        hi = nbox[2 * (size_t)r + 1];
    }
}

__device__ __forceinline__ void ref_box_coherent(uint32_t r, const float4* saabb, const float4* nbox,
                                                 float4& lo, float4& hi) {
    if (r & kLeafBit) {
        size_t p = r & ~kLeafBit;
        lo = saabb[2 * p];
        hi = saabb[2 * p + 1];
    } else {
        // written by another thread of this launch: bypass non-coherent caches
        const float* b = reinterpret_cast<const float*>(nbox + 2 * (size_t)r);
        lo.x = __hip_atomic_load(b + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        lo.y = __hip_atomic_load(b + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        lo.z = __hip_atomic_load(b + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        lo.w = 0.0f;
        hi.x = __hip_atomic_load(b + 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        hi.y = __hip_atomic_load(b + 5, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        hi.z = __hip_atomic_load(b + 6, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        hi.w = 0.0f;
    }
}

__global__ __launch_bounds__(kBlock) void k_bottom_up(int N, const uint32_t* __restrict__ leaf_parent,
                                                      const uint32_t* __restrict__ node_parent,
                                                      const uint32_t* __restrict__ left,
                                                      const uint32_t* __restrict__ right,
                                                      const uint8_t* __restrict__ is_root,
                                                      const float4* __restrict__ saabb,
                                                      float4* nbox, uint32_t* flags) {
    int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= N) return;
    uint32_t cur = leaf_parent[p];
    if (cur == kEmpty) return;  // single-triangle mesh
    // each internal node is merged by the second of its two children to arrive
    for (int guard = 0; guard < 256; guard++) {
        __threadfence();
        uint32_t old = atomicAdd(&flags[cur], 1u);
        if (old == 0u) return;
        __threadfence();
        float4 al, ah, bl, bh;
        ref_box_coherent(left[cur], saabb, nbox, al, ah);
        ref_box_coherent(right[cur], saabb, nbox, bl, bh);
        float* o = reinterpret_cast<float*>(nbox + 2 * (size_t)cur);
        __hip_atomic_store(o + 0, fminf(al.x, bl.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(o + 1, fminf(al.y, bl.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(o + 2, fminf(al.z, bl.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(o + 3, 0.0f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(o + 4, fmaxf(ah.x, bh.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(o + 5, fmaxf(ah.y, bh.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(o + 6, fmaxf(ah.z, bh.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(o + 7, 0.0f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (is_root[cur]) return;
        cur = node_parent[cur];
        if (cur == kEmpty) return;
    }
}

// ------------------------------------------------------------------------------------
// k_collapse: binary tree -> W-wide nodes, one launch per wide level
// ------------------------------------------------------------------------------------

struct CollapseArgs {
    const uint32_t* left;
    const uint32_t* right;
    const uint2* range;
    const float4* saabb;
    const float4* nbox;
    const uint32_t* segstat;
    uint32_t* wsrc;  // wide node -> binary ref
    uint32_t* wseg;  // wide node -> mesh
    uint32_t* lvl_cnt;
    uint32_t* counters;  // [0] wide nodes allocated, [1] error bits
    void* nodes;
    uint32_t capacity;
    int leaf_size;
    int level;
};

__device__ __forceinline__ uint32_t ref_count(uint32_t r, const uint2* range) {
    if (r & kLeafBit) return 1u;
    uint2 fl = range[r];
    return fl.y - fl.x + 1u;
}
__device__ __forceinline__ uint32_t ref_first(uint32_t r, const uint2* range) {
    if (r & kLeafBit) return r & ~kLeafBit;
    return range[r].x;
}
__device__ __forceinline__ float box_area(float4 lo, float4 hi) {
    float ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
    return 2.0f * (ex * ez + ex * ey + ey * ez);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_collapse(CollapseArgs a) {
    uint32_t begin = 0;
    for (int l = 0; l < a.level; l++) begin += a.lvl_cnt[l];
    uint32_t cnt = a.lvl_cnt[a.level];
    WNode<W>* nodes = reinterpret_cast<WNode<W>*>(a.nodes);
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < cnt; k += gridDim.x * kBlock) {
        uint32_t wi = begin + k;
        uint32_t bref = a.wsrc[wi];
        uint32_t seg = a.wseg[wi];
        float pad = o2f(a.segstat[seg * kStat + 6]) * 0x1p-18f;
        uint32_t refs[W];
        int n = 0;
        if (ref_count(bref, a.range) <= (uint32_t)a.leaf_size) {
            refs[n++] = bref;
        } else {
            refs[n++] = a.left[bref];
            refs[n++] = a.right[bref];
            while (n < W) {
                int best = -1;
                float best_area = -1.0f;
#pragma unroll
                for (int j = 0; j < W; j++) {
                    if (j >= n) break;
                    uint32_t r = refs[j];
                    if (ref_count(r, a.range) <= (uint32_t)a.leaf_size) continue;
                    float4 lo, hi;
                    ref_box(r, a.saabb, a.nbox, lo, hi);
                    float ar = box_area(lo, hi);
                    if (ar > best_area) { best_area = ar; best = j; }
                }
                if (best < 0) break;
                uint32_t r = refs[best];
                refs[best] = a.left[r];
                refs[n++] = a.right[r];
            }
        }
        WNode<W> out;
        uint32_t n_internal = 0;
#pragma unroll
        for (int j = 0; j < W; j++) {
            if (j < n) {
                uint32_t r = refs[j];
                float4 lo, hi;
                ref_box(r, a.saabb, a.nbox, lo, hi);
                out.lox[j] = lo.x - pad; out.loy[j] = lo.y - pad; out.loz[j] = lo.z - pad;
                out.hix[j] = hi.x + pad; out.hiy[j] = hi.y + pad; out.hiz[j] = hi.z + pad;
                uint32_t c = ref_count(r, a.range);
                out.aux[j] = c;
                if (c <= (uint32_t)a.leaf_size) out.child[j] = make_leaf(ref_first(r, a.range), c);
                else n_internal++;
            } else {
                out.lox[j] = out.loy[j] = out.loz[j] = INFINITY;
                out.hix[j] = out.hiy[j] = out.hiz[j] = -INFINITY;
                out.child[j] = kEmpty;
                out.aux[j] = 0;
            }
        }
        if (n_internal) {
            uint32_t base = atomicAdd(&a.counters[0], n_internal);
            atomicAdd(&a.lvl_cnt[a.level + 1], n_internal);
            uint32_t m = 0;
#pragma unroll
            for (int j = 0; j < W; j++) {
                if (j < n && out.aux[j] > (uint32_t)a.leaf_size) {
                    uint32_t idx = base + m++;
                    if (idx < a.capacity) {
                        a.wsrc[idx] = refs[j];
                        a.wseg[idx] = seg;
                        out.child[j] = idx;
                    } else {
                        atomicOr(&a.counters[1], 2u);
                        out.child[j] = kEmpty;
                    }
                }
            }
        }
        nodes[wi] = out;
    }
}

// level 0 of the wide tree: one node per mesh root
__global__ void k_init_levels(int n_seg, const uint32_t* __restrict__ seg_root, uint32_t* __restrict__ wsrc,
                              uint32_t* __restrict__ wseg) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_seg) {
        wsrc[i] = seg_root[i];
        wseg[i] = (uint32_t)i;
    }
}

// ------------------------------------------------------------------------------------
// host orchestration
// ------------------------------------------------------------------------------------

BuildBufs::~BuildBufs() {
    for (void* p : allocs)
        if (p) (void)hipFree(p);
}

void* BuildBufs::release(void* p) {
    for (auto& a : allocs)
        if (a == p) a = nullptr;
    return p;
}

template <typename T>
static int balloc(BuildBufs& b, T*& p, size_t count) {
    size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    void* q = nullptr;
    if (hipMalloc(&q, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(GPURT_E_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    b.allocs.push_back(q);
    p = static_cast<T*>(q);
    return GPURT_OK;
}

int build_alloc(const SegDesc* host_segs, int n_seg, int W, BuildBufs& b) {
    uint64_t N64 = 0;
    for (int i = 0; i < n_seg; i++) {
        if (host_segs[i].n_tris == 0) return set_error(GPURT_E_INVALID, "a mesh has no triangles");
        N64 += host_segs[i].n_tris;
    }
    if (n_seg <= 0 || N64 == 0 || N64 > kMaxLeafFirst - 8)
        return set_error(GPURT_E_INVALID, "triangle count out of range (1 .. 2^28-9)");
    const uint32_t N = (uint32_t)N64;
    b.N = N;
    b.n_seg = n_seg;
    b.width = W;
    b.nblocks_sort = cdiv(N, kSortTile);
    const size_t NI = N > 1 ? N - 1 : 1;
    b.capacity = (uint32_t)(NI + n_seg);
    int rc;
    if ((rc = balloc(b, b.segs, n_seg)) || (rc = balloc(b, b.aabb_lo, N)) ||
        (rc = balloc(b, b.aabb_hi, N)) || (rc = balloc(b, b.tri_v, 3 * (size_t)N)) ||
        (rc = balloc(b, b.segstat, (size_t)kStat * n_seg)) || (rc = balloc(b, b.keys0, N)) ||
        (rc = balloc(b, b.keys1, N)) || (rc = balloc(b, b.vals0, N)) || (rc = balloc(b, b.vals1, N)) ||
        (rc = balloc(b, b.hist, 256 * (size_t)b.nblocks_sort)) || (rc = balloc(b, b.saabb, 2 * (size_t)N)) ||
        (rc = balloc(b, b.left, NI)) || (rc = balloc(b, b.right, NI)) || (rc = balloc(b, b.range, NI)) ||
        (rc = balloc(b, b.nparent, NI)) || (rc = balloc(b, b.lparent, N)) ||
        (rc = balloc(b, b.seg_root, n_seg)) || (rc = balloc(b, b.is_root, NI)) ||
        (rc = balloc(b, b.nbox, 2 * NI)) || (rc = balloc(b, b.flags, NI)) ||
        (rc = balloc(b, b.wsrc, b.capacity)) || (rc = balloc(b, b.wseg, b.capacity)) ||
        (rc = balloc(b, b.lvl, kMaxLevels + 1)) || (rc = balloc(b, b.counters, 4)) ||
        (rc = balloc(b, b.tris, 3 * (size_t)N)) || (rc = balloc(b, b.leaf_order, N)))
        return rc;
    void* nodes = nullptr;
    if (hipMalloc(&nodes, node_bytes(W) * b.capacity) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(GPURT_E_OOM, "hipMalloc of BVH nodes failed");
    }
    b.allocs.push_back(nodes);
    b.nodes = nodes;
    GPURT_HIP(hipMemcpy(b.segs, host_segs, sizeof(SegDesc) * n_seg, hipMemcpyHostToDevice));
    return GPURT_OK;
}

// Enqueues the build up to (not including) the wide collapse.  No host synchronisation
// and no memset nodes, so the sequence can be captured into a HIP graph.
int build_enqueue_binary(BuildBufs& b, int treelet_size, int treelet_rounds, hipStream_t s) {
    const uint32_t N = b.N;
    const int n_seg = b.n_seg;
    const size_t NI = N > 1 ? N - 1 : 1;
    const unsigned gN = cdiv(N, kBlock);
    hipLaunchKernelGGL(k_init_build, dim3(cdiv(std::max(n_seg, kMaxLevels + 1), kBlock)), dim3(kBlock),
                       0, s, b.segs, n_seg, b.segstat, b.seg_root, b.lvl, b.counters);
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(gN, 4096u)), dim3(kBlock), 0, s, b.lparent, (size_t)N,
                       kEmpty);
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(cdiv(NI, kBlock), 4096u)), dim3(kBlock), 0, s,
                       b.nparent, NI, kEmpty);
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(cdiv(NI, kBlock), 4096u)), dim3(kBlock), 0, s, b.flags,
                       NI, 0u);
    hipLaunchKernelGGL(k_prep, dim3(gN), dim3(kBlock), 0, s, b.segs, n_seg, N, b.aabb_lo, b.aabb_hi,
                       b.tri_v, b.segstat, b.counters);
    hipLaunchKernelGGL(k_morton, dim3(gN), dim3(kBlock), 0, s, N, b.aabb_lo, b.aabb_hi, b.segstat,
                       b.keys0, b.vals0);
    // radix sort over the bits in use: 30 Morton bits + mesh-id bits
    int seg_bits = 0;
    while ((1u << seg_bits) < (uint32_t)n_seg) seg_bits++;
    const int key_bits = 30 + seg_bits;
    uint64_t* kin = b.keys0;
    uint64_t* kout = b.keys1;
    uint32_t* vin = b.vals0;
    uint32_t* vout = b.vals1;
    for (int shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(b.nblocks_sort), dim3(kBlock), 0, s, kin, N, shift, b.hist,
                           b.nblocks_sort);
        hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, s, b.hist, 256u * b.nblocks_sort);
        hipLaunchKernelGGL(k_radix_scatter, dim3(b.nblocks_sort), dim3(kBlock), 0, s, kin, vin, kout,
                           vout, N, shift, b.hist, b.nblocks_sort);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    hipLaunchKernelGGL(k_pack, dim3(gN), dim3(kBlock), 0, s, N, vin, b.aabb_lo, b.aabb_hi, b.tri_v,
                       b.saabb, b.tris, b.leaf_order);
    if (N > 1) {
        hipLaunchKernelGGL(k_karras, dim3(cdiv(N - 1, kBlock)), dim3(kBlock), 0, s, (int)N, kin, b.segs,
                           b.left, b.right, b.range, b.nparent, b.lparent, b.seg_root, b.is_root);
        hipLaunchKernelGGL(k_bottom_up, dim3(gN), dim3(kBlock), 0, s, (int)N, b.lparent, b.nparent,
                           b.left, b.right, b.is_root, b.saabb, b.nbox, b.flags);
        if (treelet_size > 0) {
            int trc = run_treelets(b, treelet_size, treelet_rounds, s);
            if (trc != GPURT_OK) return trc;
        }
    }
    hipLaunchKernelGGL(k_init_levels, dim3(cdiv(n_seg, kBlock)), dim3(kBlock), 0, s, n_seg, b.seg_root,
                       b.wsrc, b.wseg);
    GPURT_HIP(hipGetLastError());
    return GPURT_OK;
}

template <int W>
static void launch_collapse_level(const CollapseArgs& a, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL(k_collapse<W>, dim3(grid), dim3(kBlock), 0, s, a);
}

static void collapse_level(int W, const CollapseArgs& a, unsigned grid, hipStream_t s) {
    if (W == 2) launch_collapse_level<2>(a, grid, s);
    else if (W == 4) launch_collapse_level<4>(a, grid, s);
    else launch_collapse_level<8>(a, grid, s);
}

int build_enqueue_levels(BuildBufs& b, int leaf_size, int level_begin, int level_end, hipStream_t s) {
    CollapseArgs ca;
    ca.left = b.left;
    ca.right = b.right;
    ca.range = b.range;
    ca.saabb = b.saabb;
    ca.nbox = b.nbox;
    ca.segstat = b.segstat;
    ca.wsrc = b.wsrc;
    ca.wseg = b.wseg;
    ca.lvl_cnt = b.lvl;
    ca.counters = b.counters;
    ca.nodes = b.nodes;
    ca.capacity = b.capacity;
    ca.leaf_size = leaf_size;
    const unsigned cgrid = std::min(cdiv(b.capacity, kBlock), 2048u);
    for (int level = level_begin; level < level_end; level++) {
        ca.level = level;
        collapse_level(b.width, ca, cgrid, s);
    }
    GPURT_HIP(hipGetLastError());
    return GPURT_OK;
}

// Plain (single-mesh) build: enqueue, run the collapse in groups of levels until the
// tree ends (host checks between groups), then hand the products to *out.
int build_segments(gpurt_ctx ctx, const SegDesc* host_segs, int n_seg, int W, int leaf_size,
                   int treelet_size, int treelet_rounds, hipStream_t s, gpurt_bvh_s* out) {
    (void)ctx;
    BuildBufs b;
    int rc;
    if ((rc = build_alloc(host_segs, n_seg, W, b))) return rc;
    if ((rc = build_enqueue_binary(b, treelet_size, treelet_rounds, s))) return rc;
    std::vector<uint32_t> h_lvl(kMaxLevels + 1, 0);
    int level = 0;
    bool done = false;
    while (!done && level < kMaxLevels) {
        int group_end = std::min(level + 8, kMaxLevels);
        if ((rc = build_enqueue_levels(b, leaf_size, level, group_end, s))) return rc;
        level = group_end;
        GPURT_HIP(hipMemcpyAsync(h_lvl.data(), b.lvl, sizeof(uint32_t) * (kMaxLevels + 1),
                                 hipMemcpyDeviceToHost, s));
        GPURT_HIP(hipStreamSynchronize(s));
        done = h_lvl[level] == 0;
    }
    uint32_t h_counters[4] = {0, 0, 0, 0};
    GPURT_HIP(hipMemcpyAsync(h_counters, b.counters, sizeof(h_counters), hipMemcpyDeviceToHost, s));
    uint32_t st[kStat];
    GPURT_HIP(hipMemcpyAsync(st, b.segstat, sizeof(st), hipMemcpyDeviceToHost, s));
    GPURT_HIP(hipStreamSynchronize(s));
    if (!done) return set_error(GPURT_E_UNSUPPORTED, "BVH deeper than the supported level count");
    if (h_counters[1] & 1u) return set_error(GPURT_E_INVALID, "triangle index >= n_verts");
    if (h_counters[1] & 2u) return set_error(GPURT_E_DEVICE, "wide-node capacity exceeded");
    int depth = 0;
    while (depth <= kMaxLevels && h_lvl[depth] > 0) depth++;
    // shrink the node array to the nodes in use
    const uint32_t n_nodes = h_counters[0];
    void* nodes = b.nodes;
    void* tight = nullptr;
    if (hipMalloc(&tight, node_bytes(W) * n_nodes) == hipSuccess) {
        if (hipMemcpyAsync(tight, nodes, node_bytes(W) * n_nodes, hipMemcpyDeviceToDevice, s) == hipSuccess &&
            hipStreamSynchronize(s) == hipSuccess) {
            nodes = tight;  // b frees the large array
        } else {
            (void)hipFree(tight);
            (void)hipGetLastError();
            b.release(nodes);
        }
    } else {
        (void)hipGetLastError();
        b.release(nodes);
    }
    uint32_t u = st[6];
    u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float scale;
    std::memcpy(&scale, &u, 4);
    out->nodes = nodes;
    out->tris = (float4*)b.release(b.tris);
    out->leaf_order = (uint32_t*)b.release(b.leaf_order);
    out->n_nodes = n_nodes;
    out->node_capacity = n_nodes;
    out->depth = depth;
    out->width = W;
    out->leaf_size = leaf_size;
    out->n_tris = b.N;
    out->scale = scale;
    out->pad = scale * 0x1p-18f;
    out->level_cnt.assign(h_lvl.begin(), h_lvl.begin() + depth);
    return GPURT_OK;
}

}  // namespace gpurt_impl
