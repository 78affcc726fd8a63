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

// ------------------------------------------------------------------------------------
// host orchestration
// ------------------------------------------------------------------------------------

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

static int dev_alloc(DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return set_error(GPURT_E_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
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

int run_treelets(int N, int treelet_size, int rounds, const uint32_t* left, const uint32_t* right,
                 uint32_t* left_rw, uint32_t* right_rw, uint2* range, uint32_t* node_parent,
                 uint32_t* leaf_parent, const uint8_t* is_root, const float4* saabb, float4* nbox,
                 uint32_t* flags, hipStream_t s);

// Builds one BVH per segment into one node array and one packed-triangle array.
// With n_seg == 1 the result is stored in *out (plain build).
int build_segments(gpurt_ctx ctx, const SegDesc* host_segs, int n_seg, int W, int leaf_size,
                   int treelet_size, int treelet_rounds, hipStream_t s, bool /*sync_host*/,
                   gpurt_bvh_s* out) {
    uint64_t N64 = 0;
    for (int i = 0; i < n_seg; i++) N64 += host_segs[i].n_tris;
    if (N64 == 0 || N64 > kMaxLeafFirst - 8)
        return set_error(GPURT_E_INVALID, "triangle count out of range (1 .. 2^28-9)");
    const uint32_t N = (uint32_t)N64;
    (void)ctx;

    DevBuf b_segs, b_aabb_lo, b_aabb_hi, b_tri_v, b_segstat, b_keys0, b_keys1, b_vals0, b_vals1,
        b_hist, b_saabb, b_left, b_right, b_range, b_nparent, b_lparent, b_seg_root, b_is_root,
        b_nbox, b_flags, b_wsrc, b_wseg, b_lvl, b_counters;
    const uint32_t nblocks_sort = cdiv(N, kSortTile);
    const size_t NI = N > 1 ? N - 1 : 1;
    const uint32_t capacity = (uint32_t)std::max<uint64_t>(NI + n_seg, 1);
    int rc;
#define ALLOC(buf, bytes) \
    if ((rc = dev_alloc(buf, bytes)) != GPURT_OK) return rc
    ALLOC(b_segs, sizeof(SegDesc) * n_seg);
    ALLOC(b_aabb_lo, sizeof(float4) * N);
    ALLOC(b_aabb_hi, sizeof(float4) * N);
    ALLOC(b_tri_v, sizeof(float4) * 3 * (size_t)N);
    ALLOC(b_segstat, sizeof(uint32_t) * kStat * n_seg);
    ALLOC(b_keys0, sizeof(uint64_t) * N);
    ALLOC(b_keys1, sizeof(uint64_t) * N);
    ALLOC(b_vals0, sizeof(uint32_t) * N);
    ALLOC(b_vals1, sizeof(uint32_t) * N);
    ALLOC(b_hist, sizeof(uint32_t) * 256 * (size_t)nblocks_sort);
    ALLOC(b_saabb, sizeof(float4) * 2 * (size_t)N);
    ALLOC(b_left, sizeof(uint32_t) * NI);
    ALLOC(b_right, sizeof(uint32_t) * NI);
    ALLOC(b_range, sizeof(uint2) * NI);
    ALLOC(b_nparent, sizeof(uint32_t) * NI);
    ALLOC(b_lparent, sizeof(uint32_t) * N);
    ALLOC(b_seg_root, sizeof(uint32_t) * n_seg);
    ALLOC(b_is_root, NI);
    ALLOC(b_nbox, sizeof(float4) * 2 * NI);
    ALLOC(b_flags, sizeof(uint32_t) * NI);
    ALLOC(b_wsrc, sizeof(uint32_t) * capacity);
    ALLOC(b_wseg, sizeof(uint32_t) * capacity);
    ALLOC(b_lvl, sizeof(uint32_t) * (kMaxLevels + 1));
    ALLOC(b_counters, sizeof(uint32_t) * 4);

    // product buffers (owned by the bvh)
    void* nodes = nullptr;
    float4* tris = nullptr;
    uint32_t* leaf_order = nullptr;
    auto fail_free = [&]() {
        if (nodes) (void)hipFree(nodes);
        if (tris) (void)hipFree(tris);
        if (leaf_order) (void)hipFree(leaf_order);
    };
    if (hipMalloc(&nodes, node_bytes(W) * capacity) != hipSuccess ||
        hipMalloc((void**)&tris, sizeof(float4) * 3 * (size_t)N) != hipSuccess ||
        hipMalloc((void**)&leaf_order, sizeof(uint32_t) * N) != hipSuccess) {
        (void)hipGetLastError();
        fail_free();
        return set_error(GPURT_E_OOM, "hipMalloc of BVH storage failed");
    }
#undef ALLOC

    auto* segs_d = (SegDesc*)b_segs.p;
    auto* aabb_lo = (float4*)b_aabb_lo.p;
    auto* aabb_hi = (float4*)b_aabb_hi.p;
    auto* tri_v = (float4*)b_tri_v.p;
    auto* segstat = (uint32_t*)b_segstat.p;
    auto* keys0 = (uint64_t*)b_keys0.p;
    auto* keys1 = (uint64_t*)b_keys1.p;
    auto* vals0 = (uint32_t*)b_vals0.p;
    auto* vals1 = (uint32_t*)b_vals1.p;
    auto* hist = (uint32_t*)b_hist.p;
    auto* saabb = (float4*)b_saabb.p;
    auto* left = (uint32_t*)b_left.p;
    auto* right = (uint32_t*)b_right.p;
    auto* range = (uint2*)b_range.p;
    auto* nparent = (uint32_t*)b_nparent.p;
    auto* lparent = (uint32_t*)b_lparent.p;
    auto* seg_root = (uint32_t*)b_seg_root.p;
    auto* is_root = (uint8_t*)b_is_root.p;
    auto* nbox = (float4*)b_nbox.p;
    auto* flags = (uint32_t*)b_flags.p;
    auto* wsrc = (uint32_t*)b_wsrc.p;
    auto* wseg = (uint32_t*)b_wseg.p;
    auto* lvl = (uint32_t*)b_lvl.p;
    auto* counters = (uint32_t*)b_counters.p;

#define CHECK_LAUNCH()                                        \
    do {                                                      \
        hipError_t e_ = hipGetLastError();                    \
        if (e_ != hipSuccess) {                               \
            fail_free();                                      \
            return hip_error(e_, "kernel launch (build)");    \
        }                                                     \
    } while (0)

    if (hipMemcpyAsync(segs_d, host_segs, sizeof(SegDesc) * n_seg, hipMemcpyHostToDevice, s) !=
        hipSuccess) {
        fail_free();
        return hip_error(hipGetLastError(), "hipMemcpyAsync(segments)");
    }
    const unsigned gN = cdiv(N, kBlock);
    hipLaunchKernelGGL(k_init_build, dim3(cdiv(std::max(n_seg, kMaxLevels + 1), kBlock)), dim3(kBlock),
                       0, s, segs_d, n_seg, segstat, seg_root, lvl, counters);
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(gN, 4096u)), dim3(kBlock), 0, s, lparent, (size_t)N,
                       kEmpty);
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(cdiv(NI, kBlock), 4096u)), dim3(kBlock), 0, s,
                       nparent, NI, kEmpty);
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(cdiv(NI, kBlock), 4096u)), dim3(kBlock), 0, s, flags,
                       NI, 0u);
    hipLaunchKernelGGL(k_prep, dim3(gN), dim3(kBlock), 0, s, segs_d, n_seg, N, aabb_lo, aabb_hi,
                       tri_v, segstat, counters);
    hipLaunchKernelGGL(k_morton, dim3(gN), dim3(kBlock), 0, s, N, aabb_lo, aabb_hi, segstat, keys0,
                       vals0);
    CHECK_LAUNCH();

    // radix sort over the bits in use: 30 Morton bits + mesh-id bits
    int seg_bits = 0;
    while ((1u << seg_bits) < (uint32_t)n_seg) seg_bits++;
    const int key_bits = 30 + seg_bits;
    uint64_t* kin = keys0;
    uint64_t* kout = keys1;
    uint32_t* vin = vals0;
    uint32_t* vout = vals1;
    for (int shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks_sort), dim3(kBlock), 0, s, kin, N, shift, hist,
                           nblocks_sort);
        hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, s, hist, 256u * nblocks_sort);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks_sort), dim3(kBlock), 0, s, kin, vin, kout,
                           vout, N, shift, hist, nblocks_sort);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    CHECK_LAUNCH();

    hipLaunchKernelGGL(k_pack, dim3(gN), dim3(kBlock), 0, s, N, vin, aabb_lo, aabb_hi, tri_v, saabb,
                       tris, leaf_order);
    if (N > 1) {
        hipLaunchKernelGGL(k_karras, dim3(cdiv(N - 1, kBlock)), dim3(kBlock), 0, s, (int)N, kin,
                           segs_d, left, right, range, nparent, lparent, seg_root, is_root);
        hipLaunchKernelGGL(k_bottom_up, dim3(gN), dim3(kBlock), 0, s, (int)N, lparent, nparent, left,
                           right, is_root, saabb, nbox, flags);
        if (treelet_size > 0) {
            int trc = run_treelets((int)N, treelet_size, treelet_rounds, left, right, left, right,
                                   range, nparent, lparent, is_root, saabb, nbox, flags, s);
            if (trc != GPURT_OK) {
                fail_free();
                return trc;
            }
        }
    }
    CHECK_LAUNCH();

    // level 0 = one wide node per mesh root
    std::vector<uint32_t> h_roots(n_seg), h_segid(n_seg);
    if (hipMemcpyAsync(wsrc, seg_root, sizeof(uint32_t) * n_seg, hipMemcpyDeviceToDevice, s) !=
        hipSuccess) {
        fail_free();
        return hip_error(hipGetLastError(), "hipMemcpyAsync(roots)");
    }
    for (int i = 0; i < n_seg; i++) h_segid[i] = (uint32_t)i;
    if (hipMemcpyAsync(wseg, h_segid.data(), sizeof(uint32_t) * n_seg, hipMemcpyHostToDevice, s) !=
        hipSuccess) {
        fail_free();
        return hip_error(hipGetLastError(), "hipMemcpyAsync(segment ids)");
    }

    CollapseArgs ca;
    ca.left = left;
    ca.right = right;
    ca.range = range;
    ca.saabb = saabb;
    ca.nbox = nbox;
    ca.segstat = segstat;
    ca.wsrc = wsrc;
    ca.wseg = wseg;
    ca.lvl_cnt = lvl;
    ca.counters = counters;
    ca.nodes = nodes;
    ca.capacity = capacity;
    ca.leaf_size = leaf_size;
    const unsigned cgrid = std::min(cdiv(capacity, kBlock), 2048u);
    std::vector<uint32_t> h_lvl(kMaxLevels + 1, 0);
    int level = 0;
    bool done = false;
    // launch levels in groups, checking after each group whether the tree has ended
    while (!done && level < kMaxLevels) {
        int group_end = std::min(level + 8, kMaxLevels);
        for (; level < group_end; level++) {
            ca.level = level;
            collapse_level(W, ca, cgrid, s);
        }
        CHECK_LAUNCH();
        if (hipMemcpyAsync(h_lvl.data(), lvl, sizeof(uint32_t) * (kMaxLevels + 1),
                           hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            fail_free();
            return hip_error(hipGetLastError(), "reading level counts");
        }
        done = h_lvl[level] == 0;
    }
    uint32_t h_counters[4] = {0, 0, 0, 0};
    if (hipMemcpyAsync(h_counters, counters, sizeof(h_counters), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        fail_free();
        return hip_error(hipGetLastError(), "reading build counters");
    }
    if (!done) {
        fail_free();
        return set_error(GPURT_E_UNSUPPORTED, "BVH deeper than the supported level count");
    }
    if (h_counters[1] & 1u) {
        fail_free();
        return set_error(GPURT_E_INVALID, "triangle index >= n_verts");
    }
    if (h_counters[1] & 2u) {
        fail_free();
        return set_error(GPURT_E_DEVICE, "wide-node capacity exceeded");
    }
    int depth = 0;
    while (depth <= kMaxLevels && h_lvl[depth] > 0) depth++;

    // shrink the node array to the nodes in use
    const uint32_t n_nodes = h_counters[0];
    void* tight = nullptr;
    if (hipMalloc(&tight, node_bytes(W) * n_nodes) == hipSuccess) {
        if (hipMemcpyAsync(tight, nodes, node_bytes(W) * n_nodes, hipMemcpyDeviceToDevice, s) ==
                hipSuccess &&
            hipStreamSynchronize(s) == hipSuccess) {
            (void)hipFree(nodes);
            nodes = tight;
        } else {
            (void)hipFree(tight);
            (void)hipGetLastError();
        }
    } else {
        (void)hipGetLastError();
    }
    if (hipStreamSynchronize(s) != hipSuccess) {
        fail_free();
        return hip_error(hipGetLastError(), "build");
    }
    float scale = 0.f;
    {
        uint32_t st[kStat];
        if (hipMemcpy(st, segstat, sizeof(st), hipMemcpyDeviceToHost) != hipSuccess) {
            fail_free();
            return hip_error(hipGetLastError(), "reading mesh scale");
        }
        uint32_t u = st[6];
        u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
        std::memcpy(&scale, &u, 4);
    }
    out->nodes = nodes;
    out->tris = tris;
    out->leaf_order = leaf_order;
    out->n_nodes = n_nodes;
    out->node_capacity = n_nodes;
    out->depth = depth;
    out->width = W;
    out->leaf_size = leaf_size;
    out->n_tris = N;
    out->scale = scale;
    out->pad = scale * 0x1p-18f;
    return GPURT_OK;
#undef CHECK_LAUNCH
}

}  // namespace gpurt_impl
