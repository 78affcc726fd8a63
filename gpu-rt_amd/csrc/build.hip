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
//   k_boxes_level  node AABBs bottom-up, one launch per common-prefix length (no atomics)
//   k_treelet      optional SAH treelet restructuring (treelet.hip)
//   k_collapse     one launch per wide level: a node opens the child with the largest
//                  surface area until it has W children; subtrees of <= leaf_size
//                  triangles become leaves (contiguous ranges of packed triangles)
#include <algorithm>
#include <cstring>
#include <vector>

#include "build_common.h"

namespace gpurt_impl {

// ------------------------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------------------------

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


__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_init_build(const SegDesc* segs, int n_seg, uint32_t* segstat_rep, uint32_t* seg_root,
                             uint32_t* lvl_cnt, uint32_t* counters) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_seg * kRep) {
        uint32_t* st = segstat_rep + (size_t)i * kStat;
        for (int k = 0; k < 3; k++) {
            st[k] = 0xFFFFFFFFu;
            st[3 + k] = 0u;
        }
        st[6] = 0u;
        st[7] = 0u;
    }
    if (i < n_seg) seg_root[i] = kLeafBit | segs[i].tri_offset;  // single-triangle mesh: leaf root
    if (i <= kMaxLevels) lvl_cnt[i] = (i == 0) ? (uint32_t)n_seg : 0u;
    if (i < 4) counters[i] = (i == 0) ? (uint32_t)n_seg : 0u;  // [0] wide nodes, [1] error
}

// folds the kRep replicas of every mesh's statistics into segstat
__global__ void k_reduce_stats(int n_seg, const uint32_t* __restrict__ segstat_rep,
                               uint32_t* __restrict__ segstat) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_seg * kStat) return;
    const int sgi = i / kStat, k = i % kStat;
    const bool is_min = k < 3;
    uint32_t v = is_min ? 0xFFFFFFFFu : 0u;
    for (int r = 0; r < kRep; r++) {
        uint32_t x = segstat_rep[((size_t)r * n_seg + sgi) * kStat + k];
        v = is_min ? min(v, x) : max(v, x);
    }
    segstat[i] = v;
}

// ------------------------------------------------------------------------------------
// k_prep: AABBs + per-mesh centroid bounds
// ------------------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_prep(const SegDesc* __restrict__ segs, int n_seg,
                                                 uint32_t N, float4* __restrict__ aabb_lo,
                                                 float4* __restrict__ aabb_hi,
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
        bool finite = true;
        for (int k = 0; k < 3; k++) {
            const float* p = sd.verts + (size_t)vi[k] * sd.vstride;
            v[k][0] = p[0]; v[k][1] = p[1]; v[k][2] = p[2];
            finite = finite && isfinite(v[k][0]) && isfinite(v[k][1]) && isfinite(v[k][2]);
        }
        // a non-finite position is reported (the build fails with GPURT_E_INVALID); the
        // triangle's box is then the origin so every later stage works on finite boxes
        if (!finite) atomicOr(&counters[1], 4u);
        for (int a = 0; a < 3; a++) {
            for (int k = 0; k < 3; k++) {
                const float x = finite ? v[k][a] : 0.0f;
                lo[a] = fminf(lo[a], x);
                hi[a] = fmaxf(hi[a], x);
                amax = fmaxf(amax, fabsf(x));
            }
            cen[a] = (lo[a] + hi[a]) * 0.5f;
            cenx[a] = cen[a];
        }
        aabb_lo[g] = make_float4(lo[0], lo[1], lo[2], __uint_as_float((uint32_t)s));
        aabb_hi[g] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    }
    // reduce over the wave when it covers one mesh, else per lane; this block's replica
    segstat += (size_t)(blockIdx.x % kRep) * n_seg * kStat;
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

// Exclusive scan of `total` counts in three launches: per-tile sums, one block scanning
// the tile sums, then per-tile scans offset by their tile's prefix.
constexpr int kScanTile = 4096;

__global__ __launch_bounds__(kBlock) void k_scan_tiles(const uint32_t* __restrict__ a, uint32_t total,
                                                       uint32_t* __restrict__ tile_sum) {
    __shared__ uint32_t red[kBlock];
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int k = 0; k < kScanTile / kBlock; k++) {
        uint32_t i = base + k * kBlock + threadIdx.x;
        if (i < total) s += a[i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* __restrict__ t, uint32_t n) {
    __shared__ uint32_t part[1024];
    uint32_t chunk = (n + 1023) / 1024;
    uint32_t beg = threadIdx.x * chunk, end = min(beg + chunk, n);
    uint32_t s = 0;
    for (uint32_t i = beg; i < end; i++) s += t[i];
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
        uint32_t x = t[i];
        t[i] = run;
        run += x;
    }
}

// one tile of kScanTile entries per block: 16 consecutive entries per thread
__global__ __launch_bounds__(kBlock) void k_scan_apply(uint32_t* __restrict__ a, uint32_t total,
                                                       const uint32_t* __restrict__ tile_off) {
    __shared__ uint32_t part[kBlock];
    constexpr int kPer = kScanTile / kBlock;
    const uint32_t beg = blockIdx.x * kScanTile + threadIdx.x * kPer;
    uint32_t v[kPer];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        v[k] = (beg + k < total) ? a[beg + k] : 0u;
        s += v[k];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        uint32_t x = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t run = tile_off[blockIdx.x] + part[threadIdx.x] - s;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        if (beg + k < total) a[beg + k] = run;
        run += v[k];
    }
}

// Scatter of one radix pass.  The block's tile (kSortTile pairs) is first ordered by
// digit in LDS, stably, and then written out in that order, so each digit's run of the
// tile (16 pairs on average) leaves as one contiguous stretch of the output instead of
// one isolated 12-byte write per pair.  Each wave ranks its own 1024 consecutive pairs:
// per round of 64, an 8-ballot multisplit ranks the lanes of equal digit, and a
// wave-private LDS counter per digit (read by all lanes, advanced by each group's first
// lane; a wave's LDS operations complete in order) carries the rank across rounds, so the
// rounds need no block barrier.  One scan over (digit, wave) counts then gives every
// wave's first slot per digit.
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint64_t* __restrict__ kin,
                                                          const uint32_t* __restrict__ vin,
                                                          uint64_t* __restrict__ kout,
                                                          uint32_t* __restrict__ vout, uint32_t n,
                                                          int shift,
                                                          const uint32_t* __restrict__ offs,
                                                          uint32_t nblocks) {
    constexpr int kWaves = kBlock / 64;
    constexpr uint32_t kWaveItems = kSortTile / kWaves;  // consecutive pairs per wave
    static_assert(kWaveItems == 64 * kSortItems, "one pair per lane and round");
    __shared__ uint64_t skey[kSortTile];
    __shared__ uint32_t sval[kSortTile];
    __shared__ uint32_t gbase[256];          // this block's first output slot of digit d
    __shared__ uint32_t lstart[256];         // the tile's first LDS slot of digit d
    __shared__ uint32_t wcnt[kWaves][256];   // a wave's pairs of digit d, then its first slot
    __shared__ uint32_t wsum[kWaves];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t tile = blockIdx.x * kSortTile;
    const uint32_t m = min((uint32_t)kSortTile, n - tile);
    gbase[tid] = offs[(size_t)tid * nblocks + blockIdx.x];
    for (int w = 0; w < kWaves; w++) wcnt[w][tid] = 0;
    __syncthreads();
    uint32_t* const mycnt = wcnt[wave];
    const uint32_t wbase = (uint32_t)wave * kWaveItems;
    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    uint64_t key[kSortItems];
    uint32_t val[kSortItems], rank[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t i = wbase + (uint32_t)r * 64u + (uint32_t)lane;
        const bool valid = i < m;
        key[r] = valid ? kin[tile + i] : 0ull;
        val[r] = valid ? vin[tile + i] : 0u;
        const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t rk = (uint32_t)__popcll(peers & lt_mask);
        const uint32_t before = mycnt[d];
        rank[r] = before + rk;
        if (valid && rk == 0) mycnt[d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // digit tid: the tile's count and every wave's offset inside the digit's run
    uint32_t c[kWaves], tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
        c[w] = wcnt[w][tid];
        tot += c[w];
    }
    // exclusive scan of the 256 digit totals: within each wave by shuffles, then across
    uint32_t incl = tot;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t x = __shfl_up(incl, off);
        if (lane >= off) incl += x;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t start = incl - tot;
    for (int w = 0; w < wave; w++) start += wsum[w];
    lstart[tid] = start;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
        wcnt[w][tid] = start;
        start += c[w];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        if (wbase + (uint32_t)r * 64u + (uint32_t)lane < m) {
            const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
            const uint32_t pos = mycnt[d] + rank[r];
            skey[pos] = key[r];
            sval[pos] = val[r];
        }
    }
    __syncthreads();
    for (uint32_t j = tid; j < m; j += kBlock) {
        const uint64_t k = skey[j];
        const uint32_t d = (uint32_t)(k >> shift) & 255u;
        const uint32_t o = gbase[d] + (j - lstart[d]);
        kout[o] = k;
        vout[o] = sval[j];
    }
}

size_t radix_sort_scratch_words(uint32_t n) {
    const uint32_t nblocks = cdiv(n, kSortTile);
    return 256 * (size_t)nblocks + cdiv(256 * (size_t)nblocks, kScanTile);
}

bool radix_sort_pairs(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint32_t n, int key_bits,
                      uint32_t* scratch, hipStream_t s) {
    const uint32_t nblocks = cdiv(n, kSortTile);
    uint32_t* hist = scratch;  // 256 digit counts per block, digit-major
    uint32_t* scan_tmp = scratch + 256 * (size_t)nblocks;
    const uint32_t hist_n = 256u * nblocks;
    const unsigned scan_tiles = cdiv(hist_n, kScanTile);
    uint64_t *kin = k0, *kout = k1;
    uint32_t *vin = v0, *vout = v1;
    bool in1 = false;
    for (int shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kBlock), 0, s, kin, n, shift, hist, nblocks);
        hipLaunchKernelGGL(k_scan_tiles, dim3(scan_tiles), dim3(kBlock), 0, s, hist, hist_n, scan_tmp);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, scan_tmp, scan_tiles);
        hipLaunchKernelGGL(k_scan_apply, dim3(scan_tiles), dim3(kBlock), 0, s, hist, hist_n, scan_tmp);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kBlock), 0, s, kin, vin, kout, vout, n, shift, hist,
                           nblocks);
        std::swap(kin, kout);
        std::swap(vin, vout);
        in1 = !in1;
    }
    return in1;
}

// ------------------------------------------------------------------------------------
// k_pack: gather triangles and AABBs into sorted order
// ------------------------------------------------------------------------------------

// Triangles in sorted order: sorted primitive p is triangle t = vals[p] - tri_offset of
// mesh s = keys[p] >> 30.  Its vertices are read from the mesh itself (neighbours in
// Morton order share vertices, so these reads mostly hit the caches) rather than
// gathered from a per-triangle copy in input order (two box rows and three vertex rows
// at a random index each), and its box is rebuilt from them exactly as k_prep built it.
__global__ __launch_bounds__(kBlock) void k_pack(uint32_t N, const uint64_t* __restrict__ keys,
                                                 const uint32_t* __restrict__ vals,
                                                 const SegDesc* __restrict__ segs,
                                                 float4* __restrict__ saabb,
                                                 float4* __restrict__ tris,
                                                 uint32_t* __restrict__ leaf_order) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= N) return;
    const uint32_t g = vals[p];
    const uint32_t s = (uint32_t)(keys[p] >> 30);
    const SegDesc sd = segs[s];
    const uint32_t t = g - sd.tri_offset;
    float v[3][3];
    bool finite = true;
    for (int k = 0; k < 3; k++) {
        uint32_t vi = sd.idx[3 * (size_t)t + k];
        if (vi >= sd.n_verts) vi = 0;  // reported by k_prep (the build fails)
        const float* q = sd.verts + (size_t)vi * sd.vstride;
        v[k][0] = q[0]; v[k][1] = q[1]; v[k][2] = q[2];
        finite = finite && isfinite(v[k][0]) && isfinite(v[k][1]) && isfinite(v[k][2]);
    }
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int a = 0; a < 3; a++) {
        for (int k = 0; k < 3; k++) {
            const float x = finite ? v[k][a] : 0.0f;
            lo[a] = fminf(lo[a], x);
            hi[a] = fmaxf(hi[a], x);
        }
    }
    saabb[2 * (size_t)p] = make_float4(lo[0], lo[1], lo[2], __uint_as_float(s));
    saabb[2 * (size_t)p + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    tris[3 * (size_t)p + 0] = make_float4(v[0][0], v[0][1], v[0][2], __uint_as_float(t));
    tris[3 * (size_t)p + 1] = make_float4(v[1][0], v[1][1], v[1][2], 0.0f);
    tris[3 * (size_t)p + 2] = make_float4(v[2][0], v[2][1], v[2][2], 0.0f);
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
                                                   uint8_t* __restrict__ is_root,
                                                   uint8_t* __restrict__ dlev,
                                                   uint32_t* __restrict__ cnt) {
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
    cnt[i] = (uint32_t)(last - first + 1);
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
    // nodes above the mesh roots (their range spans several meshes) belong to no BVH:
    // prefix length 0xFF keeps the box and treelet passes away from them
    dlev[i] = (sf == sl) ? (uint8_t)dnode : (uint8_t)0xFF;
}

// ------------------------------------------------------------------------------------
// internal node boxes and SAH costs
// ------------------------------------------------------------------------------------

// Internal nodes bucketed by prefix length l (a counting sort, order within a bucket
// arbitrary), each length split in two: bucket 2l holds the nodes with more than
// 2 leaf_size triangles, which are the treelet roots, and 2l + 1 the others, so the box
// pass reads buckets 2l and 2l + 1 and the treelet pass only 2l (it used to load and
// skip the small roots: two dependent loads per node, which made up most of a launch on
// the deep levels).  lvl_hist[512] counts, lvl_begin[513] prefix sums, lvl_order the
// nodes.  Each block counts a contiguous range of nodes in LDS and adds to the global
// counts once per bucket it saw: with one block per 256 nodes, the same-address device
// atomics alone took over 1 ms per pass for a 25 M-node batch.  Within a block's range a
// bucket keeps roughly node order, so the per-level passes read neighbouring nodes.
constexpr unsigned kLevelBlocks = 1024;

__device__ __forceinline__ uint32_t level_bucket(uint8_t dlev, uint32_t cnt, uint32_t small_max) {
    return 2u * dlev + (cnt <= small_max ? 1u : 0u);
}

__device__ __forceinline__ void level_range(int n_int, uint32_t& lo, uint32_t& hi) {
    const uint32_t per = (((uint32_t)n_int + gridDim.x - 1) / gridDim.x + kBlock - 1) / kBlock * kBlock;
    lo = min((uint32_t)n_int, blockIdx.x * per);
    hi = min((uint32_t)n_int, lo + per);
}

static_assert(kLevelBuckets == 2 * kBlock, "two buckets per thread in the level kernels");

__global__ __launch_bounds__(kBlock) void k_level_hist(int n_int, const uint8_t* __restrict__ dlev,
                                                       const uint32_t* __restrict__ cnt, uint32_t small_max,
                                                       uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kLevelBuckets];
    h[threadIdx.x] = 0;
    h[threadIdx.x + kBlock] = 0;
    __syncthreads();
    uint32_t lo, hi;
    level_range(n_int, lo, hi);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) atomicAdd(&h[level_bucket(dlev[i], cnt[i], small_max)], 1u);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < kLevelBuckets; k += kBlock)
        if (h[k]) gatomic_add(&hist[k], h[k]);
}

__global__ void k_level_scan(const uint32_t* __restrict__ hist, uint32_t* __restrict__ begin,
                             uint32_t* __restrict__ fill) {
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t k = 0; k < kLevelBuckets; k++) {
            begin[k] = run;
            fill[k] = run;
            run += hist[k];
        }
        begin[kLevelBuckets] = run;
    }
}

__global__ __launch_bounds__(kBlock) void k_level_scatter(int n_int, const uint8_t* __restrict__ dlev,
                                                          const uint32_t* __restrict__ cnt, uint32_t small_max,
                                                          uint32_t* __restrict__ fill,
                                                          uint32_t* __restrict__ order) {
    __shared__ uint32_t h[kLevelBuckets], base[kLevelBuckets];
    h[threadIdx.x] = 0;
    h[threadIdx.x + kBlock] = 0;
    __syncthreads();
    uint32_t lo, hi;
    level_range(n_int, lo, hi);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) atomicAdd(&h[level_bucket(dlev[i], cnt[i], small_max)], 1u);
    __syncthreads();
    // reserve this block's part of every bucket, then place its nodes there
    for (uint32_t k = threadIdx.x; k < kLevelBuckets; k += kBlock) {
        base[k] = h[k] ? gatomic_add(&fill[k], h[k]) : 0u;
        h[k] = 0;
    }
    __syncthreads();
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const uint32_t l = level_bucket(dlev[i], cnt[i], small_max);
        order[base[l] + atomicAdd(&h[l], 1u)] = i;
    }
}

// Node boxes without atomics or fences, in two parts.
//
// k_boxes_local: block c owns internal nodes [256c, 256c + 256).  A node whose leaf range
// lies in one such chunk has its whole subtree there (a Karras node's index is an end of
// its range, so the subtree's internal nodes lie inside the range), so the block computes
// all of these bottom up by itself: rounds in which every node whose children are done
// merges their boxes, with the chunk's boxes and costs kept in LDS (its own global writes
// are not read back, so no cache can hand a lane a stale box) until no node is left.  A
// round always finishes the deepest pending node, so the loop ends.  This covers ~90% of
// the nodes in one launch.
__device__ __forceinline__ bool box_in_chunk(uint2 r) { return (r.x >> 8) == (r.y >> 8); }
static_assert(kBlock == 256, "k_boxes_local chunks are 256 nodes");

__global__ __launch_bounds__(kBlock) void k_boxes_local(uint32_t n_int, const uint2* __restrict__ range,
                                                        const uint32_t* __restrict__ left,
                                                        const uint32_t* __restrict__ right,
                                                        const uint32_t* __restrict__ cnt,
                                                        const uint8_t* __restrict__ dlev,
                                                        const float4* __restrict__ saabb,
                                                        float4* __restrict__ nbox, float* __restrict__ cost,
                                                        int leaf_size) {
    __shared__ float4 blo[kBlock], bhi[kBlock];
    __shared__ float bcost[kBlock];
    __shared__ uint8_t done[kBlock];
    const uint32_t c0 = blockIdx.x * kBlock;
    const uint32_t i = c0 + threadIdx.x;
    bool todo = false;
    uint32_t L = 0, R = 0;
    if (i < n_int && dlev[i] != 0xFF) {
        todo = box_in_chunk(range[i]);
        L = left[i];
        R = right[i];
    }
    done[threadIdx.x] = 0;
    __syncthreads();
    while (__syncthreads_or(todo)) {
        const bool ready = todo && ((L & kLeafBit) || done[L - c0]) && ((R & kLeafBit) || done[R - c0]);
        float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
        float c = 0.f;
        if (ready) {
            float4 al, ah, bl, bh;
            float cl, cr;
            if (L & kLeafBit) {
                const size_t p = L & ~kLeafBit;
                al = saabb[2 * p];
                ah = saabb[2 * p + 1];
                cl = kCt * box_area(al, ah);
            } else {
                al = blo[L - c0];
                ah = bhi[L - c0];
                cl = bcost[L - c0];
            }
            if (R & kLeafBit) {
                const size_t p = R & ~kLeafBit;
                bl = saabb[2 * p];
                bh = saabb[2 * p + 1];
                cr = kCt * box_area(bl, bh);
            } else {
                bl = blo[R - c0];
                bh = bhi[R - c0];
                cr = bcost[R - c0];
            }
            lo = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
            hi = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
            c = sah_node_cost(lo, hi, cnt[i], cl, cr, leaf_size);
        }
        __syncthreads();  // every read of this round before any write
        if (ready) {
            blo[threadIdx.x] = lo;
            bhi[threadIdx.x] = hi;
            bcost[threadIdx.x] = c;
            done[threadIdx.x] = 1;
            nbox[2 * (size_t)i] = lo;
            nbox[2 * (size_t)i + 1] = hi;
            cost[i] = c;
            todo = false;
        }
    }
}

// k_boxes_level: the remaining nodes (ranges across chunks).  A child's common-prefix
// length (Karras' delta of its range) is strictly greater than its parent's, so
// launching once per delta value from the largest down makes every child box final
// before its parent reads it.
__global__ __launch_bounds__(kBlock) void k_boxes_level(int level, const uint32_t* __restrict__ lvl_begin,
                                                        const uint32_t* __restrict__ lvl_order,
                                                        const uint2* __restrict__ range,
                                                        const uint32_t* __restrict__ left,
                                                        const uint32_t* __restrict__ right,
                                                        const uint32_t* __restrict__ cnt,
                                                        const float4* __restrict__ saabb,
                                                        float4* __restrict__ nbox, float* __restrict__ cost,
                                                        int leaf_size) {
    const uint32_t beg = lvl_begin[2 * level], end = lvl_begin[2 * level + 2];  // both size buckets
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t k = beg + blockIdx.x * kBlock + threadIdx.x; k < end; k += stride) {
        const uint32_t i = lvl_order[k];
        if (box_in_chunk(range[i])) continue;  // done by k_boxes_local
        float4 al, ah, bl, bh;
        const uint32_t L = left[i], R = right[i];
        ref_box(L, saabb, nbox, al, ah);
        ref_box(R, saabb, nbox, bl, bh);
        const float4 lo = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
        const float4 hi = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
        nbox[2 * (size_t)i] = lo;
        nbox[2 * (size_t)i + 1] = hi;
        cost[i] = sah_node_cost(lo, hi, cnt[i], ref_cost(L, saabb, cost), ref_cost(R, saabb, cost), leaf_size);
    }
}

// ------------------------------------------------------------------------------------
// k_collapse: binary tree -> W-wide nodes, one launch per wide level
// ------------------------------------------------------------------------------------

struct CollapseArgs {
    const uint32_t* left;
    const uint32_t* right;
    const uint2* range;
    const uint32_t* cnt;  // triangles under each internal node
    const float4* saabb;
    const float4* nbox;
    const uint32_t* segstat;
    uint32_t* wsrc;  // wide node -> binary ref
    uint32_t* wseg;  // wide node -> mesh
    uint32_t* lvl_cnt;
    uint32_t* counters;  // [0] wide nodes allocated, [1] error bits, [2] triangles re-packed
    void* nodes;
    uint32_t capacity;
    int leaf_size;
    int level;
    // re-pack (after treelet restructuring a leaf's triangles are no longer one sorted
    // range): leaves get fresh contiguous slots and their triangles are copied there
    int repack;
    const float4* tris_in;
    const uint32_t* order_in;
    float4* tris_out;
    uint32_t* order_out;
};

__device__ __forceinline__ uint32_t ref_count(uint32_t r, const uint32_t* cnt) {
    return (r & kLeafBit) ? 1u : cnt[r];
}
__device__ __forceinline__ uint32_t ref_first(uint32_t r, const uint2* range) {
    if (r & kLeafBit) return r & ~kLeafBit;
    return range[r].x;
}

// Copies the <= leaf_size triangles under binary ref r to the slots from `base`.
__device__ void repack_leaf(const CollapseArgs& a, uint32_t r, uint32_t base) {
    uint32_t stack[8];
    int sp = 0;
    uint32_t out = base;
    stack[sp++] = r;
    while (sp > 0) {
        const uint32_t x = stack[--sp];
        if (x & kLeafBit) {
            const size_t p = x & ~kLeafBit;
            a.tris_out[3 * (size_t)out] = a.tris_in[3 * p];
            a.tris_out[3 * (size_t)out + 1] = a.tris_in[3 * p + 1];
            a.tris_out[3 * (size_t)out + 2] = a.tris_in[3 * p + 2];
            a.order_out[out] = a.order_in[p];
            out++;
        } else if (sp + 2 <= 8) {
            stack[sp++] = a.right[x];
            stack[sp++] = a.left[x];
        }
    }
}

// Wave-aggregated allocation: each lane asks for n slots of *counter (any lane may ask
// for 0; all 64 lanes take part), one device atomic per wave returns the wave's block
// and the lane gets its part in lane order.  Many same-address device atomics, one per
// node, serialise at the memory side.
__device__ __forceinline__ uint32_t wave_alloc(uint32_t* counter, uint32_t n, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t x = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= (uint32_t)d) x += y;
    }
    total = __shfl(x, 63);
    uint32_t base = 0;
    if (lane == 63u && total) base = gatomic_add(counter, total);
    base = __shfl(base, 63);
    return base + x - n;
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_collapse(CollapseArgs a) {
    // the graph of a batch launches every possible level: most are empty or small, so
    // surplus blocks leave before anything else
    const uint32_t cnt = a.lvl_cnt[a.level];
    if (blockIdx.x * kBlock >= cnt) return;
    uint32_t begin = 0;
    for (int l = 0; l < a.level; l++) begin += a.lvl_cnt[l];
    WNode<W>* nodes = reinterpret_cast<WNode<W>*>(a.nodes);
    // every lane runs every iteration (slot allocation is wave-wide); lanes past the
    // level's end take part with nothing to allocate
    for (uint32_t k0 = blockIdx.x * kBlock; k0 < cnt; k0 += gridDim.x * kBlock) {
        const uint32_t k = k0 + threadIdx.x;
        const bool act = k < cnt;
        uint32_t wi = begin + (act ? k : 0u);
        uint32_t bref = a.wsrc[wi];
        uint32_t seg = a.wseg[wi];
        float pad = o2f(a.segstat[seg * kStat + 6]) * 0x1p-18f;
        // The wide node's children, kept in registers with their triangle counts and
        // boxes: a child is loaded once, together with its own children when it is an
        // internal node, so opening it (the child with the largest area among those with
        // more than leaf_size triangles, the lowest slot on ties) costs one round of
        // loads instead of re-reading every child's count and box.  Slots are addressed
        // with compile-time indices only, so nothing spills to scratch.
        uint32_t refs[W], rc[W], rl[W], rr[W];
        float4 rlo[W], rhi[W];
        float rar[W];
        auto load_child = [&](uint32_t r, uint32_t& ref, uint32_t& c, uint32_t& l, uint32_t& rgt, float4& lo,
                              float4& hi, float& ar) {
            ref = r;
            c = ref_count(r, a.cnt);
            ref_box(r, a.saabb, a.nbox, lo, hi);
            const bool internal = !(r & kLeafBit);
            l = internal ? a.left[r] : 0u;
            rgt = internal ? a.right[r] : 0u;
            ar = c > (uint32_t)a.leaf_size ? box_area(lo, hi) : -1.0f;
        };
        int n = 0;
        if (!act) {
            // nothing: n stays 0
        } else {
            uint32_t bc, bl, br;
            float4 blo, bhi;
            float bar;
            load_child(bref, refs[0], bc, bl, br, blo, bhi, bar);
            if (bc <= (uint32_t)a.leaf_size) {
                rc[0] = bc;
                rlo[0] = blo;
                rhi[0] = bhi;
                n = 1;
            } else {
                load_child(bl, refs[0], rc[0], rl[0], rr[0], rlo[0], rhi[0], rar[0]);
                load_child(br, refs[1], rc[1], rl[1], rr[1], rlo[1], rhi[1], rar[1]);
                n = 2;
                while (n < W) {
                    int best = -1;
                    float best_area = -1.0f;
#pragma unroll
                    for (int j = 0; j < W; j++) {
                        if (j < n && rar[j] > best_area) {
                            best_area = rar[j];
                            best = j;
                        }
                    }
                    if (best < 0) break;
                    uint32_t ol = 0, orr = 0;
#pragma unroll
                    for (int j = 0; j < W; j++) {
                        if (j == best) {
                            ol = rl[j];
                            orr = rr[j];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < W; j++) {
                        if (j == best) load_child(ol, refs[j], rc[j], rl[j], rr[j], rlo[j], rhi[j], rar[j]);
                        if (j == n) load_child(orr, refs[j], rc[j], rl[j], rr[j], rlo[j], rhi[j], rar[j]);
                    }
                    n++;
                }
            }
        }
        WNode<W> out;
        uint32_t n_internal = 0, n_repack = 0;
#pragma unroll
        for (int j = 0; j < W; j++) {
            if (j < n) {
                const uint32_t c = rc[j];
                if (c <= (uint32_t)a.leaf_size) n_repack += c;
                else n_internal++;
            }
        }
        uint32_t wave_total = 0;
        uint32_t rp = a.repack ? wave_alloc(&a.counters[2], n_repack, wave_total) : 0u;
#pragma unroll
        for (int j = 0; j < W; j++) {
            if (j < n) {
                const uint32_t r = refs[j];
                const float4 lo = rlo[j], hi = rhi[j];
                out.lox[j] = lo.x - pad; out.loy[j] = lo.y - pad; out.loz[j] = lo.z - pad;
                out.hix[j] = hi.x + pad; out.hiy[j] = hi.y + pad; out.hiz[j] = hi.z + pad;
                const uint32_t c = rc[j];
                out.aux[j] = c;
                if (c <= (uint32_t)a.leaf_size) {
                    uint32_t first;
                    if (a.repack) {
                        repack_leaf(a, r, rp);
                        first = rp;
                        rp += c;
                    } else {
                        first = ref_first(r, a.range);
                    }
                    out.child[j] = make_leaf(first, c);
                }
            } else {
                out.lox[j] = out.loy[j] = out.loz[j] = INFINITY;
                out.hix[j] = out.hiy[j] = out.hiz[j] = -INFINITY;
                out.child[j] = kEmpty;
                out.aux[j] = 0;
            }
        }
        const uint32_t base = wave_alloc(&a.counters[0], n_internal, wave_total);
        if ((threadIdx.x & 63u) == 63u && wave_total) gatomic_add(&a.lvl_cnt[a.level + 1], wave_total);
        if (n_internal) {
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
                        gatomic_or(&a.counters[1], 2u);
                        out.child[j] = kEmpty;
                    }
                }
            }
        }
        if (act) nodes[wi] = out;
    }
}

// SAH cost of a finished wide tree: sum over nodes of kCi * area for internal children
// and kCt * count * area for leaves, accumulated in double per wave then atomically.
template <int W>
__global__ __launch_bounds__(kBlock) void k_sah_cost(uint32_t n_nodes, const WNode<W>* __restrict__ nodes,
                                                     double* acc) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    double c = 0.0;
    if (k < n_nodes) {
        const WNode<W>& nd = nodes[k];
#pragma unroll
        for (int j = 0; j < W; j++) {
            if (nd.child[j] == kEmpty) continue;
            const float a = box_area(make_float4(nd.lox[j], nd.loy[j], nd.loz[j], 0.0f),
                                     make_float4(nd.hix[j], nd.hiy[j], nd.hiz[j], 0.0f));
            c += child_is_leaf(nd.child[j]) ? (double)kCt * leaf_count(nd.child[j]) * a : (double)kCi * a;
        }
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0 && c != 0.0) atomicAdd(acc, c);
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
        if (p) (void)(pooled ? hipFreeAsync(p, pool_stream) : hipFree(p));
}

void* BuildBufs::release(void* p) {
    for (auto& a : allocs)
        if (a == p) a = nullptr;
    return p;
}

// All scratch of a build lives in one allocation carved into 256-byte aligned pieces
// (one hipMalloc instead of ~25); the products handed to a BVH are separate.
struct Carver {
    size_t off = 0;
    char* base = nullptr;
    template <typename T>
    void take(T*& p, size_t count) {
        size_t bytes = std::max<size_t>(count * sizeof(T), 256);
        off = (off + 255) & ~size_t(255);
        p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += bytes;
    }
};

static void carve_all(BuildBufs& b, Carver& c, size_t NI) {
    const uint32_t N = b.N;
    const int n_seg = b.n_seg;
    c.take(b.segs, n_seg);
    c.take(b.aabb_lo, N);
    c.take(b.aabb_hi, N);
    c.take(b.segstat_rep, (size_t)kStat * n_seg * kRep);
    c.take(b.segstat, (size_t)kStat * n_seg);
    c.take(b.keys0, N);
    c.take(b.keys1, N);
    c.take(b.vals0, N);
    c.take(b.vals1, N);
    c.take(b.hist, radix_sort_scratch_words(N));  // the sort's digit counts and scan partials
    c.take(b.saabb, 2 * (size_t)N);
    c.take(b.left, NI);
    c.take(b.right, NI);
    c.take(b.range, NI);
    c.take(b.nparent, NI);
    c.take(b.lparent, N);
    c.take(b.seg_root, n_seg);
    c.take(b.is_root, NI);
    c.take(b.dlev, NI);
    c.take(b.cnt, NI);
    c.take(b.cost, NI);
    c.take(b.lvl_order, NI);
    c.take(b.lvl_hist, kLevelBuckets);
    c.take(b.lvl_begin, kLevelBuckets + 1);
    c.take(b.lvl_fill, kLevelBuckets);
    c.take(b.nbox, 2 * NI);
    c.take(b.wsrc, b.capacity);
    c.take(b.wseg, b.capacity);
    c.take(b.lvl, kMaxLevels + 1);
    c.take(b.counters, 4);
}

int build_alloc(const SegDesc* host_segs, int n_seg, int W, int leaf_size, int treelet_size, BuildBufs& b) {
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
    b.leaf_size = leaf_size;
    b.nblocks_sort = cdiv(N, kSortTile);
    const size_t NI = N > 1 ? N - 1 : 1;
    b.capacity = (uint32_t)(NI + n_seg);
    Carver sizing;
    carve_all(b, sizing, NI);
    void* scratch = nullptr;
    if (b.alloc(&scratch, sizing.off) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(GPURT_E_OOM, "allocation of " + std::to_string(sizing.off) + " bytes of build scratch failed");
    }
    Carver c;
    c.base = static_cast<char*>(scratch);
    carve_all(b, c, NI);
    void* nodes = nullptr;
    void* tris = nullptr;
    void* order = nullptr;
    if (b.alloc(&nodes, node_bytes(W) * b.capacity) != hipSuccess ||
        b.alloc(&tris, sizeof(float4) * 3 * (size_t)N) != hipSuccess ||
        b.alloc(&order, sizeof(uint32_t) * (size_t)N) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(GPURT_E_OOM, "allocation of BVH storage failed");  // b frees what it holds
    }
    b.nodes = nodes;
    b.tris = static_cast<float4*>(tris);
    b.leaf_order = static_cast<uint32_t*>(order);
    if (treelet_size > 0) {
        void* t2 = nullptr;
        void* o2 = nullptr;
        if (b.alloc(&t2, sizeof(float4) * 3 * (size_t)N) != hipSuccess ||
            b.alloc(&o2, sizeof(uint32_t) * (size_t)N) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(GPURT_E_OOM, "allocation of re-pack buffers failed");
        }
        b.tris2 = static_cast<float4*>(t2);
        b.leaf_order2 = static_cast<uint32_t*>(o2);
    }
    if (b.pooled) {
        GPURT_HIP(hipMemcpyAsync(b.segs, host_segs, sizeof(SegDesc) * n_seg, hipMemcpyHostToDevice, b.pool_stream));
    } else {
        GPURT_HIP(hipMemcpy(b.segs, host_segs, sizeof(SegDesc) * n_seg, hipMemcpyHostToDevice));
    }
    return GPURT_OK;
}

// Enqueues the build up to (not including) the wide collapse.  No host synchronisation
// and no memset nodes, so the sequence can be captured into a HIP graph.
int build_enqueue_binary(BuildBufs& b, int treelet_size, int treelet_rounds, hipStream_t s) {
    const uint32_t N = b.N;
    const int n_seg = b.n_seg;
    const size_t NI = N > 1 ? N - 1 : 1;
    const unsigned gN = cdiv(N, kBlock);
    hipLaunchKernelGGL(k_init_build, dim3(cdiv(std::max(n_seg * kRep, kMaxLevels + 1), kBlock)),
                       dim3(kBlock), 0, s, b.segs, n_seg, b.segstat_rep, b.seg_root, b.lvl, b.counters);
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(gN, 4096u)), dim3(kBlock), 0, s, b.lparent, (size_t)N,
                       kEmpty);
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(cdiv(NI, kBlock), 4096u)), dim3(kBlock), 0, s,
                       b.nparent, NI, kEmpty);
    hipLaunchKernelGGL(k_prep, dim3(gN), dim3(kBlock), 0, s, b.segs, n_seg, N, b.aabb_lo, b.aabb_hi,
                       b.segstat_rep, b.counters);
    hipLaunchKernelGGL(k_reduce_stats, dim3(cdiv((size_t)n_seg * kStat, kBlock)), dim3(kBlock), 0, s, n_seg,
                       b.segstat_rep, b.segstat);
    hipLaunchKernelGGL(k_morton, dim3(gN), dim3(kBlock), 0, s, N, b.aabb_lo, b.aabb_hi, b.segstat,
                       b.keys0, b.vals0);
    // radix sort over the bits in use: 30 Morton bits + mesh-id bits
    int seg_bits = 0;
    while ((1u << seg_bits) < (uint32_t)n_seg) seg_bits++;
    const int key_bits = 30 + seg_bits;
    uint64_t* kin = b.keys0;
    uint32_t* vin = b.vals0;
    if (radix_sort_pairs(b.keys0, b.vals0, b.keys1, b.vals1, N, key_bits, b.hist, s)) {
        kin = b.keys1;
        vin = b.vals1;
    }
    hipLaunchKernelGGL(k_pack, dim3(gN), dim3(kBlock), 0, s, N, kin, vin, b.segs, b.saabb, b.tris, b.leaf_order);
    if (N > 1) {
        const int n_int = (int)(N - 1);
        hipLaunchKernelGGL(k_karras, dim3(cdiv(N - 1, kBlock)), dim3(kBlock), 0, s, (int)N, kin, b.segs,
                           b.left, b.right, b.range, b.nparent, b.lparent, b.seg_root, b.is_root, b.dlev,
                           b.cnt);
        // bucket the internal nodes by prefix length, then one launch per length in use,
        // longest first; distinct keys give [64 - key_bits, 63], equal keys give
        // 64 + clz32(i ^ j) with i, j < N, i.e. [96 - ceil(log2 N), 95]
        int nbits = 0;
        while ((1u << nbits) < N) nbits++;
        const unsigned gI = cdiv(n_int, kBlock);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, s, b.lvl_hist, (size_t)kLevelBuckets, 0u);
        const unsigned gL = std::min(gI, kLevelBlocks);
        const uint32_t small_max = 2u * (uint32_t)b.leaf_size;  // subtrees no treelet restructures
        hipLaunchKernelGGL(k_level_hist, dim3(gL), dim3(kBlock), 0, s, n_int, b.dlev, b.cnt, small_max, b.lvl_hist);
        hipLaunchKernelGGL(k_level_scan, dim3(1), dim3(64), 0, s, b.lvl_hist, b.lvl_begin, b.lvl_fill);
        hipLaunchKernelGGL(k_level_scatter, dim3(gL), dim3(kBlock), 0, s, n_int, b.dlev, b.cnt, small_max, b.lvl_fill,
                           b.lvl_order);
        hipLaunchKernelGGL(k_boxes_local, dim3(gI), dim3(kBlock), 0, s, (uint32_t)n_int, b.range, b.left, b.right,
                           b.cnt, b.dlev, b.saabb, b.nbox, b.cost, b.leaf_size);
        const unsigned lgrid = std::min(gI, 1024u);
        for (int level = 95; level >= 64 - key_bits; level--) {
            if (level >= 64 && level < 96 - nbits) continue;
            hipLaunchKernelGGL(k_boxes_level, dim3(lgrid), dim3(kBlock), 0, s, level, b.lvl_begin, b.lvl_order,
                               b.range, b.left, b.right, b.cnt, b.saabb, b.nbox, b.cost, b.leaf_size);
        }
        if (treelet_size > 0) {
            int trc = run_treelets(b, treelet_size, treelet_rounds, key_bits, nbits, s);
            if (trc != GPURT_OK) return trc;
            b.repacked = true;
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
    ca.cnt = b.cnt;
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
    ca.repack = b.repacked ? 1 : 0;
    ca.tris_in = b.tris;
    ca.order_in = b.leaf_order;
    ca.tris_out = b.tris2;
    ca.order_out = b.leaf_order2;
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
                   int treelet_size, int treelet_rounds, int format, hipStream_t s, gpurt_bvh_s* out) {
    (void)ctx;
    // every buffer from the stream-ordered pool on the build stream: the products become
    // the BVH's (freed in stream order by gpurt_bvh_destroy), the scratch is freed on the
    // stream when b goes out of scope
    BuildBufs b;
    b.pooled = true;
    b.pool_stream = s;
    int rc;
    if ((rc = build_alloc(host_segs, n_seg, W, leaf_size, treelet_size, b))) return rc;
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
    if (h_counters[1] & 4u) return set_error(GPURT_E_INVALID, "non-finite vertex position");
    if (h_counters[1] & 2u) return set_error(GPURT_E_DEVICE, "wide-node capacity exceeded");
    int depth = 0;
    while (depth <= kMaxLevels && h_lvl[depth] > 0) depth++;
    // shrink the node array to the nodes in use
    const uint32_t n_nodes = h_counters[0];
    void* nodes = b.nodes;
    void* tight = nullptr;
    if (hipMallocAsync(&tight, node_bytes(W) * n_nodes, s) == hipSuccess) {
        if (hipMemcpyAsync(tight, nodes, node_bytes(W) * n_nodes, hipMemcpyDeviceToDevice, s) == hipSuccess) {
            nodes = tight;  // b frees the large array (after the copy, in stream order)
        } else {
            (void)hipFreeAsync(tight, s);
            (void)hipGetLastError();
            b.release(nodes);
        }
    } else {
        (void)hipGetLastError();
        b.release(nodes);
    }
    out->nodes = nodes;  // the caller frees the products on failure from here on
    uint32_t u = st[6];
    u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float scale;
    std::memcpy(&scale, &u, 4);
    // SAH cost relative to the root box area
    double sah = 0.0;
    {
        double* acc = reinterpret_cast<double*>(b.hist);  // scratch no longer needed
        GPURT_HIP(hipMemsetAsync(acc, 0, sizeof(double), s));
        if (W == 2) hipLaunchKernelGGL(k_sah_cost<2>, dim3(cdiv(n_nodes, kBlock)), dim3(kBlock), 0, s, n_nodes, (const WNode<2>*)nodes, acc);
        else if (W == 4) hipLaunchKernelGGL(k_sah_cost<4>, dim3(cdiv(n_nodes, kBlock)), dim3(kBlock), 0, s, n_nodes, (const WNode<4>*)nodes, acc);
        else hipLaunchKernelGGL(k_sah_cost<8>, dim3(cdiv(n_nodes, kBlock)), dim3(kBlock), 0, s, n_nodes, (const WNode<8>*)nodes, acc);
        float rl[4], rh[4];
        GPURT_HIP(hipMemcpyAsync(&sah, acc, sizeof(double), hipMemcpyDeviceToHost, s));
        GPURT_HIP(hipMemcpyAsync(rl, b.nbox, sizeof(rl), hipMemcpyDeviceToHost, s));
        GPURT_HIP(hipMemcpyAsync(rh, b.nbox + 1, sizeof(rh), hipMemcpyDeviceToHost, s));
        GPURT_HIP(hipStreamSynchronize(s));
        const float ex = rh[0] - rl[0], ey = rh[1] - rl[1], ez = rh[2] - rl[2];
        const double root_area = 2.0 * ((double)ex * ez + (double)ex * ey + (double)ey * ez);
        sah = (b.N > 1 && root_area > 0.0) ? sah / root_area : 0.0;
    }
    out->nodes = nodes;
    out->format = format;
    if (has_qnodes(format)) {
        GPURT_HIP(hipMallocAsync(&out->qnodes, qnode_bytes(W) * std::max<uint32_t>(n_nodes, 1u), s));
        launch_quantize(W, nodes, out->qnodes, n_nodes, n_nodes, nullptr, s);
        GPURT_HIP(hipGetLastError());
    }
    if (b.repacked && h_counters[2] != b.N) return set_error(GPURT_E_DEVICE, "leaf re-pack lost triangles");
    out->tris = (float4*)b.release(b.final_tris());
    out->leaf_order = (uint32_t*)b.release(b.final_order());
    out->n_nodes = n_nodes;
    out->node_capacity = n_nodes;
    out->depth = depth;
    out->width = W;
    out->leaf_size = leaf_size;
    out->n_tris = b.N;
    out->scale = scale;
    out->pad = scale * 0x1p-18f;
    out->sah_cost = (float)sah;
    out->level_cnt.assign(h_lvl.begin(), h_lvl.begin() + depth);
    return GPURT_OK;
}

}  // namespace gpurt_impl
