// gpurt_internal.h -- shared types of the gfx950 BVH builder and traversal kernels.
//
// Device layout (all 128-byte aligned allocations from hipMalloc):
//   wide nodes   WNode<W>: per child slot j: box lo/hi as six W-float arrays (SoA inside
//                the node, so one node is 6*W floats of boxes + W child words + W aux
//                words = 32*W bytes: 64 B for W=2, 128 B (one L2 line) for W=4, 256 B
//                for W=8).  Boxes are inflated by `pad` so culls stay conservative.
//   quantized    QNode<W>: the same tree with 8-bit child boxes relative to the node's
//   nodes        own box (origin + per-axis power-of-two scale), 16 + 4*W + 6*W bytes
//                rounded up to 16: 48 B for W=2, 64 B for W=4, 96 B for W=8.  Each
//                decoded box contains the FP32 box it was made from (checked in FP32
//                with the decoder's own arithmetic), so culls stay conservative.
//   triangles    3 float4 per triangle in leaf order: (a.xyz, id), (b.xyz, 0),
//                (c.xyz, 0) with id = the triangle's index in its mesh (float bits).
//   child word   internal child: wide-node index (< 2^31);
//                leaf: bit31 | (count-1) << 28 | first packed triangle (count 1..8);
//                empty slot: 0xFFFFFFFF.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "../../include/gpurt.h"

namespace gpurt_impl {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kMaxLeafFirst = 0x0FFFFFFFu;  // 2^28 - 1 packed triangles max
// Wide-node levels the collapse can emit.  A mesh's binary radix tree is at most 58
// levels deep (30 Morton bits + 28 index tie-break bits) and every wide level consumes
// at least one binary level, so 64 launches always finish the tree.
constexpr int kMaxLevels = 64;
constexpr int kBlock = 256;
// buckets of the per-prefix-length build passes: two per length (0..255)
constexpr uint32_t kLevelBuckets = 512;

__host__ __device__ inline uint32_t make_leaf(uint32_t first, uint32_t count) {
    return kLeafBit | ((count - 1u) << 28) | first;
}
__host__ __device__ inline bool child_is_leaf(uint32_t c) { return (c & kLeafBit) != 0u; }
__host__ __device__ inline uint32_t leaf_first(uint32_t c) { return c & kMaxLeafFirst; }
__host__ __device__ inline uint32_t leaf_count(uint32_t c) { return ((c >> 28) & 7u) + 1u; }

template <int W>
struct alignas(32 * W) WNode {
    static constexpr int kWidth = W;
    float lox[W], loy[W], loz[W];
    float hix[W], hiy[W], hiz[W];
    uint32_t child[W];
    uint32_t aux[W];  // triangles under the child (for cost estimates)
};
static_assert(sizeof(WNode<2>) == 64, "node size");
static_assert(sizeof(WNode<4>) == 128, "node size");
static_assert(sizeof(WNode<8>) == 256, "node size");

inline size_t node_bytes(int W) { return (size_t)32 * (size_t)W; }

// Quantized wide node, as 32-bit words.  Child j's box on axis a decodes as
//   lo = fmaf((float)qlo[a][j], s_a, origin_a),  hi = fmaf((float)qhi[a][j], s_a, origin_a)
// with s_a a power of two.  2- and 4-wide nodes store the three scales as floats (words
// 3..5: the traversal multiplies them in directly); 8-wide nodes, whose 96 bytes have no
// room for them, store the biased exponents e_a (s_a = 2^(e_a - 127), byte a of word 3).
//   words 0..2         origin x, y, z (float)
//   word 3 (.. 5)      scales (float) or exponent bytes
//   kChild0 ..         child[W]
//   kQ0 ..             q[6][W] bytes: qlo x, y, z, then qhi x, y, z; zero padding
// An empty child slot has qlo = 255 and qhi = 0 on every axis (an inverted box).
template <int W>
struct alignas(16) QNode {
    static constexpr int kWidth = W;
    static constexpr bool kFloatScale = W <= 4;
    static constexpr int kChild0 = kFloatScale ? 6 : 4;
    static constexpr int kQ0 = kChild0 + W;
    static constexpr int kWords = (kQ0 + (6 * W + 3) / 4 + 3) / 4 * 4;
    uint32_t w[kWords];
};
static_assert(sizeof(QNode<2>) == 48, "node size");
static_assert(sizeof(QNode<4>) == 64, "node size");
static_assert(sizeof(QNode<8>) == 96, "node size");

inline size_t qnode_bytes(int W) { return W == 2 ? 48 : (W == 4 ? 64 : 96); }

// Node formats of gpurt_build_opts::node_format (0 selects the library default, mixed).
constexpr int kFormatF32 = 1;    // FP32 boxes for every query
constexpr int kFormatQ8 = 2;     // 8-bit boxes for every query
constexpr int kFormatMixed = 3;  // 8-bit boxes for rays, FP32 boxes for closest points
inline bool has_qnodes(int f) { return f != kFormatF32; }
inline bool rays_use_q8(int f) { return f == kFormatQ8 || f == kFormatMixed; }
inline bool cpq_use_q8(int f) { return f == kFormatQ8; }

// Atomics on global memory through a global-address-space pointer: no flat-address
// (LDS / scratch aperture) checks are generated around them.
typedef __attribute__((address_space(1))) uint32_t global_u32;
__device__ __forceinline__ uint32_t gatomic_add(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add((global_u32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t gatomic_or(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_or((global_u32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t gatomic_max_u32(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_max((global_u32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One mesh of a (possibly single-mesh) build.
struct SegDesc {
    const float* verts;
    uint64_t vstride;  // in floats
    uint32_t n_verts;
    uint32_t n_tris;
    const uint32_t* idx;
    uint32_t tri_offset;  // first triangle of this mesh in the concatenated order
    uint32_t pad_;
};

// Build products that the traversal kernels read.
struct DeviceBvh {
    const void* nodes;   // WNode<W>[]
    const float4* tris;  // 3 per triangle
    uint32_t root;       // wide-node index of the root
    uint32_t n_tris;
};

}  // namespace gpurt_impl

struct gpurt_ctx_s {
    int device = 0;
    int n_cu = 256;
    int lds_per_cu = 160 * 1024;
    uint32_t* err = nullptr;  // sticky device error bits (bit 0: traversal stack overflow)
    // GPURT_SUMS_* mode of the backward passes' vertex-gradient sums
    int vertex_sums = 0;
    // stream on which destroyed handles release their memory, after waiting for the
    // recorded uses of the handle (stream-ordered, no host wait)
    hipStream_t release = nullptr;
    // BVHs and TLASes still open on the context: gpurt_ctx_destroy releases their device
    // memory, and destroying such a handle afterwards frees only its host part
    std::mutex live_mu;
    std::set<gpurt_bvh_s*> bvhs;
    std::set<gpurt_tlas_s*> tlases;
    // device scratch of the query and backward calls, one buffer per stream (ctx_scratch)
    struct StreamBuf {
        hipStream_t s = nullptr;
        void* mem = nullptr;
        size_t bytes = 0;
        hipEvent_t last = nullptr;  // recorded after the buffer's last use
        bool busy = false;          // a call is between ctx_scratch and ctx_scratch_done
    };
    std::mutex scratch_mu;
    std::vector<StreamBuf> scratch;
};

namespace gpurt_impl {
// Vertex-gradient sums of the backward passes: the fixed-order sum (sort by vertex, then
// segmented sums) when the context asks for it, FP32 atomics when it asks for them, and
// otherwise whichever is faster -- the fixed order from about 2 M vertex terms, where
// the atomics' device-wide rate (~20 G/s) costs more than the sort's passes (1 M
// closest-point queries = 3.1 M terms: 0.38 vs 0.48 ms; a 256-mesh batch of 19 M queries
// and rays: forward + backward 32.3 vs 36.1 ms; 262 K rays = 0.8 M terms: 0.17 vs 0.13
// ms).
constexpr size_t kFixedOrderMinTerms = size_t(1) << 21;
inline bool fixed_order_sums(const gpurt_ctx_s* c, size_t terms) {
    return c->vertex_sums == GPURT_SUMS_FIXED_ORDER ||
           (c->vertex_sums == GPURT_SUMS_AUTO && terms >= kFixedOrderMinTerms);
}
}  // namespace gpurt_impl

struct gpurt_bvh_s {
    gpurt_ctx ctx = nullptr;
    int width = 4;
    int leaf_size = 4;
    int depth = 0;
    int64_t n_tris = 0;
    int64_t n_verts = 0;
    int64_t n_nodes = 0;
    int64_t node_capacity = 0;
    float scale = 0.f;
    float pad = 0.f;
    float sah_cost = 0.f;
    void* nodes = nullptr;        // WNode<W>[node_capacity] (boxes in FP32; refit updates these)
    void* qnodes = nullptr;       // QNode<W>[n_nodes] when has_qnodes(format)
    int format = 1;
    float4* tris = nullptr;       // 3 * n_tris
    uint32_t* tri_idx = nullptr;  // 3 * n_tris, original order (for backward / refit)
    float* verts = nullptr;       // 3 * n_verts compact positions, original order
    uint32_t* leaf_order = nullptr;  // packed position -> original triangle (refit)
    std::vector<uint32_t> level_cnt;  // wide nodes per level (level l starts at sum of previous)
    // Every buffer above comes from the stream-ordered pool (hipMallocAsync on the build
    // stream), so gpurt_bvh_destroy frees it with hipFreeAsync on ctx->release after
    // waiting for the last use recorded on each stream that used the handle.
    std::mutex use_mu;
    std::vector<std::pair<hipStream_t, hipEvent_t>> uses;
    bool uses_lost = false;  // an event could not be recorded: destroy waits for the device
};

namespace gpurt_impl {

// error plumbing (abi.hip)
int set_error(int code, const std::string& msg);
int hip_error(hipError_t e, const char* what);

// use tracking of a BVH handle (abi.hip): note_use records, after work that reads `b`
// was enqueued on `s`, an event there (one event per stream, re-recorded); wait_uses
// makes stream `s` wait for every recorded use
void note_use(gpurt_bvh_s* b, hipStream_t s);
void wait_uses(gpurt_bvh_s* b, hipStream_t s);

#define GPURT_HIP(call)                                            \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return ::gpurt_impl::hip_error(e_, #call); \
    } while (0)

// validated build options (abi.hip); zero fields of the C struct select defaults
struct BuildParams {
    int width = 4, leaf_size = 4, treelet_size = 0, treelet_rounds = 0, format = 0;
};
int parse_build_opts(const gpurt_build_opts* o, BuildParams& p);

// builder (build.hip): device buffers of one build over one or more meshes
struct BuildBufs {
    uint32_t N = 0;  // triangles over all meshes
    int n_seg = 0;   // meshes
    int width = 4;
    int leaf_size = 4;
    bool repacked = false;  // treelets changed the tree: leaves re-packed into tris2
    uint32_t nblocks_sort = 0;
    uint32_t capacity = 0;  // wide nodes the node array can hold
    SegDesc* segs = nullptr;
    float4 *aabb_lo = nullptr, *aabb_hi = nullptr;
    uint32_t* segstat_rep = nullptr;  // per-mesh statistics, replicated for atomics
    uint32_t* segstat = nullptr;      // per-mesh statistics, reduced
    uint64_t *keys0 = nullptr, *keys1 = nullptr;
    uint32_t *vals0 = nullptr, *vals1 = nullptr, *hist = nullptr;
    float4* saabb = nullptr;  // sorted triangle boxes (lo, hi)
    uint32_t *left = nullptr, *right = nullptr;
    uint2* range = nullptr;
    uint32_t *nparent = nullptr, *lparent = nullptr, *seg_root = nullptr;
    uint8_t* is_root = nullptr;
    uint8_t* dlev = nullptr;  // common-prefix length (Karras delta) of each internal node
    uint32_t* cnt = nullptr;  // triangles under each internal node
    float* cost = nullptr;    // SAH cost of each internal node's subtree
    // internal nodes bucketed by prefix length l and size: bucket 2l holds those with
    // more than 2 leaf_size triangles (the treelet roots), 2l + 1 the others
    uint32_t* lvl_order = nullptr;
    uint32_t* lvl_hist = nullptr;   // kLevelBuckets counts
    uint32_t* lvl_begin = nullptr;  // bucket k = lvl_order[lvl_begin[k] .. lvl_begin[k+1])
    uint32_t* lvl_fill = nullptr;
    float4* nbox = nullptr;   // binary node boxes (lo, hi)
    uint32_t *wsrc = nullptr, *wseg = nullptr, *lvl = nullptr, *counters = nullptr;
    void* nodes = nullptr;
    float4* tris = nullptr;          // triangles in sorted (Morton) order
    uint32_t* leaf_order = nullptr;  // sorted position -> triangle
    float4* tris2 = nullptr;         // re-packed leaves (treelet builds only)
    uint32_t* leaf_order2 = nullptr;
    float4* final_tris() const { return repacked ? tris2 : tris; }
    uint32_t* final_order() const { return repacked ? leaf_order2 : leaf_order; }
    std::vector<void*> allocs;  // freed with the object unless release()d
    // stream-ordered allocations (hipMallocAsync / hipFreeAsync on pool_stream) when set,
    // else hipMalloc / hipFree (the batch keeps its buffers for its lifetime)
    hipStream_t pool_stream = nullptr;
    bool pooled = false;
    hipError_t alloc(void** p, size_t bytes) {
        *p = nullptr;
        const hipError_t e = pooled ? hipMallocAsync(p, bytes, pool_stream) : hipMalloc(p, bytes);
        if (e == hipSuccess) allocs.push_back(*p);
        return e;
    }
    BuildBufs() = default;
    BuildBufs(const BuildBufs&) = delete;
    BuildBufs& operator=(const BuildBufs&) = delete;
    ~BuildBufs();
    void* release(void* p);  // hands ownership of one allocation to the caller
};
int build_alloc(const SegDesc* host_segs, int n_seg, int width, int leaf_size, int treelet_size,
                BuildBufs& b);
int build_enqueue_binary(BuildBufs& b, int treelet_size, int treelet_rounds, hipStream_t s);
int build_enqueue_levels(BuildBufs& b, int leaf_size, int level_begin, int level_end, hipStream_t s);
int run_treelets(BuildBufs& b, int treelet_size, int rounds, int key_bits, int nbits, hipStream_t s);
// Stable LSD radix sort (build.hip) of n (uint64 key, uint32 value) pairs on the low
// key_bits bits, 8 bits per pass, ping-ponging between (k0, v0) and (k1, v1); returns
// true when the sorted pairs end in (k1, v1).  scratch: radix_sort_scratch_words(n)
// words.
bool radix_sort_pairs(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint32_t n, int key_bits,
                      uint32_t* scratch, hipStream_t s);
size_t radix_sort_scratch_words(uint32_t n);
int build_segments(gpurt_ctx ctx, const SegDesc* host_segs, int n_seg, int width, int leaf_size,
                   int treelet_size, int treelet_rounds, int format, hipStream_t stream, gpurt_bvh_s* out);

// quantized node copies (quantize.hip): nodes [0, n) with n = *n_dev when n_dev is given
// (then `capacity` threads are launched), else n_host
void launch_quantize(int width, const void* nodes, void* qnodes, uint32_t n_host, uint32_t capacity,
                     const uint32_t* n_dev, hipStream_t s);

// traversal launch planning (query.hip): stack depth, block size, grid, spill slots and
// leaf trigger for one launch of a query kind, and the per-call scratch (queue counters
// cleared on the stream + stack spill buffer)
enum class QKind { Cpq, Closest, Any };
struct TravPlan {
    int S = 8, BS = 256;
    unsigned grid = 1;
    uint32_t total = 256, spill_slots = 0, leaf_trigger = 32;
};
// Device scratch of one query or backward call.  Outside graph capture it is the
// context's buffer for the call's stream: grown when a call needs more (hipMalloc, after
// a device wait to free the old one), reused by the next call on that stream without an
// allocation, and released with the context; the stream first waits for the buffer's last recorded use,
// so a stream handle that the runtime recycles for another stream cannot overlap it.
// While the stream is being captured into a graph, or while another call on the same
// stream holds the buffer (two host threads), the call gets an allocation of its own,
// freed in stream order when it is done.  (A stream-ordered allocation per call made
// the host wait for the device on every large call: ~230 us for 1 M rays.)
struct CallScratch {
    void* mem = nullptr;
    hipStream_t s = nullptr;
    int slot = -1;  // the context's buffer index, or -1 for a call-owned allocation
};
int ctx_scratch(gpurt_ctx_s* c, hipStream_t s, size_t bytes, CallScratch& out);
void ctx_scratch_done(gpurt_ctx_s* c, CallScratch& cs);
void ctx_scratch_release(gpurt_ctx_s* c);  // after the device is idle (context destroy)

struct TravScratch {
    void* mem = nullptr;
    uint32_t* counters = nullptr;
    uint2* spill = nullptr;
    CallScratch cs;
};
// LDS per lane beyond the stack slots of the single-level kernels: the 64-bit result key
// and the member count of work donation (traverse.h)
constexpr int kGroupKeyBytes = 12;
// waves_per_simd: what the kernel is compiled for (0: the single-level kernels' WavesHint);
// lds_extra: LDS bytes per lane beyond the stack slots
int plan_traversal(gpurt_ctx ctx, int width, bool q8, QKind kind, int stack_need, const gpurt_query_opts* opts,
                   size_t n, TravPlan& p, int waves_per_simd = 0, int lds_extra = kGroupKeyBytes);
int alloc_traversal_scratch(gpurt_ctx_s* c, const TravPlan& p, hipStream_t s, TravScratch& t);
void free_traversal_scratch(gpurt_ctx_s* c, TravScratch& t);
// the same scratch in memory the caller holds
size_t traversal_scratch_bytes(const TravPlan& p);
void carve_traversal_scratch(const TravPlan& p, void* mem, TravScratch& t);
void clear_traversal_counters(const TravScratch& t, hipStream_t s);

// traversal launchers (query.hip)
int launch_cpq(gpurt_bvh_s* b, const float* q, size_t n, int32_t* prim, float* pt, float* dist,
               float* bary, const gpurt_query_opts* opts, hipStream_t s);
int launch_ray(gpurt_bvh_s* b, const gpurt_ray* rays, size_t n, int32_t* prim, float* t,
               float* uv, uint8_t* occ, bool any, const gpurt_query_opts* opts, hipStream_t s);

// gpurt_trace_closest_noalloc (query.hip): the closest-hit launch with its traversal
// scratch in a fixed per-device array of kNoAllocScratchBytes, set up once per device by
// noalloc_scratch_init (with that device current; called by gpurt_ctx_create)
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
int noalloc_scratch_init(int device);
int launch_ray_noalloc(gpurt_bvh_s* b, const gpurt_ray* rays, size_t n, int32_t* prim, float* t, float* uv,
                       const gpurt_query_opts* opts, hipStream_t s);

// backward (backward.hip)
int launch_cpq_bwd(gpurt_bvh_s* b, const float* q, size_t n, const int32_t* prim, const float* g,
                   float* g_q, float* g_verts, hipStream_t s);
int launch_ray_bwd(gpurt_bvh_s* b, const gpurt_ray* rays, size_t n, const int32_t* prim,
                   const float* t, const float* uv, const float* g, float* g_o, float* g_d,
                   float* g_verts, hipStream_t s);

int launch_refit(gpurt_bvh_s* b, const float* verts, size_t vstride_floats, hipStream_t s);

// host-side facts of a TLAS (tlas.hip): its context, and per instance the custom index
// and the BLAS
int tlas_host_info(gpurt_tlas t, gpurt_ctx* ctx, const std::vector<uint32_t>** custom,
                   const std::vector<gpurt_bvh_s*>** blas);
// TLAS traversal for wavefront callers (tlas.hip): the ray count may be read from device
// memory at launch (n_dev, at most n) and the traversal scratch may be the caller's
// (scratch of >= tlas_scratch_bytes(T, n) bytes; nullptr: allocated per call)
int tlas_trace_ex(gpurt_tlas T, const gpurt_ray* rays, size_t n, const uint32_t* n_dev, int32_t* obj,
                  int32_t* prim, float* t, float* uv, uint8_t* occ, bool any, const gpurt_query_opts* opts,
                  hipStream_t s, void* scratch, size_t scratch_bytes);
size_t tlas_scratch_bytes(gpurt_tlas T, size_t n);
// records a use of the TLAS (and through it of its BLASes) on stream s (tlas.hip)
void tlas_note_use(gpurt_tlas t, hipStream_t s);
// gpurt_ctx_destroy's part for a TLAS still open: frees its device memory on the
// context's release stream and detaches it from the context (tlas.hip)
void tlas_orphan(gpurt_tlas t);

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace gpurt_impl
