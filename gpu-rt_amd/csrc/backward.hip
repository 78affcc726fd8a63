// backward.hip -- gradients of unsigned distance and ray hit distance (plain path).
//
// Both passes reuse the forward pass's saved primitive ids and barycentrics (no
// re-traversal; the per-query math is in backward.h).  Query / ray gradients are written
// directly.  Vertex gradients are sums over every query whose triangle uses the vertex;
// they are formed without atomics and the same bits come out of every run:
//   k_*_bwd_terms   per query: its gradient, and its three vertex terms with the vertex
//                   each belongs to, as (key = vertex, value = term index) pairs
//   radix sort      of the pairs by vertex (stable: a vertex's terms stay in query order)
//   k_vertex_runs   the first pair of each vertex's run sums the run in that order and
//                   adds it to the vertex gradient
// Algorithmic bytes: a closest-point query reads q, g, prim and its triangle (12 + 4 + 4
// + 12 + 36) and writes g_q (12) and its terms; a ray reads its direction, t, uv, g and
// its triangle.  The sort moves each term's (8 + 4)-byte pair once per 8 key bits.
#include "backward.h"

namespace gpurt_impl {

// term k = 3 * item + corner: key = its vertex (n_verts when there is none, sorted last),
// value = k, terms[3k .. 3k+2] = the float3 to add
__device__ __forceinline__ void write_terms(size_t i, const CornerTerms& t, uint32_t n_verts, uint64_t* __restrict__ keys,
                                            uint32_t* __restrict__ vals, float* __restrict__ terms) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const size_t k = 3 * i + j;
        keys[k] = t.v[j] == kNoVertex ? n_verts : t.v[j];
        vals[k] = (uint32_t)k;
        terms[3 * k] = t.c[j][0];
        terms[3 * k + 1] = t.c[j][1];
        terms[3 * k + 2] = t.c[j][2];
    }
}

__global__ __launch_bounds__(kBlock) void k_cpq_bwd_terms(uint32_t n, const float* __restrict__ q,
                                                          const int32_t* __restrict__ prim,
                                                          const float* __restrict__ g,
                                                          const uint32_t* __restrict__ tri_idx, uint32_t n_tris,
                                                          const float* __restrict__ verts, uint32_t n_verts,
                                                          float* __restrict__ g_q, uint64_t* __restrict__ keys,
                                                          uint32_t* __restrict__ vals, float* __restrict__ terms) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const CornerTerms t = cpq_bwd_terms(i, q, prim, g, tri_idx, n_tris, verts, 3, g_q);
    if (keys) write_terms(i, t, n_verts, keys, vals, terms);
}

__global__ __launch_bounds__(kBlock) void k_ray_bwd_terms(uint32_t n, const float4* __restrict__ rays,
                                                          const int32_t* __restrict__ prim,
                                                          const float* __restrict__ t,
                                                          const float* __restrict__ uv,
                                                          const float* __restrict__ g,
                                                          const uint32_t* __restrict__ tri_idx, uint32_t n_tris,
                                                          const float* __restrict__ verts, uint32_t n_verts,
                                                          float* __restrict__ g_o, float* __restrict__ g_d,
                                                          uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                          float* __restrict__ terms) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const CornerTerms ct = ray_bwd_terms(i, rays, prim, t, uv, g, tri_idx, n_tris, verts, 3, g_o, g_d);
    if (keys) write_terms(i, ct, n_verts, keys, vals, terms);
}

// One thread per sorted pair; the first pair of a vertex's run adds the run, in order.
__global__ __launch_bounds__(kBlock) void k_vertex_runs(uint32_t m, const uint64_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ vals,
                                                        const float* __restrict__ terms, uint32_t n_verts,
                                                        float* __restrict__ g_verts) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= m) return;
    const uint64_t v = keys[p];
    if (v >= n_verts || (p > 0 && keys[p - 1] == v)) return;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (uint32_t r = p; r < m && keys[r] == v; r++) {
        const size_t k = vals[r];
        sx += terms[3 * k];
        sy += terms[3 * k + 1];
        sz += terms[3 * k + 2];
    }
    g_verts[3 * v] += sx;
    g_verts[3 * v + 1] += sy;
    g_verts[3 * v + 2] += sz;
}

namespace {

// scratch of the vertex-gradient sum over n items (3 terms each)
struct TermScratch {
    void* mem = nullptr;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    uint32_t *v0 = nullptr, *v1 = nullptr, *sort = nullptr;
    float* terms = nullptr;
};

int alloc_terms(size_t n, hipStream_t s, TermScratch& t) {
    if (3 * n > 0xFFFFFFFFull - kBlock) return set_error(GPURT_E_INVALID, "too many queries for one backward call");
    const size_t m = 3 * n;
    const size_t words = radix_sort_scratch_words((uint32_t)m);
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t bytes = 2 * up(8 * m) + 2 * up(4 * m) + up(12 * m) + up(4 * words) + 256;
    if (hipMallocAsync(&t.mem, bytes, s) != hipSuccess) {
        (void)hipGetLastError();
        t.mem = nullptr;
        return set_error(GPURT_E_OOM, "backward scratch allocation failed");
    }
    char* p = static_cast<char*>(t.mem);
    t.k0 = reinterpret_cast<uint64_t*>(p);
    p += up(8 * m);
    t.k1 = reinterpret_cast<uint64_t*>(p);
    p += up(8 * m);
    t.v0 = reinterpret_cast<uint32_t*>(p);
    p += up(4 * m);
    t.v1 = reinterpret_cast<uint32_t*>(p);
    p += up(4 * m);
    t.terms = reinterpret_cast<float*>(p);
    p += up(12 * m);
    t.sort = reinterpret_cast<uint32_t*>(p);
    return GPURT_OK;
}

int bits_for(uint32_t n_verts) {
    int b = 1;
    while (b < 32 && (n_verts >> b) != 0) b++;
    return b;  // keys 0 .. n_verts (n_verts = no vertex) fit in b bits
}

// sort the terms by vertex and add each vertex's run to g_verts
void sum_terms(TermScratch& t, size_t n, uint32_t n_verts, float* g_verts, hipStream_t s) {
    const uint32_t m = (uint32_t)(3 * n);
    uint64_t* k = t.k0;
    uint32_t* v = t.v0;
    if (radix_sort_pairs(t.k0, t.v0, t.k1, t.v1, m, bits_for(n_verts), t.sort, s)) {
        k = t.k1;
        v = t.v1;
    }
    hipLaunchKernelGGL(k_vertex_runs, dim3(cdiv(m, kBlock)), dim3(kBlock), 0, s, m, k, v, t.terms, n_verts, g_verts);
}

}  // namespace

int launch_cpq_bwd(gpurt_bvh_s* b, const float* q, size_t n, const int32_t* prim, const float* g,
                   float* g_q, float* g_verts, hipStream_t s) {
    if (n == 0) return GPURT_OK;
    TermScratch t;
    if (g_verts) {
        int rc = alloc_terms(n, s, t);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_cpq_bwd_terms, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, (uint32_t)n, q, prim, g,
                       b->tri_idx, (uint32_t)b->n_tris, b->verts, (uint32_t)b->n_verts, g_q, t.k0, t.v0, t.terms);
    if (g_verts) sum_terms(t, n, (uint32_t)b->n_verts, g_verts, s);
    const hipError_t e = hipGetLastError();
    if (t.mem) (void)hipFreeAsync(t.mem, s);
    if (e != hipSuccess) return hip_error(e, "closest-point backward launch");
    return GPURT_OK;
}

int launch_ray_bwd(gpurt_bvh_s* b, const gpurt_ray* rays, size_t n, const int32_t* prim,
                   const float* t, const float* uv, const float* g, float* g_o, float* g_d,
                   float* g_verts, hipStream_t s) {
    if (n == 0) return GPURT_OK;
    TermScratch ts;
    const bool vgrad = g_verts && uv;
    if (vgrad) {
        int rc = alloc_terms(n, s, ts);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_ray_bwd_terms, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, (uint32_t)n,
                       reinterpret_cast<const float4*>(rays), prim, t, uv, g, b->tri_idx, (uint32_t)b->n_tris,
                       b->verts, (uint32_t)b->n_verts, g_o, g_d, ts.k0, ts.v0, ts.terms);
    if (vgrad) sum_terms(ts, n, (uint32_t)b->n_verts, g_verts, s);
    const hipError_t e = hipGetLastError();
    if (ts.mem) (void)hipFreeAsync(ts.mem, s);
    if (e != hipSuccess) return hip_error(e, "ray backward launch");
    return GPURT_OK;
}

}  // namespace gpurt_impl
