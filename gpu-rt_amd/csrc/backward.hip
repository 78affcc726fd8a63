// backward.hip -- gradients of unsigned distance and ray hit distance (plain path).
//
// Both passes reuse the forward pass's saved primitive ids and barycentrics (no
// re-traversal; the per-query math is in backward.h).  Query / ray gradients are written
// directly.  Vertex gradients are sums over every query whose triangle uses the vertex:
//   atomic       k_cpq_bwd / k_ray_bwd add each query's three vertex terms with FP32
//                atomics: one launch, bound by the atomics (~9 random device-scope FP32
//                atomics per query; the summation order, and so the last bits, vary)
//   fixed order  k_*_bwd_terms write the terms with the vertex they belong to, a stable
//                radix sort orders them by vertex, and k_wave_segsum + k_run_sum add each
//                vertex's terms in a fixed order: the same bits on every run
// The context's mode picks one (gpurt_ctx_set_vertex_sums); by default the fixed order
// runs for calls of at least kFixedOrderMinTerms terms, where it is also the faster.
// Algorithmic bytes: a closest-point query reads q, g, prim and its triangle (12 + 4 + 4
// + 12 + 36) and writes g_q (12) and adds to three vertices (3 x 2 x 12); a ray reads its
// direction, t, uv, g and its triangle.
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
    write_terms(i, cpq_bwd_terms(i, q, prim, g, tri_idx, n_tris, verts, 3, g_q), n_verts, keys, vals, terms);
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
    write_terms(i, ray_bwd_terms(i, rays, prim, t, uv, g, tri_idx, n_tris, verts, 3, g_o, g_d), n_verts, keys, vals,
                terms);
}

__global__ __launch_bounds__(kBlock) void k_cpq_bwd(uint32_t n, const float* __restrict__ q,
                                                    const int32_t* __restrict__ prim, const float* __restrict__ g,
                                                    const uint32_t* __restrict__ tri_idx, uint32_t n_tris,
                                                    const float* __restrict__ verts, float* __restrict__ g_q,
                                                    float* g_verts) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    scatter_terms(cpq_bwd_terms(i, q, prim, g, tri_idx, n_tris, verts, 3, g_q), g_verts);
}

__global__ __launch_bounds__(kBlock) void k_ray_bwd(uint32_t n, const float4* __restrict__ rays,
                                                    const int32_t* __restrict__ prim, const float* __restrict__ t,
                                                    const float* __restrict__ uv, const float* __restrict__ g,
                                                    const uint32_t* __restrict__ tri_idx, uint32_t n_tris,
                                                    const float* __restrict__ verts, float* __restrict__ g_o,
                                                    float* __restrict__ g_d, float* g_verts) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    scatter_terms(ray_bwd_terms(i, rays, prim, t, uv, g, tri_idx, n_tris, verts, 3, g_o, g_d), uv ? g_verts : nullptr);
}

// adds a vertex's total to its gradient (each vertex has one run)
struct VertexSink {
    float* g_verts;
    __device__ void operator()(uint64_t v, float x, float y, float z) const {
        g_verts[3 * v] += x;
        g_verts[3 * v + 1] += y;
        g_verts[3 * v + 2] += z;
    }
};

namespace {

// scratch of the vertex-gradient sum over n items (3 terms each)
struct TermScratch {
    void* mem = nullptr;
    CallScratch cs;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    uint32_t *v0 = nullptr, *v1 = nullptr, *sort = nullptr;
    float *terms = nullptr, *partial = nullptr;
};

int alloc_terms(gpurt_ctx_s* c, size_t n, hipStream_t s, TermScratch& t) {
    if (3 * n > 0xFFFFFFFFull - kBlock) return set_error(GPURT_E_INVALID, "too many queries for one backward call");
    const size_t m = 3 * n;
    const size_t words = radix_sort_scratch_words((uint32_t)m);
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t bytes = 2 * up(8 * m) + 2 * up(4 * m) + 2 * up(12 * m) + up(4 * words) + 256;
    const int rc = ctx_scratch(c, s, bytes, t.cs);
    if (rc) return rc;
    t.mem = t.cs.mem;
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
    t.partial = reinterpret_cast<float*>(p);
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
    hipLaunchKernelGGL(k_wave_segsum<kBlock>, dim3(cdiv(m, kBlock)), dim3(kBlock), 0, s, m, k, v, t.terms, t.partial);
    hipLaunchKernelGGL((k_run_sum<kBlock, VertexSink>), dim3(cdiv(m, kBlock)), dim3(kBlock), 0, s, m, k, t.partial,
                       (uint64_t)n_verts, VertexSink{g_verts});
}

}  // namespace

int launch_cpq_bwd(gpurt_bvh_s* b, const float* q, size_t n, const int32_t* prim, const float* g,
                   float* g_q, float* g_verts, hipStream_t s) {
    if (n == 0) return GPURT_OK;
    if (!g_verts || !fixed_order_sums(b->ctx, 3 * n)) {
        hipLaunchKernelGGL(k_cpq_bwd, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, (uint32_t)n, q, prim, g,
                           b->tri_idx, (uint32_t)b->n_tris, b->verts, g_q, g_verts);
        GPURT_HIP(hipGetLastError());
        return GPURT_OK;
    }
    TermScratch t;
    int rc = alloc_terms(b->ctx, n, s, t);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cpq_bwd_terms, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, (uint32_t)n, q, prim, g,
                       b->tri_idx, (uint32_t)b->n_tris, b->verts, (uint32_t)b->n_verts, g_q, t.k0, t.v0, t.terms);
    sum_terms(t, n, (uint32_t)b->n_verts, g_verts, s);
    const hipError_t e = hipGetLastError();
    ctx_scratch_done(b->ctx, t.cs);
    if (e != hipSuccess) return hip_error(e, "closest-point backward launch");
    return GPURT_OK;
}

int launch_ray_bwd(gpurt_bvh_s* b, const gpurt_ray* rays, size_t n, const int32_t* prim,
                   const float* t, const float* uv, const float* g, float* g_o, float* g_d,
                   float* g_verts, hipStream_t s) {
    if (n == 0) return GPURT_OK;
    const bool vgrad = g_verts && uv;
    if (!vgrad || !fixed_order_sums(b->ctx, 3 * n)) {
        hipLaunchKernelGGL(k_ray_bwd, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, (uint32_t)n,
                           reinterpret_cast<const float4*>(rays), prim, t, uv, g, b->tri_idx, (uint32_t)b->n_tris,
                           b->verts, g_o, g_d, g_verts);
        GPURT_HIP(hipGetLastError());
        return GPURT_OK;
    }
    TermScratch ts;
    int rc = alloc_terms(b->ctx, n, s, ts);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ray_bwd_terms, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, (uint32_t)n,
                       reinterpret_cast<const float4*>(rays), prim, t, uv, g, b->tri_idx, (uint32_t)b->n_tris,
                       b->verts, (uint32_t)b->n_verts, g_o, g_d, ts.k0, ts.v0, ts.terms);
    sum_terms(ts, n, (uint32_t)b->n_verts, g_verts, s);
    const hipError_t e = hipGetLastError();
    ctx_scratch_done(b->ctx, ts.cs);
    if (e != hipSuccess) return hip_error(e, "ray backward launch");
    return GPURT_OK;
}

}  // namespace gpurt_impl
