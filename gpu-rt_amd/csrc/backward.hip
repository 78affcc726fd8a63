// backward.hip -- gradients of unsigned distance and ray hit distance (plain path).
//
// Both kernels reuse the forward pass's saved primitive ids and barycentrics (no
// re-traversal) and scatter vertex gradients with hardware FP32 atomic adds
// (global_atomic_add_f32), so their cost is one gather of the triangle plus three
// 12-byte atomic read-modify-writes per query.  The per-query math is in backward.h.
#include "backward.h"

namespace gpurt_impl {

__global__ __launch_bounds__(kBlock) void k_cpq_bwd(uint32_t n, const float* __restrict__ q,
                                                    const int32_t* __restrict__ prim,
                                                    const float* __restrict__ g,
                                                    const uint32_t* __restrict__ tri_idx,
                                                    uint32_t n_tris, const float* __restrict__ verts,
                                                    float* __restrict__ g_q, float* g_verts) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    cpq_bwd_one(i, q, prim, g, tri_idx, n_tris, verts, 3, g_q, g_verts);
}

__global__ __launch_bounds__(kBlock) void k_ray_bwd(uint32_t n, const float4* __restrict__ rays,
                                                    const int32_t* __restrict__ prim,
                                                    const float* __restrict__ t,
                                                    const float* __restrict__ uv,
                                                    const float* __restrict__ g,
                                                    const uint32_t* __restrict__ tri_idx,
                                                    uint32_t n_tris,
                                                    const float* __restrict__ verts,
                                                    float* __restrict__ g_o,
                                                    float* __restrict__ g_d, float* g_verts) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    ray_bwd_one(i, rays, prim, t, uv, g, tri_idx, n_tris, verts, 3, g_o, g_d, g_verts);
}

int launch_cpq_bwd(gpurt_bvh_s* b, const float* q, size_t n, const int32_t* prim, const float* g,
                   float* g_q, float* g_verts, hipStream_t s) {
    if (n == 0) return GPURT_OK;
    hipLaunchKernelGGL(k_cpq_bwd, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, (uint32_t)n, q, prim, g,
                       b->tri_idx, (uint32_t)b->n_tris, b->verts, g_q, g_verts);
    GPURT_HIP(hipGetLastError());
    return GPURT_OK;
}

int launch_ray_bwd(gpurt_bvh_s* b, const gpurt_ray* rays, size_t n, const int32_t* prim,
                   const float* t, const float* uv, const float* g, float* g_o, float* g_d,
                   float* g_verts, hipStream_t s) {
    if (n == 0) return GPURT_OK;
    hipLaunchKernelGGL(k_ray_bwd, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, (uint32_t)n,
                       reinterpret_cast<const float4*>(rays), prim, t, uv, g, b->tri_idx,
                       (uint32_t)b->n_tris, b->verts, g_o, g_d, g_verts);
    GPURT_HIP(hipGetLastError());
    return GPURT_OK;
}

}  // namespace gpurt_impl
