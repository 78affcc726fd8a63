// refit.hip -- BVH refit for moved vertices (same triangles, same tree).
//
// The reference only rebuilds (VK_BUILD_ACCELERATION_STRUCTURE_MODE_BUILD_KHR,
// src/vk/vulkan.cpp:832,915); an optimisation loop that moves vertices every step wants
// the cheaper update.  Steps:
//   k_refit_check   every new position finite, else GPURT_E_INVALID with the tree unchanged
//   k_refit_tris    re-gather each packed triangle from the new vertex positions (its id
//                   and leaf position are unchanged), compact copy of the positions for
//                   the backward passes, and the new coordinate scale
//   k_refit_level   one launch per wide level, deepest first: a child box is the union
//                   of a leaf's triangles (inflated by the new pad) or of the child node's
//                   own (already refitted, already inflated) child boxes
// Wide nodes of one level occupy a contiguous index range (the collapse allocates level
// l+1 while processing level l), so a level is a plain index range.
#include <algorithm>

#include "gpurt_internal.h"

namespace gpurt_impl {

__device__ __forceinline__ uint32_t f2o_r(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f_r(uint32_t u) {
    u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return __uint_as_float(u);
}

__global__ void k_refit_init(uint32_t* scale) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        scale[0] = f2o_r(0.0f);
        scale[1] = 0u;  // set when a position is not finite
    }
}

__global__ __launch_bounds__(kBlock) void k_refit_check(uint32_t n_verts, const float* __restrict__ v,
                                                        uint64_t stride, uint32_t* flag) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    bool bad = false;
    if (i < n_verts) {
        const float* p = v + (size_t)i * stride;
        bad = !(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]));
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) gatomic_or(flag, 1u);
}

__global__ __launch_bounds__(kBlock) void k_refit_verts(uint32_t n_verts, const float* __restrict__ v,
                                                        uint64_t stride, float* __restrict__ compact,
                                                        uint32_t* scale) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    float amax = 0.0f;
    if (i < n_verts) {
        const float* p = v + (size_t)i * stride;
        const float x = p[0], y = p[1], z = p[2];
        compact[3 * (size_t)i] = x;
        compact[3 * (size_t)i + 1] = y;
        compact[3 * (size_t)i + 2] = z;
        amax = fmaxf(fmaxf(fabsf(x), fabsf(y)), fabsf(z));
    }
    for (int o = 32; o > 0; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o));
    if ((threadIdx.x & 63) == 0 && amax > 0.0f) gatomic_max_u32(scale, f2o_r(amax));
}

__global__ __launch_bounds__(kBlock) void k_refit_tris(uint32_t n_tris, const uint32_t* __restrict__ leaf_order,
                                                       const uint32_t* __restrict__ tri_idx,
                                                       const float* __restrict__ compact,
                                                       float4* __restrict__ tris) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n_tris) return;
    const uint32_t g = leaf_order[p];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const size_t vi = tri_idx[3 * (size_t)g + k];
        const float w = k == 0 ? tris[3 * (size_t)p].w : 0.0f;  // keeps the triangle id
        tris[3 * (size_t)p + k] =
            make_float4(compact[3 * vi], compact[3 * vi + 1], compact[3 * vi + 2], w);
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_refit_level(uint32_t begin, uint32_t count, WNode<W>* nodes,
                                                        const float4* __restrict__ tris,
                                                        const uint32_t* __restrict__ scale) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const float pad = o2f_r(*scale) * 0x1p-18f;
    WNode<W>& nd = nodes[begin + k];
#pragma unroll
    for (int j = 0; j < W; j++) {
        const uint32_t c = nd.child[j];
        if (c == kEmpty) continue;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (child_is_leaf(c)) {
            const uint32_t f = leaf_first(c), cnt = leaf_count(c);
            for (uint32_t t = 0; t < cnt; t++)
                for (int v = 0; v < 3; v++) {
                    const float4 p = tris[3 * (size_t)(f + t) + v];
                    lo[0] = fminf(lo[0], p.x); hi[0] = fmaxf(hi[0], p.x);
                    lo[1] = fminf(lo[1], p.y); hi[1] = fmaxf(hi[1], p.y);
                    lo[2] = fminf(lo[2], p.z); hi[2] = fmaxf(hi[2], p.z);
                }
            for (int a = 0; a < 3; a++) {
                lo[a] -= pad;
                hi[a] += pad;
            }
        } else {
            const WNode<W>& cn = nodes[c];
#pragma unroll
            for (int i = 0; i < W; i++) {
                if (cn.child[i] == kEmpty) continue;
                lo[0] = fminf(lo[0], cn.lox[i]); hi[0] = fmaxf(hi[0], cn.hix[i]);
                lo[1] = fminf(lo[1], cn.loy[i]); hi[1] = fmaxf(hi[1], cn.hiy[i]);
                lo[2] = fminf(lo[2], cn.loz[i]); hi[2] = fmaxf(hi[2], cn.hiz[i]);
            }
        }
        nd.lox[j] = lo[0]; nd.loy[j] = lo[1]; nd.loz[j] = lo[2];
        nd.hix[j] = hi[0]; nd.hiy[j] = hi[1]; nd.hiz[j] = hi[2];
    }
}

template <int W>
static void refit_levels(gpurt_bvh_s* b, const uint32_t* scale, hipStream_t s) {
    std::vector<uint32_t> begin(b->level_cnt.size(), 0);
    for (size_t l = 1; l < begin.size(); l++) begin[l] = begin[l - 1] + b->level_cnt[l - 1];
    for (int l = (int)b->level_cnt.size() - 1; l >= 0; l--) {
        const uint32_t cnt = b->level_cnt[l];
        if (!cnt) continue;
        hipLaunchKernelGGL(k_refit_level<W>, dim3(cdiv(cnt, kBlock)), dim3(kBlock), 0, s, begin[l], cnt,
                           reinterpret_cast<WNode<W>*>(b->nodes), b->tris, scale);
    }
}

int launch_refit(gpurt_bvh_s* b, const float* verts, size_t vstride_floats, hipStream_t s) {
    if (!b->leaf_order || !b->tri_idx || !b->verts) return set_error(GPURT_E_INVALID, "bvh has no mesh copy");
    void* tmp = nullptr;
    GPURT_HIP(hipMallocAsync(&tmp, 256, s));
    uint32_t* scale = static_cast<uint32_t*>(tmp);
    hipLaunchKernelGGL(k_refit_init, dim3(1), dim3(64), 0, s, scale);
    // every new position must be finite; the tree is left as it was otherwise
    hipLaunchKernelGGL(k_refit_check, dim3(cdiv(b->n_verts, kBlock)), dim3(kBlock), 0, s, (uint32_t)b->n_verts,
                       verts, (uint64_t)vstride_floats, scale + 1);
    uint32_t bad = 0;
    hipError_t ce = hipMemcpyAsync(&bad, scale + 1, sizeof(bad), hipMemcpyDeviceToHost, s);
    if (ce == hipSuccess) ce = hipStreamSynchronize(s);
    if (ce != hipSuccess || bad) {
        (void)hipFreeAsync(tmp, s);
        if (ce != hipSuccess) return hip_error(ce, "refit check");
        return set_error(GPURT_E_INVALID, "non-finite vertex position");
    }
    hipLaunchKernelGGL(k_refit_verts, dim3(cdiv(b->n_verts, kBlock)), dim3(kBlock), 0, s, (uint32_t)b->n_verts,
                       verts, (uint64_t)vstride_floats, b->verts, scale);
    hipLaunchKernelGGL(k_refit_tris, dim3(cdiv(b->n_tris, kBlock)), dim3(kBlock), 0, s, (uint32_t)b->n_tris,
                       b->leaf_order, b->tri_idx, b->verts, b->tris);
    switch (b->width) {
        case 2: refit_levels<2>(b, scale, s); break;
        case 4: refit_levels<4>(b, scale, s); break;
        case 8: refit_levels<8>(b, scale, s); break;
    }
    if (has_qnodes(b->format)) launch_quantize(b->width, b->nodes, b->qnodes, (uint32_t)b->n_nodes,
                                                (uint32_t)b->n_nodes, nullptr, s);
    hipError_t e = hipGetLastError();
    (void)hipFreeAsync(tmp, s);
    if (e != hipSuccess) return hip_error(e, "refit launch");
    return GPURT_OK;
}

}  // namespace gpurt_impl
