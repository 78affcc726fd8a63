// build_common.h -- device helpers shared by the builder (build.hip) and the treelet
// restructuring (treelet.hip).
#pragma once

#include "gpurt_internal.h"

namespace gpurt_impl {

// per-mesh statistics: [cmin xyz, cmax xyz, max|coord|, unused] as ordered uints.
// k_prep accumulates into kRep replicas (block b uses replica b % kRep) so that the
// atomics of ~N/64 waves do not serialise on 7 addresses; k_reduce_stats folds them.
constexpr int kStat = 8;
constexpr int kRep = 64;

// IEEE floats mapped to uints with the same order (for atomic min/max)
__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t u) {
    u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return __uint_as_float(u);
}

// Binary references: internal node i -> i; leaf (sorted primitive p) -> kLeafBit | p.
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

__device__ __forceinline__ float box_area(float4 lo, float4 hi) {
    float ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
    return 2.0f * (ex * ez + ex * ey + ey * ez);
}

// SAH costs: an internal node costs kCi per unit of area plus its children; a leaf of
// c triangles costs kCt * c per unit of area.  Subtrees of <= leaf_size triangles are
// always collapsed into one leaf by the wide collapse, so they are priced as leaves.
constexpr float kCi = 1.2f;
constexpr float kCt = 1.0f;

__device__ __forceinline__ float ref_cost(uint32_t r, const float4* __restrict__ saabb,
                                          const float* __restrict__ cost) {
    if (r & kLeafBit) {
        size_t p = r & ~kLeafBit;
        return kCt * box_area(saabb[2 * p], saabb[2 * p + 1]);
    }
    return cost[r];
}

__device__ __forceinline__ float sah_node_cost(float4 lo, float4 hi, uint32_t count, float cl, float cr,
                                               int leaf_size) {
    const float a = box_area(lo, hi);
    if (count <= (uint32_t)leaf_size) return kCt * a * (float)count;
    return kCi * a + cl + cr;
}

}  // namespace gpurt_impl
