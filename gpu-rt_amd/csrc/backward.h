// backward.h -- per-query bodies of the backward passes, shared by the plain kernels
// (backward.hip) and the batch kernels (batch.hip).
//
//   distance  d = |q - p|, p = sum_j b_j v_j the closest point (envelope theorem, the
//             minimiser over the convex triangle is unique):
//             dd/dq = (q - p)/d,  dd/dv_j = -b_j (q - p)/d
//   hit t     t = n.(v0 - o)/(n.d), n = (v1-v0) x (v2-v0), b = (1-u-v, u, v) fixed:
//             dt/do = -n/(n.d),  dt/dd = -t n/(n.d),  dt/dv_j = b_j n/(n.d)
//
// Vertex positions are read with a stride of `vs` floats; vertex gradients are n x 3
// and accumulated with hardware FP32 atomic adds.
#pragma once

#include "gpurt_internal.h"
#include "tri_math.h"

namespace gpurt_impl {

__device__ __forceinline__ V3 vert_at(const float* __restrict__ verts, uint64_t vs, size_t v) {
    return v3(verts[vs * v], verts[vs * v + 1], verts[vs * v + 2]);
}

// Query i of a closest-point backward.  The offset q - p is re-formed relative to the
// triangle's first vertex, (q - v0) - (b1 (v1 - v0) + b2 (v2 - v0)), like the forward
// pass, instead of from the rounded closest point: its error then scales with the
// triangle, not the coordinates, which keeps the direction accurate for queries very
// close to the surface.  Without barycentrics it falls back to q - pt (and no vertex
// gradient can be formed).
__device__ __forceinline__ void cpq_bwd_one(size_t i, const float* __restrict__ q, const int32_t* __restrict__ prim,
                                            const float* __restrict__ bary, const float* __restrict__ pt,
                                            const float* __restrict__ dist, const float* __restrict__ g,
                                            const uint32_t* __restrict__ tri_idx, uint32_t n_tris,
                                            const float* __restrict__ verts, uint64_t vs, float* __restrict__ g_q,
                                            float* g_verts) {
    const int32_t p = prim[i];
    const float d = dist[i];
    float gx = 0.f, gy = 0.f, gz = 0.f;
    // ids outside the mesh (not produced by the forward pass) contribute nothing
    const bool live = p >= 0 && (uint32_t)p < n_tris && d > 0.0f && d < INFINITY;
    if (live) {
        const float s = g[i] / d;
        const V3 qq = v3(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
        V3 df;
        if (bary) {
            const V3 a = vert_at(verts, vs, tri_idx[3 * (size_t)p]);
            const V3 e1 = sub(vert_at(verts, vs, tri_idx[3 * (size_t)p + 1]), a);
            const V3 e2 = sub(vert_at(verts, vs, tri_idx[3 * (size_t)p + 2]), a);
            const float b1 = bary[3 * i + 1], b2 = bary[3 * i + 2];
            const V3 ap = sub(qq, a);
            df = v3(ap.x - (e1.x * b1 + e2.x * b2), ap.y - (e1.y * b1 + e2.y * b2), ap.z - (e1.z * b1 + e2.z * b2));
        } else {
            df = sub(qq, v3(pt[3 * i], pt[3 * i + 1], pt[3 * i + 2]));
        }
        gx = df.x * s;
        gy = df.y * s;
        gz = df.z * s;
    }
    if (g_q) {
        g_q[3 * i] = gx;
        g_q[3 * i + 1] = gy;
        g_q[3 * i + 2] = gz;
    }
    if (g_verts && bary && live) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const float b = bary[3 * i + j];
            if (b == 0.0f) continue;
            const size_t v = tri_idx[3 * (size_t)p + j];
            unsafeAtomicAdd(&g_verts[3 * v], -b * gx);
            unsafeAtomicAdd(&g_verts[3 * v + 1], -b * gy);
            unsafeAtomicAdd(&g_verts[3 * v + 2], -b * gz);
        }
    }
}

// Ray i of a hit-distance backward (misses and rays parallel to the plane contribute
// nothing).
__device__ __forceinline__ void ray_bwd_one(size_t i, const float4* __restrict__ rays, const int32_t* __restrict__ prim,
                                            const float* __restrict__ t, const float* __restrict__ uv,
                                            const float* __restrict__ g, const uint32_t* __restrict__ tri_idx,
                                            uint32_t n_tris, const float* __restrict__ verts, uint64_t vs,
                                            float* __restrict__ g_o, float* __restrict__ g_d, float* g_verts) {
    const int32_t p = prim[i];
    V3 go = v3(0.f, 0.f, 0.f), gd = v3(0.f, 0.f, 0.f);
    V3 nrm = v3(0.f, 0.f, 0.f);
    float s = 0.f;
    uint32_t vi[3] = {0, 0, 0};
    bool live = false;
    if (p >= 0 && (uint32_t)p < n_tris) {
        const float4 r1 = rays[2 * i + 1];
        V3 pv[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            vi[j] = tri_idx[3 * (size_t)p + j];
            pv[j] = vert_at(verts, vs, vi[j]);
        }
        nrm = cross(sub(pv[1], pv[0]), sub(pv[2], pv[0]));
        const float nd = dot(nrm, v3(r1.x, r1.y, r1.z));
        if (nd != 0.0f) {
            live = true;
            s = g[i] / nd;
            go = v3(-nrm.x * s, -nrm.y * s, -nrm.z * s);
            const float ts = t[i] * s;
            gd = v3(-nrm.x * ts, -nrm.y * ts, -nrm.z * ts);
        }
    }
    if (g_o) {
        g_o[3 * i] = go.x;
        g_o[3 * i + 1] = go.y;
        g_o[3 * i + 2] = go.z;
    }
    if (g_d) {
        g_d[3 * i] = gd.x;
        g_d[3 * i + 1] = gd.y;
        g_d[3 * i + 2] = gd.z;
    }
    if (g_verts && uv && live) {
        const float u = uv[2 * i], v = uv[2 * i + 1];
        const float bw[3] = {(1.0f - u) - v, u, v};
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const float w = bw[j] * s;
            if (w == 0.0f) continue;
            unsafeAtomicAdd(&g_verts[3 * (size_t)vi[j]], nrm.x * w);
            unsafeAtomicAdd(&g_verts[3 * (size_t)vi[j] + 1], nrm.y * w);
            unsafeAtomicAdd(&g_verts[3 * (size_t)vi[j] + 2], nrm.z * w);
        }
    }
}

}  // namespace gpurt_impl
