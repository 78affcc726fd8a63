// backward.h -- per-query bodies of the backward passes, shared by the plain kernels
// (backward.hip) and the batch kernels (batch.hip).
//
//   distance  d = |q - p|, p = sum_j b_j v_j the closest point (envelope theorem, the
//             minimiser over the convex triangle is unique):
//             dd/dq = (q - p)/d,  dd/dv_j = -b_j (q - p)/d
//             (p, b and q - p in FP64 on the saved triangle)
//   hit t     t = n.(v0 - o)/(n.d), n = (v1-v0) x (v2-v0), b = (1-u-v, u, v) fixed:
//             dt/do = -n/(n.d),  dt/dd = -t n/(n.d),  dt/dv_j = b_j n/(n.d)
//
// Vertex positions are read with a stride of `vs` floats; vertex gradients are n x 3.
// The terms of one query or ray are formed here; the plain (backward.hip) and batch
// (batch.hip) kernels sort them by vertex and add each vertex's terms in a fixed order,
// without atomics, so vertex gradients are the same bits on every run.
#pragma once

#include "gpurt_internal.h"
#include "tri_math.h"

namespace gpurt_impl {

__device__ __forceinline__ V3 vert_at(const float* __restrict__ verts, uint64_t vs, size_t v) {
    return v3(verts[vs * v], verts[vs * v + 1], verts[vs * v + 2]);
}

// The three vertex terms of one query or ray: vertex ids and the float3 added to each
// (weight 0 or a dead query: nothing to add, vertex id kNoVertex).
constexpr uint32_t kNoVertex = 0xFFFFFFFFu;
struct CornerTerms {
    uint32_t v[3];
    float c[3][3];
};

// Query i of a closest-point backward.  The saved id selects the triangle (no
// re-traversal); its closest point, offset q - p and barycentrics are recomputed in FP64
// (closest_pt_tri_f64_core, the computation the forward pass reports from), so the
// direction (q - p)/|q - p| stays accurate however close the query is to the surface --
// a point rounded to FP32 would be off by ~1e-7 of the coordinates, a large angle when
// |q - p| is tiny.  The saved pt / dist / bary are not needed.  Distance 0 contributes
// nothing (the gradient is undefined on the surface).  Writes g_q[i] (when g_q) and
// returns the vertex terms -b_j g_q.
__device__ __forceinline__ CornerTerms cpq_bwd_terms(size_t i, const float* __restrict__ q,
                                                     const int32_t* __restrict__ prim, const float* __restrict__ g,
                                                     const uint32_t* __restrict__ tri_idx, uint32_t n_tris,
                                                     const float* __restrict__ verts, uint64_t vs,
                                                     float* __restrict__ g_q) {
    const int32_t p = prim[i];
    float gx = 0.f, gy = 0.f, gz = 0.f;
    double b[3] = {0.0, 0.0, 0.0};
    bool live = p >= 0 && (uint32_t)p < n_tris;
    if (live) {
        const V3 qq = v3(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
        const CpF64 r = closest_pt_tri_f64_core(qq, vert_at(verts, vs, tri_idx[3 * (size_t)p]),
                                                vert_at(verts, vs, tri_idx[3 * (size_t)p + 1]),
                                                vert_at(verts, vs, tri_idx[3 * (size_t)p + 2]));
        const double d = sqrt(r.d2);
        live = d > 0.0;
        if (live) {
            const double s = (double)g[i] / d;
            gx = (float)(r.df.x * s);
            gy = (float)(r.df.y * s);
            gz = (float)(r.df.z * s);
            b[0] = r.b0, b[1] = r.b1, b[2] = r.b2;
        }
    }
    if (g_q) {
        g_q[3 * i] = gx;
        g_q[3 * i + 1] = gy;
        g_q[3 * i + 2] = gz;
    }
    CornerTerms t;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const float bj = (float)b[j];
        t.v[j] = live && bj != 0.0f ? tri_idx[3 * (size_t)p + j] : kNoVertex;
        t.c[j][0] = -bj * gx;
        t.c[j][1] = -bj * gy;
        t.c[j][2] = -bj * gz;
    }
    return t;
}

// Ray i of a hit-distance backward (misses and rays parallel to the plane contribute
// nothing).  Writes g_o[i] / g_d[i] (when given) and returns the vertex terms
// b_j n / (n.d) scaled by g (vertex ids kNoVertex without uv).
__device__ __forceinline__ CornerTerms ray_bwd_terms(size_t i, const float4* __restrict__ rays,
                                                     const int32_t* __restrict__ prim, const float* __restrict__ t,
                                                     const float* __restrict__ uv, const float* __restrict__ g,
                                                     const uint32_t* __restrict__ tri_idx, uint32_t n_tris,
                                                     const float* __restrict__ verts, uint64_t vs,
                                                     float* __restrict__ g_o, float* __restrict__ g_d) {
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
    CornerTerms ct;
    float bw[3] = {0.f, 0.f, 0.f};
    if (uv && live) {
        const float u = uv[2 * i], v = uv[2 * i + 1];
        bw[0] = (1.0f - u) - v;
        bw[1] = u;
        bw[2] = v;
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const float w = bw[j] * s;
        ct.v[j] = uv && live && w != 0.0f ? vi[j] : kNoVertex;
        ct.c[j][0] = nrm.x * w;
        ct.c[j][1] = nrm.y * w;
        ct.c[j][2] = nrm.z * w;
    }
    return ct;
}

}  // namespace gpurt_impl
