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
// The terms of one query or ray are formed here.  They are either scatter-added with
// hardware FP32 atomics (the order, and so the last bits, vary between runs) or sorted by
// vertex and added in a fixed order (below); the context's vertex-sum mode picks one.
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

// hardware FP32 atomic scatter of the terms (the order of the adds varies between runs)
__device__ __forceinline__ void scatter_terms(const CornerTerms& t, float* g_verts) {
    if (!g_verts) return;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (t.v[j] == kNoVertex) continue;
        const size_t v = t.v[j];
        unsafeAtomicAdd(&g_verts[3 * v], t.c[j][0]);
        unsafeAtomicAdd(&g_verts[3 * v + 1], t.c[j][1]);
        unsafeAtomicAdd(&g_verts[3 * v + 2], t.c[j][2]);
    }
}

// ---- fixed-order vertex sums ---------------------------------------------------------
// The terms, sorted by key (a vertex, or (mesh, vertex) in a batch; `none` for terms that
// add nothing), are summed per key in two passes with a fixed order (the same bits on
// every run), and without a serial loop over a long run:
//   k_wave_segsum  one thread per sorted term: within each 64-term wave, a segmented
//                  suffix scan over shuffles; the first term of each key in the wave keeps
//                  the sum of that key's terms in the wave (partial[p])
//   k_run_sum      the first term of each key's run adds the partials of the waves the
//                  run spans (one per 64 terms) and hands the total to the sink
template <int BS>
__global__ __launch_bounds__(BS) void k_wave_segsum(uint32_t m, const uint64_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ vals, const float* __restrict__ terms,
                                                    float* __restrict__ partial) {
    const uint32_t p = blockIdx.x * BS + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const bool in = p < m;
    const uint64_t key = in ? keys[p] : ~0ull;
    const bool head = in && (lane == 0 || keys[p - 1] != key);
    float x = 0.f, y = 0.f, z = 0.f;
    if (in) {
        const size_t k = vals[p];
        x = terms[3 * k];
        y = terms[3 * k + 1];
        z = terms[3 * k + 2];
    }
    // lanes of this key in this wave: [lane, next head)
    const uint64_t heads = __ballot(head);
    const uint64_t after = lane == 63 ? 0ull : heads >> (lane + 1);
    const uint32_t end = after ? lane + 1 + (uint32_t)__builtin_ctzll(after) : 64u;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const float ox = __shfl_down(x, d), oy = __shfl_down(y, d), oz = __shfl_down(z, d);
        if (lane + d < end) {
            x += ox;
            y += oy;
            z += oz;
        }
    }
    if (head) {
        partial[3 * (size_t)p] = x;
        partial[3 * (size_t)p + 1] = y;
        partial[3 * (size_t)p + 2] = z;
    }
}

template <int BS, typename Sink>
__global__ __launch_bounds__(BS) void k_run_sum(uint32_t m, const uint64_t* __restrict__ keys,
                                                const float* __restrict__ partial, uint64_t none, Sink sink) {
    const uint32_t p = blockIdx.x * BS + threadIdx.x;
    if (p >= m) return;
    const uint64_t key = keys[p];
    if (key >= none || (p > 0 && keys[p - 1] == key)) return;
    float x = partial[3 * (size_t)p], y = partial[3 * (size_t)p + 1], z = partial[3 * (size_t)p + 2];
    for (uint32_t q = (p | 63u) + 1u; q < m && keys[q] == key; q += 64u) {
        x += partial[3 * (size_t)q];
        y += partial[3 * (size_t)q + 1];
        z += partial[3 * (size_t)q + 2];
    }
    sink(key, x, y, z);
}

}  // namespace gpurt_impl
