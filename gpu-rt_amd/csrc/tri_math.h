// tri_math.h -- per-triangle FP32 arithmetic of the traversal kernels.
//
// The operation order below is the library's numerical contract: every product and sum
// is rounded to FP32 separately (the library is compiled with -ffp-contract=off and
// correctly rounded division/sqrt), so primitive ids, distances and hit distances are
// reproducible bit for bit by any CPU code that performs the same steps.
#pragma once

#include <hip/hip_runtime.h>

namespace gpurt_impl {

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }

// (a.x*b.x + a.y*b.y) + a.z*b.z
__device__ __forceinline__ float dot(V3 a, V3 b) {
    float s = a.x * b.x;
    s = s + a.y * b.y;
    s = s + a.z * b.z;
    return s;
}

// GLSL component order (src/shaders/rt/rtcommon.glsl uses the built-in cross)
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Closest point on triangle (a,b,c) to p: Ericson, Real-Time Collision Detection (2005)
// sec. 5.1.5, Voronoi-region tests in the book's order.  The offset p - cp is formed
// relative to the region's base vertex ((p-a) - (v*ab + w*ac), ...), so its rounding
// error scales with the triangle's size, not the coordinates.  Returns |p - cp|^2.
// All region quantities are computed before the region is chosen (less divergence);
// the tests themselves run in the book's order, so results equal the sequential form.
__device__ __forceinline__ float closest_pt_tri(V3 p, V3 a, V3 b, V3 c, V3& cp, V3& bary) {
    const V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const float d1 = dot(ab, ap);
    const float d2 = dot(ac, ap);
    V3 df;
    if (d1 <= 0.0f && d2 <= 0.0f) {  // vertex a
        df = ap;
        cp = a;
        bary = v3(1.0f, 0.0f, 0.0f);
        return dot(df, df);
    }
    const V3 bp = sub(p, b);
    const float d3 = dot(ab, bp);
    const float d4 = dot(ac, bp);
    const float vc = d1 * d4 - d3 * d2;
    const V3 cpv = sub(p, c);
    const float d5 = dot(ab, cpv);
    const float d6 = dot(ac, cpv);
    const float vb = d5 * d2 - d1 * d6;
    const float va = d3 * d6 - d5 * d4;
    const float e43 = d4 - d3, e56 = d5 - d6;
    if (d3 >= 0.0f && d4 <= d3) {  // vertex b
        df = bp;
        cp = b;
        bary = v3(0.0f, 1.0f, 0.0f);
    } else if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {  // edge ab
        const float v = d1 / (d1 - d3);
        const V3 off = v3(v * ab.x, v * ab.y, v * ab.z);
        df = sub(ap, off);
        cp = v3(a.x + off.x, a.y + off.y, a.z + off.z);
        bary = v3(1.0f - v, v, 0.0f);
    } else if (d6 >= 0.0f && d5 <= d6) {  // vertex c
        df = cpv;
        cp = c;
        bary = v3(0.0f, 0.0f, 1.0f);
    } else if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {  // edge ac
        const float w = d2 / (d2 - d6);
        const V3 off = v3(w * ac.x, w * ac.y, w * ac.z);
        df = sub(ap, off);
        cp = v3(a.x + off.x, a.y + off.y, a.z + off.z);
        bary = v3(1.0f - w, 0.0f, w);
    } else if (va <= 0.0f && e43 >= 0.0f && e56 >= 0.0f) {  // edge bc
        const float w = e43 / (e43 + e56);
        const V3 bc = sub(c, b);
        const V3 off = v3(w * bc.x, w * bc.y, w * bc.z);
        df = sub(bp, off);
        cp = v3(b.x + off.x, b.y + off.y, b.z + off.z);
        bary = v3(0.0f, 1.0f - w, w);
    } else {  // face interior
        const float denom = 1.0f / ((va + vb) + vc);
        const float v = vb * denom;
        const float w = vc * denom;
        const V3 off = v3(ab.x * v + ac.x * w, ab.y * v + ac.y * w, ab.z * v + ac.z * w);
        df = sub(ap, off);
        cp = v3(a.x + off.x, a.y + off.y, a.z + off.z);
        bary = v3((1.0f - v) - w, v, w);
    }
    return dot(df, df);
}

// Moller-Trumbore in the operation order of triangle_hit (rtcommon.glsl:181-202):
// v1=pb-pa, v2=pc-pa, p=d x v2, det=v1.p, inv=1/det, s=o-pa, u=(s.p)*inv, q=s x v1,
// v=(d.q)*inv, t=(v2.q)*inv.  Only det == 0 is rejected (no absolute-epsilon cull).
// Returns true when (u,v) lies in the closed triangle; t is then valid.
__device__ __forceinline__ bool ray_tri(V3 o, V3 d, V3 pa, V3 pb, V3 pc, float& t, float& u,
                                        float& v) {
    V3 v1 = sub(pb, pa), v2 = sub(pc, pa);
    V3 p = cross(d, v2);
    float det = dot(v1, p);
    if (det == 0.0f) return false;
    float inv = 1.0f / det;
    V3 s = sub(o, pa);
    u = dot(s, p) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    V3 q = cross(s, v1);
    v = dot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(v2, q) * inv;
    return true;
}

// The closest point of the winning triangle, recomputed in FP64 for the reported point,
// distance and barycentrics: the same region tests (Ericson 5.1.5, book order, offsets
// from the region's base vertex) on the FP32 inputs converted exactly to double, each
// product and sum rounded to FP64, then rounded once to FP32 (correctly rounded sqrt and
// division).  The FP32 search above chooses the triangle; this only sharpens what is
// reported about it (relative error ~1 ulp of FP32 instead of growing as the query
// approaches the surface).  Returns the distance.
struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3(V3 a) { return D3{(double)a.x, (double)a.y, (double)a.z}; }
__device__ __forceinline__ D3 dsub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double ddot(D3 a, D3 b) {
    double s = a.x * b.x;
    s = s + a.y * b.y;
    s = s + a.z * b.z;
    return s;
}
__device__ __forceinline__ V3 to_f32(D3 a) { return v3((float)a.x, (float)a.y, (float)a.z); }

// Result of the FP64 closest-point computation: the point q, the offset df = p - q from
// it to the query (formed from the region's base vertex), the barycentrics and |df|^2.
struct CpF64 {
    D3 q, df;
    double b0, b1, b2, d2;
};

__device__ __forceinline__ CpF64 closest_pt_tri_f64_core(V3 pf, V3 af, V3 bf, V3 cf) {
    const D3 p = d3(pf), a = d3(af), b = d3(bf), c = d3(cf);
    const D3 ab = dsub(b, a), ac = dsub(c, a), ap = dsub(p, a);
    const double d1 = ddot(ab, ap), d2 = ddot(ac, ap);
    CpF64 r;
    if (d1 <= 0.0 && d2 <= 0.0) {  // vertex a
        r.q = a, r.df = ap, r.b0 = 1.0, r.b1 = 0.0, r.b2 = 0.0;
        r.d2 = ddot(ap, ap);
        return r;
    }
    const D3 bp = dsub(p, b);
    const double d3_ = ddot(ab, bp), d4 = ddot(ac, bp);
    if (d3_ >= 0.0 && d4 <= d3_) {  // vertex b
        r.q = b, r.df = bp, r.b0 = 0.0, r.b1 = 1.0, r.b2 = 0.0;
        r.d2 = ddot(bp, bp);
        return r;
    }
    const double vc = d1 * d4 - d3_ * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3_ <= 0.0) {  // edge ab
        const double v = d1 / (d1 - d3_);
        const D3 off = D3{v * ab.x, v * ab.y, v * ab.z};
        r.df = dsub(ap, off);
        r.q = D3{a.x + off.x, a.y + off.y, a.z + off.z};
        r.b0 = 1.0 - v, r.b1 = v, r.b2 = 0.0;
        r.d2 = ddot(r.df, r.df);
        return r;
    }
    const D3 cpv = dsub(p, c);
    const double d5 = ddot(ab, cpv), d6 = ddot(ac, cpv);
    if (d6 >= 0.0 && d5 <= d6) {  // vertex c
        r.q = c, r.df = cpv, r.b0 = 0.0, r.b1 = 0.0, r.b2 = 1.0;
        r.d2 = ddot(cpv, cpv);
        return r;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {  // edge ac
        const double w = d2 / (d2 - d6);
        const D3 off = D3{w * ac.x, w * ac.y, w * ac.z};
        r.df = dsub(ap, off);
        r.q = D3{a.x + off.x, a.y + off.y, a.z + off.z};
        r.b0 = 1.0 - w, r.b1 = 0.0, r.b2 = w;
        r.d2 = ddot(r.df, r.df);
        return r;
    }
    const double va = d3_ * d6 - d5 * d4;
    const double e43 = d4 - d3_, e56 = d5 - d6;
    if (va <= 0.0 && e43 >= 0.0 && e56 >= 0.0) {  // edge bc
        const double w = e43 / (e43 + e56);
        const D3 bc = dsub(c, b);
        const D3 off = D3{w * bc.x, w * bc.y, w * bc.z};
        r.df = dsub(bp, off);
        r.q = D3{b.x + off.x, b.y + off.y, b.z + off.z};
        r.b0 = 0.0, r.b1 = 1.0 - w, r.b2 = w;
        r.d2 = ddot(r.df, r.df);
        return r;
    }
    // face interior
    const double denom = 1.0 / ((va + vb) + vc);
    const double v = vb * denom, w = vc * denom;
    const D3 off = D3{ab.x * v + ac.x * w, ab.y * v + ac.y * w, ab.z * v + ac.z * w};
    r.df = dsub(ap, off);
    r.q = D3{a.x + off.x, a.y + off.y, a.z + off.z};
    r.b0 = (1.0 - v) - w, r.b1 = v, r.b2 = w;
    r.d2 = ddot(r.df, r.df);
    return r;
}

// The reported point, barycentrics and distance, each rounded once to FP32.
__device__ __forceinline__ float closest_pt_tri_f64(V3 pf, V3 af, V3 bf, V3 cf, V3& cp, V3& bary) {
    const CpF64 r = closest_pt_tri_f64_core(pf, af, bf, cf);
    cp = to_f32(r.q);
    bary = v3((float)r.b0, (float)r.b1, (float)r.b2);
    return (float)sqrt(r.d2);
}

// The finishing pass of one closest-point query: the winner `prim[i]` (from the FP32
// search) of a mesh given as positions at a `vs`-float stride and u32 index triples;
// writes the FP64-derived point, distance and barycentrics (bary nullable).  A miss
// (prim -1, a mesh without triangles) gets distance +inf.  Indices are clamped to the
// vertex range (the build reports bad indices; this only keeps the pass in bounds).
__device__ __forceinline__ void cpq_finish_one(size_t i, const float* __restrict__ q, const int32_t* __restrict__ prim,
                                               const uint32_t* __restrict__ tri_idx, uint32_t n_tris,
                                               const float* __restrict__ verts, uint64_t vs, uint32_t n_verts,
                                               float* __restrict__ pt, float* __restrict__ dist,
                                               float* __restrict__ bary) {
    const int32_t id = prim[i];
    V3 cp = v3(0.0f, 0.0f, 0.0f), br = cp;
    float d = INFINITY;
    if (id >= 0 && (uint32_t)id < n_tris) {
        V3 t[3];
        for (int k = 0; k < 3; k++) {
            uint32_t vi = tri_idx[3 * (size_t)id + k];
            vi = vi < n_verts ? vi : 0u;
            const float* p = verts + vs * vi;
            t[k] = v3(p[0], p[1], p[2]);
        }
        d = closest_pt_tri_f64(v3(q[3 * i], q[3 * i + 1], q[3 * i + 2]), t[0], t[1], t[2], cp, br);
    }
    pt[3 * i] = cp.x;
    pt[3 * i + 1] = cp.y;
    pt[3 * i + 2] = cp.z;
    dist[i] = d;
    if (bary) {
        bary[3 * i] = br.x;
        bary[3 * i + 1] = br.y;
        bary[3 * i + 2] = br.z;
    }
}

}  // namespace gpurt_impl
