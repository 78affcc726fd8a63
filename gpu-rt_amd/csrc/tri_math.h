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

}  // namespace gpurt_impl
