// shade.h -- per-path FP32 arithmetic of the wavefront path tracer.
//
// Restates the shading code of the reference's ray-generation shader in HIP:
// the RNG (tea / lcg, src/shaders/rt/rtcommon.glsl:99-124), hemisphere and triangle
// sampling (lines 135-175), triangle_hit / triangle_pdf (181-214), make_tanspace,
// power_heuristic (218-230), Blinn-Phong (254-277) and GGX (279-349) with the MAT_*
// dispatch on the BRDF (351-369).  Every product and sum is rounded separately
// (-ffp-contract=off), vectors are combined left to right ((x*a + y*b) + z*c), dot products
// as (x*x' + y*y') + z*z', normalize(v) as v / sqrt(dot(v, v)); cos / sin / pow are the
// device library's.  Pi is the reference's 3.141592.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tri_math.h"

namespace gpurt_impl {
namespace pt {

constexpr float kPi = 3.141592f;              // M_PI, rtcommon.glsl:5
constexpr float kLargeDist = 10000000.0f;     // LARGE_DIST, rtcommon.glsl:7
constexpr float kEps = 0.00001f;              // EPS, rtcommon.glsl:8

// ---- RNG -------------------------------------------------------------------------------

// tea with the reference's constants and 16 rounds; returns the first word.
__device__ __forceinline__ uint32_t tea(uint32_t v0, uint32_t v1) {
    uint32_t s0 = 0;
    for (int n = 0; n < 16; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

__device__ __forceinline__ uint32_t lcg(uint32_t& s) {
    s = 1664525u * s + 1013904223u;
    return s & 0x00FFFFFFu;
}
__device__ __forceinline__ float randf(uint32_t& s) { return (float)lcg(s) / (float)0x01000000; }
// a + lcg % (b - a); callers guarantee b > a
__device__ __forceinline__ uint32_t randu(uint32_t& s, uint32_t a, uint32_t b) { return lcg(s) % (b - a) + a; }

// ---- vector helpers --------------------------------------------------------------------

__device__ __forceinline__ V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 mul(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 scale(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 divs(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return divs(a, length(a)); }
// x*a + y*b + z*c, left to right (GLSL vec3 sums)
__device__ __forceinline__ V3 comb3(V3 x, float a, V3 y, float b, V3 z, float c) {
    return add(add(scale(x, a), scale(y, b)), scale(z, c));
}
// GLSL reflect(I, N) = I - 2 dot(N, I) N
__device__ __forceinline__ V3 reflect(V3 i, V3 n) {
    const float k = 2.0f * dot(n, i);
    return sub(i, scale(n, k));
}
__device__ __forceinline__ bool any_positive(V3 a) { return a.x > 0.0f || a.y > 0.0f || a.z > 0.0f; }

// row-major 3x4 affine transform of a point / 3x3 linear map of a vector:
// r_i = ((m_i0 x + m_i1 y) + m_i2 z) + m_i3
__device__ __forceinline__ V3 xform_point(const float* m, V3 p) {
    V3 r;
    r.x = ((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3];
    r.y = ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7];
    r.z = ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11];
    return r;
}
__device__ __forceinline__ V3 xform_vec3(const float* m, V3 p) {
    V3 r;
    r.x = (m[0] * p.x + m[1] * p.y) + m[2] * p.z;
    r.y = (m[3] * p.x + m[4] * p.y) + m[5] * p.z;
    r.z = (m[6] * p.x + m[7] * p.y) + m[8] * p.z;
    return r;
}
// row-major 4x4 times (x, y, z, w)
__device__ __forceinline__ float4 xform4(const float* m, float x, float y, float z, float w) {
    float4 r;
    r.x = ((m[0] * x + m[1] * y) + m[2] * z) + m[3] * w;
    r.y = ((m[4] * x + m[5] * y) + m[6] * z) + m[7] * w;
    r.z = ((m[8] * x + m[9] * y) + m[10] * z) + m[11] * w;
    r.w = ((m[12] * x + m[13] * y) + m[14] * z) + m[15] * w;
    return r;
}

// ---- sampling --------------------------------------------------------------------------

__device__ __forceinline__ float radical_inverse(uint32_t bits) {
    bits = (bits << 16u) | (bits >> 16u);
    bits = ((bits & 0x55555555u) << 1u) | ((bits & 0xAAAAAAAAu) >> 1u);
    bits = ((bits & 0x33333333u) << 2u) | ((bits & 0xCCCCCCCCu) >> 2u);
    bits = ((bits & 0x0F0F0F0Fu) << 4u) | ((bits & 0xF0F0F0F0u) >> 4u);
    bits = ((bits & 0x00FF00FFu) << 8u) | ((bits & 0xFF00FF00u) >> 8u);
    return (float)bits * 2.3283064365386963e-10f;
}

// cosine-power lobe around z in the frame (x, y, z)
__device__ __forceinline__ V3 cospow_hemisphere(float exponent, uint32_t& seed, V3 x, V3 y, V3 z) {
    const float phi = (2.0f * kPi) * randf(seed);
    const float cosT = powf(randf(seed), 1.0f / (exponent + 1.0f));
    const float sinT = sqrtf(1.0f - cosT * cosT);
    return comb3(x, cosf(phi) * sinT, y, sinf(phi) * sinT, z, cosT);
}

// uniform barycentrics (sqrt warp)
__device__ __forceinline__ V3 triangle_sample(uint32_t& seed) {
    const float u = sqrtf(randf(seed));
    const float v = randf(seed);
    const float a = u * (1.0f - v);
    const float b = u * v;
    return V3{a, b, (1.0f - a) - b};
}

// Moller-Trumbore with the reference's absolute determinant threshold; accepts t >= 0
__device__ __forceinline__ bool triangle_hit(V3 o, V3 d, V3 pa, V3 pb, V3 pc, V3& hitp) {
    const V3 v1 = sub(pb, pa), v2 = sub(pc, pa);
    const V3 p = cross(d, v2);
    const float det = dot(v1, p);
    if (fabsf(det) < kEps) return false;
    const float inv = 1.0f / det;
    const V3 s = sub(o, pa);
    const float u = dot(s, p) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    const V3 q = cross(s, v1);
    const float v = dot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    const float t = dot(v2, q) * inv;
    hitp = add(o, scale(d, t));
    return t >= 0.0f;
}

// area-measure pdf of a uniform point on the triangle, converted to solid angle at o
__device__ __forceinline__ float triangle_pdf(V3 o, V3 d, V3 v0, V3 v1, V3 v2) {
    V3 hitp;
    if (!triangle_hit(o, d, v0, v1, v2, hitp)) return 0.0f;
    const V3 n = cross(sub(v1, v0), sub(v2, v0));
    const float a = 2.0f / length(n);
    const V3 dist = sub(hitp, o);
    const V3 N = normalize(n);
    const float g = dot(dist, dist) / fabsf(dot(N, d));
    return a * g;
}

__device__ __forceinline__ void make_tanspace(V3 N, V3& T, V3& B) {
    if (fabsf(N.x) > fabsf(N.y))
        T = divs(V3{N.z, 0.0f, -N.x}, sqrtf(N.x * N.x + N.z * N.z));
    else
        T = divs(V3{0.0f, -N.z, N.y}, sqrtf(N.y * N.y + N.z * N.z));
    B = cross(N, T);
}

__device__ __forceinline__ float power_heuristic(float a, float b) { return a * a / (a * a + b * b); }

// ---- materials -------------------------------------------------------------------------

struct Mat {
    V3 albedo, emissive;
    float roughness;
};
struct Shade {
    V3 wo, N, T, B;
};

// Blinn-Phong: half-vector cosine-power lobe, eval = albedo * pdf
__device__ __forceinline__ float bp_pdf(const Mat& m, const Shade& s, V3 wi) {
    const float oDn = dot(neg(s.wo), s.N);
    const float iDn = dot(wi, s.N);
    if (oDn <= 0.0f || iDn <= 0.0f) return 0.0f;
    const float e = 1.0f / m.roughness;
    const V3 H = normalize(sub(wi, s.wo));
    const float cosine = fmaxf(dot(H, s.N), 0.0f);
    const float n_pdf = (e + 1.0f) / (2.0f * kPi) * powf(cosine, e);
    return n_pdf / (4.0f * dot(neg(s.wo), H));
}
__device__ __forceinline__ V3 bp_eval(const Mat& m, const Shade& s, V3 wi) { return scale(m.albedo, bp_pdf(m, s, wi)); }
__device__ __forceinline__ bool bp_sample(uint32_t& seed, const Mat& m, const Shade& s, V3& wi) {
    const float e = 1.0f / m.roughness;
    const V3 H = cospow_hemisphere(e, seed, s.T, s.B, s.N);
    wi = reflect(s.wo, H);
    return dot(wi, s.N) > 0.0f;
}

__device__ __forceinline__ V3 ggx_F(V3 r0, float iDn) {
    const float cos5 = powf(1.0f - iDn, 5.0f);
    return V3{r0.x + (1.0f - r0.x) * cos5, r0.y + (1.0f - r0.y) * cos5, r0.z + (1.0f - r0.z) * cos5};
}
__device__ __forceinline__ float ggx_G(float oDn, float iDn, float a2) {
    const float sqr0 = sqrtf(a2 + (1.0f - a2) * iDn * iDn);
    const float sqr1 = sqrtf(a2 + (1.0f - a2) * oDn * oDn);
    return 2.0f * oDn * iDn / (oDn * sqr0 + iDn * sqr1);
}
__device__ __forceinline__ float ggx_D(float nDh, float a2) {
    const float b = nDh * nDh * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * b * b);
}
__device__ __forceinline__ float ggx_pdf(const Mat& m, const Shade& s, V3 wi) {
    const float oDn = dot(neg(s.wo), s.N);
    const float iDn = dot(wi, s.N);
    if (oDn <= 0.0f || iDn <= 0.0f) return 0.0f;
    const V3 H = normalize(sub(wi, s.wo));
    const float nDh = fmaxf(dot(H, s.N), 0.0f);
    const float oDh = fmaxf(dot(wi, H), 0.0f);
    const float a2 = m.roughness * m.roughness;
    return ggx_D(nDh, a2) * nDh / (4.0f * oDh);
}
__device__ __forceinline__ V3 ggx_eval(const Mat& m, const Shade& s, V3 wi) {
    const float oDn = dot(neg(s.wo), s.N);
    const float iDn = dot(wi, s.N);
    if (oDn <= 0.0f || iDn <= 0.0f) return V3{0.0f, 0.0f, 0.0f};
    const V3 H = normalize(sub(wi, s.wo));
    const float nDh = fmaxf(dot(H, s.N), 0.0f);
    const float a2 = m.roughness * m.roughness;
    const V3 F = ggx_F(m.albedo, iDn);
    const float D = ggx_D(nDh, a2), G = ggx_G(oDn, iDn, a2), den = 4.0f * oDn;
    return V3{F.x * D * G / den, F.y * D * G / den, F.z * D * G / den};
}
__device__ __forceinline__ bool ggx_sample(uint32_t& seed, const Mat& m, const Shade& s, V3& wi) {
    const float a2 = m.roughness * m.roughness;
    const float x0 = randf(seed);
    const float x1 = randf(seed);
    const float phi = 2.0f * kPi * x0;
    const float cosTheta = sqrtf((1.0f - x1) / (1.0f + (a2 - 1.0f) * x1));
    const float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
    const V3 H = comb3(s.T, cosf(phi) * sinTheta, s.B, sinf(phi) * sinTheta, s.N, cosTheta);
    wi = reflect(s.wo, H);
    return dot(wi, s.N) > 0.0f;
}

template <int BRDF>
__device__ __forceinline__ float mat_pdf(const Mat& m, const Shade& s, V3 wi) {
    return BRDF == 0 ? bp_pdf(m, s, wi) : ggx_pdf(m, s, wi);
}
template <int BRDF>
__device__ __forceinline__ V3 mat_eval(const Mat& m, const Shade& s, V3 wi) {
    return BRDF == 0 ? bp_eval(m, s, wi) : ggx_eval(m, s, wi);
}
template <int BRDF>
__device__ __forceinline__ bool mat_sample(uint32_t& seed, const Mat& m, const Shade& s, V3& wi) {
    return BRDF == 0 ? bp_sample(seed, m, s, wi) : ggx_sample(seed, m, s, wi);
}

}  // namespace pt
}  // namespace gpurt_impl
