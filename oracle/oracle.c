/*
 * oracle.c -- CPU restatement of the geometric-query path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or calls
 * this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, and only as the test oracle (or, for the baseline, as the timed CPU port).
 *
 * The reference (TheNumbat/GPU-RT, main branch) builds and traverses its BVH inside
 * the closed Vulkan driver (src/vk/vulkan.cpp:877 vkCmdBuildAccelerationStructuresKHR,
 * src/shaders/rt/rt.rgen:258,278 traceRayEXT), so there is no CPU BVH in it to run.
 * This file states the path's arithmetic on the CPU, from the few places the reference
 * writes it down, plus textbook algorithms that are named where used:
 *
 *   FP32 restatement (bit-exact semantics shared with the GPU kernels):
 *     orc_cp_tri_f32      closest point on a triangle: Ericson, Real-Time Collision
 *                         Detection (2005) sec. 5.1.5 ClosestPtPointTriangle, with the
 *                         operation order written out below.
 *     orc_mt_f32          ray/triangle: Moller-Trumbore in the operation order of
 *                         src/shaders/rt/rtcommon.glsl:181-202 (triangle_hit), except
 *                         that only det == 0 is rejected (the reference's 1e-5 absolute
 *                         cull only exists in its light-pdf helper; the hardware
 *                         traversal that rt.rgen:258,278 calls has no such cull), and
 *                         the accepted interval is [tmin, tmax] (rt.rgen:257-291).
 *     orc_cp_tri_refine   the reported point / distance / barycentrics of the winning
 *                         triangle: the same region tests in FP64, rounded once to FP32
 *                         (the FP32 test chooses the triangle, this sharpens the report).
 *     *_brute_f32         global (value, lowest index) minimum over all triangles;
 *                         any-hit = "some triangle has t in [tmin, tmax]"
 *                         (rt.rgen:272-291, rt.rmiss:9-11).
 *     orc_inst_*          two-level instancing (the TLAS of vulkan.cpp:777-856): each
 *                         instance's mesh is tested with the object-space ray
 *                         o' = M^-1 (o - c), d' = M^-1 d (M^-1 = FP32 rounding of the FP64
 *                         cofactor inverse); minimum over (t, instance, triangle).
 *     orc_lbvh_*          CPU LBVH: 30-bit Morton codes of AABB centroids over the
 *                         centroid bounds, stable sort, Karras (HPG 2012) radix-tree
 *                         hierarchy, leaves of <= leaf_size primitives; closest-first
 *                         stack traversal with conservative (padded) box culls, so the
 *                         answers equal the brute-force FP32 answers for any tree shape.
 *
 *   FP64 independent oracle (different formulation, used to pin the FP32 restatement):
 *     orc_cpq_brute_f64   closest point = plane projection if inside, else the best of
 *                         the three clamped segment projections.
 *     orc_ray_closest_brute_f64  plane intersection + signed edge functions.
 *     orc_*_grad_f64      closed-form gradients (envelope theorem; implicit plane hit).
 *
 * Compile with -ffp-contract=off and without fast-math: the FP32 functions must round
 * exactly like the GPU kernels.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------------------ */
/* tiny parallel-for                                                                    */
/* ------------------------------------------------------------------------------------ */

static int g_threads = 0;

ORC_API void orc_set_threads(int n) { g_threads = n; }
ORC_API int orc_num_threads(void) {
    if (g_threads > 0) return g_threads;
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

typedef void (*range_fn)(void* ctx, int64_t lo, int64_t hi);
typedef struct {
    range_fn fn;
    void* ctx;
    int64_t n, chunk;
    int64_t next;
} pfor_job;

static void* pfor_worker(void* arg) {
    pfor_job* j = (pfor_job*)arg;
    for (;;) {
        int64_t lo = __atomic_fetch_add(&j->next, j->chunk, __ATOMIC_RELAXED);
        if (lo >= j->n) break;
        int64_t hi = lo + j->chunk < j->n ? lo + j->chunk : j->n;
        j->fn(j->ctx, lo, hi);
    }
    return NULL;
}

static void parallel_for(int64_t n, int64_t chunk, range_fn fn, void* ctx) {
    if (n <= 0) return;
    int nt = orc_num_threads();
    if (nt > 256) nt = 256;
    pfor_job job = {fn, ctx, n, chunk < 1 ? 1 : chunk, 0};
    if (nt <= 1 || n <= chunk) {
        fn(ctx, 0, n);
        return;
    }
    pthread_t th[256];
    int started = 0;
    for (int i = 1; i < nt; i++)
        if (pthread_create(&th[started], NULL, pfor_worker, &job) == 0) started++;
    pfor_worker(&job);
    for (int i = 0; i < started; i++) pthread_join(th[i], NULL);
}

/* the same, for pathtrace.c */
typedef void (*orc_range_fn)(void* ctx, int64_t lo, int64_t hi);
void orc_parallel_for(int64_t n, int64_t chunk, orc_range_fn fn, void* ctx) { parallel_for(n, chunk, fn, ctx); }

/* ------------------------------------------------------------------------------------ */
/* FP32 restatement: per-primitive arithmetic                                           */
/* ------------------------------------------------------------------------------------ */

/* dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, every product and sum rounded to FP32. */
static inline float dot3f(const float* a, const float* b) {
    float s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}

/* Ericson RTCD 5.1.5, FP32.  Returns the squared distance and writes the closest point
 * cp and its barycentrics (b0,b1,b2) w.r.t. (a,b,c).  Region tests follow the book's
 * order; the offset from the query to the closest point is formed relative to the
 * region's base vertex (p - cp = (p - a) - (v*ab + w*ac), etc.), so its rounding error
 * scales with the triangle's size rather than with the coordinates' magnitude.  Every
 * rounding step below is the contract the GPU kernel reproduces. */
ORC_API float orc_cp_tri_f32(const float* p, const float* a, const float* b, const float* c,
                             float* cp, float* bary) {
    float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    float ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    float d1 = dot3f(ab, ap);
    float d2 = dot3f(ac, ap);
    float df[3];
    if (d1 <= 0.0f && d2 <= 0.0f) { /* vertex a */
        for (int k = 0; k < 3; k++) { df[k] = ap[k]; cp[k] = a[k]; }
        bary[0] = 1.0f; bary[1] = 0.0f; bary[2] = 0.0f;
        return dot3f(df, df);
    }
    float bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    float d3 = dot3f(ab, bp);
    float d4 = dot3f(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) { /* vertex b */
        for (int k = 0; k < 3; k++) { df[k] = bp[k]; cp[k] = b[k]; }
        bary[0] = 0.0f; bary[1] = 1.0f; bary[2] = 0.0f;
        return dot3f(df, df);
    }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) { /* edge ab */
        float v = d1 / (d1 - d3);
        for (int k = 0; k < 3; k++) {
            float off = v * ab[k];
            df[k] = ap[k] - off;
            cp[k] = a[k] + off;
        }
        bary[0] = 1.0f - v; bary[1] = v; bary[2] = 0.0f;
        return dot3f(df, df);
    }
    float cpv[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    float d5 = dot3f(ab, cpv);
    float d6 = dot3f(ac, cpv);
    if (d6 >= 0.0f && d5 <= d6) { /* vertex c */
        for (int k = 0; k < 3; k++) { df[k] = cpv[k]; cp[k] = c[k]; }
        bary[0] = 0.0f; bary[1] = 0.0f; bary[2] = 1.0f;
        return dot3f(df, df);
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) { /* edge ac */
        float w = d2 / (d2 - d6);
        for (int k = 0; k < 3; k++) {
            float off = w * ac[k];
            df[k] = ap[k] - off;
            cp[k] = a[k] + off;
        }
        bary[0] = 1.0f - w; bary[1] = 0.0f; bary[2] = w;
        return dot3f(df, df);
    }
    float va = d3 * d6 - d5 * d4;
    float e43 = d4 - d3, e56 = d5 - d6;
    if (va <= 0.0f && e43 >= 0.0f && e56 >= 0.0f) { /* edge bc */
        float w = e43 / (e43 + e56);
        for (int k = 0; k < 3; k++) {
            float off = w * (c[k] - b[k]);
            df[k] = bp[k] - off;
            cp[k] = b[k] + off;
        }
        bary[0] = 0.0f; bary[1] = 1.0f - w; bary[2] = w;
        return dot3f(df, df);
    }
    /* face interior */
    float denom = 1.0f / ((va + vb) + vc);
    float v = vb * denom;
    float w = vc * denom;
    for (int k = 0; k < 3; k++) {
        float off = ab[k] * v + ac[k] * w;
        df[k] = ap[k] - off;
        cp[k] = a[k] + off;
    }
    bary[0] = (1.0f - v) - w; bary[1] = v; bary[2] = w;
    return dot3f(df, df);
}

/* The reported closest point of a winning triangle: the same Ericson region tests in
 * FP64 on the FP32 inputs (exactly converted), every product and sum rounded to double
 * in the written order, the offset taken from the region's base vertex, then the point,
 * the barycentrics and sqrt(d2) each rounded once to FP32.  The FP32 search
 * (orc_cp_tri_f32) chooses the triangle; this only sharpens what is reported about it.
 * Returns the distance. */
static inline double dot3dd(const double* a, const double* b) {
    double s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}
ORC_API float orc_cp_tri_refine(const float* pf, const float* af, const float* bf, const float* cf, float* cp,
                                float* bary) {
    double p[3], a[3], b[3], c[3], ab[3], ac[3], ap[3], df[3], q[3];
    for (int k = 0; k < 3; k++) {
        p[k] = pf[k]; a[k] = af[k]; b[k] = bf[k]; c[k] = cf[k];
        ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; ap[k] = p[k] - a[k];
    }
    double d1 = dot3dd(ab, ap), d2 = dot3dd(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        for (int k = 0; k < 3; k++) cp[k] = af[k];
        bary[0] = 1.0f; bary[1] = 0.0f; bary[2] = 0.0f;
        return (float)sqrt(dot3dd(ap, ap));
    }
    double bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    double d3 = dot3dd(ab, bp), d4 = dot3dd(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
        for (int k = 0; k < 3; k++) cp[k] = bf[k];
        bary[0] = 0.0f; bary[1] = 1.0f; bary[2] = 0.0f;
        return (float)sqrt(dot3dd(bp, bp));
    }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double v = d1 / (d1 - d3);
        for (int k = 0; k < 3; k++) {
            double off = v * ab[k];
            df[k] = ap[k] - off;
            q[k] = a[k] + off;
            cp[k] = (float)q[k];
        }
        bary[0] = (float)(1.0 - v); bary[1] = (float)v; bary[2] = 0.0f;
        return (float)sqrt(dot3dd(df, df));
    }
    double cpv[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    double d5 = dot3dd(ab, cpv), d6 = dot3dd(ac, cpv);
    if (d6 >= 0.0 && d5 <= d6) {
        for (int k = 0; k < 3; k++) cp[k] = cf[k];
        bary[0] = 0.0f; bary[1] = 0.0f; bary[2] = 1.0f;
        return (float)sqrt(dot3dd(cpv, cpv));
    }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        for (int k = 0; k < 3; k++) {
            double off = w * ac[k];
            df[k] = ap[k] - off;
            q[k] = a[k] + off;
            cp[k] = (float)q[k];
        }
        bary[0] = (float)(1.0 - w); bary[1] = 0.0f; bary[2] = (float)w;
        return (float)sqrt(dot3dd(df, df));
    }
    double va = d3 * d6 - d5 * d4;
    double e43 = d4 - d3, e56 = d5 - d6;
    if (va <= 0.0 && e43 >= 0.0 && e56 >= 0.0) {
        double w = e43 / (e43 + e56);
        for (int k = 0; k < 3; k++) {
            double off = w * (c[k] - b[k]);
            df[k] = bp[k] - off;
            q[k] = b[k] + off;
            cp[k] = (float)q[k];
        }
        bary[0] = 0.0f; bary[1] = (float)(1.0 - w); bary[2] = (float)w;
        return (float)sqrt(dot3dd(df, df));
    }
    double denom = 1.0 / ((va + vb) + vc);
    double v = vb * denom, w = vc * denom;
    for (int k = 0; k < 3; k++) {
        double off = ab[k] * v + ac[k] * w;
        df[k] = ap[k] - off;
        q[k] = a[k] + off;
        cp[k] = (float)q[k];
    }
    bary[0] = (float)((1.0 - v) - w); bary[1] = (float)v; bary[2] = (float)w;
    return (float)sqrt(dot3dd(df, df));
}

/* cross(a,b) in the GLSL component order. */
static inline void cross3f(const float* a, const float* b, float* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* Moller-Trumbore in the order of rtcommon.glsl:181-202 (v1=pb-pa, v2=pc-pa, p=d x v2,
 * det=v1.p, s=o-pa, u=(s.p)*inv, q=s x v1, v=(d.q)*inv, t=(v2.q)*inv).  Returns 1 when
 * det != 0 and (u,v) is inside the closed triangle; the caller checks t. */
ORC_API int orc_mt_f32(const float* o, const float* d, const float* pa, const float* pb,
                       const float* pc, float* t_out, float* u_out, float* v_out) {
    float v1[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    float v2[3] = {pc[0] - pa[0], pc[1] - pa[1], pc[2] - pa[2]};
    float p[3];
    cross3f(d, v2, p);
    float det = dot3f(v1, p);
    if (det == 0.0f) return 0;
    float inv = 1.0f / det;
    float s[3] = {o[0] - pa[0], o[1] - pa[1], o[2] - pa[2]};
    float u = dot3f(s, p) * inv;
    if (u < 0.0f || u > 1.0f) return 0;
    float q[3];
    cross3f(s, v1, q);
    float v = dot3f(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return 0;
    *t_out = dot3f(v2, q) * inv;
    *u_out = u;
    *v_out = v;
    return 1;
}

/* Inputs the library defines as unanswerable (include/gpurt.h): a ray with a non-finite
 * origin or direction component, a zero direction or a NaN bound is a miss; a query point
 * with a non-finite coordinate finds no triangle (prim -1, distance +inf, point and
 * barycentrics 0).  The reference leaves such rays undefined (traceRayEXT requires finite
 * rays); the library and this restatement both answer them this way. */
ORC_API int orc_ray_ok(const float* r) {
    return isfinite(r[0]) && isfinite(r[1]) && isfinite(r[2]) && isfinite(r[4]) && isfinite(r[5]) &&
           isfinite(r[6]) && (r[4] != 0.0f || r[5] != 0.0f || r[6] != 0.0f) && !isnan(r[3]) && !isnan(r[7]);
}
ORC_API int orc_point_ok(const float* q) { return isfinite(q[0]) && isfinite(q[1]) && isfinite(q[2]); }

/* ------------------------------------------------------------------------------------ */
/* FP32 brute force                                                                     */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    const float* V;
    const uint32_t* F;
    int64_t nt;
    const float* Q;
    int32_t* prim;
    float* pt;
    float* dist;
    float* bary;
} cpq_bf32_ctx;

static void cpq_bf32_range(void* vctx, int64_t lo, int64_t hi) {
    cpq_bf32_ctx* c = (cpq_bf32_ctx*)vctx;
    for (int64_t i = lo; i < hi; i++) {
        const float* q = c->Q + 3 * i;
        float best = INFINITY, bcp[3] = {0, 0, 0}, bb[3] = {0, 0, 0}, bd = INFINITY;
        int64_t bi = -1;
        for (int64_t t = 0; t < c->nt && orc_point_ok(q); t++) {
            const uint32_t* f = c->F + 3 * t;
            float cp[3], br[3];
            float d2 = orc_cp_tri_f32(q, c->V + 3 * f[0], c->V + 3 * f[1], c->V + 3 * f[2], cp, br);
            if (d2 < best) { /* strict: ties keep the lower index (scan is ascending) */
                best = d2; bi = t;
            }
        }
        if (bi >= 0) {
            const uint32_t* f = c->F + 3 * bi;
            bd = orc_cp_tri_refine(q, c->V + 3 * f[0], c->V + 3 * f[1], c->V + 3 * f[2], bcp, bb);
        }
        c->prim[i] = (int32_t)bi;
        if (c->pt) memcpy(c->pt + 3 * i, bcp, sizeof bcp);
        if (c->dist) c->dist[i] = bd;
        if (c->bary) memcpy(c->bary + 3 * i, bb, sizeof bb);
    }
}

ORC_API void orc_cpq_brute_f32(const float* V, const uint32_t* F, int64_t nt, const float* Q,
                               int64_t nq, int32_t* prim, float* pt, float* dist, float* bary) {
    cpq_bf32_ctx c = {V, F, nt, Q, prim, pt, dist, bary};
    parallel_for(nq, 16, cpq_bf32_range, &c);
}

typedef struct {
    const float* V;
    const uint32_t* F;
    int64_t nt;
    const float* R;
    int32_t* prim;
    float* t;
    float* uv;
    uint8_t* occ;
} ray_bf32_ctx;

static void ray_closest_bf32_range(void* vctx, int64_t lo, int64_t hi) {
    ray_bf32_ctx* c = (ray_bf32_ctx*)vctx;
    for (int64_t i = lo; i < hi; i++) {
        const float* r = c->R + 8 * i;
        float tmin = r[3], best = r[7], bu = 0, bv = 0;
        int64_t bi = -1;
        for (int64_t k = 0; k < c->nt && orc_ray_ok(r); k++) {
            const uint32_t* f = c->F + 3 * k;
            float t, u, v;
            if (!orc_mt_f32(r, r + 4, c->V + 3 * f[0], c->V + 3 * f[1], c->V + 3 * f[2], &t, &u, &v))
                continue;
            if (!(t >= tmin)) continue;
            if (t < best || (t == best && bi < 0)) { best = t; bi = k; bu = u; bv = v; }
        }
        c->prim[i] = (int32_t)bi;
        c->t[i] = bi >= 0 ? best : INFINITY;
        if (c->uv) { c->uv[2 * i] = bi >= 0 ? bu : 0.0f; c->uv[2 * i + 1] = bi >= 0 ? bv : 0.0f; }
    }
}

static void ray_any_bf32_range(void* vctx, int64_t lo, int64_t hi) {
    ray_bf32_ctx* c = (ray_bf32_ctx*)vctx;
    for (int64_t i = lo; i < hi; i++) {
        const float* r = c->R + 8 * i;
        uint8_t hit = 0;
        for (int64_t k = 0; k < c->nt && !hit && orc_ray_ok(r); k++) {
            const uint32_t* f = c->F + 3 * k;
            float t, u, v;
            if (!orc_mt_f32(r, r + 4, c->V + 3 * f[0], c->V + 3 * f[1], c->V + 3 * f[2], &t, &u, &v))
                continue;
            if (t >= r[3] && t <= r[7]) hit = 1;
        }
        c->occ[i] = hit;
    }
}

/* rays: 8 floats each = origin.xyz, tmin, dir.xyz, tmax.  Miss: prim -1, t +inf, uv 0. */
ORC_API void orc_ray_closest_brute_f32(const float* V, const uint32_t* F, int64_t nt,
                                       const float* R, int64_t nr, int32_t* prim, float* t,
                                       float* uv) {
    ray_bf32_ctx c = {V, F, nt, R, prim, t, uv, NULL};
    parallel_for(nr, 16, ray_closest_bf32_range, &c);
}

ORC_API void orc_ray_any_brute_f32(const float* V, const uint32_t* F, int64_t nt, const float* R,
                                   int64_t nr, uint8_t* occ) {
    ray_bf32_ctx c = {V, F, nt, R, NULL, NULL, NULL, occ};
    parallel_for(nr, 16, ray_any_bf32_range, &c);
}

/* ------------------------------------------------------------------------------------ */
/* two-level instancing                                                                  */
/* ------------------------------------------------------------------------------------ */

/* FP32 inverse of the 3x3 part of a row-major 3x4 object-to-world transform: cofactors
 * over the determinant in FP64, then rounded.  Returns 0 if singular or not finite. */
ORC_API int orc_invert3(const float* t, float* minv) {
    const double m00 = t[0], m01 = t[1], m02 = t[2];
    const double m10 = t[4], m11 = t[5], m12 = t[6];
    const double m20 = t[8], m21 = t[9], m22 = t[10];
    const double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
    if (!(det != 0.0) || !isfinite(det)) return 0;
    const double inv[9] = {(m11 * m22 - m12 * m21) / det, (m02 * m21 - m01 * m22) / det, (m01 * m12 - m02 * m11) / det,
                           (m12 * m20 - m10 * m22) / det, (m00 * m22 - m02 * m20) / det, (m02 * m10 - m00 * m12) / det,
                           (m10 * m21 - m11 * m20) / det, (m01 * m20 - m00 * m21) / det, (m00 * m11 - m01 * m10) / det};
    for (int k = 0; k < 9; k++) {
        if (!isfinite(inv[k])) return 0;
        minv[k] = (float)inv[k];
    }
    return 1;
}

/* object-space ray: x = o - c, o' = rows of minv . x, d' = rows of minv . d, each row
 * summed left to right with every product rounded */
static inline void to_object_f32(const float* minv, const float* t, const float* o, const float* d, float* oo,
                                 float* od) {
    const float x[3] = {o[0] - t[3], o[1] - t[7], o[2] - t[11]};
    for (int r = 0; r < 3; r++) {
        float a = minv[3 * r] * x[0];
        a = a + minv[3 * r + 1] * x[1];
        a = a + minv[3 * r + 2] * x[2];
        oo[r] = a;
        float b = minv[3 * r] * d[0];
        b = b + minv[3 * r + 1] * d[1];
        b = b + minv[3 * r + 2] * d[2];
        od[r] = b;
    }
}

typedef struct {
    int64_t n_inst;
    const float* xf;       /* 12 per instance */
    const float* minv;     /* 9 per instance */
    const int32_t* mesh;   /* mesh of each instance */
    const uint32_t* custom;
    const uint32_t* mask;
    const float* V;        /* all meshes' vertices */
    const uint32_t* F;     /* all meshes' triangles, global vertex indices */
    const int64_t* off;    /* triangles of mesh m: [off[m], off[m+1]) */
    const float* R;
    int32_t* obj;
    int32_t* prim;
    float* t;
    float* uv;
    uint8_t* occ;
} inst_ctx;

static void inst_ray_range(void* vctx, int64_t lo, int64_t hi) {
    inst_ctx* c = (inst_ctx*)vctx;
    for (int64_t i = lo; i < hi; i++) {
        const float* r = c->R + 8 * i;
        float tmin = r[3], best = r[7], bu = 0, bv = 0;
        int64_t bi = -1, bk = -1;
        uint8_t hit = 0;
        for (int64_t n = 0; n < c->n_inst && !hit && orc_ray_ok(r); n++) {
            if ((c->mask[n] & 0xFFu) == 0) continue;
            float oo[3], od[3];
            to_object_f32(c->minv + 9 * n, c->xf + 12 * n, r, r + 4, oo, od);
            const int32_t m = c->mesh[n];
            for (int64_t k = c->off[m]; k < c->off[m + 1]; k++) {
                const uint32_t* f = c->F + 3 * k;
                float t, u, v;
                if (!orc_mt_f32(oo, od, c->V + 3 * f[0], c->V + 3 * f[1], c->V + 3 * f[2], &t, &u, &v)) continue;
                if (!(t >= tmin)) continue;
                if (c->occ) {
                    if (t <= r[7]) { hit = 1; break; }
                    continue;
                }
                if (t < best || (t == best && bi < 0)) { best = t; bi = n; bk = k - c->off[m]; bu = u; bv = v; }
            }
        }
        if (c->occ) { c->occ[i] = hit; continue; }
        c->obj[i] = bi >= 0 ? (int32_t)c->custom[bi] : -1;
        c->prim[i] = bi >= 0 ? (int32_t)bk : -1;
        c->t[i] = bi >= 0 ? best : INFINITY;
        if (c->uv) { c->uv[2 * i] = bi >= 0 ? bu : 0.0f; c->uv[2 * i + 1] = bi >= 0 ? bv : 0.0f; }
    }
}

/* Brute force over every instance and every triangle of its mesh.  Closest hit
 * (occ == NULL): minimum over (t, instance, triangle), obj = custom index, prim =
 * triangle index within the instance's mesh, -1 / +inf on a miss; any hit (occ != NULL):
 * some visible instance has a hit with t in [tmin, tmax].  Instances with mask & 0xFF
 * == 0 are invisible.  Returns 0 if a transform is singular. */
ORC_API int orc_inst_ray_brute_f32(int64_t n_inst, const float* xf, const int32_t* mesh, const uint32_t* custom,
                                   const uint32_t* mask, const float* V, const uint32_t* F, const int64_t* off,
                                   const float* R, int64_t nr, int32_t* obj, int32_t* prim, float* t, float* uv,
                                   uint8_t* occ) {
    float* minv = (float*)malloc(sizeof(float) * 9 * (size_t)(n_inst > 0 ? n_inst : 1));
    if (!minv) return 0;
    for (int64_t n = 0; n < n_inst; n++)
        if (!orc_invert3(xf + 12 * n, minv + 9 * n)) { free(minv); return 0; }
    inst_ctx c = {n_inst, xf, minv, mesh, custom, mask, V, F, off, R, obj, prim, t, uv, occ};
    parallel_for(nr, 4, inst_ray_range, &c);
    free(minv);
    return 1;
}

/* ------------------------------------------------------------------------------------ */
/* FP64 independent oracle                                                              */
/* ------------------------------------------------------------------------------------ */

static inline double dot3d(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3d(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* Closest point on segment [a,b] to p: parameter clamped to [0,1]. */
static double seg_cp_d(const double* p, const double* a, const double* b, double* s_out,
                       double* q) {
    double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double l2 = dot3d(ab, ab);
    double s = l2 > 0 ? dot3d(ap, ab) / l2 : 0.0;
    if (s < 0) s = 0;
    if (s > 1) s = 1;
    for (int k = 0; k < 3; k++) q[k] = a[k] + s * ab[k];
    double df[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
    *s_out = s;
    return dot3d(df, df);
}

/* Closest point on triangle, FP64, by a formulation independent of Ericson's region
 * tests: project onto the plane and solve the 2x2 normal equations for barycentrics;
 * if all are >= 0 that is the answer, else the answer is on the boundary, i.e. the best
 * of the three clamped edge projections (vertices are edge end points). */
static double cp_tri_f64(const double* p, const double* a, const double* b, const double* c,
                         double* cp, double* bary) {
    double e0[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double e1[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double a00 = dot3d(e0, e0), a01 = dot3d(e0, e1), a11 = dot3d(e1, e1);
    double r0 = dot3d(ap, e0), r1 = dot3d(ap, e1);
    double det = a00 * a11 - a01 * a01;
    if (det > 1e-300 * (a00 * a11 + 1e-300)) {
        double v = (a11 * r0 - a01 * r1) / det;
        double w = (a00 * r1 - a01 * r0) / det;
        if (v >= 0 && w >= 0 && v + w <= 1) {
            double q[3];
            for (int k = 0; k < 3; k++) q[k] = a[k] + v * e0[k] + w * e1[k];
            double df[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
            memcpy(cp, q, sizeof q);
            bary[0] = 1 - v - w; bary[1] = v; bary[2] = w;
            return dot3d(df, df);
        }
    }
    double qa[3], qb[3], qc[3], sa, sb, sc;
    double da = seg_cp_d(p, a, b, &sa, qa); /* edge ab */
    double db = seg_cp_d(p, b, c, &sb, qb); /* edge bc */
    double dc = seg_cp_d(p, c, a, &sc, qc); /* edge ca */
    if (da <= db && da <= dc) {
        memcpy(cp, qa, sizeof qa);
        bary[0] = 1 - sa; bary[1] = sa; bary[2] = 0;
        return da;
    }
    if (db <= dc) {
        memcpy(cp, qb, sizeof qb);
        bary[0] = 0; bary[1] = 1 - sb; bary[2] = sb;
        return db;
    }
    memcpy(cp, qc, sizeof qc);
    bary[0] = sc; bary[1] = 0; bary[2] = 1 - sc;
    return dc;
}

ORC_API double orc_cp_tri_f64(const double* p, const double* a, const double* b, const double* c,
                              double* cp, double* bary) {
    return cp_tri_f64(p, a, b, c, cp, bary);
}

typedef struct {
    const double* V;
    const uint32_t* F;
    int64_t nt;
    const double* Q;
    int32_t* prim;
    double* pt;
    double* dist;
    double* bary;
} cpq_bf64_ctx;

static void cpq_bf64_range(void* vctx, int64_t lo, int64_t hi) {
    cpq_bf64_ctx* c = (cpq_bf64_ctx*)vctx;
    for (int64_t i = lo; i < hi; i++) {
        const double* q = c->Q + 3 * i;
        double best = INFINITY, bcp[3] = {0, 0, 0}, bb[3] = {0, 0, 0};
        int64_t bi = -1;
        for (int64_t t = 0; t < c->nt; t++) {
            const uint32_t* f = c->F + 3 * t;
            double cp[3], br[3];
            double d2 = cp_tri_f64(q, c->V + 3 * f[0], c->V + 3 * f[1], c->V + 3 * f[2], cp, br);
            if (d2 < best) {
                best = d2; bi = t;
                memcpy(bcp, cp, sizeof cp); memcpy(bb, br, sizeof br);
            }
        }
        c->prim[i] = (int32_t)bi;
        if (c->pt) memcpy(c->pt + 3 * i, bcp, sizeof bcp);
        if (c->dist) c->dist[i] = bi >= 0 ? sqrt(best) : INFINITY;
        if (c->bary) memcpy(c->bary + 3 * i, bb, sizeof bb);
    }
}

ORC_API void orc_cpq_brute_f64(const double* V, const uint32_t* F, int64_t nt, const double* Q,
                               int64_t nq, int32_t* prim, double* pt, double* dist, double* bary) {
    cpq_bf64_ctx c = {V, F, nt, Q, prim, pt, dist, bary};
    parallel_for(nq, 4, cpq_bf64_range, &c);
}

/* Ray/triangle, FP64, independent of Moller-Trumbore: intersect the supporting plane,
 * then test the hit point with the three signed edge functions (sub-triangle areas
 * projected on the normal).  Barycentrics (u,v) weight (pb,pc) like rt.rchit:13. */
static int ray_tri_f64(const double* o, const double* d, const double* a, const double* b,
                       const double* c, double* t_out, double* u_out, double* v_out) {
    double e0[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double e1[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n[3];
    cross3d(e0, e1, n);
    double nd = dot3d(n, d);
    double nn = dot3d(n, n);
    if (nd == 0.0 || nn == 0.0) return 0;
    double ao[3] = {a[0] - o[0], a[1] - o[1], a[2] - o[2]};
    double t = dot3d(n, ao) / nd;
    double h[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
    double ha[3] = {a[0] - h[0], a[1] - h[1], a[2] - h[2]};
    double hb[3] = {b[0] - h[0], b[1] - h[1], b[2] - h[2]};
    double hc[3] = {c[0] - h[0], c[1] - h[1], c[2] - h[2]};
    double x[3];
    cross3d(hb, hc, x); double wa = dot3d(x, n) / nn;
    cross3d(hc, ha, x); double wb = dot3d(x, n) / nn;
    cross3d(ha, hb, x); double wc = dot3d(x, n) / nn;
    if (wa < 0 || wb < 0 || wc < 0) return 0;
    *t_out = t; *u_out = wb; *v_out = wc;
    return 1;
}

ORC_API int orc_ray_tri_f64(const double* o, const double* d, const double* a, const double* b,
                            const double* c, double* t, double* u, double* v) {
    return ray_tri_f64(o, d, a, b, c, t, u, v);
}

typedef struct {
    const double* V;
    const uint32_t* F;
    int64_t nt;
    const double* R;
    int32_t* prim;
    double* t;
    double* uv;
} ray_bf64_ctx;

static void ray_closest_bf64_range(void* vctx, int64_t lo, int64_t hi) {
    ray_bf64_ctx* c = (ray_bf64_ctx*)vctx;
    for (int64_t i = lo; i < hi; i++) {
        const double* r = c->R + 8 * i;
        double best = r[7], bu = 0, bv = 0;
        int64_t bi = -1;
        for (int64_t k = 0; k < c->nt; k++) {
            const uint32_t* f = c->F + 3 * k;
            double t, u, v;
            if (!ray_tri_f64(r, r + 4, c->V + 3 * f[0], c->V + 3 * f[1], c->V + 3 * f[2], &t, &u, &v))
                continue;
            if (!(t >= r[3])) continue;
            if (t < best || (t == best && bi < 0)) { best = t; bi = k; bu = u; bv = v; }
        }
        c->prim[i] = (int32_t)bi;
        c->t[i] = bi >= 0 ? best : INFINITY;
        if (c->uv) { c->uv[2 * i] = bu; c->uv[2 * i + 1] = bv; }
    }
}

ORC_API void orc_ray_closest_brute_f64(const double* V, const uint32_t* F, int64_t nt,
                                       const double* R, int64_t nr, int32_t* prim, double* t,
                                       double* uv) {
    ray_bf64_ctx c = {V, F, nt, R, prim, t, uv};
    parallel_for(nr, 4, ray_closest_bf64_range, &c);
}

/* ------------------------------------------------------------------------------------ */
/* FP64 closed-form gradients                                                           */
/* ------------------------------------------------------------------------------------ */

/* Unsigned distance d(q) = |q - p(q)| with p the closest point of triangle `prim`.
 * By the envelope theorem (the minimiser over the convex triangle is unique):
 *   dd/dq = (q - p)/d,   dd/dv_i = -b_i (q - p)/d.
 * g_q[i] = g[i] * dd/dq ; g_V[vertex] += g[i] * dd/dv_i.  d == 0 gives zero gradient.
 * With pt == NULL the closest point and barycentrics of `prim` are computed here by the
 * independent FP64 formulation (cp_tri_f64) instead of being taken from the caller. */
ORC_API void orc_cpq_grad_f64(const double* V, const uint32_t* F, const double* Q, int64_t nq,
                              const int32_t* prim, const double* pt, const double* bary,
                              const double* g, double* g_q, double* g_V) {
    for (int64_t i = 0; i < nq; i++) {
        double gq[3] = {0, 0, 0};
        if (prim[i] >= 0) {
            const double* q = Q + 3 * i;
            double own_p[3], own_b[3];
            const double* p = pt ? pt + 3 * i : own_p;
            const double* bi = pt ? bary + 3 * i : own_b;
            if (!pt) {
                const uint32_t* f = F + 3 * (int64_t)prim[i];
                (void)cp_tri_f64(q, V + 3 * f[0], V + 3 * f[1], V + 3 * f[2], own_p, own_b);
            }
            double df[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
            double d = sqrt(dot3d(df, df));
            if (d > 0) {
                double gi = g ? g[i] : 1.0;
                for (int k = 0; k < 3; k++) gq[k] = gi * df[k] / d;
                if (g_V) {
                    const uint32_t* f = F + 3 * (int64_t)prim[i];
                    for (int j = 0; j < 3; j++)
                        for (int k = 0; k < 3; k++) g_V[3 * (int64_t)f[j] + k] -= bi[j] * gq[k];
                }
            }
        }
        if (g_q) memcpy(g_q + 3 * i, gq, sizeof gq);
    }
}

/* Ray hit distance t = n.(v0 - o)/(n.d), n = (v1-v0) x (v2-v0).  Implicit differentiation
 * with the hit's barycentrics b = (1-u-v, u, v) held fixed gives
 *   dt/do = -n/(n.d),  dt/dd = -t n/(n.d),  dt/dv_i = b_i n/(n.d).  Misses get zero. */
ORC_API void orc_ray_grad_f64(const double* V, const uint32_t* F, const double* R, int64_t nr,
                              const int32_t* prim, const double* t, const double* uv,
                              const double* g, double* g_o, double* g_d, double* g_V) {
    for (int64_t i = 0; i < nr; i++) {
        double go[3] = {0, 0, 0}, gd[3] = {0, 0, 0};
        if (prim[i] >= 0) {
            const uint32_t* f = F + 3 * (int64_t)prim[i];
            const double *a = V + 3 * (int64_t)f[0], *b = V + 3 * (int64_t)f[1], *c = V + 3 * (int64_t)f[2];
            double e0[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            double e1[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            double n[3];
            cross3d(e0, e1, n);
            double nd = dot3d(n, R + 8 * i + 4);
            if (nd != 0.0) {
                double gi = g ? g[i] : 1.0;
                for (int k = 0; k < 3; k++) {
                    go[k] = -gi * n[k] / nd;
                    gd[k] = -gi * t[i] * n[k] / nd;
                }
                if (g_V) {
                    double u = uv[2 * i], v = uv[2 * i + 1];
                    double bw[3] = {1 - u - v, u, v};
                    for (int j = 0; j < 3; j++)
                        for (int k = 0; k < 3; k++) g_V[3 * (int64_t)f[j] + k] += bw[j] * gi * n[k] / nd;
                }
            }
        }
        if (g_o) memcpy(g_o + 3 * i, go, sizeof go);
        if (g_d) memcpy(g_d + 3 * i, gd, sizeof gd);
    }
}

/* ------------------------------------------------------------------------------------ */
/* CPU LBVH (Karras 2012) with leaves of <= leaf_size primitives                        */
/* ------------------------------------------------------------------------------------ */

/* Spread the low 10 bits of x to every third bit (bit k -> bit 3k). */
static inline uint32_t expand_bits10(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

/* 10-bit quantisation of c over [lo,hi] per axis: floor((c-lo)/(hi-lo)*1024) clamped to
 * [0,1023]; code = x<<2 | y<<1 | z interleave (x in the highest bit of each triple). */
static inline uint32_t morton30(const float* c, const float* lo, const float* inv_ext) {
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        float f = (c[k] - lo[k]) * inv_ext[k] * 1024.0f;
        int v = (int)floorf(f);
        if (!(f >= 0.0f)) v = 0; /* also catches NaN */
        if (v > 1023) v = 1023;
        q[k] = (uint32_t)v;
    }
    return (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
}

ORC_API void orc_morton_codes(const float* centroids, int64_t n, const float* lo, const float* hi,
                              uint32_t* out) {
    float inv[3];
    for (int k = 0; k < 3; k++) inv[k] = hi[k] > lo[k] ? 1.0f / (hi[k] - lo[k]) : 0.0f;
    for (int64_t i = 0; i < n; i++) out[i] = morton30(centroids + 3 * i, lo, inv);
}

typedef struct {
    float lo[3], hi[3]; /* padded (inflated) bounds */
    int32_t left, right; /* children, or -1 for a leaf */
    int32_t first, count; /* leaf: primitive range in sorted order */
} onode;

typedef struct {
    int64_t nt, n_nodes, depth, max_leaf;
    int leaf_size;
    int bad; /* set when a Karras range disagrees with the recursion (must stay 0) */
    float pad;
    onode* nodes;
    float* tris;   /* 9 floats per sorted primitive */
    int32_t* ids;  /* sorted position -> original triangle id */
} obvh;

typedef struct {
    uint32_t code;
    uint32_t idx;
} kv32;

static int kv_cmp(const void* a, const void* b) {
    const kv32* x = (const kv32*)a;
    const kv32* y = (const kv32*)b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx ? 1 : 0);
}

/* Karras' delta: length of the common prefix of keys i and j, the sorted index breaking
 * ties between equal codes; -1 outside [0,n). */
static inline int kdelta(const kv32* k, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i].code, b = k[j].code;
    if (a == b) return 32 + __builtin_clz((uint32_t)(i ^ j)); /* i != j here */
    return __builtin_clz(a ^ b);
}

typedef struct {
    int64_t first, last, split;
} krange;

static krange karras_node(const kv32* k, int64_t n, int64_t i) {
    int d = kdelta(k, n, i, i + 1) - kdelta(k, n, i, i - 1) >= 0 ? 1 : -1;
    int dmin = kdelta(k, n, i, i - d);
    int64_t lmax = 2;
    while (kdelta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (kdelta(k, n, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = kdelta(k, n, i, j);
    int64_t s = 0, t = l;
    do {
        t = (t + 1) / 2;
        if (kdelta(k, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    krange r;
    r.split = i + s * d + (d < 0 ? -1 : 0);
    r.first = i < j ? i : j;
    r.last = i < j ? j : i;
    return r;
}

typedef struct {
    obvh* b;
    krange* inner; /* per Karras internal node */
} bctx;

/* Recursively emit nodes; returns the node index. Binary Karras nodes with range
 * <= leaf_size (and single primitives) become leaves. */
static int32_t emit(bctx* c, int64_t first, int64_t last, int64_t karras_idx, int depth) {
    obvh* b = c->b;
    int32_t me = (int32_t)b->n_nodes++;
    onode* nd = &b->nodes[me];
    if (depth + 1 > b->depth) b->depth = depth + 1;
    int64_t cnt = last - first + 1;
    if (cnt <= b->leaf_size) {
        nd->left = nd->right = -1;
        nd->first = (int32_t)first;
        nd->count = (int32_t)cnt;
        if (cnt > b->max_leaf) b->max_leaf = cnt;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t p = first; p <= last; p++)
            for (int v = 0; v < 3; v++)
                for (int k = 0; k < 3; k++) {
                    float x = b->tris[9 * p + 3 * v + k];
                    if (x < lo[k]) lo[k] = x;
                    if (x > hi[k]) hi[k] = x;
                }
        for (int k = 0; k < 3; k++) { nd->lo[k] = lo[k]; nd->hi[k] = hi[k]; }
        return me;
    }
    krange r = c->inner[karras_idx];
    if (r.first != first || r.last != last) b->bad = 1;
    int64_t sp = r.split;
    /* left child covers [first, sp], right child [sp+1, last]; a child that is an
     * internal Karras node is the one whose index is its range end nearest the split */
    int32_t L = emit(c, first, sp, sp, depth + 1);
    int32_t R = emit(c, sp + 1, last, sp + 1, depth + 1);
    nd = &b->nodes[me];
    nd->left = L;
    nd->right = R;
    nd->first = (int32_t)first;
    nd->count = (int32_t)cnt;
    for (int k = 0; k < 3; k++) {
        float a = b->nodes[L].lo[k], bb = b->nodes[R].lo[k];
        nd->lo[k] = a < bb ? a : bb;
        a = b->nodes[L].hi[k]; bb = b->nodes[R].hi[k];
        nd->hi[k] = a > bb ? a : bb;
    }
    return me;
}

ORC_API void* orc_lbvh_build(const float* V, int64_t nv, const uint32_t* F, int64_t nt,
                             int leaf_size) {
    if (nt <= 0 || leaf_size < 1) return NULL;
    for (int64_t t = 0; t < 3 * nt; t++)
        if (F[t] >= (uint64_t)nv) return NULL;
    obvh* b = (obvh*)calloc(1, sizeof(obvh));
    b->nt = nt;
    b->leaf_size = leaf_size;
    float* cen = (float*)malloc(sizeof(float) * 3 * nt);
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float scale = 0.0f;
    for (int64_t t = 0; t < nt; t++) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int v = 0; v < 3; v++)
            for (int k = 0; k < 3; k++) {
                float x = V[3 * (int64_t)F[3 * t + v] + k];
                if (x < lo[k]) lo[k] = x;
                if (x > hi[k]) hi[k] = x;
                float ax = fabsf(x);
                if (ax > scale) scale = ax;
            }
        for (int k = 0; k < 3; k++) {
            float c = (lo[k] + hi[k]) * 0.5f;
            cen[3 * t + k] = c;
            if (c < clo[k]) clo[k] = c;
            if (c > chi[k]) chi[k] = c;
        }
    }
    /* boxes are inflated by pad = 2^-18 * max|coordinate| so that culls stay
     * conservative under FP32 rounding of the box and triangle arithmetic */
    b->pad = scale * 0x1p-18f;
    float inv[3];
    for (int k = 0; k < 3; k++) inv[k] = chi[k] > clo[k] ? 1.0f / (chi[k] - clo[k]) : 0.0f;
    kv32* keys = (kv32*)malloc(sizeof(kv32) * nt);
    for (int64_t t = 0; t < nt; t++) {
        keys[t].code = morton30(cen + 3 * t, clo, inv);
        keys[t].idx = (uint32_t)t;
    }
    free(cen);
    qsort(keys, (size_t)nt, sizeof(kv32), kv_cmp);
    b->tris = (float*)malloc(sizeof(float) * 9 * nt);
    b->ids = (int32_t*)malloc(sizeof(int32_t) * nt);
    for (int64_t p = 0; p < nt; p++) {
        int64_t t = keys[p].idx;
        b->ids[p] = (int32_t)t;
        for (int v = 0; v < 3; v++)
            for (int k = 0; k < 3; k++) b->tris[9 * p + 3 * v + k] = V[3 * (int64_t)F[3 * t + v] + k];
    }
    bctx c;
    c.b = b;
    c.inner = (krange*)malloc(sizeof(krange) * (nt > 1 ? nt - 1 : 1));
    for (int64_t i = 0; i + 1 < nt; i++) c.inner[i] = karras_node(keys, nt, i);
    b->nodes = (onode*)malloc(sizeof(onode) * (2 * nt));
    b->n_nodes = 0;
    emit(&c, 0, nt - 1, 0, 0);
    for (int64_t i = 0; i < b->n_nodes; i++)
        for (int k = 0; k < 3; k++) {
            b->nodes[i].lo[k] -= b->pad;
            b->nodes[i].hi[k] += b->pad;
        }
    free(c.inner);
    free(keys);
    return b;
}

ORC_API void orc_lbvh_free(void* h) {
    obvh* b = (obvh*)h;
    if (!b) return;
    free(b->nodes);
    free(b->tris);
    free(b->ids);
    free(b);
}

/* out: n_nodes, depth, max_leaf, n_leaves, bad */
ORC_API void orc_lbvh_info(void* h, int64_t* out) {
    obvh* b = (obvh*)h;
    int64_t leaves = 0;
    for (int64_t i = 0; i < b->n_nodes; i++) leaves += b->nodes[i].left < 0;
    out[0] = b->n_nodes;
    out[1] = b->depth;
    out[2] = b->max_leaf;
    out[3] = leaves;
    out[4] = b->bad;
}

/* sorted position -> original triangle id, for tests of the Morton order */
ORC_API void orc_lbvh_order(void* h, int32_t* out) {
    obvh* b = (obvh*)h;
    memcpy(out, b->ids, sizeof(int32_t) * b->nt);
}

static inline float box_d2(const onode* n, const float* q) {
    float s = 0.0f;
    for (int k = 0; k < 3; k++) {
        float d = 0.0f;
        if (q[k] < n->lo[k]) d = n->lo[k] - q[k];
        else if (q[k] > n->hi[k]) d = q[k] - n->hi[k];
        s += d * d;
    }
    return s;
}

#define ORC_STACK 256

typedef struct {
    obvh* b;
    const float* Q;
    int32_t* prim;
    float* pt;
    float* dist;
    float* bary;
    int64_t* counters;
    const float* R;
    float* t;
    float* uv;
    uint8_t* occ;
} qctx;

static void lbvh_cpq_range(void* vctx, int64_t lo, int64_t hi) {
    qctx* c = (qctx*)vctx;
    obvh* b = c->b;
    int64_t nodes_visited = 0, tris_tested = 0;
    int32_t st_n[ORC_STACK];
    float st_d[ORC_STACK];
    for (int64_t i = lo; i < hi; i++) {
        const float* q = c->Q + 3 * i;
        float best = INFINITY, bcp[3] = {0, 0, 0}, bb[3] = {0, 0, 0}, bd = INFINITY;
        const float* btr = NULL;
        int32_t bi = -1;
        int sp = 0;
        if (orc_point_ok(q)) { st_n[sp] = 0; st_d[sp] = 0.0f; sp++; }
        while (sp > 0) {
            sp--;
            int32_t ni = st_n[sp];
            /* prune when the box bound exceeds the best distance by a relative margin */
            if (st_d[sp] > best * (1.0f + 0x1p-16f)) continue;
            const onode* nd = &b->nodes[ni];
            nodes_visited++;
            if (nd->left < 0) {
                for (int32_t p = nd->first; p < nd->first + nd->count; p++) {
                    const float* tr = b->tris + 9 * (int64_t)p;
                    float cp[3], br[3];
                    float d2 = orc_cp_tri_f32(q, tr, tr + 3, tr + 6, cp, br);
                    int32_t id = b->ids[p];
                    tris_tested++;
                    if (d2 < best || (d2 == best && id < bi)) {
                        best = d2; bi = id; btr = tr;
                    }
                }
                continue;
            }
            float dl = box_d2(&b->nodes[nd->left], q);
            float dr = box_d2(&b->nodes[nd->right], q);
            int32_t nl = nd->left, nr = nd->right;
            if (dl > dr) { float tf = dl; dl = dr; dr = tf; int32_t ti = nl; nl = nr; nr = ti; }
            float thr = best * (1.0f + 0x1p-16f);
            if (dr <= thr && sp < ORC_STACK) { st_n[sp] = nr; st_d[sp] = dr; sp++; }
            if (dl <= thr && sp < ORC_STACK) { st_n[sp] = nl; st_d[sp] = dl; sp++; }
        }
        if (btr) bd = orc_cp_tri_refine(q, btr, btr + 3, btr + 6, bcp, bb);
        c->prim[i] = bi;
        if (c->pt) memcpy(c->pt + 3 * i, bcp, sizeof bcp);
        if (c->dist) c->dist[i] = bd;
        if (c->bary) memcpy(c->bary + 3 * i, bb, sizeof bb);
    }
    if (c->counters) {
        __atomic_fetch_add(&c->counters[0], nodes_visited, __ATOMIC_RELAXED);
        __atomic_fetch_add(&c->counters[1], tris_tested, __ATOMIC_RELAXED);
    }
}

ORC_API void orc_lbvh_cpq(void* h, const float* Q, int64_t nq, int32_t* prim, float* pt,
                          float* dist, float* bary, int64_t* counters) {
    qctx c;
    memset(&c, 0, sizeof c);
    c.b = (obvh*)h; c.Q = Q; c.prim = prim; c.pt = pt; c.dist = dist; c.bary = bary;
    c.counters = counters;
    parallel_for(nq, 256, lbvh_cpq_range, &c);
}

/* conservative slab test (boxes are already inflated by pad): returns entry distance or
 * +inf when the box misses [tmin, tcur] */
static inline float box_ray(const onode* n, const float* o, const float* inv, float tmin,
                            float tcur) {
    float tn = tmin, tf = tcur;
    for (int k = 0; k < 3; k++) {
        float t0 = (n->lo[k] - o[k]) * inv[k];
        float t1 = (n->hi[k] - o[k]) * inv[k];
        float a = t0 < t1 ? t0 : t1, bb = t0 < t1 ? t1 : t0;
        if (a > tn) tn = a;
        if (bb < tf) tf = bb;
    }
    return tn <= tf + fabsf(tf) * 0x1p-20f ? tn : INFINITY;
}

static inline void ray_inv(const float* d, float* inv) {
    for (int k = 0; k < 3; k++) {
        float x = d[k];
        if (fabsf(x) < 1e-30f) x = x < 0.0f ? -1e-30f : 1e-30f;
        inv[k] = 1.0f / x;
    }
}

static void lbvh_ray_range(void* vctx, int64_t lo, int64_t hi) {
    qctx* c = (qctx*)vctx;
    obvh* b = c->b;
    int any = c->occ != NULL;
    int64_t nodes_visited = 0, tris_tested = 0;
    int32_t st_n[ORC_STACK];
    float st_t[ORC_STACK];
    for (int64_t i = lo; i < hi; i++) {
        const float* r = c->R + 8 * i;
        const float* o = r;
        const float* d = r + 4;
        float tmin = r[3], best = r[7], bu = 0, bv = 0, inv[3];
        int32_t bi = -1;
        uint8_t hit = 0;
        ray_inv(d, inv);
        int sp = 0;
        float t0 = box_ray(&b->nodes[0], o, inv, tmin, best);
        if (t0 < INFINITY && orc_ray_ok(r)) { st_n[sp] = 0; st_t[sp] = t0; sp++; }
        while (sp > 0 && !hit) {
            sp--;
            int32_t ni = st_n[sp];
            if (st_t[sp] > best) continue;
            const onode* nd = &b->nodes[ni];
            nodes_visited++;
            if (nd->left < 0) {
                for (int32_t p = nd->first; p < nd->first + nd->count; p++) {
                    const float* tr = b->tris + 9 * (int64_t)p;
                    float t, u, v;
                    tris_tested++;
                    if (!orc_mt_f32(o, d, tr, tr + 3, tr + 6, &t, &u, &v)) continue;
                    if (!(t >= tmin && t <= best)) continue;
                    int32_t id = b->ids[p];
                    if (any) { hit = 1; break; }
                    if (t < best || bi < 0 || id < bi) { best = t; bi = id; bu = u; bv = v; }
                }
                continue;
            }
            float tl = box_ray(&b->nodes[nd->left], o, inv, tmin, best);
            float tr = box_ray(&b->nodes[nd->right], o, inv, tmin, best);
            int32_t nl = nd->left, nr = nd->right;
            if (tl > tr) { float tf = tl; tl = tr; tr = tf; int32_t ti = nl; nl = nr; nr = ti; }
            if (tr < INFINITY && sp < ORC_STACK) { st_n[sp] = nr; st_t[sp] = tr; sp++; }
            if (tl < INFINITY && sp < ORC_STACK) { st_n[sp] = nl; st_t[sp] = tl; sp++; }
        }
        if (any) {
            c->occ[i] = hit;
        } else {
            c->prim[i] = bi;
            c->t[i] = bi >= 0 ? best : INFINITY;
            if (c->uv) { c->uv[2 * i] = bi >= 0 ? bu : 0.0f; c->uv[2 * i + 1] = bi >= 0 ? bv : 0.0f; }
        }
    }
    if (c->counters) {
        __atomic_fetch_add(&c->counters[0], nodes_visited, __ATOMIC_RELAXED);
        __atomic_fetch_add(&c->counters[1], tris_tested, __ATOMIC_RELAXED);
    }
}

ORC_API void orc_lbvh_ray_closest(void* h, const float* R, int64_t nr, int32_t* prim, float* t,
                                  float* uv, int64_t* counters) {
    qctx c;
    memset(&c, 0, sizeof c);
    c.b = (obvh*)h; c.R = R; c.prim = prim; c.t = t; c.uv = uv; c.counters = counters;
    parallel_for(nr, 256, lbvh_ray_range, &c);
}

ORC_API void orc_lbvh_ray_any(void* h, const float* R, int64_t nr, uint8_t* occ,
                              int64_t* counters) {
    qctx c;
    memset(&c, 0, sizeof c);
    c.b = (obvh*)h; c.R = R; c.occ = occ; c.counters = counters;
    parallel_for(nr, 256, lbvh_ray_range, &c);
}
