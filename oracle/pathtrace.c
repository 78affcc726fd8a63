/*
 * pathtrace.c -- CPU restatement of the reference's path tracer (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/ and bench.py's cpu_baseline leg use it, as the test oracle.  It restates the
 * ray-generation shader src/shaders/rt/rt.rgen in its own per-pixel form: for every pixel
 * a loop over samples and bounces (main, rt.rgen:611-677) calling
 *   make_camera_ray (597-609), trace_ray / visibility (257-291) -- here the FP32
 *     instanced brute force of oracle.c (orc_mt_f32 on object-space rays), hit_info /
 *     mat_info / shade_info (62-150, factors only: no textures or normal maps),
 *   integrate_direct (393-414), integrate_mats (362-391), integrate_mis (303-360) with
 *     light_sample / light_sample_dir / light_pdf (152-255, light_pdf as the reference's
 *     loop over every light triangle),
 *   Russian roulette (660-665), the per-frame mean and blend (667-681),
 * and the helpers of src/shaders/rt/rtcommon.glsl: tea / lcg / randf / randu (99-124),
 * radical_inverse / hammersley (126-137), cospow_hemisphere (163-171), triangle_sample
 * (175-181), triangle_hit / triangle_pdf (181-214), make_tanspace / power_heuristic
 * (218-230), Blinn-Phong (254-277) and GGX (279-349).  Pi is the reference's 3.141592.
 * Every product and sum is rounded to FP32 separately in the written order (compile with
 * -ffp-contract=off); cos / sin / pow are the C library's, so a GPU implementation with
 * the same operation order agrees to within their last-bit differences.
 *
 * The per-light bounding-box test of light_pdf (rt.rgen:233) is left out: it only skips
 * lights whose triangles triangle_hit rejects anyway.  A scene without lights contributes
 * no light samples (the reference would take a remainder by zero); the four random draws
 * of a light sample are still consumed.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* from oracle.c */
int orc_mt_f32(const float* o, const float* d, const float* pa, const float* pb, const float* pc, float* t_out,
               float* u_out, float* v_out);
int orc_invert3(const float* t, float* minv);
typedef void (*orc_range_fn)(void* ctx, int64_t lo, int64_t hi);
void orc_parallel_for(int64_t n, int64_t chunk, orc_range_fn fn, void* ctx);

#define PT_PI 3.141592f
#define PT_EPS 0.00001f
#define PT_LARGE 10000000.0f

typedef struct {
    float x, y, z;
} f3;

static inline f3 mk(float x, float y, float z) {
    f3 r = {x, y, z};
    return r;
}
static inline f3 add3(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub3(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 mul3(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 scl3(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline f3 div3(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline f3 neg3(f3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(f3 a, f3 b) {
    float s = a.x * b.x;
    s = s + a.y * b.y;
    s = s + a.z * b.z;
    return s;
}
static inline f3 crs(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float len3(f3 a) { return sqrtf(dot(a, a)); }
static inline f3 nrm3(f3 a) { return div3(a, len3(a)); }
/* x*a + y*b + z*c summed left to right */
static inline f3 lin3(f3 x, float a, f3 y, float b, f3 z, float c) {
    return add3(add3(scl3(x, a), scl3(y, b)), scl3(z, c));
}
static inline f3 reflect3(f3 i, f3 n) { return sub3(i, scl3(n, 2.0f * dot(n, i))); }
static inline f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

/* mat4 (row-major) times a point / a direction, each row summed left to right */
static inline f3 xpoint(const float* m, f3 p) {
    return mk(((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3], ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7],
              ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11]);
}
static inline f3 xvec9(const float* m, f3 p) {
    return mk((m[0] * p.x + m[1] * p.y) + m[2] * p.z, (m[3] * p.x + m[4] * p.y) + m[5] * p.z,
              (m[6] * p.x + m[7] * p.y) + m[8] * p.z);
}
static inline void x4(const float* m, float x, float y, float z, float w, float* r) {
    for (int i = 0; i < 4; i++) r[i] = ((m[4 * i] * x + m[4 * i + 1] * y) + m[4 * i + 2] * z) + m[4 * i + 3] * w;
}

/* ---- RNG (rtcommon.glsl:99-124) ---- */

ORC_API uint32_t orc_tea_rounds(uint32_t v0, uint32_t v1, const uint32_t* key, int rounds) {
    uint32_t s0 = 0;
    for (int n = 0; n < rounds; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + key[0]) ^ (v1 + s0) ^ ((v1 >> 5) + key[1]);
        v1 += ((v0 << 4) + key[2]) ^ (v0 + s0) ^ ((v0 >> 5) + key[3]);
    }
    return v0;
}
static const uint32_t k_tea_key[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};
ORC_API uint32_t orc_tea(uint32_t v0, uint32_t v1) { return orc_tea_rounds(v0, v1, k_tea_key, 16); }
static inline uint32_t lcg(uint32_t* s) {
    *s = 1664525u * *s + 1013904223u;
    return *s & 0x00FFFFFFu;
}
static inline float randf(uint32_t* s) { return (float)lcg(s) / (float)0x01000000; }
static inline uint32_t randu(uint32_t* s, uint32_t a, uint32_t b) { return lcg(s) % (b - a) + a; }
/* n successive randf draws from seed (for the RNG tests) */
ORC_API void orc_randf_seq(uint32_t seed, int64_t n, float* out) {
    for (int64_t i = 0; i < n; i++) out[i] = randf(&seed);
}

static inline float radical_inverse(uint32_t bits) {
    bits = (bits << 16u) | (bits >> 16u);
    bits = ((bits & 0x55555555u) << 1u) | ((bits & 0xAAAAAAAAu) >> 1u);
    bits = ((bits & 0x33333333u) << 2u) | ((bits & 0xCCCCCCCCu) >> 2u);
    bits = ((bits & 0x0F0F0F0Fu) << 4u) | ((bits & 0xF0F0F0F0u) >> 4u);
    bits = ((bits & 0x00FF00FFu) << 8u) | ((bits & 0xFF00FF00u) >> 8u);
    return (float)bits * 2.3283064365386963e-10f;
}

/* ---- sampling and materials ---- */

static inline f3 cospow_hemisphere(float e, uint32_t* seed, f3 x, f3 y, f3 z) {
    float phi = (2.0f * PT_PI) * randf(seed);
    float cosT = powf(randf(seed), 1.0f / (e + 1.0f));
    float sinT = sqrtf(1.0f - cosT * cosT);
    return lin3(x, cosf(phi) * sinT, y, sinf(phi) * sinT, z, cosT);
}
static inline f3 triangle_sample(uint32_t* seed) {
    float u = sqrtf(randf(seed));
    float v = randf(seed);
    float a = u * (1.0f - v);
    float b = u * v;
    return mk(a, b, (1.0f - a) - b);
}
static inline int triangle_hit(f3 o, f3 d, f3 pa, f3 pb, f3 pc, f3* hitp) {
    f3 v1 = sub3(pb, pa), v2 = sub3(pc, pa);
    f3 p = crs(d, v2);
    float det = dot(v1, p);
    if (fabsf(det) < PT_EPS) return 0;
    float inv = 1.0f / det;
    f3 s = sub3(o, pa);
    float u = dot(s, p) * inv;
    if (u < 0.0f || u > 1.0f) return 0;
    f3 q = crs(s, v1);
    float v = dot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return 0;
    float t = dot(v2, q) * inv;
    *hitp = add3(o, scl3(d, t));
    return t >= 0.0f;
}
static inline float triangle_pdf(f3 o, f3 d, f3 v0, f3 v1, f3 v2) {
    f3 hitp;
    if (!triangle_hit(o, d, v0, v1, v2, &hitp)) return 0.0f;
    f3 n = crs(sub3(v1, v0), sub3(v2, v0));
    float a = 2.0f / len3(n);
    f3 dist = sub3(hitp, o);
    f3 N = nrm3(n);
    float g = dot(dist, dist) / fabsf(dot(N, d));
    return a * g;
}
static inline void make_tanspace(f3 N, f3* T, f3* B) {
    if (fabsf(N.x) > fabsf(N.y))
        *T = div3(mk(N.z, 0.0f, -N.x), sqrtf(N.x * N.x + N.z * N.z));
    else
        *T = div3(mk(0.0f, -N.z, N.y), sqrtf(N.y * N.y + N.z * N.z));
    *B = crs(N, *T);
}
static inline float power_heuristic(float a, float b) { return a * a / (a * a + b * b); }

typedef struct {
    f3 albedo, emissive;
    float roughness;
} mat_t;
typedef struct {
    f3 wo, N, T, B;
} shade_t;

static float bp_pdf(const mat_t* m, const shade_t* s, f3 wi) {
    float oDn = dot(neg3(s->wo), s->N);
    float iDn = dot(wi, s->N);
    if (oDn <= 0.0f || iDn <= 0.0f) return 0.0f;
    float e = 1.0f / m->roughness;
    f3 H = nrm3(sub3(wi, s->wo));
    float cosine = fmaxf(dot(H, s->N), 0.0f);
    float n_pdf = (e + 1.0f) / (2.0f * PT_PI) * powf(cosine, e);
    return n_pdf / (4.0f * dot(neg3(s->wo), H));
}
static int bp_sample(uint32_t* seed, const mat_t* m, const shade_t* s, f3* wi) {
    float e = 1.0f / m->roughness;
    f3 H = cospow_hemisphere(e, seed, s->T, s->B, s->N);
    *wi = reflect3(s->wo, H);
    return dot(*wi, s->N) > 0.0f;
}
static float ggx_D(float nDh, float a2) {
    float b = nDh * nDh * (a2 - 1.0f) + 1.0f;
    return a2 / (PT_PI * b * b);
}
static float ggx_G(float oDn, float iDn, float a2) {
    float sqr0 = sqrtf(a2 + (1.0f - a2) * iDn * iDn);
    float sqr1 = sqrtf(a2 + (1.0f - a2) * oDn * oDn);
    return 2.0f * oDn * iDn / (oDn * sqr0 + iDn * sqr1);
}
static float ggx_pdf(const mat_t* m, const shade_t* s, f3 wi) {
    float oDn = dot(neg3(s->wo), s->N);
    float iDn = dot(wi, s->N);
    if (oDn <= 0.0f || iDn <= 0.0f) return 0.0f;
    f3 H = nrm3(sub3(wi, s->wo));
    float nDh = fmaxf(dot(H, s->N), 0.0f);
    float oDh = fmaxf(dot(wi, H), 0.0f);
    float a2 = m->roughness * m->roughness;
    return ggx_D(nDh, a2) * nDh / (4.0f * oDh);
}
static f3 ggx_eval(const mat_t* m, const shade_t* s, f3 wi) {
    float oDn = dot(neg3(s->wo), s->N);
    float iDn = dot(wi, s->N);
    if (oDn <= 0.0f || iDn <= 0.0f) return mk(0.0f, 0.0f, 0.0f);
    f3 H = nrm3(sub3(wi, s->wo));
    float nDh = fmaxf(dot(H, s->N), 0.0f);
    float a2 = m->roughness * m->roughness;
    float cos5 = powf(1.0f - iDn, 5.0f);
    f3 F = mk(m->albedo.x + (1.0f - m->albedo.x) * cos5, m->albedo.y + (1.0f - m->albedo.y) * cos5,
              m->albedo.z + (1.0f - m->albedo.z) * cos5);
    float D = ggx_D(nDh, a2), G = ggx_G(oDn, iDn, a2), den = 4.0f * oDn;
    return mk(F.x * D * G / den, F.y * D * G / den, F.z * D * G / den);
}
static int ggx_sample(uint32_t* seed, const mat_t* m, const shade_t* s, f3* wi) {
    float a2 = m->roughness * m->roughness;
    float x0 = randf(seed);
    float x1 = randf(seed);
    float phi = 2.0f * PT_PI * x0;
    float cosTheta = sqrtf((1.0f - x1) / (1.0f + (a2 - 1.0f) * x1));
    float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
    f3 H = lin3(s->T, cosf(phi) * sinTheta, s->B, sinf(phi) * sinTheta, s->N, cosTheta);
    *wi = reflect3(s->wo, H);
    return dot(*wi, s->N) > 0.0f;
}
static float mat_pdf(int brdf, const mat_t* m, const shade_t* s, f3 wi) {
    return brdf == 0 ? bp_pdf(m, s, wi) : ggx_pdf(m, s, wi);
}
static f3 mat_eval(int brdf, const mat_t* m, const shade_t* s, f3 wi) {
    return brdf == 0 ? scl3(m->albedo, bp_pdf(m, s, wi)) : ggx_eval(m, s, wi);
}
static int mat_sample(int brdf, uint32_t* seed, const mat_t* m, const shade_t* s, f3* wi) {
    return brdf == 0 ? bp_sample(seed, m, s, wi) : ggx_sample(seed, m, s, wi);
}

/* Exposed for the material tests: MAT_pdf / MAT_eval / MAT_sample at a shading frame
 * built from N (wo is the incoming direction).  out = {ok, wi.xyz, pdf(wi), eval(wi).xyz}
 * per sample, n samples drawn from seed. */
ORC_API void orc_mat_samples(int brdf, const float* albedo, float roughness, const float* wo, const float* N,
                             uint32_t seed, int64_t n, float* out) {
    mat_t m = {ld3(albedo), mk(0, 0, 0), roughness};
    shade_t s;
    s.wo = ld3(wo);
    s.N = ld3(N);
    make_tanspace(s.N, &s.T, &s.B);
    for (int64_t i = 0; i < n; i++) {
        f3 wi = mk(0, 0, 0);
        int ok = mat_sample(brdf, &seed, &m, &s, &wi);
        f3 ev = mat_eval(brdf, &m, &s, wi);
        float* o = out + 8 * i;
        o[0] = (float)ok;
        o[1] = wi.x;
        o[2] = wi.y;
        o[3] = wi.z;
        o[4] = mat_pdf(brdf, &m, &s, wi);
        o[5] = ev.x;
        o[6] = ev.y;
        o[7] = ev.z;
    }
}
/* MAT_pdf / MAT_eval for given directions: out = {pdf, eval.xyz} per direction */
ORC_API void orc_mat_eval(int brdf, const float* albedo, float roughness, const float* wo, const float* N,
                          const float* wi, int64_t n, float* out) {
    mat_t m = {ld3(albedo), mk(0, 0, 0), roughness};
    shade_t s;
    s.wo = ld3(wo);
    s.N = ld3(N);
    make_tanspace(s.N, &s.T, &s.B);
    for (int64_t i = 0; i < n; i++) {
        f3 w = ld3(wi + 3 * i);
        f3 ev = mat_eval(brdf, &m, &s, w);
        out[4 * i] = mat_pdf(brdf, &m, &s, w);
        out[4 * i + 1] = ev.x;
        out[4 * i + 2] = ev.y;
        out[4 * i + 3] = ev.z;
    }
}

/* ---- scene ---- */

typedef struct {
    int64_t n_obj;
    const float* model;    /* 12 per object: rows of the 3x4 object-to-world */
    float* minv;           /* 9 per object (orc_invert3) */
    float* nit;            /* 9 per object: transpose of minv (modelIT) */
    const int64_t* voff;   /* vertices of object i: [voff[i], voff[i+1]) */
    const float* V;        /* positions, 3 per vertex */
    const float* N;        /* normals, 3 per vertex */
    const int64_t* foff;   /* triangles of object i: [foff[i], foff[i+1]) */
    const uint32_t* F;     /* object-local vertex numbers, 3 per triangle */
    const float* mat;      /* 8 per object: albedo.rgb, emissive.rgb, metalness, roughness */
    int64_t n_lights;
    int64_t* lights;       /* object of each light */
    float* lworld;         /* world positions of every light triangle, 9 each, light order */
    int64_t* lfirst;       /* first light triangle of each light */
} scene_t;

static inline int is_emissive(const float* m) { return m[3] != 0.0f || m[4] != 0.0f || m[5] != 0.0f; }

/* FP32 closest hit (occ == NULL) or any hit of one world ray over every object */
static int cast(const scene_t* sc, f3 o, float tmin, f3 d, float tmax, int any, int64_t* obj, int64_t* prim,
                float* uo, float* vo) {
    const float ow[3] = {o.x, o.y, o.z}, dw[3] = {d.x, d.y, d.z};
    float best = tmax, bu = 0.0f, bv = 0.0f;
    int64_t bi = -1, bk = -1;
    for (int64_t n = 0; n < sc->n_obj; n++) {
        const float* t = sc->model + 12 * n;
        const float* mi = sc->minv + 9 * n;
        const float x[3] = {ow[0] - t[3], ow[1] - t[7], ow[2] - t[11]};
        float oo[3], od[3];
        for (int r = 0; r < 3; r++) {
            float a = mi[3 * r] * x[0];
            a = a + mi[3 * r + 1] * x[1];
            a = a + mi[3 * r + 2] * x[2];
            oo[r] = a;
            float b = mi[3 * r] * dw[0];
            b = b + mi[3 * r + 1] * dw[1];
            b = b + mi[3 * r + 2] * dw[2];
            od[r] = b;
        }
        const float* V = sc->V + 3 * sc->voff[n];
        for (int64_t k = sc->foff[n]; k < sc->foff[n + 1]; k++) {
            const uint32_t* f = sc->F + 3 * k;
            float tt, u, v;
            if (!orc_mt_f32(oo, od, V + 3 * f[0], V + 3 * f[1], V + 3 * f[2], &tt, &u, &v)) continue;
            if (!(tt >= tmin)) continue;
            if (any) {
                if (tt <= tmax) return 1;
                continue;
            }
            if (tt < best || (tt == best && bi < 0)) {
                best = tt;
                bi = n;
                bk = k - sc->foff[n];
                bu = u;
                bv = v;
            }
        }
    }
    if (any) return 0;
    *obj = bi;
    *prim = bk;
    *uo = bu;
    *vo = bv;
    return bi >= 0;
}

static float light_pdf(const scene_t* sc, f3 p, f3 d) {
    float oacc = 0.0f;
    for (int64_t l = 0; l < sc->n_lights; l++) {
        const int64_t o = sc->lights[l];
        const int64_t nt = sc->foff[o + 1] - sc->foff[o];
        float tacc = 0.0f;
        for (int64_t t = 0; t < nt; t++) {
            const float* w = sc->lworld + 9 * (sc->lfirst[l] + t);
            tacc = tacc + triangle_pdf(p, d, ld3(w), ld3(w + 3), ld3(w + 6));
        }
        oacc = oacc + tacc / (float)nt;
    }
    return oacc / (float)sc->n_lights;
}

typedef struct {
    f3 pos, emissive;
    float pdf;
} lsamp_t;

static lsamp_t light_sample(const scene_t* sc, uint32_t* seed, f3 p) {
    lsamp_t s;
    uint32_t l = randu(seed, 0u, (uint32_t)sc->n_lights);
    int64_t o = sc->lights[l];
    uint32_t nt = (uint32_t)(sc->foff[o + 1] - sc->foff[o]);
    uint32_t t = randu(seed, 0u, nt);
    const float* w = sc->lworld + 9 * (sc->lfirst[l] + t);
    f3 w0 = ld3(w), w1 = ld3(w + 3), w2 = ld3(w + 6);
    f3 b = triangle_sample(seed);
    s.pos = lin3(w0, b.x, w1, b.y, w2, b.z);
    s.emissive = ld3(sc->mat + 8 * o + 3);
    f3 narea = crs(sub3(w1, w0), sub3(w2, w0));
    float a = 2.0f / len3(narea);
    f3 dist = sub3(s.pos, p);
    f3 N = nrm3(narea);
    f3 dd = nrm3(dist);
    float g = dot(dist, dist) / fabsf(dot(N, dd));
    s.pdf = a * g / (float)(nt * (uint32_t)sc->n_lights);
    return s;
}

static f3 light_sample_dir(const scene_t* sc, uint32_t* seed, f3 p) {
    uint32_t l = randu(seed, 0u, (uint32_t)sc->n_lights);
    int64_t o = sc->lights[l];
    uint32_t nt = (uint32_t)(sc->foff[o + 1] - sc->foff[o]);
    uint32_t t = randu(seed, 0u, nt);
    const uint32_t* f = sc->F + 3 * (sc->foff[o] + t);
    const float* V = sc->V + 3 * sc->voff[o];
    f3 b = triangle_sample(seed);
    f3 point = xpoint(sc->model + 12 * o, lin3(ld3(V + 3 * f[0]), b.x, ld3(V + 3 * f[1]), b.y, ld3(V + 3 * f[2]), b.z));
    return nrm3(sub3(point, p));
}

/* ---- render ---- */

typedef struct {
    const scene_t* sc;
    const float* fp;      /* iV[16], iP[16], clear[3], env[3] */
    const int32_t* ip;    /* width, height, frame, samples, max_frame, qmc, max_depth, use_metalness,
                             integrator, brdf, use_rr, seed */
    float* image;
    float* gpos;
    float* gnorm;
    float* galb;
} render_ctx;

static void render_range(void* vctx, int64_t lo, int64_t hi) {
    const render_ctx* c = (const render_ctx*)vctx;
    const scene_t* sc = c->sc;
    const float *iV = c->fp, *iP = c->fp + 16;
    const f3 clear = ld3(c->fp + 32), env = ld3(c->fp + 35);
    const uint32_t w = (uint32_t)c->ip[0], h = (uint32_t)c->ip[1], frame = (uint32_t)c->ip[2];
    const uint32_t samples = (uint32_t)c->ip[3], max_frame = (uint32_t)c->ip[4];
    const int qmc = c->ip[5], max_depth = c->ip[6], use_metal = c->ip[7], integ = c->ip[8], brdf = c->ip[9];
    const int use_rr = c->ip[10];
    const uint32_t seed0 = (uint32_t)c->ip[11];
    for (int64_t pix = lo; pix < hi; pix++) {
        const uint32_t x = (uint32_t)pix % w, y = (uint32_t)pix / w;
        uint32_t seed = orc_tea((uint32_t)pix, seed0);
        f3 acc_pix = mk(0, 0, 0), gp = mk(0, 0, 0), gn = mk(0, 0, 0), ga = mk(0, 0, 0);
        float org[4];
        x4(iV, 0.0f, 0.0f, 0.0f, 1.0f, org);
        for (uint32_t s = 0; s < samples; s++) {
            /* make_camera_ray */
            float jx, jy;
            if (qmc == 0) {
                if (frame == 0) {
                    jx = 0.5f;
                    jy = 0.5f;
                } else {
                    jx = randf(&seed);
                    jy = randf(&seed);
                }
            } else {
                uint32_t i = s + samples * frame, N = samples * max_frame;
                jx = (float)i / (float)N;
                jy = radical_inverse(i);
            }
            const float ux = ((float)x + jx) / (float)w, uy = ((float)y + jy) / (float)h;
            float target[4], dir[4];
            x4(iP, ux * 2.0f - 1.0f, uy * 2.0f - 1.0f, 0.0f, 1.0f, target);
            x4(iV, target[0], target[1], target[2], 0.0f, dir);
            f3 o = mk(org[0], org[1], org[2]);
            f3 d = nrm3(mk(dir[0], dir[1], dir[2]));
            f3 acc = mk(0, 0, 0), thr = mk(1, 1, 1);
            float mis = 1.0f;
            for (int depth = 0; depth < max_depth; depth++) {
                int64_t obj, prim;
                float u, v;
                if (!cast(sc, o, PT_EPS, d, PT_LARGE, 0, &obj, &prim, &u, &v)) {
                    if (depth == 0)
                        acc = clear;
                    else
                        acc = add3(acc, mul3(env, thr));
                    break;
                }
                /* hit_info */
                const float b0 = (1.0f - u) - v;
                const uint32_t* f = sc->F + 3 * (sc->foff[obj] + prim);
                const float* V = sc->V + 3 * sc->voff[obj];
                const float* Nv = sc->N + 3 * sc->voff[obj];
                f3 n = lin3(ld3(Nv + 3 * f[0]), b0, ld3(Nv + 3 * f[1]), u, ld3(Nv + 3 * f[2]), v);
                f3 normal = nrm3(xvec9(sc->nit + 9 * obj, n));
                f3 pos = xpoint(sc->model + 12 * obj, lin3(ld3(V + 3 * f[0]), b0, ld3(V + 3 * f[1]), u, ld3(V + 3 * f[2]), v));
                /* mat_info */
                const float* mm = sc->mat + 8 * obj;
                mat_t m = {ld3(mm), ld3(mm + 3), mm[7]};
                if (use_metal) {
                    const float mt = mm[6];
                    const float k = 0.04f * (1.0f - mt);
                    m.albedo = add3(mk(k, k, k), scl3(m.albedo, mt));
                }
                /* shade_info */
                shade_t s_;
                s_.wo = d;
                s_.N = dot(s_.wo, normal) > 0.0f ? neg3(normal) : normal;
                make_tanspace(s_.N, &s_.T, &s_.B);
                if (s == 0 && depth == 0) {
                    gp = pos;
                    gn = s_.N;
                    ga = m.albedo;
                }
                int term = 0;
                const int emissive = m.emissive.x > 0.0f || m.emissive.y > 0.0f || m.emissive.z > 0.0f;
                f3 nd = d;
                if (integ == 0) {
                    term = 1;
                    if (emissive) {
                        acc = add3(acc, m.emissive);
                    } else if (m.roughness != 0.0f) {
                        if (sc->n_lights > 0) {
                            lsamp_t ls = light_sample(sc, &seed, pos);
                            f3 wi = nrm3(sub3(ls.pos, pos));
                            f3 atten = mat_eval(brdf, &m, &s_, wi);
                            if (ls.pdf != 0.0f) {
                                f3 dv = sub3(ls.pos, pos);
                                float dl = len3(dv);
                                float shadow = cast(sc, pos, PT_EPS, div3(dv, dl), dl - PT_EPS, 1, NULL, NULL, NULL, NULL)
                                                   ? 0.0f
                                                   : 1.0f;
                                acc = add3(acc, scl3(mul3(div3(atten, ls.pdf), ls.emissive), shadow));
                            }
                        } else {
                            for (int k = 0; k < 4; k++) (void)lcg(&seed);
                        }
                    }
                } else if (integ == 1) {
                    if (emissive) {
                        acc = add3(acc, mul3(m.emissive, thr));
                        term = 1;
                    } else if (m.roughness == 0.0f) {
                        nd = reflect3(s_.wo, s_.N);
                        thr = mul3(thr, m.albedo);
                    } else {
                        f3 wi;
                        if (!mat_sample(brdf, &seed, &m, &s_, &wi)) {
                            term = 1;
                        } else {
                            float pdf = mat_pdf(brdf, &m, &s_, wi);
                            f3 atten = mat_eval(brdf, &m, &s_, wi);
                            if (pdf != 0.0f) {
                                thr = mul3(thr, div3(atten, pdf));
                                nd = wi;
                            } else {
                                term = 1;
                            }
                        }
                    }
                } else {
                    if (emissive) {
                        acc = add3(acc, mul3(scl3(thr, mis), m.emissive));
                        term = 1;
                    } else if (m.roughness == 0.0f) {
                        nd = reflect3(s_.wo, s_.N);
                        thr = mul3(thr, m.albedo);
                        mis = 1.0f;
                    } else {
                        f3 wl = mk(0, 0, 0);
                        float pdf_ll = 0.0f;
                        if (sc->n_lights > 0) {
                            wl = light_sample_dir(sc, &seed, pos);
                            pdf_ll = light_pdf(sc, pos, wl);
                        } else {
                            for (int k = 0; k < 4; k++) (void)lcg(&seed);
                        }
                        if (pdf_ll != 0.0f) {
                            float pdf_lm = mat_pdf(brdf, &m, &s_, wl);
                            f3 atten = mat_eval(brdf, &m, &s_, wl);
                            f3 weight = scl3(div3(atten, pdf_ll), power_heuristic(pdf_ll, pdf_lm));
                            int64_t lo_, lp_;
                            float lu_, lv_;
                            f3 L = env;
                            if (cast(sc, pos, PT_EPS, wl, PT_LARGE, 0, &lo_, &lp_, &lu_, &lv_)) L = ld3(sc->mat + 8 * lo_ + 3);
                            acc = add3(acc, mul3(mul3(thr, weight), L));
                        }
                        f3 wb;
                        if (!mat_sample(brdf, &seed, &m, &s_, &wb)) {
                            term = 1;
                        } else {
                            float pdf_bm = mat_pdf(brdf, &m, &s_, wb);
                            if (pdf_bm != 0.0f) {
                                float pdf_bl = light_pdf(sc, pos, wb);
                                f3 atten = mat_eval(brdf, &m, &s_, wb);
                                thr = mul3(thr, div3(atten, pdf_bm));
                                mis = power_heuristic(pdf_bm, pdf_bl);
                                nd = wb;
                            } else {
                                term = 1;
                            }
                        }
                    }
                }
                if (use_rr) {
                    float pcont = fminf(fmaxf(thr.x, fmaxf(thr.y, thr.z)) + 0.001f, 0.95f);
                    if (randf(&seed) >= pcont) break;
                    thr = div3(thr, pcont);
                }
                if (term) break;
                o = pos;
                d = nd;
            }
            acc_pix = add3(acc_pix, acc);
        }
        const float ns = (float)samples;
        f3 avg = div3(acc_pix, ns);
        float* im = c->image + 4 * pix;
        if (frame > 0) {
            const float a = 1.0f / (float)(frame + 1);
            avg = mk(im[0] * (1.0f - a) + avg.x * a, im[1] * (1.0f - a) + avg.y * a, im[2] * (1.0f - a) + avg.z * a);
        }
        im[0] = avg.x;
        im[1] = avg.y;
        im[2] = avg.z;
        im[3] = 1.0f;
        if (c->gpos) {
            float* g = c->gpos + 4 * pix;
            g[0] = gp.x, g[1] = gp.y, g[2] = gp.z, g[3] = 1.0f;
            g = c->gnorm + 4 * pix;
            g[0] = gn.x, g[1] = gn.y, g[2] = gn.z, g[3] = 1.0f;
            g = c->galb + 4 * pix;
            g[0] = ga.x, g[1] = ga.y, g[2] = ga.z, g[3] = 1.0f;
        }
    }
}

static int scene_init(scene_t* sc, int64_t n_obj, const float* model, const int64_t* voff, const float* V,
                      const float* N, const int64_t* foff, const uint32_t* F, const float* mat) {
    memset(sc, 0, sizeof(*sc));
    sc->n_obj = n_obj;
    sc->model = model;
    sc->voff = voff;
    sc->V = V;
    sc->N = N;
    sc->foff = foff;
    sc->F = F;
    sc->mat = mat;
    sc->minv = (float*)malloc(sizeof(float) * 9 * (size_t)n_obj);
    sc->nit = (float*)malloc(sizeof(float) * 9 * (size_t)n_obj);
    sc->lights = (int64_t*)malloc(sizeof(int64_t) * (size_t)n_obj);
    sc->lfirst = (int64_t*)malloc(sizeof(int64_t) * (size_t)n_obj);
    if (!sc->minv || !sc->nit || !sc->lights || !sc->lfirst) return 0;
    int64_t nl = 0;
    for (int64_t i = 0; i < n_obj; i++) {
        if (!orc_invert3(model + 12 * i, sc->minv + 9 * i)) return 0;
        for (int r = 0; r < 3; r++)
            for (int k = 0; k < 3; k++) sc->nit[9 * i + 3 * r + k] = sc->minv[9 * i + 3 * k + r];
        if (is_emissive(mat + 8 * i) && foff[i + 1] > foff[i]) {
            sc->lfirst[sc->n_lights] = nl;
            sc->lights[sc->n_lights++] = i;
            nl += foff[i + 1] - foff[i];
        }
    }
    sc->lworld = (float*)malloc(sizeof(float) * 9 * (size_t)(nl > 0 ? nl : 1));
    if (!sc->lworld) return 0;
    for (int64_t l = 0; l < sc->n_lights; l++) {
        const int64_t o = sc->lights[l];
        const float* Vo = V + 3 * voff[o];
        for (int64_t t = 0; t < foff[o + 1] - foff[o]; t++)
            for (int c = 0; c < 3; c++) {
                f3 w = xpoint(model + 12 * o, ld3(Vo + 3 * F[3 * (foff[o] + t) + c]));
                float* dst = sc->lworld + 9 * (sc->lfirst[l] + t) + 3 * c;
                dst[0] = w.x, dst[1] = w.y, dst[2] = w.z;
            }
    }
    return 1;
}

static void scene_free(scene_t* sc) {
    free(sc->minv);
    free(sc->nit);
    free(sc->lights);
    free(sc->lfirst);
    free(sc->lworld);
}

/* Render one frame.  model: 12 floats per object (row-major 3x4 object-to-world); V / N:
 * positions / normals of all objects, object i's at [voff[i], voff[i+1]); F: object-local
 * triangles, object i's at [foff[i], foff[i+1]); mat: 8 floats per object.  fp / ip as in
 * render_ctx.  image (h*w*4) is read when frame > 0.  Returns 0 on a singular transform. */
ORC_API int orc_render(int64_t n_obj, const float* model, const int64_t* voff, const float* V, const float* N,
                       const int64_t* foff, const uint32_t* F, const float* mat, const float* fp, const int32_t* ip,
                       float* image, float* gpos, float* gnorm, float* galb) {
    scene_t sc;
    if (!scene_init(&sc, n_obj, model, voff, V, N, foff, F, mat)) {
        scene_free(&sc);
        return 0;
    }
    render_ctx c = {&sc, fp, ip, image, gpos, gnorm, galb};
    orc_parallel_for((int64_t)ip[0] * ip[1], 16, render_range, &c);
    scene_free(&sc);
    return 1;
}

/* light_pdf for n (point, direction) pairs: the reference's loop over every light triangle */
ORC_API int orc_light_pdf(int64_t n_obj, const float* model, const int64_t* voff, const float* V,
                          const int64_t* foff, const uint32_t* F, const float* mat, const float* P,
                          const float* D, int64_t n, float* out) {
    scene_t sc;
    if (!scene_init(&sc, n_obj, model, voff, V, V, foff, F, mat)) {
        scene_free(&sc);
        return 0;
    }
    for (int64_t i = 0; i < n; i++) out[i] = sc.n_lights ? light_pdf(&sc, ld3(P + 3 * i), ld3(D + 3 * i)) : 0.0f;
    scene_free(&sc);
    return 1;
}
