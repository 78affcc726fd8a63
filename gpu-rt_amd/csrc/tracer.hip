// tracer.hip -- wavefront path tracer over a TLAS (gpurt_scene_*, gpurt_render).
//
// The reference renders with one ray-generation shader per pixel that loops over samples
// and bounces (src/shaders/rt/rt.rgen:611-677).  Here a frame is a sequence of
// stream-ordered launches over all live paths at once:
//
//   frame init    seed[p] = tea(p, seed), clear the per-pixel sum and G-buffer
//   per sample    camera rays for every pixel (make_camera_ray, rt.rgen:597-609)
//     per bounce  closest hit through the TLAS for the live paths
//                 shade: hit_info / mat_info / shade_info (rt.rgen:62-150), one of the
//                   integrators (303-414), Russian roulette (660-665); appends the next
//                   rays and the shadow / light rays to compacted lists (one atomic per
//                   wave)
//                 any hit (direct: visibility, 272-291) or closest hit (MIS:
//                   direct_light, 293-301) for the shadow / light rays, then their
//                   contributions are added
//     sum         pixel += this sample's radiance
//   resolve       image = mean over samples, blended with the previous frame's image
//
// Per pixel the RNG is drawn in the shader's order, so every path sees the reference's
// random numbers for the same seed.  light_pdf (rt.rgen:222-255) is answered by a BVH over
// the world-space light triangles instead of a loop over all of them.  Path state lives
// in HBM as SoA float4 records indexed by pixel; rays are compacted, so each launch
// covers only live paths.  The list lengths stay on the device: every launch reads its
// count there (the traversal's persistent blocks and the shading blocks past the count
// exit at once), so a frame is enqueued without any host synchronisation.
#include <cmath>
#include <string>
#include <vector>

#include "gpurt_internal.h"
#include "shade.h"

namespace gpurt_impl {
namespace {

using pt::kEps;
using pt::kLargeDist;

constexpr int kShadeBlock = 256;
constexpr int kCountSlots = 2 * 64 + 2;  // two list lengths per bounce, up to 64 bounces
// Pixel groups rendered concurrently, by integrator (measured at 1024x1024 on MI355X with
// HIP's 4 hardware queues per process: direct 4.56 / 4.22 / 4.57 ms for 1 / 2 / 3 groups,
// material sampling 14.8 / 14.4 / 14.2 / 13.8 ms for 1..4, MIS 21.4 / 26.9 ms for 1 / 2 --
// MIS already runs two streams per group, and more streams than hardware queues
// serialise)
constexpr int kDefaultGroups[3] = {2, 4, 1};
constexpr int kLightStack = 64;  // light BVH traversal stack (checked against its depth)
constexpr float kExitSlack = 0x1p-20f;  // relative slack on box exits (conservative culls)
constexpr uint32_t kLoopLightTris = 32;  // up to this many light triangles light_pdf loops

// per object: Scene_Desc (src/vk/rt.h:67-78) without textures
struct ObjDev {
    float m[12];    // object-to-world, rows of the 3x4 affine part
    float nit[9];   // inverse transpose of the 3x3 part (modelIT, rt.cpp:35), rows
    float albedo[3];
    float emissive[3];
    float mr[2];    // metalness, roughness
    uint32_t vbase;  // first vertex of the object in the concatenated vertex array
    uint32_t ibase;  // first index of the object in the concatenated index array
    uint32_t n_tris;
};

// per light: Scene_Light (rt.h:79-84) with the offset of its triangles in the light list
struct LightDev {
    uint32_t obj, n_tris, first, pad_;
};

struct SceneDev {
    const ObjDev* objs;
    const float4* verts;    // 3 float4 per vertex (Mesh::Vertex)
    const uint32_t* idx;    // 3 object-local vertex numbers per triangle
    const LightDev* lights;
    uint32_t n_lights;
    uint32_t n_ltris;
    const float4* lworld;   // 3 float4 (xyz, 0) per light triangle: world positions
    const WNode<4>* lnodes;  // light BVH (FP32 nodes, width 4)
    const float4* ltris;    // its packed triangles (id = light-triangle index)
    const float* ltri_n;    // triangle count of the light owning each light triangle
    float n_tris0;          // triangle count of light 0 (single-light normalisation)
};

// ---- light sampling and light_pdf ------------------------------------------------------

__device__ __forceinline__ V3 xyz(float4 a) { return v3(a.x, a.y, a.z); }

// light_pdf (rt.rgen:222-255): every light triangle that triangle_hit accepts for the ray
// (p, d, t >= 0) contributes triangle_pdf / n_tris(light); the sum is divided by the
// light count.  Beyond kLoopLightTris light triangles the terms are found through the
// light BVH: its box culls are conservative (boxes padded by the builder, slack on the
// exit distance), so the set of contributing triangles is the reference's; with one light
// the raw sum is divided once by its triangle count, as the reference does.
__device__ float light_pdf(const SceneDev& sc, V3 p, V3 d) {
    if (sc.n_lights == 0) return 0.0f;
    if (sc.n_ltris <= kLoopLightTris) {
        // few light triangles: the reference's own loop (per light, per triangle, in
        // order), cheaper than any traversal and the same sums bit for bit
        float oacc = 0.0f;
        for (uint32_t l = 0; l < sc.n_lights; l++) {
            const LightDev L = sc.lights[l];
            float tacc = 0.0f;
            for (uint32_t t = 0; t < L.n_tris; t++) {
                const size_t k = 3 * (size_t)(L.first + t);
                tacc = tacc + pt::triangle_pdf(p, d, xyz(sc.lworld[k]), xyz(sc.lworld[k + 1]), xyz(sc.lworld[k + 2]));
            }
            oacc = oacc + tacc / (float)L.n_tris;
        }
        return oacc / (float)sc.n_lights;
    }
    const float x = fabsf(d.x) < 1e-30f ? copysignf(1e-30f, d.x) : d.x;
    const float y = fabsf(d.y) < 1e-30f ? copysignf(1e-30f, d.y) : d.y;
    const float z = fabsf(d.z) < 1e-30f ? copysignf(1e-30f, d.z) : d.z;
    const V3 inv = v3(1.0f / x, 1.0f / y, 1.0f / z);
    const float oix = -p.x * inv.x, oiy = -p.y * inv.y, oiz = -p.z * inv.z;
    const bool one = sc.n_lights == 1;
    uint32_t stack[kLightStack];
    int sp = 0;
    stack[sp++] = 0u;
    float sum = 0.0f;
    while (sp > 0) {
        const WNode<4>& nd = sc.lnodes[stack[--sp]];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t c = nd.child[j];
            if (c == kEmpty) continue;
            const float t0x = fmaf(nd.lox[j], inv.x, oix), t1x = fmaf(nd.hix[j], inv.x, oix);
            const float t0y = fmaf(nd.loy[j], inv.y, oiy), t1y = fmaf(nd.hiy[j], inv.y, oiy);
            const float t0z = fmaf(nd.loz[j], inv.z, oiz), t1z = fmaf(nd.hiz[j], inv.z, oiz);
            const float lo = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), 0.0f));
            const float hi = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
            if (!(lo <= hi + fabsf(hi) * kExitSlack)) continue;
            if (!child_is_leaf(c)) {
                if (sp < kLightStack) stack[sp++] = c;
                continue;
            }
            const uint32_t f = leaf_first(c), cnt = leaf_count(c);
            for (uint32_t k = 0; k < cnt; k++) {
                const float4 A = sc.ltris[3 * (size_t)(f + k)];
                const float4 B = sc.ltris[3 * (size_t)(f + k) + 1];
                const float4 C = sc.ltris[3 * (size_t)(f + k) + 2];
                const float pdf = pt::triangle_pdf(p, d, xyz(A), xyz(B), xyz(C));
                if (pdf != 0.0f) sum = sum + (one ? pdf : pdf / sc.ltri_n[__float_as_uint(A.w)]);
            }
        }
    }
    return one ? sum / sc.n_tris0 : sum / (float)sc.n_lights;
}

struct LightSample {
    V3 pos, emissive;
    float pdf;
};

// light_sample (rt.rgen:152-200): uniform light, uniform triangle, uniform point
__device__ __forceinline__ LightSample light_sample(const SceneDev& sc, uint32_t& seed, V3 p) {
    LightSample s;
    const uint32_t l = pt::randu(seed, 0u, sc.n_lights);
    const LightDev L = sc.lights[l];
    const uint32_t t = pt::randu(seed, 0u, L.n_tris);
    const size_t k = 3 * (size_t)(L.first + t);
    const V3 w0 = xyz(sc.lworld[k]), w1 = xyz(sc.lworld[k + 1]), w2 = xyz(sc.lworld[k + 2]);
    const V3 b = pt::triangle_sample(seed);
    s.pos = pt::comb3(w0, b.x, w1, b.y, w2, b.z);
    const ObjDev& O = sc.objs[L.obj];
    s.emissive = v3(O.emissive[0], O.emissive[1], O.emissive[2]);
    const V3 narea = cross(sub(w1, w0), sub(w2, w0));
    const float a = 2.0f / pt::length(narea);
    const V3 dist = sub(s.pos, p);
    const V3 N = pt::normalize(narea);
    const V3 dd = pt::normalize(dist);
    const float g = dot(dist, dist) / fabsf(dot(N, dd));
    s.pdf = a * g / (float)(L.n_tris * sc.n_lights);
    return s;
}

// light_sample_dir (rt.rgen:202-220): the point is interpolated in object space, then
// placed by the model matrix
__device__ __forceinline__ V3 light_sample_dir(const SceneDev& sc, uint32_t& seed, V3 p) {
    const uint32_t l = pt::randu(seed, 0u, sc.n_lights);
    const LightDev L = sc.lights[l];
    const uint32_t t = pt::randu(seed, 0u, L.n_tris);
    const ObjDev& O = sc.objs[L.obj];
    const uint32_t* ix = sc.idx + O.ibase + 3 * (size_t)t;
    const V3 v0 = xyz(sc.verts[3 * (size_t)(O.vbase + ix[0])]);
    const V3 v1 = xyz(sc.verts[3 * (size_t)(O.vbase + ix[1])]);
    const V3 v2 = xyz(sc.verts[3 * (size_t)(O.vbase + ix[2])]);
    const V3 b = pt::triangle_sample(seed);
    const V3 point = pt::xform_point(O.m, pt::comb3(v0, b.x, v1, b.y, v2, b.z));
    return pt::normalize(sub(point, p));
}

// ---- wave-level append ----------------------------------------------------------------

// Slot of this lane in a list that grows by the wave's set predicates (one atomic per
// wave).  Every lane of the wave must call it.
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t* counter) {
    const unsigned long long m = __ballot(pred ? 1 : 0);
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    const unsigned long long live = __ballot(1);
    const int leader = __ffsll((long long)live) - 1;
    uint32_t base = 0;
    if ((int)__lane_id() == leader && m != 0ull) base = gatomic_add(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    return base + below;
}

// ---- kernels ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_frame_init(uint32_t P, uint32_t seed, uint32_t* __restrict__ seeds,
                                                       float4* __restrict__ pix, float4* gpos, float4* gnorm,
                                                       float4* galb) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    seeds[p] = pt::tea(p, seed);
    pix[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 z = make_float4(0.f, 0.f, 0.f, 1.f);
    if (gpos) gpos[p] = z;
    if (gnorm) gnorm[p] = z;
    if (galb) galb[p] = z;
}

struct CamArgs {
    float iv[16], ip[16];
    uint32_t w, h, frame, qmc, sample, samples, max_frame;
    uint32_t p0, np;  // the pixel group [p0, p0 + np)
};

__global__ __launch_bounds__(kBlock) void k_camera(CamArgs c, uint32_t* __restrict__ seeds, float4* __restrict__ rays,
                                                   uint32_t* __restrict__ ids, float4* __restrict__ thrmis,
                                                   float4* __restrict__ accs) {
    const uint32_t p = c.p0 + blockIdx.x * kBlock + threadIdx.x;
    if (p >= c.p0 + c.np) return;
    const uint32_t x = p % c.w, y = p / c.w;
    uint32_t seed = seeds[p];
    float jx, jy;
    if (c.qmc == 0) {
        if (c.frame == 0) {
            jx = 0.5f;
            jy = 0.5f;
        } else {
            jx = pt::randf(seed);
            jy = pt::randf(seed);
        }
    } else {
        const uint32_t i = c.sample + c.samples * c.frame, N = c.samples * c.max_frame;
        jx = (float)i / (float)N;
        jy = pt::radical_inverse(i);
    }
    const float ux = ((float)x + jx) / (float)c.w;
    const float uy = ((float)y + jy) / (float)c.h;
    const float4 target = pt::xform4(c.ip, ux * 2.0f - 1.0f, uy * 2.0f - 1.0f, 0.0f, 1.0f);
    const float4 dir = pt::xform4(c.iv, target.x, target.y, target.z, 0.0f);
    const float4 org = pt::xform4(c.iv, 0.0f, 0.0f, 0.0f, 1.0f);
    const V3 d = pt::normalize(v3(dir.x, dir.y, dir.z));
    const uint32_t slot = p - c.p0;  // rays / ids point at the group's list
    rays[2 * (size_t)slot] = make_float4(org.x, org.y, org.z, kEps);
    rays[2 * (size_t)slot + 1] = make_float4(d.x, d.y, d.z, kLargeDist);
    ids[slot] = p;
    thrmis[p] = make_float4(1.f, 1.f, 1.f, 1.f);
    accs[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    seeds[p] = seed;
}

struct WaveArgs {
    SceneDev sc;
    float clear[3], env[3];
    int depth, max_depth, sample, use_rr, use_metalness;
    const uint32_t* n_in;  // live paths (rays_in / ids_in entries)
    const uint32_t* ids_in;
    const float4* rays_in;
    const int32_t* hobj;
    const int32_t* hprim;
    const float* huv;
    uint32_t* seeds;
    float4* thrmis;
    float4* accs;
    float4* rays_out;
    uint32_t* ids_out;
    float4* sh_rays;
    uint32_t* sh_ids;
    float4* sh_contrib;
    uint32_t* n_next;    // appended next rays
    uint32_t* n_shadow;  // appended shadow / light rays
    float4* gpos;
    float4* gnorm;
    float4* galb;
};

// One bounce of every live path: INTEG 0 direct (integrate_direct), 1 material sampling
// (integrate_mats), 2 MIS (integrate_mis); BRDF 0 Blinn-Phong, 1 GGX.
template <int INTEG, int BRDF>
__global__ __launch_bounds__(kShadeBlock) void k_shade(WaveArgs a) {
    const uint32_t n = *a.n_in;
    if (blockIdx.x * kShadeBlock >= n) return;  // whole block past the live paths
    const uint32_t i = blockIdx.x * kShadeBlock + threadIdx.x;
    bool next = false, shadow = false;
    V3 npos = v3(0.f, 0.f, 0.f), ndir = npos, sdir = npos, contrib = npos;
    float stmax = 0.0f;
    uint32_t pid = 0;
    if (i < n) {
        pid = a.ids_in[i];
        const float4 r1 = a.rays_in[2 * (size_t)i + 1];
        const V3 d = v3(r1.x, r1.y, r1.z);
        uint32_t seed = a.seeds[pid];
        const float4 tm = a.thrmis[pid];
        V3 thr = v3(tm.x, tm.y, tm.z);
        float mis = tm.w;
        const float4 ac = a.accs[pid];
        V3 acc = v3(ac.x, ac.y, ac.z);
        const int32_t obj = a.hobj[i];
        if (obj < 0) {
            // miss: the clear colour for camera rays, environment light after a bounce
            if (a.depth == 0)
                acc = v3(a.clear[0], a.clear[1], a.clear[2]);
            else
                acc = pt::add(acc, pt::mul(v3(a.env[0], a.env[1], a.env[2]), thr));
        } else {
            const SceneDev& sc = a.sc;
            const ObjDev& O = sc.objs[obj];
            const uint32_t prim = (uint32_t)a.hprim[i];
            const float u = a.huv[2 * (size_t)i], v = a.huv[2 * (size_t)i + 1];
            const float b0 = (1.0f - u) - v;
            // hit_info (rt.rgen:62-95): bary = (1-u-v, u, v) on the triangle's vertices
            const uint32_t* ix = sc.idx + O.ibase + 3 * (size_t)prim;
            const size_t q0 = 3 * (size_t)(O.vbase + ix[0]), q1 = 3 * (size_t)(O.vbase + ix[1]),
                         q2 = 3 * (size_t)(O.vbase + ix[2]);
            const V3 n = pt::comb3(xyz(sc.verts[q0 + 1]), b0, xyz(sc.verts[q1 + 1]), u, xyz(sc.verts[q2 + 1]), v);
            const V3 normal = pt::normalize(pt::xform_vec3(O.nit, n));
            const V3 lp = pt::comb3(xyz(sc.verts[q0]), b0, xyz(sc.verts[q1]), u, xyz(sc.verts[q2]), v);
            const V3 pos = pt::xform_point(O.m, lp);
            // mat_info (rt.rgen:97-133) from the object's factors
            pt::Mat m;
            m.albedo = v3(O.albedo[0], O.albedo[1], O.albedo[2]);
            m.emissive = v3(O.emissive[0], O.emissive[1], O.emissive[2]);
            m.roughness = O.mr[1];
            if (a.use_metalness) {
                const float mt = O.mr[0];
                m.albedo = pt::add(v3(0.04f * (1.0f - mt), 0.04f * (1.0f - mt), 0.04f * (1.0f - mt)),
                                   pt::scale(m.albedo, mt));
            }
            // shade_info (rt.rgen:135-150): face the normal against the incoming ray
            pt::Shade s;
            s.wo = d;
            s.N = dot(s.wo, normal) > 0.0f ? pt::neg(normal) : normal;
            pt::make_tanspace(s.N, s.T, s.B);
            if (a.sample == 0 && a.depth == 0 && a.gpos) {
                a.gpos[pid] = make_float4(pos.x, pos.y, pos.z, 1.f);
                a.gnorm[pid] = make_float4(s.N.x, s.N.y, s.N.z, 1.f);
                a.galb[pid] = make_float4(m.albedo.x, m.albedo.y, m.albedo.z, 1.f);
            }
            bool term = false;
            const bool emissive = pt::any_positive(m.emissive);
            if (INTEG == 0) {
                // integrate_direct (rt.rgen:393-414)
                term = true;
                if (emissive) {
                    acc = pt::add(acc, m.emissive);
                } else if (m.roughness != 0.0f) {
                    if (sc.n_lights > 0) {
                        const LightSample ls = light_sample(sc, seed, pos);
                        const V3 wi = pt::normalize(sub(ls.pos, pos));
                        const V3 atten = pt::mat_eval<BRDF>(m, s, wi);
                        if (ls.pdf != 0.0f) {
                            const V3 dir = sub(ls.pos, pos);
                            const float dl = pt::length(dir);
                            sdir = pt::divs(dir, dl);
                            stmax = dl - kEps;
                            contrib = pt::mul(pt::divs(atten, ls.pdf), ls.emissive);
                            shadow = true;
                        }
                    } else {
                        for (int k = 0; k < 4; k++) (void)pt::lcg(seed);  // the draws of light_sample
                    }
                }
            } else if (INTEG == 1) {
                // integrate_mats (rt.rgen:362-391)
                if (emissive) {
                    acc = pt::add(acc, pt::mul(m.emissive, thr));
                    term = true;
                } else if (m.roughness == 0.0f) {
                    ndir = pt::reflect(s.wo, s.N);
                    thr = pt::mul(thr, m.albedo);
                } else {
                    V3 wi;
                    if (!pt::mat_sample<BRDF>(seed, m, s, wi)) {
                        term = true;
                    } else {
                        const float pdf = pt::mat_pdf<BRDF>(m, s, wi);
                        const V3 atten = pt::mat_eval<BRDF>(m, s, wi);
                        if (pdf != 0.0f) {
                            thr = pt::mul(thr, pt::divs(atten, pdf));
                            ndir = wi;
                        } else {
                            term = true;
                        }
                    }
                }
            } else {
                // integrate_mis (rt.rgen:303-360)
                if (emissive) {
                    acc = pt::add(acc, pt::mul(pt::scale(thr, mis), m.emissive));
                    term = true;
                } else if (m.roughness == 0.0f) {
                    ndir = pt::reflect(s.wo, s.N);
                    thr = pt::mul(thr, m.albedo);
                    mis = 1.0f;
                } else {
                    V3 wl = v3(0.f, 0.f, 0.f);
                    float pdf_ll = 0.0f;
                    if (sc.n_lights > 0) {
                        wl = light_sample_dir(sc, seed, pos);
                        pdf_ll = light_pdf(sc, pos, wl);
                    } else {
                        for (int k = 0; k < 4; k++) (void)pt::lcg(seed);
                    }
                    if (pdf_ll != 0.0f) {
                        const float pdf_lm = pt::mat_pdf<BRDF>(m, s, wl);
                        const V3 atten = pt::mat_eval<BRDF>(m, s, wl);
                        const V3 weight = pt::scale(pt::divs(atten, pdf_ll), pt::power_heuristic(pdf_ll, pdf_lm));
                        contrib = pt::mul(thr, weight);
                        sdir = wl;
                        stmax = kLargeDist;
                        shadow = true;
                    }
                    V3 wb;
                    if (!pt::mat_sample<BRDF>(seed, m, s, wb)) {
                        term = true;
                    } else {
                        const float pdf_bm = pt::mat_pdf<BRDF>(m, s, wb);
                        if (pdf_bm != 0.0f) {
                            const float pdf_bl = light_pdf(sc, pos, wb);
                            const V3 atten = pt::mat_eval<BRDF>(m, s, wb);
                            thr = pt::mul(thr, pt::divs(atten, pdf_bm));
                            mis = pt::power_heuristic(pdf_bm, pdf_bl);
                            ndir = wb;
                        } else {
                            term = true;
                        }
                    }
                }
            }
            // Russian roulette after every hit (rt.rgen:660-665)
            bool alive = !term;
            if (a.use_rr) {
                const float pcont = fminf(fmaxf(thr.x, fmaxf(thr.y, thr.z)) + 0.001f, 0.95f);
                if (pt::randf(seed) >= pcont)
                    alive = false;
                else
                    thr = pt::divs(thr, pcont);
            }
            next = alive && a.depth + 1 < a.max_depth;
            npos = pos;
        }
        a.seeds[pid] = seed;
        a.thrmis[pid] = make_float4(thr.x, thr.y, thr.z, mis);
        a.accs[pid] = make_float4(acc.x, acc.y, acc.z, 0.f);
    }
    // shadow rays start at the hit point too
    const uint32_t ns = wave_append(next, a.n_next);
    const uint32_t ss = wave_append(shadow, a.n_shadow);
    if (next) {
        a.rays_out[2 * (size_t)ns] = make_float4(npos.x, npos.y, npos.z, kEps);
        a.rays_out[2 * (size_t)ns + 1] = make_float4(ndir.x, ndir.y, ndir.z, kLargeDist);
        a.ids_out[ns] = pid;
    }
    if (shadow) {
        a.sh_rays[2 * (size_t)ss] = make_float4(npos.x, npos.y, npos.z, kEps);
        a.sh_rays[2 * (size_t)ss + 1] = make_float4(sdir.x, sdir.y, sdir.z, stmax);
        a.sh_ids[ss] = pid;
        a.sh_contrib[ss] = make_float4(contrib.x, contrib.y, contrib.z, 0.f);
    }
}

// direct lighting: acc += contrib * (visible ? 1 : 0)
__global__ __launch_bounds__(kBlock) void k_resolve_shadow(const uint32_t* __restrict__ n_dev,
                                                           const uint32_t* __restrict__ ids,
                                                           const float4* __restrict__ contrib,
                                                           const uint8_t* __restrict__ occ, float4* accs) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= *n_dev) return;
    const uint32_t pid = ids[i];
    const float sh = occ[i] ? 0.0f : 1.0f;
    const float4 c = contrib[i];
    float4 a = accs[pid];
    a.x = a.x + c.x * sh;
    a.y = a.y + c.y * sh;
    a.z = a.z + c.z * sh;
    accs[pid] = a;
}

// MIS light rays: acc += contrib * (emissive of the closest hit, or the environment)
__global__ __launch_bounds__(kBlock) void k_resolve_light(const uint32_t* __restrict__ n_dev,
                                                          const uint32_t* __restrict__ ids,
                                                          const float4* __restrict__ contrib,
                                                          const int32_t* __restrict__ obj,
                                                          const ObjDev* __restrict__ objs, float ex, float ey,
                                                          float ez, float4* accs) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= *n_dev) return;
    const uint32_t pid = ids[i];
    const int32_t o = obj[i];
    float lx = ex, ly = ey, lz = ez;
    if (o >= 0) {
        lx = objs[o].emissive[0];
        ly = objs[o].emissive[1];
        lz = objs[o].emissive[2];
    }
    const float4 c = contrib[i];
    float4 a = accs[pid];
    a.x = a.x + c.x * lx;
    a.y = a.y + c.y * ly;
    a.z = a.z + c.z * lz;
    accs[pid] = a;
}

__global__ __launch_bounds__(kBlock) void k_sum_sample(uint32_t P, const float4* __restrict__ accs,
                                                       float4* __restrict__ pix) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    const float4 a = accs[p];
    float4 s = pix[p];
    s.x = s.x + a.x;
    s.y = s.y + a.y;
    s.z = s.z + a.z;
    pix[p] = s;
}

// avg = sum / samples; frame > 0 blends mix(old, avg, 1/(frame+1)) (rt.rgen:673-681)
__global__ __launch_bounds__(kBlock) void k_resolve(uint32_t P, uint32_t samples, uint32_t frame,
                                                    const float4* __restrict__ pix, float4* image) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    const float4 s = pix[p];
    const float ns = (float)samples;
    float4 avg = make_float4(s.x / ns, s.y / ns, s.z / ns, 1.0f);
    if (frame > 0) {
        const float w = 1.0f / (float)(frame + 1);
        const float4 o = image[p];
        avg.x = o.x * (1.0f - w) + avg.x * w;
        avg.y = o.y * (1.0f - w) + avg.y * w;
        avg.z = o.z * (1.0f - w) + avg.z * w;
    }
    image[p] = avg;
}

__global__ __launch_bounds__(kBlock) void k_light_pdf(uint32_t n, const float* __restrict__ p,
                                                      const float* __restrict__ d, float* __restrict__ out,
                                                      SceneDev sc) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    out[i] = light_pdf(sc, v3(p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]),
                       v3(d[3 * (size_t)i], d[3 * (size_t)i + 1], d[3 * (size_t)i + 2]));
}

// world positions of the light triangles (the FP32 model * vec4(v, 1) of rt.rgen:180-182),
// three float4 per light triangle, plus the flat xyz copy the light BVH is built over
__global__ __launch_bounds__(kBlock) void k_light_world(uint32_t n, const uint2* __restrict__ src,
                                                        const ObjDev* __restrict__ objs,
                                                        const float4* __restrict__ verts,
                                                        const uint32_t* __restrict__ idx, float4* __restrict__ lworld,
                                                        float* __restrict__ flat, uint32_t* __restrict__ flat_idx) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const uint2 st = src[k];  // (object, triangle)
    const ObjDev& O = objs[st.x];
    for (int c = 0; c < 3; c++) {
        const uint32_t vi = idx[O.ibase + 3 * (size_t)st.y + c];
        const V3 w = pt::xform_point(O.m, xyz(verts[3 * (size_t)(O.vbase + vi)]));
        lworld[3 * (size_t)k + c] = make_float4(w.x, w.y, w.z, 0.f);
        flat[9 * (size_t)k + 3 * c] = w.x;
        flat[9 * (size_t)k + 3 * c + 1] = w.y;
        flat[9 * (size_t)k + 3 * c + 2] = w.z;
        flat_idx[3 * (size_t)k + c] = 3 * k + c;
    }
}

// bad[0] |= 1 if some index of [idx, idx + n) is >= n_verts
__global__ __launch_bounds__(kBlock) void k_check_indices(size_t n, const uint32_t* __restrict__ idx, uint32_t n_verts,
                                                          uint32_t* bad) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (idx[i] >= n_verts) gatomic_or(bad, 1u);
}

// per-sample list lengths: counts[2d] = rays of bounce d (bounce 0: every pixel),
// counts[2d + 1] = shadow / light rays of bounce d
__global__ void k_sample_counts(uint32_t* counts, uint32_t n_entries, uint32_t P) {
    for (uint32_t k = threadIdx.x; k < n_entries; k += blockDim.x) counts[k] = k == 0 ? P : 0u;
}

// ray statistics: tally[0] += bounce rays, tally[1] += shadow / light rays of one sample
__global__ void k_tally(const uint32_t* counts, uint32_t depth, unsigned long long* tally) {
    if (threadIdx.x != 0) return;
    unsigned long long b = 0, sh = 0;
    for (uint32_t d = 0; d < depth; d++) {
        b += counts[2 * d];
        sh += counts[2 * d + 1];
    }
    atomicAdd(&tally[0], b);  // several pixel groups add concurrently
    atomicAdd(&tally[1], sh);
}

struct DevScope {
    int prev = -1;
    bool ok = false;
    explicit DevScope(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        ok = hipSetDevice(d) == hipSuccess;
    }
    ~DevScope() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// 3x3 inverse transpose in FP64 (rows), false if singular or not finite
bool inverse_transpose3(const float* m16, float out[9]) {
    const double a[3][3] = {{m16[0], m16[1], m16[2]}, {m16[4], m16[5], m16[6]}, {m16[8], m16[9], m16[10]}};
    double c[3][3];
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 3; k++) {
            const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, k1 = (k + 1) % 3, k2 = (k + 2) % 3;
            c[r][k] = a[r1][k1] * a[r2][k2] - a[r1][k2] * a[r2][k1];
        }
    const double det = a[0][0] * c[0][0] + a[0][1] * c[0][1] + a[0][2] * c[0][2];
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
    // inverse = cofactor^T / det, so its transpose is cofactor / det
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 3; k++) out[3 * r + k] = (float)(c[r][k] / det);
    return true;
}

}  // namespace
}  // namespace gpurt_impl

using namespace gpurt_impl;

namespace {

// One pixel group's streams and events (the main chain on `st`; shadow / light rays on
// `aux`), traversal scratch for each, and list lengths (two sets of kCountSlots).
struct Lane {
    hipStream_t st = nullptr, aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, done[2] = {nullptr, nullptr}, end = nullptr;
    void* trav = nullptr;
    void* trav2 = nullptr;
    uint32_t* counts = nullptr;
    bool create() {
        return hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
               hipStreamCreateWithFlags(&aux, hipStreamNonBlocking) == hipSuccess &&
               hipEventCreateWithFlags(&fork, hipEventDisableTiming) == hipSuccess &&
               hipEventCreateWithFlags(&join, hipEventDisableTiming) == hipSuccess &&
               hipEventCreateWithFlags(&done[0], hipEventDisableTiming) == hipSuccess &&
               hipEventCreateWithFlags(&done[1], hipEventDisableTiming) == hipSuccess &&
               hipEventCreateWithFlags(&end, hipEventDisableTiming) == hipSuccess &&
               hipMalloc((void**)&counts, 2 * kCountSlots * sizeof(uint32_t)) == hipSuccess;
    }
    void free_scratch() {
        for (void** q : {&trav, &trav2})
            if (*q) (void)hipFree(*q);
        trav = trav2 = nullptr;
    }
    void release() {
        if (st) (void)hipStreamSynchronize(st);
        if (aux) (void)hipStreamSynchronize(aux);
        free_scratch();
        if (counts) (void)hipFree(counts);
        counts = nullptr;
        for (hipEvent_t e : {fork, join, done[0], done[1], end})
            if (e) (void)hipEventDestroy(e);
        fork = join = done[0] = done[1] = end = nullptr;
        for (hipStream_t q : {st, aux})
            if (q) (void)hipStreamDestroy(q);
        st = aux = nullptr;
    }
};

}  // namespace

struct gpurt_scene_s {
    gpurt_ctx ctx = nullptr;
    int device = 0;  // the context's device (destroy does not read the context)
    gpurt_tlas tlas = nullptr;
    size_t n_objs = 0;
    uint32_t n_lights = 0, n_ltris = 0;
    float n_tris0 = 1.0f;
    ObjDev* objs = nullptr;
    float4* verts = nullptr;
    uint32_t* idx = nullptr;
    LightDev* lights = nullptr;
    float4* lworld = nullptr;
    float* ltri_n = nullptr;
    gpurt_bvh lbvh = nullptr;
    // wavefront state (per pixel and per list slot), grown on demand
    size_t cap = 0;
    void* wf = nullptr;
    // pixel groups rendered concurrently, each on its own streams with its own traversal
    // scratch and list lengths
    std::vector<Lane> lanes;
    size_t lane_cap = 0;  // pixels per group the traversal scratch is sized for
    size_t trav_bytes = 0;
    hipEvent_t ev_start = nullptr;
    unsigned long long* tally = nullptr;  // device [2]: rays traced (ray_counts)
    ~gpurt_scene_s() {
        if (lbvh) gpurt_bvh_destroy(lbvh);
        for (void* p : {(void*)objs, (void*)verts, (void*)idx, (void*)lights, (void*)lworld, (void*)ltri_n, wf,
                        (void*)tally})
            if (p) (void)hipFree(p);
        for (Lane& l : lanes) l.release();
        if (ev_start) (void)hipEventDestroy(ev_start);
    }
    SceneDev dev() const {
        SceneDev s{};
        s.objs = objs;
        s.verts = verts;
        s.idx = idx;
        s.lights = lights;
        s.n_lights = n_lights;
        s.n_ltris = n_ltris;
        s.lworld = lworld;
        s.lnodes = lbvh ? (const WNode<4>*)lbvh->nodes : nullptr;
        s.ltris = lbvh ? lbvh->tris : nullptr;
        s.ltri_n = ltri_n;
        s.n_tris0 = n_tris0;
        return s;
    }
};

namespace {

// carve of the wavefront buffers for P pixels
struct Wave {
    uint32_t* seeds;
    float4 *thrmis, *accs, *pix;
    float4* rays[2];
    uint32_t* ids[2];
    int32_t *hobj, *hprim;
    float *ht, *huv;
    float4 *sh_rays, *sh_contrib;
    uint32_t* sh_ids;
    uint8_t* occ;
    int32_t *sh_obj, *sh_prim;  // closest hits of the MIS light rays
    float* sh_t;
    // direct lighting keeps two samples in flight: the second set of the per-sample
    // buffers its shadow-ray stage reads
    float4 *accs2, *sh_rays2, *sh_contrib2;
    uint32_t* sh_ids2;
    uint8_t* occ2;
};

size_t carve(void* base, size_t P, Wave* w) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        void* p = base ? (char*)base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    Wave t;
    t.seeds = (uint32_t*)take(4 * P);
    t.thrmis = (float4*)take(16 * P);
    t.accs = (float4*)take(16 * P);
    t.pix = (float4*)take(16 * P);
    for (int k = 0; k < 2; k++) {
        t.rays[k] = (float4*)take(32 * P);
        t.ids[k] = (uint32_t*)take(4 * P);
    }
    t.hobj = (int32_t*)take(4 * P);
    t.hprim = (int32_t*)take(4 * P);
    t.ht = (float*)take(4 * P);
    t.huv = (float*)take(8 * P);
    t.sh_rays = (float4*)take(32 * P);
    t.sh_contrib = (float4*)take(16 * P);
    t.sh_ids = (uint32_t*)take(4 * P);
    t.occ = (uint8_t*)take(P);
    t.sh_obj = (int32_t*)take(4 * P);
    t.sh_prim = (int32_t*)take(4 * P);
    t.accs2 = (float4*)take(16 * P);
    t.sh_rays2 = (float4*)take(32 * P);
    t.sh_contrib2 = (float4*)take(16 * P);
    t.sh_ids2 = (uint32_t*)take(4 * P);
    t.occ2 = (uint8_t*)take(P);
    t.sh_t = (float*)take(4 * P);
    if (w) *w = t;
    return off;
}

template <int I, int B>
void launch_shade(const WaveArgs& a, uint32_t capacity, hipStream_t s) {
    hipLaunchKernelGGL((k_shade<I, B>), dim3(cdiv(capacity, kShadeBlock)), dim3(kShadeBlock), 0, s, a);
}

}  // namespace

// One pixel group [p0, p0 + np) of a frame, every sample of it, on the group's streams.
// The group's ray lists are the slots [p0, p0 + np) of the list arrays; per-pixel state is
// indexed by pixel.  Samples of a pixel run in order (the shader's RNG order).
static int render_group(gpurt_scene S, Lane& L, const gpurt_render_params& r, const Wave& w, WaveArgs a, CamArgs c,
                        uint32_t p0, uint32_t np) {
    hipStream_t s = L.st;
    c.p0 = p0;
    c.np = np;
    const size_t o = p0;  // list offset of this group
    float4* rays[2] = {w.rays[0] + 2 * o, w.rays[1] + 2 * o};
    uint32_t* ids[2] = {w.ids[0] + o, w.ids[1] + o};
    int32_t* hobj = w.hobj + o;
    int32_t* hprim = w.hprim + o;
    float* huv = w.huv + 2 * o;
    a.hobj = hobj;
    a.hprim = hprim;
    a.huv = huv;
    int rc = GPURT_OK;
    if (r.integrator == 0 && r.max_depth >= 1) {
        // Direct lighting ends every path at its first hit, so a sample is camera rays ->
        // closest hit -> shade -> shadow rays.  Sample s's shadow rays, their
        // contributions and its pixel sum run on `aux` while sample s+1's camera rays are
        // traced and shaded (the next sample only needs the RNG states the shading left);
        // the per-sample buffers the shadow stage reads alternate between two sets.
        for (int smp = 0; smp < r.samples && rc == GPURT_OK; smp++) {
            const int par = smp & 1;
            float4* accs = par ? w.accs2 : w.accs;
            float4* shr = (par ? w.sh_rays2 : w.sh_rays) + 2 * o;
            float4* shc = (par ? w.sh_contrib2 : w.sh_contrib) + o;
            uint32_t* shi = (par ? w.sh_ids2 : w.sh_ids) + o;
            uint8_t* occ = (par ? w.occ2 : w.occ) + o;
            uint32_t* cnt = L.counts + par * kCountSlots;
            if (smp >= 2) GPURT_HIP(hipStreamWaitEvent(s, L.done[par], 0));  // sample smp-2 is done with them
            c.sample = (uint32_t)smp;
            hipLaunchKernelGGL(k_sample_counts, dim3(1), dim3(64), 0, s, cnt, 4u, np);
            hipLaunchKernelGGL(k_camera, dim3(cdiv(np, kBlock)), dim3(kBlock), 0, s, c, w.seeds, rays[0], ids[0],
                               w.thrmis, accs);
            rc = tlas_trace_ex(S->tlas, (const gpurt_ray*)rays[0], np, cnt, hobj, hprim, w.ht + o, huv, nullptr,
                               false, nullptr, s, L.trav, S->trav_bytes);
            if (rc) break;
            a.depth = 0;
            a.sample = smp;
            a.n_in = cnt;
            a.n_next = cnt + 2;
            a.n_shadow = cnt + 1;
            a.ids_in = ids[0];
            a.rays_in = rays[0];
            a.rays_out = rays[1];
            a.ids_out = ids[1];
            a.accs = accs;
            a.sh_rays = shr;
            a.sh_ids = shi;
            a.sh_contrib = shc;
            if (r.brdf == 0)
                launch_shade<0, 0>(a, np, s);
            else
                launch_shade<0, 1>(a, np, s);
            if (hipGetLastError() != hipSuccess || hipEventRecord(L.fork, s) != hipSuccess ||
                hipStreamWaitEvent(L.aux, L.fork, 0) != hipSuccess) {
                rc = hip_error(hipGetLastError(), "shade launch");
                break;
            }
            rc = tlas_trace_ex(S->tlas, (const gpurt_ray*)shr, np, cnt + 1, nullptr, nullptr, nullptr, nullptr, occ,
                               true, nullptr, L.aux, L.trav2, S->trav_bytes);
            if (rc) break;
            hipLaunchKernelGGL(k_resolve_shadow, dim3(cdiv(np, kBlock)), dim3(kBlock), 0, L.aux, cnt + 1, shi, shc, occ,
                               accs);
            if (r.ray_counts) hipLaunchKernelGGL(k_tally, dim3(1), dim3(64), 0, L.aux, cnt, 1u, S->tally);
            hipLaunchKernelGGL(k_sum_sample, dim3(cdiv(np, kBlock)), dim3(kBlock), 0, L.aux, np, accs + o, w.pix + o);
            if (hipGetLastError() != hipSuccess || hipEventRecord(L.done[par], L.aux) != hipSuccess) {
                rc = hip_error(hipGetLastError(), "shadow stage");
                break;
            }
        }
        (void)hipStreamWaitEvent(s, L.done[0], 0);
        (void)hipStreamWaitEvent(s, L.done[1], 0);
        return rc;
    }
    a.accs = w.accs;
    a.sh_rays = w.sh_rays + 2 * o;
    a.sh_ids = w.sh_ids + o;
    a.sh_contrib = w.sh_contrib + o;
    const int bounces = r.max_depth;
    const uint32_t n_counts = 2u * (uint32_t)r.max_depth + 2u;
    // MIS: the light rays of bounce d are traced on `aux` while bounce d+1's rays are
    // traced on `s`; their contributions are added before bounce d+1 is shaded (the
    // shader's order of additions)
    auto resolve_light = [&](int d) {
        GPURT_HIP(hipStreamWaitEvent(s, L.join, 0));
        hipLaunchKernelGGL(k_resolve_light, dim3(cdiv(np, kBlock)), dim3(kBlock), 0, s, L.counts + 2 * d + 1,
                           w.sh_ids + o, w.sh_contrib + o, w.sh_obj + o, S->objs, r.env_light[0], r.env_light[1],
                           r.env_light[2], w.accs);
        return GPURT_OK;
    };
    for (int smp = 0; smp < r.samples && rc == GPURT_OK; smp++) {
        c.sample = (uint32_t)smp;
        hipLaunchKernelGGL(k_sample_counts, dim3(1), dim3(64), 0, s, L.counts, n_counts, np);
        hipLaunchKernelGGL(k_camera, dim3(cdiv(np, kBlock)), dim3(kBlock), 0, s, c, w.seeds, rays[0], ids[0], w.thrmis,
                           w.accs);
        int cur = 0;
        bool pending_light = false;
        for (int depth = 0; depth < bounces; depth++) {
            const uint32_t* n_in = L.counts + 2 * depth;
            uint32_t* n_sh = L.counts + 2 * depth + 1;
            rc = tlas_trace_ex(S->tlas, (const gpurt_ray*)rays[cur], np, n_in, hobj, hprim, w.ht + o, huv,
                               nullptr, false, nullptr, s, L.trav, S->trav_bytes);
            if (rc) break;
            if (pending_light) {
                if ((rc = resolve_light(depth - 1))) break;
                pending_light = false;
            }
            a.depth = depth;
            a.sample = smp;
            a.n_in = n_in;
            a.n_next = L.counts + 2 * (depth + 1);
            a.n_shadow = n_sh;
            a.ids_in = ids[cur];
            a.rays_in = rays[cur];
            a.rays_out = rays[cur ^ 1];
            a.ids_out = ids[cur ^ 1];
            switch (r.integrator * 2 + r.brdf) {
                case 0: launch_shade<0, 0>(a, np, s); break;
                case 1: launch_shade<0, 1>(a, np, s); break;
                case 2: launch_shade<1, 0>(a, np, s); break;
                case 3: launch_shade<1, 1>(a, np, s); break;
                case 4: launch_shade<2, 0>(a, np, s); break;
                default: launch_shade<2, 1>(a, np, s); break;
            }
            if (hipGetLastError() != hipSuccess) {
                rc = hip_error(hipGetLastError(), "shade launch");
                break;
            }
            if (r.integrator == 2) {
                if (hipEventRecord(L.fork, s) != hipSuccess || hipStreamWaitEvent(L.aux, L.fork, 0) != hipSuccess) {
                    rc = hip_error(hipGetLastError(), "light-ray stream fork");
                    break;
                }
                rc = tlas_trace_ex(S->tlas, (const gpurt_ray*)(w.sh_rays + 2 * o), np, n_sh, w.sh_obj + o,
                                   w.sh_prim + o, w.sh_t + o, nullptr, nullptr, false, nullptr, L.aux, L.trav2,
                                   S->trav_bytes);
                if (rc) break;
                if (hipEventRecord(L.join, L.aux) != hipSuccess) {
                    rc = hip_error(hipGetLastError(), "light-ray stream join");
                    break;
                }
                pending_light = true;
            } else if (r.integrator == 0) {  // direct lighting with max_depth 0 never gets here
                rc = tlas_trace_ex(S->tlas, (const gpurt_ray*)(w.sh_rays + 2 * o), np, n_sh, nullptr, nullptr,
                                   nullptr, nullptr, w.occ + o, true, nullptr, s, L.trav, S->trav_bytes);
                if (rc) break;
                hipLaunchKernelGGL(k_resolve_shadow, dim3(cdiv(np, kBlock)), dim3(kBlock), 0, s, n_sh, w.sh_ids + o,
                                   w.sh_contrib + o, w.occ + o, w.accs);
            }
            cur ^= 1;
        }
        if (rc == GPURT_OK && pending_light) rc = resolve_light(bounces - 1);
        if (rc) break;
        if (r.ray_counts) hipLaunchKernelGGL(k_tally, dim3(1), dim3(64), 0, s, L.counts, (uint32_t)bounces, S->tally);
        hipLaunchKernelGGL(k_sum_sample, dim3(cdiv(np, kBlock)), dim3(kBlock), 0, s, np, w.accs + o, w.pix + o);
    }
    (void)hipStreamWaitEvent(s, L.join, 0);  // no light-ray work outlives the group's stream order
    return rc;
}

extern "C" {

int gpurt_scene_create(gpurt_tlas tlas, const gpurt_object* objects, size_t n_objects, void* stream,
                       gpurt_scene* out) {
    if (!out) return set_error(GPURT_E_INVALID, "out is null");
    *out = nullptr;
    if (!tlas || !objects) return set_error(GPURT_E_INVALID, "null tlas or object array");
    if (n_objects == 0 || n_objects > ((size_t)1 << 24)) return set_error(GPURT_E_INVALID, "object count must be 1..2^24");
    gpurt_ctx ctx = nullptr;
    const std::vector<uint32_t>* custom = nullptr;
    const std::vector<gpurt_bvh_s*>* blas = nullptr;
    int rc = tlas_host_info(tlas, &ctx, &custom, &blas);
    if (rc) return rc;
    for (size_t i = 0; i < custom->size(); i++) {
        const uint32_t c = (*custom)[i];
        const std::string at = " (instance " + std::to_string(i) + ")";
        if (c >= n_objects) return set_error(GPURT_E_INVALID, "instance custom index is not an object index" + at);
        if ((uint64_t)(*blas)[i]->n_tris != objects[c].n_tris)
            return set_error(GPURT_E_INVALID, "instance BLAS and object triangle counts differ" + at);
    }
    std::vector<ObjDev> od(n_objects);
    std::vector<LightDev> lights;
    std::vector<uint2> lsrc;
    std::vector<float> ltri_n;
    uint64_t nv = 0, ni = 0;
    for (size_t i = 0; i < n_objects; i++) {
        const gpurt_object& o = objects[i];
        const std::string at = " (object " + std::to_string(i) + ")";
        if ((o.n_verts && !o.verts) || (o.n_tris && !o.indices))
            return set_error(GPURT_E_INVALID, "null vertex or index pointer" + at);
        if (o.n_tris && !o.n_verts) return set_error(GPURT_E_INVALID, "triangles without vertices" + at);
        for (int k = 0; k < 16; k++)
            if (!std::isfinite(o.model[k])) return set_error(GPURT_E_INVALID, "non-finite model matrix" + at);
        if (o.model[12] != 0.f || o.model[13] != 0.f || o.model[14] != 0.f || o.model[15] != 1.f)
            return set_error(GPURT_E_INVALID, "model matrix is not affine" + at);
        for (int k = 0; k < 3; k++)
            if (!std::isfinite(o.albedo[k]) || !std::isfinite(o.emissive[k]) || o.emissive[k] < 0.f)
                return set_error(GPURT_E_INVALID, "material factors must be finite, emission >= 0" + at);
        if (!std::isfinite(o.metal_rough[0]) || !std::isfinite(o.metal_rough[1]))
            return set_error(GPURT_E_INVALID, "non-finite metal/roughness" + at);
        ObjDev& d = od[i];
        for (int r = 0; r < 3; r++)
            for (int k = 0; k < 4; k++) d.m[4 * r + k] = o.model[4 * r + k];
        if (!inverse_transpose3(o.model, d.nit)) return set_error(GPURT_E_INVALID, "singular model matrix" + at);
        for (int k = 0; k < 3; k++) {
            d.albedo[k] = o.albedo[k];
            d.emissive[k] = o.emissive[k];
        }
        d.mr[0] = o.metal_rough[0];
        d.mr[1] = o.metal_rough[1];
        d.vbase = (uint32_t)nv;
        d.ibase = (uint32_t)ni;
        d.n_tris = o.n_tris;
        nv += o.n_verts;
        ni += 3 * (uint64_t)o.n_tris;
        if (nv >= ((uint64_t)1 << 30) || ni >= ((uint64_t)1 << 32) - 3)
            return set_error(GPURT_E_INVALID, "scene too large (2^30 vertices, 2^32 indices)");
        // lights: objects with emission (rt.cpp:47-60); objects without triangles emit nothing
        if ((o.emissive[0] != 0.f || o.emissive[1] != 0.f || o.emissive[2] != 0.f) && o.n_tris > 0) {
            LightDev L{(uint32_t)i, o.n_tris, (uint32_t)lsrc.size(), 0u};
            lights.push_back(L);
            for (uint32_t t = 0; t < o.n_tris; t++) {
                lsrc.push_back(make_uint2((uint32_t)i, t));
                ltri_n.push_back((float)o.n_tris);
            }
            if (lsrc.size() >= ((size_t)1 << 28)) return set_error(GPURT_E_INVALID, "too many light triangles");
        }
    }
    DevScope dg(ctx->device);
    if (!dg.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    hipStream_t s = (hipStream_t)stream;
    auto* S = new gpurt_scene_s();
    S->ctx = ctx;
    S->device = ctx->device;
    S->tlas = tlas;
    S->n_objs = n_objects;
    S->n_lights = (uint32_t)lights.size();
    S->n_ltris = (uint32_t)lsrc.size();
    S->n_tris0 = lights.empty() ? 1.0f : (float)lights[0].n_tris;
    const size_t nl = lsrc.size();
    float* flat = nullptr;
    uint32_t* flat_idx = nullptr;
    uint2* src_d = nullptr;
    uint32_t* bad = nullptr;
    auto fail = [&](int code) {
        for (void* p : {(void*)flat, (void*)flat_idx, (void*)src_d, (void*)bad})
            if (p) (void)hipFree(p);
        delete S;
        return code;
    };
    if (hipMalloc((void**)&S->objs, sizeof(ObjDev) * n_objects) != hipSuccess ||
        hipMalloc((void**)&S->verts, 48 * std::max<uint64_t>(nv, 1)) != hipSuccess ||
        hipMalloc((void**)&S->idx, 4 * std::max<uint64_t>(ni, 1)) != hipSuccess ||
        hipMalloc((void**)&S->tally, 2 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc((void**)&bad, sizeof(uint32_t)) != hipSuccess ||
        (nl && (hipMalloc((void**)&S->lights, sizeof(LightDev) * lights.size()) != hipSuccess ||
                hipMalloc((void**)&S->lworld, 48 * nl) != hipSuccess ||
                hipMalloc((void**)&S->ltri_n, 4 * nl) != hipSuccess ||
                hipMalloc((void**)&flat, 36 * nl) != hipSuccess ||
                hipMalloc((void**)&flat_idx, 12 * nl) != hipSuccess ||
                hipMalloc((void**)&src_d, sizeof(uint2) * nl) != hipSuccess))) {
        (void)hipGetLastError();
        return fail(set_error(GPURT_E_OOM, "scene allocation failed"));
    }
    rc = GPURT_OK;
    if (hipMemcpy(S->objs, od.data(), sizeof(ObjDev) * n_objects, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemsetAsync(bad, 0, sizeof(uint32_t), s) != hipSuccess)
        rc = hip_error(hipGetLastError(), "scene upload");
    for (size_t i = 0; i < n_objects && rc == GPURT_OK; i++) {
        const gpurt_object& o = objects[i];
        if (o.n_verts &&
            hipMemcpyAsync((char*)S->verts + 48 * (size_t)od[i].vbase, o.verts, 48 * (size_t)o.n_verts,
                           hipMemcpyDeviceToDevice, s) != hipSuccess)
            rc = hip_error(hipGetLastError(), "vertex copy");
        if (rc == GPURT_OK && o.n_tris) {
            if (hipMemcpyAsync(S->idx + od[i].ibase, o.indices, 12 * (size_t)o.n_tris, hipMemcpyDeviceToDevice, s) !=
                hipSuccess) {
                rc = hip_error(hipGetLastError(), "index copy");
            } else {
                hipLaunchKernelGGL(k_check_indices, dim3(cdiv(3 * (size_t)o.n_tris, kBlock)), dim3(kBlock), 0, s,
                                   3 * (size_t)o.n_tris, S->idx + od[i].ibase, o.n_verts, bad);
            }
        }
    }
    uint32_t bad_h = 0;
    if (rc == GPURT_OK && (hipMemcpyAsync(&bad_h, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
                           hipStreamSynchronize(s) != hipSuccess))
        rc = hip_error(hipGetLastError(), "index check");
    if (rc == GPURT_OK && bad_h) rc = set_error(GPURT_E_INVALID, "a triangle index is >= its object's vertex count");
    if (rc == GPURT_OK && nl) {
        if (hipMemcpy(S->lights, lights.data(), sizeof(LightDev) * lights.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(S->ltri_n, ltri_n.data(), 4 * nl, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(src_d, lsrc.data(), sizeof(uint2) * nl, hipMemcpyHostToDevice) != hipSuccess) {
            rc = hip_error(hipGetLastError(), "light upload");
        } else {
            hipLaunchKernelGGL(k_light_world, dim3(cdiv(nl, kBlock)), dim3(kBlock), 0, s, (uint32_t)nl, src_d, S->objs,
                               S->verts, S->idx, S->lworld, flat, flat_idx);
            gpurt_build_opts bo{};
            bo.node_width = 4;
            bo.leaf_size = 4;
            bo.node_format = kFormatF32;
            rc = gpurt_bvh_build(ctx, flat, 12, 3 * nl, flat_idx, nl, &bo, stream, &S->lbvh);
            if (rc == GPURT_OK && 3 * S->lbvh->depth + 1 > kLightStack)
                rc = set_error(GPURT_E_UNSUPPORTED, "light BVH too deep for the light_pdf stack");
        }
    }
    if (rc == GPURT_OK && hipStreamSynchronize(s) != hipSuccess) rc = hip_error(hipGetLastError(), "scene build");
    if (rc) return fail(rc);
    for (void* p : {(void*)flat, (void*)flat_idx, (void*)src_d, (void*)bad}) (void)hipFree(p);
    *out = S;
    return GPURT_OK;
}

void gpurt_scene_destroy(gpurt_scene scene) {
    if (!scene) return;
    DevScope dg(scene->device);
    (void)hipDeviceSynchronize();
    delete scene;
}

int gpurt_scene_light_pdf(gpurt_scene S, const float* p, const float* d, size_t n, float* pdf, void* stream) {
    if (!S) return set_error(GPURT_E_INVALID, "null scene");
    if (n == 0) return GPURT_OK;
    if (!p || !d || !pdf) return set_error(GPURT_E_INVALID, "null point/direction/output pointer");
    if (n >= ((size_t)1 << 32)) return set_error(GPURT_E_INVALID, "too many queries");
    DevScope dg(S->ctx->device);
    if (!dg.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    hipLaunchKernelGGL(k_light_pdf, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, (hipStream_t)stream, (uint32_t)n, p, d,
                       pdf, S->dev());
    GPURT_HIP(hipGetLastError());
    return GPURT_OK;
}

int gpurt_render(gpurt_scene S, const gpurt_render_params* prm, float* image, float* gbuf_pos, float* gbuf_norm,
                 float* gbuf_albedo, void* stream) {
    if (!S || !prm || !image) return set_error(GPURT_E_INVALID, "null scene, parameters or image");
    const gpurt_render_params& r = *prm;
    if (r.width < 1 || r.height < 1 || (int64_t)r.width * r.height > ((int64_t)1 << 28))
        return set_error(GPURT_E_INVALID, "image size must be 1..2^28 pixels");
    if (r.samples < 1 || r.samples > 65536) return set_error(GPURT_E_INVALID, "samples must be 1..65536");
    if (r.max_depth < 0 || r.max_depth > 64) return set_error(GPURT_E_INVALID, "max_depth must be 0..64");
    if (r.frame < 0) return set_error(GPURT_E_INVALID, "frame must be >= 0");
    if (r.integrator == 3 || r.integrator == 4)
        return set_error(GPURT_E_UNSUPPORTED, "ReSTIR integrators are not supported");
    if (r.integrator < 0 || r.integrator > 2) return set_error(GPURT_E_INVALID, "integrator must be 0, 1 or 2");
    if (r.brdf < 0 || r.brdf > 1) return set_error(GPURT_E_INVALID, "brdf must be 0 or 1");
    if (r.qmc && r.max_frame < 1) return set_error(GPURT_E_INVALID, "qmc needs max_frame >= 1");
    if (r.concurrency < 0 || r.concurrency > 8) return set_error(GPURT_E_INVALID, "concurrency must be 0..8");
    if ((gbuf_pos != nullptr) != (gbuf_norm != nullptr) || (gbuf_pos != nullptr) != (gbuf_albedo != nullptr))
        return set_error(GPURT_E_INVALID, "give all three G-buffer images or none");
    for (const float* q : {image, gbuf_pos, gbuf_norm, gbuf_albedo})
        if (((uintptr_t)q & 15u) != 0) return set_error(GPURT_E_INVALID, "images must be 16-byte aligned");
    for (int k = 0; k < 16; k++)
        if (!std::isfinite(r.inv_view[k]) || !std::isfinite(r.inv_proj[k]))
            return set_error(GPURT_E_INVALID, "non-finite camera matrix");
    DevScope dg(S->ctx->device);
    if (!dg.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    hipStream_t s = (hipStream_t)stream;
    const uint32_t P = (uint32_t)r.width * (uint32_t)r.height;
    const int G = std::max(1, std::min<int>(r.concurrency > 0 ? r.concurrency : kDefaultGroups[r.integrator],
                                            (int)std::min<uint32_t>(P, 8u)));
    if (!S->ev_start && hipEventCreateWithFlags(&S->ev_start, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(GPURT_E_DEVICE, "creating a HIP event failed");
    }
    while ((int)S->lanes.size() < G) {
        S->lanes.emplace_back();
        if (!S->lanes.back().create()) {
            (void)hipGetLastError();
            S->lanes.back().release();
            S->lanes.pop_back();
            return set_error(GPURT_E_DEVICE, "creating the streams of a pixel group failed");
        }
    }
    const uint32_t per = (uint32_t)cdiv(P, G);  // pixels per group (the last may have fewer)
    if (P > S->cap || per > S->lane_cap) {
        GPURT_HIP(hipStreamSynchronize(s));
        for (Lane& l : S->lanes) {
            GPURT_HIP(hipStreamSynchronize(l.st));
            GPURT_HIP(hipStreamSynchronize(l.aux));
            l.free_scratch();
        }
        if (S->wf) (void)hipFree(S->wf);
        S->wf = nullptr;
        S->cap = S->lane_cap = 0;
        S->trav_bytes = tlas_scratch_bytes(S->tlas, per);
        if (S->trav_bytes == 0) return set_error(GPURT_E_INVALID, "cannot plan the traversal");
        bool ok = hipMalloc(&S->wf, carve(nullptr, P, nullptr)) == hipSuccess;
        for (Lane& l : S->lanes)
            ok = ok && hipMalloc(&l.trav, S->trav_bytes) == hipSuccess && hipMalloc(&l.trav2, S->trav_bytes) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            for (Lane& l : S->lanes) l.free_scratch();
            if (S->wf) (void)hipFree(S->wf);
            S->wf = nullptr;
            return set_error(GPURT_E_OOM, "wavefront state allocation failed");
        }
        S->cap = P;
        S->lane_cap = per;
    }
    for (int g = 0; g < G; g++)  // groups created since the last sizing
        if (!S->lanes[g].trav && (hipMalloc(&S->lanes[g].trav, S->trav_bytes) != hipSuccess ||
                                  hipMalloc(&S->lanes[g].trav2, S->trav_bytes) != hipSuccess)) {
            (void)hipGetLastError();
            return set_error(GPURT_E_OOM, "traversal scratch allocation failed");
        }
    Wave w;
    carve(S->wf, S->cap, &w);
    hipLaunchKernelGGL(k_frame_init, dim3(cdiv(P, kBlock)), dim3(kBlock), 0, s, P, r.seed, w.seeds, w.pix,
                       (float4*)gbuf_pos, (float4*)gbuf_norm, (float4*)gbuf_albedo);
    if (r.ray_counts) GPURT_HIP(hipMemsetAsync(S->tally, 0, 2 * sizeof(unsigned long long), s));
    GPURT_HIP(hipEventRecord(S->ev_start, s));
    CamArgs c{};
    for (int k = 0; k < 16; k++) {
        c.iv[k] = r.inv_view[k];
        c.ip[k] = r.inv_proj[k];
    }
    c.w = (uint32_t)r.width;
    c.h = (uint32_t)r.height;
    c.frame = (uint32_t)r.frame;
    c.qmc = r.qmc ? 1u : 0u;
    c.samples = (uint32_t)r.samples;
    c.max_frame = (uint32_t)std::max(r.max_frame, 1);
    WaveArgs a{};
    a.sc = S->dev();
    for (int k = 0; k < 3; k++) {
        a.clear[k] = r.clear_col[k];
        a.env[k] = r.env_light[k];
    }
    a.max_depth = r.max_depth;
    a.use_rr = r.use_rr ? 1 : 0;
    a.use_metalness = r.use_metalness ? 1 : 0;
    a.seeds = w.seeds;
    a.thrmis = w.thrmis;
    a.gpos = (float4*)gbuf_pos;
    a.gnorm = (float4*)gbuf_norm;
    a.galb = (float4*)gbuf_albedo;
    int rc = GPURT_OK;
    for (int g = 0; g < G && rc == GPURT_OK; g++) {
        Lane& L = S->lanes[g];
        const uint32_t p0 = (uint32_t)g * per, np = std::min(P, p0 + per) - std::min(P, p0);
        if (np == 0) continue;
        if (hipStreamWaitEvent(L.st, S->ev_start, 0) != hipSuccess) {
            rc = hip_error(hipGetLastError(), "pixel group fork");
            break;
        }
        rc = render_group(S, L, r, w, a, c, p0, np);
        if (hipEventRecord(L.end, L.st) != hipSuccess && rc == GPURT_OK) rc = hip_error(hipGetLastError(), "join");
    }
    for (int g = 0; g < G; g++) (void)hipStreamWaitEvent(s, S->lanes[g].end, 0);
    tlas_note_use(S->tlas, s);  // every traversal of the frame is ordered before this point
    if (rc) return rc;
    hipLaunchKernelGGL(k_resolve, dim3(cdiv(P, kBlock)), dim3(kBlock), 0, s, P, (uint32_t)r.samples, (uint32_t)r.frame,
                       w.pix, (float4*)image);
    GPURT_HIP(hipGetLastError());
    if (r.ray_counts) {  // the only wait of a frame, and only when statistics are asked for
        unsigned long long t[2] = {0, 0};
        GPURT_HIP(hipMemcpyAsync(t, S->tally, sizeof(t), hipMemcpyDeviceToHost, s));
        GPURT_HIP(hipStreamSynchronize(s));
        r.ray_counts[0] += t[0];
        r.ray_counts[1] += t[1];
    }
    return GPURT_OK;
}

}  // extern "C"
