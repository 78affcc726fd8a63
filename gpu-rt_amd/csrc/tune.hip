// tune.hip -- the autotuner (gpurt_tune_batch, gpurt_tune_queries) and its result cache.
//
// The traversal answers are exact for any tree and any launch shape, so the parameters
// that shape the tree (node width, treelet size, node encoding) and the launch (LDS
// stack depth, resident waves per CU) change speed only.  The tuner times every
// candidate on the caller's own workload and stream with HIP events (median of `reps`
// runs after one untimed run) and keeps the fastest; the library defaults are timed too,
// so the gain is reported.  Results can be kept in a text cache keyed by the device and
// by log2 buckets of the workload's sizes, read back by C++ callers
// (gpurt_tune_cache_load) without timing anything.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gpurt_internal.h"

namespace gpurt_impl {
namespace {

struct TuneDevGuard {
    int prev = -1;
    bool ok = false;
    explicit TuneDevGuard(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        ok = hipSetDevice(d) == hipSuccess;
    }
    ~TuneDevGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

std::vector<int32_t> values(const int32_t* v, int32_t n, std::vector<int32_t> dflt) {
    if (!v || n <= 0) return dflt;
    return std::vector<int32_t>(v, v + n);
}

int bucket(size_t n) {
    int b = 0;
    while (((size_t)2 << b) <= n) b++;
    return n ? b : -1;
}

std::string device_tag(int device) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) {
        (void)hipGetLastError();
        return "device" + std::to_string(device);
    }
    std::string arch = p.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
    return arch + "x" + std::to_string(p.multiProcessorCount);
}

void set_key(gpurt_tune_result* out, const std::string& k) {
    std::snprintf(out->key, sizeof(out->key), "%s", k.c_str());
}

// median of `reps` timed runs of fn on stream s, after one untimed run
template <typename Fn>
int time_median(Fn fn, int reps, hipStream_t s, float& ms) {
    int rc = fn();
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return hip_error(hipGetLastError(), "tuner events");
    }
    std::vector<float> t;
    for (int r = 0; r < reps && rc == GPURT_OK; r++) {
        float x = 0.f;
        if (hipEventRecord(e0, s) != hipSuccess) rc = hip_error(hipGetLastError(), "tuner event");
        if (rc == GPURT_OK) rc = fn();
        if (rc == GPURT_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                               hipEventElapsedTime(&x, e0, e1) != hipSuccess))
            rc = hip_error(hipGetLastError(), "tuner timing");
        t.push_back(x);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    std::sort(t.begin(), t.end());
    ms = t[t.size() / 2];
    return GPURT_OK;
}

struct Candidate {
    gpurt_build_opts b{};
    gpurt_query_opts q{};
};

}  // namespace
}  // namespace gpurt_impl

using namespace gpurt_impl;

extern "C" {

int gpurt_tune_cache_load(const char* path, const char* key, gpurt_tune_result* out) {
    if (!path || !key || !out) return set_error(GPURT_E_INVALID, "null argument");
    std::ifstream f(path);
    if (!f) return set_error(GPURT_E_INVALID, std::string("cannot read tuner cache ") + path);
    std::string line;
    bool found = false;
    gpurt_tune_result r{};
    while (std::getline(f, line)) {
        const size_t tab = line.find('\t');
        if (tab == std::string::npos || line.compare(0, tab, key) != 0 || tab != std::strlen(key)) continue;
        std::istringstream in(line.substr(tab + 1));
        gpurt_tune_result x{};
        if (in >> x.build.node_width >> x.build.leaf_size >> x.build.treelet_size >> x.build.treelet_rounds >>
            x.build.node_format >> x.query.stack_depth >> x.query.waves_per_cu >> x.query.leaf_trigger >>
            x.query.refill_trigger >> x.best_ms >> x.default_ms) {
            r = x;
            found = true;
        }
    }
    if (!found) return set_error(GPURT_E_INVALID, std::string("no tuned result for ") + key);
    r.from_cache = 1;
    r.candidates = 0;
    set_key(&r, key);
    *out = r;
    return GPURT_OK;
}

int gpurt_tune_cache_store(const char* path, const gpurt_tune_result* r) {
    if (!path || !r) return set_error(GPURT_E_INVALID, "null argument");
    if (!r->key[0] || std::memchr(r->key, '\t', sizeof(r->key)) || std::memchr(r->key, '\n', sizeof(r->key)))
        return set_error(GPURT_E_INVALID, "bad tuner cache key");
    std::ofstream f(path, std::ios::app);
    if (!f) return set_error(GPURT_E_INVALID, std::string("cannot write tuner cache ") + path);
    f << r->key << '\t' << r->build.node_width << ' ' << r->build.leaf_size << ' ' << r->build.treelet_size << ' '
      << r->build.treelet_rounds << ' ' << r->build.node_format << ' ' << r->query.stack_depth << ' '
      << r->query.waves_per_cu << ' ' << r->query.leaf_trigger << ' ' << r->query.refill_trigger << ' '
      << r->best_ms << ' ' << r->default_ms << '\n';
    if (!f) return set_error(GPURT_E_INVALID, std::string("cannot write tuner cache ") + path);
    return GPURT_OK;
}

// every combination, build options outermost (consecutive candidates share a build)
static std::vector<Candidate> candidates(const gpurt_tune_space* sp) {
    const std::vector<int32_t> W = values(sp ? sp->node_width : nullptr, sp ? sp->n_node_width : 0, {2, 4, 8});
    const std::vector<int32_t> S = values(sp ? sp->stack_depth : nullptr, sp ? sp->n_stack_depth : 0, {8, 16, 32, 64});
    const std::vector<int32_t> V = values(sp ? sp->waves_per_cu : nullptr, sp ? sp->n_waves_per_cu : 0, {0, 16, 24});
    const std::vector<int32_t> T = values(sp ? sp->treelet_size : nullptr, sp ? sp->n_treelet_size : 0, {0, 6});
    const std::vector<int32_t> F = values(sp ? sp->node_format : nullptr, sp ? sp->n_node_format : 0, {1, 2});
    std::vector<Candidate> out;
    for (int32_t w : W)
        for (int32_t t : T)
            for (int32_t f : F)
                for (int32_t s : S)
                    for (int32_t v : V) {
                        Candidate c;
                        c.b.node_width = w;
                        c.b.treelet_size = t;
                        c.b.node_format = f;
                        c.q.stack_depth = s;
                        c.q.waves_per_cu = v;
                        out.push_back(c);
                    }
    return out;
}

int gpurt_tune_batch(gpurt_ctx ctx, const gpurt_batch_mesh* meshes, size_t n_meshes, const gpurt_tune_space* sp,
                     void* stream, gpurt_tune_result* out) {
    if (!ctx || !meshes || !out) return set_error(GPURT_E_INVALID, "null context, mesh array or result");
    if (n_meshes == 0) return set_error(GPURT_E_INVALID, "no meshes");
    if (sp && sp->reps < 0) return set_error(GPURT_E_INVALID, "reps must be >= 0");
    TuneDevGuard dg(ctx->device);
    if (!dg.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    size_t nt = 0, nq = 0, nr = 0;
    for (size_t m = 0; m < n_meshes; m++) {
        nt += meshes[m].n_tris;
        nq += meshes[m].n_q;
        nr += meshes[m].n_r;
    }
    const std::string key = device_tag(ctx->device) + "|batch|m" + std::to_string(bucket(n_meshes)) + "|t" +
                            std::to_string(bucket(nt)) + "|q" + std::to_string(bucket(nq)) + "|r" +
                            std::to_string(bucket(nr));
    if (sp && sp->cache_path && gpurt_tune_cache_load(sp->cache_path, key.c_str(), out) == GPURT_OK) return GPURT_OK;
    const int reps = sp && sp->reps ? sp->reps : 3;
    hipStream_t s = (hipStream_t)stream;
    auto run_one = [&](const Candidate& c, float& ms) {
        gpurt_batch b = nullptr;
        int rc = gpurt_batch_create(ctx, meshes, n_meshes, &c.b, &c.q, &b);
        if (rc) return rc;
        rc = time_median([&] { return gpurt_batch_run(b, stream); }, reps, s, ms);
        gpurt_batch_destroy(b);
        return rc;
    };
    gpurt_tune_result r{};
    float dms = 0.f;
    int rc = run_one(Candidate{}, dms);
    if (rc) return rc;
    r.default_ms = dms;
    r.best_ms = INFINITY;
    for (const Candidate& c : candidates(sp)) {
        float ms = 0.f;
        if ((rc = run_one(c, ms))) return rc;
        r.candidates++;
        if (ms < r.best_ms) {
            r.best_ms = ms;
            r.build = c.b;
            r.query = c.q;
        }
    }
    set_key(&r, key);
    if (sp && sp->cache_path && (rc = gpurt_tune_cache_store(sp->cache_path, &r))) return rc;
    *out = r;
    return GPURT_OK;
}

int gpurt_tune_queries(gpurt_ctx ctx, const float* verts, size_t vstride_bytes, size_t n_verts, const uint32_t* tri_idx,
                       size_t n_tris, const float* q, size_t n_q, const gpurt_ray* rays, size_t n_r,
                       const gpurt_tune_space* sp, void* stream, gpurt_tune_result* out) {
    if (!ctx || !verts || !tri_idx || !out) return set_error(GPURT_E_INVALID, "null argument");
    if ((n_q && !q) || (n_r && !rays)) return set_error(GPURT_E_INVALID, "null query or ray pointer");
    if (n_q == 0 && n_r == 0) return set_error(GPURT_E_INVALID, "nothing to time: no queries and no rays");
    if (sp && (sp->reps < 0 || !(sp->builds_per_run >= 0.f)))
        return set_error(GPURT_E_INVALID, "reps and builds_per_run must be >= 0");
    TuneDevGuard dg(ctx->device);
    if (!dg.ok) return set_error(GPURT_E_DEVICE, "hipSetDevice failed");
    const float bpr = sp ? sp->builds_per_run : 0.f;
    char bb[32];
    std::snprintf(bb, sizeof(bb), "b%g", bpr);
    const std::string kind = std::string(n_q ? "cpq" : "") + (n_q && n_r ? "+" : "") + (n_r ? "ray" : "");
    const std::string key = device_tag(ctx->device) + "|" + kind + "|t" + std::to_string(bucket(n_tris)) + "|q" +
                            std::to_string(bucket(n_q)) + "|r" + std::to_string(bucket(n_r)) + "|" + bb;
    if (sp && sp->cache_path && gpurt_tune_cache_load(sp->cache_path, key.c_str(), out) == GPURT_OK) return GPURT_OK;
    const int reps = sp && sp->reps ? sp->reps : 3;
    hipStream_t s = (hipStream_t)stream;
    // outputs of the trial queries (stream-ordered, freed on the stream at the end)
    void* mem = nullptr;
    const size_t qbytes = n_q * (4 + 12 + 4), rbytes = n_r * (4 + 4 + 8);
    if (hipMallocAsync(&mem, std::max<size_t>(qbytes + rbytes, 256), s) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(GPURT_E_OOM, "tuner output allocation failed");
    }
    char* p = static_cast<char*>(mem);
    int32_t* qp = reinterpret_cast<int32_t*>(p);
    float* qpt = reinterpret_cast<float*>(p + 4 * n_q);
    float* qd = reinterpret_cast<float*>(p + 16 * n_q);
    int32_t* rp = reinterpret_cast<int32_t*>(p + qbytes);
    float* rt = reinterpret_cast<float*>(p + qbytes + 4 * n_r);
    float* ruv = reinterpret_cast<float*>(p + qbytes + 8 * n_r);
    auto query = [&](gpurt_bvh b, const gpurt_query_opts& qo) {
        int rc = GPURT_OK;
        if (n_q) rc = gpurt_closest_point(b, q, n_q, qp, qpt, qd, nullptr, &qo, stream);
        if (rc == GPURT_OK && n_r) rc = gpurt_trace_closest(b, rays, n_r, rp, rt, ruv, &qo, stream);
        return rc;
    };
    auto build = [&](const gpurt_build_opts& bo, gpurt_bvh& b, float& build_ms) {
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = gpurt_bvh_build(ctx, verts, vstride_bytes, n_verts, tri_idx, n_tris, &bo, stream, &b);
        build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return rc;
    };
    gpurt_tune_result r{};
    int rc = GPURT_OK;
    {
        gpurt_bvh b = nullptr;
        float bms = 0.f, qms = 0.f;
        rc = build(gpurt_build_opts{}, b, bms);
        if (rc == GPURT_OK) rc = time_median([&] { return query(b, gpurt_query_opts{}); }, reps, s, qms);
        gpurt_bvh_destroy(b);
        r.default_ms = qms + bpr * bms;
    }
    r.best_ms = INFINITY;
    const std::vector<Candidate> cands = candidates(sp);
    for (size_t i = 0; i < cands.size() && rc == GPURT_OK; i++) {
        // candidates are ordered build-major: one build per distinct build option set
        const Candidate& c = cands[i];
        if (i > 0 && std::memcmp(&cands[i - 1].b, &c.b, sizeof(c.b)) == 0) continue;
        gpurt_bvh b = nullptr;
        float bms = 0.f;
        if ((rc = build(c.b, b, bms))) break;
        for (size_t k = i; k < cands.size() && std::memcmp(&cands[k].b, &c.b, sizeof(c.b)) == 0 && rc == GPURT_OK; k++) {
            float qms = 0.f;
            rc = time_median([&] { return query(b, cands[k].q); }, reps, s, qms);
            r.candidates++;
            const float obj = qms + bpr * bms;
            if (rc == GPURT_OK && obj < r.best_ms) {
                r.best_ms = obj;
                r.build = c.b;
                r.query = cands[k].q;
            }
        }
        gpurt_bvh_destroy(b);
    }
    (void)hipFreeAsync(mem, s);
    if (rc) return rc;
    set_key(&r, key);
    if (sp && sp->cache_path && (rc = gpurt_tune_cache_store(sp->cache_path, &r))) return rc;
    *out = r;
    return GPURT_OK;
}

}  // extern "C"
