// accel_demo.cpp -- the C++ host layer (include/gpurt.hpp) end to end: a Mesh with the
// reference's 48-byte vertices, an Accel built from it, closest-hit / any-hit rays and
// closest-point queries with known answers.  Prints "accel_demo ok" and exits 0 when
// every answer matches, else prints the first mismatch and exits 1.
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/gpurt.hpp"

using gpurt::Mesh;

static int fail(const char* what, int i) {
    std::printf("accel_demo mismatch: %s at %d\n", what, i);
    return 1;
}

int main() {
    try {
        gpurt::Context ctx(0);
        // unit square in the z = 0 plane, split along its diagonal, plus a second square
        // at z = -1 under it
        std::vector<Mesh::Vertex> v(8);
        const float xy[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
        for (int k = 0; k < 8; k++) {
            v[k].pos = {xy[k % 4][0], xy[k % 4][1], k < 4 ? 0.0f : -1.0f, 0.0f};
            v[k].norm = {0, 0, 1, 0};
        }
        std::vector<Mesh::Index> idx = {0, 1, 2, 1, 3, 2, 4, 5, 6, 5, 7, 6};
        Mesh mesh(std::move(v), std::move(idx));
        gpurt::Accel accel(ctx, mesh);
        if (accel.info().n_tris != 4) return fail("n_tris", 0);

        // rays straight down from z = 1 at (x, y); hits the upper square at t = 1
        const int n = 4;
        std::vector<gpurt_ray> rays(n);
        const float px[n][2] = {{0.25f, 0.25f}, {0.75f, 0.75f}, {0.5f, 0.25f}, {2.0f, 2.0f}};
        for (int i = 0; i < n; i++) rays[i] = gpurt_ray{{px[i][0], px[i][1], 1.0f}, 1e-5f, {0, 0, -1.0f}, 1e7f};
        gpurt::DeviceArray<gpurt_ray> d_rays;
        d_rays.upload(rays.data(), n);
        gpurt::DeviceArray<int32_t> prim(n);
        gpurt::DeviceArray<float> t(n), uv(2 * n);
        gpurt::DeviceArray<uint8_t> occ(n);
        accel.trace_closest(d_rays.data(), n, prim.data(), t.data(), uv.data());
        // visibility with tmax just short of the plane: nothing is hit
        for (auto& r : rays) r.tmax = 0.999f;
        gpurt::DeviceArray<gpurt_ray> d_short;
        d_short.upload(rays.data(), n);
        accel.trace_any(d_short.data(), n, occ.data());
        gpurt::check_hip(hipDeviceSynchronize(), "sync");
        std::vector<int32_t> hp(n);
        std::vector<float> ht(n);
        std::vector<uint8_t> ho(n);
        prim.download(hp.data());
        t.download(ht.data());
        occ.download(ho.data());
        const int expect_prim[n] = {0, 1, 0, -1};
        for (int i = 0; i < n; i++) {
            if (hp[i] != expect_prim[i]) return fail("closest-hit prim", i);
            if (hp[i] >= 0 && ht[i] != 1.0f) return fail("closest-hit t", i);
            if (ho[i] != 0) return fail("any-hit", i);
        }

        // closest point of (0.25, 0.25, 0.5): on triangle 0 at distance 0.5
        std::vector<float> q = {0.25f, 0.25f, 0.5f, 0.25f, 0.25f, -3.0f};
        gpurt::DeviceArray<float> d_q;
        d_q.upload(q.data(), q.size());
        gpurt::DeviceArray<int32_t> cprim(2);
        gpurt::DeviceArray<float> pt(6), dist(2), bary(6);
        accel.closest_point(d_q.data(), 2, cprim.data(), pt.data(), dist.data(), bary.data());
        gpurt::check_hip(hipDeviceSynchronize(), "sync");
        std::vector<int32_t> cp(2);
        std::vector<float> cd(2);
        cprim.download(cp.data());
        dist.download(cd.data());
        if (cp[0] != 0 || cd[0] != 0.5f) return fail("closest point", 0);
        if (cp[1] != 2 || cd[1] != 2.0f) return fail("closest point", 1);

        // errors are exceptions, not exits
        bool threw = false;
        try {
            gpurt_build_opts bad{3, 4, 0, 0};
            gpurt::Accel a2(ctx, mesh, &bad);
        } catch (const gpurt::Error& e) {
            threw = e.code() == GPURT_E_UNSUPPORTED;
        }
        if (!threw) return fail("error path", 0);

        // autotuner: a small space over these rays and queries, stored in a cache file and
        // read back; a BVH built with the tuned options gives the same answers
        char cache[] = "/tmp/gpurt_tune_XXXXXX";
        const int fd = mkstemp(cache);
        if (fd < 0) return fail("cache file", 0);
        close(fd);
        const int32_t widths[] = {2, 4}, stacks[] = {8, 16}, formats[] = {1, 2};
        gpurt_tune_space sp{};
        sp.node_width = widths;
        sp.n_node_width = 2;
        sp.stack_depth = stacks;
        sp.n_stack_depth = 2;
        sp.node_format = formats;
        sp.n_node_format = 2;
        const int32_t zero[] = {0};
        sp.treelet_size = zero;
        sp.n_treelet_size = 1;
        sp.waves_per_cu = zero;
        sp.n_waves_per_cu = 1;
        sp.cache_path = cache;
        for (auto& r : rays) r.tmax = 1e7f;
        d_rays.upload(rays.data(), n);
        const gpurt_tune_result tuned = gpurt::tune_queries(ctx, mesh, d_q.data(), 2, d_rays.data(), n, &sp);
        if (tuned.candidates != 8 || tuned.from_cache) return fail("tuner candidates", tuned.candidates);
        gpurt_tune_result loaded;
        if (!gpurt::load_tuned(cache, tuned.key, loaded)) return fail("tuner cache load", 0);
        if (loaded.build.node_width != tuned.build.node_width || loaded.query.stack_depth != tuned.query.stack_depth ||
            loaded.build.node_format != tuned.build.node_format || !loaded.from_cache)
            return fail("tuner cache contents", 0);
        const gpurt_tune_result again = gpurt::tune_queries(ctx, mesh, d_q.data(), 2, d_rays.data(), n, &sp);
        if (!again.from_cache || again.candidates != 0) return fail("tuner cache hit", 0);
        std::remove(cache);
        gpurt::Accel tuned_accel(ctx, mesh, &loaded.build);
        tuned_accel.trace_closest(d_rays.data(), n, prim.data(), t.data(), uv.data(), &loaded.query);
        gpurt::check_hip(hipDeviceSynchronize(), "sync");
        prim.download(hp.data());
        for (int i = 0; i < n; i++)
            if (hp[i] != expect_prim[i]) return fail("tuned closest-hit prim", i);
        std::printf("accel_demo ok (tuned: width %d, format %d, stack %d; %.3f ms vs %.3f ms default)\n",
                    loaded.build.node_width, loaded.build.node_format, loaded.query.stack_depth, tuned.best_ms,
                    tuned.default_ms);
        return 0;
    } catch (const std::exception& e) {
        std::printf("accel_demo exception: %s\n", e.what());
        return 1;
    }
}
