// tracer_demo.cpp -- the C++ host layer's ray-tracing pipeline (gpurt::Tlas, gpurt::RTPipe)
// end to end, with answers the integrators fix exactly: a camera that sees only an
// emitting wall renders every pixel as the wall's emission under each integrator (direct:
// acc += emissive; material sampling: emissive * throughput 1; MIS: throughput 1 * MIS
// weight 1 * emissive), a camera that sees nothing renders the clear colour, and the
// G-buffer holds the wall's depth.  Prints "tracer_demo ok" and exits 0 when every pixel
// matches, else prints the first mismatch and exits 1.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/gpurt.hpp"

using gpurt::Mesh;

static int fail(const char* what, int i, float got, float want) {
    std::printf("tracer_demo mismatch: %s at %d: %g != %g\n", what, i, got, want);
    return 1;
}

// a quad of half-size h in the z = 0 plane (normal +z), as Mesh::Vertex records
static Mesh quad(float h) {
    std::vector<Mesh::Vertex> v(4);
    const float xy[4][2] = {{-h, -h}, {h, -h}, {h, h}, {-h, h}};
    for (int k = 0; k < 4; k++) {
        v[k].pos = {xy[k][0], xy[k][1], 0.0f, 0.0f};
        v[k].norm = {0.0f, 0.0f, 1.0f, 0.0f};
    }
    return Mesh(std::move(v), std::vector<Mesh::Index>{0, 1, 2, 0, 2, 3});
}

int main() {
    try {
        gpurt::Context ctx(0);
        Mesh wall = quad(6.0f), tile = quad(1.0f);
        gpurt::Accel b0(ctx, wall), b1(ctx, tile);
        // the wall 5 units in front of the camera (-z), a small grey tile behind the camera
        const std::array<float, 12> T0 = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, -5};
        const std::array<float, 12> T1 = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 5};
        gpurt::Tlas tlas(ctx, {&b0, &b1}, {T0, T1});
        std::vector<gpurt_object> objs(2);
        const Mesh* meshes[2] = {&wall, &tile};
        const std::array<float, 12>* T[2] = {&T0, &T1};
        const float emit[2][3] = {{1.5f, 2.5f, 3.5f}, {0.0f, 0.0f, 0.0f}};
        for (int i = 0; i < 2; i++) {
            gpurt_object& o = objs[i];
            o.verts = reinterpret_cast<const float*>(meshes[i]->device_verts());
            o.indices = meshes[i]->device_inds();
            o.n_verts = (uint32_t)meshes[i]->verts().size();
            o.n_tris = (uint32_t)meshes[i]->inds().size() / 3;
            for (int k = 0; k < 12; k++) o.model[k] = (*T[i])[k];
            o.model[12] = o.model[13] = o.model[14] = 0.0f;
            o.model[15] = 1.0f;
            for (int c = 0; c < 3; c++) {
                o.albedo[c] = 0.5f;
                o.emissive[c] = emit[i][c];
            }
            o.metal_rough[0] = 0.0f;
            o.metal_rough[1] = 0.5f;
        }
        gpurt::RTPipe pipe(tlas, objs);

        const int w = 24, h = 16;
        gpurt_render_params p{};
        for (int k = 0; k < 16; k++) p.inv_view[k] = (k % 5 == 0) ? 1.0f : 0.0f;  // camera at the origin, looking -z
        const float t = 1.0f / std::tan(0.5f * 0.6f), aspect = (float)w / (float)h, znear = 0.1f;
        p.inv_proj[0] = aspect / t;
        p.inv_proj[5] = 1.0f / t;
        p.inv_proj[11] = -1.0f;
        p.inv_proj[14] = 1.0f / znear;
        p.clear_col[0] = 0.25f, p.clear_col[1] = 0.5f, p.clear_col[2] = 0.75f;
        p.width = w, p.height = h, p.samples = 3, p.max_frame = 16, p.max_depth = 4, p.use_rr = 1, p.seed = 7;

        gpurt::DeviceArray<float> img(4 * w * h), gpos(4 * w * h), gnorm(4 * w * h), galb(4 * w * h);
        std::vector<float> out(4 * w * h), pos(4 * w * h);
        for (int integ = 0; integ < 3; integ++) {
            for (int brdf = 0; brdf < 2; brdf++) {
                p.integrator = integ, p.brdf = brdf;
                pipe.trace(p, img.data(), gpos.data(), gnorm.data(), galb.data());
                gpurt::check_hip(hipDeviceSynchronize(), "render");
                img.download(out.data());
                gpos.download(pos.data());
                for (int i = 0; i < w * h; i++) {
                    for (int c = 0; c < 3; c++)
                        if (out[4 * i + c] != emit[0][c]) return fail("emission", i, out[4 * i + c], emit[0][c]);
                    if (out[4 * i + 3] != 1.0f) return fail("alpha", i, out[4 * i + 3], 1.0f);
                    if (pos[4 * i + 2] != -5.0f) return fail("G-buffer depth", i, pos[4 * i + 2], -5.0f);
                }
            }
        }
        // turn the camera around the y axis by 90 degrees: it sees neither object
        for (int k = 0; k < 16; k++) p.inv_view[k] = 0.0f;
        p.inv_view[2] = 1.0f, p.inv_view[5] = 1.0f, p.inv_view[8] = -1.0f, p.inv_view[15] = 1.0f;
        p.integrator = 2;
        pipe.trace(p, img.data());
        gpurt::check_hip(hipDeviceSynchronize(), "render");
        img.download(out.data());
        for (int i = 0; i < w * h; i++)
            for (int c = 0; c < 3; c++)
                if (out[4 * i + c] != p.clear_col[c]) return fail("clear colour", i, out[4 * i + c], p.clear_col[c]);
        std::printf("tracer_demo ok\n");
        return 0;
    } catch (const gpurt::Error& e) {
        std::printf("tracer_demo error %d: %s\n", e.code(), e.what());
        return 1;
    }
}
