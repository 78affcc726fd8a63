// gpurt.hpp -- C++ host layer over the C ABI (gpurt.h), shaped like the reference's
// VK::Mesh / VK::Accel so a caller of those can switch with few edits.
//
//   gpurt::Mesh   mirrors VK::Mesh (src/vk/mesh.h:14-58): owns host vertex/index arrays
//                 (48-byte Vertex {pos, norm, tang} with u in pos.w and v in norm.w,
//                 scene.cpp:310) and their device copies, computes the bounding box on
//                 recreate (mesh.cpp:73-75) and uploads lazily in sync() (mesh.cpp:78-85).
//   gpurt::Accel  mirrors VK::Accel's bottom-level constructor (vulkan.h:256-284,
//                 vulkan.cpp:881-936): move-only, recreate(mesh), destroy(), a public
//                 handle `accel`.  Queries replace traceRayEXT (rt.rgen:257-291).
//
// Errors throw gpurt::Error (the reference calls die()/exit, log.h:46-49; a library must
// not end the process).  Query arguments are device pointers; calls are asynchronous on
// the given stream.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gpurt.h"

namespace gpurt {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc) {
    if (rc != GPURT_OK) throw Error(rc, gpurt_last_error());
}

inline void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(GPURT_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

struct Vec4 {
    float x = 0, y = 0, z = 0, w = 0;
};

struct BBox {  // lib/bbox.h:15-42 semantics: empty = (FLT_MAX, -FLT_MAX)
    float min[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float max[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void enclose(const float* p) {
        for (int k = 0; k < 3; k++) {
            min[k] = std::min(min[k], p[k]);
            max[k] = std::max(max[k], p[k]);
        }
    }
};

class Context {
public:
    explicit Context(int device = 0) { check(gpurt_ctx_create(device, &h_)); }
    ~Context() { gpurt_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    Context(Context&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    Context& operator=(Context&& o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    gpurt_ctx handle() const { return h_; }

private:
    gpurt_ctx h_ = nullptr;
};

template <typename T>
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { resize(n); }
    ~DeviceArray() { reset(); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
    DeviceArray& operator=(DeviceArray&& o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    void resize(size_t n) {
        if (n == n_) return;
        reset();
        if (n) check_hip(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "hipMalloc");
        n_ = n;
    }
    void reset() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void upload(const T* src, size_t n) {
        resize(n);
        if (n) check_hip(hipMemcpy(p_, src, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    void download(T* dst) const {
        if (n_) check_hip(hipMemcpy(dst, p_, n_ * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
    T* data() const { return p_; }
    size_t size() const { return n_; }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// Mirrors VK::Mesh.
class Mesh {
public:
    typedef uint32_t Index;
    struct Vertex {
        Vec4 pos;   // xyz = position, w = texture u
        Vec4 norm;  // xyz = normal, w = texture v
        Vec4 tang;
    };
    static_assert(sizeof(Vertex) == 48, "Mesh::Vertex is 48 bytes like the reference's");

    Mesh() = default;
    Mesh(std::vector<Vertex>&& vertices, std::vector<Index>&& indices) {
        recreate(std::move(vertices), std::move(indices));
    }
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;
    Mesh(Mesh&&) = default;
    Mesh& operator=(Mesh&&) = default;

    void recreate(std::vector<Vertex>&& vertices, std::vector<Index>&& indices) {
        if (indices.size() % 3 != 0) throw Error(GPURT_E_INVALID, "index count is not a multiple of 3");
        verts_ = std::move(vertices);
        idxs_ = std::move(indices);
        dirty_ = true;
        BBox box;
        for (auto& v : verts_) box.enclose(&v.pos.x);
        bbox_ = box;
    }
    // upload if the host arrays changed since the last upload
    void sync() const {
        if (!dirty_) return;
        vbuf_.upload(verts_.data(), verts_.size());
        ibuf_.upload(idxs_.data(), idxs_.size());
        dirty_ = false;
    }
    const std::vector<Vertex>& verts() const { return verts_; }
    const std::vector<Index>& inds() const { return idxs_; }
    BBox bbox() const { return bbox_; }
    const float* device_verts() const {
        sync();
        return reinterpret_cast<const float*>(vbuf_.data());
    }
    const uint32_t* device_inds() const {
        sync();
        return ibuf_.data();
    }

private:
    std::vector<Vertex> verts_;
    std::vector<Index> idxs_;
    BBox bbox_;
    mutable DeviceArray<Vertex> vbuf_;
    mutable DeviceArray<Index> ibuf_;
    mutable bool dirty_ = false;
};

// Mirrors VK::Accel (bottom level).  `accel` is the public handle like
// VkAccelerationStructureKHR accel (vulkan.h:273).
class Accel {
public:
    Accel() = default;
    Accel(const Context& ctx, const Mesh& mesh, const gpurt_build_opts* opts = nullptr,
          hipStream_t stream = nullptr) {
        recreate(ctx, mesh, opts, stream);
    }
    ~Accel() { destroy(); }
    Accel(const Accel&) = delete;
    Accel& operator=(const Accel&) = delete;
    Accel(Accel&& o) noexcept : accel(std::exchange(o.accel, nullptr)) {}
    Accel& operator=(Accel&& o) noexcept {
        std::swap(accel, o.accel);
        return *this;
    }

    // Accel::recreate(const Mesh&) (vulkan.cpp:881-936): sync the mesh, build, wait.
    void recreate(const Context& ctx, const Mesh& mesh, const gpurt_build_opts* opts = nullptr,
                  hipStream_t stream = nullptr) {
        destroy();
        mesh.sync();
        check(gpurt_bvh_build(ctx.handle(), mesh.device_verts(), sizeof(Mesh::Vertex), mesh.verts().size(),
                              mesh.device_inds(), mesh.inds().size() / 3, opts, stream, &accel));
    }
    // moved vertices, same topology
    void refit(const Mesh& mesh, hipStream_t stream = nullptr) {
        mesh.sync();
        check(gpurt_bvh_refit(accel, mesh.device_verts(), sizeof(Mesh::Vertex), stream));
    }
    void destroy() {
        gpurt_bvh_destroy(accel);
        accel = nullptr;
    }
    gpurt_bvh_info info() const {
        gpurt_bvh_info i;
        check(gpurt_bvh_get_info(accel, &i));
        return i;
    }

    // trace_ray (rt.rgen:257-270): closest hit, payload = (prim, t, uv) per ray
    void trace_closest(const gpurt_ray* rays, size_t n, int32_t* prim, float* t, float* uv,
                       const gpurt_query_opts* opts = nullptr, hipStream_t stream = nullptr) const {
        check(gpurt_trace_closest(accel, rays, n, prim, t, uv, opts, stream));
    }
    // visibility (rt.rgen:272-291): occluded[i] = 1 when something is hit in [tmin, tmax]
    void trace_any(const gpurt_ray* rays, size_t n, uint8_t* occluded, const gpurt_query_opts* opts = nullptr,
                   hipStream_t stream = nullptr) const {
        check(gpurt_trace_any(accel, rays, n, occluded, opts, stream));
    }
    void closest_point(const float* q, size_t n, int32_t* prim, float* pt, float* dist, float* bary,
                       const gpurt_query_opts* opts = nullptr, hipStream_t stream = nullptr) const {
        check(gpurt_closest_point(accel, q, n, prim, pt, dist, bary, opts, stream));
    }

    gpurt_bvh accel = nullptr;
};

// Autotuner: the build and query options that run a workload fastest (gpurt_tune_queries
// over one mesh's closest-point queries and rays; gpurt_tune_batch over a batch), and the
// tuned-result cache a program reads at start-up instead of tuning again.  The reference
// sets these knobs by hand (RTPipe's host fields, src/vk/rt.h:38-53).
//
//   gpurt_tune_space sp{};  sp.cache_path = "tuned.txt";
//   gpurt_tune_result r = gpurt::tune_queries(ctx, mesh, q, nq, rays, nr, &sp);
//   gpurt::Accel accel(ctx, mesh, &r.build);
//   accel.trace_closest(rays, nr, prim, t, uv, &r.query);
inline gpurt_tune_result tune_queries(const Context& ctx, const Mesh& mesh, const float* q, size_t n_q,
                                      const gpurt_ray* rays, size_t n_r, const gpurt_tune_space* space = nullptr,
                                      hipStream_t stream = nullptr) {
    mesh.sync();
    gpurt_tune_result r;
    check(gpurt_tune_queries(ctx.handle(), mesh.device_verts(), sizeof(Mesh::Vertex), mesh.verts().size(),
                             mesh.device_inds(), mesh.inds().size() / 3, q, n_q, rays, n_r, space, stream, &r));
    return r;
}

inline gpurt_tune_result tune_batch(const Context& ctx, const std::vector<gpurt_batch_mesh>& meshes,
                                    const gpurt_tune_space* space = nullptr, hipStream_t stream = nullptr) {
    gpurt_tune_result r;
    check(gpurt_tune_batch(ctx.handle(), meshes.data(), meshes.size(), space, stream, &r));
    return r;
}

// The cached result for `key` (gpurt_tune_result::key of an earlier tuning run), or false.
inline bool load_tuned(const std::string& cache_path, const std::string& key, gpurt_tune_result& out) {
    return gpurt_tune_cache_load(cache_path.c_str(), key.c_str(), &out) == GPURT_OK;
}

// Mirrors the top-level VK::Accel(const std::vector<Drop<Accel>>&, const std::vector<Mat4>&)
// (vulkan.h:259, vulkan.cpp:777-856): instance i places blas[i] by the row-major 3x4
// object-to-world transform T[i] with custom index i and mask 0xFF.  The BLASes must
// outlive it.
class Tlas {
public:
    Tlas() = default;
    Tlas(const Context& ctx, const std::vector<const Accel*>& blas, const std::vector<std::array<float, 12>>& T,
         hipStream_t stream = nullptr) {
        recreate(ctx, blas, T, stream);
    }
    ~Tlas() { destroy(); }
    Tlas(const Tlas&) = delete;
    Tlas& operator=(const Tlas&) = delete;
    Tlas(Tlas&& o) noexcept : accel(std::exchange(o.accel, nullptr)) {}
    Tlas& operator=(Tlas&& o) noexcept {
        std::swap(accel, o.accel);
        return *this;
    }
    void recreate(const Context& ctx, const std::vector<const Accel*>& blas,
                  const std::vector<std::array<float, 12>>& T, hipStream_t stream = nullptr) {
        destroy();
        if (blas.size() != T.size()) throw Error(GPURT_E_INVALID, "one transform per BLAS");
        std::vector<gpurt_instance> inst(blas.size());
        for (size_t i = 0; i < blas.size(); i++) {
            std::copy(T[i].begin(), T[i].end(), inst[i].transform);
            inst[i].custom_index = (uint32_t)i;
            inst[i].mask = 0xFF;
            inst[i].blas = blas[i]->accel;
        }
        check(gpurt_tlas_build(ctx.handle(), inst.data(), inst.size(), nullptr, stream, &accel));
    }
    void destroy() {
        gpurt_tlas_destroy(accel);
        accel = nullptr;
    }
    // trace_ray through both levels: obj = custom index of the hit instance (-1: miss)
    void trace_closest(const gpurt_ray* rays, size_t n, int32_t* obj, int32_t* prim, float* t, float* uv,
                       const gpurt_query_opts* opts = nullptr, hipStream_t stream = nullptr) const {
        check(gpurt_tlas_trace_closest(accel, rays, n, obj, prim, t, uv, opts, stream));
    }
    void trace_any(const gpurt_ray* rays, size_t n, uint8_t* occluded, const gpurt_query_opts* opts = nullptr,
                   hipStream_t stream = nullptr) const {
        check(gpurt_tlas_trace_any(accel, rays, n, occluded, opts, stream));
    }

    gpurt_tlas accel = nullptr;
};

// Mirrors VK::RTPipe (src/vk/rt.h): build_desc + use_accel at construction, trace() per
// frame.  object i of `objects` is the TLAS instance with custom index i.
class RTPipe {
public:
    RTPipe() = default;
    RTPipe(const Tlas& tlas, const std::vector<gpurt_object>& objects, hipStream_t stream = nullptr) {
        check(gpurt_scene_create(tlas.accel, objects.data(), objects.size(), stream, &scene));
    }
    ~RTPipe() { gpurt_scene_destroy(scene); }
    RTPipe(const RTPipe&) = delete;
    RTPipe& operator=(const RTPipe&) = delete;
    RTPipe(RTPipe&& o) noexcept : scene(std::exchange(o.scene, nullptr)) {}
    RTPipe& operator=(RTPipe&& o) noexcept {
        std::swap(scene, o.scene);
        return *this;
    }
    // one frame into `image` (height x width float4, device); G-buffers optional
    void trace(const gpurt_render_params& p, float* image, float* gbuf_pos = nullptr, float* gbuf_norm = nullptr,
               float* gbuf_albedo = nullptr, hipStream_t stream = nullptr) const {
        check(gpurt_render(scene, &p, image, gbuf_pos, gbuf_norm, gbuf_albedo, stream));
    }

    gpurt_scene scene = nullptr;
};

}  // namespace gpurt
