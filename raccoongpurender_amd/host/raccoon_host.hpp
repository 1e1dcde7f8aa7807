// C++ host surface over the C ABI (include/raccoon_gpu.h), mirroring the reference's host
// classes for the wavefront path (namespace rt):
//   SceneManager            common/scene_manager.hpp:330-650 (geometry, materials, envmap,
//                           buildBVH, createBuffer/createMaterialBuffer/createEnvmapBuffer)
//   WavefrontLane           common/wavefront_path_tracing.hpp:339-1153 (setup, initialize, step,
//                           update_intermediate, merge_from, finalize_color, re_start,
//                           step_count, stat_avg_sample)
//   WavefrontPathTracing    common/wavefront_path_tracing.hpp:1159-1470 (continuous rendering with
//                           a step thread and an image thread, launch_fixed, re_start,
//                           create_color_image, onColorRecieved) for one GPU
//   StacklessBVH/ThreadedBVH  the two traversals of common/stackless_bvh.hpp and
//                           common/threaded_bvh.hpp as GPU ray casters over one built tree
// Scene ingestion takes arrays instead of Alembic objects (the reference's addPolymesh/addPoint
// read houdini_alembic objects). Where the reference asserts and aborts (RAC_ASSERT,
// common/raccoon_ocl.hpp:20-28) these classes throw rt::GpuError with the library's message.
#pragma once

#include <stdint.h>

#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <mutex>
#include <thread>
#include <sstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/raccoon_gpu.h"

namespace rt {

class GpuError : public std::runtime_error {
public:
    GpuError(int code, const std::string& what) : std::runtime_error(what), code(code) {}
    int code;
};

inline void check(int rc, const rgpu_ctx* ctx, const char* what) {
    if (rc == RGPU_OK) return;
    std::string msg = what;
    if (ctx) {
        const char* detail = rgpu_last_error_string(ctx);
        if (detail && *detail) msg += std::string(": ") + detail;
    }
    throw GpuError(rc, msg + " (status " + std::to_string(rc) + ")");
}

// One GPU: the OpenCLContext + OpenCLLane pair of common/raccoon_ocl.hpp for a single device.
class GpuContext {
public:
    explicit GpuContext(int device_index = 0) {
        rgpu_config cfg = {RGPU_ABI_VERSION, 0};
        check(rgpu_create(device_index, &cfg, &h_), nullptr, "rgpu_create");
    }
    ~GpuContext() {
        if (h_) rgpu_destroy(h_);
    }
    GpuContext(const GpuContext&) = delete;
    GpuContext& operator=(const GpuContext&) = delete;
    static int deviceCount() {
        int n = 0;
        return rgpu_device_count(&n) == RGPU_OK ? n : 0;
    }
    rgpu_ctx* get() const { return h_; }
    void synchronize() const { check(rgpu_synchronize(h_, nullptr), h_, "rgpu_synchronize"); }

private:
    rgpu_ctx* h_ = nullptr;
};

// RAII handle for the library's opaque objects.
template <class T, int (*Destroy)(T*)>
class Handle {
public:
    Handle() = default;
    explicit Handle(T* p) : p_(p) {}
    ~Handle() { reset(); }
    Handle(Handle&& o) noexcept : p_(o.p_) { o.p_ = nullptr; }
    Handle& operator=(Handle&& o) noexcept {
        if (this != &o) {
            reset();
            p_ = o.p_;
            o.p_ = nullptr;
        }
        return *this;
    }
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
    void reset() {
        if (p_) Destroy(p_);
        p_ = nullptr;
    }
    T* get() const { return p_; }
    explicit operator bool() const { return p_ != nullptr; }

private:
    T* p_ = nullptr;
};

using SceneBuffer = Handle<rgpu_scene, rgpu_scene_destroy>;
using MaterialBuffer = Handle<rgpu_materials, rgpu_materials_destroy>;
using EnvmapBuffer = Handle<rgpu_envmap, rgpu_envmap_destroy>;
using BVHHandle = Handle<rgpu_bvh, rgpu_bvh_destroy>;
using WavefrontHandle = Handle<rgpu_wavefront, rgpu_wavefront_destroy>;

// A second HIP stream of a context at the default priority (the reference's per-lane queues,
// wavefront_path_tracing.hpp:346-348).
class GpuStream {
public:
    explicit GpuStream(GpuContext& ctx) : ctx_(ctx) {
        check(rgpu_stream_create(ctx.get(), &s_), ctx.get(), "rgpu_stream_create");
    }
    ~GpuStream() {
        if (s_) rgpu_stream_destroy(ctx_.get(), s_);
    }
    GpuStream(const GpuStream&) = delete;
    GpuStream& operator=(const GpuStream&) = delete;
    rgpu_stream get() const { return s_; }
    void synchronize() const { check(rgpu_synchronize(ctx_.get(), s_), ctx_.get(), "rgpu_synchronize"); }

private:
    GpuContext& ctx_;
    rgpu_stream s_ = nullptr;
};

// Device memory of a context (an OpenCLBuffer).
class DeviceMemory {
public:
    DeviceMemory(GpuContext& ctx, size_t bytes) : ctx_(ctx), bytes_(bytes) {
        check(rgpu_alloc(ctx.get(), bytes, &p_), ctx.get(), "rgpu_alloc");
    }
    ~DeviceMemory() {
        if (p_) rgpu_free(ctx_.get(), p_);
    }
    DeviceMemory(const DeviceMemory&) = delete;
    DeviceMemory& operator=(const DeviceMemory&) = delete;
    void* get() const { return p_; }
    size_t bytes() const { return bytes_; }

private:
    GpuContext& ctx_;
    size_t bytes_ = 0;
    void* p_ = nullptr;
};

// 8-bit colour of the finalised image (RGBA8ValueType, wavefront_path_tracing.hpp:87-92).
struct RGBA8ValueType {
    uint8_t r, g, b, a;
};

// Camera setting of a render: the look-at inputs of standardCamera()
// (wavefront_path_tracing.hpp:103-118); lens radius 0.0045 as kernels/new_path.cl:61.
struct CameraSetting {
    float eye[3] = {0.0f, 0.0f, 5.0f};
    float target[3] = {0.0f, 0.0f, 0.0f};
    float up[3] = {0.0f, 1.0f, 0.0f};
    float fov_y_deg = 45.0f;
    float focus_distance = 0.0f;
    uint32_t width = 256, height = 256;
    float lens_radius = 0.0045f;
};

inline rgpu_camera standardCamera(const CameraSetting& c) {
    rgpu_camera cam;
    check(rgpu_camera_look_at(c.eye, c.target, c.up, c.fov_y_deg, c.focus_distance, c.width, c.height,
                              c.lens_radius, &cam),
          nullptr, "rgpu_camera_look_at");
    return cam;
}

// A material of the scene: its type (material_type.hpp) and index in the type's table.
struct MaterialRef {
    int32_t type = RGPU_MAT_LAMBERTIAN;
    int32_t index = 0;
};

// SceneManager (common/scene_manager.hpp:330-650): triangles (primitives 0..ntris-1) then
// spheres, one material per primitive, material tables per type, an optional envmap.
// Triangle mesh of a Wavefront OBJ file: `v` and `f` records (corners `i`, `i/t`, `i//n`,
// `i/t/n`; negative indices count back from the last vertex; polygons split into a fan
// (0, k, k+1)); other records are skipped. OBJ faces are counter-clockwise seen from the front;
// with `clockwise` the corners are reversed to addPolymesh's clockwise order. Coordinates are
// read as doubles and rounded to float. Throws GpuError(RGPU_ERR_INVALID_ARG) naming the line of
// a bad record.
struct ObjMesh {
    std::vector<float> points;     // xyz
    std::vector<uint32_t> indices;  // 3 per triangle
};
inline ObjMesh parse_obj(std::istream& in, bool clockwise = true) {
    ObjMesh m;
    std::string line;
    size_t ln = 0;
    auto bad = [&](const std::string& what) {
        return GpuError(RGPU_ERR_INVALID_ARG, "OBJ line " + std::to_string(ln) + ": " + what);
    };
    std::vector<uint32_t> corner;
    while (std::getline(in, line)) {
        ++ln;
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag)) continue;
        if (tag == "v") {
            std::string c[3];
            if (!(ss >> c[0] >> c[1] >> c[2])) throw bad("a vertex needs three coordinates");
            for (const std::string& s : c) {
                char* end = nullptr;
                const double d = std::strtod(s.c_str(), &end);
                if (end == s.c_str() || *end != 0) throw bad("bad vertex coordinate");
                m.points.push_back((float)d);
            }
        } else if (tag == "f") {
            corner.clear();
            std::string c;
            const long nv = (long)(m.points.size() / 3);
            while (ss >> c) {
                const std::string head = c.substr(0, c.find('/'));
                char* end = nullptr;
                const long i = std::strtol(head.c_str(), &end, 10);
                if (head.empty() || *end != 0) throw bad("bad face corner '" + c + "'");
                const long j = i > 0 ? i - 1 : nv + i;
                if (i == 0 || j < 0 || j >= nv) throw bad("vertex index " + head + " out of range");
                corner.push_back((uint32_t)j);
            }
            if (corner.size() < 3) throw bad("a face needs three corners");
            for (size_t k = 1; k + 1 < corner.size(); ++k) {
                m.indices.push_back(corner[0]);
                m.indices.push_back(clockwise ? corner[k + 1] : corner[k]);
                m.indices.push_back(clockwise ? corner[k] : corner[k + 1]);
            }
        }
    }
    return m;
}
inline ObjMesh load_obj(const std::string& path, bool clockwise = true) {
    std::ifstream in(path);
    if (!in) throw GpuError(RGPU_ERR_INVALID_ARG, "cannot open OBJ file " + path);
    return parse_obj(in, clockwise);
}

class SceneManager {
public:
    MaterialRef addLambertian(const float Le[3], const float R[3], bool backEmission = false) {
        rgpu_lambertian m = {{Le[0], Le[1], Le[2]}, {R[0], R[1], R[2]}, backEmission ? 1 : 0};
        lambertians_.push_back(m);
        return {RGPU_MAT_LAMBERTIAN, (int32_t)lambertians_.size() - 1};
    }
    MaterialRef addWard(float alpha_x, float alpha_y, const float reflectance[3], const float edgetint[3],
                        float falloff) {
        rgpu_ward w = {alpha_x, alpha_y, {reflectance[0], reflectance[1], reflectance[2]},
                       {edgetint[0], edgetint[1], edgetint[2]}, falloff};
        wards_.push_back(w);
        return {RGPU_MAT_WARD, (int32_t)wards_.size() - 1};
    }
    MaterialRef addSpecular() { return {RGPU_MAT_SPECULAR, 0}; }
    MaterialRef addDielectric(float ior = 1.5f) {
        dielectrics_.push_back(rgpu_dielectric{ior});
        return {RGPU_MAT_DIELECTRIC, (int32_t)dielectrics_.size() - 1};
    }
    // HomogeneousVolume boundary: C = extinction, R = scattering colour
    MaterialRef addVolume(float C, const float R[3]) {
        volumes_.push_back(rgpu_volume{C, {R[0], R[1], R[2]}});
        return {RGPU_MAT_VOLUME, (int32_t)volumes_.size() - 1};
    }

    // addPolymesh (scene_manager.hpp:339-370): triangles over `points` (xyz), clockwise as seen
    // from the front, all with material `m`.
    // Every index is checked before anything is added, so a refused mesh leaves the scene as it
    // was.
    void addPolymesh(const float* xyz, size_t npoints, const uint32_t* tri, size_t ntris, MaterialRef m) {
        for (size_t i = 0; i < 3 * ntris; ++i)
            if (tri[i] >= npoints) throw GpuError(RGPU_ERR_INVALID_ARG, "addPolymesh: index out of range");
        const uint32_t base = (uint32_t)(points_.size() / 3);
        points_.insert(points_.end(), xyz, xyz + 3 * npoints);
        indices_.reserve(indices_.size() + 3 * ntris);
        for (size_t i = 0; i < 3 * ntris; ++i) indices_.push_back(base + tri[i]);
        tri_mats_.insert(tri_mats_.end(), ntris, m);
        bvh_.reset();
    }
    // A Wavefront OBJ mesh (load_obj) as one polymesh with material `m`.
    void addObj(const std::string& path, MaterialRef m, bool clockwise = true) {
        const ObjMesh o = load_obj(path, clockwise);
        addPolymesh(o.points.data(), o.points.size() / 3, o.indices.data(), o.indices.size() / 3, m);
    }
    // Sphere primitives of addPointPrimitive (scene_manager.hpp:372-430).
    void addSphere(const float center[3], float radius, MaterialRef m) {
        spheres_.insert(spheres_.end(), {center[0], center[1], center[2], radius});
        sphere_mats_.push_back(m);
        bvh_.reset();
    }
    // ImageEnvmap of addPoint (scene_manager.hpp:432-497): rgb(a) floats, row 0 at the zenith.
    void setEnvmap(const float* rgb, uint32_t channels, uint32_t width, uint32_t height) {
        env_.assign(rgb, rgb + (size_t)channels * width * height);
        env_channels_ = channels;
        env_w_ = width;
        env_h_ = height;
    }

    // buildBVH (scene_manager.hpp:517-577): binned SAH over triangles then spheres.
    void buildBVH(uint32_t max_leaf = 4, uint32_t bins = 16, uint32_t max_depth = 63) {
        rgpu_bvh_options opt = {RGPU_ABI_VERSION, max_leaf, bins, max_depth};
        rgpu_bvh* b = nullptr;
        check(rgpu_bvh_build(points_.data(), 3, (uint32_t)(points_.size() / 3), indices_.data(),
                             (uint32_t)(indices_.size() / 3), spheres_.data(), (uint32_t)(spheres_.size() / 4), &opt,
                             &b),
              nullptr, "rgpu_bvh_build");
        bvh_ = BVHHandle(b);
    }
    // The same tree built on the GPU of `gpu` (rgpu_bvh_build_device: bins <= 32, finite
    // coordinates).
    void buildBVH(GpuContext& gpu, uint32_t max_leaf = 4, uint32_t bins = 16, uint32_t max_depth = 63) {
        rgpu_bvh_options opt = {RGPU_ABI_VERSION, max_leaf, bins, max_depth};
        rgpu_bvh* b = nullptr;
        check(rgpu_bvh_build_device(gpu.get(), points_.data(), 3, (uint32_t)(points_.size() / 3), indices_.data(),
                                    (uint32_t)(indices_.size() / 3), spheres_.data(),
                                    (uint32_t)(spheres_.size() / 4), &opt, &b, nullptr),
              gpu.get(), "rgpu_bvh_build_device");
        bvh_ = BVHHandle(b);
    }
    rgpu_bvh_stats bvhStats() const {
        rgpu_bvh_stats st{};
        check(bvh_ ? rgpu_bvh_get_stats(bvh_.get(), &st) : RGPU_ERR_INVALID_ARG, nullptr, "bvhStats (build first)");
        return st;
    }

    // createBuffer / createMaterialBuffer / createEnvmapBuffer (scene_manager.hpp:579-650)
    SceneBuffer createBuffer(GpuContext& ctx) const {
        if (!bvh_) throw GpuError(RGPU_ERR_INVALID_ARG, "createBuffer: buildBVH() first");
        rgpu_scene* s = nullptr;
        check(rgpu_scene_create(ctx.get(), bvh_.get(), points_.data(), 3, (uint32_t)(points_.size() / 3),
                                indices_.data(), (uint32_t)(indices_.size() / 3), spheres_.data(),
                                (uint32_t)(spheres_.size() / 4), &s),
              ctx.get(), "rgpu_scene_create");
        return SceneBuffer(s);
    }
    MaterialBuffer createMaterialBuffer(GpuContext& ctx) const {
        std::vector<int32_t> types, idx;
        for (const std::vector<MaterialRef>* list : {&tri_mats_, &sphere_mats_})
            for (const MaterialRef& m : *list) {
                types.push_back(m.type);
                idx.push_back(m.index);
            }
        rgpu_materials* m = nullptr;
        check(rgpu_materials_create(ctx.get(), (uint32_t)types.size(), types.data(), idx.data(), lambertians_.data(),
                                    (uint32_t)lambertians_.size(), wards_.data(), (uint32_t)wards_.size(),
                                    dielectrics_.data(), (uint32_t)dielectrics_.size(), volumes_.data(),
                                    (uint32_t)volumes_.size(), &m),
              ctx.get(), "rgpu_materials_create");
        return MaterialBuffer(m);
    }
    EnvmapBuffer createEnvmapBuffer(GpuContext& ctx) const {
        static const float kBlack[3] = {0.0f, 0.0f, 0.0f};
        const bool none = env_.empty();
        rgpu_envmap* e = nullptr;
        check(rgpu_envmap_create(ctx.get(), none ? kBlack : env_.data(), none ? 3 : env_channels_, none ? 1 : env_w_,
                                 none ? 1 : env_h_, &e),
              ctx.get(), "rgpu_envmap_create");
        return EnvmapBuffer(e);
    }

    size_t triangleCount() const { return indices_.size() / 3; }
    size_t pointCount() const { return points_.size() / 3; }
    size_t sphereCount() const { return spheres_.size() / 4; }
    const rgpu_bvh* bvh() const { return bvh_.get(); }

private:
    std::vector<float> points_;
    std::vector<uint32_t> indices_;
    std::vector<float> spheres_;
    std::vector<MaterialRef> tri_mats_, sphere_mats_;
    std::vector<rgpu_lambertian> lambertians_;
    std::vector<rgpu_ward> wards_;
    std::vector<rgpu_dielectric> dielectrics_;
    std::vector<rgpu_volume> volumes_;
    std::vector<float> env_;
    uint32_t env_channels_ = 3, env_w_ = 0, env_h_ = 0;
    BVHHandle bvh_;
};

// Options of a lane beyond the reference's (which hard-codes them).
struct LaneOptions {
    uint64_t seed = 0;
    uint32_t max_depth = 0;        // 0 = unlimited, as the reference
    float radiance_clamp = 0.0f;   // 0 = 1000, as the reference
    uint32_t flags = 0;            // RGPU_WF_*
    int variant = RGPU_VARIANT_STACKLESS;
};

// WavefrontLane (wavefront_path_tracing.hpp:339-1153) on one GPU. Steps run on the context's
// stream; the finalize side (update_intermediate, merge_from, finalize_color) runs on the lane's
// own finalize stream, so an image thread can work beside a step thread.
class WavefrontLane {
public:
    WavefrontLane(GpuContext& lane, int wavefrontPathCount, LaneOptions options = LaneOptions())
        : lane_(lane), path_count_(wavefrontPathCount), options_(options), finalize_(lane) {}

    // setup(camera, sceneManager) (wavefront_path_tracing.hpp:451-530): device buffers of the
    // scene, materials and envmap, and the resident path state.
    void setup(const rgpu_camera& camera, const SceneManager& sceneManager) {
        wf_.reset();
        scene_ = sceneManager.createBuffer(lane_);
        materials_ = sceneManager.createMaterialBuffer(lane_);
        envmap_ = sceneManager.createEnvmapBuffer(lane_);
        rgpu_wavefront_config cfg = {RGPU_ABI_VERSION, (uint32_t)path_count_, options_.seed, 0, options_.max_depth,
                                     options_.radiance_clamp, 0.0f, options_.flags, (uint32_t)options_.variant, 0};
        rgpu_wavefront* w = nullptr;
        check(rgpu_wavefront_create(lane_.get(), scene_.get(), materials_.get(), envmap_.get(), &cfg, &w), lane_.get(),
              "rgpu_wavefront_create");
        wf_ = WavefrontHandle(w);
        camera_ = camera;
    }
    // initialize(lane_index, num_device) (wavefront_path_tracing.hpp:532-570)
    void initialize(int lane_index, int num_device) {
        std::lock_guard<std::mutex> g(finalize_mu_);
        check(rgpu_wavefront_initialize(wf_.get(), &camera_, (uint32_t)lane_index, (uint32_t)num_device),
              lane_.get(), "rgpu_wavefront_initialize");
        lane_index_ = lane_index;
        num_device_ = num_device;
        step_count_ = 0;
        size_images();
    }
    // Previews beside running steps: every step publishes the intermediate image for
    // update_intermediate through the library's weak device mutex (kernels/mutex.cl).
    void enable_previews(uint32_t every_steps = 1) {
        std::lock_guard<std::mutex> g(finalize_mu_);
        check(rgpu_wavefront_set_publish(wf_.get(), every_steps), lane_.get(), "rgpu_wavefront_set_publish");
        previews_ = every_steps != 0;
    }
    // step() (wavefront_path_tracing.hpp:591-939): one bounce of every resident path. A pending
    // re_start is applied first, on the stepping thread, as the reference does (592-604).
    void step() {
        {
            std::lock_guard<std::mutex> g(restart_mu_);
            if (restart_pending_) {
                restart_pending_ = false;
                camera_ = restart_camera_;
                initialize(restart_lane_, num_device_);
            }
        }
        check(rgpu_wavefront_step(wf_.get(), nullptr, 1), lane_.get(), "rgpu_wavefront_step");
        step_count_.fetch_add(1);
    }
    // update_intermediate() (wavefront_path_tracing.hpp:942-960): the lane's intermediate image
    // (RGB16 means and counts) on the finalize stream. With previews on, it is the image the
    // last step published, copied if the step side does not hold the lock (otherwise the
    // previous intermediate stays); without, it is made from the accumulation after the steps
    // queued so far. Returns whether the intermediate was refreshed.
    bool update_intermediate() {
        std::lock_guard<std::mutex> g(finalize_mu_);
        if (previews_) {
            check(rgpu_wavefront_update_intermediate(wf_.get(), finalize_.get(), intermediate(),
                                                     static_cast<int*>(took_->get())),
                  lane_.get(), "rgpu_wavefront_update_intermediate");
            int took = 0;
            check(rgpu_download(lane_.get(), finalize_.get(), &took, took_->get(), sizeof(int)), lane_.get(),
                  "rgpu_download");
            return took != 0;
        }
        lane_.synchronize();  // the steps queued on the context's stream
        check(rgpu_wavefront_intermediate(wf_.get(), finalize_.get(), intermediate()), lane_.get(),
              "rgpu_wavefront_intermediate");
        finalize_.synchronize();
        return true;
    }
    // merge_from(other) (wavefront_path_tracing.hpp:962-991): this intermediate becomes the
    // count-weighted mean of both lanes' (on one GPU, device to device).
    void merge_from(WavefrontLane& other) {
        std::lock_guard<std::mutex> g(finalize_mu_);
        other.finalize_.synchronize();
        check(rgpu_merge_intermediate(lane_.get(), finalize_.get(), intermediate(), other.intermediate(), npixels()),
              lane_.get(), "rgpu_merge_intermediate");
    }
    // finalize_color() (wavefront_path_tracing.hpp:993-1007): tonemap of the intermediate, read back.
    std::vector<RGBA8ValueType> finalize_color() {
        std::lock_guard<std::mutex> g(finalize_mu_);
        std::vector<RGBA8ValueType> img(npixels());
        check(rgpu_tonemap_intermediate(lane_.get(), finalize_.get(), intermediate(), npixels(),
                                        static_cast<uint8_t*>(final_color_->get())),
              lane_.get(), "rgpu_tonemap_intermediate");
        check(rgpu_download(lane_.get(), finalize_.get(), img.data(), final_color_->get(), img.size() * 4),
              lane_.get(), "rgpu_download");
        return img;
    }
    // The intermediate on the host (12 bytes per pixel: half-float means and a 32-bit count).
    std::vector<rgpu_rgb16> read_intermediate() {
        std::lock_guard<std::mutex> g(finalize_mu_);
        std::vector<rgpu_rgb16> out(npixels());
        check(rgpu_download(lane_.get(), finalize_.get(), out.data(), intermediate(), out.size() * sizeof(rgpu_rgb16)),
              lane_.get(), "rgpu_download");
        return out;
    }
    // re_start(lane_index, camera) (wavefront_path_tracing.hpp:1013-1018): applied by the next step.
    void re_start(int lane_index, const rgpu_camera& camera) {
        std::lock_guard<std::mutex> g(restart_mu_);
        restart_pending_ = true;
        restart_lane_ = lane_index;
        restart_camera_ = camera;
    }
    int step_count() const { return step_count_.load(); }
    // Snapshot of the lane that resume() continues exactly (no counterpart in the reference,
    // whose lane state lives only in device buffers).
    struct Checkpoint {
        rgpu_camera camera;
        int step_count;
        std::vector<uint8_t> bytes;
    };
    Checkpoint checkpoint() const {
        uint64_t n = 0;
        check(rgpu_wavefront_checkpoint_size(wf_.get(), &n), lane_.get(), "rgpu_wavefront_checkpoint_size");
        Checkpoint c{camera_, step_count_.load(), std::vector<uint8_t>(n)};
        check(rgpu_wavefront_checkpoint(wf_.get(), c.bytes.data(), n), lane_.get(), "rgpu_wavefront_checkpoint");
        return c;
    }
    void resume(const Checkpoint& c) {
        std::lock_guard<std::mutex> g(finalize_mu_);
        check(rgpu_wavefront_resume(wf_.get(), c.bytes.data(), c.bytes.size()), lane_.get(), "rgpu_wavefront_resume");
        camera_ = c.camera;
        step_count_ = c.step_count;
        size_images();
    }
    // stat_avg_sample(): samples per pixel so far (kernels/stat.cl)
    float stat_avg_sample() const {
        uint64_t total = 0;
        check(rgpu_wavefront_stat(wf_.get(), &total), lane_.get(), "rgpu_wavefront_stat");
        return (float)((double)total / ((double)camera_.width * camera_.height));
    }
    const rgpu_camera& camera() const { return camera_; }
    rgpu_wavefront* handle() const { return wf_.get(); }
    uint32_t npixels() const { return camera_.width * camera_.height; }

private:
    rgpu_rgb16* intermediate() const { return static_cast<rgpu_rgb16*>(intermediate_->get()); }
    // device images of the finalize side, sized to the camera
    void size_images() {
        const size_t P = npixels();
        if (!intermediate_ || intermediate_->bytes() != P * sizeof(rgpu_rgb16)) {
            intermediate_.reset(new DeviceMemory(lane_, P * sizeof(rgpu_rgb16)));
            final_color_.reset(new DeviceMemory(lane_, P * 4));
            check(rgpu_fill_u32(lane_.get(), finalize_.get(), intermediate_->get(), 0u, P * 3), lane_.get(),
                  "rgpu_fill_u32");
            finalize_.synchronize();
        }
        if (!took_) took_.reset(new DeviceMemory(lane_, sizeof(int)));
    }

    GpuContext& lane_;
    int path_count_;
    LaneOptions options_;
    rgpu_camera camera_{};
    int lane_index_ = 0, num_device_ = 1;
    std::atomic<int> step_count_{0};
    SceneBuffer scene_;
    MaterialBuffer materials_;
    EnvmapBuffer envmap_;
    WavefrontHandle wf_;
    GpuStream finalize_;
    std::unique_ptr<DeviceMemory> intermediate_, final_color_, took_;
    bool previews_ = false;
    std::mutex finalize_mu_;  // the finalize side against (re)initialisation
    std::mutex restart_mu_;
    bool restart_pending_ = false;
    int restart_lane_ = 0;
    rgpu_camera restart_camera_{};
};

// WavefrontPathTracing (wavefront_path_tracing.hpp:1159-1470) on one GPU (RenderMode_SingleGPU).
class WavefrontPathTracing {
public:
    // Fixed launches: the caller steps (launch_fixed) and asks for images (create_color_image).
    WavefrontPathTracing(const SceneManager& sceneManager, const rgpu_camera& camera,
                         std::function<void(RGBA8ValueType*, int, int)> colorRecieved, int wavefrontPathCount = 1 << 22,
                         int device_index = 0, LaneOptions options = LaneOptions())
        : onColorRecieved(std::move(colorRecieved)), context_(device_index) {
        lane_.reset(new WavefrontLane(context_, wavefrontPathCount, options));
        lane_->setup(camera, sceneManager);
        lane_->initialize(0, 1);
    }
    // Continuous rendering (wavefront_path_tracing.hpp:1286-1313): a worker thread steps the lane
    // until stop() or destruction, and an image thread, every margin_period_ms once the lane has
    // taken more than 4 steps, makes an image through the preview handoff (update_intermediate
    // on its own stream while steps run) and passes it to onColorRecieved.
    WavefrontPathTracing(const SceneManager& sceneManager, const rgpu_camera& camera, int margin_period_ms,
                         std::function<void(RGBA8ValueType*, int, int)> colorRecieved, int wavefrontPathCount = 1 << 22,
                         int device_index = 0, LaneOptions options = LaneOptions())
        : onColorRecieved(std::move(colorRecieved)), context_(device_index) {
        lane_.reset(new WavefrontLane(context_, wavefrontPathCount, options));
        lane_->setup(camera, sceneManager);
        lane_->initialize(0, 1);
        lane_->enable_previews(1);
        continue_ = true;
        workers_.emplace_back([this]() {
            guarded([this]() {
                int queued = 0;
                while (continue_) {
                    lane_->step();
                    // bound the work queued ahead of the GPU (the reference throttles on markers)
                    if (++queued % 4 == 0) context_.synchronize();
                }
                context_.synchronize();
            });
        });
        try {
            workers_.emplace_back([this, margin_period_ms]() {
                guarded([this, margin_period_ms]() {
                    while (continue_) {
                        if (4 < lane_->step_count()) create_color_image();
                        std::this_thread::sleep_for(std::chrono::milliseconds(margin_period_ms));
                    }
                });
            });
        } catch (...) {
            join_workers();  // the step thread must not outlive a constructor that failed
            throw;
        }
    }
    // Joins the threads; a worker's error that stop() has not rethrown is dropped, since a
    // destructor must not throw.
    ~WavefrontPathTracing() { join_workers(); }
    // Ends continuous rendering (joins the threads); rethrows a worker's error.
    void stop() {
        join_workers();
        std::lock_guard<std::mutex> g(error_mu_);
        if (error_) {
            std::exception_ptr e = error_;
            error_ = nullptr;
            std::rethrow_exception(e);
        }
    }
    bool running() const { return continue_.load(); }
    // launch_fixed(steps) (wavefront_path_tracing.hpp:1443-1458)
    void launch_fixed(int steps) {
        for (int i = 0; i < steps; ++i) lane_->step();
        create_color_image();
    }
    // re_start(camera) (wavefront_path_tracing.hpp:1460-1464): the next step restarts the lane.
    void re_start(const rgpu_camera& camera) { lane_->re_start(0, camera); }
    // create_color_image() (wavefront_path_tracing.hpp:1345-1408): with one lane there is
    // nothing to merge; the finalised image goes to onColorRecieved.
    void create_color_image() {
        lane_->update_intermediate();
        std::vector<RGBA8ValueType> img = lane_->finalize_color();
        images_.fetch_add(1);
        if (onColorRecieved) onColorRecieved(img.data(), (int)lane_->camera().width, (int)lane_->camera().height);
    }
    int images() const { return images_.load(); }
    WavefrontLane& lane() { return *lane_; }
    GpuContext& context() { return context_; }

    std::function<void(RGBA8ValueType*, int, int)> onColorRecieved;

private:
    void join_workers() {
        continue_ = false;
        for (std::thread& t : workers_)
            if (t.joinable()) t.join();
        workers_.clear();
    }
    template <class F>
    void guarded(F f) {
        try {
            f();
        } catch (...) {
            std::lock_guard<std::mutex> g(error_mu_);
            if (!error_) error_ = std::current_exception();
            continue_ = false;
        }
    }

    GpuContext context_;
    std::unique_ptr<WavefrontLane> lane_;
    std::atomic<bool> continue_{false};
    std::atomic<int> images_{0};
    std::vector<std::thread> workers_;
    std::mutex error_mu_;
    std::exception_ptr error_;
};

// A built tree cast against on the GPU with one of the traversals: the stackless walk of
// common/stackless_bvh.hpp (kernels/extension_ray_cast_stackless.cl), the multi-threaded walk of
// common/threaded_bvh.hpp (kernels/extension_ray_cast.cl) or its single-table form
// (TBVHNode / intersect_tbvh, common/threaded_bvh.hpp:269-396).
class RayCaster {
public:
    RayCaster(GpuContext& ctx, const SceneManager& sceneManager, int variant)
        : ctx_(ctx), scene_(sceneManager.createBuffer(ctx)), variant_(variant) {}
    // Closest hits of host rays (prim -1 and t = tmax on a miss).
    std::vector<rgpu_hit> intersect(const std::vector<rgpu_ray>& rays) const {
        std::vector<rgpu_hit> hits(rays.size());
        if (rays.empty()) return hits;
        void* dr = nullptr;
        void* dh = nullptr;
        check(rgpu_alloc(ctx_.get(), rays.size() * sizeof(rgpu_ray), &dr), ctx_.get(), "rgpu_alloc");
        int rc = rgpu_alloc(ctx_.get(), rays.size() * sizeof(rgpu_hit), &dh);
        if (rc == RGPU_OK)
            rc = rgpu_upload(ctx_.get(), nullptr, dr, rays.data(), rays.size() * sizeof(rgpu_ray));
        if (rc == RGPU_OK)
            rc = rgpu_extension_ray_cast(ctx_.get(), nullptr, scene_.get(), variant_, (const rgpu_ray*)dr,
                                         (rgpu_hit*)dh, (uint32_t)rays.size(), 0);
        if (rc == RGPU_OK) rc = rgpu_download(ctx_.get(), nullptr, hits.data(), dh, hits.size() * sizeof(rgpu_hit));
        if (dh) rgpu_free(ctx_.get(), dh);
        rgpu_free(ctx_.get(), dr);
        check(rc, ctx_.get(), "RayCaster::intersect");
        return hits;
    }

private:
    GpuContext& ctx_;
    SceneBuffer scene_;
    int variant_;
};

class StacklessBVH : public RayCaster {
public:
    StacklessBVH(GpuContext& ctx, const SceneManager& sm) : RayCaster(ctx, sm, RGPU_VARIANT_STACKLESS) {}
};
class ThreadedBVH : public RayCaster {
public:
    ThreadedBVH(GpuContext& ctx, const SceneManager& sm) : RayCaster(ctx, sm, RGPU_VARIANT_THREADED) {}
};
class SingleThreadedBVH : public RayCaster {
public:
    SingleThreadedBVH(GpuContext& ctx, const SceneManager& sm) : RayCaster(ctx, sm, RGPU_VARIANT_THREADED_SINGLE) {}
};

}  // namespace rt
