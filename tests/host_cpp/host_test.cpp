// Exercises the C++ host surface (raccoongpurender_amd/host/raccoon_host.hpp) the way the
// reference's host code drives its lane: SceneManager -> buildBVH -> WavefrontPathTracing
// (launch_fixed, re_start) and the two BVH ray casters.
//
//   host_test cpu          scene + BVH on the host; GpuContext must report a missing device
//   host_test gpu <dir>    render + ray casts; writes the scene, the image and the hits to
//                          <dir> so that the Python binding can redo them through the C ABI
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <chrono>
#include <mutex>
#include <thread>

#include "../../raccoongpurender_amd/host/raccoon_host.hpp"

namespace {

struct Scene {
    rt::SceneManager sm;
    // copies of what went in, for the dump
    std::vector<float> points;
    std::vector<uint32_t> indices;
    std::vector<float> spheres;
    std::vector<int32_t> types, idx;
    std::vector<float> lambertians;  // Le, R, back
    std::vector<float> iors;
    std::vector<float> env;
};

void quad(Scene& s, const float a[3], const float b[3], const float c[3], const float d[3], rt::MaterialRef m) {
    const float p[12] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], d[0], d[1], d[2]};
    const uint32_t t[6] = {0, 1, 2, 0, 2, 3};
    s.sm.addPolymesh(p, 4, t, 2, m);
    const uint32_t base = (uint32_t)(s.points.size() / 3);
    s.points.insert(s.points.end(), p, p + 12);
    for (uint32_t i : t) s.indices.push_back(base + i);
    for (int k = 0; k < 2; ++k) {
        s.types.push_back(m.type);
        s.idx.push_back(m.index);
    }
}

rt::MaterialRef lambert(Scene& s, float le, float r, float g, float b) {
    const float Le[3] = {le, le * 0.8f, le * 0.5f}, R[3] = {r, g, b};
    s.lambertians.insert(s.lambertians.end(), {Le[0], Le[1], Le[2], r, g, b, 0.0f});
    return s.sm.addLambertian(Le, R);
}

// A closed box (walls clockwise seen from inside), a ceiling light, a lambertian and a glass
// sphere primitive, under a small sky.
void build(Scene& s) {
    auto white = lambert(s, 0.0f, 0.75f, 0.75f, 0.75f);
    auto red = lambert(s, 0.0f, 0.63f, 0.06f, 0.04f);
    auto green = lambert(s, 0.0f, 0.14f, 0.45f, 0.09f);
    auto light = lambert(s, 17.0f, 0.0f, 0.0f, 0.0f);
    const float A[3] = {-1, -1, -1}, B[3] = {1, -1, -1}, C[3] = {1, -1, 1}, D[3] = {-1, -1, 1};
    const float E[3] = {-1, 1, -1}, F[3] = {1, 1, -1}, G[3] = {1, 1, 1}, H[3] = {-1, 1, 1};
    quad(s, A, B, C, D, white);  // floor
    quad(s, H, G, F, E, white);  // ceiling
    quad(s, E, F, B, A, white);  // back
    quad(s, A, D, H, E, red);    // left
    quad(s, B, F, G, C, green);  // right
    const float L0[3] = {-0.25f, 0.99f, -0.25f}, L1[3] = {0.25f, 0.99f, -0.25f}, L2[3] = {0.25f, 0.99f, 0.25f},
                L3[3] = {-0.25f, 0.99f, 0.25f};
    quad(s, L3, L2, L1, L0, light);
    const float c0[4] = {-0.4f, -0.6f, -0.2f, 0.4f}, c1[4] = {0.45f, -0.65f, 0.3f, 0.35f};
    s.sm.addSphere(c0, c0[3], white);
    s.spheres.insert(s.spheres.end(), c0, c0 + 4);
    s.types.push_back(white.type);
    s.idx.push_back(white.index);
    auto glass = s.sm.addDielectric(1.5f);
    s.iors.push_back(1.5f);
    s.sm.addSphere(c1, c1[3], glass);
    s.spheres.insert(s.spheres.end(), c1, c1 + 4);
    s.types.push_back(glass.type);
    s.idx.push_back(glass.index);
    for (int y = 0; y < 8; ++y)
        for (int x = 0; x < 16; ++x) s.env.insert(s.env.end(), {0.2f + 0.05f * y, 0.3f, 0.4f + 0.02f * x});
    s.sm.setEnvmap(s.env.data(), 3, 16, 8);
    s.sm.buildBVH();
}

template <class T>
void dump(const std::string& path, const std::vector<T>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    if (mode == "obj" && argc > 2) {
        // host_test obj <file>: the mesh rt::load_obj reads, floats as their bit patterns
        try {
            const rt::ObjMesh m = rt::load_obj(argv[2]);
            std::printf("{\"points_bits\": [");
            for (size_t i = 0; i < m.points.size(); ++i) {
                uint32_t b;
                std::memcpy(&b, &m.points[i], 4);
                std::printf("%s%u", i ? ", " : "", b);
            }
            std::printf("], \"indices\": [");
            for (size_t i = 0; i < m.indices.size(); ++i) std::printf("%s%u", i ? ", " : "", m.indices[i]);
            std::printf("]}\n");
            return 0;
        } catch (const rt::GpuError& e) {
            std::printf("{\"error\": %d}\n", e.code);
            return 0;
        }
    }
    try {
        Scene s;
        build(s);
        const rgpu_bvh_stats st = s.sm.bvhStats();
        if (mode == "cpu") {
            bool no_device = false;
            int code = 0;
            try {
                rt::GpuContext ctx(0);
            } catch (const rt::GpuError& e) {
                no_device = true;
                code = e.code;
            }
            // a refused mesh (its last index is out of range) leaves the scene as it was
            bool bad_index = false;
            rt::SceneManager sm;
            const float good_p[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
            const uint32_t good_t[3] = {0, 1, 2};
            sm.addPolymesh(good_p, 3, good_t, 1, rt::MaterialRef{});
            try {
                const float p[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
                const uint32_t t[6] = {0, 1, 2, 0, 1, 5};
                sm.addPolymesh(p, 3, t, 2, rt::MaterialRef{});
            } catch (const rt::GpuError& e) {
                bad_index = e.code == RGPU_ERR_INVALID_ARG && sm.triangleCount() == 1 && sm.pointCount() == 3;
            }
            std::printf("{\"triangles\": %zu, \"spheres\": %zu, \"nodes\": %u, \"leaves\": %u, \"slots\": %u, "
                        "\"no_device\": %s, \"code\": %d, \"bad_index_rejected\": %s}\n",
                        s.sm.triangleCount(), s.sm.sphereCount(), st.nodes, st.leaves, st.slots,
                        no_device ? "true" : "false", code, bad_index ? "true" : "false");
            return 0;
        }
        if (mode != "gpu" || argc < 3) {
            std::fprintf(stderr, "usage: host_test cpu | host_test gpu <dir>\n");
            return 2;
        }
        const std::string dir = argv[2];
        rt::CameraSetting cs;
        const float eye[3] = {0.0f, 0.0f, 3.3f}, target[3] = {0.0f, 0.0f, 0.0f};
        std::memcpy(cs.eye, eye, sizeof(eye));
        std::memcpy(cs.target, target, sizeof(target));
        cs.width = 64;
        cs.height = 48;
        const rgpu_camera cam = rt::standardCamera(cs);
        rt::LaneOptions opt;
        opt.seed = 7;
        std::vector<rt::RGBA8ValueType> image;
        int w = 0, h = 0;
        rt::WavefrontPathTracing pt(
            s.sm, cam,
            [&](rt::RGBA8ValueType* px, int width, int height) {
                image.assign(px, px + (size_t)width * height);
                w = width;
                h = height;
            },
            64 * 48 * 4, 0, opt);
        pt.launch_fixed(12);
        const float spp = pt.lane().stat_avg_sample();
        const int steps = pt.lane().step_count();
        const std::vector<rt::RGBA8ValueType> first = image;
        // a restart clears the accumulation and starts again from the same seeds
        pt.re_start(cam);
        pt.launch_fixed(12);
        const std::vector<rt::RGBA8ValueType> again = image;
        // a checkpoint resumed after more steps continues as the uninterrupted lane: the sample
        // totals (integer counts) after the same further steps agree exactly
        {
            rt::WavefrontLane& lane = pt.lane();
            const rt::WavefrontLane::Checkpoint ck = lane.checkpoint();
            for (int i = 0; i < 3; ++i) lane.step();
            const float spp_a = lane.stat_avg_sample();
            lane.resume(ck);
            for (int i = 0; i < 3; ++i) lane.step();
            if (lane.stat_avg_sample() != spp_a || lane.step_count() != ck.step_count + 3)
                throw std::runtime_error("resumed lane diverged from the uninterrupted one");
        }
        // continuous rendering (wavefront_path_tracing.hpp:1286-1313): a step thread and an image
        // thread; previews come through the handoff while steps run, their sample totals never
        // decrease, and stop() ends both threads
        int live_images = 0, live_steps = 0;
        bool live_monotonic = true, live_alpha = true;
        uint64_t live_last_total = 0, live_final_total = 0, live_stat_total = 0;
        {
            std::mutex m;
            rt::WavefrontPathTracing* self = nullptr;
            rt::WavefrontPathTracing live(
                s.sm, cam, 2,
                [&](rt::RGBA8ValueType* px, int width, int height) {
                    std::lock_guard<std::mutex> g(m);
                    for (int i = 0; i < width * height; ++i) live_alpha = live_alpha && px[i].a == 255;
                    if (!self) return;
                    uint64_t total = 0;
                    for (const rgpu_rgb16& v : self->lane().read_intermediate()) total += v.count;
                    live_monotonic = live_monotonic && total >= live_last_total;
                    live_last_total = total;
                },
                64 * 48 * 4, 0, opt);
            {
                std::lock_guard<std::mutex> g(m);
                self = &live;
            }
            const auto t0 = std::chrono::steady_clock::now();
            while (live.images() < 8 && std::chrono::steady_clock::now() - t0 < std::chrono::seconds(30))
                std::this_thread::sleep_for(std::chrono::milliseconds(5));
            live.stop();
            live_images = live.images();
            live_steps = live.lane().step_count();
            // with the step thread gone nothing contends: the update takes the published image
            if (!live.lane().update_intermediate()) throw std::runtime_error("an uncontended preview did not take the lock");
            for (const rgpu_rgb16& v : live.lane().read_intermediate()) live_final_total += v.count;
            live_stat_total = (uint64_t)((double)live.lane().stat_avg_sample() * 64 * 48 + 0.5);
        }
        // ray casts through both BVH walks
        std::vector<rgpu_ray> rays;
        for (int i = 0; i < 1000; ++i) {
            const float a = 0.0061f * i, b = 0.0173f * i;
            rgpu_ray r = {{0.1f * std::sin(b), 0.2f * std::cos(a), 0.05f}, 1e30f,
                          {std::cos(a) * std::sin(b), std::sin(a), std::cos(a) * std::cos(b)}, 0};
            rays.push_back(r);
        }
        // the same scene's tree built on the GPU: the host build's node and leaf counts and depth
        {
            Scene s2;
            build(s2);
            s2.sm.buildBVH(pt.context());
            const rgpu_bvh_stats d = s2.sm.bvhStats();
            if (d.nodes != st.nodes || d.leaves != st.leaves || d.depth != st.depth || d.slots != st.slots)
                throw std::runtime_error("device BVH build differs from the host build");
        }
        std::vector<rgpu_hit> hs = rt::StacklessBVH(pt.context(), s.sm).intersect(rays);
        std::vector<rgpu_hit> ht = rt::ThreadedBVH(pt.context(), s.sm).intersect(rays);
        std::vector<rgpu_hit> h1 = rt::SingleThreadedBVH(pt.context(), s.sm).intersect(rays);
        dump(dir + "/points.f32", s.points);
        dump(dir + "/indices.u32", s.indices);
        dump(dir + "/spheres.f32", s.spheres);
        dump(dir + "/types.i32", s.types);
        dump(dir + "/idx.i32", s.idx);
        dump(dir + "/lambertians.f32", s.lambertians);
        dump(dir + "/iors.f32", s.iors);
        dump(dir + "/env.f32", s.env);
        const rgpu_camera* cp = &cam;
        std::vector<float> camv;
        for (const float* q : {cp->eye, cp->forward, cp->up, cp->right, cp->imageplane_o, cp->imageplane_r,
                               cp->imageplane_b})
            camv.insert(camv.end(), q, q + 3);
        camv.push_back(cp->lens_radius);
        dump(dir + "/camera.f32", camv);
        dump(dir + "/image.u8", std::vector<uint8_t>((const uint8_t*)first.data(),
                                                     (const uint8_t*)first.data() + 4 * first.size()));
        dump(dir + "/image_again.u8", std::vector<uint8_t>((const uint8_t*)again.data(),
                                                           (const uint8_t*)again.data() + 4 * again.size()));
        dump(dir + "/rays.f32", std::vector<float>((const float*)rays.data(), (const float*)rays.data() + 8 * rays.size()));
        dump(dir + "/hits_stackless.i32",
             std::vector<int32_t>((const int32_t*)hs.data(), (const int32_t*)hs.data() + 4 * hs.size()));
        dump(dir + "/hits_threaded.i32",
             std::vector<int32_t>((const int32_t*)ht.data(), (const int32_t*)ht.data() + 4 * ht.size()));
        dump(dir + "/hits_threaded_single.i32",
             std::vector<int32_t>((const int32_t*)h1.data(), (const int32_t*)h1.data() + 4 * h1.size()));
        std::printf("{\"width\": %d, \"height\": %d, \"spp\": %.6f, \"steps\": %d, \"nodes\": %u, \"paths\": %d, "
                    "\"live_images\": %d, \"live_steps\": %d, \"live_monotonic\": %s, \"live_alpha\": %s, "
                    "\"live_last_total\": %llu, \"live_final_total\": %llu, \"live_stat_total\": %llu}\n",
                    w, h, spp, steps, st.nodes, 64 * 48 * 4, live_images, live_steps, live_monotonic ? "true" : "false",
                    live_alpha ? "true" : "false", (unsigned long long)live_last_total,
                    (unsigned long long)live_final_total, (unsigned long long)live_stat_total);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "host_test: %s\n", e.what());
        return 1;
    }
}
