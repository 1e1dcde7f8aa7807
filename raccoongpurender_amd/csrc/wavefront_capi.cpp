// C ABI of the wavefront path tracer: materials, envmap, camera, and the
// WavefrontLane-style lifecycle (setup / initialize / step / read back).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/raccoon_gpu.h"
#include "envmap_build.h"
#include "fingerprint.h"
#include "rgpu_ctx.h"
#include "rgpu_scene.h"
#include "rgpu_wavefront.h"

using namespace rgpu;


struct rgpu_materials {
    rgpu_ctx* ctx = nullptr;
    DevMaterials dev;
    void* bufs[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    uint32_t nprims = 0;
    uint64_t fingerprint = 0;  // of the material records (checkpoints compare it)
};

struct rgpu_envmap {
    rgpu_ctx* ctx = nullptr;
    EnvmapTables host;
    DevEnvmap dev;
    void* bufs[4] = {nullptr, nullptr, nullptr, nullptr};
    uint64_t fingerprint = 0;  // of the texels (checkpoints compare it)
};

struct rgpu_wavefront {
    rgpu_ctx* ctx = nullptr;
    const rgpu_scene* scene = nullptr;
    const rgpu_materials* mats = nullptr;
    const rgpu_envmap* env = nullptr;
    WFParams params{};
    uint64_t seed = 0;
    int variant = 0;
    int queue_blocks = 0;
    Camera cam{};
    bool initialized = false;
    uint32_t npixels = 0;
    DevState st;
    std::vector<void*> owned;
    unsigned long long* stat_total = nullptr;
    uint8_t* tonemap_buf = nullptr;
    int trace_block = 256;  // 0 would select persistent extension-ray-cast waves
    hipGraphExec_t graph = nullptr;  // captured bounce loop
    uint32_t graph_steps = 0;
    DevGuide guide;                  // network + per-slot guiding buffers (guide.enabled == 0: plain)
    struct rgpu_guide* guide_obj = nullptr;
    // batch of render jobs (rgpu_wavefront_create_batch); empty for a single render
    struct Job {
        const rgpu_scene* scene;
        const rgpu_materials* mats;
        const rgpu_envmap* env;
        Camera cam;
        float* accum;
        uint32_t npixels, spp, first_block, nblocks;
        uint64_t seed;
        GuideFrame gframe;
    };
    std::vector<Job> jobs;
    DevJobs djobs;
    bool is_batch() const { return !jobs.empty(); }
    // preview handoff (rgpu_wavefront_set_publish): the published intermediate image and the
    // weak-mutex words (mutex, step side's flag, preview side's flag)
    uint32_t publish_every = 0;
    uint64_t publish_tick = 0;
    RGB16* pub_image = nullptr;
    uint32_t pub_pixels = 0;
    int* pub_words = nullptr;
};

struct rgpu_guide {
    rgpu_ctx* ctx = nullptr;
    DevGuide dev;                   // network, Adam state and training buffer (per-slot fields unused)
    std::vector<void*> owned;
    uint32_t step = 0;              // Adam steps taken
    float lr = 1e-3f;
    uint32_t minibatch = 16384;
    float select_p = 0.5f;
    std::vector<rgpu_wavefront*> users;  // wavefronts it is attached to (destroy refuses while any)
};

namespace {
void detach_from_guide(rgpu_wavefront* w) {
    if (!w->guide_obj) return;
    auto& u = w->guide_obj->users;
    u.erase(std::remove(u.begin(), u.end(), w), u.end());
    w->guide_obj = nullptr;
}
}  // namespace

namespace {
void drop_graph(rgpu_wavefront* w) {
    if (w->graph) (void)hipGraphExecDestroy(w->graph);
    w->graph = nullptr;
    w->graph_steps = 0;
}
}  // namespace

#define RGPU_HIP(ctx, call)                                     \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call);  \
    } while (0)

namespace {

template <class T>
int upload_vec(rgpu_ctx* ctx, const T* src, size_t n, void** dst) {
    *dst = nullptr;
    if (n == 0) {
        // keep a valid pointer for kernels that index nothing
        RGPU_HIP(ctx, hipMalloc(dst, 16));
        return RGPU_OK;
    }
    RGPU_HIP(ctx, hipMalloc(dst, n * sizeof(T)));
    RGPU_HIP(ctx, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return RGPU_OK;
}

int wf_alloc(rgpu_wavefront* w, size_t bytes, void** out) {
    RGPU_HIP(w->ctx, hipMalloc(out, bytes < 16 ? 16 : bytes));
    w->owned.push_back(*out);
    return RGPU_OK;
}

hipError_t step_once(rgpu_wavefront* w, hipStream_t s, hipEvent_t* marks = nullptr) {
    if (w->is_batch())
        return wf_step(w->st, w->params, TraceScene{}, DevMaterials{}, DevEnvmap{}, Camera{}, w->guide, w->variant,
                       w->queue_blocks, w->trace_block, s, &w->djobs, marks);
    return wf_step(w->st, w->params, w->scene->dev, w->mats->dev, w->env->dev, w->cam, w->guide, w->variant,
                   w->queue_blocks, w->trace_block, s, nullptr, marks);
}

Camera to_camera(const rgpu_camera* cam) {
    Camera c{};
    for (int k = 0; k < 3; ++k) {
        c.eye[k] = cam->eye[k];
        c.forward[k] = cam->forward[k];
        c.up[k] = cam->up[k];
        c.right[k] = cam->right[k];
        c.o[k] = cam->imageplane_o[k];
        c.r[k] = cam->imageplane_r[k];
        c.b[k] = cam->imageplane_b[k];
    }
    c.width = cam->width;
    c.height = cam->height;
    c.lens_radius = cam->lens_radius;
    return c;
}

// Options shared by single renders and batches.
int parse_config(rgpu_ctx* ctx, const rgpu_wavefront_config* cfg, rgpu_wavefront* w) {
    if (cfg->abi_version != RGPU_ABI_VERSION) return fail(ctx, RGPU_ERR_INVALID_ARG, "abi_version mismatch");
    if (!valid_variant((int)cfg->variant))
        return fail(ctx, RGPU_ERR_INVALID_ARG, "unknown traversal variant");
    WFParams& p = w->params;
    p.rr_start = cfg->rr_start ? cfg->rr_start : 25u;
    p.max_depth = cfg->max_depth;
    p.accum64 = (cfg->flags & RGPU_WF_ACCUM64) ? 1u : 0u;
    p.quirks = (cfg->flags & RGPU_WF_NO_REFERENCE_QUIRKS) ? 0u : 1u;
    p.consistent = (cfg->flags & RGPU_WF_CONSISTENT_NORMALS) ? 1u : 0u;
    p.clamp = cfg->radiance_clamp > 0.0f ? cfg->radiance_clamp : 1000.0f;
    p.bxdf_p = cfg->bxdf_select_p > 0.0f ? cfg->bxdf_select_p : 0.5f;
    w->seed = cfg->seed;
    w->variant = (int)cfg->variant;
    w->queue_blocks = cfg->queue_blocks ? (int)cfg->queue_blocks : (ctx->cu_count > 0 ? ctx->cu_count * 8 : 2048);
    return RGPU_OK;
}

// Slot state for n path slots.
int alloc_slots(rgpu_wavefront* w, size_t n) {
    w->params.n = (uint32_t)n;
    DevState& s = w->st;
    s.nblocks = (uint32_t)((n + 255) / 256);
    int rc = hipSetDevice(w->ctx->device) == hipSuccess ? RGPU_OK : RGPU_ERR_HIP;
    void* ptr;
    auto A = [&](size_t bytes) -> void* {
        if (rc != RGPU_OK) return nullptr;
        rc = wf_alloc(w, bytes, &ptr);
        return rc == RGPU_OK ? ptr : nullptr;
    };
    s.ray = (TraceRay*)A(n * sizeof(TraceRay));
    s.path = (PathRec*)A(n * sizeof(PathRec));
    s.ext = (ExtRec*)A(n * sizeof(ExtRec));
    s.shade = (ShadeRec*)A(n * sizeof(ShadeRec));
    s.vol = (VolList*)A(n * sizeof(VolList));
    s.rng = (uint4*)A(n * sizeof(uint4));
    s.qsel = (uint8_t*)A(n);
    s.items = (uint32_t*)A(n * kNumQueues * sizeof(uint32_t));
    s.counts = (uint32_t*)A(kNumQueues * sizeof(uint32_t));
    s.block_counts = (uint32_t*)A((size_t)s.nblocks * (kNumQueues + 1) * sizeof(uint32_t));
    s.next_pixel = (unsigned long long*)A(sizeof(unsigned long long));
    w->stat_total = (unsigned long long*)A(sizeof(unsigned long long));
    return rc;
}

// (Re)start every job of a batch: cursors to zero, slots seeded, accumulation cleared.
int restart_batch(rgpu_wavefront* w) {
    rgpu_ctx* ctx = w->ctx;
    std::vector<JobDesc> table(w->jobs.size());
    for (size_t j = 0; j < w->jobs.size(); ++j) {
        const rgpu_wavefront::Job& J = w->jobs[j];
        JobDesc& d = table[j];
        d.sc = J.scene->dev;
        d.m = J.mats->dev;
        d.env = J.env->dev;
        d.cam = J.cam;
        d.accum = J.accum;
        d.next_pixel = 0;
        d.budget = (uint64_t)J.npixels * J.spp;
        d.seed = J.seed;
        d.first_block = J.first_block;
        d.nblocks = J.nblocks;
        d.gframe = J.gframe;
    }
    RGPU_HIP(ctx, hipMemcpy(w->djobs.jobs, table.data(), table.size() * sizeof(JobDesc), hipMemcpyHostToDevice));
    RGPU_HIP(ctx, wf_initialize_jobs(w->st, w->params, w->djobs, ctx->stream));
    for (const rgpu_wavefront::Job& J : w->jobs) {
        DevState js = w->st;
        js.accum = J.accum;
        RGPU_HIP(ctx, wf_clear_accum(js, J.npixels, false, ctx->stream));
    }
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RGPU_OK;
}

int restart(rgpu_wavefront* w, const rgpu_camera* camera) {
    if (w->is_batch()) {
        drop_graph(w);
        return restart_batch(w);
    }
    if (!camera) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "a camera is needed to restart a single render");
    return rgpu_wavefront_initialize(w, camera, 0, 1);
}

}  // namespace

extern "C" {

int rgpu_materials_create(rgpu_ctx* ctx, uint32_t nprims, const int32_t* types, const int32_t* indices,
                          const rgpu_lambertian* lambertians, uint32_t nl, const rgpu_ward* wards, uint32_t nw,
                          const rgpu_dielectric* dielectrics, uint32_t nd, const rgpu_volume* volumes, uint32_t nv,
                          rgpu_materials** out) {
    if (!ctx || !out || (nprims && (!types || !indices))) return RGPU_ERR_INVALID_ARG;
    *out = nullptr;
    std::vector<Material> mat(nprims);
    for (uint32_t i = 0; i < nprims; ++i) {
        int32_t t = types[i], k = indices[i];
        uint32_t lim = t == kMatLambertian ? nl : t == kMatWard ? nw : t == kMatDielectric ? nd : t == kMatVolume ? nv : 1;
        if (t < kMatLambertian || t > kMatVolume)
            return fail(ctx, RGPU_ERR_INVALID_ARG, "material type must be 1..5 (lambertian..volume)");
        if (t != kMatSpecular && (k < 0 || (uint32_t)k >= lim))
            return fail(ctx, RGPU_ERR_INVALID_ARG, "material index out of range");
        mat[i] = Material{t, k};
    }
    std::vector<Lambertian> L(nl);
    for (uint32_t i = 0; i < nl; ++i) {
        std::memcpy(L[i].Le, lambertians[i].Le, 12);
        std::memcpy(L[i].R, lambertians[i].R, 12);
        L[i].back_emission = lambertians[i].back_emission;
        L[i].pad = 0;
    }
    std::vector<Ward> W(nw);
    for (uint32_t i = 0; i < nw; ++i) {
        W[i].alpha_x = wards[i].alpha_x;
        W[i].alpha_y = wards[i].alpha_y;
        std::memcpy(W[i].reflectance, wards[i].reflectance, 12);
        std::memcpy(W[i].edgetint, wards[i].edgetint, 12);
        W[i].falloff = wards[i].falloff;
        W[i].pad = 0;
        if (!(W[i].alpha_x > 0.0f) || !(W[i].alpha_y > 0.0f) || !(W[i].falloff > 0.0f))
            return fail(ctx, RGPU_ERR_INVALID_ARG, "Ward roughness and falloff must be positive");
    }
    std::vector<Dielectric> D(nd);
    for (uint32_t i = 0; i < nd; ++i) D[i].ior = dielectrics[i].ior;
    std::vector<Volume> V(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        V[i].C = volumes[i].C;
        std::memcpy(V[i].R, volumes[i].R, 12);
        if (!(V[i].C > 0.0f)) return fail(ctx, RGPU_ERR_INVALID_ARG, "volume extinction must be positive");
    }
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    rgpu_materials* m = new (std::nothrow) rgpu_materials();
    if (!m) return RGPU_ERR_OUT_OF_MEMORY;
    m->ctx = ctx;
    m->nprims = nprims;
    uint64_t fp = fingerprint64(mat.data(), mat.size() * sizeof(Material), nprims);
    fp = fingerprint64(L.data(), L.size() * sizeof(Lambertian), fp ^ nl);
    fp = fingerprint64(W.data(), W.size() * sizeof(Ward), fp ^ nw);
    fp = fingerprint64(D.data(), D.size() * sizeof(Dielectric), fp ^ nd);
    m->fingerprint = fingerprint64(V.data(), V.size() * sizeof(Volume), fp ^ nv);
    int rc = upload_vec(ctx, mat.data(), mat.size(), &m->bufs[0]);
    if (rc == RGPU_OK) rc = upload_vec(ctx, L.data(), L.size(), &m->bufs[1]);
    if (rc == RGPU_OK) rc = upload_vec(ctx, W.data(), W.size(), &m->bufs[2]);
    if (rc == RGPU_OK) rc = upload_vec(ctx, D.data(), D.size(), &m->bufs[3]);
    if (rc == RGPU_OK) rc = upload_vec(ctx, V.data(), V.size(), &m->bufs[4]);
    if (rc != RGPU_OK) {
        rgpu_materials_destroy(m);
        return rc;
    }
    m->dev.mat = (const Material*)m->bufs[0];
    m->dev.lambertian = (const Lambertian*)m->bufs[1];
    m->dev.ward = (const Ward*)m->bufs[2];
    m->dev.dielectric = (const Dielectric*)m->bufs[3];
    m->dev.volume = (const Volume*)m->bufs[4];
    *out = m;
    return RGPU_OK;
}

int rgpu_materials_destroy(rgpu_materials* m) {
    if (!m) return RGPU_ERR_INVALID_ARG;
    (void)hipSetDevice(m->ctx->device);
    for (void* p : m->bufs)
        if (p) (void)hipFree(p);
    delete m;
    return RGPU_OK;
}

int rgpu_envmap_create(rgpu_ctx* ctx, const float* rgb, uint32_t channels, uint32_t width, uint32_t height,
                       rgpu_envmap** out) {
    if (!out || !rgb || width == 0 || height == 0 || (channels != 3 && channels != 4)) return RGPU_ERR_INVALID_ARG;
    *out = nullptr;
    if ((uint64_t)width * height >= (1ull << 31)) return fail(ctx, RGPU_ERR_INVALID_ARG, "envmap too large");
    rgpu_envmap* e = new (std::nothrow) rgpu_envmap();
    if (!e) return RGPU_ERR_OUT_OF_MEMORY;
    e->ctx = ctx;
    e->fingerprint = fingerprint64(rgb, (size_t)width * height * channels * sizeof(float),
                                   ((uint64_t)width << 32) | height);
    try {
        build_envmap_tables(rgb, channels, width, height, &e->host);
    } catch (const std::exception& ex) {
        delete e;
        return fail(ctx, RGPU_ERR_INVALID_ARG, ex.what());
    }
    if (!ctx) {  // host tables only
        *out = e;
        return RGPU_OK;
    }
    const size_t P = (size_t)width * height;
    std::vector<EnvTexelRec> rec(P);
    std::vector<AliasBucket> alias(7 * P);
    for (size_t i = 0; i < P; ++i) {
        EnvTexelRec& r = rec[i];
        for (int t = 0; t < 7; ++t) r.pdf[t] = e->host.pdf[t][i];
        r.pdf[7] = 0.0f;
        r.frag = e->host.fragments[i];
        r.pad[0] = r.pad[1] = r.pad[2] = r.pad[3] = 0.0f;
    }
    for (int t = 0; t < 7; ++t)
        std::memcpy(alias.data() + t * P, e->host.alias[t].data(), P * sizeof(AliasBucket));
    int rc = hipSetDevice(ctx->device) == hipSuccess ? RGPU_OK : RGPU_ERR_HIP;
    if (rc == RGPU_OK) rc = upload_vec(ctx, e->host.texels.data(), e->host.texels.size(), &e->bufs[0]);
    if (rc == RGPU_OK) rc = upload_vec(ctx, rec.data(), rec.size(), &e->bufs[1]);
    if (rc == RGPU_OK) rc = upload_vec(ctx, alias.data(), alias.size(), &e->bufs[2]);
    if (rc == RGPU_OK) rc = upload_vec(ctx, e->host.pdf[0].data(), P, &e->bufs[3]);
    if (rc != RGPU_OK) {
        rgpu_envmap_destroy(e);
        return rc;
    }
    e->dev.texels = (const float4*)e->bufs[0];
    e->dev.rec = (const EnvTexelRec*)e->bufs[1];
    e->dev.alias = (const AliasBucket*)e->bufs[2];
    e->dev.pdf0 = (const float*)e->bufs[3];
    e->dev.width = width;
    e->dev.height = height;
    *out = e;
    return RGPU_OK;
}

int rgpu_envmap_destroy(rgpu_envmap* e) {
    if (!e) return RGPU_ERR_INVALID_ARG;
    if (e->ctx) (void)hipSetDevice(e->ctx->device);
    for (void* p : e->bufs)
        if (p) (void)hipFree(p);
    delete e;
    return RGPU_OK;
}

int rgpu_envmap_export(const rgpu_envmap* e, int table, float* pdf, float* heights, int32_t* aliases,
                       float* fragments, int32_t* degenerate) {
    if (!e || table < 0 || table > 6) return RGPU_ERR_INVALID_ARG;
    const size_t P = (size_t)e->host.width * e->host.height;
    for (size_t i = 0; i < P; ++i) {
        if (pdf) pdf[i] = e->host.pdf[table][i];
        if (heights) heights[i] = e->host.alias[table][i].height;
        if (aliases) aliases[i] = e->host.alias[table][i].alias;
        if (fragments) {
            const EnvFragment& f = e->host.fragments[i];
            fragments[4 * i + 0] = f.beg_y;
            fragments[4 * i + 1] = f.end_y;
            fragments[4 * i + 2] = f.beg_phi;
            fragments[4 * i + 3] = f.end_phi;
        }
    }
    if (degenerate) *degenerate = e->host.degenerate[table] ? 1 : 0;
    return RGPU_OK;
}

int rgpu_envmap_sample(rgpu_ctx* ctx, rgpu_stream stream, const rgpu_envmap* e, uint32_t* states,
                       const float* normals, uint32_t count, float* out) {
    if (!ctx || !e) return RGPU_ERR_INVALID_ARG;
    if (e->ctx != ctx) return fail(ctx, RGPU_ERR_INVALID_ARG, "envmap has no device tables on this context");
    if (count && (!states || !out)) return fail(ctx, RGPU_ERR_INVALID_ARG, "null state or output buffer");
    if ((((uintptr_t)states | (uintptr_t)out) & 15) || ((uintptr_t)normals & 3))
        return fail(ctx, RGPU_ERR_INVALID_ARG, "states and out need 16-byte alignment, normals 4-byte");
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, envmap_sample(e->dev, states, normals, count, out, stream_of(ctx, stream)));
    return RGPU_OK;
}

int rgpu_envmap_pdf(rgpu_ctx* ctx, rgpu_stream stream, const rgpu_envmap* e, const float* dirs, const float* normals,
                    uint32_t count, float* pdfs) {
    if (!ctx || !e) return RGPU_ERR_INVALID_ARG;
    if (e->ctx != ctx) return fail(ctx, RGPU_ERR_INVALID_ARG, "envmap has no device tables on this context");
    if (count && (!dirs || !pdfs)) return fail(ctx, RGPU_ERR_INVALID_ARG, "null direction or pdf buffer");
    if (((uintptr_t)dirs | (uintptr_t)normals | (uintptr_t)pdfs) & 3)
        return fail(ctx, RGPU_ERR_INVALID_ARG, "directions, normals and pdfs need 4-byte alignment");
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, envmap_pdf(e->dev, dirs, normals, count, pdfs, stream_of(ctx, stream)));
    return RGPU_OK;
}

int rgpu_camera_look_at(const float* eye, const float* target, const float* up, float fov_y_deg,
                        float focus_distance, uint32_t width, uint32_t height, float lens_radius, rgpu_camera* out) {
    if (!eye || !target || !up || !out || width == 0 || height == 0 || !(fov_y_deg > 0.0f) || !(fov_y_deg < 180.0f))
        return RGPU_ERR_INVALID_ARG;
    double f[3], r[3], u[3];
    for (int k = 0; k < 3; ++k) f[k] = (double)target[k] - eye[k];
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (!(fl > 0.0)) return RGPU_ERR_INVALID_ARG;
    for (double& x : f) x /= fl;
    r[0] = f[1] * up[2] - f[2] * up[1];
    r[1] = f[2] * up[0] - f[0] * up[2];
    r[2] = f[0] * up[1] - f[1] * up[0];
    double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (!(rl > 0.0)) return RGPU_ERR_INVALID_ARG;
    for (double& x : r) x /= rl;
    u[0] = r[1] * f[2] - r[2] * f[1];
    u[1] = r[2] * f[0] - r[0] * f[2];
    u[2] = r[0] * f[1] - r[1] * f[0];
    double dist = focus_distance > 0.0f ? focus_distance : fl;
    double ph = 2.0 * dist * std::tan(fov_y_deg * 3.14159265358979323846 / 360.0);  // object-plane height
    double pw = ph * width / height;
    std::memset(out, 0, sizeof(*out));
    for (int k = 0; k < 3; ++k) {
        out->eye[k] = eye[k];
        out->forward[k] = (float)f[k];
        out->up[k] = (float)u[k];
        out->right[k] = (float)r[k];
        // image plane: top-left corner, then right / down steps per pixel (standardCamera,
        // common/wavefront_path_tracing.hpp:103-118)
        out->imageplane_o[k] = (float)(eye[k] + f[k] * dist - r[k] * pw * 0.5 + u[k] * ph * 0.5);
        out->imageplane_r[k] = (float)(r[k] * pw / width);
        out->imageplane_b[k] = (float)(-u[k] * ph / height);
    }
    out->width = width;
    out->height = height;
    out->lens_radius = lens_radius;
    return RGPU_OK;
}

// (Re)allocates the accumulation and tonemap buffers of a single render for P pixels.
// The published image and mutex of the preview handoff, sized to the lane's image and reset (mutex
// free, nothing published). Called when publishing is switched on and whenever the lane is
// (re)initialised; never while a preview call may run.
static int ensure_publish(rgpu_wavefront* w) {
    if (!w->publish_every || !w->npixels) return RGPU_OK;
    rgpu_ctx* ctx = w->ctx;
    void* ptr;
    int rc;
    if (!w->pub_words) {
        if ((rc = wf_alloc(w, 16, &ptr)) != RGPU_OK) return rc;
        w->pub_words = (int*)ptr;
    }
    if (w->pub_pixels != w->npixels) {
        for (auto it = w->owned.begin(); it != w->owned.end(); ++it)
            if (*it == (void*)w->pub_image) {
                (void)hipFree(*it);
                w->owned.erase(it);
                break;
            }
        w->pub_image = nullptr;
        w->pub_pixels = 0;
        if ((rc = wf_alloc(w, (size_t)w->npixels * sizeof(RGB16), &ptr)) != RGPU_OK) return rc;
        w->pub_image = (RGB16*)ptr;
        w->pub_pixels = w->npixels;
    }
    RGPU_HIP(ctx, hipMemsetAsync(w->pub_image, 0, (size_t)w->pub_pixels * sizeof(RGB16), ctx->stream));
    RGPU_HIP(ctx, handoff_init(w->pub_words, ctx->stream));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    w->publish_tick = 0;
    return RGPU_OK;
}

static int ensure_accum(rgpu_wavefront* w, uint32_t P) {
    if (P == w->npixels) return RGPU_OK;
    for (void* q : {(void*)w->st.accum, (void*)w->st.accum64, (void*)w->st.accum_count, (void*)w->tonemap_buf}) {
        if (!q) continue;
        for (auto it = w->owned.begin(); it != w->owned.end(); ++it)
            if (*it == q) {
                (void)hipFree(q);
                w->owned.erase(it);
                break;
            }
    }
    w->st.accum = nullptr;
    w->st.accum64 = nullptr;
    w->st.accum_count = nullptr;
    w->tonemap_buf = nullptr;
    w->npixels = 0;
    void* ptr;
    int rc;
    if (w->params.accum64) {
        if ((rc = wf_alloc(w, (size_t)P * 3 * sizeof(double), &ptr)) != RGPU_OK) return rc;
        w->st.accum64 = (double*)ptr;
        if ((rc = wf_alloc(w, (size_t)P * sizeof(uint32_t), &ptr)) != RGPU_OK) return rc;
        w->st.accum_count = (uint32_t*)ptr;
    } else {
        if ((rc = wf_alloc(w, (size_t)P * 4 * sizeof(float), &ptr)) != RGPU_OK) return rc;
        w->st.accum = (float*)ptr;
    }
    if ((rc = wf_alloc(w, (size_t)P * 4, &ptr)) != RGPU_OK) return rc;
    w->tonemap_buf = (uint8_t*)ptr;
    w->npixels = P;
    return RGPU_OK;
}

int rgpu_wavefront_destroy(rgpu_wavefront* w) {
    if (!w) return RGPU_ERR_INVALID_ARG;
    (void)hipSetDevice(w->ctx->device);
    (void)hipStreamSynchronize(w->ctx->stream);
    drop_graph(w);
    detach_from_guide(w);
    for (void* p : w->owned) (void)hipFree(p);
    delete w;
    return RGPU_OK;
}

int rgpu_wavefront_create(rgpu_ctx* ctx, const rgpu_scene* scene, const rgpu_materials* mats, const rgpu_envmap* env,
                          const rgpu_wavefront_config* cfg, rgpu_wavefront** out) {
    if (!ctx || !scene || !mats || !env || !cfg || !out) return RGPU_ERR_INVALID_ARG;
    *out = nullptr;
    if (!env->ctx) return fail(ctx, RGPU_ERR_INVALID_ARG, "envmap was built without a context (host tables only)");
    if (cfg->path_count == 0 || cfg->path_count > (1u << 28))
        return fail(ctx, RGPU_ERR_INVALID_ARG, "path_count must be in 1..2^28");
    if (mats->nprims != scene->nprims)
        return fail(ctx, RGPU_ERR_INVALID_ARG, "materials must give exactly one entry per primitive");
    rgpu_wavefront* w = new (std::nothrow) rgpu_wavefront();
    if (!w) return RGPU_ERR_OUT_OF_MEMORY;
    w->ctx = ctx;
    w->scene = scene;
    w->mats = mats;
    w->env = env;
    int rc = parse_config(ctx, cfg, w);
    if (rc == RGPU_OK) rc = alloc_slots(w, cfg->path_count);
    if (rc == RGPU_OK) {
        // the persistent trace of a single render casts the extension rays through the ray-cast
        // kernel: 32 + 16 bytes per slot
        void* ptr = nullptr;
        rc = wf_alloc(w, (size_t)cfg->path_count * sizeof(TraceRay), &ptr);
        w->st.xray = (TraceRay*)ptr;
        if (rc == RGPU_OK) rc = wf_alloc(w, (size_t)cfg->path_count * sizeof(TraceHit), &ptr);
        w->st.xhit = (TraceHit*)ptr;
    }
    if (rc != RGPU_OK) {
        rgpu_wavefront_destroy(w);
        return rc;
    }
    *out = w;
    return RGPU_OK;
}

int rgpu_wavefront_create_batch(rgpu_ctx* ctx, const rgpu_batch_job* jobs, uint32_t njobs,
                                const rgpu_wavefront_config* cfg, rgpu_wavefront** out) {
    if (!ctx || !jobs || njobs == 0 || !cfg || !out) return RGPU_ERR_INVALID_ARG;
    *out = nullptr;
    if (njobs > 65535) return fail(ctx, RGPU_ERR_INVALID_ARG, "at most 65535 jobs in a batch");
    if (cfg->flags & RGPU_WF_ACCUM64) return fail(ctx, RGPU_ERR_UNSUPPORTED, "batches accumulate in float");
    uint64_t blocks = 0;
    for (uint32_t j = 0; j < njobs; ++j) {
        const rgpu_batch_job& b = jobs[j];
        if (!b.scene || !b.materials || !b.envmap || b.path_count == 0 || b.camera.width == 0 || b.camera.height == 0)
            return fail(ctx, RGPU_ERR_INVALID_ARG, "batch job needs a scene, materials, envmap, camera and slots");
        if (!b.envmap->ctx) return fail(ctx, RGPU_ERR_INVALID_ARG, "envmap was built without a context");
        if (b.materials->nprims != b.scene->nprims)
            return fail(ctx, RGPU_ERR_INVALID_ARG, "materials must give exactly one entry per primitive");
        if ((uint64_t)b.camera.width * b.camera.height >= (1ull << 31))
            return fail(ctx, RGPU_ERR_INVALID_ARG, "job resolution too large");
        blocks += (b.path_count + kJobBlock - 1) / kJobBlock;
    }
    if (blocks * kJobBlock > (1ull << 28)) return fail(ctx, RGPU_ERR_INVALID_ARG, "batch exceeds 2^28 slots");
    rgpu_wavefront* w = new (std::nothrow) rgpu_wavefront();
    if (!w) return RGPU_ERR_OUT_OF_MEMORY;
    w->ctx = ctx;
    int rc = parse_config(ctx, cfg, w);
    if (rc == RGPU_OK) rc = alloc_slots(w, (size_t)blocks * kJobBlock);
    std::vector<uint16_t> job_of_block(blocks);
    std::vector<const rgpu_scene*> classes;  // distinct scenes in order of first use
    uint32_t fb = 0;
    for (uint32_t j = 0; j < njobs && rc == RGPU_OK; ++j) {
        const rgpu_batch_job& b = jobs[j];
        rgpu_wavefront::Job J{};
        // guide inputs: positions normalised to the scene's root box (grown by 1e-3 on each
        // side), the scene's class code 1..kGClasses; scenes past the kGClasses-th get none
        uint32_t cls = 0;
        for (size_t c = 0; c < classes.size(); ++c)
            if (classes[c] == b.scene) cls = (uint32_t)c + 1;
        if (cls == 0 && classes.size() < (size_t)kGClasses) {
            classes.push_back(b.scene);
            cls = (uint32_t)classes.size();
        }
        J.gframe.cls = cls;
        for (int k = 0; k < 3; ++k) {
            const float lo = b.scene->lo[k] - 1e-3f, ext = b.scene->hi[k] + 1e-3f - lo;
            J.gframe.lo[k] = lo;
            J.gframe.inv_ext[k] = ext > 0.0f ? 1.0f / ext : 1.0f;
        }
        J.scene = b.scene;
        J.mats = b.materials;
        J.env = b.envmap;
        J.cam = to_camera(&b.camera);
        J.npixels = b.camera.width * b.camera.height;
        J.spp = b.spp;
        J.seed = b.seed;
        J.first_block = fb;
        J.nblocks = (b.path_count + kJobBlock - 1) / kJobBlock;
        void* p = nullptr;
        rc = wf_alloc(w, (size_t)J.npixels * 4 * sizeof(float), &p);
        J.accum = (float*)p;
        for (uint32_t k = 0; k < J.nblocks; ++k) job_of_block[fb + k] = (uint16_t)j;
        fb += J.nblocks;
        w->jobs.push_back(J);
    }
    void* p = nullptr;
    if (rc == RGPU_OK) rc = wf_alloc(w, njobs * sizeof(JobDesc), &p);
    w->djobs.jobs = (JobDesc*)p;
    if (rc == RGPU_OK) rc = wf_alloc(w, blocks * sizeof(uint16_t), &p);
    if (rc == RGPU_OK) {
        w->djobs.job_of_block = (const uint16_t*)p;
        w->djobs.njobs = njobs;
        if (hipMemcpy(p, job_of_block.data(), blocks * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(ctx, RGPU_ERR_HIP, "upload of the job map failed");
    }
    if (rc == RGPU_OK) rc = restart_batch(w);
    if (rc != RGPU_OK) {
        rgpu_wavefront_destroy(w);
        return rc;
    }
    w->initialized = true;
    *out = w;
    return RGPU_OK;
}

int rgpu_wavefront_job_count(const rgpu_wavefront* w, uint32_t* njobs) {
    if (!w || !njobs) return RGPU_ERR_INVALID_ARG;
    *njobs = (uint32_t)w->jobs.size();
    return RGPU_OK;
}

// A job's queue run in queue q spans the scanned offsets of its first block and of the next
// job's first block (the queue's total for the last job).
int rgpu_wavefront_job_queue_counts(rgpu_wavefront* w, uint32_t job, uint32_t* counts7, uint64_t* samples_started) {
    if (!w || !counts7 || job >= w->jobs.size()) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t nb = w->st.nblocks;
    const bool last = job + 1 == w->jobs.size();
    uint32_t totals[kNumQueues];
    RGPU_HIP(ctx, hipMemcpy(totals, w->st.counts, sizeof(totals), hipMemcpyDeviceToHost));
    for (int q = 0; q < kNumQueues; ++q) {
        uint32_t lo = 0, hi = totals[q];
        RGPU_HIP(ctx, hipMemcpy(&lo, w->st.block_counts + (size_t)q * nb + w->jobs[job].first_block, 4,
                                hipMemcpyDeviceToHost));
        if (!last)
            RGPU_HIP(ctx, hipMemcpy(&hi, w->st.block_counts + (size_t)q * nb + w->jobs[job + 1].first_block, 4,
                                    hipMemcpyDeviceToHost));
        counts7[q] = hi - lo;
    }
    if (samples_started) {
        unsigned long long c = 0;
        RGPU_HIP(ctx, hipMemcpy(&c, &w->djobs.jobs[job].next_pixel, 8, hipMemcpyDeviceToHost));
        const uint64_t budget = (uint64_t)w->jobs[job].npixels * w->jobs[job].spp;
        *samples_started = budget != 0 && c > budget ? budget : c;
    }
    return RGPU_OK;
}

int rgpu_wavefront_job_accumulation(rgpu_wavefront* w, uint32_t job, float* rgb, uint32_t* counts) {
    if (!w || job >= w->jobs.size()) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t P = w->jobs[job].npixels;
    std::vector<float> a(P * 4);
    RGPU_HIP(ctx, hipMemcpy(a.data(), w->jobs[job].accum, P * 16, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < P; ++i) {
        if (rgb) {
            rgb[3 * i] = a[4 * i];
            rgb[3 * i + 1] = a[4 * i + 1];
            rgb[3 * i + 2] = a[4 * i + 2];
        }
        if (counts) std::memcpy(&counts[i], &a[4 * i + 3], 4);
    }
    return RGPU_OK;
}

int rgpu_wavefront_job_stat(rgpu_wavefront* w, uint32_t job, uint64_t* total_samples) {
    if (!w || job >= w->jobs.size() || !total_samples) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    DevState js = w->st;
    js.accum = w->jobs[job].accum;
    RGPU_HIP(ctx, hipMemsetAsync(w->stat_total, 0, 8, ctx->stream));
    RGPU_HIP(ctx, wf_stat(js, w->jobs[job].npixels, false, w->stat_total, ctx->stream));
    unsigned long long t = 0;
    RGPU_HIP(ctx, hipMemcpyAsync(&t, w->stat_total, 8, hipMemcpyDeviceToHost, ctx->stream));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *total_samples = t;
    return RGPU_OK;
}

int rgpu_wavefront_job_tonemap(rgpu_wavefront* w, uint32_t job, uint8_t* rgba8) {
    if (!w || job >= w->jobs.size() || !rgba8) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t P = w->jobs[job].npixels;
    void* buf = nullptr;
    RGPU_HIP(ctx, hipMalloc(&buf, (size_t)P * 4));
    DevState js = w->st;
    js.accum = w->jobs[job].accum;
    hipError_t e = wf_tonemap(js, P, false, buf, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rgba8, buf, (size_t)P * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return hip_fail(ctx, e, "rgpu_wavefront_job_tonemap");
    return RGPU_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ tonemap without allocation
// Staging for rgpu_wavefront_job_tonemap_noalloc: the RGBA8 image of one job of at most
// kNoAllocMaxPixels pixels (4 MiB). Every device the library runs on holds its own copy until
// the process ends. Host code never passes `scratch` itself to a kernel or a copy (on the host
// the name is only a stand-in); it uses the per-device address that hipGetSymbolAddress gives.
constexpr uint32_t kNoAllocMaxPixels = 1024u * 1024u;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) uint8_t scratch[(size_t)kNoAllocMaxPixels * 4];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on `device`, which must be the current device. Looked up on
// the first call for that device and kept. Allocates nothing.
hipError_t scratch_address(int device, void** out) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}
}  // namespace

extern "C" {

int rgpu_wavefront_job_tonemap_noalloc(rgpu_wavefront* w, uint32_t job, uint8_t* rgba8) {
    if (!w || job >= w->jobs.size() || !rgba8) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = w->ctx;
    const uint32_t P = w->jobs[job].npixels;
    if (P > kNoAllocMaxPixels)
        return fail(ctx, RGPU_ERR_INVALID_ARG,
                    "rgpu_wavefront_job_tonemap_noalloc: the job has " + std::to_string(P) +
                        " pixels; the static staging holds at most 1048576 (1024x1024)");
    if (ctx->device < 0 || ctx->device >= kNoAllocMaxDevices)
        return fail(ctx, RGPU_ERR_INVALID_ARG, "rgpu_wavefront_job_tonemap_noalloc: device index above 63");
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    void* buf = nullptr;
    RGPU_HIP(ctx, scratch_address(ctx->device, &buf));
    DevState js = w->st;
    js.accum = w->jobs[job].accum;
    // every pixel of the staging is written before it is read back: nothing to clear
    RGPU_HIP(ctx, wf_tonemap(js, P, false, buf, ctx->stream));
    RGPU_HIP(ctx, hipMemcpyAsync(rgba8, buf, (size_t)P * 4, hipMemcpyDeviceToHost, ctx->stream));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RGPU_OK;
}

// Steps until no slot is live (every job has started and finished its samples) or
// max_steps, checking every `check_every` steps.
int rgpu_wavefront_run_batch(rgpu_wavefront* w, uint32_t max_steps, uint32_t check_every, uint32_t* steps_done) {
    if (!w || !w->is_batch() || check_every == 0) return RGPU_ERR_INVALID_ARG;
    for (int j = 0; j < (int)w->jobs.size(); ++j)
        if (w->jobs[j].spp == 0) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "every job of a run needs an spp budget");
    uint32_t done = 0;
    while (done < max_steps) {
        const uint32_t k = std::min(check_every, max_steps - done);
        int rc = rgpu_wavefront_step(w, nullptr, k);
        if (rc != RGPU_OK) return rc;
        done += k;
        if (w->guide_obj && w->guide.train) {
            rc = rgpu_wavefront_guide_train(w, w->guide_obj->dev.capacity, nullptr);
            if (rc != RGPU_OK) return rc;
        }
        uint32_t c[kNumQueues];
        rc = rgpu_wavefront_queue_counts(w, c, nullptr);
        if (rc != RGPU_OK) return rc;
        uint64_t live = 0;
        for (int q = 0; q < kNumQueues; ++q) live += c[q];
        if (live == 0) break;
    }
    if (steps_done) *steps_done = done;
    return RGPU_OK;
}

int rgpu_wavefront_initialize(rgpu_wavefront* w, const rgpu_camera* cam, uint32_t lane_index, uint32_t num_lanes) {
    if (!w || !cam || num_lanes == 0 || lane_index >= num_lanes || cam->width == 0 || cam->height == 0)
        return RGPU_ERR_INVALID_ARG;
    if (w->is_batch()) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "a batch is initialised at creation");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(w);  // a captured graph holds the old camera and buffers by value
    const uint32_t P = cam->width * cam->height;
    int arc = ensure_accum(w, P);
    if (arc != RGPU_OK) return arc;
    w->cam = to_camera(cam);
    // WavefrontLane::initialize (common/wavefront_path_tracing.hpp:532-570): pixel offset
    // lane * (P / lanes), RNG offset lane * N, every slot in the new-path queue.
    const uint64_t next_pixel = (uint64_t)lane_index * (P / num_lanes);
    const uint64_t seed_offset = w->seed + (uint64_t)lane_index * w->params.n;
    RGPU_HIP(ctx, wf_initialize(w->st, w->params, seed_offset, next_pixel, P, ctx->stream));
    RGPU_HIP(ctx, wf_clear_accum(w->st, P, w->params.accum64 != 0, ctx->stream));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if ((arc = ensure_publish(w)) != RGPU_OK) return arc;
    w->initialized = true;
    return RGPU_OK;
}

int rgpu_wavefront_step(rgpu_wavefront* w, rgpu_stream stream, uint32_t steps) {
    if (!w) return RGPU_ERR_INVALID_ARG;
    if (!w->initialized) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "rgpu_wavefront_initialize was not called");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = stream_of(ctx, stream);
    for (uint32_t i = 0; i < steps; ++i) {
        RGPU_HIP(ctx, step_once(w, s));
        // the end of a step publishes the intermediate if the preview side does not hold it
        // (wavefront_path_tracing.hpp:877-898)
        if (w->publish_every && w->pub_image && ++w->publish_tick % w->publish_every == 0)
            RGPU_HIP(ctx, handoff_publish(w->st, w->npixels, w->params.accum64 != 0, w->pub_image, w->pub_words, s));
    }
    return RGPU_OK;
}

int rgpu_wavefront_set_publish(rgpu_wavefront* w, uint32_t every_steps) {
    if (!w) return RGPU_ERR_INVALID_ARG;
    if (w->is_batch()) return fail(w->ctx, RGPU_ERR_UNSUPPORTED, "previews cover single renders, not batches");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(w);  // a captured graph holds the previous publishing choice
    w->publish_every = every_steps;
    return ensure_publish(w);
}

int rgpu_wavefront_update_intermediate(rgpu_wavefront* w, rgpu_stream stream, rgpu_rgb16* dst, int* took) {
    if (!w || !dst) return RGPU_ERR_INVALID_ARG;
    if (!w->initialized) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "rgpu_wavefront_initialize was not called");
    if (!w->publish_every || !w->pub_image)
        return fail(w->ctx, RGPU_ERR_INVALID_ARG, "publishing is off (rgpu_wavefront_set_publish)");
    if (((uintptr_t)dst | (uintptr_t)took) & 3) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "buffers need 4-byte alignment");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, handoff_take(w->pub_image, reinterpret_cast<RGB16*>(dst), w->pub_pixels, w->pub_words, took,
                               stream_of(ctx, stream)));
    return RGPU_OK;
}

int rgpu_wavefront_graph_capture(rgpu_wavefront* w, uint32_t steps) {
    if (!w || steps == 0) return RGPU_ERR_INVALID_ARG;
    if (!w->initialized) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "rgpu_wavefront_initialize was not called");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(w);
    hipStream_t cs;
    RGPU_HIP(ctx, hipStreamCreate(&cs));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    for (uint32_t i = 0; i < steps && e == hipSuccess; ++i) e = step_once(w, cs);
    // with publishing on, a replay publishes once, after its last step
    if (e == hipSuccess && w->publish_every && w->pub_image)
        e = handoff_publish(w->st, w->npixels, w->params.accum64 != 0, w->pub_image, w->pub_words, cs);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(cs, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&w->graph, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) {
        w->graph = nullptr;
        return hip_fail(ctx, e, "rgpu_wavefront_graph_capture");
    }
    w->graph_steps = steps;
    return RGPU_OK;
}

int rgpu_wavefront_graph_launch(rgpu_wavefront* w, rgpu_stream stream, uint32_t launches) {
    if (!w) return RGPU_ERR_INVALID_ARG;
    if (!w->graph) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "no captured graph (rgpu_wavefront_graph_capture)");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    for (uint32_t i = 0; i < launches; ++i) RGPU_HIP(ctx, hipGraphLaunch(w->graph, stream_of(ctx, stream)));
    return RGPU_OK;
}

int rgpu_wavefront_set_launch(rgpu_wavefront* w, int variant, uint32_t trace_block, uint32_t queue_blocks) {
    if (!w) return RGPU_ERR_INVALID_ARG;
    if (!valid_variant(variant))
        return fail(w->ctx, RGPU_ERR_INVALID_ARG, "unknown traversal variant");
    if (trace_block != 0 && trace_block != 64 && trace_block != 128 && trace_block != 256)
        return fail(w->ctx, RGPU_ERR_INVALID_ARG, "trace_block must be 0 (persistent), 64, 128 or 256");
    if (queue_blocks == 0 || queue_blocks > 65536)
        return fail(w->ctx, RGPU_ERR_INVALID_ARG, "queue_blocks must be in 1..65536");
    drop_graph(w);
    w->variant = variant;
    w->trace_block = (int)trace_block;
    w->queue_blocks = (int)queue_blocks;
    return RGPU_OK;
}

int rgpu_wavefront_get_launch(const rgpu_wavefront* w, int* variant, uint32_t* trace_block, uint32_t* queue_blocks) {
    if (!w) return RGPU_ERR_INVALID_ARG;
    if (variant) *variant = w->variant;
    if (trace_block) *trace_block = (uint32_t)w->trace_block;
    if (queue_blocks) *queue_blocks = (uint32_t)w->queue_blocks;
    return RGPU_OK;
}

// Times every launch configuration on the live wavefront (re-initialised with `camera`
// before each trial) and keeps the fastest; the wavefront is left initialised with it.
int rgpu_wavefront_autotune(rgpu_wavefront* w, const rgpu_camera* camera, uint32_t warmup, uint32_t steps,
                            double* best_ms_per_step) {
    if (!w || (!camera && !w->is_batch()) || steps == 0) return RGPU_ERR_INVALID_ARG;
    const int cu = w->ctx->cu_count > 0 ? w->ctx->cu_count : 256;
    const int variants[3] = {RGPU_VARIANT_STACKLESS, RGPU_VARIANT_THREADED, RGPU_VARIANT_THREADED_SINGLE};
    const uint32_t blocks[4] = {0, 64, 128, 256};
    const uint32_t qmul[4] = {2, 4, 8, 16};
    double best = 1e300;
    int bv = w->variant;
    uint32_t bb = (uint32_t)w->trace_block, bq = (uint32_t)w->queue_blocks;
    for (int v : variants)
        for (uint32_t b : blocks)
            for (uint32_t q : qmul) {
                if (b == 0 && w->is_batch()) continue;  // a batch traces 0 as 256
                int rc = rgpu_wavefront_set_launch(w, v, b, (uint32_t)cu * q);
                if (rc == RGPU_OK) rc = restart(w, camera);
                if (rc == RGPU_OK && warmup) rc = rgpu_wavefront_step(w, nullptr, warmup);
                double ms = 0.0;
                if (rc == RGPU_OK) rc = rgpu_wavefront_time_steps(w, steps, &ms);
                if (rc != RGPU_OK) return rc;
                if (ms < best) {
                    best = ms;
                    bv = v;
                    bb = b;
                    bq = (uint32_t)cu * q;
                }
            }
    int rc = rgpu_wavefront_set_launch(w, bv, bb, bq);
    if (rc == RGPU_OK) rc = restart(w, camera);
    if (rc == RGPU_OK && best_ms_per_step) *best_ms_per_step = best;
    return rc;
}

int rgpu_wavefront_time_steps(rgpu_wavefront* w, uint32_t steps, double* ms_per_step) {
    if (!w || !ms_per_step || steps == 0) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    hipEvent_t a, b;
    RGPU_HIP(ctx, hipEventCreate(&a));
    RGPU_HIP(ctx, hipEventCreate(&b));
    (void)hipEventRecord(a, ctx->stream);
    int rc = rgpu_wavefront_step(w, nullptr, steps);
    (void)hipEventRecord(b, ctx->stream);
    hipError_t e = hipEventSynchronize(b);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc != RGPU_OK) return rc;
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventElapsedTime");
    *ms_per_step = (double)ms / steps;
    return RGPU_OK;
}

int rgpu_wavefront_time_stages(rgpu_wavefront* w, uint32_t steps, double* ms_per_stage) {
    if (!w || !ms_per_stage || steps == 0 || steps > 4096) return RGPU_ERR_INVALID_ARG;
    if (!w->initialized) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "rgpu_wavefront_initialize was not called");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<hipEvent_t> ev((size_t)steps * kStageMarks, nullptr);
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < ev.size() && e == hipSuccess; ++i) e = hipEventCreate(&ev[i]);
    for (uint32_t k = 0; k < steps && e == hipSuccess; ++k) e = step_once(w, ctx->stream, ev.data() + (size_t)k * kStageMarks);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    std::vector<double> sum(kStageMarks - 1, 0.0);
    for (uint32_t k = 0; k < steps && e == hipSuccess; ++k)
        for (int j = 0; j + 1 < kStageMarks && e == hipSuccess; ++j) {
            float ms = 0.0f;
            e = hipEventElapsedTime(&ms, ev[(size_t)k * kStageMarks + j], ev[(size_t)k * kStageMarks + j + 1]);
            sum[j] += ms;
        }
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e != hipSuccess) return hip_fail(ctx, e, "rgpu_wavefront_time_stages");
    for (int j = 0; j + 1 < kStageMarks; ++j) ms_per_stage[j] = sum[j] / steps;
    return RGPU_OK;
}

int rgpu_wavefront_queue_counts(rgpu_wavefront* w, uint32_t* counts7, uint64_t* next_pixel) {
    if (!w || !counts7) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipMemcpyAsync(counts7, w->st.counts, kNumQueues * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (next_pixel)
        RGPU_HIP(ctx, hipMemcpyAsync(next_pixel, w->st.next_pixel, 8, hipMemcpyDeviceToHost, ctx->stream));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RGPU_OK;
}

int rgpu_wavefront_read_accumulation(rgpu_wavefront* w, double* rgb, uint32_t* counts) {
    if (!w || !w->initialized) return RGPU_ERR_INVALID_ARG;
    if (w->is_batch()) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "use the rgpu_wavefront_job_* readers of a batch");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t P = w->npixels;
    if (w->params.accum64) {
        std::vector<double> a(3 * P);
        std::vector<uint32_t> c(P);
        RGPU_HIP(ctx, hipMemcpy(a.data(), w->st.accum64, 3 * P * 8, hipMemcpyDeviceToHost));
        RGPU_HIP(ctx, hipMemcpy(c.data(), w->st.accum_count, P * 4, hipMemcpyDeviceToHost));
        if (rgb) std::memcpy(rgb, a.data(), 3 * P * 8);
        if (counts) std::memcpy(counts, c.data(), P * 4);
    } else {
        std::vector<float> a(4 * P);
        RGPU_HIP(ctx, hipMemcpy(a.data(), w->st.accum, 4 * P * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < P; ++i) {
            if (rgb)
                for (int k = 0; k < 3; ++k) rgb[3 * i + k] = a[4 * i + k];
            if (counts) std::memcpy(&counts[i], &a[4 * i + 3], 4);
        }
    }
    return RGPU_OK;
}

int rgpu_wavefront_tonemap(rgpu_wavefront* w, uint8_t* rgba8) {
    if (!w || !w->initialized || !rgba8) return RGPU_ERR_INVALID_ARG;
    if (w->is_batch()) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "use the rgpu_wavefront_job_* readers of a batch");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, wf_tonemap(w->st, w->npixels, w->params.accum64 != 0, w->tonemap_buf, ctx->stream));
    RGPU_HIP(ctx, hipMemcpyAsync(rgba8, w->tonemap_buf, (size_t)w->npixels * 4, hipMemcpyDeviceToHost, ctx->stream));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RGPU_OK;
}

int rgpu_wavefront_stat(rgpu_wavefront* w, uint64_t* total_samples) {
    if (!w || !w->initialized || !total_samples) return RGPU_ERR_INVALID_ARG;
    if (w->is_batch()) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "use the rgpu_wavefront_job_* readers of a batch");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipMemsetAsync(w->stat_total, 0, 8, ctx->stream));
    RGPU_HIP(ctx, wf_stat(w->st, w->npixels, w->params.accum64 != 0, w->stat_total, ctx->stream));
    unsigned long long t = 0;
    RGPU_HIP(ctx, hipMemcpyAsync(&t, w->stat_total, 8, hipMemcpyDeviceToHost, ctx->stream));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *total_samples = t;
    return RGPU_OK;
}

int rgpu_wavefront_read_state(rgpu_wavefront* w, uint32_t* rng, float* rays, float* paths, void* ext, float* shade,
                              int32_t* vol, uint32_t* items) {
    if (!w || !w->initialized) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t n = w->params.n;
    if (rng) RGPU_HIP(ctx, hipMemcpy(rng, w->st.rng, n * 16, hipMemcpyDeviceToHost));
    if (rays) RGPU_HIP(ctx, hipMemcpy(rays, w->st.ray, n * 32, hipMemcpyDeviceToHost));
    if (paths) RGPU_HIP(ctx, hipMemcpy(paths, w->st.path, n * 32, hipMemcpyDeviceToHost));
    if (ext) RGPU_HIP(ctx, hipMemcpy(ext, w->st.ext, n * 32, hipMemcpyDeviceToHost));
    if (shade) RGPU_HIP(ctx, hipMemcpy(shade, w->st.shade, n * 32, hipMemcpyDeviceToHost));
    if (vol) RGPU_HIP(ctx, hipMemcpy(vol, w->st.vol, n * 32, hipMemcpyDeviceToHost));
    if (items) RGPU_HIP(ctx, hipMemcpy(items, w->st.items, n * kNumQueues * 4, hipMemcpyDeviceToHost));
    return RGPU_OK;
}

// ------------------------------------------------------------------ intermediate images
int rgpu_wavefront_intermediate(rgpu_wavefront* w, rgpu_stream stream, rgpu_rgb16* dst) {
    if (!w || !dst) return RGPU_ERR_INVALID_ARG;
    if (!w->initialized) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "rgpu_wavefront_initialize was not called");
    if (w->is_batch()) return fail(w->ctx, RGPU_ERR_UNSUPPORTED, "a batch has one image per job");
    if ((uintptr_t)dst & 3) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "dst needs 4-byte alignment");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, wf_intermediate(w->st, w->npixels, w->params.accum64 != 0, reinterpret_cast<RGB16*>(dst),
                                  stream_of(ctx, stream)));
    return RGPU_OK;
}

int rgpu_merge_intermediate(rgpu_ctx* ctx, rgpu_stream stream, rgpu_rgb16* a, const rgpu_rgb16* b, uint32_t npixels) {
    if (!ctx) return RGPU_ERR_INVALID_ARG;
    if (npixels && (!a || !b)) return fail(ctx, RGPU_ERR_INVALID_ARG, "null intermediate");
    if (((uintptr_t)a | (uintptr_t)b) & 3) return fail(ctx, RGPU_ERR_INVALID_ARG, "intermediates need 4-byte alignment");
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, merge_intermediate(reinterpret_cast<RGB16*>(a), reinterpret_cast<const RGB16*>(b), npixels,
                                     stream_of(ctx, stream)));
    return RGPU_OK;
}

int rgpu_tonemap_intermediate(rgpu_ctx* ctx, rgpu_stream stream, const rgpu_rgb16* src, uint32_t npixels,
                              uint8_t* rgba8) {
    if (!ctx) return RGPU_ERR_INVALID_ARG;
    if (npixels && (!src || !rgba8)) return fail(ctx, RGPU_ERR_INVALID_ARG, "null buffer");
    if (((uintptr_t)src | (uintptr_t)rgba8) & 3) return fail(ctx, RGPU_ERR_INVALID_ARG, "buffers need 4-byte alignment");
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, tonemap_intermediate(reinterpret_cast<const RGB16*>(src), npixels, rgba8, stream_of(ctx, stream)));
    return RGPU_OK;
}

// ------------------------------------------------------------------ checkpoint / resume
namespace {
struct CheckpointHeader {
    char magic[8];  // "RGPUCKPT"
    uint32_t version;
    uint32_t path_count;
    uint32_t npixels;
    uint32_t accum64;
    Camera cam;
    uint64_t seed;     // the lane's configured seed (provenance; the RNG states are in the payload)
    uint64_t payload;  // bytes after the header
    // what the payload's indices and records refer to: a resume into anything else is refused
    uint32_t nprims;
    uint32_t pad;
    uint64_t scene_fp, materials_fp, envmap_fp;
    WFParams params;
    uint64_t payload_fp;  // fingerprint64 of the header (this field zero) and the payload
};

uint64_t checkpoint_fingerprint(CheckpointHeader h, const uint8_t* payload) {
    h.payload_fp = 0;
    return fingerprint64(payload, h.payload, fingerprint64(&h, sizeof(h)));
}
constexpr uint32_t kCheckpointVersion = 2;

struct Section {
    void* dev;
    size_t bytes;
};

// The device buffers a single render carries from one step to the next, in checkpoint order.
std::vector<Section> checkpoint_sections(const rgpu_wavefront* w, uint32_t npixels) {
    const size_t n = w->params.n, P = npixels;
    const DevState& s = w->st;
    std::vector<Section> v = {
        {s.ray, n * sizeof(TraceRay)},   {s.path, n * sizeof(PathRec)},  {s.ext, n * sizeof(ExtRec)},
        {s.shade, n * sizeof(ShadeRec)}, {s.vol, n * sizeof(VolList)},   {s.rng, n * sizeof(uint4)},
        {s.qsel, n},                     {s.items, n * kNumQueues * 4},  {s.counts, kNumQueues * 4},
        {s.block_counts, (size_t)s.nblocks * (kNumQueues + 1) * 4},      {s.next_pixel, 8},
    };
    if (w->params.accum64) {
        v.push_back({s.accum64, P * 3 * sizeof(double)});
        v.push_back({s.accum_count, P * 4});
    } else {
        v.push_back({s.accum, P * 4 * sizeof(float)});
    }
    return v;
}

uint64_t checkpoint_bytes(const rgpu_wavefront* w, uint32_t npixels) {
    uint64_t b = sizeof(CheckpointHeader);
    for (const Section& s : checkpoint_sections(w, npixels)) b += s.bytes;
    return b;
}

int checkpoint_usable(rgpu_wavefront* w) {
    if (w->is_batch()) return fail(w->ctx, RGPU_ERR_UNSUPPORTED, "checkpoints cover single renders, not batches");
    if (w->guide.enabled) return fail(w->ctx, RGPU_ERR_UNSUPPORTED, "checkpoints do not cover an attached guide");
    return RGPU_OK;
}

// Host-side range checks of the indices a step reads from a payload before it writes them
// itself: queue lengths and items (slots), pixel indices (accumulation), hit primitives and
// medium lists (material records). Returns an empty string when every index is in range.
std::string payload_index_error(const rgpu_wavefront* w, const uint8_t* p, uint32_t npixels, uint32_t nprims) {
    const size_t n = w->params.n;
    const uint8_t* ray = p;
    const uint8_t* ext = ray + n * sizeof(TraceRay) + n * sizeof(PathRec);
    const uint8_t* vol = ext + n * sizeof(ExtRec) + n * sizeof(ShadeRec);
    const uint8_t* items = vol + n * sizeof(VolList) + n * sizeof(uint4) + n;
    const uint8_t* counts = items + n * kNumQueues * 4;
    for (size_t i = 0; i < n; ++i) {
        TraceRay r;  // pad holds the slot's pixel
        std::memcpy(&r, ray + i * sizeof(TraceRay), sizeof(r));
        if (r.pad >= npixels && r.pad != kDeadPixel) return "a slot's pixel index is outside the image";
        ExtRec e;
        std::memcpy(&e, ext + i * sizeof(ExtRec), sizeof(e));
        if (e.hit_surface < -1 || e.hit_surface >= (int64_t)nprims || e.hit_volume < -1 ||
            e.hit_volume >= (int64_t)nprims)
            return "a slot's hit primitive is outside the scene";
        VolList v;
        std::memcpy(&v, vol + i * sizeof(VolList), sizeof(v));
        if (v.count < 0 || v.count > 7) return "a slot's medium list has a bad length";
        for (int k = 0; k < v.count; ++k)
            if (v.volumes[k] < 0 || v.volumes[k] >= (int64_t)nprims) return "a slot's medium list names no primitive";
    }
    for (int q = 0; q < kNumQueues; ++q) {
        uint32_t c;
        std::memcpy(&c, counts + 4 * q, 4);
        if (c > n) return "a queue is longer than the lane";
        for (uint32_t i = 0; i < c; ++i) {
            uint32_t slot;
            std::memcpy(&slot, items + ((size_t)q * n + i) * 4, 4);
            if (slot >= n) return "a queue item names no slot";
        }
    }
    return std::string();
}
}  // namespace

int rgpu_wavefront_checkpoint_size(rgpu_wavefront* w, uint64_t* bytes) {
    if (!w || !bytes) return RGPU_ERR_INVALID_ARG;
    if (!w->initialized) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "rgpu_wavefront_initialize was not called");
    int rc = checkpoint_usable(w);
    if (rc != RGPU_OK) return rc;
    *bytes = checkpoint_bytes(w, w->npixels);
    return RGPU_OK;
}

int rgpu_wavefront_checkpoint(rgpu_wavefront* w, void* host, uint64_t bytes) {
    if (!w || !host) return RGPU_ERR_INVALID_ARG;
    if (!w->initialized) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "rgpu_wavefront_initialize was not called");
    int rc = checkpoint_usable(w);
    if (rc != RGPU_OK) return rc;
    const uint64_t need = checkpoint_bytes(w, w->npixels);
    if (bytes < need) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "checkpoint buffer too small");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CheckpointHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, "RGPUCKPT", 8);
    h.version = kCheckpointVersion;
    h.path_count = w->params.n;
    h.npixels = w->npixels;
    h.accum64 = w->params.accum64 ? 1u : 0u;
    h.cam = w->cam;
    h.seed = w->seed;
    h.payload = need - sizeof(h);
    h.nprims = w->scene->nprims;
    h.scene_fp = w->scene->fingerprint;
    h.materials_fp = w->mats->fingerprint;
    h.envmap_fp = w->env->fingerprint;
    h.params = w->params;
    uint8_t* payload = static_cast<uint8_t*>(host) + sizeof(h);
    uint8_t* p = payload;
    for (const Section& s : checkpoint_sections(w, w->npixels)) {
        RGPU_HIP(ctx, hipMemcpyAsync(p, s.dev, s.bytes, hipMemcpyDeviceToHost, ctx->stream));
        p += s.bytes;
    }
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    h.payload_fp = checkpoint_fingerprint(h, payload);
    std::memcpy(host, &h, sizeof(h));
    return RGPU_OK;
}

int rgpu_wavefront_resume(rgpu_wavefront* w, const void* host, uint64_t bytes) {
    if (!w || !host) return RGPU_ERR_INVALID_ARG;
    int rc = checkpoint_usable(w);
    if (rc != RGPU_OK) return rc;
    CheckpointHeader h;
    if (bytes < sizeof(h)) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "not a checkpoint (too short)");
    std::memcpy(&h, host, sizeof(h));
    if (std::memcmp(h.magic, "RGPUCKPT", 8) != 0 || h.version != kCheckpointVersion)
        return fail(w->ctx, RGPU_ERR_INVALID_ARG, "not a checkpoint of this library version");
    if (h.path_count != w->params.n || h.accum64 != (w->params.accum64 ? 1u : 0u))
        return fail(w->ctx, RGPU_ERR_INVALID_ARG, "checkpoint of a lane with another path count or accumulation");
    if (h.npixels == 0 || (uint64_t)h.cam.width * h.cam.height != h.npixels)
        return fail(w->ctx, RGPU_ERR_INVALID_ARG, "checkpoint camera does not match its image size");
    if (bytes < sizeof(h) + h.payload || sizeof(h) + h.payload != checkpoint_bytes(w, h.npixels))
        return fail(w->ctx, RGPU_ERR_INVALID_ARG, "checkpoint truncated or inconsistent");
    if (h.nprims != w->scene->nprims || h.scene_fp != w->scene->fingerprint)
        return fail(w->ctx, RGPU_ERR_INVALID_ARG, "checkpoint of another scene");
    if (h.materials_fp != w->mats->fingerprint)
        return fail(w->ctx, RGPU_ERR_INVALID_ARG, "checkpoint of other materials");
    if (h.envmap_fp != w->env->fingerprint)
        return fail(w->ctx, RGPU_ERR_INVALID_ARG, "checkpoint of another envmap");
    if (std::memcmp(&h.params, &w->params, sizeof(WFParams)) != 0)
        return fail(w->ctx, RGPU_ERR_INVALID_ARG,
                    "checkpoint of a lane with other settings (depth, clamp, roulette, flags)");
    const uint8_t* payload = static_cast<const uint8_t*>(host) + sizeof(h);
    if (checkpoint_fingerprint(h, payload) != h.payload_fp)
        return fail(w->ctx, RGPU_ERR_INVALID_ARG, "checkpoint is corrupted (fingerprint mismatch)");
    const std::string bad = payload_index_error(w, payload, h.npixels, w->scene->nprims);
    if (!bad.empty()) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "checkpoint rejected: " + bad);
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(w);  // a captured graph holds the old camera by value
    w->initialized = false;  // until every buffer holds the checkpoint
    if ((rc = ensure_accum(w, h.npixels)) != RGPU_OK) return rc;
    const uint8_t* p = payload;
    for (const Section& s : checkpoint_sections(w, h.npixels)) {
        RGPU_HIP(ctx, hipMemcpyAsync(s.dev, p, s.bytes, hipMemcpyHostToDevice, ctx->stream));
        p += s.bytes;
    }
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    w->cam = h.cam;
    if ((rc = ensure_publish(w)) != RGPU_OK) return rc;
    w->initialized = true;
    return RGPU_OK;
}

// ------------------------------------------------------------------ path guiding
static int guide_alloc(rgpu_guide* g, size_t bytes, void** out) {
    RGPU_HIP(g->ctx, hipMalloc(out, bytes < 16 ? 16 : bytes));
    g->owned.push_back(*out);
    return RGPU_OK;
}

int rgpu_guide_destroy(rgpu_guide* g) {
    if (!g) return RGPU_ERR_INVALID_ARG;
    if (!g->users.empty())
        return fail(g->ctx, RGPU_ERR_INVALID_ARG,
                    "rgpu_guide_destroy: the guide is attached to " + std::to_string(g->users.size()) +
                        " wavefront(s); detach it (rgpu_wavefront_attach_guide with NULL) or destroy them first");
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    for (void* p : g->owned) (void)hipFree(p);
    delete g;
    return RGPU_OK;
}

int rgpu_guide_create(rgpu_ctx* ctx, const rgpu_guide_config* cfg, rgpu_guide** out) {
    if (!ctx || !cfg || !out) return RGPU_ERR_INVALID_ARG;
    *out = nullptr;
    if (cfg->abi_version != RGPU_ABI_VERSION) return fail(ctx, RGPU_ERR_INVALID_ARG, "abi_version mismatch");
    if (!(cfg->select_p > 0.0f && cfg->select_p < 1.0f))
        return fail(ctx, RGPU_ERR_INVALID_ARG, "guide selection probability must be in (0, 1)");
    rgpu_guide* g = new (std::nothrow) rgpu_guide();
    if (!g) return RGPU_ERR_OUT_OF_MEMORY;
    g->ctx = ctx;
    g->lr = cfg->learning_rate > 0.0f ? cfg->learning_rate : 1e-3f;
    g->minibatch = cfg->minibatch ? cfg->minibatch : 16384u;
    g->select_p = cfg->select_p;
    DevGuide& d = g->dev;
    d.capacity = cfg->capacity ? cfg->capacity : (1u << 18);
    d.select_p = cfg->select_p;
    d.max_depth = cfg->max_depth;
    if (!(cfg->weight_cap >= 0.0f) || !(cfg->log_kappa_max == 0.0f ||
                                        (cfg->log_kappa_max > kGLogKappaMin && cfg->log_kappa_max <= 12.0f))) {
        delete g;
        return fail(ctx, RGPU_ERR_INVALID_ARG, "weight_cap must be >= 0 and log_kappa_max in (-3, 12]");
    }
    d.weight_cap = cfg->weight_cap;
    d.log_kappa_max = cfg->log_kappa_max != 0.0f ? cfg->log_kappa_max : kGLogKappaMax;
    if (cfg->learn_select > 1u) {
        delete g;
        return fail(ctx, RGPU_ERR_INVALID_ARG, "learn_select must be 0 or 1");
    }
    d.learn_select = cfg->learn_select;
    for (int k = 0; k < 3; ++k) {
        float ext = cfg->scene_hi[k] - cfg->scene_lo[k];
        d.lo[k] = cfg->scene_lo[k];
        d.inv_ext[k] = ext > 0.0f ? 1.0f / ext : 1.0f;
    }
    int rc = hipSetDevice(ctx->device) == hipSuccess ? RGPU_OK : RGPU_ERR_HIP;
    void* p = nullptr;
    auto A = [&](size_t bytes) -> void* {
        if (rc != RGPU_OK) return nullptr;
        rc = guide_alloc(g, bytes, &p);
        return rc == RGPU_OK ? p : nullptr;
    };
    d.params = (float*)A(kGParams * 4);
    d.w_bf = (uint16_t*)A(kGParams * 2);
    d.wt_bf = (uint16_t*)A(kGParams * 2);
    d.grad = (float*)A(kGParams * 4);
    d.partials = (float*)A((size_t)kGTrainGroups * kGParams * 4);
    g->minibatch = std::min(g->minibatch, d.capacity);
    d.act_rows = (g->minibatch + 31u) / 32u * 32u;
    d.act = (uint16_t*)A((size_t)d.act_rows * kGActBytesPerRow);
    d.adam_m = (float*)A(kGParams * 4);
    d.adam_v = (float*)A(kGParams * 4);
    d.samples = (GuideSample*)A((size_t)d.capacity * sizeof(GuideSample));
    d.sample_count = (uint32_t*)A(4);
    if (rc != RGPU_OK) {
        rgpu_guide_destroy(g);
        return rc;
    }
    // He-uniform hidden layers; a small output layer so the initial mixture is broad and even.
    std::vector<float> w(kGParams, 0.0f);
    uint64_t x = cfg->seed ^ 0x5DEECE66Dull;
    auto uni = [&]() {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return (double)(z >> 11) / 9007199254740992.0 * 2.0 - 1.0;
    };
    const int offs[4] = {kGoW0, kGoW1, kGoW2, kGoW3}, nin[4] = {kGIn, kGHid, kGHid, kGHid},
              nout[4] = {kGHid, kGHid, kGHid, kGOut};
    for (int l = 0; l < 4; ++l) {
        double a = std::sqrt(6.0 / (l == 0 ? 78 : nin[l])) * (l == 3 ? 0.1 : 1.0);
        for (int o = 0; o < nout[l]; ++o)
            for (int i = 0; i < nin[l]; ++i) w[offs[l] + o * nin[l] + i] = (l == 0 && i >= 78) ? 0.0f : (float)(a * uni());
    }
    std::vector<float> zeros(kGParams, 0.0f);
    rc = RGPU_OK;
    if (hipMemcpy(d.params, w.data(), kGParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.adam_m, zeros.data(), kGParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.adam_v, zeros.data(), kGParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(d.sample_count, 0, 4) != hipSuccess || guide_refresh(d, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = RGPU_ERR_HIP;
    if (rc != RGPU_OK) {
        rgpu_guide_destroy(g);
        return fail(ctx, rc, "guide initialisation failed");
    }
    *out = g;
    return RGPU_OK;
}

int rgpu_guide_params(rgpu_guide* g, float* get, const float* set) {
    if (!g) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = g->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (set) {
        RGPU_HIP(ctx, hipMemcpy(g->dev.params, set, kGParams * 4, hipMemcpyHostToDevice));
        RGPU_HIP(ctx, guide_refresh(g->dev, ctx->stream));
        RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (get) RGPU_HIP(ctx, hipMemcpy(get, g->dev.params, kGParams * 4, hipMemcpyDeviceToHost));
    return RGPU_OK;
}

int rgpu_guide_param_count(void) { return kGParams; }

// Test hooks: `samples` are rows of 16 floats (p01, n, wo, wi, weight, pad).
int rgpu_guide_forward(rgpu_guide* g, const float* samples, uint32_t rows, float* out) {
    if (!g || (rows && (!samples || !out))) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = g->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    void *din = nullptr, *dout = nullptr;
    RGPU_HIP(ctx, hipMalloc(&din, (size_t)rows * 64 + 16));
    hipError_t e = hipMalloc(&dout, (size_t)rows * kGOut * 4 + 16);
    if (e == hipSuccess) e = hipMemcpy(din, samples, (size_t)rows * 64, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = guide_forward(g->dev, (const GuideSample*)din, rows, (float*)dout, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, dout, (size_t)rows * kGOut * 4, hipMemcpyDeviceToHost);
    (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return hip_fail(ctx, e, "rgpu_guide_forward");
    return RGPU_OK;
}

// Gradient of the mean guiding loss over `rows` samples (they replace the training buffer).
int rgpu_guide_gradients(rgpu_guide* g, const float* samples, uint32_t rows, float* grad) {
    if (!g || !samples || !grad || rows == 0 || rows > g->dev.act_rows) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = g->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipMemcpy(g->dev.samples, samples, (size_t)rows * 64, hipMemcpyHostToDevice));
    RGPU_HIP(ctx, hipMemset(g->dev.grad, 0, kGParams * 4));
    RGPU_HIP(ctx, guide_gradients(g->dev, rows, ctx->stream));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    RGPU_HIP(ctx, hipMemcpy(grad, g->dev.grad, kGParams * 4, hipMemcpyDeviceToHost));
    RGPU_HIP(ctx, hipMemset(g->dev.grad, 0, kGParams * 4));
    return RGPU_OK;
}

// Train on host-provided samples: minibatches of the configured size, one Adam step each.
int rgpu_guide_train_samples(rgpu_guide* g, const float* samples, uint32_t rows) {
    if (!g || !samples || rows == 0 || rows > g->dev.capacity) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = g->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipMemcpy(g->dev.samples, samples, (size_t)rows * 64, hipMemcpyHostToDevice));
    for (uint32_t first = 0; first < rows; first += g->minibatch) {
        uint32_t n = std::min(g->minibatch, rows - first);
        RGPU_HIP(ctx, guide_train_batch(g->dev, first, n, g->lr, ++g->step, ctx->stream));
    }
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RGPU_OK;
}

int rgpu_wavefront_attach_guide(rgpu_wavefront* w, rgpu_guide* g, int sampling, int training) {
    if (!w) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(w);
    if (g && g->ctx->device != ctx->device)
        return fail(ctx, RGPU_ERR_INVALID_ARG, "the guide lives on another device than the wavefront");
    // Detach first: the lane renders plain and keeps only its per-slot guiding buffers (it owns
    // them) for a later attach, also when this attach fails below.
    detach_from_guide(w);
    DevGuide keep;
    keep.lobes = w->guide.lobes;
    keep.sel = w->guide.sel;
    keep.rec_geo = w->guide.rec_geo;
    keep.rec_T = w->guide.rec_T;
    keep.rec_L = w->guide.rec_L;
    keep.emit_w = w->guide.emit_w;
    keep.rec_count = w->guide.rec_count;
    keep.nslots = w->guide.nslots;
    w->guide = keep;
    if (!g) return RGPU_OK;
    DevGuide d = g->dev;
    const size_t n = w->params.n;
    if (!w->guide.lobes) {
        void* p;
        int rc;
        if ((rc = wf_alloc(w, n * kGSlotFloats * 4, &p)) != RGPU_OK) return rc;
        d.lobes = (float*)p;
        if ((rc = wf_alloc(w, n * 4, &p)) != RGPU_OK) return rc;
        d.sel = (float*)p;
        if ((rc = wf_alloc(w, n * sizeof(GuideVertex) * kGRecords, &p)) != RGPU_OK) return rc;
        d.rec_geo = (GuideVertex*)p;
        if ((rc = wf_alloc(w, n * sizeof(float4) * kGRecords, &p)) != RGPU_OK) return rc;
        d.rec_T = (float4*)p;
        if ((rc = wf_alloc(w, n * sizeof(float4) * kGRecords, &p)) != RGPU_OK) return rc;
        d.rec_L = (float4*)p;
        if ((rc = wf_alloc(w, n * sizeof(float4), &p)) != RGPU_OK) return rc;
        d.emit_w = (float4*)p;
        if ((rc = wf_alloc(w, n * 4, &p)) != RGPU_OK) return rc;
        d.rec_count = (uint32_t*)p;
    } else {
        d.lobes = w->guide.lobes;
        d.sel = w->guide.sel;
        d.rec_geo = w->guide.rec_geo;
        d.rec_T = w->guide.rec_T;
        d.rec_L = w->guide.rec_L;
        d.emit_w = w->guide.emit_w;
        d.rec_count = w->guide.rec_count;
    }
    // paths in flight start recording afresh: vertices recorded for another guide, or before a
    // detach, do not describe the rest of the path
    RGPU_HIP(ctx, hipMemset(d.rec_count, 0, n * 4));
    d.nslots = (uint32_t)n;
    // inference: one 16-wave workgroup per CU (the staged weights and the waves' activation
    // buffers fill the CU's LDS), each wave looping over 16-row blocks
    d.infer_blocks = (uint32_t)(ctx->cu_count > 0 ? ctx->cu_count : 256);
    d.enabled = sampling ? 1u : 0u;
    d.train = training ? 1u : 0u;
    w->guide = d;
    w->guide_obj = g;
    g->users.push_back(w);
    return RGPU_OK;
}

// Test hook: lobe records [first, first + count) as the last inference wrote them, in vertex
// order (the lambertian queue, then the Ward queue); 48 floats each: 8 lobes of mean xyz,
// kappa, weight, log(weight C(kappa)).
int rgpu_wavefront_guide_lobes(rgpu_wavefront* w, uint32_t first, uint32_t count, float* out) {
    if (!w || (count && !out) || !w->guide.lobes) return RGPU_ERR_INVALID_ARG;
    if ((uint64_t)first + count > w->params.n) return fail(w->ctx, RGPU_ERR_INVALID_ARG, "record range outside the lane");
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (count) {
        // the device holds twelve planes of float4 (plane k: floats 4k..4k+3 of every record)
        const size_t planes = kGSlotFloats / 4, n = w->guide.nslots;
        std::vector<float> stage(planes * 4 * (size_t)count);
        for (size_t k = 0; k < planes; ++k)
            RGPU_HIP(ctx, hipMemcpy(stage.data() + k * 4 * count, w->guide.lobes + 4 * (k * n + first), 16 * (size_t)count,
                                    hipMemcpyDeviceToHost));
        for (size_t r = 0; r < count; ++r)
            for (size_t k = 0; k < planes; ++k)
                for (int q = 0; q < 4; ++q) out[r * kGSlotFloats + 4 * k + q] = stage[(k * count + r) * 4 + q];
    }
    return RGPU_OK;
}

// Number of training samples the attached guide has collected (may exceed its capacity;
// samples beyond it were dropped).
int rgpu_wavefront_guide_pending(rgpu_wavefront* w, uint32_t* pending) {
    if (!w || !pending || !w->guide_obj) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = w->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipMemcpyAsync(pending, w->guide.sample_count, 4, hipMemcpyDeviceToHost, ctx->stream));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RGPU_OK;
}

// Copies the first `rows` buffered training samples to the host; *buffered = samples buffered
// (capped at the capacity).
int rgpu_guide_read_samples(rgpu_guide* g, float* out, uint32_t rows, uint32_t* buffered) {
    if (!g || (rows && !out)) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = g->ctx;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t n = 0;
    RGPU_HIP(ctx, hipMemcpy(&n, g->dev.sample_count, 4, hipMemcpyDeviceToHost));
    n = std::min(n, g->dev.capacity);
    if (buffered) *buffered = n;
    if (rows > n) return fail(ctx, RGPU_ERR_INVALID_ARG, "fewer samples buffered than asked for");
    if (rows)
        RGPU_HIP(ctx, hipMemcpy(out, g->dev.samples, (size_t)rows * sizeof(GuideSample), hipMemcpyDeviceToHost));
    return RGPU_OK;
}

// One training pass if at least `min_samples` samples are buffered: minibatch Adam steps over
// the buffer, then the buffer is emptied. *trained = samples used (0 if none).
int rgpu_wavefront_guide_train(rgpu_wavefront* w, uint32_t min_samples, uint32_t* trained) {
    if (!w || !w->guide_obj) return RGPU_ERR_INVALID_ARG;
    rgpu_guide* g = w->guide_obj;
    rgpu_ctx* ctx = w->ctx;
    uint32_t pending = 0;
    int rc = rgpu_wavefront_guide_pending(w, &pending);
    if (rc != RGPU_OK) return rc;
    uint32_t rows = std::min(pending, g->dev.capacity);
    if (trained) *trained = 0;
    if (rows == 0 || rows < min_samples) return RGPU_OK;
    for (uint32_t first = 0; first < rows; first += g->minibatch) {
        uint32_t n = std::min(g->minibatch, rows - first);
        RGPU_HIP(ctx, guide_train_batch(g->dev, first, n, g->lr, ++g->step, ctx->stream));
    }
    RGPU_HIP(ctx, hipMemsetAsync(g->dev.sample_count, 0, 4, ctx->stream));
    if (trained) *trained = rows;
    return RGPU_OK;
}

}  // extern "C"
