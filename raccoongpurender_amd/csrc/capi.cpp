// C ABI (include/raccoon_gpu.h): context, memory, host BVH build, device scene and
// extension ray casting on one gfx950 GPU.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/raccoon_gpu.h"
#include "bvh_build.h"
#include "bvh_gpu.h"
#include "fingerprint.h"
#include "rgpu_traverse.h"
#include "rgpu_ctx.h"
#include "rgpu_scene.h"

using namespace rgpu;

static_assert(RGPU_VARIANT_STACKLESS == kVariantStackless && RGPU_VARIANT_THREADED == kVariantThreaded &&
                  RGPU_VARIANT_THREADED_SINGLE == kVariantSingleThreaded,
              "public traversal variant ids must equal the kernels'");

struct rgpu_bvh {
    BVHTree tree;
};

namespace rgpu {

int fail(rgpu_ctx* ctx, int code, const std::string& msg) {
    if (ctx) {
        std::lock_guard<std::mutex> g(ctx->err_mu);
        ctx->last_error = msg;
    }
    return code;
}

int hip_fail(rgpu_ctx* ctx, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    if (e == hipErrorOutOfMemory) return fail(ctx, RGPU_ERR_OUT_OF_MEMORY, m);
    if (e == hipErrorNoBinaryForGpu || e == hipErrorInvalidImage || e == hipErrorInvalidDeviceFunction)
        return fail(ctx, RGPU_ERR_NO_CODE_OBJECT, m);
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return fail(ctx, RGPU_ERR_NO_DEVICE, m);
    return fail(ctx, RGPU_ERR_HIP, m);
}

hipStream_t stream_of(rgpu_ctx* ctx, void* s) { return s ? (hipStream_t)s : ctx->stream; }

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);

}  // namespace rgpu

#define RGPU_HIP(ctx, call)                                  \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

#define RGPU_GUARD_BEGIN try {
#define RGPU_GUARD_END(ctx)                                                         \
    }                                                                               \
    catch (const std::bad_alloc&) {                                                 \
        return fail(ctx, RGPU_ERR_OUT_OF_MEMORY, "host allocation failed");         \
    }                                                                               \
    catch (const std::invalid_argument& ex) {                                       \
        return fail(ctx, RGPU_ERR_INVALID_ARG, ex.what());                          \
    }                                                                               \
    catch (const std::exception& ex) {                                              \
        return fail(ctx, RGPU_ERR_HIP, ex.what());                                  \
    }

extern "C" {

int rgpu_device_count(int* out) {
    if (!out) return RGPU_ERR_INVALID_ARG;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *out = 0;
        return RGPU_ERR_NO_DEVICE;
    }
    *out = n;
    return RGPU_OK;
}

int rgpu_create(int device_index, const rgpu_config* cfg, rgpu_ctx** out) {
    if (!out) return RGPU_ERR_INVALID_ARG;
    *out = nullptr;
    if (cfg && cfg->abi_version != RGPU_ABI_VERSION) return RGPU_ERR_INVALID_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return RGPU_ERR_NO_DEVICE;
    if (device_index < 0 || device_index >= n) return RGPU_ERR_INVALID_ARG;
    rgpu_ctx* ctx = new (std::nothrow) rgpu_ctx();
    if (!ctx) return RGPU_ERR_OUT_OF_MEMORY;
    ctx->device = device_index;
    if (hipSetDevice(device_index) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return RGPU_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_index) == hipSuccess) {
        ctx->cu_count = prop.multiProcessorCount;
        ctx->arch = prop.gcnArchName;
    }
    *out = ctx;
    return RGPU_OK;
}

int rgpu_destroy(rgpu_ctx* ctx) {
    if (!ctx) return RGPU_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return RGPU_OK;
}

const char* rgpu_last_error_string(const rgpu_ctx* ctx) {
    if (!ctx) return "null context";
    return ctx->last_error.c_str();
}

int rgpu_alloc(rgpu_ctx* ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return RGPU_ERR_INVALID_ARG;
    *dptr = nullptr;
    if (bytes == 0) return RGPU_OK;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipMalloc(dptr, bytes));
    return RGPU_OK;
}

int rgpu_free(rgpu_ctx* ctx, void* dptr) {
    if (!ctx) return RGPU_ERR_INVALID_ARG;
    if (!dptr) return RGPU_OK;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipFree(dptr));
    return RGPU_OK;
}

int rgpu_upload(rgpu_ctx* ctx, rgpu_stream stream, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return RGPU_ERR_INVALID_ARG;
    if (!bytes) return RGPU_OK;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = stream_of(ctx, stream);
    RGPU_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    RGPU_HIP(ctx, hipStreamSynchronize(s));
    return RGPU_OK;
}

int rgpu_download(rgpu_ctx* ctx, rgpu_stream stream, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return RGPU_ERR_INVALID_ARG;
    if (!bytes) return RGPU_OK;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = stream_of(ctx, stream);
    RGPU_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    RGPU_HIP(ctx, hipStreamSynchronize(s));
    return RGPU_OK;
}

int rgpu_fill_u32(rgpu_ctx* ctx, rgpu_stream stream, void* dst, uint32_t value, size_t count) {
    if (!ctx || (count && !dst)) return RGPU_ERR_INVALID_ARG;
    if (!count) return RGPU_OK;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipMemsetD32Async((hipDeviceptr_t)dst, (int)value, count, stream_of(ctx, stream)));
    return RGPU_OK;
}

int rgpu_copy(rgpu_ctx* ctx, rgpu_stream stream, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return RGPU_ERR_INVALID_ARG;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, launch_copy(dst, src, bytes, stream_of(ctx, stream)));
    return RGPU_OK;
}

int rgpu_stream_create(rgpu_ctx* ctx, rgpu_stream* out) {
    if (!ctx || !out) return RGPU_ERR_INVALID_ARG;
    *out = nullptr;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = nullptr;
    RGPU_HIP(ctx, hipStreamCreate(&s));  // default priority
    *out = (rgpu_stream)s;
    return RGPU_OK;
}

int rgpu_stream_destroy(rgpu_ctx* ctx, rgpu_stream stream) {
    if (!ctx || !stream) return RGPU_ERR_INVALID_ARG;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize((hipStream_t)stream));
    RGPU_HIP(ctx, hipStreamDestroy((hipStream_t)stream));
    return RGPU_OK;
}

int rgpu_synchronize(rgpu_ctx* ctx, rgpu_stream stream) {
    if (!ctx) return RGPU_ERR_INVALID_ARG;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    RGPU_HIP(ctx, hipStreamSynchronize(stream_of(ctx, stream)));
    return RGPU_OK;
}

// ------------------------------------------------------------------ BVH
int rgpu_bvh_build(const float* points, uint32_t stride, uint32_t npoints, const uint32_t* indices, uint32_t ntris,
                   const float* spheres, uint32_t nspheres, const rgpu_bvh_options* opt, rgpu_bvh** out) {
    if (!out) return RGPU_ERR_INVALID_ARG;
    *out = nullptr;
    if (ntris && (!points || !indices || stride < 3)) return RGPU_ERR_INVALID_ARG;
    if (nspheres && !spheres) return RGPU_ERR_INVALID_ARG;
    if ((uint64_t)ntris + nspheres >= (1ull << 28)) return RGPU_ERR_INVALID_ARG;  // leaf slot field width
    BuildOptions bo;
    if (opt) {
        if (opt->abi_version != RGPU_ABI_VERSION) return RGPU_ERR_INVALID_ARG;
        if (opt->max_leaf) bo.max_leaf = opt->max_leaf;
        if (opt->bins) bo.bins = opt->bins;
        if (opt->max_depth) bo.max_depth = opt->max_depth;
        if (!(opt->traversal_cost >= 0.0f && opt->traversal_cost <= 64.0f)) return RGPU_ERR_INVALID_ARG;
        if (opt->traversal_cost > 0.0f) bo.traversal_cost = opt->traversal_cost;
    }
    try {
        rgpu_bvh* b = new rgpu_bvh();
        try {
            build_bvh(points, stride, npoints, indices, ntris, spheres, nspheres, bo, &b->tree);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    } catch (const std::bad_alloc&) {
        return RGPU_ERR_OUT_OF_MEMORY;
    } catch (const std::exception&) {
        return RGPU_ERR_INVALID_ARG;
    }
    return RGPU_OK;
}

int rgpu_bvh_build_device(rgpu_ctx* ctx, const float* points, uint32_t stride, uint32_t npoints,
                          const uint32_t* indices, uint32_t ntris, const float* spheres, uint32_t nspheres,
                          const rgpu_bvh_options* opt, rgpu_bvh** out, double* device_ms) {
    if (!ctx || !out) return RGPU_ERR_INVALID_ARG;
    *out = nullptr;
    if (ntris && (!points || !indices || stride < 3)) return fail(ctx, RGPU_ERR_INVALID_ARG, "bad triangle arrays");
    if (nspheres && !spheres) return fail(ctx, RGPU_ERR_INVALID_ARG, "bad sphere array");
    if ((uint64_t)ntris + nspheres >= (1ull << 28)) return fail(ctx, RGPU_ERR_INVALID_ARG, "too many primitives");
    BuildOptions bo;
    if (opt) {
        if (opt->abi_version != RGPU_ABI_VERSION) return fail(ctx, RGPU_ERR_INVALID_ARG, "ABI version mismatch");
        if (opt->max_leaf) bo.max_leaf = opt->max_leaf;
        if (opt->bins) bo.bins = opt->bins;
        if (opt->max_depth) bo.max_depth = opt->max_depth;
        if (!(opt->traversal_cost >= 0.0f && opt->traversal_cost <= 64.0f))
            return fail(ctx, RGPU_ERR_INVALID_ARG, "traversal_cost must be in [0, 64]");
        if (opt->traversal_cost > 0.0f) bo.traversal_cost = opt->traversal_cost;
    }
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    rgpu_bvh* b = new (std::nothrow) rgpu_bvh();
    if (!b) return RGPU_ERR_OUT_OF_MEMORY;
    BuildTimes t;
    int rc = RGPU_OK;
    try {
        build_bvh_device(points, stride, npoints, indices, ntris, spheres, nspheres, bo, ctx->stream, &b->tree, &t);
    } catch (const std::bad_alloc&) {
        rc = fail(ctx, RGPU_ERR_OUT_OF_MEMORY, "device BVH build: out of memory");
    } catch (const DeviceBuildError& e) {
        rc = fail(ctx, RGPU_ERR_HIP, e.what());
    } catch (const std::exception& e) {
        rc = fail(ctx, RGPU_ERR_INVALID_ARG, e.what());
    }
    if (rc != RGPU_OK) {
        delete b;
        return rc;
    }
    if (device_ms) *device_ms = t.device_ms;
    *out = b;
    return RGPU_OK;
}

int rgpu_bvh_destroy(rgpu_bvh* bvh) {
    if (!bvh) return RGPU_ERR_INVALID_ARG;
    delete bvh;
    return RGPU_OK;
}

int rgpu_bvh_get_stats(const rgpu_bvh* bvh, rgpu_bvh_stats* out) {
    if (!bvh || !out) return RGPU_ERR_INVALID_ARG;
    out->nodes = (uint32_t)bvh->tree.nodes.size();
    out->leaves = bvh->tree.leaf_count;
    out->depth = bvh->tree.depth;
    out->slots = (uint32_t)bvh->tree.prim_ids.size();
    out->sah_cost = bvh->tree.sah_cost;
    return RGPU_OK;
}

int rgpu_bvh_export(const rgpu_bvh* bvh, float* bounds, int32_t* left, int32_t* right, uint32_t* prim_beg,
                    uint32_t* prim_cnt, uint32_t* prim_ids) {
    if (!bvh) return RGPU_ERR_INVALID_ARG;
    const auto& t = bvh->tree;
    for (size_t i = 0; i < t.nodes.size(); ++i) {
        const BuildNode& n = t.nodes[i];
        if (bounds)
            for (int k = 0; k < 3; ++k) {
                bounds[6 * i + k] = n.bmin[k];
                bounds[6 * i + 3 + k] = n.bmax[k];
            }
        if (left) left[i] = n.left;
        if (right) right[i] = n.right;
        if (prim_beg) prim_beg[i] = n.prim_beg;
        if (prim_cnt) prim_cnt[i] = n.prim_cnt;
    }
    if (prim_ids && !t.prim_ids.empty()) std::memcpy(prim_ids, t.prim_ids.data(), t.prim_ids.size() * 4);
    return RGPU_OK;
}

// ------------------------------------------------------------------ scene
static int upload_new(rgpu_ctx* ctx, const void* src, size_t bytes, void** dst) {
    *dst = nullptr;
    if (bytes == 0) return RGPU_OK;
    RGPU_HIP(ctx, hipMalloc(dst, bytes));
    hipError_t e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(*dst);
        *dst = nullptr;
        return hip_fail(ctx, e, "hipMemcpy(scene upload)");
    }
    return RGPU_OK;
}

int rgpu_scene_destroy(rgpu_scene* s) {
    if (!s) return RGPU_ERR_INVALID_ARG;
    if (s->ctx) (void)hipSetDevice(s->ctx->device);
    for (void* p : {s->sl, s->mt, s->prims, s->points, s->indices, s->spheres})
        if (p) (void)hipFree(p);
    delete s;
    return RGPU_OK;
}

int rgpu_scene_create(rgpu_ctx* ctx, const rgpu_bvh* bvh, const float* points, uint32_t stride, uint32_t npoints,
                      const uint32_t* indices, uint32_t ntris, const float* spheres, uint32_t nspheres,
                      rgpu_scene** out) {
    if (!ctx || !bvh || !out) return RGPU_ERR_INVALID_ARG;
    *out = nullptr;
    if (ntris && (!points || !indices || stride < 3)) return fail(ctx, RGPU_ERR_INVALID_ARG, "bad triangle arrays");
    if (nspheres && !spheres) return fail(ctx, RGPU_ERR_INVALID_ARG, "bad sphere array");
    if (bvh->tree.prim_ids.size() != (size_t)ntris + nspheres)
        return fail(ctx, RGPU_ERR_INVALID_ARG, "BVH was built for a different primitive count");
    if (bvh->tree.nodes.size() >= (size_t)kChildLeaf)
        return fail(ctx, RGPU_ERR_INVALID_ARG, "BVH has 2^30 nodes or more");
    RGPU_GUARD_BEGIN
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<SLNode> sl;
    std::vector<MTNode> mt;
    std::vector<PrimRec> pr;
    pack_stackless(bvh->tree, &sl);
    number_top_breadth_first(&sl, kTreeletMaxNodes);
    mark_leaf_links(&sl);
    pack_threaded(bvh->tree, &mt);
    pack_prims(bvh->tree, points, stride, indices, ntris, spheres, nspheres, &pr);
    std::vector<float> p4((size_t)npoints * 4, 0.0f);
    for (uint32_t i = 0; i < npoints; ++i)
        for (int k = 0; k < 3; ++k) p4[4 * (size_t)i + k] = points[(size_t)i * stride + k];
    rgpu_scene* s = new rgpu_scene();
    s->ctx = ctx;
    int rc = RGPU_OK;
    if (rc == RGPU_OK) rc = upload_new(ctx, sl.data(), sl.size() * sizeof(SLNode), &s->sl);
    if (rc == RGPU_OK) rc = upload_new(ctx, mt.data(), mt.size() * sizeof(MTNode), &s->mt);
    if (rc == RGPU_OK) rc = upload_new(ctx, pr.data(), pr.size() * sizeof(PrimRec), &s->prims);
    if (rc == RGPU_OK) rc = upload_new(ctx, p4.data(), p4.size() * sizeof(float), &s->points);
    if (rc == RGPU_OK) rc = upload_new(ctx, indices, (size_t)ntris * 3 * sizeof(uint32_t), &s->indices);
    if (rc == RGPU_OK) rc = upload_new(ctx, spheres, (size_t)nspheres * 4 * sizeof(float), &s->spheres);
    if (rc != RGPU_OK) {
        rgpu_scene_destroy(s);
        return rc;
    }
    s->dev.sl_nodes = (const SLNode*)s->sl;
    s->dev.mt_nodes = (const MTNode*)s->mt;
    s->dev.prims = (const PrimRec*)s->prims;
    s->dev.points = (const float4*)s->points;
    s->dev.indices = (const uint32_t*)s->indices;
    s->dev.spheres = (const float4*)s->spheres;
    s->dev.nnodes = (uint32_t)bvh->tree.nodes.size();
    s->dev.sphere_begin = (int32_t)ntris;
    bool finite = true;
    for (const BuildNode& b : bvh->tree.nodes)
        for (int k = 0; k < 3; ++k) finite = finite && std::isfinite(b.bmin[k]) && std::isfinite(b.bmax[k]);
    s->dev.boxes_finite = finite ? 1u : 0u;
    s->nprims = ntris + nspheres;
    s->fingerprint = fingerprint64(pr.data(), pr.size() * sizeof(PrimRec),
                                   fingerprint64(sl.data(), sl.size() * sizeof(SLNode), ntris));
    if (!bvh->tree.nodes.empty())
        for (int k = 0; k < 3; ++k) {
            s->lo[k] = bvh->tree.nodes[0].bmin[k];
            s->hi[k] = bvh->tree.nodes[0].bmax[k];
        }
    s->bytes[0] = sl.size() * sizeof(SLNode);
    s->bytes[1] = mt.size() * sizeof(MTNode);
    s->bytes[2] = pr.size() * sizeof(PrimRec);
    s->bytes[3] = p4.size() * sizeof(float) + (size_t)ntris * 12 + (size_t)nspheres * 16;
    *out = s;
    return RGPU_OK;
    RGPU_GUARD_END(ctx)
}

int rgpu_scene_bytes(const rgpu_scene* s, uint64_t* out4) {
    if (!s || !out4) return RGPU_ERR_INVALID_ARG;
    for (int i = 0; i < 4; ++i) out4[i] = s->bytes[i];
    return RGPU_OK;
}

// ------------------------------------------------------------------ ray cast
int rgpu_extension_ray_cast(rgpu_ctx* ctx, rgpu_stream stream, const rgpu_scene* scene, int variant,
                            const rgpu_ray* rays, rgpu_hit* hits, uint32_t count, uint32_t block) {
    if (!ctx || !scene) return RGPU_ERR_INVALID_ARG;
    if (count && (!rays || !hits)) return fail(ctx, RGPU_ERR_INVALID_ARG, "null ray or hit buffer");
    if (!valid_variant(variant))
        return fail(ctx, RGPU_ERR_INVALID_ARG, "unknown traversal variant");
    if (block != 0 && (block % 64 != 0 || block > 1024))
        return fail(ctx, RGPU_ERR_INVALID_ARG, "block must be a multiple of 64 up to 1024");
    if (block > 256) return fail(ctx, RGPU_ERR_INVALID_ARG, "traversal kernels are built for blocks <= 256");
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    TraceTuning tune;
    tune.cu_count = (uint32_t)ctx->cu_count;
    tune.waves_per_simd = ctx->trace_waves;
    tune.refill = ctx->trace_refill;
    tune.leaf_batch = ctx->trace_leaf_batch;
    tune.treelet = ctx->trace_treelet;
    RGPU_HIP(ctx, launch_trace(scene->dev, variant, (const TraceRay*)rays, (TraceHit*)hits, count, block,
                               stream_of(ctx, stream), nullptr, &tune));
    return RGPU_OK;
}

int rgpu_set_trace_tuning(rgpu_ctx* ctx, uint32_t waves_per_simd, uint32_t refill, uint32_t leaf_batch,
                          uint32_t treelet) {
    if (!ctx || waves_per_simd > 16 || refill > 64 || leaf_batch > 64) return RGPU_ERR_INVALID_ARG;
    if (treelet > kTreeletMaxNodes && treelet != RGPU_TREELET_OFF) return RGPU_ERR_INVALID_ARG;
    ctx->trace_waves = waves_per_simd;
    ctx->trace_refill = refill;
    ctx->trace_leaf_batch = leaf_batch;
    ctx->trace_treelet = treelet;
    return RGPU_OK;
}

int rgpu_extension_ray_cast_stats(rgpu_ctx* ctx, const rgpu_scene* scene, int variant, const rgpu_ray* rays,
                                  rgpu_hit* hits, uint32_t count, uint64_t* counters5) {
    if (!ctx || !scene || !counters5) return RGPU_ERR_INVALID_ARG;
    if (count && (!rays || !hits)) return fail(ctx, RGPU_ERR_INVALID_ARG, "null ray or hit buffer");
    if (!valid_variant(variant))
        return fail(ctx, RGPU_ERR_INVALID_ARG, "unknown traversal variant");
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    unsigned long long* d = nullptr;
    RGPU_HIP(ctx, hipMalloc(&d, 5 * sizeof(unsigned long long)));
    hipError_t e = hipMemsetAsync(d, 0, 5 * sizeof(unsigned long long), ctx->stream);
    if (e == hipSuccess)
        e = launch_trace(scene->dev, variant, (const TraceRay*)rays, (TraceHit*)hits, count, 256, ctx->stream, d);
    if (e == hipSuccess)
        e = hipMemcpyAsync(counters5, d, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(ctx, e, "rgpu_extension_ray_cast_stats");
    return RGPU_OK;
}

int rgpu_time_extension_ray_cast(rgpu_ctx* ctx, const rgpu_scene* scene, int variant, const rgpu_ray* rays,
                                 rgpu_hit* hits, uint32_t count, uint32_t block, int iters, double* ms) {
    if (!ctx || !scene || !ms || iters < 1) return RGPU_ERR_INVALID_ARG;
    RGPU_HIP(ctx, hipSetDevice(ctx->device));
    hipEvent_t a, b;
    RGPU_HIP(ctx, hipEventCreate(&a));
    RGPU_HIP(ctx, hipEventCreate(&b));
    int rc = RGPU_OK;
    (void)hipEventRecord(a, ctx->stream);
    for (int i = 0; i < iters && rc == RGPU_OK; ++i)
        rc = rgpu_extension_ray_cast(ctx, nullptr, scene, variant, rays, hits, count, block);
    (void)hipEventRecord(b, ctx->stream);
    hipError_t e = hipEventSynchronize(b);
    float total = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&total, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc != RGPU_OK) return rc;
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventElapsedTime");
    *ms = (double)total / iters;
    return RGPU_OK;
}

}  // extern "C"
