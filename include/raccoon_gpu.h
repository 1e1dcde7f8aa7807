/*
 * raccoon_gpu.h — C ABI of the MI355X (gfx950) wavefront path-tracing core.
 *
 * This is the drop-in boundary that replaces the reference's OpenCL runtime layer
 * (common/raccoon_ocl.hpp) and its kernel launch sites for the path-tracing hot path
 * (common/wavefront_path_tracing.hpp:591-939). Plain pointers and sizes only; no C++
 * or torch types cross it. Every entry point returns an int status: RGPU_OK (0) or a
 * negative code; the message of the last failure on a context is available from
 * rgpu_last_error_string(). Nothing aborts the process: where the reference asserts
 * (RAC_ASSERT, common/raccoon_ocl.hpp:20-28) this layer returns a status.
 *
 * Threading: a context may be driven from several host threads provided each uses its
 * own stream; calls on the same stream from two threads at once are not supported.
 * Device pointers passed to the rgpu_* calls must come from rgpu_alloc on the same
 * context (or another allocator on the same device).
 */
#ifndef RACCOON_GPU_H
#define RACCOON_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RGPU_ABI_VERSION 1

/* status codes */
#define RGPU_OK 0
#define RGPU_ERR_INVALID_ARG (-1)
#define RGPU_ERR_OUT_OF_MEMORY (-2)
#define RGPU_ERR_HIP (-3)
#define RGPU_ERR_NO_DEVICE (-4)
#define RGPU_ERR_NO_CODE_OBJECT (-5)
#define RGPU_ERR_UNSUPPORTED (-6)

/* traversal variants */
#define RGPU_VARIANT_STACKLESS 0 /* parent-pointer BVH2, bitstack (kernels/extension_ray_cast_stackless.cl) */
#define RGPU_VARIANT_THREADED 1  /* skip-pointer BVH2, 6 direction tables (kernels/extension_ray_cast.cl) */

typedef struct rgpu_ctx rgpu_ctx;
typedef struct rgpu_bvh rgpu_bvh;
typedef struct rgpu_scene rgpu_scene;
typedef void* rgpu_stream; /* a hipStream_t; NULL selects the context's own stream */

/* 32-byte ray: origin, far clip (the reference's incoming ExtensionResult.tmin), direction. */
typedef struct rgpu_ray {
    float ro[3];
    float tmax;
    float rd[3];
    uint32_t pad;
} rgpu_ray;

/* 16-byte hit: primitive id (-1 on a miss), t (tmax on a miss), barycentrics u, v. */
typedef struct rgpu_hit {
    int32_t prim;
    float t;
    float u;
    float v;
} rgpu_hit;

typedef struct rgpu_config {
    uint32_t abi_version; /* RGPU_ABI_VERSION */
    uint32_t flags;       /* reserved, 0 */
} rgpu_config;

typedef struct rgpu_bvh_options {
    uint32_t abi_version; /* RGPU_ABI_VERSION */
    uint32_t max_leaf;    /* primitives per leaf, 1..8 (default 4) */
    uint32_t bins;        /* SAH bins per axis (default 16) */
    uint32_t max_depth;   /* deepest leaf, <= 64 (default 63) */
} rgpu_bvh_options;

typedef struct rgpu_bvh_stats {
    uint32_t nodes;
    uint32_t leaves;
    uint32_t depth;
    uint32_t slots; /* leaf primitive slots (= primitives) */
    double sah_cost;
} rgpu_bvh_stats;

/* ---------------------------------------------------------------- lifecycle
 * Replaces OpenCLContext + lane_initialize (common/raccoon_ocl.hpp:568-790). */
int rgpu_create(int device_index, const rgpu_config* cfg, rgpu_ctx** out);
int rgpu_destroy(rgpu_ctx* ctx);
const char* rgpu_last_error_string(const rgpu_ctx* ctx);
int rgpu_device_count(int* out);

/* ---------------------------------------------------------------- memory
 * Replaces OpenCLBuffer create/fill/read/write (common/raccoon_ocl.hpp:353-501). */
int rgpu_alloc(rgpu_ctx* ctx, size_t bytes, void** dptr);
int rgpu_free(rgpu_ctx* ctx, void* dptr);
int rgpu_upload(rgpu_ctx* ctx, rgpu_stream stream, void* dst, const void* src, size_t bytes);
int rgpu_download(rgpu_ctx* ctx, rgpu_stream stream, void* dst, const void* src, size_t bytes);
/* Fill `count` 32-bit words with `value` (OpenCLBuffer::fill, raccoon_ocl.hpp:427-434). */
int rgpu_fill_u32(rgpu_ctx* ctx, rgpu_stream stream, void* dst, uint32_t value, size_t count);
int rgpu_synchronize(rgpu_ctx* ctx, rgpu_stream stream);
/* Device-to-device copy with the library's own copy kernel (copy_if_locked without the
 * lock, kernels/mutex.cl:19-24); asynchronous on `stream`. */
int rgpu_copy(rgpu_ctx* ctx, rgpu_stream stream, void* dst, const void* src, size_t bytes);

/* ---------------------------------------------------------------- timing
 * Replaces profiling events (OpenCLEvent::wait, common/raccoon_ocl.hpp:172-189).
 * Records a pair of events around `iters` launches of rgpu_extension_ray_cast on the
 * context stream and returns the mean milliseconds per launch. */
int rgpu_time_extension_ray_cast(rgpu_ctx* ctx, const rgpu_scene* scene, int variant, const rgpu_ray* rays,
                                 rgpu_hit* hits, uint32_t count, uint32_t block, int iters, double* ms_per_launch);

/* ---------------------------------------------------------------- BVH (host)
 * Replaces SceneManager::buildBVH -> create_embreeBVH -> create_stackless_bvh
 * (common/scene_manager.hpp:517-577, common/stackless_bvh.hpp:108-130,328-341):
 * binned-SAH build on the host. Triangles are primitives 0..ntris-1 and spheres
 * follow (sphereBegin = ntris), as in the reference. points: `stride` floats per point
 * (>= 3); spheres: cx, cy, cz, r. */
int rgpu_bvh_build(const float* points, uint32_t stride, uint32_t npoints, const uint32_t* indices, uint32_t ntris,
                   const float* spheres, uint32_t nspheres, const rgpu_bvh_options* opt, rgpu_bvh** out);
int rgpu_bvh_destroy(rgpu_bvh* bvh);
int rgpu_bvh_get_stats(const rgpu_bvh* bvh, rgpu_bvh_stats* out);
/* The built tree in depth-first pre-order: per node its box (6 floats: min xyz, max xyz),
 * children (-1 for a leaf) and leaf slot range; prim_ids maps slots to primitives.
 * Arrays are sized from rgpu_bvh_get_stats (nodes, slots). Any output may be NULL. */
int rgpu_bvh_export(const rgpu_bvh* bvh, float* bounds, int32_t* left, int32_t* right, uint32_t* prim_beg,
                    uint32_t* prim_cnt, uint32_t* prim_ids);

/* ---------------------------------------------------------------- scene (device)
 * Replaces SceneManager::createBuffer (common/scene_manager.hpp:579-591) and the
 * SceneBuffer it fills (common/scene_manager.hpp:284-294): packs the stackless and
 * threaded node records and the primitive records and uploads them with the
 * vertices. Host arrays are copied; the caller keeps ownership. */
int rgpu_scene_create(rgpu_ctx* ctx, const rgpu_bvh* bvh, const float* points, uint32_t stride, uint32_t npoints,
                      const uint32_t* indices, uint32_t ntris, const float* spheres, uint32_t nspheres,
                      rgpu_scene** out);
int rgpu_scene_destroy(rgpu_scene* scene);
/* Device bytes held by the scene per table: [stackless nodes, threaded tables, primitive records, vertices]. */
int rgpu_scene_bytes(const rgpu_scene* scene, uint64_t* out4);

/* ---------------------------------------------------------------- extension ray cast
 * Replaces the extension_ray_cast kernel and its launch
 * (kernels/extension_ray_cast_stackless.cl:202-246, kernels/extension_ray_cast.cl:96-132;
 * call site common/wavefront_path_tracing.hpp:832-844) for a flat ray batch:
 * closest hit in [0, ray.tmax] for `count` device-resident rays. block = 0 picks the
 * default workgroup size. */
int rgpu_extension_ray_cast(rgpu_ctx* ctx, rgpu_stream stream, const rgpu_scene* scene, int variant,
                            const rgpu_ray* rays, rgpu_hit* hits, uint32_t count, uint32_t block);

/* Instrumented extension ray cast: the same traversal (identical hits) with per-ray event
 * counting, reduced to 5 totals {stackless branch visits, stackless leaf visits, stackless
 * parent steps, primitive tests, threaded node visits}. The reference measures kernels with
 * profiling events only (common/raccoon_ocl.hpp:172-189); these counts give the algorithmic
 * bytes the roofline is priced on. Synchronous; allocates a 40-byte device buffer. */
int rgpu_extension_ray_cast_stats(rgpu_ctx* ctx, const rgpu_scene* scene, int variant, const rgpu_ray* rays,
                                  rgpu_hit* hits, uint32_t count, uint64_t* counters5);

#ifdef __cplusplus
}
#endif

#endif /* RACCOON_GPU_H */
