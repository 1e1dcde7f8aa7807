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

#define RGPU_ABI_VERSION 3

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
#define RGPU_VARIANT_THREADED 1  /* skip-pointer BVH2, a table per ray-direction class (kernels/extension_ray_cast.cl) */
#define RGPU_VARIANT_THREADED_SINGLE 2 /* skip-pointer BVH2, one table in the build's child order
                                        * (TBVHNode / intersect_tbvh, common/threaded_bvh.hpp:269-396) */

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
    float traversal_cost; /* SAH cost of a node visit relative to a primitive test; 0 = 1 */
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
/* A stream of the context's device at the default priority (the reference's per-lane step and
 * finalize queues, common/wavefront_path_tracing.hpp:346-348), and its release (after it drains). */
int rgpu_stream_create(rgpu_ctx* ctx, rgpu_stream* out);
int rgpu_stream_destroy(rgpu_ctx* ctx, rgpu_stream stream);
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
/* The same build on the GPU of ctx: level by level binned SAH, giving rgpu_bvh_build's tree node
 * for node (same boxes, children, leaves and primitive order). Bins are limited to 2..32 and
 * coordinates must be finite. device_ms (may be NULL): time on the device from the primitive
 * boxes to the last level, uploads and the final renumbering excluded. Replaces the Embree build
 * of SceneManager::buildBVH (common/scene_manager.hpp:517-577) as rgpu_bvh_build does; the host
 * build stays for scenes outside those limits. Synchronous. */
int rgpu_bvh_build_device(rgpu_ctx* ctx, const float* points, uint32_t stride, uint32_t npoints,
                          const uint32_t* indices, uint32_t ntris, const float* spheres, uint32_t nspheres,
                          const rgpu_bvh_options* opt, rgpu_bvh** out, double* device_ms);
int rgpu_bvh_destroy(rgpu_bvh* bvh);
int rgpu_bvh_get_stats(const rgpu_bvh* bvh, rgpu_bvh_stats* out);
/* The built tree in depth-first pre-order: per node its box (6 floats: min xyz, max xyz),
 * children (-1 for a leaf) and leaf slot range; prim_ids maps slots to primitives.
 * Arrays are sized from rgpu_bvh_get_stats (nodes, slots). Any output may be NULL. */
int rgpu_bvh_export(const rgpu_bvh* bvh, float* bounds, int32_t* left, int32_t* right, uint32_t* prim_beg,
                    uint32_t* prim_cnt, uint32_t* prim_ids);

/* ---------------------------------------------------------------- scene (device)
 * Replaces SceneManager::createBuffer (common/scene_manager.hpp:579-591) and the
 * SceneBuffer it fills (common/scene_manager.hpp:284-294): packs the stackless node
 * records, the threaded node tables (one per ray-direction class and one for the single-table
 * walk, each in its own pre-order) and the primitive records, and uploads them with the
 * vertices. Host arrays are copied; the caller keeps ownership. A BVH of 2^30 nodes or more
 * (over 2^29 primitives) is refused with RGPU_ERR_INVALID_ARG: the walk's node ids carry two
 * flag bits. */
int rgpu_scene_create(rgpu_ctx* ctx, const rgpu_bvh* bvh, const float* points, uint32_t stride, uint32_t npoints,
                      const uint32_t* indices, uint32_t ntris, const float* spheres, uint32_t nspheres,
                      rgpu_scene** out);
int rgpu_scene_destroy(rgpu_scene* scene);
/* Device bytes held by the scene per table: [stackless nodes, the seven threaded tables,
 * primitive records, vertices]. */
int rgpu_scene_bytes(const rgpu_scene* scene, uint64_t* out4);

/* ---------------------------------------------------------------- extension ray cast
 * Replaces the extension_ray_cast kernel and its launch
 * (kernels/extension_ray_cast_stackless.cl:202-246, kernels/extension_ray_cast.cl:96-132;
 * call site common/wavefront_path_tracing.hpp:832-844) for a flat ray batch:
 * closest hit in [0, ray.tmax] for `count` device-resident rays. block = 0 runs persistent
 * waves that refill idle lanes from per-wave ray chunks (best for incoherent rays); block =
 * 64/128/256 runs one thread per ray in workgroups of that size (best for coherent rays). The
 * results are bit-identical either way. */
int rgpu_extension_ray_cast(rgpu_ctx* ctx, rgpu_stream stream, const rgpu_scene* scene, int variant,
                            const rgpu_ray* rays, rgpu_hit* hits, uint32_t count, uint32_t block);
/* Launch shape of the persistent ray-cast kernel used when block == 0: resident waves per
 * SIMD (grid = CUs * waves_per_simd workgroups of 256; 1..16), the number of idle lanes of a
 * wave that makes it hand out its next rays (1..64), and the number of lanes waiting at leaves
 * that makes it test their primitives (1..64). 0 keeps the default (8 waves; refill at 24 idle
 * lanes for stackless, 48 for threaded; leaves at 8 waiting lanes for stackless, 6 for threaded,
 * or once as many lanes wait as walk on). leaf_batch also applies to the one-thread-per-ray kernels (block != 0), whose
 * default is 64 (leaves only once as many lanes wait as walk on) and where 1 selects the plain
 * one-ray walk. treelet: the number of top-of-tree stackless node records each
 * workgroup of the persistent kernel copies to LDS and reads from there (1..1024, 64 bytes
 * each; 0 = default 320; RGPU_TREELET_OFF = none). Results do not depend on any of them. No
 * counterpart in the reference (it leaves launch shapes to the OpenCL driver,
 * common/raccoon_ocl.hpp:906). */
#define RGPU_TREELET_OFF 0xFFFFFFFFu
int rgpu_set_trace_tuning(rgpu_ctx* ctx, uint32_t waves_per_simd, uint32_t refill, uint32_t leaf_batch,
                          uint32_t treelet);

/* Instrumented extension ray cast: the same traversal (identical hits) with per-ray event
 * counting, reduced to 5 totals {stackless branch visits, stackless leaf visits, stackless
 * parent steps, primitive tests, threaded node visits}. The reference measures kernels with
 * profiling events only (common/raccoon_ocl.hpp:172-189); these counts give the algorithmic
 * bytes the roofline is priced on. Synchronous; allocates a 40-byte device buffer. */
int rgpu_extension_ray_cast_stats(rgpu_ctx* ctx, const rgpu_scene* scene, int variant, const rgpu_ray* rays,
                                  rgpu_hit* hits, uint32_t count, uint64_t* counters5);

/* ================================================================ wavefront path tracing */

typedef struct rgpu_materials rgpu_materials;
typedef struct rgpu_envmap rgpu_envmap;
typedef struct rgpu_wavefront rgpu_wavefront;

/* material types (kernels/types.cl:4-10) */
#define RGPU_MAT_LAMBERTIAN 1
#define RGPU_MAT_SPECULAR 2
#define RGPU_MAT_DIELECTRIC 3
#define RGPU_MAT_WARD 4
#define RGPU_MAT_VOLUME 5 /* boundary of a homogeneous medium */

/* Lambertian (kernels/types.cl:78-82). back_emission is stored but, as in the reference,
 * emission is two-sided (kernels/lambertian.cl:113). */
typedef struct rgpu_lambertian {
    float Le[3];
    float R[3];
    int32_t back_emission;
} rgpu_lambertian;
/* Ward (kernels/types.cl:92-97) with separate roughness per tangent axis; alpha_x == alpha_y
 * evaluates the reference's isotropic formulas. */
typedef struct rgpu_ward {
    float alpha_x, alpha_y;
    float reflectance[3];
    float edgetint[3];
    float falloff;
} rgpu_ward;
typedef struct rgpu_dielectric {
    float ior; /* the reference hard-codes 1.5 (kernels/delta_materials.cl:95-96) */
} rgpu_dielectric;
/* HomogeneousVolume (kernels/types.cl:99-102): C = extinction, R = scattering colour. */
typedef struct rgpu_volume {
    float C;
    float R[3];
} rgpu_volume;

/* StandardCamera (kernels/camera.cl:4-22, common/wavefront_path_tracing.hpp:39-48) plus the
 * thin-lens radius the reference hard-codes as 0.0045 (kernels/new_path.cl:61). */
typedef struct rgpu_camera {
    float eye[3], forward[3], up[3], right[3];
    float imageplane_o[3], imageplane_r[3], imageplane_b[3];
    uint32_t width, height;
    float lens_radius;
} rgpu_camera;

#define RGPU_WF_ACCUM64 1u              /* accumulate pixel sums in double, 32-bit counts */
#define RGPU_WF_NO_REFERENCE_QUIRKS 2u  /* skip the stale-volume strategy draws of new paths */
/* Sample the lambertian cosine lobe and evaluate envmap pdfs about the normal flipped toward
 * the outgoing direction. The reference uses the unflipped normal there
 * (kernels/lambertian.cl:59-64, kernels/envmap_sampling.cl:263-265) while it shades and samples
 * the envmap with the flipped one: on a back face its cosine-lobe samples point into the
 * surface and return nothing, and a Ward vertex prices envmap directions on the wrong
 * hemisphere. */
#define RGPU_WF_CONSISTENT_NORMALS 4u

typedef struct rgpu_wavefront_config {
    uint32_t abi_version;   /* RGPU_ABI_VERSION */
    uint32_t path_count;    /* resident path slots (kWavefrontPathCountGPU = 2^22 in the reference) */
    uint64_t seed;          /* added to the splitmix64 seed of every slot (reference: 0) */
    uint32_t rr_start;      /* Russian roulette after this many bounces; 0 = 25 (kernels/logic.cl:77) */
    uint32_t max_depth;     /* terminate a path at this bounce; 0 = unlimited (reference) */
    float radiance_clamp;   /* drop samples with a channel >= clamp; 0 = 1000 (kernels/logic.cl:126) */
    float bxdf_select_p;    /* BSDF selection probability; 0 = 0.5 (kernels/mixture_density.cl:40) */
    uint32_t flags;         /* RGPU_WF_* */
    uint32_t variant;       /* RGPU_VARIANT_* traversal used by the extension ray cast */
    uint32_t queue_blocks;  /* workgroups of the queue-driven kernels; 0 = 8 per CU */
} rgpu_wavefront_config;

/* Per-primitive materials: the primitive index is the material index (kernels/
 * extension_ray_cast_stackless.cl:242), so `types`/`indices` have one entry per primitive,
 * triangles first, then spheres. Replaces SceneManager::createMaterialBuffer
 * (common/scene_manager.hpp:593-602) and its MaterialBuffer. */
int rgpu_materials_create(rgpu_ctx* ctx, uint32_t nprims, const int32_t* types, const int32_t* indices,
                          const rgpu_lambertian* lambertians, uint32_t nlambertian, const rgpu_ward* wards,
                          uint32_t nward, const rgpu_dielectric* dielectrics, uint32_t ndielectric,
                          const rgpu_volume* volumes, uint32_t nvolume, rgpu_materials** out);
int rgpu_materials_destroy(rgpu_materials* m);

/* Environment map and its 7 alias tables (1 uniform-weight + 6 six-axis). Replaces the
 * ImageEnvmap construction in SceneManager::addPoint and createEnvmapBuffer
 * (common/scene_manager.hpp:432-497, 604-650). rgb: width*height*channels floats, row 0 at
 * the zenith. ctx may be NULL to build the host tables only (rgpu_envmap_export). */
int rgpu_envmap_create(rgpu_ctx* ctx, const float* rgb, uint32_t channels, uint32_t width, uint32_t height,
                       rgpu_envmap** out);
int rgpu_envmap_destroy(rgpu_envmap* e);
/* Host copy of table `table` (0 = uniform, 1..6 = +X,-X,+Y,-Y,+Z,-Z): pdf per texel, alias
 * heights and aliases, fragments (4 floats: beg_y, end_y, beg_phi, end_phi), and whether the
 * table fell back from all-zero weights. Outputs may be NULL. */
int rgpu_envmap_export(const rgpu_envmap* e, int table, float* pdf, float* heights, int32_t* aliases,
                       float* fragments, int32_t* degenerate);
/* Flat-batch environment sampling: replaces sample_envmap_6axis_stage and sample_envmap_stage
 * (kernels/envmap_sampling.cl:169-248) over `count` device-resident samples with no queue
 * indirection. states: 4 words per sample (the xoshiro128** state), advanced in place by the
 * draws of one sample. normals: 3 floats per sample selecting the six-axis tables (the caller
 * flips them to the outgoing side, as the reference stage does at :199-203), or NULL for the
 * uniform-weight table. out: 4 floats per sample (direction x, y, z, pdf in solid angle).
 * states and out 16-byte aligned. The envmap must have been created on ctx. */
int rgpu_envmap_sample(rgpu_ctx* ctx, rgpu_stream stream, const rgpu_envmap* e, uint32_t* states,
                       const float* normals, uint32_t count, float* out);
/* The matching densities of given directions (3 floats each): replaces
 * evaluate_envmap_6axis_pdf_stage and evaluate_envmap_pdf_stage
 * (kernels/envmap_sampling.cl:250-283). normals as above. */
int rgpu_envmap_pdf(rgpu_ctx* ctx, rgpu_stream stream, const rgpu_envmap* e, const float* dirs, const float* normals,
                    uint32_t count, float* pdfs);

/* Pinhole/thin-lens camera from a look-at (the standardCamera image-plane construction of
 * common/wavefront_path_tracing.hpp:103-118). focus_distance 0 = distance to target. */
int rgpu_camera_look_at(const float* eye, const float* target, const float* up, float fov_y_deg,
                        float focus_distance, uint32_t width, uint32_t height, float lens_radius,
                        rgpu_camera* out);

/* WavefrontLane (common/wavefront_path_tracing.hpp:339-1153): resident path state, stage
 * queues and accumulation on one GPU. The scene, materials and envmap must outlive it. */
int rgpu_wavefront_create(rgpu_ctx* ctx, const rgpu_scene* scene, const rgpu_materials* materials,
                          const rgpu_envmap* envmap, const rgpu_wavefront_config* cfg, rgpu_wavefront** out);
int rgpu_wavefront_destroy(rgpu_wavefront* w);
/* WavefrontLane::initialize(lane_index, num_device) (wavefront_path_tracing.hpp:532-570):
 * RNG seeding, every slot a new path, pixel cursor at lane * (P / lanes), cleared
 * accumulation. Also the restart path of re_start() (wavefront_path_tracing.hpp:592-604). */
int rgpu_wavefront_initialize(rgpu_wavefront* w, const rgpu_camera* camera, uint32_t lane_index,
                              uint32_t num_lanes);
/* `steps` calls of WavefrontLane::step() (wavefront_path_tracing.hpp:591-939), asynchronous. */
int rgpu_wavefront_step(rgpu_wavefront* w, rgpu_stream stream, uint32_t steps);
/* Runs `steps` steps between two events on the context stream; mean ms per step. */
int rgpu_wavefront_time_steps(rgpu_wavefront* w, uint32_t steps, double* ms_per_step);
/* Runs `steps` steps (<= 4096) with HIP events between their stages on the context stream and
 * returns the mean ms per step of each of the RGPU_STAGES stages, in the order of
 * WavefrontLane::step (wavefront_path_tracing.hpp:606-875): new path; guide inference (0 when no
 * guide samples); lambertian; Ward; mirror and dielectric; medium boundary and inside; extension
 * ray cast with the medium free flight; logic; queue scan and scatter. Synchronous. */
#define RGPU_STAGES 9
int rgpu_wavefront_time_stages(rgpu_wavefront* w, uint32_t steps, double* ms_per_stage);
/* Queue lengths after the last step (new path, lambertian, specular, dielectric, ward,
 * volume boundary, volume inside) and the pixel cursor next_pixel_index. Synchronous. */
int rgpu_wavefront_queue_counts(rgpu_wavefront* w, uint32_t* counts7, uint64_t* next_pixel);
/* Per-pixel radiance sums (3 doubles) and sample counts. Synchronous. */
int rgpu_wavefront_read_accumulation(rgpu_wavefront* w, double* rgb_sums, uint32_t* counts);
/* Mean radiance -> RGBA8 through the half-float intermediate and gamma 2.2
 * (kernels/accumlation.cl:6-25,68-79). Synchronous. */
int rgpu_wavefront_tonemap(rgpu_wavefront* w, uint8_t* rgba8);
/* Total accumulated samples (kernels/stat.cl). Synchronous. */
int rgpu_wavefront_stat(rgpu_wavefront* w, uint64_t* total_samples);
/* Per-slot state for testing: RNG states (4 u32), rays (ro, tmax, rd, pixel), paths (T, logic_i,
 * L, pad), extension results (hit_surface, hit_volume, t, Ng, pad), shading (Le, T, pad),
 * in-volume lists (7 ids + count), queue items (7 x path_count). Any may be NULL. */
int rgpu_wavefront_read_state(rgpu_wavefront* w, uint32_t* rng, float* rays, float* paths, void* ext,
                              float* shade, int32_t* vol, uint32_t* items);

/* Intermediate image of a render: per pixel the half-float mean radiance (IEEE binary16 bits,
 * round to nearest even) and the sample count, widened from the reference's 16 bits to 32
 * (RGB16IntermediateValueType, kernels/types.cl; accumlation_to_intermediate,
 * kernels/accumlation.cl:6-25, without the preview mutex). 12 bytes. */
typedef struct rgpu_rgb16 {
    uint16_t r, g, b, pad;
    uint32_t count;
} rgpu_rgb16;
/* The lane's intermediate image into dst (device memory, one rgpu_rgb16 per pixel). Single
 * renders only. */
int rgpu_wavefront_intermediate(rgpu_wavefront* w, rgpu_stream stream, rgpu_rgb16* dst);
/* Preview handoff between a step thread and a preview thread (the weak device mutex of
 * kernels/mutex.cl with accumlation_to_intermediate and copy_if_locked, kernels/accumlation.cl:6-25;
 * WavefrontLane::step and update_intermediate, common/wavefront_path_tracing.hpp:877-898,942-960).
 * set_publish(every_steps): with every_steps > 0, every that many rgpu_wavefront_step steps (and
 * once after each replay of a captured graph) end by trying a device-side lock and, if they get
 * it, refreshing the lane's published intermediate image; 0 switches publishing off. Single
 * renders only; synchronous.
 * update_intermediate: tries the same lock on `stream` and, if it gets it, copies the published
 * image into dst (device memory, one rgpu_rgb16 per pixel); otherwise dst keeps its previous
 * contents. took (device memory, may be NULL) receives 1 or 0. Asynchronous on `stream`.
 * Neither side ever waits for the other: a step that finds the lock held skips publishing, a
 * preview that finds it held keeps its image, so a copy is never torn and pixel counts in dst
 * never decrease. Intended use: one thread calls rgpu_wavefront_step on one stream while another
 * calls update_intermediate (then rgpu_merge_intermediate / rgpu_tonemap_intermediate) on a
 * second stream. initialize, resume, set_publish and destroy must not overlap a preview call. */
int rgpu_wavefront_set_publish(rgpu_wavefront* w, uint32_t every_steps);
int rgpu_wavefront_update_intermediate(rgpu_wavefront* w, rgpu_stream stream, rgpu_rgb16* dst, int* took);
/* a = count-weighted mean of a and b, counts summed (merge_intermediate,
 * kernels/accumlation.cl:27-50; the merge of several lanes in
 * common/wavefront_path_tracing.hpp:962-991). Device buffers of npixels entries. */
int rgpu_merge_intermediate(rgpu_ctx* ctx, rgpu_stream stream, rgpu_rgb16* a, const rgpu_rgb16* b, uint32_t npixels);
/* RGBA8 of an intermediate: pow(mean, 1/2.2) * 256 clamped, alpha 255 (tonemap,
 * kernels/accumlation.cl:68-79). rgba8: device memory, 4 bytes per pixel. */
int rgpu_tonemap_intermediate(rgpu_ctx* ctx, rgpu_stream stream, const rgpu_rgb16* src, uint32_t npixels,
                              uint8_t* rgba8);

/* Checkpoint / resume of a single render (no batch, no attached guide): a host snapshot of
 * everything a lane carries from one step to the next (rays, path, extension and shading
 * records, medium lists, RNG states, queues, pixel cursor, accumulation, camera). Resuming it
 * into a lane of the same path count and settings (the same lane or a new one over the same
 * scene, materials and envmap) continues the render exactly where the snapshot was taken: the
 * following steps give the slot states the uninterrupted lane would have. The reference has
 * none (its state lives only in device buffers, common/wavefront_path_tracing.hpp:592-604);
 * re_start there re-initialises. Synchronous. checkpoint_size: bytes the snapshot needs.
 * The snapshot records fingerprints of the scene, materials and envmap, the lane's settings
 * (path count, depth limit, roulette start, clamp, BSDF selection, flags) and a fingerprint of
 * its own bytes. resume refuses (RGPU_ERR_INVALID_ARG) before touching the device when any of
 * them differs, when the camera does not match the image size, or when a queue item, pixel
 * index, hit primitive or medium list in the payload is out of range; the lane is then
 * unchanged. A resume that fails later (a HIP error) leaves the lane uninitialised. */
int rgpu_wavefront_checkpoint_size(rgpu_wavefront* w, uint64_t* bytes);
int rgpu_wavefront_checkpoint(rgpu_wavefront* w, void* host, uint64_t bytes);
int rgpu_wavefront_resume(rgpu_wavefront* w, const void* host, uint64_t bytes);

/* hipGraph of the bounce loop: captures `steps` consecutive steps once; each launch replays
 * them. Replaces the reference's per-step host launch list (wavefront_path_tracing.hpp:606-875).
 * The graph holds the camera and buffers by value; initialize() and set_launch() drop it.
 * Calls on one wavefront that overlap on different streams race. */
int rgpu_wavefront_graph_capture(rgpu_wavefront* w, uint32_t steps);
int rgpu_wavefront_graph_launch(rgpu_wavefront* w, rgpu_stream stream, uint32_t launches);

/* Launch configuration: traversal variant, extension-ray-cast workgroup size (64/128/256 threads,
 * one per slot, 256 by default; 0 = the persistent ray-cast kernel over the slots' rays, for a
 * single render; a batch traces 0 as 256) and grid of the queue-driven kernels. Results do not
 * depend on it. The reference leaves workgroup sizes to the driver
 * (common/raccoon_ocl.hpp:906). */
int rgpu_wavefront_set_launch(rgpu_wavefront* w, int variant, uint32_t trace_block, uint32_t queue_blocks);
int rgpu_wavefront_get_launch(const rgpu_wavefront* w, int* variant, uint32_t* trace_block,
                              uint32_t* queue_blocks);
/* Autotuner: times every (variant, trace_block, queue_blocks) configuration for `steps` steps
 * after `warmup`, keeps the fastest and leaves the wavefront initialised with `camera`. */
int rgpu_wavefront_autotune(rgpu_wavefront* w, const rgpu_camera* camera, uint32_t warmup, uint32_t steps,
                            double* best_ms_per_step);

/* ================================================================ batches of render jobs
 * Several render jobs (own scene, materials, envmap, camera, resolution, samples per pixel and
 * seed) in one persistent wavefront, generalising the reference's single pixel cursor
 * next_pixel_index (kernels/new_path.cl:41-43) to one cursor per job. Job j owns
 * ceil(path_count / 256) blocks of 256 consecutive slots; its queue items stay one slot-ordered
 * run in every stage queue, so each job's slots evolve exactly as a single render of that job
 * with the same path_count (rounded up to 256), seed, scene and camera. A job with spp > 0
 * starts width * height * spp samples (every pixel spp times), then parks its slots. */
typedef struct rgpu_batch_job {
    const rgpu_scene* scene;
    const rgpu_materials* materials;
    const rgpu_envmap* envmap;
    rgpu_camera camera;
    uint32_t path_count;    /* slots of this job; rounded up to a multiple of 256 */
    uint32_t spp;           /* samples per pixel to start; 0 = unlimited */
    uint64_t seed;          /* RNG seed, as rgpu_wavefront_config.seed of a single render */
} rgpu_batch_job;

/* Creates and initialises the batch (cfg->path_count and cfg->seed are ignored; RGPU_WF_ACCUM64
 * is not supported). step, time_steps, queue_counts (totals), read_state, graph_capture/launch,
 * set_launch, autotune (camera may be NULL) and guiding apply to a batch; initialize and the
 * single-render readers do not. The scenes, materials and envmaps must outlive it. A guide
 * attached to a batch is one network for all jobs: a vertex's position is normalised to its
 * job's scene box (the BVH root box grown by 1e-3) and its scene is an input (one-hot class,
 * in order of first use; scenes past the 8th share "no class"). */
int rgpu_wavefront_create_batch(rgpu_ctx* ctx, const rgpu_batch_job* jobs, uint32_t njobs,
                                const rgpu_wavefront_config* cfg, rgpu_wavefront** out);
int rgpu_wavefront_job_count(const rgpu_wavefront* w, uint32_t* njobs);
/* A job's queue lengths after the last step and the samples it has started. Synchronous. */
int rgpu_wavefront_job_queue_counts(rgpu_wavefront* w, uint32_t job, uint32_t* counts7, uint64_t* samples_started);
/* A job's per-pixel radiance sums (3 floats) and sample counts; either may be NULL. Synchronous. */
int rgpu_wavefront_job_accumulation(rgpu_wavefront* w, uint32_t job, float* rgb_sums, uint32_t* counts);
int rgpu_wavefront_job_stat(rgpu_wavefront* w, uint32_t job, uint64_t* total_samples);
int rgpu_wavefront_job_tonemap(rgpu_wavefront* w, uint32_t job, uint8_t* rgba8);
/* rgpu_wavefront_job_tonemap without any device allocation, for callers that cannot allocate
 * (serving loops that preview many jobs, harnesses that forbid hipMalloc). No counterpart in
 * the reference. It stages the image in a static device array the library holds per device
 * until the process ends (4 MiB), so a job may have at most 1024 x 1024 pixels; larger jobs
 * return RGPU_ERR_INVALID_ARG. Output equals rgpu_wavefront_job_tonemap's bit for bit.
 * Synchronous. Every call on one device shares that staging: calls that overlap on one device
 * (from two threads or two contexts) race. */
int rgpu_wavefront_job_tonemap_noalloc(rgpu_wavefront* w, uint32_t job, uint8_t* rgba8);
/* Steps (check_every at a time) until every job has finished its samples or max_steps have
 * run; every job needs spp > 0. With a training guide attached, a training pass runs after
 * every check_every steps once `capacity` samples are buffered. */
int rgpu_wavefront_run_batch(rgpu_wavefront* w, uint32_t max_steps, uint32_t check_every, uint32_t* steps_done);

/* ================================================================ path guiding
 * No counterpart in the reference (it has no path guiding). A small MLP maps a vertex (12-band
 * frequency-encoded position in the scene box, normal, outgoing direction) to an 8-lobe von
 * Mises-Fisher mixture; at lambertian and Ward vertices it is a third strategy of the
 * reference's one-sample mixture pdf (kernels/types.cl:58-71, kernels/mixture_density.cl),
 * chosen with probability select_p. It trains online (Adam) on samples the tracer records. */
typedef struct rgpu_guide rgpu_guide;
typedef struct rgpu_guide_config {
    uint32_t abi_version;
    float select_p;        /* guide selection probability in (0, 1), e.g. 0.5 */
    uint32_t capacity;     /* training samples per pass; 0 = 2^18 */
    uint32_t minibatch;    /* rows per Adam step; 0 = 16384 */
    float learning_rate;   /* 0 = 1e-3 */
    uint32_t seed;         /* weight initialisation */
    float scene_lo[3];     /* scene box that normalises positions */
    float scene_hi[3];
    uint32_t max_depth;    /* guide the first max_depth vertices of a path; 0 = every vertex */
    float weight_cap;      /* training weights F/p above it are clamped to it; 0 = no clamp.
                              Affects what the network learns only: the estimator stays unbiased */
    float log_kappa_max;   /* upper clamp of ln(kappa) of a lobe, in (-3, 12]; 0 = 7.5 */
    uint32_t learn_select; /* 1: each vertex selects the guide with a probability the network learns
                              (output 40, clamped to [0.05, 0.95]; Mueller et al. 2019), trained on
                              the log density of the whole one-sample mixture; select_p is then
                              unused. 0: the fixed select_p. No reference counterpart (the
                              reference has no path guiding). */
} rgpu_guide_config;

int rgpu_guide_create(rgpu_ctx* ctx, const rgpu_guide_config* cfg, rgpu_guide** out);
/* Lifetime: a guide must outlive the wavefronts it is attached to. rgpu_guide_destroy refuses
 * (RGPU_ERR_INVALID_ARG, guide unchanged) while any wavefront still has it attached; detach it
 * first (rgpu_wavefront_attach_guide with g = NULL) or destroy those wavefronts. */
int rgpu_guide_destroy(rgpu_guide* g);
int rgpu_guide_param_count(void);
/* Copy the fp32 weights out (get) and/or in (set; refreshes the bf16 copies). */
int rgpu_guide_params(rgpu_guide* g, float* get, const float* set);
/* Test hooks over host rows of 16 floats (position in [0,1]^3, normal, wo, wi, weight, scene
 * class code (0 = none, 1..8: one-hot input), the density of the other strategies at wi (0 = not
 * known), pad 1):
 * raw head outputs (64 per row), the gradient of the mean loss -weight*log q(wi) (plus
 * -weight*log of the one-sample mixture density with learn_select, for rows with a known other
 * density; at most minibatch rows), and training (minibatch Adam steps over at most capacity
 * rows). */
int rgpu_guide_forward(rgpu_guide* g, const float* samples, uint32_t rows, float* out);
int rgpu_guide_gradients(rgpu_guide* g, const float* samples, uint32_t rows, float* grad);
int rgpu_guide_train_samples(rgpu_guide* g, const float* samples, uint32_t rows);
/* Test hook: copy the first `rows` rows (16 floats, as above) of the training buffer a
 * wavefront has filled; *buffered = samples in it (capped at capacity). Synchronous. */
int rgpu_guide_read_samples(rgpu_guide* g, float* out, uint32_t rows, uint32_t* buffered);
/* Attach (or detach with g = NULL) a guide: `sampling` enables guided sampling, `training`
 * records vertices and collects training samples. Attaching replaces any guide attached before.
 * The lane keeps its per-slot guiding buffers across a detach and reuses them on the next
 * attach; paths in flight start recording afresh. Guide and wavefront must share a device. */
int rgpu_wavefront_attach_guide(rgpu_wavefront* w, rgpu_guide* g, int sampling, int training);
int rgpu_wavefront_guide_pending(rgpu_wavefront* w, uint32_t* pending);
/* Test hook: the lobe records the last guide inference wrote, records [first, first + count) in
 * the order of the bounce's vertices (the lambertian queue, then the Ward queue; rgpu_wavefront_
 * state's queue items give their slots): 48 floats each, 8 lobes of (mean x, y, z, kappa,
 * weight, log(weight C(kappa))). */
int rgpu_wavefront_guide_lobes(rgpu_wavefront* w, uint32_t first, uint32_t count, float* out);
/* If at least min_samples are buffered: one pass of minibatch Adam steps, buffer emptied. */
int rgpu_wavefront_guide_train(rgpu_wavefront* w, uint32_t min_samples, uint32_t* trained);

#ifdef __cplusplus
}
#endif

#endif /* RACCOON_GPU_H */
