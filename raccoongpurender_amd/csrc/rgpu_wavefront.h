// Device layouts and launch interface of the wavefront path-tracing stages.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "envmap_build.h"
#include "guide_device.h"
#include "rgpu_traverse.h"

namespace rgpu {

// Material type ids (kernels/types.cl:4-10, common/material_type.hpp:14-19).
enum : int32_t {
    kMatNone = 0,
    kMatLambertian = 1,
    kMatSpecular = 2,
    kMatDielectric = 3,
    kMatWard = 4,
    kMatVolume = 5,        // boundary of a homogeneous medium
    kMatVolumeInside = 6,  // scattering event inside a medium (queue id only)
};
static const int kNumQueues = 7;  // new path + the six material queues of kernels/logic.cl:232-246

struct Material {
    int32_t type;
    int32_t index;
};
struct Lambertian {  // kernels/types.cl:78-82
    float Le[3];
    float R[3];
    int32_t back_emission;
    int32_t pad;
};
struct Ward {  // kernels/types.cl:92-97, with separate alpha_x / alpha_y
    float alpha_x, alpha_y;
    float reflectance[3];
    float edgetint[3];
    float falloff;
    float pad;
};
struct Dielectric {
    float ior;  // the reference hard-codes 1.5 (kernels/delta_materials.cl:95-96)
};
struct Volume {  // kernels/types.cl:99-102
    float C;     // extinction used for free flight
    float R[3];  // single-scattering colour
};

// Per-slot state records (32 B each; the reference's AoS structs of kernels/types.cl).
struct PathRec {  // WavefrontPath T, L, logic_i (ro/rd/pixel live in the TraceRay)
    float T[3];
    uint32_t logic_i;
    float L[3];
    uint32_t pad;
};
struct ExtRec {  // ExtensionResult
    int32_t hit_surface;
    int32_t hit_volume;
    float tmin;
    float ngx;
    float ngy;
    float ngz;
    float pad0, pad1;
};
struct ShadeRec {  // ShadingResult
    float Le[3];
    float T[3];
    float pad0, pad1;
};
struct VolList {  // InVolumeList
    int32_t volumes[7];
    int32_t count;
};

struct Camera {  // StandardCamera (kernels/camera.cl:4-22) + lens radius
    float eye[3], forward[3], up[3], right[3];
    float o[3], r[3], b[3];  // image plane origin and per-pixel steps
    uint32_t width, height;
    float lens_radius;
};

struct WFParams {
    uint32_t n;               // resident path slots
    uint32_t rr_start;        // Russian roulette after this many bounces (25)
    uint32_t max_depth;       // 0 = unlimited
    uint32_t accum64;         // accumulate in double
    uint32_t quirks;          // reproduce the reference's stale-volume strategy draws
    float clamp;              // per-sample radiance clamp (1000)
    float bxdf_p;             // selection probability of BSDF sampling (0.5)
    uint32_t consistent;      // flip the normal toward wo in BSDF sampling and envmap pdfs too
};

struct DevMaterials {
    const Material* mat = nullptr;  // one per primitive
    const Lambertian* lambertian = nullptr;
    const Ward* ward = nullptr;
    const Dielectric* dielectric = nullptr;
    const Volume* volume = nullptr;
};

// One 64-byte record per texel: the pdfs of the seven tables (uniform, +X, -X, +Y, -Y, +Z, -Z;
// the eighth word unused) and the texel's sampling fragment, so that drawing a direction or
// pricing one gathers a single cache line for the texel instead of one per table.
struct alignas(64) EnvTexelRec {
    float pdf[8];
    EnvFragment frag;
    float pad[4];
};
static_assert(sizeof(EnvTexelRec) == 64, "texel record is one 64-byte line");

struct DevEnvmap {
    const float4* texels = nullptr;       // width*height RGBA
    const EnvTexelRec* rec = nullptr;     // width*height
    const float* pdf0 = nullptr;          // width*height: the uniform table's pdfs again, densely
                                          // packed for pricing arbitrary directions (16 per line)
    const AliasBucket* alias = nullptr;   // 7 tables of width*height
    uint32_t width = 0, height = 0;
};

struct DevState {
    TraceRay* ray = nullptr;  // ro, tmax, rd, pixel index
    PathRec* path = nullptr;
    ExtRec* ext = nullptr;
    ShadeRec* shade = nullptr;
    VolList* vol = nullptr;
    uint4* rng = nullptr;
    uint8_t* qsel = nullptr;          // queue chosen by logic
    uint32_t* items = nullptr;        // kNumQueues * n queue items
    uint32_t* counts = nullptr;       // kNumQueues counts
    uint32_t* block_counts = nullptr; // kNumQueues * nblocks
    unsigned long long* next_pixel = nullptr;
    float* accum = nullptr;           // 4 floats per pixel (r, g, b, count bits) or
    double* accum64 = nullptr;        // 3 doubles per pixel + counts in accum_count
    uint32_t* accum_count = nullptr;
    TraceRay* xray = nullptr;         // single render, persistent trace: extension rays (far clip, medium)
    TraceHit* xhit = nullptr;         // and their hits
    uint32_t nblocks = 0;             // logic blocks (n / 256 rounded up)
};

// ---------------------------------------------------------------- batches of render jobs
// A batch runs several render jobs in one wavefront: job j owns the slots of 256-slot blocks
// [first_block, first_block + nblocks), so its queue items form one contiguous, slot-ordered
// run in every queue and the per-block queue offsets locate that run. Each job has its own
// scene, materials, envmap, camera, accumulation buffer, sample cursor and sample budget.
constexpr uint32_t kJobBlock = 256;          // slot alignment of a job (the logic block)
constexpr uint32_t kDeadPixel = 0xFFFFFFFFu;  // pixel of a slot whose job has started all its samples

struct JobDesc {
    TraceScene sc;
    DevMaterials m;
    DevEnvmap env;
    Camera cam;
    float* accum = nullptr;               // 4 floats per pixel of this job
    unsigned long long next_pixel = 0;    // samples started (new_path cursor)
    unsigned long long budget = 0;        // samples to start; 0 = unlimited
    unsigned long long seed = 0;          // RNG seed of the job's first slot
    uint32_t first_block = 0;
    uint32_t nblocks = 0;
    GuideFrame gframe = {{0.0f, 0.0f, 0.0f}, {1.0f, 1.0f, 1.0f}, 0u};  // guide inputs of the job's vertices
};

struct DevJobs {
    JobDesc* jobs = nullptr;               // device table
    const uint16_t* job_of_block = nullptr;
    uint32_t njobs = 0;
};

hipError_t wf_initialize(const DevState& st, const WFParams& p, uint64_t seed_offset, uint64_t next_pixel,
                         uint32_t npixels, hipStream_t s);
hipError_t wf_initialize_jobs(const DevState& st, const WFParams& p, const DevJobs& jobs, hipStream_t s);
// One bounce. persistent_blocks: grid of the queue-driven kernels; trace_block: workgroup
// size of the extension ray cast (64, 128 or 256; 0 = persistent waves with lane refill). jobs: null for a single render (scene,
// materials, envmap and camera as given), else the batch's job table (those are ignored).
// marks (optional): kStageMarks events recorded between the stages of the bounce, in the order
// of rgpu_wavefront_time_stages.
constexpr int kStageMarks = 10;
hipError_t wf_step(const DevState& st, const WFParams& p, const TraceScene& sc, const DevMaterials& m,
                   const DevEnvmap& env, const Camera& cam, const DevGuide& g, int variant, int persistent_blocks,
                   int trace_block, hipStream_t s, const DevJobs* jobs = nullptr, hipEvent_t* marks = nullptr);
// Path-guiding network (guide.hip).
hipError_t guide_infer(const DevState& st, const WFParams& p, const DevGuide& g, const DevJobs& jobs,
                       hipStream_t s);
hipError_t guide_train_batch(const DevGuide& g, uint32_t first, uint32_t rows, float lr, uint32_t step,
                             hipStream_t s);
hipError_t guide_refresh(const DevGuide& g, hipStream_t s);
hipError_t guide_forward(const DevGuide& g, const GuideSample* in, uint32_t rows, float* out, hipStream_t s);
hipError_t guide_gradients(const DevGuide& g, uint32_t rows, hipStream_t s);
hipError_t wf_clear_accum(const DevState& st, uint32_t npixels, bool accum64, hipStream_t s);
hipError_t wf_tonemap(const DevState& st, uint32_t npixels, bool accum64, void* rgba8, hipStream_t s);
hipError_t wf_stat(const DevState& st, uint32_t npixels, bool accum64, unsigned long long* total, hipStream_t s);
// Intermediate image (kernels/accumlation.cl): half-float means and a 32-bit sample count per
// pixel, 12 bytes (rgpu_rgb16 in the public header).
struct RGB16 {
    uint16_t r, g, b, pad;
    uint32_t count;
};
static_assert(sizeof(RGB16) == 12, "intermediate pixel is 12 bytes");
hipError_t wf_intermediate(const DevState& st, uint32_t npixels, bool accum64, RGB16* out, hipStream_t s);
hipError_t merge_intermediate(RGB16* a, const RGB16* b, uint32_t npixels, hipStream_t s);
// Preview handoff (kernels/mutex.cl): words[0] the weak mutex, words[1] the step side's flag,
// words[2] the preview side's flag. publish refreshes `shared` from the accumulation when the
// step side takes the mutex; take copies `shared` into dst when the preview side takes it (and
// writes 1 or 0 to *took, device memory, if given). Neither waits.
hipError_t handoff_init(int* words, hipStream_t s);
hipError_t handoff_publish(const DevState& st, uint32_t npixels, bool accum64, RGB16* shared, int* words,
                           hipStream_t s);
hipError_t handoff_take(const RGB16* shared, RGB16* dst, uint32_t npixels, int* words, int* took, hipStream_t s);
hipError_t tonemap_intermediate(const RGB16* in, uint32_t npixels, void* rgba8, hipStream_t s);
// Flat-batch environment sampling: states (4 words per sample) advance in place; normals (3 floats
// per sample) select the six-axis tables, null the uniform-weight one; out holds 4 floats per
// sample (direction, pdf). envmap_pdf evaluates the same densities for given directions.
hipError_t envmap_sample(const DevEnvmap& e, uint32_t* states, const float* normals, uint32_t n, float* out,
                         hipStream_t s);
hipError_t envmap_pdf(const DevEnvmap& e, const float* dirs, const float* normals, uint32_t n, float* pdfs,
                      hipStream_t s);

}  // namespace rgpu
