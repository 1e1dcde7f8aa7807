// Internal interface between the C ABI and the traversal kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rgpu_types.h"

namespace rgpu {

// 32-byte ray: origin, far clip, direction (rgpu_ray in the public header).
struct alignas(16) TraceRay {
    float ro[3];
    float tmax;
    float rd[3];
    uint32_t pad;
};
// 16-byte hit: primitive id (-1 on a miss), t (the far clip on a miss), u, v.
struct alignas(16) TraceHit {
    int32_t prim;
    float t;
    float u;
    float v;
};

// kVariantThreaded: the multi-threaded BVH (a link table per ray-direction class);
// kVariantSingleThreaded: the single-table threaded BVH. Both walk the same box records.
enum { kVariantStackless = 0, kVariantThreaded = 1, kVariantSingleThreaded = 2 };

// Device-resident scene as the kernels see it.
struct TraceScene {
    const SLNode* sl_nodes = nullptr;
    const MTNode* mt_nodes = nullptr;  // kMTOrders tables of nnodes records
    const PrimRec* prims = nullptr;
    const float4* points = nullptr;    // original vertices (for geometric normals)
    const uint32_t* indices = nullptr;  // 3 per triangle
    const float4* spheres = nullptr;
    uint32_t nnodes = 0;
    int32_t sphere_begin = 0;
    uint32_t boxes_finite = 0;  // every node box finite (the slab test may drop its NaN guard)
};

// Launch shape of the persistent ray-cast kernel: CUs to fill, resident waves per SIMD (grid =
// cu_count * waves_per_simd workgroups of 256), the idle-lane count that triggers a refill and
// the count of lanes waiting at leaves that triggers their primitive tests. Zeros take the
// defaults (256 CUs, 8 waves; refill at 24 lanes for the stackless walk, 48 for the threaded
// ones; leaves at 8 waiting lanes for the stackless walk, 6 for the threaded ones), measured on
// 4M incoherent rays over 1.3M triangles.
struct TraceTuning {
    uint32_t cu_count = 0;
    uint32_t waves_per_simd = 0;
    uint32_t refill = 0;
    uint32_t leaf_batch = 0;
    uint32_t treelet = 0;  // stackless records staged in LDS per workgroup; kTreeletOff: none
};
constexpr uint32_t kDefaultLeafBatch = 6;
constexpr uint32_t kDefaultLeafBatchStackless = 8;
// One thread per ray (coherent rays, the wavefront's per-slot trace): leaves wait until as many
// lanes wait as walk on.
constexpr uint32_t kPerRayLeafBatch = 64;
// The scene packer numbers this many top records breadth-first, so that any LDS treelet of up
// to this size is a prefix of the node array (64 KB of LDS).
constexpr uint32_t kTreeletMaxNodes = 1024;
constexpr uint32_t kTreeletOff = 0xFFFFFFFFu;
constexpr uint32_t kDefaultTreelet = 320;

// counters: null for the production kernels; otherwise 5 device u64 totals
// {branch visits, leaf visits, parent steps, primitive tests, threaded node visits}.
// block 0 without counters: the persistent kernel; 64/128/256: one thread per ray.
hipError_t launch_trace(const TraceScene& sc, int variant, const TraceRay* rays, TraceHit* hits, uint32_t n,
                        uint32_t block, hipStream_t stream, unsigned long long* counters = nullptr,
                        const TraceTuning* tune = nullptr);

}  // namespace rgpu
