// The device scene behind the opaque rgpu_scene handle of the C ABI.
#pragma once
#include <stdint.h>

#include "rgpu_ctx.h"
#include "rgpu_traverse.h"

struct rgpu_scene {
    rgpu_ctx* ctx = nullptr;
    rgpu::TraceScene dev;
    void* sl = nullptr;
    void* mt = nullptr;
    void* prims = nullptr;
    void* points = nullptr;
    void* indices = nullptr;
    void* spheres = nullptr;
    uint32_t nprims = 0;  // triangles + spheres
    uint64_t fingerprint = 0;  // of the packed node and primitive records (checkpoints compare it)
    uint64_t bytes[4] = {0, 0, 0, 0};
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {1.0f, 1.0f, 1.0f};  // root box of the BVH
};
