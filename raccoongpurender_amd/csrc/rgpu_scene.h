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
    uint64_t bytes[4] = {0, 0, 0, 0};
};
