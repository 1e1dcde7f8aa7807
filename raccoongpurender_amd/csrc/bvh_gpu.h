// Device BVH build (bvh_gpu.hip): the host builder's tree, built level by level on the GPU.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "bvh_build.h"

namespace rgpu {

struct BuildTimes {
    double device_ms = 0.0;  // first primitive box to last level on the stream (uploads excluded)
};

struct DeviceBuildError : std::runtime_error {
    hipError_t err;
    DeviceBuildError(hipError_t e, const char* what)
        : std::runtime_error(std::string(what) + ": " + hipGetErrorString(e)), err(e) {}
};

// Same inputs and options as build_bvh (bins limited to 2..32, finite coordinates only); the
// tree equals build_bvh's node for node, with its primitive order. Runs on stream s of the
// current device; throws std::invalid_argument, std::bad_alloc or DeviceBuildError.
void build_bvh_device(const float* points, uint32_t stride, uint32_t npoints, const uint32_t* indices,
                      uint32_t ntris, const float* spheres, uint32_t nspheres, const BuildOptions& opt,
                      hipStream_t s, BVHTree* out, BuildTimes* times);

}  // namespace rgpu
