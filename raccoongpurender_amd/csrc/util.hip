// Small utility kernels: a plain vectorised device copy (the reference's copy_if_locked,
// kernels/mutex.cl:19-24, without the lock), also used to calibrate HBM byte counters.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rgpu {

__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}

__global__ void __launch_bounds__(256) k_copy1(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream) {
    if (bytes == 0) return hipSuccess;
    size_t n16 = bytes / 16;
    if (((uintptr_t)dst | (uintptr_t)src) % 16 == 0 && n16 > 0) {
        size_t blocks = (n16 + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, stream, (const uint4*)src, (uint4*)dst, n16);
        size_t tail = bytes - n16 * 16;
        if (tail)
            hipLaunchKernelGGL(k_copy1, dim3(1), dim3(256), 0, stream, (const uint8_t*)src + n16 * 16,
                               (uint8_t*)dst + n16 * 16, tail);
    } else {
        size_t blocks = (bytes + 255) / 256;
        hipLaunchKernelGGL(k_copy1, dim3((unsigned)blocks), dim3(256), 0, stream, (const uint8_t*)src, (uint8_t*)dst,
                           bytes);
    }
    return hipGetLastError();
}

}  // namespace rgpu
