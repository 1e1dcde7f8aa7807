// Extension ray casting on gfx950: stackless (parent-pointer) and threaded
// (skip-pointer, six direction tables) BVH2 traversal.
//
// Reference behaviour (kernels/extension_ray_cast_stackless.cl:81-246,
// kernels/extension_ray_cast.cl:44-94, kernels/slab.cl, kernels/intersect_triangle.cl):
//  - closest hit in [0, tmax] with the incoming tmax as far clip;
//  - Moller-Trumbore with |det| < 1e-8 rejected, u,v in [0,1], t == tmin accepted
//    (a later primitive wins an exact tie), t == 0 accepted;
//  - NaN-safe slabs; near child first, exact-equality tie rule on the entry distance.
// Every float expression is evaluated in the same order as the CPU oracle and the
// library is built with -ffp-contract=off, so hit index, t, u and v are bit-identical
// to the oracle on the same tree.
//
// Layout (rgpu_types.h): stackless records are 64 B (a leaf or a backtrack step reads
// only the last 16 B), threaded records are 32 B in one pre-ordered table per ray
// direction class, primitives are 48 B records in leaf order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rgpu_traverse.h"
#include "rgpu_types.h"
#include "trace_device.h"

namespace rgpu {

__device__ __forceinline__ void load_ray(const TraceRay* rays, uint32_t i, F3* ro, F3* rd, float* tmax) {
    const float4* r = reinterpret_cast<const float4*>(rays + i);
    float4 a = r[0];
    float4 b = r[1];
    *ro = f3(a.x, a.y, a.z);
    *tmax = a.w;
    *rd = f3(b.x, b.y, b.z);
}

__device__ __forceinline__ void store_hit(TraceHit* hits, uint32_t i, int32_t prim, float t, float u, float v) {
    int4 h;
    h.x = prim;
    h.y = __float_as_int(t);
    h.z = __float_as_int(prim >= 0 ? u : 0.0f);
    h.w = __float_as_int(prim >= 0 ? v : 0.0f);
    *reinterpret_cast<int4*>(hits + i) = h;
}

// Sum a per-lane count over the wavefront and add it to a global 64-bit counter.
__device__ __forceinline__ void wave_add(unsigned long long* dst, uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(dst, (unsigned long long)v);
}

__device__ __forceinline__ void flush_counts(unsigned long long* counters, const Cnt& c) {
    wave_add(counters + 0, c.branch);
    wave_add(counters + 1, c.leaf);
    wave_add(counters + 2, c.parent);
    wave_add(counters + 3, c.prim);
    wave_add(counters + 4, c.node);
}

template <bool kCount>
__global__ void __launch_bounds__(256) k_trace_stackless(const TraceRay* __restrict__ rays,
                                                         TraceHit* __restrict__ hits, uint32_t n,
                                                         const SLNode* __restrict__ nodes,
                                                         const PrimRec* __restrict__ prims,
                                                         int32_t sphere_begin, unsigned long long* counters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Cnt c;
    if (i < n) {
        F3 ro, rd;
        float t;
        load_ray(rays, i, &ro, &rd, &t);
        float u, v;
        int32_t prim = trace_stackless<kCount>(nodes, prims, sphere_begin, ro, rd, &t, &u, &v, &c);
        store_hit(hits, i, prim, t, u, v);
    }
    if (kCount) flush_counts(counters, c);
}

template <bool kCount>
__global__ void __launch_bounds__(256) k_trace_threaded(const TraceRay* __restrict__ rays,
                                                        TraceHit* __restrict__ hits, uint32_t n,
                                                        const MTNode* __restrict__ tables, uint32_t nnodes,
                                                        const PrimRec* __restrict__ prims,
                                                        int32_t sphere_begin, unsigned long long* counters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Cnt c;
    if (i < n) {
        F3 ro, rd;
        float t;
        load_ray(rays, i, &ro, &rd, &t);
        float u, v;
        int32_t prim = trace_threaded<kCount>(tables, nnodes, prims, sphere_begin, ro, rd, &t, &u, &v, &c);
        store_hit(hits, i, prim, t, u, v);
    }
    if (kCount) flush_counts(counters, c);
}

// Empty scene: every ray misses and keeps its far clip.
__global__ void __launch_bounds__(256) k_trace_empty(const TraceRay* __restrict__ rays, TraceHit* __restrict__ hits,
                                                     uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 ro, rd;
    float t;
    load_ray(rays, i, &ro, &rd, &t);
    store_hit(hits, i, -1, t, 0.0f, 0.0f);
}

hipError_t launch_trace(const TraceScene& sc, int variant, const TraceRay* rays, TraceHit* hits, uint32_t n,
                        uint32_t block, hipStream_t stream, unsigned long long* counters) {
    if (n == 0) return hipSuccess;
    if (block == 0) block = 256;
    dim3 grid((n + block - 1) / block);
    if (sc.nnodes == 0) {
        hipLaunchKernelGGL(k_trace_empty, grid, dim3(block), 0, stream, rays, hits, n);
    } else if (variant == kVariantThreaded) {
        if (counters)
            hipLaunchKernelGGL(k_trace_threaded<true>, grid, dim3(block), 0, stream, rays, hits, n, sc.mt_nodes,
                               sc.nnodes, sc.prims, sc.sphere_begin, counters);
        else
            hipLaunchKernelGGL(k_trace_threaded<false>, grid, dim3(block), 0, stream, rays, hits, n, sc.mt_nodes,
                               sc.nnodes, sc.prims, sc.sphere_begin, counters);
    } else {
        if (counters)
            hipLaunchKernelGGL(k_trace_stackless<true>, grid, dim3(block), 0, stream, rays, hits, n, sc.sl_nodes,
                               sc.prims, sc.sphere_begin, counters);
        else
            hipLaunchKernelGGL(k_trace_stackless<false>, grid, dim3(block), 0, stream, rays, hits, n, sc.sl_nodes,
                               sc.prims, sc.sphere_begin, counters);
    }
    return hipGetLastError();
}

}  // namespace rgpu
