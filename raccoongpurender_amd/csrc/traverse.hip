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

#include <type_traits>

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

// One thread per ray, in three forms of the same walk (same visits, same results): plain, the
// walk as one loop (best for coherent rays, whose lanes reach leaves together anyway); count,
// the same with event counters; batched, leaves batched across the wave (better for incoherent
// rays, chosen by a leaf_batch above 1).
enum { kWalkPlain = 0, kWalkCount = 1, kWalkBatched = 2 };

template <int kMode>
__global__ void __launch_bounds__(256) k_trace_stackless(const TraceRay* __restrict__ rays,
                                                         TraceHit* __restrict__ hits, uint32_t n, TraceScene sc,
                                                         uint32_t leaf_batch, unsigned long long* counters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Cnt c;
    if (i < n) {
        F3 ro, rd;
        float t;
        load_ray(rays, i, &ro, &rd, &t);
        float u, v;
        int32_t prim = kMode == kWalkBatched ? trace_batched<kVariantStackless>(sc, ro, rd, &t, &u, &v, leaf_batch)
                                         : trace_stackless<kMode == kWalkCount>(sc.sl_nodes, sc.prims, sc.sphere_begin,
                                                                             sc.boxes_finite != 0, ro, rd, &t, &u,
                                                                             &v, &c);
        store_hit(hits, i, prim, t, u, v);
    }
    if (kMode == kWalkCount) flush_counts(counters, c);
}

// kVariant: kVariantThreaded (a table per direction class) or kVariantSingleThreaded.
template <int kMode, int kVariant>
__global__ void __launch_bounds__(256) k_trace_threaded(const TraceRay* __restrict__ rays,
                                                        TraceHit* __restrict__ hits, uint32_t n, TraceScene sc,
                                                        uint32_t leaf_batch, unsigned long long* counters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Cnt c;
    if (i < n) {
        F3 ro, rd;
        float t;
        load_ray(rays, i, &ro, &rd, &t);
        float u, v;
        int32_t prim = kMode == kWalkBatched
                           ? trace_batched<kVariant>(sc, ro, rd, &t, &u, &v, leaf_batch)
                           : trace_threaded<kMode == kWalkCount, kVariant>(sc.mt_nodes, sc.nnodes, sc.prims,
                                                                           sc.sphere_begin, sc.boxes_finite != 0, ro,
                                                                           rd, &t, &u, &v, &c);
        store_hit(hits, i, prim, t, u, v);
    }
    if (kMode == kWalkCount) flush_counts(counters, c);
}

// Persistent wave64 traversal: wave w owns rays [w * chunk, (w + 1) * chunk) and hands them to
// its lanes as lanes finish, refilling whenever at least `refill` lanes are idle (all at once
// when the wave is idle), so that lanes do not wait for the slowest ray of the wave. No
// atomics: a wave's cursor is wave-uniform.
// Leaves are batched: a lane that enters a leaf stops there, and the wave tests primitives only
// once `leaf_batch` lanes wait at leaves or at least as many lanes wait as walk on, so that the
// primitive tests run with many lanes at once instead of taxing every node visit of the wave. Each ray still
// visits its nodes and tests its primitives in the order of the one-ray walk.
// Stackless: the first `top` node records (the top of the tree, numbered breadth-first) are
// copied to LDS at the start (top * 64 bytes of dynamic shared memory) and read from there.
template <int kVariant>
__global__ void __launch_bounds__(256) k_trace_persistent(const TraceRay* __restrict__ rays,
                                                          TraceHit* __restrict__ hits, uint32_t n, TraceScene sc,
                                                          uint32_t chunk, uint32_t refill, uint32_t leaf_batch,
                                                          uint32_t top) {
    const SLTreelet src{reinterpret_cast<const float4*>(sc.sl_nodes), top};
    if constexpr (kVariant == kVariantStackless) {
        for (uint32_t j = threadIdx.x; j < 4 * top; j += blockDim.x) sl_treelet[j] = src.g[j];
        __syncthreads();
    }
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    uint64_t next = (uint64_t)wave * chunk;
    const uint64_t end = next + chunk < (uint64_t)n ? next + chunk : (uint64_t)n;
    uint32_t ray = kInvalid;
    using State = typename std::conditional<kVariant == kVariantStackless, StacklessRay, ThreadedRay>::type;
    State st;
    st.node = kInvalid;
    if constexpr (kVariant != kVariantStackless) st.leaf = 0;
    const bool finite = sc.boxes_finite != 0;
    for (;;) {
        const bool idle = !walk_live(st);
        const unsigned long long idle_mask = __ballot(idle);
        const uint32_t nidle = (uint32_t)__popcll(idle_mask);
        if (nidle >= refill || nidle == 64u) {
            if (idle) {
                const uint64_t r = next + (uint64_t)__popcll(idle_mask & below);
                if (r < end) {
                    ray = (uint32_t)r;
                    F3 ro, rd;
                    float tmax;
                    load_ray(rays, ray, &ro, &rd, &tmax);
                    if constexpr (kVariant != kVariantStackless)
                        threaded_begin<kVariant>(st, sc.mt_nodes, sc.nnodes, ro, rd, tmax);
                    else
                        stackless_begin(st, ro, rd, tmax);
                }
            }
            next += nidle;
        }
        if (__ballot(walk_live(st)) == 0) break;
        const bool pending = walk_pending(st);
        const bool walking = walk_live(st) && !pending;
        const uint32_t npending = (uint32_t)__popcll(__ballot(pending));
        const uint32_t nwalk = (uint32_t)__popcll(__ballot(walking));
        const bool leaf_phase = npending >= leaf_batch || npending >= nwalk;
        if (leaf_phase ? pending : walking) {
            if (leaf_phase) {
                if constexpr (kVariant != kVariantStackless)
                    threaded_leaf(st, sc.prims, sc.sphere_begin);
                else
                    stackless_leaf(st, src, sc.prims, sc.sphere_begin);
            } else {
                if constexpr (kVariant != kVariantStackless)
                    threaded_visit(st, finite);
                else
                    stackless_visit(st, src, finite);
            }
            if (!walk_live(st)) store_hit(hits, ray, st.prim, st.tmin, st.u, st.v);
        }
    }
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

template <int kVariant>
void launch_threaded(int mode, dim3 grid, uint32_t block, hipStream_t stream, const TraceRay* rays, TraceHit* hits,
                     uint32_t n, const TraceScene& sc, uint32_t leaf_batch, unsigned long long* counters) {
    if (mode == kWalkCount)
        hipLaunchKernelGGL((k_trace_threaded<kWalkCount, kVariant>), grid, dim3(block), 0, stream, rays, hits, n, sc,
                           leaf_batch, counters);
    else if (mode == kWalkBatched)
        hipLaunchKernelGGL((k_trace_threaded<kWalkBatched, kVariant>), grid, dim3(block), 0, stream, rays, hits, n, sc,
                           leaf_batch, counters);
    else
        hipLaunchKernelGGL((k_trace_threaded<kWalkPlain, kVariant>), grid, dim3(block), 0, stream, rays, hits, n, sc,
                           leaf_batch, counters);
}

hipError_t launch_trace(const TraceScene& sc, int variant, const TraceRay* rays, TraceHit* hits, uint32_t n,
                        uint32_t block, hipStream_t stream, unsigned long long* counters, const TraceTuning* tune) {
    if (n == 0) return hipSuccess;
    if (block == 0 && !counters && sc.nnodes != 0) {
        // persistent kernel: as many resident waves as fit, each with a chunk of the rays
        const uint32_t cus = tune && tune->cu_count ? tune->cu_count : 256u;
        const uint32_t wps = tune && tune->waves_per_simd ? tune->waves_per_simd : 8u;
        // measured on the 1.3M-triangle incoherent set: refills at 24 (stackless) or 48 (threaded)
        // idle lanes, primitive tests once 8 (stackless) or 6 (threaded) lanes wait at leaves
        const uint32_t refill =
            tune && tune->refill ? tune->refill : (variant != kVariantStackless ? 48u : 24u);
        const uint32_t leaf_batch = tune && tune->leaf_batch ? tune->leaf_batch
                                    : variant == kVariantStackless ? kDefaultLeafBatchStackless
                                                                   : kDefaultLeafBatch;
        const uint32_t blocks = cus * wps;  // 4 waves (one per SIMD) per 256-thread block
        const uint32_t waves = blocks * 4u;
        const uint32_t chunk = (n + waves - 1) / waves;
        if (variant == kVariantThreaded) {
            hipLaunchKernelGGL(k_trace_persistent<kVariantThreaded>, dim3(blocks), dim3(256), 0, stream, rays, hits,
                               n, sc, chunk, refill, leaf_batch, 0u);
        } else if (variant == kVariantSingleThreaded) {
            hipLaunchKernelGGL(k_trace_persistent<kVariantSingleThreaded>, dim3(blocks), dim3(256), 0, stream, rays,
                               hits, n, sc, chunk, refill, leaf_batch, 0u);
        } else {
            // the top of the tree in LDS: 320 records (20 KB) let the 8 workgroups per CU of the
            // default launch share the CU's 160 KB
            uint32_t top = tune && tune->treelet ? tune->treelet : kDefaultTreelet;
            if (top == kTreeletOff) top = 0;
            top = top < sc.nnodes ? top : sc.nnodes;
            top = top < kTreeletMaxNodes ? top : kTreeletMaxNodes;
            hipLaunchKernelGGL(k_trace_persistent<kVariantStackless>, dim3(blocks), dim3(256), (size_t)top * 64u,
                               stream, rays, hits, n, sc, chunk, refill, leaf_batch, top);
        }
        return hipGetLastError();
    }
    if (block == 0) block = 256;
    dim3 grid((n + block - 1) / block);
    // one thread per ray: leaves batched by default only once as many lanes wait at leaves as
    // walk on (leaf_batch 64; leaf_batch 1 asks for the plain walk)
    const uint32_t leaf_batch = tune && tune->leaf_batch ? tune->leaf_batch : kPerRayLeafBatch;
    const int mode = counters ? kWalkCount : (leaf_batch > 1 ? kWalkBatched : kWalkPlain);
    if (sc.nnodes == 0) {
        hipLaunchKernelGGL(k_trace_empty, grid, dim3(block), 0, stream, rays, hits, n);
    } else if (variant == kVariantThreaded) {
        launch_threaded<kVariantThreaded>(mode, grid, block, stream, rays, hits, n, sc, leaf_batch, counters);
    } else if (variant == kVariantSingleThreaded) {
        launch_threaded<kVariantSingleThreaded>(mode, grid, block, stream, rays, hits, n, sc, leaf_batch, counters);
    } else {
        if (mode == kWalkCount)
            hipLaunchKernelGGL(k_trace_stackless<kWalkCount>, grid, dim3(block), 0, stream, rays, hits, n, sc,
                               leaf_batch, counters);
        else if (mode == kWalkBatched)
            hipLaunchKernelGGL(k_trace_stackless<kWalkBatched>, grid, dim3(block), 0, stream, rays, hits, n, sc,
                               leaf_batch, counters);
        else
            hipLaunchKernelGGL(k_trace_stackless<kWalkPlain>, grid, dim3(block), 0, stream, rays, hits, n, sc,
                               leaf_batch, counters);
    }
    return hipGetLastError();
}

}  // namespace rgpu
