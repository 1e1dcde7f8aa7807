// Device BVH build: the host builder's binned-SAH top-down split (bvh_build.cpp), one tree
// level per round of kernels, producing the same tree.
//
// Every quantity the split decisions read is order-independent: node and centroid boxes are
// min/max reductions (atomics on an order-preserving integer encoding of the floats), bin counts
// are sums, and the SAH search uses the host builder's exact expressions (float centroids and
// bin indices, double areas and costs), by one wave per task (prefix and suffix scans, then the
// first minimum in the host's axis and bin order) or one lane per tiny task. Primitives move by a
// stable partition (a prefix sum over the level's flags), as the host builder's
// std::stable_partition does. Nodes are numbered level by level on the device and renumbered
// into the host builder's depth-first pre-order on the host; the leaves' primitive order is then
// the order of the final id array, as it is on the host.
//
// Replaces Embree's build in SceneManager::buildBVH (common/scene_manager.hpp:517-577).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "bvh_build.h"
#include "bvh_gpu.h"

namespace rgpu {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kMaxBins = 32;
// Tasks above this many primitives are bounded and binned by the position-parallel kernels (bins
// in device memory at the task's big-task rank); the others by one wave each (k_small_tasks),
// which also decides their split, except those of at most kTinyTask primitives, which take one
// lane each (k_tiny_tasks).
constexpr uint32_t kSmallTask = 2048;
constexpr uint32_t kTinyTask = 8;

struct PBox {
    float lo[3], hi[3];
};

// Order-preserving map of finite floats onto uint32 (a < b  <=>  enc(a) < enc(b)).
__device__ __forceinline__ uint32_t fenc(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fdec(uint32_t e) {
    return __uint_as_float((e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e);
}
constexpr uint32_t kEncMin = 0x00000000u;  // below every encoded float: start of a max
constexpr uint32_t kEncMax = 0xFFFFFFFFu;  // above every encoded float: start of a min

struct Task {
    uint32_t beg, end, node, depth;
};

// What a level decides for a task.
enum : uint32_t { kLeaf = 0, kSplit = 1, kMedian = 2 };
struct Decision {
    uint32_t kind, axis, bin, mid;  // mid: first position of the right child (split and median)
};

struct Params {
    uint32_t n;          // primitives
    uint32_t bins;       // B
    uint32_t max_leaf;
    uint32_t max_depth;
    double traversal_cost;
};

// ------------------------------------------------------------------ primitive boxes
__global__ void k_prim_boxes(const float* __restrict__ pts, uint32_t stride, const uint32_t* __restrict__ idx,
                             uint32_t ntris, const float* __restrict__ sph, uint32_t nsph, PBox* __restrict__ pb,
                             float* __restrict__ cen, uint32_t* __restrict__ ids) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntris + nsph) return;
    PBox b;
    if (i < ntris) {
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = FLT_MAX;
            b.hi[k] = -FLT_MAX;
        }
        for (int j = 0; j < 3; ++j) {
            const float* p = pts + (size_t)idx[3 * (size_t)i + j] * stride;
            for (int k = 0; k < 3; ++k) {
                b.lo[k] = fminf(b.lo[k], p[k]);
                b.hi[k] = fmaxf(b.hi[k], p[k]);
            }
        }
    } else {
        const float* s = sph + 4 * (size_t)(i - ntris);
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = s[k] - s[3];
            b.hi[k] = s[k] + s[3];
        }
    }
    for (int k = 0; k < 3; ++k) {
        b.lo[k] += 0.0f;  // -0 as +0, as the host builder does
        b.hi[k] += 0.0f;
        cen[3 * (size_t)i + k] = 0.5f * (b.lo[k] + b.hi[k]);
    }
    pb[i] = b;
    ids[i] = i;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Task bounds: 12 encoded words per task (box lo, box hi, centroid lo, centroid hi); the lows
// start at kEncMax, the highs at kEncMin.
// Also lists the big tasks by their slot (bslot: the exclusive scan of the big-task flags), so
// that k_split runs one wave per big task.
__global__ void k_clear_bounds(uint32_t* __restrict__ tb, uint32_t ntasks, const Task* __restrict__ tasks,
                               const uint32_t* __restrict__ bslot, uint32_t* __restrict__ big_task) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 12 * ntasks) return;
    const uint32_t w = i % 12;
    tb[i] = (w < 3 || (w >= 6 && w < 9)) ? kEncMax : kEncMin;
    if (i < ntasks && tasks[i].end - tasks[i].beg > kSmallTask) big_task[bslot[i]] = i;
}
// Bins: per task, 3 axes x B bins x 7 words (box lo, box hi, count).
__global__ void k_clear_bins(uint32_t* __restrict__ bins, size_t words) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= words) return;
    const uint32_t w = (uint32_t)(i % 7);
    bins[i] = w < 3 ? kEncMax : (w < 6 ? kEncMin : 0u);
}

// The task of every position (kNone: a finished leaf); tasks are sorted by their first position.
// pos_big: the same for the tasks of more than kSmallTask primitives, kNone elsewhere.
__global__ void k_task_of(const Task* __restrict__ tasks, uint32_t ntasks, uint32_t n, uint32_t* __restrict__ pos_task,
                          uint32_t* __restrict__ pos_big) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t lo = 0, hi = ntasks;  // first task with beg > i
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if (tasks[m].beg <= i) lo = m + 1;
        else hi = m;
    }
    uint32_t t = kNone, big = kNone;
    if (lo > 0 && i < tasks[lo - 1].end) {
        t = lo - 1;
        if (tasks[t].end - tasks[t].beg > kSmallTask) big = t;
    }
    pos_task[i] = t;
    pos_big[i] = big;
}

// Min (kMin) or max of v over the wave64.
template <bool kMin>
__device__ __forceinline__ uint32_t wave_red(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(v, off, 64);
        v = kMin ? min(v, o) : max(v, o);
    }
    return v;
}

// A workgroup of k_bounds / k_bin covers kSubBlocks runs of 256 positions. A run inside one
// task is reduced in the workgroup and added to a running partial, which goes to device memory
// with one atomic per word when the task changes and at the end, so the top levels (few tasks,
// every workgroup on the same words) take 1/kSubBlocks of the global atomics. In a run across
// task ends, or with positions of small tasks, k_bounds reduces each task segment of a wave
// and takes one atomic per segment and word, and k_bin accumulates the run's tasks in its LDS
// bins one after the other.
constexpr uint32_t kSubBlocks = 8;

__device__ __forceinline__ void flush_bounds(const uint32_t* acc, uint32_t t, uint32_t (*part)[12], uint32_t* tb) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int w = 0; w < 12; ++w) {
        const bool mn = w < 3 || (w >= 6 && w < 9);
        const uint32_t r = mn ? wave_red<true>(acc[w]) : wave_red<false>(acc[w]);
        if (lane == 0) part[wave][w] = r;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        const uint32_t w = threadIdx.x;
        const bool mn = w < 3 || (w >= 6 && w < 9);
        uint32_t r = part[0][w];
        for (int q = 1; q < 4; ++q) r = mn ? min(r, part[q][w]) : max(r, part[q][w]);
        if (mn) atomicMin(tb + 12 * (size_t)t + w, r);
        else atomicMax(tb + 12 * (size_t)t + w, r);
    }
    __syncthreads();  // part is written again by the next flush
}

__global__ void __launch_bounds__(256) k_bounds(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ pos_task,
                                                uint32_t n, const PBox* __restrict__ pb, const float* __restrict__ cen,
                                                uint32_t* __restrict__ tb) {
    __shared__ uint32_t part[4][12];
    uint32_t acc[12];
    uint32_t acc_t = kNone;  // the task of the running partial (the same in every thread)
    for (uint32_t j = 0; j < kSubBlocks; ++j) {
        const uint32_t first = (blockIdx.x * kSubBlocks + j) * blockDim.x;
        if (first >= n) break;
        const uint32_t i = first + threadIdx.x;
        const uint32_t last = min(first + blockDim.x, n) - 1;
        const uint32_t t0 = pos_task[first], t1 = pos_task[last];
        const bool uniform = t0 != kNone && t0 == t1;  // task segments are contiguous
        uint32_t v[12];
        const uint32_t t = i < n ? pos_task[i] : kNone;
        if (t != kNone) {
            const uint32_t p = ids[i];
            const PBox b = pb[p];
            for (int k = 0; k < 3; ++k) {
                v[k] = fenc(b.lo[k]);
                v[3 + k] = fenc(b.hi[k]);
                const uint32_t c = fenc(cen[3 * (size_t)p + k]);
                v[6 + k] = c;
                v[9 + k] = c;
            }
        } else {
            for (int k = 0; k < 3; ++k) {
                v[k] = kEncMax;
                v[3 + k] = kEncMin;
                v[6 + k] = kEncMax;
                v[9 + k] = kEncMin;
            }
        }
        if (uniform) {
            if (t0 != acc_t) {
                if (acc_t != kNone) flush_bounds(acc, acc_t, part, tb);
                for (int k = 0; k < 3; ++k) {
                    acc[k] = kEncMax;
                    acc[3 + k] = kEncMin;
                    acc[6 + k] = kEncMax;
                    acc[9 + k] = kEncMin;
                }
                acc_t = t0;
            }
            for (int w = 0; w < 12; ++w) {
                const bool mn = w < 3 || (w >= 6 && w < 9);
                acc[w] = mn ? min(acc[w], v[w]) : max(acc[w], v[w]);
            }
        } else {
            // A task's positions are contiguous, so a wave of this run holds a few task segments:
            // each segment is reduced across its lanes (an inclusive scan whose steps combine
            // only lanes of the same task), and its last lane alone takes the atomics. A wave with
            // no position of a big task has nothing to add (most runs of the deep levels).
            if (__ballot(t != kNone) == 0) continue;
            const uint32_t lane = threadIdx.x & 63;
            bool same[6];
            for (int k = 0; k < 6; ++k) {
                const uint32_t o = __shfl_up(t, 1u << k, 64);
                same[k] = lane >= (1u << k) && o == t;
            }
            for (int w = 0; w < 12; ++w) {
                const bool mn = w < 3 || (w >= 6 && w < 9);
                uint32_t x = v[w];
                for (int k = 0; k < 6; ++k) {
                    const uint32_t o = __shfl_up(x, 1u << k, 64);
                    if (same[k]) x = mn ? min(x, o) : max(x, o);
                }
                v[w] = x;
            }
            const uint32_t tn = __shfl_down(t, 1u, 64);
            if (t != kNone && (lane == 63 || tn != t)) {
                for (int w = 0; w < 12; ++w) {
                    const bool mn = w < 3 || (w >= 6 && w < 9);
                    if (mn) atomicMin(tb + 12 * (size_t)t + w, v[w]);
                    else atomicMax(tb + 12 * (size_t)t + w, v[w]);
                }
            }
        }
    }
    if (acc_t != kNone) flush_bounds(acc, acc_t, part, tb);
}

// Bin of centroid coordinate c on an axis whose centroid range starts at clo with
// scale = B / extent: the host builder's expression.
__device__ __forceinline__ uint32_t bin_of(float c, float clo, float scale, uint32_t B) {
    int b = (int)((c - clo) * scale);
    b = min(max(b, 0), (int)B - 1);
    return (uint32_t)b;
}

__device__ __forceinline__ void init_bins(uint32_t* lb, uint32_t words) {
    for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) {
        const uint32_t r = w % 7;
        lb[w] = r < 3 ? kEncMax : (r < 6 ? kEncMin : 0u);
    }
}

__device__ __forceinline__ void flush_bins(const uint32_t* lb, uint32_t* g, uint32_t words) {
    for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) {
        const uint32_t r = w % 7;
        const uint32_t v = lb[w];
        if (r < 3) {
            if (v != kEncMax) atomicMin(g + w, v);
        } else if (r < 6) {
            if (v != kEncMin) atomicMax(g + w, v);
        } else if (v) {
            atomicAdd(g + w, v);
        }
    }
}

// One primitive's centroid bin on each axis of its task: box min/max and count, by atomics
// on dst (LDS bins or the task's bins in device memory).
__device__ __forceinline__ void bin_primitive(uint32_t p, uint32_t t, const PBox* __restrict__ pb,
                                              const float* __restrict__ cen, const uint32_t* __restrict__ tb,
                                              uint32_t B, uint32_t* dst) {
    const PBox b = pb[p];
    for (int k = 0; k < 3; ++k) {
        const float clo = fdec(tb[12 * (size_t)t + 6 + k]), chi = fdec(tb[12 * (size_t)t + 9 + k]);
        const float ext = chi - clo;
        if (!(ext > 0.0f)) continue;
        const float scale = (float)B / ext;
        const uint32_t bi = bin_of(cen[3 * (size_t)p + k], clo, scale, B);
        uint32_t* q = dst + (k * B + bi) * 7;
        for (int a = 0; a < 3; ++a) {
            atomicMin(q + a, fenc(b.lo[a]));
            atomicMax(q + 3 + a, fenc(b.hi[a]));
        }
        atomicAdd(q + 6, 1u);
    }
}

__global__ void __launch_bounds__(256) k_bin(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ pos_task,
                                             uint32_t n, const PBox* __restrict__ pb, const float* __restrict__ cen,
                                             const uint32_t* __restrict__ tb, uint32_t B, const uint32_t* __restrict__ slot,
                                             uint32_t* __restrict__ bins) {
    __shared__ uint32_t lb[3 * kMaxBins * 7];
    const uint32_t words = 3 * B * 7;
    uint32_t lb_t = kNone;  // the task the LDS bins hold (the same in every thread)
    for (uint32_t j = 0; j < kSubBlocks; ++j) {
        const uint32_t first = (blockIdx.x * kSubBlocks + j) * blockDim.x;
        if (first >= n) break;
        const uint32_t i = first + threadIdx.x;
        const uint32_t last = min(first + blockDim.x, n) - 1;
        const uint32_t t0 = pos_task[first], t1 = pos_task[last];
        const uint32_t t = i < n ? pos_task[i] : kNone;
        // The run's first and last tasks (both kNone-free ones) are accumulated in the LDS bins in
        // turn. A task of this kernel holds more than kSmallTask >= 256 positions, so no other
        // task can lie inside a run; were one to, its primitives would go to device memory.
        for (int side = 0; side < 2; ++side) {
            const uint32_t tt = side == 0 ? t0 : t1;  // the same in every thread
            if (tt == kNone || (side == 1 && tt == t0)) continue;
            if (tt != lb_t) {
                __syncthreads();  // the LDS atomics of earlier runs have landed
                if (lb_t != kNone) {
                    flush_bins(lb, bins + (size_t)slot[lb_t] * words, words);
                    __syncthreads();
                }
                init_bins(lb, words);
                __syncthreads();
                lb_t = tt;
            }
            if (t == tt) bin_primitive(ids[i], t, pb, cen, tb, B, lb);
        }
        if (t != kNone && t != t0 && t != t1) bin_primitive(ids[i], t, pb, cen, tb, B, bins + (size_t)slot[t] * words);
    }
    if (lb_t != kNone) {
        __syncthreads();
        flush_bins(lb, bins + (size_t)slot[lb_t] * words, words);
    }
}

// Surface area of a decoded box, in double: the host builder's Box::area.
__device__ __forceinline__ double box_area(const float* lo, const float* hi) {
    const double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

__device__ __forceinline__ uint32_t levels_needed(uint32_t n, uint32_t max_leaf) {
    const uint32_t leaves = (n + max_leaf - 1) / max_leaf;
    uint32_t l = 0;
    while ((1u << l) < leaves) ++l;
    return l;
}

// Best SAH split over the bins: the first (axis, bin) in the host builder's loop order whose
// cost is below every earlier one.
struct Best {
    int axis;
    uint32_t bin;
    double cost;
};


// Wave form (every lane of the wave calls it; bins in LDS): lane b holds bin b of an axis, the
// prefix and suffix boxes and counts are wave scans (min, max and sums, exact in any order),
// and the first minimum is a wave reduction over (cost, bin).
// All three axes at once when 3 W <= 64 lanes (W = B rounded up to a power of two): lane
// k W + b holds bin b of axis k, the scans run within segments of W lanes, and the lane order
// is the host's (axis, bin) loop order, so the first minimum is the lowest lane.
template <uint32_t W>
__device__ Best wave_best_axes(const uint32_t* L, uint32_t B, const float* clo, const float* chi, uint32_t lane) {
    const uint32_t k = lane / W, b = lane % W;
    const bool on = k < 3 && b < B && chi[k < 3 ? k : 0] - clo[k < 3 ? k : 0] > 0.0f;
    float pl[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, ph[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    uint32_t pc = 0;
    if (on) {
        const uint32_t* q = L + (k * B + b) * 7;
        pc = q[6];
        if (pc)
            for (int a = 0; a < 3; ++a) {
                pl[a] = fdec(q[a]);
                ph[a] = fdec(q[3 + a]);
            }
    }
    float sl[3] = {pl[0], pl[1], pl[2]}, sh[3] = {ph[0], ph[1], ph[2]};
    uint32_t sc = pc;
    for (uint32_t off = 1; off < W; off <<= 1) {
        for (int a = 0; a < 3; ++a) {
            const float ul = __shfl_up(pl[a], off, W), uh = __shfl_up(ph[a], off, W);
            const float dl = __shfl_down(sl[a], off, W), dh = __shfl_down(sh[a], off, W);
            if (b >= off) {
                pl[a] = fminf(pl[a], ul);
                ph[a] = fmaxf(ph[a], uh);
            }
            if (b + off < W) {
                sl[a] = fminf(sl[a], dl);
                sh[a] = fmaxf(sh[a], dh);
            }
        }
        const uint32_t uc = __shfl_up(pc, off, W), dc = __shfl_down(sc, off, W);
        if (b >= off) pc += uc;
        if (b + off < W) sc += dc;
    }
    float rl[3], rh[3];
    for (int a = 0; a < 3; ++a) {
        rl[a] = __shfl_down(sl[a], 1, W);
        rh[a] = __shfl_down(sh[a], 1, W);
    }
    const uint32_t rc = __shfl_down(sc, 1, W);
    double c = DBL_MAX;
    if (on && b + 1 < B && pc != 0 && rc != 0) c = box_area(pl, ph) * pc + box_area(rl, rh) * rc;
    uint32_t idx = lane;
    for (int off = 32; off > 0; off >>= 1) {
        const double oc = __shfl_xor(c, off, 64);
        const uint32_t oi = __shfl_xor(idx, off, 64);
        if (oc < c || (oc == c && oi < idx)) {
            c = oc;
            idx = oi;
        }
    }
    if (!(c < DBL_MAX)) return Best{-1, 0, DBL_MAX};
    return Best{(int)(idx / W), idx % W, c};
}

__device__ Best wave_best(const uint32_t* L, uint32_t B, const float* clo, const float* chi, uint32_t lane) {
    if (B <= 4) return wave_best_axes<4>(L, B, clo, chi, lane);
    if (B <= 8) return wave_best_axes<8>(L, B, clo, chi, lane);
    if (B <= 16) return wave_best_axes<16>(L, B, clo, chi, lane);
    Best best{-1, 0, DBL_MAX};
    for (int k = 0; k < 3; ++k) {
        const float ext = chi[k] - clo[k];
        if (!(ext > 0.0f)) continue;
        float pl[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, ph[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        uint32_t pc = 0;
        if (lane < B) {
            const uint32_t* q = L + (k * B + lane) * 7;
            pc = q[6];
            if (pc)
                for (int a = 0; a < 3; ++a) {
                    pl[a] = fdec(q[a]);
                    ph[a] = fdec(q[3 + a]);
                }
        }
        float sl[3] = {pl[0], pl[1], pl[2]}, sh[3] = {ph[0], ph[1], ph[2]};
        uint32_t sc = pc;
        for (int off = 1; off < 64; off <<= 1) {
            for (int a = 0; a < 3; ++a) {
                const float ul = __shfl_up(pl[a], off, 64), uh = __shfl_up(ph[a], off, 64);
                const float dl = __shfl_down(sl[a], off, 64), dh = __shfl_down(sh[a], off, 64);
                if (lane >= (uint32_t)off) {
                    pl[a] = fminf(pl[a], ul);
                    ph[a] = fmaxf(ph[a], uh);
                }
                if (lane + off < 64) {
                    sl[a] = fminf(sl[a], dl);
                    sh[a] = fmaxf(sh[a], dh);
                }
            }
            const uint32_t uc = __shfl_up(pc, off, 64), dc = __shfl_down(sc, off, 64);
            if (lane >= (uint32_t)off) pc += uc;
            if (lane + off < 64) sc += dc;
        }
        // the right side of the split after bin `lane` is the suffix from lane + 1
        float rl[3], rh[3];
        for (int a = 0; a < 3; ++a) {
            rl[a] = __shfl_down(sl[a], 1, 64);
            rh[a] = __shfl_down(sh[a], 1, 64);
        }
        const uint32_t rc = __shfl_down(sc, 1, 64);
        double c = DBL_MAX;
        if (lane + 1 < B && pc != 0 && rc != 0) c = box_area(pl, ph) * pc + box_area(rl, rh) * rc;
        uint32_t b = lane;
        for (int off = 32; off > 0; off >>= 1) {
            const double oc = __shfl_xor(c, off, 64);
            const uint32_t ob = __shfl_xor(b, off, 64);
            if (oc < c || (oc == c && ob < b)) {
                c = oc;
                b = ob;
            }
        }
        if (c < best.cost) best = Best{k, b, c};
    }
    return best;
}

// The host builder's split decision for task t (Builder::build in bvh_build.cpp), from its 12
// encoded bound words tbw and the best binned split a wave found over its bins.
__device__ void decide_task(uint32_t t, const Task tk, const uint32_t* tbw, const Best& bb,
                            const Params prm, Decision* __restrict__ dec, float* __restrict__ node_box,
                            uint32_t* __restrict__ node_leaf, double* __restrict__ node_sah,
                            uint32_t* __restrict__ node_depth, uint32_t* __restrict__ any_median) {
    const uint32_t n = tk.end - tk.beg;
    float lo[3], hi[3], clo[3], chi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = fdec(tbw[k]);
        hi[k] = fdec(tbw[3 + k]);
        clo[k] = fdec(tbw[6 + k]);
        chi[k] = fdec(tbw[9 + k]);
        node_box[6 * (size_t)tk.node + k] = lo[k];
        node_box[6 * (size_t)tk.node + 3 + k] = hi[k];
    }
    node_depth[tk.node] = tk.depth;
    const double parea = box_area(lo, hi);
    Decision d{kLeaf, 0, 0, 0};
    int best_axis = -1;
    uint32_t best_bin = 0;
    double best_cost = DBL_MAX;
    const bool force_median = tk.depth + levels_needed(n, prm.max_leaf) + 1 >= prm.max_depth;
    if (n > 1 && !force_median) {
        best_axis = bb.axis;
        best_bin = bb.bin;
        best_cost = bb.cost;
    }
    bool leaf = n == 1;
    if (!leaf) {
        if (best_axis >= 0) {
            const double split_cost = prm.traversal_cost + (parea > 0.0 ? best_cost / parea : (double)n);
            if (n <= prm.max_leaf && (double)n <= split_cost) leaf = true;
            else d = Decision{kSplit, (uint32_t)best_axis, best_bin, 0};
        } else if (n <= prm.max_leaf) {
            leaf = true;
        } else {
            // object median on the widest centroid axis
            uint32_t k = 0;
            float w = -1.0f;
            for (uint32_t a = 0; a < 3; ++a) {
                const float e = chi[a] - clo[a];
                if (e > w) {
                    w = e;
                    k = a;
                }
            }
            d = Decision{kMedian, k, 0, tk.beg + n / 2};
            atomicOr(any_median, 1u);
        }
    }
    if (leaf) {
        node_leaf[2 * (size_t)tk.node] = tk.beg;
        node_leaf[2 * (size_t)tk.node + 1] = n;
        node_sah[tk.node] = parea * n;
    } else {
        node_leaf[2 * (size_t)tk.node] = 0;
        node_leaf[2 * (size_t)tk.node + 1] = 0;
        node_sah[tk.node] = parea;
    }
    dec[t] = d;
}

// One wave per big task (slot order): the SAH search over its bins in device memory as in
// k_small_tasks, then the split decision by lane 0.
__global__ void __launch_bounds__(256) k_split(const Task* __restrict__ tasks, uint32_t ntasks,
                                               const uint32_t* __restrict__ tb, const uint32_t* __restrict__ slot,
                                               const uint32_t* __restrict__ big_task, const uint32_t* __restrict__ bins,
                                               Params prm, Decision* __restrict__ dec, float* __restrict__ node_box,
                                               uint32_t* __restrict__ node_leaf, double* __restrict__ node_sah,
                                               uint32_t* __restrict__ node_depth, uint32_t* __restrict__ any_median) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t sl = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (sl >= slot[ntasks]) return;  // whole waves; slot[ntasks] is the number of big tasks
    const uint32_t t = big_task[sl];
    const Task tk = tasks[t];
    const uint32_t* tbw = tb + 12 * (size_t)t;
    const uint32_t* L = bins + (size_t)sl * 3 * prm.bins * 7;
    float cl[3], ch[3];
    for (int k = 0; k < 3; ++k) {
        cl[k] = fdec(tbw[6 + k]);
        ch[k] = fdec(tbw[9 + k]);
    }
    const Best best = wave_best(L, prm.bins, cl, ch, lane);
    if (lane == 0) decide_task(t, tk, tbw, best, prm, dec, node_box, node_leaf, node_sah, node_depth, any_median);
}

// One wave per small task: bounds (wave reductions), bins (in the wave's LDS), then the split
// decision by lane 0 from the LDS bins. No atomics outside LDS.
__global__ void __launch_bounds__(256) k_small_tasks(const Task* __restrict__ tasks, uint32_t ntasks,
                                                     const uint32_t* __restrict__ ids, const PBox* __restrict__ pb,
                                                     const float* __restrict__ cen, Params prm,
                                                     uint32_t* __restrict__ tb, Decision* __restrict__ dec,
                                                     float* __restrict__ node_box, uint32_t* __restrict__ node_leaf,
                                                     double* __restrict__ node_sah, uint32_t* __restrict__ node_depth,
                                                     uint32_t* __restrict__ any_median) {
    __shared__ uint32_t wl[4][3 * kMaxBins * 7];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (t >= ntasks) return;  // whole waves
    const Task tk = tasks[t];
    if (tk.end - tk.beg > kSmallTask || tk.end - tk.beg <= kTinyTask) return;
    uint32_t v[12];
    for (int k = 0; k < 3; ++k) {
        v[k] = kEncMax;
        v[3 + k] = kEncMin;
        v[6 + k] = kEncMax;
        v[9 + k] = kEncMin;
    }
    for (uint32_t i = tk.beg + lane; i < tk.end; i += 64) {
        const uint32_t p = ids[i];
        const PBox b = pb[p];
        for (int k = 0; k < 3; ++k) {
            v[k] = min(v[k], fenc(b.lo[k]));
            v[3 + k] = max(v[3 + k], fenc(b.hi[k]));
            const uint32_t c = fenc(cen[3 * (size_t)p + k]);
            v[6 + k] = min(v[6 + k], c);
            v[9 + k] = max(v[9 + k], c);
        }
    }
    for (int q = 0; q < 12; ++q) {
        const bool mn = q < 3 || (q >= 6 && q < 9);
        v[q] = mn ? wave_red<true>(v[q]) : wave_red<false>(v[q]);  // every lane holds the result
    }
    if (lane == 0)
        for (int q = 0; q < 12; ++q) tb[12 * (size_t)t + q] = v[q];
    const uint32_t B = prm.bins, words = 3 * B * 7;
    uint32_t* L = wl[w];
    for (uint32_t j = lane; j < words; j += 64) {
        const uint32_t r = j % 7;
        L[j] = r < 3 ? kEncMax : (r < 6 ? kEncMin : 0u);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    float clo[3], scale[3];
    bool use[3];
    for (int k = 0; k < 3; ++k) {
        clo[k] = fdec(v[6 + k]);
        const float ext = fdec(v[9 + k]) - clo[k];
        use[k] = ext > 0.0f;
        scale[k] = use[k] ? (float)B / ext : 0.0f;
    }
    for (uint32_t i = tk.beg + lane; i < tk.end; i += 64) {
        const uint32_t p = ids[i];
        const PBox b = pb[p];
        for (int k = 0; k < 3; ++k) {
            if (!use[k]) continue;
            const uint32_t bi = bin_of(cen[3 * (size_t)p + k], clo[k], scale[k], B);
            uint32_t* q = L + (k * B + bi) * 7;
            for (int a = 0; a < 3; ++a) {
                atomicMin(q + a, fenc(b.lo[a]));
                atomicMax(q + 3 + a, fenc(b.hi[a]));
            }
            atomicAdd(q + 6, 1u);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    float cl[3], ch[3];
    for (int k = 0; k < 3; ++k) {
        cl[k] = fdec(v[6 + k]);
        ch[k] = fdec(v[9 + k]);
    }
    const Best best = wave_best(L, B, cl, ch, lane);
    if (lane == 0)
        decide_task(t, tk, v, best, prm, dec, node_box, node_leaf, node_sah, node_depth, any_median);
}

// Left flag of every position of a split (binned) or median task.
// One lane per task of at most kTinyTask primitives. With so few primitives the bin scan comes
// down to the splits after the occupied bins: an empty bin repeats the cost of the split after
// the occupied bin below it, and the scan keeps the first minimum in (axis, bin) order, so the
// candidate for an occupied bin b is the split after b, with the primitives in bins <= b on the
// left. Boxes are reduced in the encoded form the bins hold, so each cost is the bin scan's to
// the bit.
__global__ void __launch_bounds__(256) k_tiny_tasks(const Task* __restrict__ tasks, uint32_t ntasks,
                                                    const uint32_t* __restrict__ ids, const PBox* __restrict__ pb,
                                                    const float* __restrict__ cen, Params prm,
                                                    uint32_t* __restrict__ tb, Decision* __restrict__ dec,
                                                    float* __restrict__ node_box, uint32_t* __restrict__ node_leaf,
                                                    double* __restrict__ node_sah, uint32_t* __restrict__ node_depth,
                                                    uint32_t* __restrict__ any_median) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntasks) return;
    const Task tk = tasks[t];
    const uint32_t n = tk.end - tk.beg;
    if (n > kTinyTask) return;
    uint32_t lo[kTinyTask][3], hi[kTinyTask][3], v[12];
    float c[kTinyTask][3];
    for (int k = 0; k < 3; ++k) {
        v[k] = kEncMax;
        v[3 + k] = kEncMin;
        v[6 + k] = kEncMax;
        v[9 + k] = kEncMin;
    }
#pragma unroll
    for (uint32_t i = 0; i < kTinyTask; ++i) {
        for (int k = 0; k < 3; ++k) {
            lo[i][k] = kEncMax;
            hi[i][k] = kEncMin;
            c[i][k] = 0.0f;
        }
        if (i < n) {
            const uint32_t p = ids[tk.beg + i];
            const PBox b = pb[p];
            for (int k = 0; k < 3; ++k) {
                lo[i][k] = fenc(b.lo[k]);
                hi[i][k] = fenc(b.hi[k]);
                c[i][k] = cen[3 * (size_t)p + k];
                const uint32_t e = fenc(c[i][k]);
                v[k] = min(v[k], lo[i][k]);
                v[3 + k] = max(v[3 + k], hi[i][k]);
                v[6 + k] = min(v[6 + k], e);
                v[9 + k] = max(v[9 + k], e);
            }
        }
    }
    for (int q = 0; q < 12; ++q) tb[12 * (size_t)t + q] = v[q];
    const uint32_t B = prm.bins;
    Best best{-1, 0, DBL_MAX};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float clo = fdec(v[6 + k]);
        const float ext = fdec(v[9 + k]) - clo;
        if (!(ext > 0.0f)) continue;
        const float scale = (float)B / ext;
        uint32_t bi[kTinyTask];
#pragma unroll
        for (uint32_t i = 0; i < kTinyTask; ++i) bi[i] = i < n ? bin_of(c[i][k], clo, scale, B) : kNone;
        uint32_t kbin = 0;
        double kcost = DBL_MAX;
#pragma unroll
        for (uint32_t j = 0; j < kTinyTask; ++j) {
            if (j >= n) continue;
            const uint32_t s = bi[j];
            uint32_t L[6] = {kEncMax, kEncMax, kEncMax, kEncMin, kEncMin, kEncMin};
            uint32_t R[6] = {kEncMax, kEncMax, kEncMax, kEncMin, kEncMin, kEncMin};
            uint32_t lc = 0, rc = 0;
#pragma unroll
            for (uint32_t i = 0; i < kTinyTask; ++i) {
                if (i >= n) continue;
                const bool left = bi[i] <= s;
                for (int a = 0; a < 3; ++a) {
                    if (left) {
                        L[a] = min(L[a], lo[i][a]);
                        L[3 + a] = max(L[3 + a], hi[i][a]);
                    } else {
                        R[a] = min(R[a], lo[i][a]);
                        R[3 + a] = max(R[3 + a], hi[i][a]);
                    }
                }
                lc += left ? 1u : 0u;
                rc += left ? 0u : 1u;
            }
            if (rc == 0) continue;  // the highest occupied bin: nothing on the right
            float ll[3], lh[3], rl[3], rh[3];
            for (int a = 0; a < 3; ++a) {
                ll[a] = fdec(L[a]);
                lh[a] = fdec(L[3 + a]);
                rl[a] = fdec(R[a]);
                rh[a] = fdec(R[3 + a]);
            }
            const double cost = box_area(ll, lh) * lc + box_area(rl, rh) * rc;
            if (cost < kcost || (cost == kcost && s < kbin)) {
                kcost = cost;
                kbin = s;
            }
        }
        if (kcost < best.cost) best = Best{k, kbin, kcost};
    }
    decide_task(t, tk, v, best, prm, dec, node_box, node_leaf, node_sah, node_depth, any_median);
}

__global__ void k_flags(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ pos_task, uint32_t n,
                        const float* __restrict__ cen, const uint32_t* __restrict__ tb, const Decision* __restrict__ dec,
                        uint32_t B, uint32_t* __restrict__ flag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    uint32_t f = 0;
    const uint32_t t = i < n ? pos_task[i] : kNone;
    if (t != kNone) {
        const Decision d = dec[t];
        if (d.kind == kSplit) {
            const uint32_t k = d.axis;
            const float clo = fdec(tb[12 * (size_t)t + 6 + k]), chi = fdec(tb[12 * (size_t)t + 9 + k]);
            const float scale = (float)B / (chi - clo);
            f = bin_of(cen[3 * (size_t)ids[i] + k], clo, scale, B) <= d.bin ? 1u : 0u;
        } else if (d.kind == kMedian) {
            f = i < d.mid ? 1u : 0u;  // the segment is already sorted by (centroid, id)
        }
    }
    flag[i] = f;
}

// Stable partition of every split task's segment; other positions keep their ids.
__global__ void k_scatter(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ pos_task, uint32_t n,
                          const Task* __restrict__ tasks, const Decision* __restrict__ dec,
                          const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pre,
                          uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = pos_task[i];
    uint32_t dst = i;
    if (t != kNone && dec[t].kind != kLeaf) {
        const Task tk = tasks[t];
        const uint32_t nleft = pre[tk.end] - pre[tk.beg];
        const uint32_t lbefore = pre[i] - pre[tk.beg];
        dst = flag[i] ? tk.beg + lbefore : tk.beg + nleft + (i - tk.beg - lbefore);
    }
    out[dst] = ids[i];
}

// Children of the split tasks: next level's tasks (in order) and the new node ids.
__global__ void k_children(const Task* __restrict__ tasks, uint32_t ntasks, Decision* __restrict__ dec,
                           const uint32_t* __restrict__ split_pre, const uint32_t* __restrict__ pre,
                           uint32_t node_base, Task* __restrict__ next, uint32_t* __restrict__ node_child) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntasks) return;
    const Task tk = tasks[t];
    Decision d = dec[t];
    if (d.kind == kLeaf) return;
    const uint32_t n = tk.end - tk.beg;
    uint32_t mid = tk.beg + (pre[tk.end] - pre[tk.beg]);
    if (mid == tk.beg || mid == tk.end) mid = tk.beg + n / 2;  // never an empty child
    const uint32_t r = split_pre[t];
    const uint32_t l = node_base + 2 * r;
    node_child[2 * (size_t)tk.node] = l;
    node_child[2 * (size_t)tk.node + 1] = l + 1;
    next[2 * r] = Task{tk.beg, mid, l, tk.depth + 1};
    next[2 * r + 1] = Task{mid, tk.end, l + 1, tk.depth + 1};
}

// Big-task flag, for the bins slots of the position-parallel kernels.
__global__ void k_big_flags(const Task* __restrict__ tasks, uint32_t ntasks, uint32_t* __restrict__ f) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntasks) return;
    f[t] = (t < ntasks && tasks[t].end - tasks[t].beg > kSmallTask) ? 1u : 0u;
}

// Split flag (1 for split and median tasks) for the child numbering.
__global__ void k_split_flags(const Decision* __restrict__ dec, uint32_t ntasks, uint32_t* __restrict__ f) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntasks) return;
    f[t] = (t < ntasks && dec[t].kind != kLeaf) ? 1u : 0u;
}

// ------------------------------------------------------------------ exclusive scan (u32)
// 2048 elements per workgroup of 256 (8 per thread); block totals scanned recursively.
constexpr uint32_t kScanPer = 2048;

__global__ void __launch_bounds__(256) k_scan_block(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                    size_t n, uint32_t* __restrict__ sums) {
    __shared__ uint32_t wsum[4];
    const size_t base = (size_t)blockIdx.x * kScanPer + (size_t)threadIdx.x * 8;
    uint32_t v[8], run = 0;
    for (int j = 0; j < 8; ++j) {
        v[j] = base + j < n ? in[base + j] : 0u;
        const uint32_t x = v[j];
        v[j] = run;
        run += x;
    }
    // inclusive scan of the thread totals across the wave
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = run;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(inc, off, 64);
        if (lane >= (uint32_t)off) inc += o;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t woff = 0;
    for (uint32_t q = 0; q < wave; ++q) woff += wsum[q];
    const uint32_t toff = woff + inc - run;
    for (int j = 0; j < 8; ++j)
        if (base + j < n) out[base + j] = v[j] + toff;
    if (threadIdx.x == 255) sums[blockIdx.x] = woff + inc;
}

__global__ void k_scan_add(uint32_t* __restrict__ out, size_t n, const uint32_t* __restrict__ offs) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += offs[i / kScanPer];
}

size_t scan_temp_words(size_t n) {
    size_t w = 0;
    while (n > kScanPer) {
        n = (n + kScanPer - 1) / kScanPer;
        w += 2 * n;
    }
    return w + 2;
}

// out[i] = sum of in[0..i) for i < n; tmp: scan_temp_words(n) words.
hipError_t scan_exclusive(const uint32_t* in, uint32_t* out, size_t n, uint32_t* tmp, hipStream_t s) {
    const size_t nb = (n + kScanPer - 1) / kScanPer;
    uint32_t* sums = tmp;
    hipLaunchKernelGGL(k_scan_block, dim3((uint32_t)nb), dim3(256), 0, s, in, out, n, sums);
    if (nb > 1) {
        uint32_t* offs = tmp + nb;
        hipError_t e = scan_exclusive(sums, offs, nb, tmp + 2 * nb, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_scan_add, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, out, n, offs);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ host driver
// A pair of events owned for the span of one build.
struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
    Events() = default;
    Events(const Events&) = delete;
    Events& operator=(const Events&) = delete;
    ~Events() {
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

struct DevBuf {
    std::vector<void*> ptrs;
    ~DevBuf() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    T* get(size_t count) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count * sizeof(T), 16)) != hipSuccess) throw std::bad_alloc();
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw DeviceBuildError(e, what);
}

inline dim3 grid_of(size_t n, uint32_t b = 256) { return dim3((uint32_t)((n + b - 1) / b)); }

}  // namespace

void build_bvh_device(const float* points, uint32_t stride, uint32_t npoints, const uint32_t* indices,
                      uint32_t ntris, const float* spheres, uint32_t nspheres, const BuildOptions& opt,
                      hipStream_t s, BVHTree* out, BuildTimes* times) {
    if (opt.max_leaf < 1 || opt.max_leaf > 8) throw std::invalid_argument("max_leaf must be in 1..8");
    if (opt.bins < 2 || opt.bins > kMaxBins) throw std::invalid_argument("the device builder takes 2..32 bins");
    if (opt.max_depth < 1 || opt.max_depth > 64) throw std::invalid_argument("max_depth must be in 1..64");
    *out = BVHTree();
    const uint32_t n = ntris + nspheres;
    if (n == 0) return;
    for (size_t i = 0; i < 3 * (size_t)ntris; ++i)
        if (indices[i] >= npoints) throw std::invalid_argument("triangle index out of range");
    // the float encoding orders finite values only
    for (uint32_t i = 0; i < npoints; ++i)
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(points[(size_t)i * stride + k]))
                throw std::invalid_argument("the device builder takes finite vertices only");
    for (size_t i = 0; i < 4 * (size_t)nspheres; ++i)
        if (!std::isfinite(spheres[i])) throw std::invalid_argument("the device builder takes finite spheres only");

    Events ev;  // destroyed on every exit, the throwing ones included
    check(hipEventCreate(&ev.e[0]), "hipEventCreate");
    check(hipEventCreate(&ev.e[1]), "hipEventCreate");
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    DevBuf m;
    const uint32_t B = opt.bins;
    float* d_pts = m.get<float>((size_t)npoints * stride);
    uint32_t* d_idx = m.get<uint32_t>(3 * (size_t)ntris);
    float* d_sph = m.get<float>(4 * (size_t)nspheres);
    if (npoints) check(hipMemcpyAsync(d_pts, points, (size_t)npoints * stride * 4, hipMemcpyHostToDevice, s), "copy");
    if (ntris) check(hipMemcpyAsync(d_idx, indices, 12 * (size_t)ntris, hipMemcpyHostToDevice, s), "copy");
    if (nspheres) check(hipMemcpyAsync(d_sph, spheres, 16 * (size_t)nspheres, hipMemcpyHostToDevice, s), "copy");
    check(hipEventRecord(e0, s), "hipEventRecord");

    PBox* pb = m.get<PBox>(n);
    float* cen = m.get<float>(3 * (size_t)n);
    uint32_t* ids = m.get<uint32_t>(n);
    uint32_t* ids2 = m.get<uint32_t>(n);
    uint32_t* pos_task = m.get<uint32_t>(n);
    uint32_t* pos_big = m.get<uint32_t>(n);
    uint32_t* flag = m.get<uint32_t>((size_t)n + 1);
    uint32_t* pre = m.get<uint32_t>((size_t)n + 1);
    uint32_t* scan_tmp = m.get<uint32_t>(scan_temp_words((size_t)n + 1));
    // a tree over n primitives with leaves of >= 1 has at most 2n - 1 nodes; a level has at most
    // n / 2 + 1 tasks of two or more... and leaves of one, so bound tasks by n
    const size_t max_nodes = 2 * (size_t)n;
    const uint32_t max_tasks = n;
    Task* tasks = m.get<Task>(max_tasks);
    Task* next = m.get<Task>(max_tasks);
    Decision* dec = m.get<Decision>(max_tasks);
    uint32_t* tb = m.get<uint32_t>(12 * (size_t)max_tasks);
    uint32_t* sflag = m.get<uint32_t>((size_t)max_tasks + 1);
    uint32_t* spre = m.get<uint32_t>((size_t)max_tasks + 1);
    uint32_t* stmp = m.get<uint32_t>(scan_temp_words((size_t)max_tasks + 1));
    float* node_box = m.get<float>(6 * max_nodes);
    uint32_t* node_leaf = m.get<uint32_t>(2 * max_nodes);
    uint32_t* node_child = m.get<uint32_t>(2 * max_nodes);
    double* node_sah = m.get<double>(max_nodes);
    uint32_t* node_depth = m.get<uint32_t>(max_nodes);
    uint32_t* any_median = m.get<uint32_t>(1);
    // bins of the big tasks (more than kSmallTask primitives each, so at most n / kSmallTask)
    const size_t max_big = (size_t)n / (kSmallTask + 1) + 1;
    const size_t bwords = max_big * 3 * B * 7;
    uint32_t* bins = m.get<uint32_t>(bwords);
    uint32_t* bslot = m.get<uint32_t>((size_t)max_tasks + 1);
    uint32_t* big_task = m.get<uint32_t>(max_big);

    hipLaunchKernelGGL(k_prim_boxes, grid_of(n), dim3(256), 0, s, d_pts, stride, d_idx, ntris, d_sph, nspheres, pb,
                       cen, ids);
    const Task root{0, n, 0, 0};
    check(hipMemcpyAsync(tasks, &root, sizeof(Task), hipMemcpyHostToDevice, s), "copy");
    uint32_t ntasks = 1, nnodes = 1;
    Params prm{n, B, opt.max_leaf, opt.max_depth, opt.traversal_cost};
    std::vector<float> hcen;  // host centroids, for a median split (computed on first need)
    while (ntasks) {
        hipLaunchKernelGGL(k_task_of, grid_of(n), dim3(256), 0, s, tasks, ntasks, n, pos_task, pos_big);
        hipLaunchKernelGGL(k_big_flags, grid_of((size_t)ntasks + 1), dim3(256), 0, s, tasks, ntasks, sflag);
        check(scan_exclusive(sflag, bslot, (size_t)ntasks + 1, stmp, s), "scan");
        hipLaunchKernelGGL(k_clear_bounds, grid_of(12 * (size_t)ntasks), dim3(256), 0, s, tb, ntasks, tasks, bslot,
                           big_task);
        hipLaunchKernelGGL(k_bounds, grid_of(n, 256 * kSubBlocks), dim3(256), 0, s, ids, pos_big, n, pb, cen, tb);
        hipLaunchKernelGGL(k_clear_bins, grid_of(bwords), dim3(256), 0, s, bins, bwords);
        hipLaunchKernelGGL(k_bin, grid_of(n, 256 * kSubBlocks), dim3(256), 0, s, ids, pos_big, n, pb, cen, tb, B, bslot,
                           bins);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, any_median, (size_t)1, 0u);
        hipLaunchKernelGGL(k_split, grid_of(64 * max_big), dim3(256), 0, s, tasks, ntasks, tb, bslot, big_task, bins,
                           prm, dec, node_box, node_leaf, node_sah, node_depth, any_median);
        hipLaunchKernelGGL(k_small_tasks, grid_of(64 * (size_t)ntasks), dim3(256), 0, s, tasks, ntasks, ids, pb, cen,
                           prm, tb, dec, node_box, node_leaf, node_sah, node_depth, any_median);
        hipLaunchKernelGGL(k_tiny_tasks, grid_of(ntasks), dim3(256), 0, s, tasks, ntasks, ids, pb, cen, prm, tb, dec,
                           node_box, node_leaf, node_sah, node_depth, any_median);
        check(hipGetLastError(), "level kernels");
        // The level's partition (flags, scan, stable scatter into ids2) and the numbering of its
        // splits are queued at once, and the host learns in one round trip both the split count
        // and whether a task takes an object median. Median splits (near the depth limit, or no
        // centroid extent) are rare: the host then sorts those segments of ids by centroid and
        // the partition runs again; the split count does not depend on it.
        auto partition = [&]() {
            hipLaunchKernelGGL(k_flags, grid_of((size_t)n + 1), dim3(256), 0, s, ids, pos_task, n, cen, tb, dec, B,
                               flag);
            check(scan_exclusive(flag, pre, (size_t)n + 1, scan_tmp, s), "scan");
            hipLaunchKernelGGL(k_scatter, grid_of(n), dim3(256), 0, s, ids, pos_task, n, tasks, dec, flag, pre, ids2);
        };
        partition();
        hipLaunchKernelGGL(k_split_flags, grid_of((size_t)ntasks + 1), dim3(256), 0, s, dec, ntasks, sflag);
        check(scan_exclusive(sflag, spre, (size_t)ntasks + 1, stmp, s), "scan");
        uint32_t hany = 0, nsplit = 0;
        check(hipMemcpyAsync(&hany, any_median, 4, hipMemcpyDeviceToHost, s), "copy");
        check(hipMemcpyAsync(&nsplit, spre + ntasks, 4, hipMemcpyDeviceToHost, s), "copy");
        check(hipStreamSynchronize(s), "sync");
        if (hany) {
            std::vector<Decision> hdec(ntasks);
            std::vector<Task> htask(ntasks);
            check(hipMemcpyAsync(hdec.data(), dec, ntasks * sizeof(Decision), hipMemcpyDeviceToHost, s), "copy");
            check(hipMemcpyAsync(htask.data(), tasks, ntasks * sizeof(Task), hipMemcpyDeviceToHost, s), "copy");
            check(hipStreamSynchronize(s), "sync");
            for (uint32_t t = 0; t < ntasks; ++t) {
                if (hdec[t].kind != kMedian) continue;
                if (hcen.empty()) {
                    hcen.resize(3 * (size_t)n);
                    check(hipMemcpy(hcen.data(), cen, 12 * (size_t)n, hipMemcpyDeviceToHost), "copy");
                }
                const Task& tk = htask[t];
                std::vector<uint32_t> seg(tk.end - tk.beg);
                check(hipMemcpy(seg.data(), ids + tk.beg, 4 * seg.size(), hipMemcpyDeviceToHost), "copy");
                const uint32_t k = hdec[t].axis;
                std::sort(seg.begin(), seg.end(), [&](uint32_t x, uint32_t y) {
                    const float cx = hcen[3 * (size_t)x + k], cy = hcen[3 * (size_t)y + k];
                    return cx < cy || (cx == cy && x < y);
                });
                check(hipMemcpy(ids + tk.beg, seg.data(), 4 * seg.size(), hipMemcpyHostToDevice), "copy");
            }
            partition();
        }
        std::swap(ids, ids2);
        if (nnodes + 2 * (size_t)nsplit > max_nodes || 2 * (size_t)nsplit > max_tasks)
            throw std::runtime_error("device BVH build: node count out of bounds");
        hipLaunchKernelGGL(k_children, grid_of(ntasks), dim3(256), 0, s, tasks, ntasks, dec, spre, pre, nnodes, next,
                           node_child);
        check(hipGetLastError(), "children");
        std::swap(tasks, next);
        nnodes += 2 * nsplit;
        ntasks = 2 * nsplit;
    }
    check(hipEventRecord(e1, s), "hipEventRecord");

    // back to the host: level-numbered nodes -> depth-first pre-order
    std::vector<float> hbox(6 * (size_t)nnodes);
    std::vector<uint32_t> hleaf(2 * (size_t)nnodes), hchild(2 * (size_t)nnodes), hdepth(nnodes);
    std::vector<double> hsah(nnodes);
    out->prim_ids.resize(n);
    check(hipMemcpyAsync(hbox.data(), node_box, hbox.size() * 4, hipMemcpyDeviceToHost, s), "copy");
    check(hipMemcpyAsync(hleaf.data(), node_leaf, hleaf.size() * 4, hipMemcpyDeviceToHost, s), "copy");
    check(hipMemcpyAsync(hchild.data(), node_child, hchild.size() * 4, hipMemcpyDeviceToHost, s), "copy");
    check(hipMemcpyAsync(hdepth.data(), node_depth, hdepth.size() * 4, hipMemcpyDeviceToHost, s), "copy");
    check(hipMemcpyAsync(hsah.data(), node_sah, hsah.size() * 8, hipMemcpyDeviceToHost, s), "copy");
    check(hipMemcpyAsync(out->prim_ids.data(), ids, 4 * (size_t)n, hipMemcpyDeviceToHost, s), "copy");
    check(hipStreamSynchronize(s), "sync");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");

    // children always have larger level ids than their parent
    std::vector<uint32_t> size(nnodes, 1), pre_id(nnodes, 0);
    for (size_t x = nnodes; x-- > 0;)
        if (hleaf[2 * x + 1] == 0) size[x] = 1 + size[hchild[2 * x]] + size[hchild[2 * x + 1]];
    out->nodes.resize(nnodes);
    double sah = 0.0;
    uint32_t depth = 0, leaves = 0;
    for (size_t x = 0; x < nnodes; ++x) {
        const uint32_t me = pre_id[x];
        BuildNode& b = out->nodes[me];
        for (int k = 0; k < 3; ++k) {
            b.bmin[k] = hbox[6 * x + k];
            b.bmax[k] = hbox[6 * x + 3 + k];
        }
        if (hleaf[2 * x + 1] != 0) {
            b.left = b.right = -1;
            b.prim_beg = hleaf[2 * x];
            b.prim_cnt = hleaf[2 * x + 1];
            depth = std::max(depth, hdepth[x]);
            ++leaves;
        } else {
            const uint32_t l = hchild[2 * x], r = hchild[2 * x + 1];
            pre_id[l] = me + 1;
            pre_id[r] = me + 1 + size[l];
            b.left = (int32_t)pre_id[l];
            b.right = (int32_t)pre_id[r];
            b.prim_beg = 0;
            b.prim_cnt = 0;
        }
        sah += hsah[x];
    }
    out->depth = depth;
    out->leaf_count = leaves;
    const float* rb = &hbox[0];
    double dx = (double)rb[3] - rb[0], dy = (double)rb[4] - rb[1], dz = (double)rb[5] - rb[2];
    const double ra = (dx < 0 || dy < 0 || dz < 0) ? 0.0 : 2.0 * (dx * dy + dy * dz + dz * dx);
    out->sah_cost = ra > 0.0 ? sah / ra : 0.0;
    if (times) times->device_ms = ms;
}

}  // namespace rgpu
