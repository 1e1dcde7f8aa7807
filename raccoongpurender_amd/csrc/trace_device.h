// Device-side traversal routines shared by the ray-cast kernels and the wavefront stages.
// See traverse.hip for the reference behaviour they restate.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "rgpu_traverse.h"
#include "rgpu_types.h"

namespace rgpu {

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 sub3(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross3(F3 a, F3 b) {
    return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// One axis of the slab test: entry/exit parameters with the NaN substitution of
// kernels/slab.cl:15-16 (t0 first, then t1 from the updated t0). kPlain leaves the
// substitution out; callers choose it only where neither product can be NaN, so both forms
// give the same bits there.
template <bool kPlain = false>
__device__ __forceinline__ void slab_axis(float lo, float hi, float o, float inv, float* tn, float* tf) {
    float t0 = (lo - o) * inv;
    float t1 = (hi - o) * inv;
    if (!kPlain) {
        t0 = __builtin_isnan(t0) ? -t1 : t0;
        t1 = __builtin_isnan(t1) ? -t0 : t1;
    }
    *tn = fminf(t0, t1);
    *tf = fmaxf(t0, t1);
}

// (plane - o) * inv is never NaN when the planes and the origin are finite and every inverse
// direction component is finite and non-zero: a finite or overflowed difference times a finite
// non-zero factor. Then the NaN substitution of the slab test never fires. boxes_finite is a
// property of the scene (all node boxes finite).
__device__ __forceinline__ bool plain_slabs(bool boxes_finite, F3 o, F3 inv) {
    const bool fo = __builtin_isfinite(o.x) && __builtin_isfinite(o.y) && __builtin_isfinite(o.z);
    const bool fi = __builtin_isfinite(inv.x) && __builtin_isfinite(inv.y) && __builtin_isfinite(inv.z) &&
                    inv.x != 0.0f && inv.y != 0.0f && inv.z != 0.0f;
    return boxes_finite && fo && fi;
}

template <bool kPlain = false>
__device__ __forceinline__ bool slab_hit(const float* lo, const float* hi, F3 o, F3 inv, float farclip,
                                         float* h0, float* h1) {
    float nx, fx, ny, fy, nz, fz;
    slab_axis<kPlain>(lo[0], hi[0], o.x, inv.x, &nx, &fx);
    slab_axis<kPlain>(lo[1], hi[1], o.y, inv.y, &ny, &fy);
    slab_axis<kPlain>(lo[2], hi[2], o.z, inv.z, &nz, &fz);
    float rmin = fmaxf(fmaxf(nx, ny), nz);
    float rmax = fminf(fminf(fx, fy), fz);
    *h0 = fmaxf(0.0f, rmin);
    *h1 = fminf(rmax, farclip);
    return rmin <= rmax && 0.0f <= rmax && rmin <= farclip;
}

// Both child boxes of a stackless node; the wave takes the plain slab form when every active
// lane may (a wave-uniform branch, so no lane runs both forms).
__device__ __forceinline__ void slab_pair(const float* lmin, const float* lmax, const float* rmin, const float* rmax,
                                          F3 o, F3 inv, float farclip, bool plain, bool* L, bool* R, float* Lh0,
                                          float* Lh1, float* Rh0, float* Rh1) {
    if (__all(plain)) {
        *L = slab_hit<true>(lmin, lmax, o, inv, farclip, Lh0, Lh1);
        *R = slab_hit<true>(rmin, rmax, o, inv, farclip, Rh0, Rh1);
    } else {
        *L = slab_hit<false>(lmin, lmax, o, inv, farclip, Lh0, Lh1);
        *R = slab_hit<false>(rmin, rmax, o, inv, farclip, Rh0, Rh1);
    }
}

__device__ __forceinline__ bool slab_one(const float* lo, const float* hi, F3 o, F3 inv, float farclip, bool plain,
                                         float* h0, float* h1) {
    if (__all(plain)) return slab_hit<true>(lo, hi, o, inv, farclip, h0, h1);
    return slab_hit<false>(lo, hi, o, inv, farclip, h0, h1);
}

// Test one primitive record; updates (tmin, u, v, prim) when it is the closest so far.
__device__ __forceinline__ void test_prim(const PrimRec* __restrict__ prims, uint32_t slot, int32_t sphere_begin,
                                          F3 ro, F3 rd, float* tmin, float* uo, float* vo, int32_t* prim) {
    const float4* rec = reinterpret_cast<const float4*>(prims + slot);
    float4 q0 = rec[0];
    float4 q2 = rec[2];
    int32_t pid = (int32_t)__float_as_uint(q2.y);
    if (pid < sphere_begin) {
        float4 q1 = rec[1];
        F3 v0 = f3(q0.x, q0.y, q0.z);
        F3 e1 = f3(q0.w, q1.x, q1.y);
        F3 e2 = f3(q1.z, q1.w, q2.x);
        F3 pvec = cross3(rd, e2);
        float det = dot3(e1, pvec);
        if (fabsf(det) < 1.0e-8f) return;
        float invDet = 1.0f / det;
        F3 tvec = sub3(ro, v0);
        float u = dot3(tvec, pvec) * invDet;
        if (u < 0.0f || u > 1.0f) return;
        F3 qvec = cross3(tvec, e1);
        float v = dot3(rd, qvec) * invDet;
        if (v < 0.0f || u + v > 1.0f) return;
        float t = dot3(e2, qvec) * invDet;
        if (t < 0.0f) return;
        if (*tmin < t) return;
        *tmin = t;
        *uo = u;
        *vo = v;
        *prim = pid;
    } else {
        // kernels/extension_ray_cast_stackless.cl:52-76
        F3 c = f3(q0.x, q0.y, q0.z);
        float r = q0.w;
        float r_sq = r * r;
        F3 S = sub3(ro, c);
        F3 crs = cross3(S, rd);
        float d_sq = dot3(crs, crs);
        float B = dot3(S, rd);
        float D = r_sq - d_sq;
        if (D < 0.0f) return;
        float Ds = __builtin_sqrtf(D);
        float t0 = -B - Ds;
        float t1 = -B + Ds;
        if (0.0f < t0 && t0 < *tmin) {
            *tmin = t0;
            *prim = pid;
            return;
        }
        if (0.0f < t1 && t1 < *tmin) {
            *tmin = t1;
            *prim = pid;
        }
    }
}

// Per-ray traversal event counts (the instrumented kernels only).
struct Cnt {
    uint32_t branch = 0, leaf = 0, parent = 0, prim = 0, node = 0;
};

// The far children a stackless walk has deferred, newest first, in six registers. The
// newest deferred child always belongs to the lowest set bit of the bitstack, so a backtrack
// can jump to it without climbing the parent pointers; once more than six are pending the
// oldest fall out, and a backtrack that finds the stack empty climbs as the reference does.
// Only loads are saved: the visit order is the reference's.
struct ShortStack {
    uint32_t s0, s1, s2, s3, s4, s5, n;
    __device__ __forceinline__ void clear() { n = 0; }
    __device__ __forceinline__ void push(uint32_t x) {
        s5 = s4;
        s4 = s3;
        s3 = s2;
        s2 = s1;
        s1 = s0;
        s0 = x;
        n = n < 6 ? n + 1 : 6;
    }
    __device__ __forceinline__ uint32_t pop() {
        const uint32_t x = s0;
        s0 = s1;
        s1 = s2;
        s2 = s3;
        s3 = s4;
        s4 = s5;
        --n;
        return x;
    }
};

// Stackless traversal of one ray (kernels/extension_ray_cast_stackless.cl:81-200).
template <bool kCount>
__device__ __forceinline__ int32_t trace_stackless(const SLNode* __restrict__ nodes,
                                                   const PrimRec* __restrict__ prims, int32_t sphere_begin,
                                                   bool boxes_finite, F3 ro, F3 rd, float* tmin_io, float* uo,
                                                   float* vo, Cnt* cnt) {
    const F3 inv = f3(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    const bool plain = plain_slabs(boxes_finite, ro, inv);
    float tmin = *tmin_io;
    float u = 0.0f, v = 0.0f;
    int32_t prim = -1;
    uint64_t bitstack = 0;
    uint32_t node = 0;
    ShortStack stk;
    stk.clear();
    const uint4* q3p = reinterpret_cast<const uint4*>(nodes);
    for (;;) {
        // the whole record in one round trip: the boxes are needed unless the node is a leaf,
        // and waiting for {a, b, parent, sibling} before asking for them would double the
        // latency of every branch visit
        const float4* rec = reinterpret_cast<const float4*>(nodes + node);
        const float4 q0 = rec[0], q1 = rec[1], q2 = rec[2];
        const uint4 q3 = q3p[4 * node + 3];  // {a, b, parent, sibling}
        if ((q3.x & kLeafFlag) == 0) {
            if (kCount) cnt->branch++;
            float lmin[3] = {q0.x, q0.y, q0.z}, lmax[3] = {q0.w, q1.x, q1.y};
            float rmin[3] = {q1.z, q1.w, q2.x}, rmax[3] = {q2.y, q2.z, q2.w};
            float Lh0, Lh1, Rh0, Rh1;
            bool L, R;
            slab_pair(lmin, lmax, rmin, rmax, ro, inv, tmin, plain, &L, &R, &Lh0, &Lh1, &Rh0, &Rh1);
            if (L || R) {
                bitstack = bitstack << 1;
                if (L && R) {
                    bool nearL = (Lh0 == Rh0) ? (Lh1 < Rh1) : (Lh0 < Rh0);
                    node = (nearL ? q3.x : q3.y) & ~kChildLeaf;
                    stk.push(nearL ? q3.y : q3.x);
                    bitstack = bitstack | 1;
                } else {
                    node = (L ? q3.x : q3.y) & ~kChildLeaf;
                }
                continue;
            }
        } else {
            const uint32_t beg = q3.x & ~kLeafFlag, end = q3.y;
            if (kCount) {
                cnt->leaf++;
                cnt->prim += end - beg;
            }
            for (uint32_t s = beg; s < end; ++s) test_prim(prims, s, sphere_begin, ro, rd, &tmin, &u, &v, &prim);
        }
        // Backtrack: climb while the top bit is 0, then step to the sibling.
        if (bitstack == 0) break;
        if (stk.n != 0) {
            // the climb the reference makes, one bit per parent, as one shift
            const uint32_t z = (uint32_t)__builtin_ctzll(bitstack);
            if (kCount) cnt->parent += z;
            bitstack = bitstack >> z;
            node = stk.pop() & ~kChildLeaf;
            bitstack = bitstack ^ 1;
            continue;
        }
        bool done = false;
        uint4 q = q3;
        while ((bitstack & 1) == 0) {
            if (bitstack == 0) {
                done = true;
                break;
            }
            node = q.z;
            if (kCount) cnt->parent++;
            bitstack = bitstack >> 1;
            if (bitstack != 0) q = q3p[4 * node + 3];
        }
        if (done) break;
        node = q.w & ~kChildLeaf;
        bitstack = bitstack ^ 1;
    }
    *tmin_io = tmin;
    *uo = u;
    *vo = v;
    return prim;
}

// Ray-direction class of kernels/extension_ray_cast.cl:45-58.
__device__ __forceinline__ uint32_t direction_class(F3 rd) {
    float ax = fabsf(rd.x), ay = fabsf(rd.y), az = fabsf(rd.z);
    float maxYZ = fmaxf(ay, az);
    float maxXZ = fmaxf(ax, az);
    if (maxYZ < ax) return 0.0f < rd.x ? 0u : 1u;
    if (maxXZ < ay) return 0.0f < rd.y ? 2u : 3u;
    return 0.0f < rd.z ? 4u : 5u;
}

// The table a ray walks: its direction class for the multi-threaded BVH, the single table
// otherwise (common/threaded_bvh.hpp:527-547, 366-395).
template <int kVariant>
__device__ __forceinline__ const MTNode* threaded_table(const MTNode* tables, uint32_t nnodes, F3 rd) {
    const uint32_t order = kVariant == kVariantSingleThreaded ? (uint32_t)kMTSingleOrder : direction_class(rd);
    return tables + (size_t)order * nnodes;
}

// Threaded traversal of one ray (common/threaded_bvh.hpp:527-569, 366-395).
template <bool kCount, int kVariant>
__device__ __forceinline__ int32_t trace_threaded(const MTNode* __restrict__ tables, uint32_t nnodes,
                                                  const PrimRec* __restrict__ prims, int32_t sphere_begin,
                                                  bool boxes_finite, F3 ro, F3 rd, float* tmin_io, float* uo,
                                                  float* vo, Cnt* cnt) {
    const MTNode* tab = threaded_table<kVariant>(tables, nnodes, rd);
    const F3 inv = f3(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    const bool plain = plain_slabs(boxes_finite, ro, inv);
    float tmin = *tmin_io;
    float u = 0.0f, v = 0.0f;
    int32_t prim = -1;
    uint32_t node = 0;
    while (node != kInvalid) {
        const float4* rec = reinterpret_cast<const float4*>(tab + node);
        float4 q0 = rec[0], q1 = rec[1];
        float lo[3] = {q0.x, q0.y, q0.z}, hi[3] = {q0.w, q1.x, q1.y};
        uint32_t miss = __float_as_uint(q1.z), leaf = __float_as_uint(q1.w);
        float h0, h1;
        if (kCount) cnt->node++;
        bool hit = slab_one(lo, hi, ro, inv, tmin, plain, &h0, &h1);
        if (hit && leaf != 0) {
            const uint32_t beg = leaf & 0x0FFFFFFFu;
            const uint32_t end = beg + ((leaf >> 28) & 7u) + 1u;
            if (kCount) cnt->prim += end - beg;
            for (uint32_t s = beg; s < end; ++s) test_prim(prims, s, sphere_begin, ro, rd, &tmin, &u, &v, &prim);
        }
        node = (hit && leaf == 0) ? node + 1 : miss;
    }
    *tmin_io = tmin;
    *uo = u;
    *vo = v;
    return prim;
}

// ---------------------------------------------------------------- resumable traversal
// The same two traversals cut into single steps over explicit per-lane state, so that a lane
// whose ray has finished can take the next ray while the rest of its wave keeps going. Each
// ray still visits nodes and tests primitives in exactly the order above.

struct ThreadedRay {
    const MTNode* tab;
    F3 ro, rd, inv;
    float tmin, u, v;
    int32_t prim;
    bool plain;     // origin and inverse direction allow the plain slab form (set once per ray)
    uint32_t node;  // kInvalid when finished
    uint32_t leaf;  // primitive word of an entered leaf not yet tested (0: none)
};

template <int kVariant>
__device__ __forceinline__ void threaded_begin(ThreadedRay& r, const MTNode* tables, uint32_t nnodes, F3 ro, F3 rd,
                                               float tmax) {
    r.tab = threaded_table<kVariant>(tables, nnodes, rd);
    r.ro = ro;
    r.rd = rd;
    r.inv = f3(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    r.plain = plain_slabs(true, ro, r.inv);
    r.tmin = tmax;
    r.u = r.v = 0.0f;
    r.prim = -1;
    r.node = 0;
    r.leaf = 0;
}

// Box test of the current node of trace_threaded: a branch moves on (hit: next record, miss:
// miss link); an entered leaf is marked pending with its primitive word kept in r.leaf.
__device__ __forceinline__ void threaded_visit(ThreadedRay& r, bool boxes_finite) {
    const float4* rec = reinterpret_cast<const float4*>(r.tab + r.node);
    float4 q0 = rec[0], q1 = rec[1];
    float lo[3] = {q0.x, q0.y, q0.z}, hi[3] = {q0.w, q1.x, q1.y};
    uint32_t miss = __float_as_uint(q1.z), leaf = __float_as_uint(q1.w);
    float h0, h1;
    bool hit = slab_one(lo, hi, r.ro, r.inv, r.tmin, boxes_finite && r.plain, &h0, &h1);
    if (hit && leaf != 0) r.leaf = leaf;  // the miss link is where the walk goes on afterwards
    r.node = (hit && leaf == 0) ? r.node + 1 : miss;
}

// The pending leaf of a lane: its primitives; the walk goes on at the miss link already in node.
__device__ __forceinline__ void threaded_leaf(ThreadedRay& r, const PrimRec* __restrict__ prims, int32_t sphere_begin) {
    const uint32_t beg = r.leaf & 0x0FFFFFFFu;
    const uint32_t end = beg + ((r.leaf >> 28) & 7u) + 1u;
    for (uint32_t s = beg; s < end; ++s) test_prim(prims, s, sphere_begin, r.ro, r.rd, &r.tmin, &r.u, &r.v, &r.prim);
    r.leaf = 0;
}


struct StacklessRay {
    F3 ro, rd, inv;
    float tmin, u, v;
    int32_t prim;
    bool plain;     // origin and inverse direction allow the plain slab form (set once per ray)
    uint64_t bitstack;
    uint32_t node;  // kInvalid when finished
    ShortStack stk;
};

__device__ __forceinline__ void stackless_begin(StacklessRay& r, F3 ro, F3 rd, float tmax) {
    r.ro = ro;
    r.rd = rd;
    r.inv = f3(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    r.plain = plain_slabs(true, ro, r.inv);
    r.tmin = tmax;
    r.u = r.v = 0.0f;
    r.prim = -1;
    r.bitstack = 0;
    r.node = 0;
    r.stk.clear();
}

// Pending-leaf mark of a resumable walk: the lane has reached a leaf and waits for its wave to
// run the leaves together (node ids stay below 2^30; kInvalid is never marked).
static const uint32_t kPendingLeaf = 0x80000000u;
static_assert(kPendingLeaf == kChildLeaf << 1, "a leaf link becomes a pending leaf by one shift");

// The walk state for a child or sibling link: the node id, marked pending when the link names a
// leaf (its record need not be visited to learn that).
__device__ __forceinline__ uint32_t enter_link(uint32_t id) { return (id & ~kChildLeaf) | ((id & kChildLeaf) << 1); }

// Where the resumable stackless walk reads node records from. SLGlobal: the node array in
// device memory. SLTreelet: the first `top` records (the top levels of the tree, numbered
// breadth-first by the scene packer) from a copy in LDS that the persistent kernel stages at
// its start, the rest from device memory; a lane's record comes from one or the other, the
// values are the same.
extern __shared__ float4 sl_treelet[];

struct SLGlobal {
    const float4* __restrict__ g;
    __device__ __forceinline__ void record(uint32_t i, float4* q0, float4* q1, float4* q2, uint4* q3) const {
        const float4* p = g + 4 * (size_t)i;
        *q0 = p[0];
        *q1 = p[1];
        *q2 = p[2];
        *q3 = reinterpret_cast<const uint4*>(p)[3];
    }
    __device__ __forceinline__ uint4 links(uint32_t i) const { return reinterpret_cast<const uint4*>(g)[4 * (size_t)i + 3]; }
};

struct SLTreelet {
    const float4* __restrict__ g;
    uint32_t top;
    __device__ __forceinline__ void record(uint32_t i, float4* q0, float4* q1, float4* q2, uint4* q3) const {
        if (i < top) {
            const float4* p = sl_treelet + 4 * i;
            *q0 = p[0];
            *q1 = p[1];
            *q2 = p[2];
            const float4 w = p[3];
            *q3 = make_uint4(__float_as_uint(w.x), __float_as_uint(w.y), __float_as_uint(w.z), __float_as_uint(w.w));
        } else {
            const float4* p = g + 4 * (size_t)i;
            *q0 = p[0];
            *q1 = p[1];
            *q2 = p[2];
            *q3 = reinterpret_cast<const uint4*>(p)[3];
        }
    }
    __device__ __forceinline__ uint4 links(uint32_t i) const {
        if (i < top) {
            const float4 w = sl_treelet[4 * i + 3];
            return make_uint4(__float_as_uint(w.x), __float_as_uint(w.y), __float_as_uint(w.z), __float_as_uint(w.w));
        }
        return reinterpret_cast<const uint4*>(g)[4 * (size_t)i + 3];
    }
};

// The end of a stackless step: the backtrack from the current node (record word q3) to the
// next sibling, or the end of the walk.
template <class Src>
__device__ __forceinline__ void stackless_backtrack(StacklessRay& r, const Src& src, uint4 q3) {
    if (r.bitstack == 0) {
        r.node = kInvalid;
        return;
    }
    if (r.stk.n != 0) {
        // the climb the reference makes, one bit per parent, as one shift (bitstack != 0)
        r.bitstack = r.bitstack >> __builtin_ctzll(r.bitstack);
        r.node = enter_link(r.stk.pop());
        r.bitstack = r.bitstack ^ 1;
        return;
    }
    uint4 q = q3;
    while ((r.bitstack & 1) == 0) {
        if (r.bitstack == 0) {
            r.node = kInvalid;
            return;
        }
        r.node = q.z;
        r.bitstack = r.bitstack >> 1;
        if (r.bitstack != 0) q = src.links(r.node);
    }
    r.node = enter_link(q.w);
    r.bitstack = r.bitstack ^ 1;
}

// Visit of a branch node of trace_stackless (descent, or backtrack when neither child is
// entered). Reaching a leaf marks it pending instead of testing its primitives.
template <class Src>
__device__ __forceinline__ void stackless_visit(StacklessRay& r, const Src& src, bool boxes_finite) {
    float4 q0, q1, q2;
    uint4 q3;
    src.record(r.node, &q0, &q1, &q2, &q3);  // the whole record in one round trip
    if (q3.x & kLeafFlag) {
        r.node |= kPendingLeaf;
        return;
    }
    float lmin[3] = {q0.x, q0.y, q0.z}, lmax[3] = {q0.w, q1.x, q1.y};
    float rmin[3] = {q1.z, q1.w, q2.x}, rmax[3] = {q2.y, q2.z, q2.w};
    float Lh0, Lh1, Rh0, Rh1;
    bool L, R;
    slab_pair(lmin, lmax, rmin, rmax, r.ro, r.inv, r.tmin, boxes_finite && r.plain, &L, &R, &Lh0, &Lh1, &Rh0, &Rh1);
    if (L || R) {
        r.bitstack = r.bitstack << 1;
        if (L && R) {
            bool nearL = (Lh0 == Rh0) ? (Lh1 < Rh1) : (Lh0 < Rh0);
            r.node = enter_link(nearL ? q3.x : q3.y);
            r.stk.push(nearL ? q3.y : q3.x);
            r.bitstack = r.bitstack | 1;
        } else {
            r.node = enter_link(L ? q3.x : q3.y);
        }
        return;
    }
    stackless_backtrack(r, src, q3);
}

// The pending leaf of a lane: its primitives, then the backtrack.
template <class Src>
__device__ __forceinline__ void stackless_leaf(StacklessRay& r, const Src& src, const PrimRec* __restrict__ prims,
                                               int32_t sphere_begin) {
    r.node &= ~kPendingLeaf;
    const uint4 q3 = src.links(r.node);
    const uint32_t beg = q3.x & ~kLeafFlag, end = q3.y;
    for (uint32_t s = beg; s < end; ++s) test_prim(prims, s, sphere_begin, r.ro, r.rd, &r.tmin, &r.u, &r.v, &r.prim);
    stackless_backtrack(r, src, q3);
}


// Lane states of the resumable walks: live (a ray in progress), pending (stopped at an entered
// leaf whose primitives are not yet tested).
__device__ __forceinline__ bool walk_live(const StacklessRay& r) { return r.node != kInvalid; }
__device__ __forceinline__ bool walk_pending(const StacklessRay& r) {
    return r.node != kInvalid && (r.node & kPendingLeaf) != 0;
}
__device__ __forceinline__ bool walk_live(const ThreadedRay& r) { return r.node != kInvalid || r.leaf != 0; }
__device__ __forceinline__ bool walk_pending(const ThreadedRay& r) { return r.leaf != 0; }

// One ray per lane with leaves batched across the wave: a lane that enters a leaf waits until
// leaf_batch lanes wait, or at least as many lanes wait as still walk, then the wave tests their
// primitives together. The lane's own visits and tests are those of trace_stackless / trace_threaded.
// Lanes whose ray has ended stay in the loop, masked, until the wave is done.
// kTreelet (stackless): the kernel has staged the first `top` node records in LDS.
template <int kVariant, bool kTreelet = false>
__device__ __forceinline__ int32_t trace_batched(const TraceScene& sc, F3 ro, F3 rd, float* tmin_io, float* uo,
                                                 float* vo, uint32_t leaf_batch, uint32_t top = 0) {
    const SLTreelet tsrc{reinterpret_cast<const float4*>(sc.sl_nodes), top};
    const SLGlobal gsrc{reinterpret_cast<const float4*>(sc.sl_nodes)};
    using State = typename std::conditional<kVariant == kVariantStackless, StacklessRay, ThreadedRay>::type;
    State st;
    if constexpr (kVariant != kVariantStackless)
        threaded_begin<kVariant>(st, sc.mt_nodes, sc.nnodes, ro, rd, *tmin_io);
    else
        stackless_begin(st, ro, rd, *tmin_io);
    const bool finite = sc.boxes_finite != 0;
    for (;;) {
        const bool live = walk_live(st);
        if (__ballot(live) == 0) break;
        const bool pending = walk_pending(st);
        const bool walking = live && !pending;
        const uint32_t npending = (uint32_t)__popcll(__ballot(pending));
        if (npending >= leaf_batch || npending >= (uint32_t)__popcll(__ballot(walking))) {
            if (pending) {
                if constexpr (kVariant != kVariantStackless)
                    threaded_leaf(st, sc.prims, sc.sphere_begin);
                else if constexpr (kTreelet)
                    stackless_leaf(st, tsrc, sc.prims, sc.sphere_begin);
                else
                    stackless_leaf(st, gsrc, sc.prims, sc.sphere_begin);
            }
        } else if (walking) {
            if constexpr (kVariant != kVariantStackless)
                threaded_visit(st, finite);
            else if constexpr (kTreelet)
                stackless_visit(st, tsrc, finite);
            else
                stackless_visit(st, gsrc, finite);
        }
    }
    *tmin_io = st.tmin;
    *uo = st.u;
    *vo = st.v;
    return st.prim;
}

}  // namespace rgpu
