// ORACLE — test infrastructure only. A CPU restatement of the reference's
// RNG and extension-ray-cast path, used as the checker by tests/, by
// __graft_entry__.smoke() and by the cpu_baseline leg of bench.py. Nothing in the
// product links, imports or calls this file.
//
// Parity pinning (see tests/test_oracle_*.py):
//   splitmix64 / xoshiro128**      published known-answer values; the GF(2) jump
//                                  identity (jump == 2^64 steps of the linear engine)
//   slabs / triangle / sphere      closed-form cases and brute force
//   stackless / threaded traversal brute force over every primitive with the same
//                                  triangle routine (the method the reference's own
//                                  BVH checkers use, BVH/bin/data/bvh_groundTruth.hiplc)
//   reference BVH layouts          structural invariants (links, pre-order, bounds)
//
// Built with -ffp-contract=off and IEEE denormals, like the HIP kernels.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- RNG
// kernels/peseudo_random.cl:37-46 (splitmix64)
inline uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
// kernels/peseudo_random.cl:9-23 (xoshiro128**)
inline uint32_t xoshiro_next(uint32_t* s) {
    const uint32_t result = rotl32(s[0] * 5u, 7) * 9u;
    const uint32_t t = s[1] << 9;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl32(s[3], 11);
    return result;
}
// kernels/peseudo_random.cl:25-28
inline float uniform01(uint32_t* s) {
    uint32_t bits = (xoshiro_next(s) >> 9) | 0x3f800000u;
    float f;
    memcpy(&f, &bits, 4);
    return f - 1.0f;
}
// kernels/peseudo_random.cl:34-39
inline uint64_t uniform_integer(uint32_t* s) {
    uint64_t a = xoshiro_next(s) >> 1;
    uint64_t b = xoshiro_next(s) >> 1;
    return (a << 31) | b;
}

// ---------------------------------------------------------------- vector helpers
struct V3 {
    float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float fmn(float a, float b) { return std::fmin(a, b); }
inline float fmx(float a, float b) { return std::fmax(a, b); }

// kernels/extension_ray_cast_stackless.cl:35-50 and kernels/slab.cl:10-21
inline bool slabs_with_hit(V3 p0, V3 p1, V3 ro, V3 inv, float farclip, float* h0, float* h1) {
    V3 t0 = mul(sub(p0, ro), inv);
    V3 t1 = mul(sub(p1, ro), inv);
    // select(t0, -t1, isnan(t0)); then select(t1, -t0, isnan(t1)) with the updated t0
    if (std::isnan(t0.x)) t0.x = -t1.x;
    if (std::isnan(t0.y)) t0.y = -t1.y;
    if (std::isnan(t0.z)) t0.z = -t1.z;
    if (std::isnan(t1.x)) t1.x = -t0.x;
    if (std::isnan(t1.y)) t1.y = -t0.y;
    if (std::isnan(t1.z)) t1.z = -t0.z;
    V3 tmin = {fmn(t0.x, t1.x), fmn(t0.y, t1.y), fmn(t0.z, t1.z)};
    V3 tmax = {fmx(t0.x, t1.x), fmx(t0.y, t1.y), fmx(t0.z, t1.z)};
    float rmin = fmx(fmx(tmin.x, tmin.y), tmin.z);
    float rmax = fmn(fmn(tmax.x, tmax.y), tmax.z);
    *h0 = fmx(0.0f, rmin);
    *h1 = fmn(rmax, farclip);
    return rmin <= rmax && 0.0f <= rmax && rmin <= farclip;
}
inline bool slabs(V3 p0, V3 p1, V3 ro, V3 inv, float farclip) {
    float a, b;
    return slabs_with_hit(p0, p1, ro, inv, farclip, &a, &b);
}

// kernels/intersect_triangle.cl:7-45 (float u, as the GPU kernel computes it)
inline bool intersect_ray_triangle(V3 ro, V3 rd, V3 v0, V3 v1, V3 v2, float* tmin, float* uo, float* vo) {
    const float kEpsilon = 1.0e-8f;
    V3 v0v1 = sub(v1, v0);
    V3 v0v2 = sub(v2, v0);
    V3 pvec = cross(rd, v0v2);
    float det = dot(v0v1, pvec);
    if (std::fabs(det) < kEpsilon) return false;
    float invDet = 1.0f / det;
    V3 tvec = sub(ro, v0);
    float u = dot(tvec, pvec) * invDet;
    if (u < 0.0f || u > 1.0f) return false;
    V3 qvec = cross(tvec, v0v1);
    float v = dot(rd, qvec) * invDet;
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = dot(v0v2, qvec) * invDet;
    if (t < 0.0f) return false;
    if (*tmin < t) return false;
    *tmin = t;
    *uo = u;
    *vo = v;
    return true;
}

// kernels/extension_ray_cast_stackless.cl:52-76
inline bool intersect_sphere(V3 ro, V3 rd, V3 o, float r, float* tmin) {
    float r_sq = r * r;
    V3 S = sub(ro, o);
    V3 crs = cross(S, rd);
    float d_sq = dot(crs, crs);
    float B = dot(S, rd);
    float D = r_sq - d_sq;
    if (D < 0.0f) return false;
    float Ds = std::sqrt(D);
    float t0 = -B - Ds;
    float t1 = -B + Ds;
    if (0.0f < t0 && t0 < *tmin) {
        *tmin = t0;
        return true;
    }
    if (0.0f < t1 && t1 < *tmin) {
        *tmin = t1;
        return true;
    }
    return false;
}

// ---------------------------------------------------------------- scene access
struct Scene {
    const float* points;
    uint32_t stride;
    const uint32_t* indices;
    int32_t sphere_begin;
    const float* spheres;
    V3 pt(uint32_t i) const {
        const float* p = points + (size_t)i * stride;
        return {p[0], p[1], p[2]};
    }
    // Test one primitive; returns true if it became the closest hit.
    bool test(uint32_t prim, V3 ro, V3 rd, float* tmin, float* u, float* v) const {
        if ((int32_t)prim < sphere_begin) {
            const uint32_t* ix = indices + 3 * (size_t)prim;
            return intersect_ray_triangle(ro, rd, pt(ix[0]), pt(ix[1]), pt(ix[2]), tmin, u, v);
        }
        const float* s = spheres + 4 * (size_t)(prim - sphere_begin);
        return intersect_sphere(ro, rd, {s[0], s[1], s[2]}, s[3], tmin);
    }
};

struct Ray {
    float ro[3];
    float tmax;
    float rd[3];
    uint32_t pad;
};
struct Hit {
    int32_t prim;
    float t;
    float u;
    float v;
};
// Per-ray event counts, the basis of the algorithmic-bytes figure.
struct Counters {
    uint64_t branch_visits;  // stackless: branch records read (both child boxes)
    uint64_t leaf_visits;    // stackless: leaf records read
    uint64_t parent_steps;   // stackless: backtrack moves to a parent
    uint64_t prim_tests;     // primitives tested
    uint64_t node_visits;    // threaded: node records read
};

// ---------------------------------------------------------------- reference stackless layout
// common/stackless_bvh.hpp:264-275 with OpenCLFloat3 = 16-byte float3.
struct RefAABB {
    float lower[4];
    float upper[4];
};
struct RefStacklessNode {
    RefAABB L_bounds;
    RefAABB R_bounds;
    uint32_t link_parent;
    uint32_t link_L;
    uint32_t link_R;
    uint32_t link_sibling;
    uint32_t primitive_indices_beg;
    uint32_t primitive_indices_end;
    uint32_t pad[2];
};
static_assert(sizeof(RefStacklessNode) == 96, "reference node is 96 bytes");

// A generic binary tree handed in by the caller.
struct Tree {
    uint32_t nn;
    const int32_t* left;
    const int32_t* right;
    const float* bounds;  // 6 per node: min xyz, max xyz
    const uint32_t* prim_beg;
    const uint32_t* prim_cnt;
    const uint32_t* prim_ids;
    RefAABB box(int32_t n) const {
        RefAABB b;
        for (int k = 0; k < 3; ++k) {
            b.lower[k] = bounds[6 * (size_t)n + k];
            b.upper[k] = bounds[6 * (size_t)n + 3 + k];
        }
        b.lower[3] = b.upper[3] = 0.0f;
        return b;
    }
};

// common/stackless_bvh.hpp:283-292 (allocate_stacklessBVH): pre-order numbering.
void allocate_preorder(const Tree& t, int32_t node, std::vector<uint32_t>* index, uint32_t* size) {
    (*index)[node] = (*size)++;
    if (t.left[node] >= 0) {
        allocate_preorder(t, t.left[node], index, size);
        allocate_preorder(t, t.right[node], index, size);
    }
}
// common/stackless_bvh.hpp:294-327 (build_stacklessBVH)
void build_stackless(const Tree& t, int32_t node, int32_t parent, const std::vector<uint32_t>& index,
                     std::vector<RefStacklessNode>* nodes, std::vector<uint32_t>* prim_ids) {
    uint32_t me = index[node];
    if (parent >= 0) (*nodes)[me].link_parent = index[parent];
    if (t.left[node] >= 0) {
        uint32_t L = index[t.left[node]];
        uint32_t R = index[t.right[node]];
        (*nodes)[me].link_L = L;
        (*nodes)[me].link_R = R;
        (*nodes)[me].L_bounds = t.box(t.left[node]);
        (*nodes)[me].R_bounds = t.box(t.right[node]);
        (*nodes)[L].link_sibling = R;
        (*nodes)[R].link_sibling = L;
        build_stackless(t, t.left[node], node, index, nodes, prim_ids);
        build_stackless(t, t.right[node], node, index, nodes, prim_ids);
    } else {
        (*nodes)[me].primitive_indices_beg = (uint32_t)prim_ids->size();
        for (uint32_t i = 0; i < t.prim_cnt[node]; ++i) prim_ids->push_back(t.prim_ids[t.prim_beg[node] + i]);
        (*nodes)[me].primitive_indices_end = (uint32_t)prim_ids->size();
    }
}

// kernels/extension_ray_cast_stackless.cl:81-200 (intersect_bvh), exact-equality tie rule.
Hit intersect_stackless(const RefStacklessNode* nodes, const uint32_t* prim_ids, const Scene& sc,
                        const Ray& ray, Counters* c) {
    V3 ro = {ray.ro[0], ray.ro[1], ray.ro[2]};
    V3 rd = {ray.rd[0], ray.rd[1], ray.rd[2]};
    float tmin = ray.tmax;
    V3 inv = {1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z};
    float u = 0.0f, v = 0.0f;
    int32_t prim = -1;
    uint64_t bitstack = 0;
    uint32_t node = 0;
    for (;;) {
        const RefStacklessNode& nd = nodes[node];
        bool branch = nd.primitive_indices_end == 0;
        if (branch) {
            c->branch_visits++;
            float Lh0, Lh1, Rh0, Rh1;
            V3 Llo = {nd.L_bounds.lower[0], nd.L_bounds.lower[1], nd.L_bounds.lower[2]};
            V3 Lhi = {nd.L_bounds.upper[0], nd.L_bounds.upper[1], nd.L_bounds.upper[2]};
            V3 Rlo = {nd.R_bounds.lower[0], nd.R_bounds.lower[1], nd.R_bounds.lower[2]};
            V3 Rhi = {nd.R_bounds.upper[0], nd.R_bounds.upper[1], nd.R_bounds.upper[2]};
            bool L = slabs_with_hit(Llo, Lhi, ro, inv, tmin, &Lh0, &Lh1);
            bool R = slabs_with_hit(Rlo, Rhi, ro, inv, tmin, &Rh0, &Rh1);
            if (L || R) {
                bitstack = bitstack << 1;
                if (L && R) {
                    uint32_t near_node;
                    if (Lh0 == Rh0) near_node = Lh1 < Rh1 ? nd.link_L : nd.link_R;
                    else near_node = Lh0 < Rh0 ? nd.link_L : nd.link_R;
                    node = near_node;
                    bitstack = bitstack | 1;
                } else if (L) {
                    node = nd.link_L;
                } else {
                    node = nd.link_R;
                }
                continue;
            }
        } else {
            c->leaf_visits++;
            for (uint32_t i = nd.primitive_indices_beg; i < nd.primitive_indices_end; ++i) {
                c->prim_tests++;
                uint32_t p = prim_ids[i];
                float tu = u, tv = v;
                if (sc.test(p, ro, rd, &tmin, &tu, &tv)) {
                    prim = (int32_t)p;
                    if ((int32_t)p < sc.sphere_begin) {
                        u = tu;
                        v = tv;
                    }
                }
            }
        }
        // backtrack
        while ((bitstack & 1) == 0) {
            if (bitstack == 0) {
                Hit h;
                h.prim = prim;
                h.t = tmin;
                h.u = prim >= 0 ? u : 0.0f;
                h.v = prim >= 0 ? v : 0.0f;
                return h;
            }
            node = nodes[node].link_parent;
            c->parent_steps++;
            bitstack = bitstack >> 1;
        }
        node = nodes[node].link_sibling;
        bitstack = bitstack ^ 1;
    }
}

// ---------------------------------------------------------------- reference threaded (MTBVH) layout
// common/threaded_bvh.hpp:398-404 (node) and kernels/extension_ray_cast.cl:27-39 (links).
struct RefMTBVHNode {
    RefAABB bounds;
    int32_t primitive_indices_beg;
    int32_t primitive_indices_end;
    int32_t pad[2];
};

struct MutTree {
    std::vector<int32_t> L, R;
    const Tree* t;
};

// common/threaded_bvh.hpp:494-516 (sortBranchForMultiThreadedBVH)
void sort_for_direction(MutTree* m, int32_t node, int axis_dir) {
    if (m->L[node] < 0) return;
    int axis = axis_dir / 2;
    RefAABB a = m->t->box(m->L[node]);
    RefAABB b = m->t->box(m->R[node]);
    // AABB_center = mix(lower, upper, 0.5f) (common/stackless_bvh.hpp:26-28)
    float Lc = a.lower[axis] * 0.5f + a.upper[axis] * 0.5f;
    float Rc = b.lower[axis] * 0.5f + b.upper[axis] * 0.5f;
    if (axis_dir & 1) std::swap(Lc, Rc);
    if (!(Lc < Rc)) std::swap(m->L[node], m->R[node]);
    sort_for_direction(m, m->L[node], axis_dir);
    sort_for_direction(m, m->R[node], axis_dir);
}

// common/threaded_bvh.hpp:417-484 (linkMultiThreadedBVH), links stored as
// links[dir][node][hit, miss] like kernels/extension_ray_cast.cl:27-39.
void link_direction(const MutTree& m, int dir, int32_t node, int32_t sibling, int32_t parent_sibling,
                    const std::vector<uint32_t>& index, std::vector<int32_t>* links, uint32_t nn) {
    int32_t* hm = links->data() + (size_t)dir * nn * 2;
    uint32_t me = index[node];
    if (m.L[node] >= 0) {
        hm[2 * me + 0] = (int32_t)index[m.L[node]];
        if (sibling >= 0) hm[2 * me + 1] = (int32_t)index[sibling];
        else hm[2 * me + 1] = parent_sibling >= 0 ? (int32_t)index[parent_sibling] : -1;
        link_direction(m, dir, m.L[node], m.R[node], m.R[node], index, links, nn);
        link_direction(m, dir, m.R[node], -1, parent_sibling, index, links, nn);
    } else {
        int32_t link_node = sibling >= 0 ? sibling : parent_sibling;
        int32_t l = link_node >= 0 ? (int32_t)index[link_node] : -1;
        hm[2 * me + 0] = l;
        hm[2 * me + 1] = l;
    }
}

// common/threaded_bvh.hpp:269-364 (TBVHNode, allocateThreadedBVH, linkThreadedBVH): the single
// threaded table over the build's own child order, stored as the seventh link table. The hit
// link is always the array neighbour (-1 past the end); a branch's miss link is its sibling (L
// or root) or its parent's sibling (R); a leaf's miss link equals its hit link.
void link_single(const Tree& t, int32_t node, int32_t sibling, int32_t parent_sibling,
                 const std::vector<uint32_t>& index, std::vector<int32_t>* links, uint32_t nn) {
    int32_t* hm = links->data() + (size_t)6 * nn * 2;
    uint32_t me = index[node];
    uint32_t neighbor = me + 1;
    hm[2 * me + 0] = neighbor < nn ? (int32_t)neighbor : -1;
    if (t.left[node] >= 0) {
        if (sibling >= 0) hm[2 * me + 1] = (int32_t)index[sibling];
        else hm[2 * me + 1] = parent_sibling >= 0 ? (int32_t)index[parent_sibling] : -1;
        link_single(t, t.left[node], t.right[node], t.right[node], index, links, nn);
        link_single(t, t.right[node], -1, parent_sibling, index, links, nn);
    } else {
        hm[2 * me + 1] = hm[2 * me + 0];
    }
}

// The walk of intersect_mtbvh and intersect_tbvh over one link table hm.
Hit walk_links(const RefMTBVHNode* nodes, const int32_t* hm, const uint32_t* prim_ids, const Scene& sc,
               const Ray& ray, Counters* c) {
    V3 ro = {ray.ro[0], ray.ro[1], ray.ro[2]};
    V3 rd = {ray.rd[0], ray.rd[1], ray.rd[2]};
    V3 inv = {1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z};
    float tmin = ray.tmax;
    float u = 0.0f, v = 0.0f;
    int32_t prim = -1;
    int32_t node = 0;
    while (0 <= node) {
        c->node_visits++;
        const RefMTBVHNode& nd = nodes[node];
        V3 lo = {nd.bounds.lower[0], nd.bounds.lower[1], nd.bounds.lower[2]};
        V3 hi = {nd.bounds.upper[0], nd.bounds.upper[1], nd.bounds.upper[2]};
        if (slabs(lo, hi, ro, inv, tmin)) {
            for (int32_t i = nd.primitive_indices_beg; i < nd.primitive_indices_end; ++i) {
                c->prim_tests++;
                uint32_t p = prim_ids[i];
                float tu = u, tv = v;
                if (sc.test(p, ro, rd, &tmin, &tu, &tv)) {
                    prim = (int32_t)p;
                    if ((int32_t)p < sc.sphere_begin) {
                        u = tu;
                        v = tv;
                    }
                }
            }
            node = hm[2 * node + 0];
        } else {
            node = hm[2 * node + 1];
        }
    }
    Hit h;
    h.prim = prim;
    h.t = tmin;
    h.u = prim >= 0 ? u : 0.0f;
    h.v = prim >= 0 ? v : 0.0f;
    return h;
}

// common/threaded_bvh.hpp:527-569 (intersect_mtbvh) and kernels/extension_ray_cast.cl:44-94,
// with the incoming far clip honoured (the GPU stackless kernel reads it from ExtensionResult.tmin).
Hit intersect_mtbvh(const RefMTBVHNode* nodes, const int32_t* links, uint32_t nn, const uint32_t* prim_ids,
                    const Scene& sc, const Ray& ray, Counters* c) {
    float ax = std::fabs(ray.rd[0]), ay = std::fabs(ray.rd[1]), az = std::fabs(ray.rd[2]);
    float maxYZ = fmx(ay, az);
    float maxXZ = fmx(ax, az);
    int dir;
    if (maxYZ < ax) dir = 0.0f < ray.rd[0] ? 0 : 1;
    else if (maxXZ < ay) dir = 0.0f < ray.rd[1] ? 2 : 3;
    else dir = 0.0f < ray.rd[2] ? 4 : 5;
    return walk_links(nodes, links + (size_t)dir * nn * 2, prim_ids, sc, ray, c);
}

// common/threaded_bvh.hpp:366-395 (intersect_tbvh): the same walk over the single table, for
// every ray, with the incoming far clip honoured.
Hit intersect_tbvh(const RefMTBVHNode* nodes, const int32_t* links, uint32_t nn, const uint32_t* prim_ids,
                   const Scene& sc, const Ray& ray, Counters* c) {
    return walk_links(nodes, links + (size_t)6 * nn * 2, prim_ids, sc, ray, c);
}

Hit intersect_brute(uint32_t nprims, const Scene& sc, const Ray& ray, Counters* c) {
    V3 ro = {ray.ro[0], ray.ro[1], ray.ro[2]};
    V3 rd = {ray.rd[0], ray.rd[1], ray.rd[2]};
    float tmin = ray.tmax;
    float u = 0.0f, v = 0.0f;
    int32_t prim = -1;
    for (uint32_t p = 0; p < nprims; ++p) {
        c->prim_tests++;
        float tu = u, tv = v;
        if (sc.test(p, ro, rd, &tmin, &tu, &tv)) {
            prim = (int32_t)p;
            if ((int32_t)p < sc.sphere_begin) {
                u = tu;
                v = tv;
            }
        }
    }
    Hit h;
    h.prim = prim;
    h.t = tmin;
    h.u = prim >= 0 ? u : 0.0f;
    h.v = prim >= 0 ? v : 0.0f;
    return h;
}

template <class F>
void parallel_rays(uint32_t nrays, int nthreads, F f) {
    if (nthreads < 1) nthreads = 1;
    if ((uint32_t)nthreads > nrays) nthreads = nrays > 0 ? (int)nrays : 1;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) {
        th.emplace_back([=]() {
            uint32_t b = (uint32_t)((uint64_t)nrays * t / nthreads);
            uint32_t e = (uint32_t)((uint64_t)nrays * (t + 1) / nthreads);
            f(b, e, t);
        });
    }
    for (auto& x : th) x.join();
}

void sum_counters(const std::vector<Counters>& per, uint64_t* out) {
    if (!out) return;
    Counters s = {};
    for (auto& c : per) {
        s.branch_visits += c.branch_visits;
        s.leaf_visits += c.leaf_visits;
        s.parent_steps += c.parent_steps;
        s.prim_tests += c.prim_tests;
        s.node_visits += c.node_visits;
    }
    out[0] = s.branch_visits;
    out[1] = s.leaf_visits;
    out[2] = s.parent_steps;
    out[3] = s.prim_tests;
    out[4] = s.node_visits;
}

}  // namespace

extern "C" {

// ---- RNG
void orc_splitmix64_seq(uint64_t seed, uint32_t n, uint64_t* out) {
    uint64_t x = seed;
    for (uint32_t i = 0; i < n; ++i) out[i] = splitmix64(&x);
}
// kernels/peseudo_random.cl:48-64 (random_initialize): state of slot gid from
// splitmix64 seeded with seed_offset + gid.
void orc_random_initialize(uint64_t seed_offset, uint32_t n, uint32_t* states) {
    for (uint32_t gid = 0; gid < n; ++gid) {
        uint64_t x = seed_offset + gid;
        uint64_t r0 = splitmix64(&x);
        uint64_t r1 = splitmix64(&x);
        uint32_t* s = states + 4 * (size_t)gid;
        s[0] = (uint32_t)(r0 & 0xFFFFFFFFu);
        s[1] = (uint32_t)(r0 >> 32);
        s[2] = (uint32_t)(r1 & 0xFFFFFFFFu);
        s[3] = (uint32_t)(r1 >> 32);
        if ((s[0] | s[1] | s[2] | s[3]) == 0) s[0] = 1;
    }
}
void orc_xoshiro_next(uint32_t* state4, uint32_t n, uint32_t* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = xoshiro_next(state4);
}
void orc_xoshiro_uniform(uint32_t* state4, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = uniform01(state4);
}
void orc_xoshiro_uniform_integer(uint32_t* state4, uint32_t n, uint64_t* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = uniform_integer(state4);
}
// common/share/peseudo_random.hpp:93-116 (jump): published jump polynomial.
void orc_xoshiro_jump(uint32_t* s) {
    static const uint32_t JUMP[] = {0x8764000b, 0xf542d2d3, 0x6fa035c3, 0x77f2db5b};
    uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int i = 0; i < 4; i++)
        for (int b = 0; b < 32; b++) {
            if (JUMP[i] & (1u << b)) {
                s0 ^= s[0];
                s1 ^= s[1];
                s2 ^= s[2];
                s3 ^= s[3];
            }
            xoshiro_next(s);
        }
    s[0] = s0;
    s[1] = s1;
    s[2] = s2;
    s[3] = s3;
}

// ---- geometry primitives (for closed-form tests)
int orc_slabs_with_hit(const float* p0, const float* p1, const float* ro, const float* rd, float farclip,
                       float* h0h1) {
    V3 inv = {1.0f / rd[0], 1.0f / rd[1], 1.0f / rd[2]};
    return slabs_with_hit({p0[0], p0[1], p0[2]}, {p1[0], p1[1], p1[2]}, {ro[0], ro[1], ro[2]}, inv, farclip,
                          &h0h1[0], &h0h1[1])
               ? 1
               : 0;
}
int orc_intersect_triangle(const float* ro, const float* rd, const float* v0, const float* v1, const float* v2,
                           float* tuv /* in: tuv[0]=tmin; out: t,u,v */) {
    float u = 0, v = 0;
    bool h = intersect_ray_triangle({ro[0], ro[1], ro[2]}, {rd[0], rd[1], rd[2]}, {v0[0], v0[1], v0[2]},
                                    {v1[0], v1[1], v1[2]}, {v2[0], v2[1], v2[2]}, &tuv[0], &u, &v);
    if (h) {
        tuv[1] = u;
        tuv[2] = v;
    }
    return h ? 1 : 0;
}

// ---- reference stackless layout from a generic tree (root = node 0 of the input)
// out_nodes: nn * 96 bytes; out_prim_ids: total leaf slots.
int orc_stackless_build(uint32_t nn, const int32_t* left, const int32_t* right, const float* bounds,
                        const uint32_t* prim_beg, const uint32_t* prim_cnt, const uint32_t* prim_ids,
                        void* out_nodes, uint32_t* out_prim_ids) {
    if (nn == 0) return 0;
    Tree t{nn, left, right, bounds, prim_beg, prim_cnt, prim_ids};
    std::vector<uint32_t> index(nn, 0);
    uint32_t size = 0;
    allocate_preorder(t, 0, &index, &size);
    if (size != nn) return -1;
    std::vector<RefStacklessNode> nodes(nn);
    memset(nodes.data(), 0, nn * sizeof(RefStacklessNode));
    std::vector<uint32_t> pids;
    build_stackless(t, 0, -1, index, &nodes, &pids);
    memcpy(out_nodes, nodes.data(), nn * sizeof(RefStacklessNode));
    memcpy(out_prim_ids, pids.data(), pids.size() * sizeof(uint32_t));
    return (int)pids.size();
}

// rays: 8 floats each (ro, tmax, rd, pad); hits: 4 words each (prim, t, u, v)
// counters (optional): 5 totals {branch_visits, leaf_visits, parent_steps, prim_tests, node_visits}
int orc_stackless_intersect(const void* nodes, const uint32_t* prim_ids, const float* points, uint32_t stride,
                            const uint32_t* indices, int32_t sphere_begin, const float* spheres, uint32_t nrays,
                            const void* rays, void* hits, uint64_t* counters, int nthreads) {
    Scene sc{points, stride, indices, sphere_begin, spheres};
    const RefStacklessNode* nd = (const RefStacklessNode*)nodes;
    const Ray* r = (const Ray*)rays;
    Hit* h = (Hit*)hits;
    std::vector<Counters> per(nthreads > 0 ? nthreads : 1, Counters{});
    parallel_rays(nrays, nthreads, [&](uint32_t b, uint32_t e, int t) {
        Counters c = {};
        for (uint32_t i = b; i < e; ++i) h[i] = intersect_stackless(nd, prim_ids, sc, r[i], &c);
        per[t] = c;
    });
    sum_counters(per, counters);
    return 0;
}

// ---- reference threaded (MTBVH) layout from a generic tree.
// out_nodes: nn * 48 bytes (RefMTBVHNode); out_links: 7 * nn * 2 int32 (the six direction
// classes, then the single threaded table); out_prim_ids: leaf slots.
int orc_mtbvh_build(uint32_t nn, const int32_t* left, const int32_t* right, const float* bounds,
                    const uint32_t* prim_beg, const uint32_t* prim_cnt, const uint32_t* prim_ids,
                    void* out_nodes, int32_t* out_links, uint32_t* out_prim_ids) {
    if (nn == 0) return 0;
    Tree t{nn, left, right, bounds, prim_beg, prim_cnt, prim_ids};
    // common/threaded_bvh.hpp:406-415 (allocateMultiThreadedBVH): pre-order indices
    std::vector<uint32_t> index(nn, 0);
    uint32_t size = 0;
    allocate_preorder(t, 0, &index, &size);
    if (size != nn) return -1;
    std::vector<RefMTBVHNode> nodes(nn);
    memset(nodes.data(), 0, nn * sizeof(RefMTBVHNode));
    // Node bounds: the root holds the union of its children (threaded_bvh.hpp:428-433);
    // every other node holds the box its parent stored for it, i.e. its own box.
    for (uint32_t n = 0; n < nn; ++n) nodes[index[n]].bounds = t.box(n);
    if (left[0] >= 0) {
        RefAABB a = t.box(left[0]), b = t.box(right[0]);
        for (int k = 0; k < 3; ++k) {
            nodes[0].bounds.lower[k] = std::min(a.lower[k], b.lower[k]);
            nodes[0].bounds.upper[k] = std::max(a.upper[k], b.upper[k]);
        }
    }
    // Primitive ranges: one concatenation in pre-order (the reference re-appends the
    // same ranges for each direction; the ranges it keeps are equivalent).
    std::vector<uint32_t> pids;
    std::vector<int32_t> order(nn);
    for (uint32_t n = 0; n < nn; ++n) order[index[n]] = (int32_t)n;
    for (uint32_t i = 0; i < nn; ++i) {
        int32_t n = order[i];
        if (left[n] < 0) {
            nodes[i].primitive_indices_beg = (int32_t)pids.size();
            for (uint32_t k = 0; k < prim_cnt[n]; ++k) pids.push_back(prim_ids[prim_beg[n] + k]);
            nodes[i].primitive_indices_end = (int32_t)pids.size();
        }
    }
    MutTree m;
    m.t = &t;
    m.L.assign(left, left + nn);
    m.R.assign(right, right + nn);
    std::vector<int32_t> links((size_t)7 * nn * 2, -1);
    for (int dir = 0; dir < 6; ++dir) {
        sort_for_direction(&m, 0, dir);
        link_direction(m, dir, 0, -1, -1, index, &links, nn);
    }
    link_single(t, 0, -1, -1, index, &links, nn);
    memcpy(out_nodes, nodes.data(), nn * sizeof(RefMTBVHNode));
    memcpy(out_links, links.data(), links.size() * sizeof(int32_t));
    memcpy(out_prim_ids, pids.data(), pids.size() * sizeof(uint32_t));
    return (int)pids.size();
}

int orc_mtbvh_intersect(const void* nodes, const int32_t* links, uint32_t nn, const uint32_t* prim_ids,
                        const float* points, uint32_t stride, const uint32_t* indices, int32_t sphere_begin,
                        const float* spheres, uint32_t nrays, const void* rays, void* hits, uint64_t* counters,
                        int nthreads) {
    Scene sc{points, stride, indices, sphere_begin, spheres};
    const RefMTBVHNode* nd = (const RefMTBVHNode*)nodes;
    const Ray* r = (const Ray*)rays;
    Hit* h = (Hit*)hits;
    std::vector<Counters> per(nthreads > 0 ? nthreads : 1, Counters{});
    parallel_rays(nrays, nthreads, [&](uint32_t b, uint32_t e, int t) {
        Counters c = {};
        for (uint32_t i = b; i < e; ++i) h[i] = intersect_mtbvh(nd, links, nn, prim_ids, sc, r[i], &c);
        per[t] = c;
    });
    sum_counters(per, counters);
    return 0;
}

int orc_tbvh_intersect(const void* nodes, const int32_t* links, uint32_t nn, const uint32_t* prim_ids,
                       const float* points, uint32_t stride, const uint32_t* indices, int32_t sphere_begin,
                       const float* spheres, uint32_t nrays, const void* rays, void* hits, uint64_t* counters,
                       int nthreads) {
    Scene sc{points, stride, indices, sphere_begin, spheres};
    const RefMTBVHNode* nd = (const RefMTBVHNode*)nodes;
    const Ray* r = (const Ray*)rays;
    Hit* h = (Hit*)hits;
    std::vector<Counters> per(nthreads > 0 ? nthreads : 1, Counters{});
    parallel_rays(nrays, nthreads, [&](uint32_t b, uint32_t e, int t) {
        Counters c = {};
        for (uint32_t i = b; i < e; ++i) h[i] = intersect_tbvh(nd, links, nn, prim_ids, sc, r[i], &c);
        per[t] = c;
    });
    sum_counters(per, counters);
    return 0;
}

int orc_bruteforce_intersect(uint32_t nprims, const float* points, uint32_t stride, const uint32_t* indices,
                             int32_t sphere_begin, const float* spheres, uint32_t nrays, const void* rays,
                             void* hits, int nthreads) {
    Scene sc{points, stride, indices, sphere_begin, spheres};
    const Ray* r = (const Ray*)rays;
    Hit* h = (Hit*)hits;
    parallel_rays(nrays, nthreads, [&](uint32_t b, uint32_t e, int) {
        Counters c = {};
        for (uint32_t i = b; i < e; ++i) h[i] = intersect_brute(nprims, sc, r[i], &c);
    });
    return 0;
}

// Geometric normal of a hit: CW winding normalize(cross(v2-v1, v1-v0))
// (kernels/extension_ray_cast_stackless.cl:224-239); spheres: radial.
void orc_hit_normals(uint32_t nhits, const void* rays, const void* hits, const float* points, uint32_t stride,
                     const uint32_t* indices, int32_t sphere_begin, const float* spheres, float* out_ng) {
    Scene sc{points, stride, indices, sphere_begin, spheres};
    const Ray* r = (const Ray*)rays;
    const Hit* h = (const Hit*)hits;
    for (uint32_t i = 0; i < nhits; ++i) {
        float* o = out_ng + 3 * (size_t)i;
        if (h[i].prim < 0) {
            o[0] = o[1] = o[2] = 0.0f;
            continue;
        }
        V3 n;
        if (h[i].prim < sphere_begin) {
            const uint32_t* ix = indices + 3 * (size_t)h[i].prim;
            V3 v0 = sc.pt(ix[0]), v1 = sc.pt(ix[1]), v2 = sc.pt(ix[2]);
            n = cross(sub(v2, v1), sub(v1, v0));
        } else {
            const float* s = spheres + 4 * (size_t)(h[i].prim - sphere_begin);
            V3 ro = {r[i].ro[0], r[i].ro[1], r[i].ro[2]};
            V3 rd = {r[i].rd[0], r[i].rd[1], r[i].rd[2]};
            V3 p = {ro.x + rd.x * h[i].t, ro.y + rd.y * h[i].t, ro.z + rd.z * h[i].t};
            n = sub(p, {s[0], s[1], s[2]});
        }
        float inv = 1.0f / std::sqrt(dot(n, n));
        o[0] = n.x * inv;
        o[1] = n.y * inv;
        o[2] = n.z * inv;
    }
}

}  // extern "C"
