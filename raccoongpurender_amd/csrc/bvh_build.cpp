// Binned-SAH BVH2 builder and the packers for the device node layouts.
//
// The reference builds its tree with Embree's rtcBuildBVH (quality HIGH, binary,
// leaves of 1..5 primitives; common/stackless_bvh.hpp:108-130). This library does not
// use Embree: this is a top-down binned SAH builder (16 bins per axis, leaves of
// at most 4 primitives) with a depth cap so that every leaf fits the 64-bit
// bitstack of the stackless traversal. Traversal results are defined on whatever
// tree is built, so the tree itself carries no parity obligation; it is checked by
// invariants (each primitive in exactly one leaf, children inside parents, depth).
#include "bvh_build.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <deque>
#include <future>
#include <memory>
#include <mutex>
#include <stdexcept>

namespace rgpu {
namespace {

struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    double area() const {
        double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct TNode {
    Box box;
    TNode* l = nullptr;
    TNode* r = nullptr;
    uint32_t beg = 0, cnt = 0;
    uint32_t subtree = 1;  // node count of the subtree rooted here
};

struct Arena {
    std::mutex mu;
    std::vector<std::unique_ptr<std::deque<TNode>>> pools;
    std::deque<TNode>* make() {
        std::lock_guard<std::mutex> g(mu);
        pools.emplace_back(new std::deque<TNode>());
        return pools.back().get();
    }
};

struct Builder {
    const BuildOptions& opt;
    std::vector<Box> pbox;
    std::vector<float> cen;  // 3 per primitive
    std::vector<uint32_t> ids;
    Arena arena;

    explicit Builder(const BuildOptions& o) : opt(o) {}

    static uint32_t levels_needed(uint32_t n, uint32_t max_leaf) {
        uint32_t leaves = (n + max_leaf - 1) / max_leaf;
        uint32_t l = 0;
        while ((1u << l) < leaves) ++l;
        return l;
    }

    TNode* build(uint32_t beg, uint32_t end, uint32_t depth, std::deque<TNode>* pool,
                 uint32_t* max_depth, uint32_t* leaves, double* sah_sum) {
        pool->emplace_back();
        TNode* node = &pool->back();
        const uint32_t n = end - beg;
        Box cb;
        for (uint32_t i = beg; i < end; ++i) {
            node->box.grow(pbox[ids[i]]);
            cb.grow(&cen[3 * (size_t)ids[i]]);
        }
        const double parea = node->box.area();

        auto make_leaf = [&]() {
            node->beg = beg;
            node->cnt = n;
            *max_depth = std::max(*max_depth, depth);
            *leaves += 1;
            *sah_sum += parea * n;
            return node;
        };
        if (n == 1) return make_leaf();

        // Binned SAH over the three centroid axes.
        const uint32_t B = opt.bins;
        int best_axis = -1;
        uint32_t best_bin = 0;
        double best_cost = DBL_MAX;
        bool force_median = depth + levels_needed(n, opt.max_leaf) + 1 >= opt.max_depth;
        if (!force_median) {
            std::vector<Box> bb(B);
            std::vector<uint32_t> bc(B);
            std::vector<double> lA(B), rA(B);
            std::vector<uint32_t> lN(B), rN(B);
            for (int k = 0; k < 3; ++k) {
                float ext = cb.hi[k] - cb.lo[k];
                if (!(ext > 0.0f)) continue;
                float scale = (float)B / ext;
                std::fill(bb.begin(), bb.end(), Box());
                std::fill(bc.begin(), bc.end(), 0u);
                for (uint32_t i = beg; i < end; ++i) {
                    uint32_t p = ids[i];
                    int b = (int)((cen[3 * (size_t)p + k] - cb.lo[k]) * scale);
                    b = std::min(std::max(b, 0), (int)B - 1);
                    bb[b].grow(pbox[p]);
                    bc[b]++;
                }
                Box acc;
                uint32_t cnt = 0;
                for (uint32_t b = 0; b < B; ++b) {
                    acc.grow(bb[b]);
                    cnt += bc[b];
                    lA[b] = acc.area();
                    lN[b] = cnt;
                }
                acc = Box();
                cnt = 0;
                for (int b = (int)B - 1; b >= 0; --b) {
                    acc.grow(bb[b]);
                    cnt += bc[b];
                    rA[b] = acc.area();
                    rN[b] = cnt;
                }
                for (uint32_t b = 0; b + 1 < B; ++b) {
                    if (lN[b] == 0 || rN[b + 1] == 0) continue;
                    double c = lA[b] * lN[b] + rA[b + 1] * rN[b + 1];
                    if (c < best_cost) {
                        best_cost = c;
                        best_axis = k;
                        best_bin = b;
                    }
                }
            }
        }

        uint32_t mid;
        if (best_axis >= 0) {
            double split_cost = opt.traversal_cost + (parea > 0.0 ? best_cost / parea : (double)n);
            if (n <= opt.max_leaf && (double)n <= split_cost) return make_leaf();
            const int k = best_axis;
            const float scale = (float)B / (cb.hi[k] - cb.lo[k]);
            const float lo = cb.lo[k];
            // stable, so that the order of a leaf's primitives is a function of the input alone
            // (the device builder partitions the same way)
            auto it = std::stable_partition(ids.begin() + beg, ids.begin() + end, [&](uint32_t p) {
                int b = (int)((cen[3 * (size_t)p + k] - lo) * scale);
                b = std::min(std::max(b, 0), (int)B - 1);
                return (uint32_t)b <= best_bin;
            });
            mid = (uint32_t)(it - ids.begin());
        } else {
            if (n <= opt.max_leaf) return make_leaf();
            // Object median along the widest centroid axis (also the depth guard).
            int k = 0;
            float w = -1.0f;
            for (int a = 0; a < 3; ++a) {
                float e = cb.hi[a] - cb.lo[a];
                if (e > w) { w = e; k = a; }
            }
            // sorted by (centroid, id), a total order: both halves come out in one defined order
            mid = beg + n / 2;
            std::sort(ids.begin() + beg, ids.begin() + end, [&](uint32_t x, uint32_t y) {
                float cx = cen[3 * (size_t)x + k], cy = cen[3 * (size_t)y + k];
                return cx < cy || (cx == cy && x < y);
            });
        }
        if (mid == beg || mid == end) mid = beg + n / 2;  // defensive: never an empty child

        *sah_sum += parea;  // traversal cost of this branch
        if (n > 65536 && depth < 8) {
            std::deque<TNode>* lpool = arena.make();
            uint32_t ldepth = 0, lleaves = 0;
            double lsah = 0.0;
            auto fut = std::async(std::launch::async, [&]() {
                return build(beg, mid, depth + 1, lpool, &ldepth, &lleaves, &lsah);
            });
            node->r = build(mid, end, depth + 1, pool, max_depth, leaves, sah_sum);
            node->l = fut.get();
            *max_depth = std::max(*max_depth, ldepth);
            *leaves += lleaves;
            *sah_sum += lsah;
        } else {
            node->l = build(beg, mid, depth + 1, pool, max_depth, leaves, sah_sum);
            node->r = build(mid, end, depth + 1, pool, max_depth, leaves, sah_sum);
        }
        node->subtree = 1 + node->l->subtree + node->r->subtree;
        return node;
    }
};

void flatten(const TNode* n, const std::vector<uint32_t>& ids, BVHTree* out) {
    // Iterative pre-order walk: index = position of first visit.
    std::vector<const TNode*> stack;
    stack.push_back(n);
    while (!stack.empty()) {
        const TNode* t = stack.back();
        stack.pop_back();
        BuildNode b;
        for (int k = 0; k < 3; ++k) {
            b.bmin[k] = t->box.lo[k];
            b.bmax[k] = t->box.hi[k];
        }
        uint32_t me = (uint32_t)out->nodes.size();
        if (t->l) {
            b.left = (int32_t)(me + 1);
            b.right = (int32_t)(me + 1 + t->l->subtree);
            b.prim_beg = 0;
            b.prim_cnt = 0;
            out->nodes.push_back(b);
            stack.push_back(t->r);
            stack.push_back(t->l);
        } else {
            b.left = -1;
            b.right = -1;
            b.prim_beg = (uint32_t)out->prim_ids.size();
            b.prim_cnt = t->cnt;
            for (uint32_t i = 0; i < t->cnt; ++i) out->prim_ids.push_back(ids[t->beg + i]);
            out->nodes.push_back(b);
        }
    }
}

}  // namespace

void build_bvh(const float* points, uint32_t stride, uint32_t npoints,
               const uint32_t* indices, uint32_t ntris,
               const float* spheres, uint32_t nspheres,
               const BuildOptions& opt, BVHTree* out) {
    if (opt.max_leaf < 1 || opt.max_leaf > 8) throw std::invalid_argument("max_leaf must be in 1..8");
    if (opt.bins < 2 || opt.bins > 256) throw std::invalid_argument("bins must be in 2..256");
    if (opt.max_depth < 1 || opt.max_depth > 64) throw std::invalid_argument("max_depth must be in 1..64");
    *out = BVHTree();
    const uint32_t n = ntris + nspheres;
    if (n == 0) return;
    Builder b(opt);
    b.pbox.resize(n);
    b.cen.resize(3 * (size_t)n);
    b.ids.resize(n);
    for (uint32_t i = 0; i < ntris; ++i) {
        Box bx;
        for (int j = 0; j < 3; ++j) {
            uint32_t v = indices[3 * (size_t)i + j];
            if (v >= npoints) throw std::invalid_argument("triangle index out of range");
            bx.grow(points + (size_t)v * stride);
        }
        for (int k = 0; k < 3; ++k) {  // -0 as +0: box planes do not depend on which corner came first
            bx.lo[k] += 0.0f;
            bx.hi[k] += 0.0f;
        }
        b.pbox[i] = bx;
    }
    for (uint32_t s = 0; s < nspheres; ++s) {
        const float* sp = spheres + 4 * (size_t)s;
        Box bx;
        for (int k = 0; k < 3; ++k) {
            bx.lo[k] = (sp[k] - sp[3]) + 0.0f;
            bx.hi[k] = (sp[k] + sp[3]) + 0.0f;
        }
        b.pbox[ntris + s] = bx;
    }
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) b.cen[3 * (size_t)i + k] = 0.5f * (b.pbox[i].lo[k] + b.pbox[i].hi[k]);
        b.ids[i] = i;
    }
    std::deque<TNode>* pool = b.arena.make();
    uint32_t depth = 0, leaves = 0;
    double sah = 0.0;
    TNode* root = b.build(0, n, 0, pool, &depth, &leaves, &sah);
    out->nodes.reserve(root->subtree);
    out->prim_ids.reserve(n);
    flatten(root, b.ids, out);
    out->depth = depth;
    out->leaf_count = leaves;
    double ra = root->box.area();
    out->sah_cost = ra > 0.0 ? sah / ra : 0.0;
}

void pack_stackless(const BVHTree& t, std::vector<SLNode>* out) {
    const size_t nn = t.nodes.size();
    out->assign(nn, SLNode());
    for (size_t i = 0; i < nn; ++i) {
        SLNode& s = (*out)[i];
        std::memset(&s, 0, sizeof(s));
    }
    for (size_t i = 0; i < nn; ++i) {
        const BuildNode& b = t.nodes[i];
        SLNode& s = (*out)[i];
        if (b.left >= 0) {
            const BuildNode& L = t.nodes[b.left];
            const BuildNode& R = t.nodes[b.right];
            for (int k = 0; k < 3; ++k) {
                s.lmin[k] = L.bmin[k];
                s.lmax[k] = L.bmax[k];
                s.rmin[k] = R.bmin[k];
                s.rmax[k] = R.bmax[k];
            }
            s.a = (uint32_t)b.left;
            s.b = (uint32_t)b.right;
            (*out)[b.left].parent = (uint32_t)i;
            (*out)[b.right].parent = (uint32_t)i;
            (*out)[b.left].sibling = (uint32_t)b.right;
            (*out)[b.right].sibling = (uint32_t)b.left;
        } else {
            s.a = kLeafFlag | b.prim_beg;
            s.b = b.prim_beg + b.prim_cnt;
        }
    }
}

void number_top_breadth_first(std::vector<SLNode>* nodes, uint32_t top) {
    std::vector<SLNode>& sl = *nodes;
    const size_t nn = sl.size();
    if (nn == 0 || top <= 1) return;
    std::vector<uint32_t> order;
    order.reserve(top);
    order.push_back(0);
    for (size_t h = 0; h < order.size() && order.size() < top; ++h) {
        const SLNode& s = sl[order[h]];
        if (s.a & kLeafFlag) continue;
        order.push_back(s.a);
        if (order.size() < top) order.push_back(s.b);
    }
    const uint32_t none = 0xFFFFFFFFu;
    std::vector<uint32_t> id(nn, none);
    for (size_t k = 0; k < order.size(); ++k) id[order[k]] = (uint32_t)k;
    uint32_t next = (uint32_t)order.size();
    for (size_t i = 0; i < nn; ++i)
        if (id[i] == none) id[i] = next++;
    std::vector<SLNode> out(nn);
    for (size_t i = 0; i < nn; ++i) {
        SLNode s = sl[i];
        if (!(s.a & kLeafFlag)) {
            s.a = id[s.a];
            s.b = id[s.b];
        }
        s.parent = id[s.parent];
        s.sibling = id[s.sibling];
        out[id[i]] = s;
    }
    sl.swap(out);
}

void mark_leaf_links(std::vector<SLNode>* nodes) {
    std::vector<SLNode>& sl = *nodes;
    auto leaf = [&](uint32_t i) { return (sl[i].a & kLeafFlag) != 0; };
    for (size_t i = 0; i < sl.size(); ++i) {
        SLNode& s = sl[i];
        if (!(s.a & kLeafFlag)) {
            if (leaf(s.a)) s.a |= kChildLeaf;
            if (leaf(s.b)) s.b |= kChildLeaf;
        }
        if (i != 0 && leaf(s.sibling)) s.sibling |= kChildLeaf;  // the root has no sibling
    }
}

void pack_threaded(const BVHTree& t, std::vector<MTNode>* out) {
    const uint32_t nn = (uint32_t)t.nodes.size();
    out->assign((size_t)kMTOrders * nn, MTNode());
    if (nn == 0) return;
    // Mutable child order; the six direction sorts are applied one after another to
    // the same tree, as common/threaded_bvh.hpp:518-525 does, so a child pair whose
    // centres tie is swapped by every sort. The single table keeps the build's order.
    std::vector<int32_t> L(nn), R(nn);
    for (uint32_t i = 0; i < nn; ++i) {
        L[i] = t.nodes[i].left;
        R[i] = t.nodes[i].right;
    }
    std::vector<uint32_t> size(nn, 1);
    for (int32_t i = (int32_t)nn - 1; i >= 0; --i)
        if (L[i] >= 0) size[i] = 1 + size[L[i]] + size[R[i]];
    std::vector<std::pair<uint32_t, uint32_t>> stack;  // (node, position in this direction's order)
    for (int dir = 0; dir < kMTOrders; ++dir) {
        if (dir < 6) {
            const int axis = dir / 2;
            const bool negative = (dir & 1) != 0;
            for (uint32_t i = 0; i < nn; ++i) {
                if (L[i] < 0) continue;
                const BuildNode& a = t.nodes[L[i]];
                const BuildNode& b = t.nodes[R[i]];
                float lc = 0.5f * (a.bmin[axis] + a.bmax[axis]);
                float rc = 0.5f * (b.bmin[axis] + b.bmax[axis]);
                if (negative) std::swap(lc, rc);
                if (!(lc < rc)) std::swap(L[i], R[i]);
            }
        } else {
            for (uint32_t i = 0; i < nn; ++i) {
                L[i] = t.nodes[i].left;
                R[i] = t.nodes[i].right;
            }
        }
        MTNode* tab = out->data() + (size_t)dir * nn;
        stack.clear();
        stack.push_back({0u, 0u});
        while (!stack.empty()) {
            auto [node, pos] = stack.back();
            stack.pop_back();
            const BuildNode& b = t.nodes[node];
            MTNode& m = tab[pos];
            for (int k = 0; k < 3; ++k) {
                m.bmin[k] = b.bmin[k];
                m.bmax[k] = b.bmax[k];
            }
            uint32_t next = pos + size[node];
            m.miss = next < nn ? next : kInvalid;
            if (L[node] >= 0) {
                m.leaf = 0;
                stack.push_back({(uint32_t)R[node], pos + 1 + size[L[node]]});
                stack.push_back({(uint32_t)L[node], pos + 1});
            } else {
                m.leaf = kLeafFlag | ((b.prim_cnt - 1) << 28) | b.prim_beg;
            }
        }
    }
}

void pack_prims(const BVHTree& t, const float* points, uint32_t stride,
                const uint32_t* indices, uint32_t ntris,
                const float* spheres, uint32_t nspheres, std::vector<PrimRec>* out) {
    (void)nspheres;
    out->assign(t.prim_ids.size(), PrimRec());
    for (size_t s = 0; s < t.prim_ids.size(); ++s) {
        uint32_t p = t.prim_ids[s];
        PrimRec& r = (*out)[s];
        std::memset(&r, 0, sizeof(r));
        r.prim = p;
        if (p < ntris) {
            const float* v0 = points + (size_t)indices[3 * (size_t)p + 0] * stride;
            const float* v1 = points + (size_t)indices[3 * (size_t)p + 1] * stride;
            const float* v2 = points + (size_t)indices[3 * (size_t)p + 2] * stride;
            float e1[3], e2[3];
            for (int k = 0; k < 3; ++k) {
                e1[k] = v1[k] - v0[k];
                e2[k] = v2[k] - v0[k];
            }
            r.q0[0] = v0[0]; r.q0[1] = v0[1]; r.q0[2] = v0[2]; r.q0[3] = e1[0];
            r.q1[0] = e1[1]; r.q1[1] = e1[2]; r.q1[2] = e2[0]; r.q1[3] = e2[1];
            r.e2z = e2[2];
        } else {
            const float* sp = spheres + 4 * (size_t)(p - ntris);
            r.q0[0] = sp[0]; r.q0[1] = sp[1]; r.q0[2] = sp[2]; r.q0[3] = sp[3];
        }
    }
}

}  // namespace rgpu
