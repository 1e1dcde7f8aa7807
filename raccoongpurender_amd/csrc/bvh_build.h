// Host-side BVH construction and packing into the device layouts of rgpu_types.h.
#pragma once
#include <stdint.h>
#include <vector>
#include "rgpu_types.h"

namespace rgpu {

// One node of the built binary tree, stored in depth-first pre-order
// (root = 0, then the left subtree, then the right subtree), which is the order
// common/stackless_bvh.hpp:283-292 assigns to Embree's tree.
struct BuildNode {
    float bmin[3];
    float bmax[3];
    int32_t left;   // -1 for a leaf
    int32_t right;  // -1 for a leaf
    uint32_t prim_beg;  // leaf: first index into BVHTree::prim_ids
    uint32_t prim_cnt;  // leaf: number of primitives (>= 1); 0 for a branch
};

struct BuildOptions {
    uint32_t max_leaf = 4;    // largest leaf
    uint32_t bins = 16;       // SAH bins per axis
    uint32_t max_depth = 63;  // deepest leaf; the stackless bitstack holds 64 levels
    double traversal_cost = 1.0;  // SAH cost of a node visit (a primitive test costs 1)
};

struct BVHTree {
    std::vector<BuildNode> nodes;
    std::vector<uint32_t> prim_ids;  // leaf slots -> primitive id
    uint32_t depth = 0;              // depth of the deepest leaf (root = 0)
    uint32_t leaf_count = 0;
    double sah_cost = 0.0;           // C_trav=1, C_isect=1, normalised by root area
};

// Primitive bounds: triangles first (ids 0..ntris-1), then spheres
// (ids ntris..ntris+nspheres-1), as common/scene_manager.hpp:517-572 orders them.
// points: xyz with `stride` floats per point. spheres: (cx, cy, cz, r) per sphere.
void build_bvh(const float* points, uint32_t stride, uint32_t npoints,
               const uint32_t* indices, uint32_t ntris,
               const float* spheres, uint32_t nspheres,
               const BuildOptions& opt, BVHTree* out);

// Device layouts.
void pack_stackless(const BVHTree& t, std::vector<SLNode>* out);
// Renumbers the stackless records so that the first `top` ids are the top of the tree in
// breadth-first order (root stays 0); the other records keep their relative order. Child,
// parent and sibling links follow; the walk and its results do not change.
void number_top_breadth_first(std::vector<SLNode>* nodes, uint32_t top);
// Sets kChildLeaf on the child and sibling ids that name leaves (after any renumbering).
void mark_leaf_links(std::vector<SLNode>* nodes);
// out: kMTOrders tables of t.nodes.size() records, order-major (+X,-X,+Y,-Y,+Z,-Z, then the
// single table in the build's child order).
void pack_threaded(const BVHTree& t, std::vector<MTNode>* out);
void pack_prims(const BVHTree& t, const float* points, uint32_t stride,
                const uint32_t* indices, uint32_t ntris,
                const float* spheres, uint32_t nspheres, std::vector<PrimRec>* out);

}  // namespace rgpu
