// Device-resident record layouts shared by the host packers and the gfx950 kernels.
//
// Every record is a multiple of 16 bytes so a lane fetches it with dwordx4 loads
// (global_load_dwordx4), and records never straddle a 64-byte boundary.
#pragma once
#include <stdint.h>

namespace rgpu {

// Stackless (parent-pointer) BVH2 record, one per node, indexed by the node's
// depth-first pre-order id (the numbering of common/stackless_bvh.hpp:283-292).
// Replaces the reference's 96-byte StacklessBVHNode (common/stackless_bvh.hpp:264-275,
// kernels/extension_ray_cast_stackless.cl:21-32): the child boxes are stored as
// 6 floats each instead of two padded float3 pairs, and a leaf is marked by a
// flag bit rather than by primitive_indices_end == 0.
//
//   q0 = { Lmin.x, Lmin.y, Lmin.z, Lmax.x }
//   q1 = { Lmax.y, Lmax.z, Rmin.x, Rmin.y }
//   q2 = { Rmin.z, Rmax.x, Rmax.y, Rmax.z }
//   q3 = { a, b, parent, sibling }
//
// branch: a = left child id, b = right child id (bit 31 clear in a)
// leaf:   a = kLeafFlag | first primitive slot, b = one past the last slot
// A child id (a, b of a branch) and a sibling id carry kChildLeaf when that node is a leaf, so
// a walk that enters it goes straight to its primitives; node ids stay below 2^30.
// A branch visit reads q0..q3 (64 B); a leaf visit and every backtrack step read q3 only.
struct alignas(16) SLNode {
    float lmin[3];
    float lmax[3];
    float rmin[3];
    float rmax[3];
    uint32_t a;
    uint32_t b;
    uint32_t parent;
    uint32_t sibling;
};
static_assert(sizeof(SLNode) == 64, "stackless node record must be 64 bytes");

static const uint32_t kLeafFlag = 0x80000000u;
static const uint32_t kChildLeaf = 0x40000000u;
static const uint32_t kInvalid = 0xFFFFFFFFu;

// Threaded (skip-pointer) node, one table per traversal order: the six ray-direction classes
// (+X,-X,+Y,-Y,+Z,-Z) of the multi-threaded BVH, then the single-table threaded BVH in the
// build's own child order (common/threaded_bvh.hpp:269-364). Each table is the tree re-laid out
// in the depth-first pre-order of that order's child ordering (common/threaded_bvh.hpp:494-516
// for the directions), so the hit link of a branch is always the next record and need not be
// stored; the miss link is the record that follows the node's subtree
// (common/threaded_bvh.hpp:417-484, 309-360).
//   q0 = { bmin.x, bmin.y, bmin.z, bmax.x }
//   q1 = { bmax.y, bmax.z, miss, leaf }
// leaf == 0 for a branch; otherwise kLeafFlag | ((count-1) << 28) | first slot.
// A leaf's hit and miss links are equal (both = miss).
struct alignas(16) MTNode {
    float bmin[3];
    float bmax[3];
    uint32_t miss;
    uint32_t leaf;
};
static_assert(sizeof(MTNode) == 32, "threaded node record must be 32 bytes");
constexpr int kMTOrders = 7;  // tables: 6 direction classes, then the single table
constexpr int kMTSingleOrder = 6;

// Primitive record, one per leaf slot, in leaf order. The triangle edges are the
// exact float differences v1-v0 and v2-v0 that the Moller-Trumbore test forms
// (kernels/intersect_triangle.cl:9-10), so precomputing them changes no bit.
//   triangle: q0 = { v0.x, v0.y, v0.z, e1.x }, q1 = { e1.y, e1.z, e2.x, e2.y },
//             q2 = { e2.z, prim_id, 0, 0 }
//   sphere:   q0 = { c.x, c.y, c.z, r }, q1 = 0, q2 = { 0, prim_id, 0, 0 }
struct alignas(16) PrimRec {
    float q0[4];
    float q1[4];
    float e2z;
    uint32_t prim;
    uint32_t pad0;
    uint32_t pad1;
};
static_assert(sizeof(PrimRec) == 48, "primitive record must be 48 bytes");

}  // namespace rgpu
