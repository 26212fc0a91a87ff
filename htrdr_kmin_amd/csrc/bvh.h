/*
 * Ground triangle mesh and its bounding volume hierarchy, built on the host.
 *
 * Replaces the Star-3D / Embree scene behind htrdr_geometry_trace_ray
 * (src/core/htrdr_geometry.c:686-718). The BVH is a binary tree laid out in
 * depth-first order with skip links, so that the device walks it without a
 * stack (a stack would cost the wavefront kernel a few dozen registers per
 * lane): a node's first child is the next node, and `skip' is the index of
 * the node that follows its subtree. Leaves hold up to 4 triangles.
 *
 * Children are listed lower side first along the split axis. Laying the tree
 * out once per octant of ray directions (near child first) was measured on a
 * 2M-triangle terrain: 8 % fewer nodes visited, but 8 times the node bytes
 * competing for the caches, and no faster.
 *
 * Triangles are stored in leaf order as (v0, e1 = v1 - v0, e2 = v2 - v0) in
 * FP32: the reference traces the ground in FP32 too (the ray is converted at
 * src/core/htrdr_geometry.c:698-700).
 */
#ifndef HTGX_BVH_H
#define HTGX_BVH_H

#include <stdint.h>
#include <vector>

namespace htgx {

struct BvhNode {      /* 32 bytes */
  float lo[3];
  float hi[3];
  uint32_t skip;      /* next node after this subtree (== #nodes at the end) */
  uint32_t tris;      /* leaf: first triangle << 3 | count (1-4); inner: 0 */
};
static_assert(sizeof(BvhNode) == 32, "BvhNode is 32 bytes");

struct BvhTri {       /* 36 bytes */
  float v0[3];
  float e1[3];
  float e2[3];
};
static_assert(sizeof(BvhTri) == 36, "BvhTri is 36 bytes");

struct BvhHost {
  std::vector<BvhNode> nodes;   /* depth-first, skip links */
  std::vector<BvhTri> tris;     /* leaf order */
  std::vector<uint32_t> prim;   /* leaf order -> input triangle index */
  std::vector<uint32_t> mtl;    /* leaf order -> material */
  float lo[3], hi[3];           /* scene bounds */
  uint32_t depth = 0;
};

/* Build from verts[nverts][3], tris[ntris][3], tri_mtl[ntris]; returns an
 * htgx_res code. Degenerate triangles are kept (they never report a hit). */
int bvh_build(const float* verts, uint32_t nverts, const uint32_t* tris,
              const uint32_t* tri_mtl, uint32_t ntris, BvhHost& out);

} /* namespace htgx */

#endif /* HTGX_BVH_H */
