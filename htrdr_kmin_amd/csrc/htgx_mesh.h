/*
 * Ground mesh ray casting in FP32, host and device.
 *
 * Semantics of htrdr_geometry_trace_ray (src/core/htrdr_geometry.c:686-718)
 * with its self-hit filter (:96-150): the closest triangle hit in the open
 * range ]r0, r1[, double-sided, whose hit
 *  - is not on the triangle the ray leaves, and
 *  - is not a near-origin (distance <= 0.1) hit with both the hit and the
 *    previous hit on an edge or a vertex (barycentrics within 1e-4 of 0 or 1),
 * which the reference treats as self-intersections across shared edges.
 * The ray is converted to FP32 as the reference does (:698-700). The
 * intersection is Moller-Trumbore (J. Graphics Tools 2(1), 1997); the BVH
 * walk is the stackless skip-link order of bvh.h.
 */
#ifndef HTGX_MESH_H
#define HTGX_MESH_H

#include "bvh.h"
#include "htgx_rng.h" /* HTGX_HD */

#include <cfloat>
#include <cmath>

#include <float.h>

namespace htgx {

struct MeshHit {
  float t, u, v;
  int32_t tri;   /* leaf-order triangle index; -1: no hit */
  uint32_t nodes, ntris; /* BVH nodes and triangles visited */
};

HTGX_HD bool mesh_on_edge(float u, float v)
{
  const float eps = 1e-4f;
  const float w = 1.f - u - v;
  return fabsf(u) <= eps || fabsf(u - 1.f) <= eps || fabsf(v) <= eps || fabsf(v - 1.f) <= eps
      || fabsf(w) <= eps || fabsf(w - 1.f) <= eps;
}

/* Moller-Trumbore, double-sided; false if parallel or outside */
HTGX_HD bool mesh_tri_hit(const BvhTri& T, const float o[3], const float d[3], float* t, float* u,
                          float* v)
{
  const float p0 = d[1] * T.e2[2] - d[2] * T.e2[1];
  const float p1 = d[2] * T.e2[0] - d[0] * T.e2[2];
  const float p2 = d[0] * T.e2[1] - d[1] * T.e2[0];
  const float det = T.e1[0] * p0 + T.e1[1] * p1 + T.e1[2] * p2;
  if(det == 0.f) return false;
  const float inv = 1.f / det;
  const float s0 = o[0] - T.v0[0], s1 = o[1] - T.v0[1], s2 = o[2] - T.v0[2];
  const float uu = (s0 * p0 + s1 * p1 + s2 * p2) * inv;
  if(!(uu >= 0.f && uu <= 1.f)) return false;
  const float q0 = s1 * T.e1[2] - s2 * T.e1[1];
  const float q1 = s2 * T.e1[0] - s0 * T.e1[2];
  const float q2 = s0 * T.e1[1] - s1 * T.e1[0];
  const float vv = (d[0] * q0 + d[1] * q1 + d[2] * q2) * inv;
  if(!(vv >= 0.f && uu + vv <= 1.f)) return false;
  *t = (T.e2[0] * q0 + T.e2[1] * q1 + T.e2[2] * q2) * inv;
  *u = uu;
  *v = vv;
  return true;
}

/* The self-hit filter of src/core/htrdr_geometry.c:116-150 */
HTGX_HD bool mesh_self_hit(float t, float u, float v, int32_t tri, float r0, float r1,
                           int32_t prev_tri, bool prev_edge)
{
  if(t <= r0 || t >= r1) return true;
  if(tri == prev_tri) return true;
  if(prev_tri >= 0 && fabsf(t) <= 1e-1f && prev_edge && mesh_on_edge(u, v)) return true;
  return false;
}

/* Closest accepted hit along o + t d, t in ]r0, r1[. With `any', the first
 * accepted hit ends the walk (occlusion queries). */
HTGX_HD MeshHit mesh_trace(const BvhNode* nodes, uint32_t nnodes, const BvhTri* tris,
                           const float o[3], const float d[3], float r0, float r1,
                           int32_t prev_tri, bool prev_edge, bool any)
{
  MeshHit h;
  h.t = r1;
  h.u = h.v = 0.f;
  h.tri = -1;
  h.nodes = h.ntris = 0;
  float inv[3];
#pragma unroll
  for(int a = 0; a < 3; ++a) inv[a] = 1.f / d[a];
  uint32_t i = 0;
  while(i < nnodes) {
    const BvhNode& N = nodes[i];
    h.nodes++;
    float tn = r0, tf = h.t;
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      const float ta = (N.lo[a] - o[a]) * inv[a];
      const float tb = (N.hi[a] - o[a]) * inv[a];
      tn = fmaxf(tn, fminf(ta, tb));
      tf = fminf(tf, fmaxf(ta, tb));
    }
    if(!(tn <= tf)) { i = N.skip; continue; }
    if(!N.tris) { i += 1; continue; }
    const uint32_t first = N.tris >> 3, count = N.tris & 7u;
    for(uint32_t k = 0; k < count; ++k) {
      float t, u, v;
      const int32_t tri = (int32_t)(first + k);
      h.ntris++;
      if(!mesh_tri_hit(tris[first + k], o, d, &t, &u, &v)) continue;
      if(t >= h.t || mesh_self_hit(t, u, v, tri, r0, r1, prev_tri, prev_edge)) continue;
      h.t = t;
      h.u = u;
      h.v = v;
      h.tri = tri;
      if(any) return h;
    }
    i = N.skip;
  }
  return h;
}

/* The closest accepted hit on a mesh, or with `repeat' on its copies along
 * X and Y with the period of its bounding box [lo, hi]: the walk of
 * htrdr_slab_trace_ray (src/core/htrdr_slab.c:35-137) from the copy where
 * the ray enters the box's Z range, across at most 32 copies, each traced
 * with the origin relative to it and the whole range; the first copy with a
 * hit holds the closest one (htrdr_atmosphere_ground_trace_ray,
 * src/atmosphere/htrdr_atmosphere_ground.c:185-210). The previous hit is the
 * same triangle in every copy, as in the reference. Real is the precision of
 * the walk; the ray is converted to FP32 per copy, as the reference does
 * (src/core/htrdr_geometry.c:698-700). One trace site serves both cases, so
 * the BVH walk is inlined once. The returned tri is -1 without a hit. */
template<class Real>
HTGX_HD MeshHit mesh_trace_copies(const BvhNode* nodes, uint32_t nnodes, const BvhTri* tris,
                                  bool repeat, const Real lo[3], const Real hi[3], const Real o[3],
                                  const Real d[3], Real r0, Real r1, int32_t prev_tri,
                                  bool prev_edge, bool any)
{
  Real sz[2] = {0, 0}, tmax[3] = {0, 0, 0}, tdelta[2] = {0, 0};
  int64_t xy[2] = {0, 0};
  int incr[2] = {1, 1};
  if(repeat) {
    Real tz0 = (lo[2] - o[2]) / d[2], tz1 = (hi[2] - o[2]) / d[2];
    if(tz0 > tz1) { const Real t = tz0; tz0 = tz1; tz1 = t; }
    tz0 = tz0 > r0 ? tz0 : r0;
    tz1 = tz1 < r1 ? tz1 : r1;
    if(tz0 > tz1) {
      MeshHit none;
      none.t = 0.f;
      none.u = none.v = 0.f;
      none.tri = -1;
      none.nodes = none.ntris = 0;
      return none;
    }
    tmax[2] = tz1;
#pragma unroll
    for(int a = 0; a < 2; ++a) {
      sz[a] = hi[a] - lo[a];
      xy[a] = (int64_t)floor((o[a] + tz0 * d[a] - lo[a]) / sz[a]);
      incr[a] = d[a] < 0 ? -1 : 1;
      const Real low = lo[a] + (Real)xy[a] * sz[a];
      tmax[a] = ((d[a] < 0 ? low : low + sz[a]) - o[a]) / d[a];
      tdelta[a] = (d[a] < 0 ? -sz[a] : sz[a]) / d[a];
    }
  }
  const float df[3] = {(float)d[0], (float)d[1], (float)d[2]};
  const float r0f = (float)r0;
  const float r1f = (double)r1 > (double)FLT_MAX ? INFINITY : (float)r1;
  uint32_t nodes_acc = 0, tris_acc = 0;
  MeshHit m;
  for(int step = 0; step < 32; ++step) {
    /* without repetition xy and sz are 0: the origin as given */
    const float of[3] = {(float)(o[0] - (Real)xy[0] * sz[0]), (float)(o[1] - (Real)xy[1] * sz[1]),
                         (float)o[2]};
    m = mesh_trace(nodes, nnodes, tris, of, df, r0f, r1f, prev_tri, prev_edge, any);
    nodes_acc += m.nodes;
    tris_acc += m.ntris;
    if(m.tri >= 0 || !repeat) break;
    const int ax = tmax[0] < tmax[1] ? (tmax[0] < tmax[2] ? 0 : 2) : (tmax[1] < tmax[2] ? 1 : 2);
    if(ax == 2 || tmax[ax] >= r1) break;
    tmax[ax] += tdelta[ax];
    xy[ax] += incr[ax];
  }
  m.nodes = nodes_acc;
  m.ntris = tris_acc;
  return m;
}

} /* namespace htgx */

#endif /* HTGX_MESH_H */
