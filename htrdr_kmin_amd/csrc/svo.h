/*
 * Flattened, pointerless k_min/k_max octree over the cloud grid.
 *
 * Replaces the star-vx octrees htsky builds per (band, k-point)
 * (src/atmosphere/htrdr_atmosphere.c:366-382, "each octree node stores the
 * minimum and maximum extinction coefficients", doc/htrdr-atmosphere.1.in:
 * 226-228). One tree serves every band: nodes store the FP32 min/max of the
 * liquid water content, and the property bounds of a band are obtained at
 * traversal time as c(prop, band) * LWC_min/max (c > 0, so min/max commute
 * with the scaling) plus the gas value of the crossed layer.
 *
 * Levels: level 0 holds the leaf blocks of BxBxB cells; level l+1 has
 * ceil(dims_l / 2) nodes per axis. A node is mergeable if it is a block, or
 * if all its children are mergeable and
 *     cext_max * (LWC_max - LWC_min) * edge <= optical_thickness
 * with edge the node's largest nominal edge (B * 2^l cells). A leaf is the
 * highest mergeable ancestor of a block. The leaf word of each block stores
 * that leaf's level and global id, so traversal never descends the tree.
 */
#ifndef HTGX_SVO_H
#define HTGX_SVO_H

#include "htgx.h"

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

namespace htgx {

constexpr uint32_t kMaxLevels = 16;

struct SvoHost {
  uint32_t B = 1;
  uint32_t nlevels = 0;
  uint32_t dims[kMaxLevels][3] = {};
  uint64_t offset[kMaxLevels] = {};
  double cs[3] = {0, 0, 0}, inv_cs[3] = {0, 0, 0};
  double bs[3] = {0, 0, 0}, inv_bs[3] = {0, 0, 0};
  double cext_max = 0;
  std::vector<float2> bounds;    /* all nodes, level by level */
  std::vector<uint8_t> mergeable; /* all nodes */
  std::vector<uint32_t> lmap;    /* per level-0 block: level << 28 | leaf id */
  uint64_t nnodes = 0;
  uint64_t nleaves = 0;
};

/* Returns HTGX_BAD_ARG if the cloud definition is not a multiple of the leaf
 * block or if the tree has more than 2^28 nodes. */
int svo_build(const htgx_scene_desc& scene, SvoHost& svo);

/* The tree's shape only (block and cell sizes, level dimensions and offsets,
 * cext_max): what svo_build computes before filling the arrays, and all a
 * device build needs besides the LWC grid */
int svo_layout(const htgx_scene_desc& scene, SvoHost& svo);

/* The same tree built on the current device from the LWC grid already in
 * device memory (lwc, def[2] x def[1] x def[0] floats), into bounds
 * [svo.nnodes] and lmap [blocks], with the same min/max and merge rule, bit
 * for bit (svo holds the layout from svo_layout; its nleaves is set). */
int svo_build_device(const SvoHost& layout, const float* lwc, const uint32_t def[3],
                     double optical_thickness, float2* bounds, uint32_t* lmap, uint64_t* nleaves);

/* Octree cache of option -O (svo_cache.cpp): the tree and the build
 * parameters it was made with */
struct SvoCache {
  uint32_t leaf_block = 0, nlevels = 0;
  double optical_thickness = 0;
  uint64_t nleaves = 0;
  std::vector<uint32_t> lmap;
  std::vector<float2> bounds;
};

/* Digest of the inputs a cache depends on (bands, gas, clouds, cloud optics) */
uint64_t svo_cache_digest(const htgx_scene_desc& scene);
/* *found = false and HTGX_OK if the file does not exist; HTGX_BAD_ARG (err
 * says why) if it is not a cache, was built from other inputs, is truncated,
 * does not fit the scene's cloud grid or holds an invalid tree (leaf words,
 * tiling, bounds checked against the scene's LWC); HTGX_IO_ERR if it cannot
 * be opened */
int svo_cache_read(const char* path, const htgx_scene_desc& scene, uint64_t digest, SvoCache& out,
                   bool* found, std::string& err);
int svo_cache_write(const char* path, uint64_t digest, const SvoCache& cache, std::string& err);

} /* namespace htgx */

#endif /* HTGX_SVO_H */
