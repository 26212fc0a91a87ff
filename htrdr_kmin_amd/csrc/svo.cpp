/* Bottom-up build of the flattened k_min/k_max octree (see svo.h). */
#include "svo.h"

#include <algorithm>
#include <cfloat>
#include <cstring>

namespace htgx {

int svo_layout(const htgx_scene_desc& scene, SvoHost& svo)
{
  const htgx_cloud_desc& cl = scene.clouds;
  const uint32_t B = scene.leaf_block ? scene.leaf_block : 1;
  svo = SvoHost();
  svo.B = B;
  for(int a = 0; a < 3; ++a) {
    if(!cl.def[a] || cl.def[a] % B) return HTGX_BAD_ARG;
    if(!(cl.upper[a] > cl.lower[a])) return HTGX_BAD_ARG;
    svo.cs[a] = (cl.upper[a] - cl.lower[a]) / (double)cl.def[a];
    svo.inv_cs[a] = 1.0 / svo.cs[a];
    svo.bs[a] = svo.cs[a] * (double)B;
    svo.inv_bs[a] = 1.0 / svo.bs[a];
    svo.dims[0][a] = cl.def[a] / B;
  }
  svo.nlevels = 1;
  svo.offset[0] = 0;
  while(svo.dims[svo.nlevels - 1][0] > 1 || svo.dims[svo.nlevels - 1][1] > 1
     || svo.dims[svo.nlevels - 1][2] > 1) {
    const uint32_t l = svo.nlevels;
    if(l >= kMaxLevels) return HTGX_BAD_ARG;
    for(int a = 0; a < 3; ++a) svo.dims[l][a] = (svo.dims[l - 1][a] + 1) / 2;
    svo.offset[l] = svo.offset[l - 1]
      + (uint64_t)svo.dims[l - 1][0] * svo.dims[l - 1][1] * svo.dims[l - 1][2];
    svo.nlevels++;
  }
  const uint32_t top = svo.nlevels - 1;
  svo.nnodes = svo.offset[top] + 1;
  if(svo.nnodes > (1ull << 28)) return HTGX_BAD_ARG;

  svo.cext_max = 0;
  for(uint32_t b = 0; b < scene.spectral.nbands; ++b)
    svo.cext_max = std::max(svo.cext_max, cl.cext[b]);
  return HTGX_OK;
}

int svo_build(const htgx_scene_desc& scene, SvoHost& svo)
{
  int res = svo_layout(scene, svo);
  if(res != HTGX_OK) return res;
  const htgx_cloud_desc& cl = scene.clouds;
  const uint32_t B = svo.B;
  svo.bounds.resize(svo.nnodes);
  svo.mergeable.resize(svo.nnodes);

  /* Level 0: exact FP32 min/max over the cells of each block */
  {
    const uint32_t* d = svo.dims[0];
    const int64_t n = (int64_t)d[0] * d[1] * d[2];
    #pragma omp parallel for schedule(static)
    for(int64_t i = 0; i < n; ++i) {
      const uint32_t x = (uint32_t)(i % d[0]);
      const uint32_t y = (uint32_t)((i / d[0]) % d[1]);
      const uint32_t z = (uint32_t)(i / ((int64_t)d[0] * d[1]));
      float mn = FLT_MAX, mx = -FLT_MAX;
      for(uint32_t k = z * B; k < (z + 1) * B; ++k)
      for(uint32_t j = y * B; j < (y + 1) * B; ++j) {
        const float* row = cl.lwc + ((size_t)k * cl.def[1] + j) * cl.def[0];
        for(uint32_t ii = x * B; ii < (x + 1) * B; ++ii) {
          mn = std::min(mn, row[ii]);
          mx = std::max(mx, row[ii]);
        }
      }
      svo.bounds[i] = make_float2(mn, mx);
      svo.mergeable[i] = 1;
    }
  }

  /* Upper levels */
  for(uint32_t l = 1; l < svo.nlevels; ++l) {
    const uint32_t* d = svo.dims[l];
    const uint32_t* dc = svo.dims[l - 1];
    double edge = 0;
    for(int a = 0; a < 3; ++a) edge = std::max(edge, svo.bs[a] * (double)(1u << l));
    const int64_t n = (int64_t)d[0] * d[1] * d[2];
    #pragma omp parallel for schedule(static)
    for(int64_t i = 0; i < n; ++i) {
      const uint32_t x = (uint32_t)(i % d[0]);
      const uint32_t y = (uint32_t)((i / d[0]) % d[1]);
      const uint32_t z = (uint32_t)(i / ((int64_t)d[0] * d[1]));
      float mn = FLT_MAX, mx = -FLT_MAX;
      bool all = true;
      for(int c = 0; c < 8; ++c) {
        const uint32_t cx = 2 * x + (c & 1), cy = 2 * y + ((c >> 1) & 1), cz = 2 * z + ((c >> 2) & 1);
        if(cx >= dc[0] || cy >= dc[1] || cz >= dc[2]) continue;
        const uint64_t ci = svo.offset[l - 1] + ((uint64_t)cz * dc[1] + cy) * dc[0] + cx;
        mn = std::min(mn, svo.bounds[ci].x);
        mx = std::max(mx, svo.bounds[ci].y);
        all = all && svo.mergeable[ci];
      }
      const uint64_t ni = svo.offset[l] + (uint64_t)i;
      svo.bounds[ni] = make_float2(mn, mx);
      svo.mergeable[ni] = all
        && svo.cext_max * ((double)mx - (double)mn) * edge <= scene.optical_thickness;
    }
  }

  /* Leaf word of every block: highest mergeable ancestor */
  {
    const uint32_t* d = svo.dims[0];
    const int64_t n = (int64_t)d[0] * d[1] * d[2];
    svo.lmap.resize((size_t)n);
    uint64_t nleaves = 0;
    #pragma omp parallel for schedule(static) reduction(+:nleaves)
    for(int64_t i = 0; i < n; ++i) {
      const uint32_t x = (uint32_t)(i % d[0]);
      const uint32_t y = (uint32_t)((i / d[0]) % d[1]);
      const uint32_t z = (uint32_t)(i / ((int64_t)d[0] * d[1]));
      uint32_t L = 0;
      uint64_t node = (uint64_t)i;
      while(L + 1 < svo.nlevels) {
        const uint32_t* dp = svo.dims[L + 1];
        const uint64_t p = svo.offset[L + 1]
          + ((uint64_t)(z >> (L + 1)) * dp[1] + (y >> (L + 1))) * dp[0] + (x >> (L + 1));
        if(!svo.mergeable[p]) break;
        node = p;
        ++L;
      }
      svo.lmap[(size_t)i] = (L << 28) | (uint32_t)node;
      /* count each leaf once, from its first block */
      const uint32_t m = (1u << L) - 1u;
      if(!(x & m) && !(y & m) && !(z & m)) nleaves += 1;
    }
    svo.nleaves = nleaves;
  }
  return HTGX_OK;
}

} /* namespace htgx */
