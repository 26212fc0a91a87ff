/* Device build of the flattened k_min/k_max octree (svo.h): the same tree as
 * svo_build, from the LWC grid already in HBM. Three passes, one thread per
 * node: the finest blocks' exact FP32 LWC min/max over their cells, each
 * upper level from its (at most 8) children with the merge rule, and the
 * leaf word of every block (its highest mergeable ancestor). Min/max compare
 * as std::min/std::max do and the merge test evaluates the same FP64
 * expression in the same order (no contraction), so the result is bit for
 * bit the host's. */
#include "svo.h"

#include <hip/hip_runtime.h>

#include <cfloat>

namespace htgx {
namespace {

/* bits of the leaf id in a level-map word (level << 28 | leaf id), as the
 * solvers read it */
constexpr uint32_t kLeafIdBits = 28;

struct Levels {
  uint32_t n;
  uint32_t dims[kMaxLevels][3];
  uint64_t offset[kMaxLevels];
};

__global__ void
svo_blocks_kernel(const float* __restrict__ lwc, uint32_t def0, uint32_t def1, uint32_t B,
                  uint32_t bx, uint32_t by, uint32_t bz, float2* __restrict__ bounds,
                  uint8_t* __restrict__ mergeable)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n = (uint64_t)bx * by * bz;
  if(i >= n) return;
  const uint32_t x = (uint32_t)(i % bx), y = (uint32_t)((i / bx) % by), z = (uint32_t)(i / ((uint64_t)bx * by));
  float mn = FLT_MAX, mx = -FLT_MAX;
  for(uint32_t k = z * B; k < (z + 1) * B; ++k)
    for(uint32_t j = y * B; j < (y + 1) * B; ++j) {
      const float* row = lwc + ((size_t)k * def1 + j) * def0;
      for(uint32_t ii = x * B; ii < (x + 1) * B; ++ii) {
        const float v = row[ii];
        mn = v < mn ? v : mn; /* std::min(mn, v) */
        mx = mx < v ? v : mx; /* std::max(mx, v) */
      }
    }
  bounds[i] = make_float2(mn, mx);
  mergeable[i] = 1;
}

__global__ void
svo_level_kernel(Levels L, uint32_t l, double cext_max, double edge, double thickness,
                 float2* __restrict__ bounds, uint8_t* __restrict__ mergeable)
{
  const uint32_t* d = L.dims[l];
  const uint32_t* dc = L.dims[l - 1];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= (uint64_t)d[0] * d[1] * d[2]) return;
  const uint32_t x = (uint32_t)(i % d[0]), y = (uint32_t)((i / d[0]) % d[1]);
  const uint32_t z = (uint32_t)(i / ((uint64_t)d[0] * d[1]));
  float mn = FLT_MAX, mx = -FLT_MAX;
  bool all = true;
  for(int c = 0; c < 8; ++c) {
    const uint32_t cx = 2 * x + (c & 1), cy = 2 * y + ((c >> 1) & 1), cz = 2 * z + ((c >> 2) & 1);
    if(cx >= dc[0] || cy >= dc[1] || cz >= dc[2]) continue;
    const uint64_t ci = L.offset[l - 1] + ((uint64_t)cz * dc[1] + cy) * dc[0] + cx;
    const float2 b = bounds[ci];
    mn = b.x < mn ? b.x : mn;
    mx = mx < b.y ? b.y : mx;
    all = all && mergeable[ci];
  }
  const uint64_t ni = L.offset[l] + i;
  bounds[ni] = make_float2(mn, mx);
  mergeable[ni] = all && cext_max * ((double)mx - (double)mn) * edge <= thickness;
}

__global__ void
svo_lmap_kernel(Levels L, const uint8_t* __restrict__ mergeable, uint32_t* __restrict__ lmap,
                unsigned long long* __restrict__ nleaves)
{
  const uint32_t* d = L.dims[0];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < (uint64_t)d[0] * d[1] * d[2];
  bool first = false;
  if(valid) {
    const uint32_t x = (uint32_t)(i % d[0]), y = (uint32_t)((i / d[0]) % d[1]);
    const uint32_t z = (uint32_t)(i / ((uint64_t)d[0] * d[1]));
    uint32_t lv = 0;
    uint64_t node = i;
    while(lv + 1 < L.n) {
      const uint32_t* dp = L.dims[lv + 1];
      const uint64_t p = L.offset[lv + 1]
        + ((uint64_t)(z >> (lv + 1)) * dp[1] + (y >> (lv + 1))) * dp[0] + (x >> (lv + 1));
      if(!mergeable[p]) break;
      node = p;
      ++lv;
    }
    lmap[i] = (lv << kLeafIdBits) | (uint32_t)node;
    const uint32_t m = (1u << lv) - 1u;
    first = !(x & m) && !(y & m) && !(z & m); /* each leaf counted from its first block */
  }
  const uint64_t ballot = __ballot(first);
  if(ballot && (threadIdx.x & 63u) == (uint32_t)(__ffsll((unsigned long long)ballot) - 1))
    atomicAdd(nleaves, (unsigned long long)__popcll(ballot));
}

} /* namespace */

#define SVO_OK(x) do { if((x) != hipSuccess) { res = HTGX_UNKNOWN_ERR; goto out; } } while(0)

int svo_build_device(const SvoHost& svo, const float* lwc, const uint32_t def[3],
                     double optical_thickness, float2* bounds, uint32_t* lmap, uint64_t* nleaves)
{
  int res = HTGX_OK;
  uint8_t* merge = nullptr;
  unsigned long long* count = nullptr;
  unsigned long long n_host = 0;
  Levels L;
  L.n = svo.nlevels;
  for(uint32_t l = 0; l < kMaxLevels; ++l)
    for(int a = 0; a < 3; ++a) L.dims[l][a] = svo.dims[l][a];
  for(uint32_t l = 0; l < kMaxLevels; ++l) L.offset[l] = svo.offset[l];
  const uint64_t nblocks = (uint64_t)svo.dims[0][0] * svo.dims[0][1] * svo.dims[0][2];

  SVO_OK(hipMalloc(&merge, svo.nnodes));
  SVO_OK(hipMalloc(&count, sizeof(unsigned long long)));
  SVO_OK(hipMemset(count, 0, sizeof(unsigned long long)));
  svo_blocks_kernel<<<(uint32_t)((nblocks + 255) / 256), 256>>>(
    lwc, def[0], def[1], svo.B, svo.dims[0][0], svo.dims[0][1], svo.dims[0][2], bounds, merge);
  SVO_OK(hipGetLastError());
  for(uint32_t l = 1; l < svo.nlevels; ++l) {
    /* the node's largest nominal edge, as svo_build computes it */
    double edge = 0;
    for(int a = 0; a < 3; ++a) {
      const double e = svo.bs[a] * (double)(1u << l);
      edge = edge < e ? e : edge;
    }
    const uint64_t n = (uint64_t)svo.dims[l][0] * svo.dims[l][1] * svo.dims[l][2];
    svo_level_kernel<<<(uint32_t)((n + 255) / 256), 256>>>(L, l, svo.cext_max, edge,
                                                           optical_thickness, bounds, merge);
    SVO_OK(hipGetLastError());
  }
  svo_lmap_kernel<<<(uint32_t)((nblocks + 255) / 256), 256>>>(L, merge, lmap, count);
  SVO_OK(hipGetLastError());
  SVO_OK(hipMemcpy(&n_host, count, sizeof(n_host), hipMemcpyDeviceToHost));
  *nleaves = n_host;
out:
  if(merge) (void)hipFree(merge);
  if(count) (void)hipFree(count);
  return res;
}

} /* namespace htgx */
