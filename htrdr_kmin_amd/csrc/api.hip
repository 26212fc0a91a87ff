/*
 * C-ABI of libhtgx.so: scene upload, plain FP64 solver, finalisation and
 * debug entry points (declarations and reference citations in htgx.h).
 */
#include "htgx.h"
#include "htgx_device.h"
#include "htgx_internal.h"
#include "bvh.h"
#include "svo.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include <unistd.h>

using namespace htgx;

#define HIP_OK(expr) do {                                                     \
  hipError_t e_ = (expr);                                                    \
  if(e_ != hipSuccess) {                                                     \
    fprintf(stderr, "htgx: %s failed: %s (%s:%d)\n", #expr,                  \
      hipGetErrorString(e_), __FILE__, __LINE__);                            \
    return e_ == hipErrorOutOfMemory ? HTGX_MEM_ERR : HTGX_UNKNOWN_ERR;       \
  }                                                                           \
} while(0)

/*******************************************************************************
 * Kernels
 ******************************************************************************/
namespace {

/* Pixel of item i of a realisation slice: 8x8 tiles (one wavefront per tile,
 * the tile size of src/core/htrdr_draw_map.c:45), pixels of a tile in Morton
 * order like draw_tile (src/core/htrdr_draw_map.c:301-319). */
__device__ __forceinline__ void
tile_pixel(uint32_t i, uint32_t ntx, uint32_t* x, uint32_t* y)
{
  const uint32_t tile = i >> 6, m = i & 63u;
  *x = (tile % ntx) * 8u + ((m & 1u) | ((m >> 1) & 2u) | ((m >> 2) & 4u));
  *y = (tile / ntx) * 8u + (((m >> 1) & 1u) | ((m >> 2) & 2u) | ((m >> 3) & 4u));
}

__device__ __forceinline__ uint64_t
wave_sum(uint64_t v)
{
  for(int off = 32; off > 0; off >>= 1)
    v += (uint64_t)__shfl_xor((unsigned long long)v, off, 64);
  return v;
}

__device__ __forceinline__ void
flush_counters(const Counters& c, unsigned long long* counters)
{
  const uint64_t a = wave_sum(c.nleaf), b = wave_sum(c.nseg), d = wave_sum(c.ncoll),
                 e = wave_sum(c.ncell), f = wave_sum(c.nev);
  if((threadIdx.x & 63u) == 0) {
    atomicAdd(counters + 1, (unsigned long long)a);
    atomicAdd(counters + 2, (unsigned long long)b);
    atomicAdd(counters + 3, (unsigned long long)d);
    atomicAdd(counters + 4, (unsigned long long)e);
    atomicAdd(counters + 5, (unsigned long long)f);
  }
}

/* One thread = one pixel x one slice of its realisations, one path at a
 * time, FP64 throughout: the "plain" solver and the FP64 GPU oracle. It is
 * the device counterpart of draw_pixel_xwave / draw_pixel_flux
 * (src/atmosphere/htrdr_atmosphere_draw_map.c:210-436) driven by the tile
 * loop of draw_map (src/core/htrdr_draw_map.c:324-491). */
template<bool COUNT>
__global__ void __launch_bounds__(256)
plain_kernel(DevScene S, DevMap M, uint32_t ntx, uint32_t items_per_slice,
             uint32_t nslices, uint32_t spt, htgx_mc_accum* partials,
             unsigned long long* counters)
{
  const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Counters cnt = {0, 0, 0, 0, 0};
  uint32_t x = 0, y = 0, slice = 0;
  bool active = item < (uint64_t)items_per_slice * nslices;
  if(active) {
    slice = (uint32_t)(item / items_per_slice);
    tile_pixel((uint32_t)(item % items_per_slice), ntx, &x, &y);
    active = x < M.width && y < M.height;
  }
  if(active) {
    const uint64_t npix = (uint64_t)M.width * M.height;
    const uint64_t pixel = (uint64_t)y * M.width + x;
    const uint32_t rb = slice * spt;
    const uint32_t re = min(rb + spt, M.spp);
    Counters* pc = COUNT ? &cnt : nullptr;
    htgx_mc_accum a;
    a.sum = a.sum_c = a.sum2 = a.sum2_c = a.tsum = a.tsum2 = 0;
    a.n = a.nfail = 0;
    Path P;
    Stream rng;
    TrackOut out;
    uint32_t r = rb;
    uint64_t t0 = wall_clock64();
    real_begin(S, M, P, rng, pixel, (uint64_t)M.spp_offset + r);
    /* Flat loop: a lane whose realisation ends starts its next one at once
     * instead of waiting at a per-realisation barrier for the other lanes. */
    while(r < re) {
      while(P.stage == ST_WALK_DONE) walk_done(S, M, P);
      if(P.stage == ST_REAL_DONE) {
        const uint64_t t1 = wall_clock64();
        if(P.rejected) {
          a.nfail++;
        } else {
          const double us = (double)(t1 - t0) * M.us_per_tick;
          nsum(a.sum, a.sum_c, P.weight);
          nsum(a.sum2, a.sum2_c, P.weight * P.weight);
          a.tsum += us;
          a.tsum2 += us * us;
          a.n++;
        }
        if(++r >= re) break;
        t0 = t1;
        real_begin(S, M, P, rng, pixel, (uint64_t)M.spp_offset + r);
        continue;
      }
      int mode, prop;
      double o[3], d[3], r1;
      stage_query(P, &mode, &prop, o, d, &r1);
      track(S, rng, mode, prop, P.iband, P.iquad, o, d, r1, out, pc);
      stage_consume(S, P, rng, out, pc);
    }
    partials[(size_t)slice * npix + pixel] = a;
  }
  if(COUNT) flush_counters(cnt, counters);
}

/* Fold the slices of each pixel, in slice order, into its accumulator */
__global__ void __launch_bounds__(256)
merge_kernel(const htgx_mc_accum* partials, uint32_t nslices, uint64_t npix,
             htgx_mc_accum* accum)
{
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(p >= npix) return;
  htgx_mc_accum a = accum[p];
  for(uint32_t s = 0; s < nslices; ++s) {
    const htgx_mc_accum q = partials[(size_t)s * npix + p];
    nsum(a.sum, a.sum_c, q.sum);
    a.sum_c += q.sum_c;
    nsum(a.sum2, a.sum2_c, q.sum2);
    a.sum2_c += q.sum2_c;
    a.tsum += q.tsum;
    a.tsum2 += q.tsum2;
    a.n += q.n;
    a.nfail += q.nfail;
  }
  accum[p] = a;
}

__global__ void
trace_segments_kernel(DevScene S, const double* rays, uint32_t nrays, int prop,
                      uint32_t iband, uint32_t iquad, htgx_segment* segs,
                      uint32_t max_segs, uint32_t* nsegs)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= nrays) return;
  const double* r = rays + (size_t)i * 8;
  SegIter it;
  Segment seg;
  uint32_t n = 0;
  it.init(S, r, r + 3, r[6], r[7], nullptr);
  while(it.next(S, prop, iband, iquad, seg, nullptr)) {
    if(n < max_segs) {
      htgx_segment& o = segs[(size_t)i * max_segs + n];
      o.t0 = seg.t0;
      o.t1 = seg.t1;
      o.kmin = seg.kmin;
      o.kmax = seg.kmax;
      o.leaf = seg.in_cloud ? (int64_t)seg.leaf : -1;
      o.layer = seg.layer;
      o.pad = 0;
    }
    n++;
  }
  nsegs[i] = n;
}

__global__ void __launch_bounds__(64)
realisations_kernel(DevScene S, DevMap M, const uint64_t* pix, const uint64_t* real,
                    uint32_t n, double* out)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  Path P;
  Stream rng;
  TrackOut o;
  real_begin(S, M, P, rng, pix[i], real[i]);
  run_until_done(S, M, P, rng, o, nullptr);
  out[i] = P.rejected ? __longlong_as_double(0x7ff8000000000000LL) : P.weight;
}

__global__ void
threefry_kernel(const uint64_t* ctr, const uint64_t* key, uint32_t n, uint64_t* out)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  threefry4x64_20(ctr + 4 * (size_t)i, key + 4 * (size_t)i, out + 4 * (size_t)i);
}

__global__ void
threefry2x32_kernel(const uint32_t* ctr, const uint32_t* key, uint32_t n, uint32_t* out)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  threefry2x32_20(ctr + 2 * (size_t)i, key + 2 * (size_t)i, out + 2 * (size_t)i);
}

__global__ void
sample_dirs_kernel(DevScene S, int kind, double param, double dx, double dy, double dz,
                   const double* u, uint32_t n, double* out)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const double din[3] = {dx, dy, dz};
  const double u0 = u[2 * (size_t)i], u1 = u[2 * (size_t)i + 1];
  double* o = out + 3 * (size_t)i;
  switch(kind) {
    case 0: local_to_world(din, hg_cos_sample(param, u0), 2.0 * kPi * u1, o); break;
    case 1: local_to_world(din, rayleigh_cos_sample(u0), 2.0 * kPi * u1, o); break;
    case 2: {
      const double cos_t = 1.0 - u0 * (1.0 - S.sun_cos_half);
      local_to_world(S.sun_dir, cos_t, 2.0 * kPi * u1, o);
    } break;
    default: cos_hemisphere(din, u0, u1, o); break;
  }
}

/* One ground-mesh ray cast as the FP64 solver casts it (ground_trace):
 * ray[8] = org, dir, range; the ray is converted to FP32 */
HTGX_HD void mesh_trace_ray(const BvhNode* nodes, uint32_t nnodes, const BvhTri* tris,
                            const double* ray, int32_t prev, bool prev_edge, bool any,
                            double* t, int32_t* tri, uint32_t* nvisited)
{
  const float o[3] = {(float)ray[0], (float)ray[1], (float)ray[2]};
  const float d[3] = {(float)ray[3], (float)ray[4], (float)ray[5]};
  const float r1 = ray[7] > (double)FLT_MAX ? INFINITY : (float)ray[7];
  const MeshHit m = mesh_trace(nodes, nnodes, tris, o, d, (float)ray[6], r1, prev, prev_edge, any);
  *t = m.tri >= 0 ? (double)m.t : INFINITY;
  *tri = m.tri;
  *nvisited = m.nodes;
}

__global__ void
mesh_trace_kernel(DevScene S, const double* rays, const int32_t* prev, const int32_t* prev_edge,
                  uint32_t n, int any, double* t, int32_t* tri, uint32_t* nvisited)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  mesh_trace_ray(S.bvh, S.nbvh, S.tris, rays + 8 * (size_t)i, prev[i], prev_edge[i] != 0,
                 any != 0, t + i, tri + i, nvisited + i);
}

/* Plain streaming copy, the reference traffic of the HBM counters */
__global__ void __launch_bounds__(256)
copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
      i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

template<class T>
int device_alloc(htgx_ctx* ctx, size_t n, T** out)
{
  void* d = nullptr;
  HIP_OK(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
  ctx->bufs.push_back(d);
  *out = (T*)d;
  return HTGX_OK;
}

template<class T>
int upload(htgx_ctx* ctx, const T* host, size_t n, const T** out)
{
  void* d = nullptr;
  if(!n) { *out = nullptr; return HTGX_OK; }
  HIP_OK(hipMalloc(&d, n * sizeof(T)));
  ctx->bufs.push_back(d);
  HIP_OK(hipMemcpy(d, host, n * sizeof(T), hipMemcpyHostToDevice));
  *out = (const T*)d;
  return HTGX_OK;
}

} /* namespace */

/*******************************************************************************
 * Context
 ******************************************************************************/
/* Ground materials: grey albedos, or BRDF lists (sorted, reflectivities in
 * [0, 1]); interfaces whose materials exist and of which one reflects */
static int
check_materials(const htgx_ground_desc& g)
{
  if(!g.nmtls) return HTGX_BAD_ARG;
  if(g.mtl_brdf_first) {
    if(!g.mtl_brdf_list || g.mtl_brdf_first[0] != 0) return HTGX_BAD_ARG;
    if(g.mtl_brdf_first[g.nmtls] && !g.brdfs) return HTGX_BAD_ARG;
    for(uint32_t m = 0; m < g.nmtls; ++m) {
      const uint32_t b0 = g.mtl_brdf_first[m], b1 = g.mtl_brdf_first[m + 1];
      const int list = g.mtl_brdf_list[m];
      if(b1 < b0) return HTGX_BAD_ARG;
      if(list != HTGX_BRDF_LIST_WAVELENGTHS && list != HTGX_BRDF_LIST_BANDS) return HTGX_BAD_ARG;
      for(uint32_t i = b0; i < b1; ++i) {
        const htgx_brdf& e = g.brdfs[i];
        if(e.type != HTGX_BRDF_LAMBERTIAN && e.type != HTGX_BRDF_SPECULAR) return HTGX_BAD_ARG;
        if(!(e.reflectivity >= 0 && e.reflectivity <= 1)) return HTGX_BAD_ARG;
        if(!(e.wlen[0] >= 0)) return HTGX_BAD_ARG;
        if(list == HTGX_BRDF_LIST_WAVELENGTHS ? e.wlen[1] != e.wlen[0] : !(e.wlen[1] >= e.wlen[0]))
          return HTGX_BAD_ARG;
        if(i > b0) {
          const htgx_brdf& p = g.brdfs[i - 1];
          if(list == HTGX_BRDF_LIST_WAVELENGTHS ? !(e.wlen[0] > p.wlen[0]) : !(e.wlen[0] >= p.wlen[1]))
            return HTGX_BAD_ARG;
        }
      }
    }
  } else {
    if(!g.mtl_albedo) return HTGX_BAD_ARG;
  }
  if(g.interfaces) {
    if(!g.ninterfaces) return HTGX_BAD_ARG;
    if(g.sky_mtl < -1 || g.sky_mtl >= (int32_t)g.nmtls) return HTGX_BAD_ARG;
    auto reflects = [&](int32_t m) {
      return !g.mtl_brdf_first || g.mtl_brdf_first[m + 1] > g.mtl_brdf_first[m];
    };
    for(uint32_t i = 0; i < g.ninterfaces; ++i) {
      const int32_t* f = g.interfaces + 3 * (size_t)i;
      if(f[0] < 0 || f[0] >= (int32_t)g.nmtls || f[1] < 0 || f[1] >= (int32_t)g.nmtls) return HTGX_BAD_ARG;
      if(f[2] < -1 || f[2] >= (int32_t)g.nmtls) return HTGX_BAD_ARG;
      if(!reflects(f[0]) && !reflects(f[1]) && !(f[2] >= 0 && reflects(f[2]))) return HTGX_BAD_ARG;
    }
  }
  return HTGX_OK;
}

static int
check_scene(const htgx_scene_desc* s)
{
  if(!s) return HTGX_BAD_ARG;
  const htgx_spectral_desc& sp = s->spectral;
  if(!sp.nbands || !sp.band_lo || !sp.band_hi || !sp.nquads || !sp.quad_weights)
    return HTGX_BAD_ARG;
  for(uint32_t b = 0; b < sp.nbands; ++b) {
    if(!sp.nquads[b] || !(sp.band_hi[b] >= sp.band_lo[b])) return HTGX_BAD_ARG;
    if(b && sp.band_lo[b] < sp.band_hi[b - 1]) return HTGX_BAD_ARG;
  }
  /* a scene may have no gas layers (nlayers = 0: tests, cloud-only media) */
  if(s->gas.nlayers && (!(s->gas.layer_thickness > 0) || !s->gas.ka || !s->gas.ks))
    return HTGX_BAD_ARG;
  /* coefficients finite and in range: NaN or infinite extinctions would
   * leave the tracking loops with no progress to make */
  auto nonneg = [](const double* v, size_t n) {
    for(size_t i = 0; i < n; ++i) if(!(v[i] >= 0 && v[i] <= DBL_MAX)) return false;
    return true;
  };
  uint64_t nq = 0;
  for(uint32_t b = 0; b < sp.nbands; ++b) nq += sp.nquads[b];
  if(s->gas.nlayers) {
    if(!nonneg(s->gas.ka, nq * s->gas.nlayers) || !nonneg(s->gas.ks, (size_t)sp.nbands * s->gas.nlayers))
      return HTGX_BAD_ARG;
    if(s->gas.temperature && !nonneg(s->gas.temperature, s->gas.nlayers)) return HTGX_BAD_ARG;
  }
  const htgx_cloud_desc& c = s->clouds;
  const bool clouds = c.def[0] || c.def[1] || c.def[2];
  if(clouds && (!c.def[0] || !c.def[1] || !c.def[2] || !c.lwc || !c.cext || !c.ssa || !c.hg_g))
    return HTGX_BAD_ARG;
  if(clouds) {
    for(int a = 0; a < 3; ++a)
      if(!(c.upper[a] > c.lower[a]) || !std::isfinite(c.lower[a]) || !std::isfinite(c.upper[a]))
        return HTGX_BAD_ARG;
    if(!nonneg(c.cext, sp.nbands)) return HTGX_BAD_ARG;
    for(uint32_t b = 0; b < sp.nbands; ++b)
      if(!(c.ssa[b] >= 0 && c.ssa[b] <= 1) || !(c.hg_g[b] > -1 && c.hg_g[b] < 1)) return HTGX_BAD_ARG;
    const size_t ncells = (size_t)c.def[0] * c.def[1] * c.def[2];
    bool ok = true;
    for(size_t i = 0; i < ncells; ++i) ok &= c.lwc[i] >= 0.f && c.lwc[i] <= FLT_MAX;
    if(!ok) return HTGX_BAD_ARG;
  }
  if(s->ground.type == HTGX_GROUND_PLANE || s->ground.type == HTGX_GROUND_MESH) {
    const htgx_ground_desc& g = s->ground;
    int res = check_materials(g);
    if(res != HTGX_OK) return res;
    const uint32_t nsurf = g.interfaces ? g.ninterfaces : g.nmtls;
    if(g.type == HTGX_GROUND_PLANE) {
      if(g.plane_mtl >= nsurf) return HTGX_BAD_ARG;
    } else {
      if(!g.nverts || !g.ntris || !g.verts || !g.tris || !g.tri_mtl) return HTGX_BAD_ARG;
      for(uint32_t t = 0; t < g.ntris; ++t) {
        if(g.tri_mtl[t] >= nsurf) return HTGX_BAD_ARG;
        for(int k = 0; k < 3; ++k) if(g.tris[3 * (size_t)t + k] >= g.nverts) return HTGX_BAD_ARG;
      }
    }
  } else if(s->ground.type != HTGX_GROUND_NONE) {
    return HTGX_BAD_ARG;
  }
  if(!(s->optical_thickness >= 0)) return HTGX_BAD_ARG;
  return HTGX_OK;
}

extern "C" int
htgx_octree_cache_prepare(const htgx_scene_desc* scene, const char* path, int* loaded, char* err,
                          size_t errlen)
{
  if(loaded) *loaded = 0;
  if(err && errlen) err[0] = '\0';
  if(!path) return HTGX_BAD_ARG;
  int res;
  if((res = check_scene(scene)) != HTGX_OK) return res;
  const htgx_cloud_desc& cl = scene->clouds;
  if(!(cl.def[0] && cl.def[1] && cl.def[2])) return HTGX_OK; /* no octree, no cache */
  std::string msg;
  const uint64_t digest = svo_cache_digest(*scene);
  SvoCache cache;
  bool found = false;
  res = svo_cache_read(path, *scene, digest, cache, &found, msg);
  if(res == HTGX_OK && !found) {
    SvoHost svo;
    if((res = svo_build(*scene, svo)) == HTGX_OK) {
      cache.leaf_block = svo.B;
      cache.nlevels = svo.nlevels;
      cache.optical_thickness = scene->optical_thickness;
      cache.nleaves = svo.nleaves;
      cache.lmap = std::move(svo.lmap);
      cache.bounds = std::move(svo.bounds);
      res = svo_cache_write(path, digest, cache, msg);
    } else {
      msg = "the octree cannot be built for this cloud grid and leaf block";
    }
  }
  if(loaded) *loaded = found ? 1 : 0;
  if(res != HTGX_OK && err && errlen) {
    std::strncpy(err, msg.c_str(), errlen - 1);
    err[errlen - 1] = '\0';
  }
  return res;
}

extern "C" int
htgx_leaf_block_for_definition(const uint32_t def[3], const uint32_t max_def[3],
                               uint32_t* leaf_block)
{
  if(!def || !max_def || !leaf_block) return HTGX_BAD_ARG;
  uint32_t B = 1, dmin = UINT32_MAX;
  for(int a = 0; a < 3; ++a) {
    if(!def[a] || !max_def[a]) return HTGX_BAD_ARG;
    B = std::max(B, (def[a] + max_def[a] - 1) / max_def[a]);
    dmin = std::min(dmin, def[a]);
  }
  for(; B <= dmin; ++B)
    if(def[0] % B == 0 && def[1] % B == 0 && def[2] % B == 0) { *leaf_block = B; return HTGX_OK; }
  return HTGX_BAD_ARG;
}

extern "C" int
htgx_debug_octree(const htgx_scene_desc* scene, int on_device, uint32_t* lmap, float* bounds,
                  uint64_t* nblocks, uint64_t* nnodes, uint64_t* nleaves, double* build_ms)
{
  int res;
  if(!nblocks || !nnodes || !nleaves || !build_ms) return HTGX_BAD_ARG;
  if((res = check_scene(scene)) != HTGX_OK) return res;
  const htgx_cloud_desc& cl = scene->clouds;
  if(!(cl.def[0] && cl.def[1] && cl.def[2])) return HTGX_BAD_OP;
  SvoHost svo;
  if((res = svo_layout(*scene, svo)) != HTGX_OK) return res;
  const uint64_t nb = (uint64_t)svo.dims[0][0] * svo.dims[0][1] * svo.dims[0][2];
  *nblocks = nb;
  *nnodes = svo.nnodes;
  *nleaves = 0;
  *build_ms = 0;
  if(!lmap || !bounds) return HTGX_OK;
  if(!on_device) {
    const auto t0 = std::chrono::steady_clock::now();
    if((res = svo_build(*scene, svo)) != HTGX_OK) return res;
    *build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::memcpy(lmap, svo.lmap.data(), nb * sizeof(uint32_t));
    std::memcpy(bounds, svo.bounds.data(), svo.nnodes * sizeof(float2));
    *nleaves = svo.nleaves;
    return HTGX_OK;
  }
  const size_t ncells = (size_t)cl.def[0] * cl.def[1] * cl.def[2];
  float* d_lwc = nullptr;
  float2* d_bounds = nullptr;
  uint32_t* d_lmap = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  res = HTGX_UNKNOWN_ERR;
  if(hipMalloc(&d_lwc, ncells * sizeof(float)) == hipSuccess &&
     hipMalloc(&d_bounds, svo.nnodes * sizeof(float2)) == hipSuccess &&
     hipMalloc(&d_lmap, nb * sizeof(uint32_t)) == hipSuccess &&
     hipMemcpy(d_lwc, cl.lwc, ncells * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
     hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
     hipEventRecord(e0, nullptr) == hipSuccess) {
    uint64_t n = 0;
    res = svo_build_device(svo, d_lwc, cl.def, scene->optical_thickness, d_bounds, d_lmap, &n);
    float ms = 0.f;
    if(res == HTGX_OK &&
       (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess ||
        hipMemcpy(lmap, d_lmap, nb * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(bounds, d_bounds, svo.nnodes * sizeof(float2), hipMemcpyDeviceToHost) != hipSuccess))
      res = HTGX_UNKNOWN_ERR;
    *nleaves = n;
    *build_ms = ms;
  }
  if(e0) (void)hipEventDestroy(e0);
  if(e1) (void)hipEventDestroy(e1);
  if(d_lwc) (void)hipFree(d_lwc);
  if(d_bounds) (void)hipFree(d_bounds);
  if(d_lmap) (void)hipFree(d_lmap);
  return res;
}

/* Option -d: the octree's leaves per band and k-point as legacy VTK */
extern "C" int
htgx_dump_octrees_vtk(const htgx_scene_desc* scene, int fd)
{
  int res;
  if(fd < 0) return HTGX_BAD_ARG;
  if((res = check_scene(scene)) != HTGX_OK) return res;
  const htgx_scene_desc& sc = *scene;
  const htgx_cloud_desc& cl = sc.clouds;
  if(!(cl.def[0] && cl.def[1] && cl.def[2])) return HTGX_BAD_OP;
  SvoHost svo;
  if((res = svo_build(sc, svo)) != HTGX_OK) return res;

  /* the leaves: (level, node) from the first block of each */
  struct Leaf { uint32_t lo[3], hi[3]; float2 b; };
  std::vector<Leaf> leaves;
  const uint32_t* d0 = svo.dims[0];
  for(uint32_t z = 0; z < d0[2]; ++z)
    for(uint32_t y = 0; y < d0[1]; ++y)
      for(uint32_t x = 0; x < d0[0]; ++x) {
        const uint32_t w = svo.lmap[((size_t)z * d0[1] + y) * d0[0] + x];
        const uint32_t L = w >> kLeafIdBits, m = (1u << L) - 1u;
        if((x & m) || (y & m) || (z & m)) continue;
        Leaf f;
        const uint32_t c[3] = {x, y, z};
        for(int a = 0; a < 3; ++a) { f.lo[a] = c[a]; f.hi[a] = std::min(c[a] + (1u << L), d0[a]); }
        f.b = svo.bounds[w & kLeafIdMask];
        leaves.push_back(f);
      }

  const htgx_spectral_desc& sp = sc.spectral;
  const htgx_gas_desc& gas = sc.gas;
  std::string out;
  auto flush = [&]() {
    size_t off = 0;
    while(off < out.size()) {
      const ssize_t n = ::write(fd, out.data() + off, out.size() - off);
      if(n <= 0) return false;
      off += (size_t)n;
    }
    out.clear();
    return true;
  };
  auto put = [&](const char* fmt, auto... v) {
    char buf[256];
    const int n = snprintf(buf, sizeof(buf), fmt, v...);
    out.append(buf, (size_t)n);
  };
  uint32_t q0 = 0;
  for(uint32_t b = 0; b < sp.nbands; ++b) {
    for(uint32_t q = 0; q < sp.nquads[b]; ++q) {
      const uint32_t iq = q0 + q;
      out.append("# vtk DataFile Version 2.0\n");
      put("octree of band %u (%g-%g nm), k-point %u: kext bounds of its leaves\n", b,
          sp.band_lo[b], sp.band_hi[b], q);
      put("ASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS %zu double\n", leaves.size() * 8);
      for(const Leaf& f : leaves) {
        double p[2][3];
        for(int a = 0; a < 3; ++a) {
          p[0][a] = cl.lower[a] + f.lo[a] * svo.bs[a];
          p[1][a] = cl.lower[a] + f.hi[a] * svo.bs[a];
        }
        for(int k = 0; k < 8; ++k) /* VTK_VOXEL corner order: x fastest */
          put("%.9g %.9g %.9g\n", p[k & 1][0], p[(k >> 1) & 1][1], p[(k >> 2) & 1][2]);
        if(out.size() > (1u << 22) && !flush()) return HTGX_IO_ERR;
      }
      put("CELLS %zu %zu\n", leaves.size(), leaves.size() * 9);
      for(size_t i = 0; i < leaves.size(); ++i) {
        const size_t o = 8 * i;
        put("8 %zu %zu %zu %zu %zu %zu %zu %zu\n", o, o + 1, o + 2, o + 3, o + 4, o + 5, o + 6, o + 7);
        if(out.size() > (1u << 22) && !flush()) return HTGX_IO_ERR;
      }
      put("CELL_TYPES %zu\n", leaves.size());
      for(size_t i = 0; i < leaves.size(); ++i) {
        out.append("11\n");
        if(out.size() > (1u << 22) && !flush()) return HTGX_IO_ERR;
      }
      put("CELL_DATA %zu\n", leaves.size());
      for(int which = 0; which < 2; ++which) {
        put("SCALARS %s double 1\nLOOKUP_TABLE default\n", which ? "kext_max" : "kext_min");
        for(const Leaf& f : leaves) {
          /* gas extinction bounds over the layers the leaf overlaps (0 where
           * it reaches outside the layers) */
          double gmin = INFINITY, gmax = -INFINITY;
          const double z0 = cl.lower[2] + f.lo[2] * svo.bs[2], z1 = cl.lower[2] + f.hi[2] * svo.bs[2];
          if(!gas.nlayers || z0 < gas.z_bottom || z1 > gas.z_bottom + gas.nlayers * gas.layer_thickness) {
            gmin = 0; gmax = std::max(gmax, 0.0);
          }
          for(uint32_t j = 0; j < gas.nlayers; ++j) {
            const double lz0 = gas.z_bottom + j * gas.layer_thickness, lz1 = lz0 + gas.layer_thickness;
            if(lz1 <= z0 || lz0 >= z1) continue;
            const double k = gas.ka[(size_t)iq * gas.nlayers + j] + gas.ks[(size_t)b * gas.nlayers + j];
            gmin = std::min(gmin, k);
            gmax = std::max(gmax, k);
          }
          const double kc = cl.cext[b] * (which ? (double)f.b.y : (double)f.b.x);
          put("%.9g\n", kc + (which ? gmax : gmin));
          if(out.size() > (1u << 22) && !flush()) return HTGX_IO_ERR;
        }
      }
      out.append("---\n");
      if(!flush()) return HTGX_IO_ERR;
    }
    q0 += sp.nquads[b];
  }
  return HTGX_OK;
}

extern "C" int
htgx_ctx_create(const htgx_scene_desc* scene, htgx_ctx** out)
{
  int res;
  if(!out) return HTGX_BAD_ARG;
  if((res = check_scene(scene)) != HTGX_OK) return res;
  htgx_ctx* ctx = new(std::nothrow) htgx_ctx();
  if(!ctx) return HTGX_MEM_ERR;
  *out = nullptr;
  res = htgx_ctx_init(ctx, *scene);
  if(res != HTGX_OK) { htgx_ctx_destroy(ctx); return res; }
  *out = ctx;
  return HTGX_OK;
}

extern "C" int
htgx_ctx_create_cached(const htgx_scene_desc* scene, const char* cache_path, int* loaded,
                       char* err, size_t err_len, htgx_ctx** out)
{
  int res;
  if(!out || !cache_path || !*cache_path) return HTGX_BAD_ARG;
  if(err && err_len) err[0] = 0;
  if((res = check_scene(scene)) != HTGX_OK) return res;
  htgx_ctx* ctx = new(std::nothrow) htgx_ctx();
  if(!ctx) return HTGX_MEM_ERR;
  *out = nullptr;
  ctx->cache_path = cache_path;
  res = htgx_ctx_init(ctx, *scene);
  ctx->cache_path = nullptr;
  if(loaded) *loaded = ctx->cache_loaded ? 1 : 0;
  if(err && err_len) std::snprintf(err, err_len, "%s", ctx->cache_err.c_str());
  if(res != HTGX_OK) { htgx_ctx_destroy(ctx); return res; }
  *out = ctx;
  return HTGX_OK;
}

int
htgx_ctx_init(htgx_ctx* ctx, const htgx_scene_desc& sc)
{
  int res;
  DevScene& S = ctx->S;
  std::memset(&S, 0, sizeof(S));
  HIP_OK(hipGetDevice(&ctx->device));
  {
    int rate_khz = 0;
    HIP_OK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
    ctx->us_per_tick = rate_khz > 0 ? 1000.0 / (double)rate_khz : 0.01;
  }

  /* Spectral tables */
  const uint32_t nb = sc.spectral.nbands;
  std::vector<uint32_t> qoff(nb + 1);
  uint32_t nq = 0;
  for(uint32_t b = 0; b < nb; ++b) { qoff[b] = nq; nq += sc.spectral.nquads[b]; }
  qoff[nb] = nq;
  std::vector<double> qcdf(nq);
  for(uint32_t b = 0; b < nb; ++b) {
    double acc = 0;
    for(uint32_t q = 0; q < sc.spectral.nquads[b]; ++q) {
      acc += sc.spectral.quad_weights[qoff[b] + q];
      qcdf[qoff[b] + q] = acc;
    }
    qcdf[qoff[b] + sc.spectral.nquads[b] - 1] = 1.0;
  }
  ctx->band_lo.assign(sc.spectral.band_lo, sc.spectral.band_lo + nb);
  ctx->band_hi.assign(sc.spectral.band_hi, sc.spectral.band_hi + nb);
  ctx->nquads.assign(sc.spectral.nquads, sc.spectral.nquads + nb);
  S.nbands = nb;
  if((res = upload(ctx, sc.spectral.band_hi, nb, &S.band_hi))) return res;
  if((res = upload(ctx, qoff.data(), qoff.size(), &S.quad_off))) return res;
  if((res = upload(ctx, qcdf.data(), qcdf.size(), &S.quad_cdf))) return res;

  /* Gas */
  const uint32_t nl = sc.gas.nlayers;
  S.nlayers = nl;
  S.zb = sc.gas.z_bottom;
  S.dzl = nl ? sc.gas.layer_thickness : 1.0;
  S.inv_dzl = 1.0 / S.dzl;
  if((res = upload(ctx, sc.gas.ka, (size_t)nq * nl, &S.gas_ka))) return res;
  if((res = upload(ctx, sc.gas.ks, (size_t)nb * nl, &S.gas_ks))) return res;
  if(sc.gas.temperature && nl) {
    if((res = upload(ctx, sc.gas.temperature, nl, &S.gas_T))) return res;
    ctx->has_temperature = true;
  }

  /* Clouds + octree */
  const htgx_cloud_desc& cl = sc.clouds;
  if(cl.def[0] && cl.def[1] && cl.def[2]) {
    /* the octree is built on the device from the uploaded LWC grid
     * (svo_device.hip); only its shape is computed here */
    SvoHost svo;
    SvoCache cache;
    uint64_t digest = 0;
    htgx_scene_desc lsc = sc;
    if(ctx->cache_path) {
      digest = svo_cache_digest(sc);
      bool found = false;
      if((res = svo_cache_read(ctx->cache_path, sc, digest, cache, &found, ctx->cache_err)))
        return res;
      if(found) { /* the cache's own threshold and leaf block (-T, -V ignored) */
        lsc.leaf_block = cache.leaf_block;
        lsc.optical_thickness = cache.optical_thickness;
        ctx->cache_loaded = true;
      }
    }
    if((res = svo_layout(lsc, svo)) != HTGX_OK) {
      if(ctx->cache_loaded)
        ctx->cache_err = std::string("the leaf block of the cache `") + ctx->cache_path
                       + "' does not fit the cloud grid";
      return res;
    }
    S.has_clouds = 1;
    S.B = svo.B;
    for(int a = 0; a < 3; ++a) {
      S.def[a] = cl.def[a];
      S.dims0[a] = svo.dims[0][a];
      S.lower[a] = cl.lower[a];
      S.bs[a] = svo.bs[a];
      S.inv_bs[a] = svo.inv_bs[a];
      S.inv_cs[a] = svo.inv_cs[a];
    }
    S.repeat = sc.repeat_clouds ? 1 : 0;
    for(int a = 0; a < 2; ++a) S.period[a] = (double)svo.dims[0][a] * svo.bs[a];
    const size_t ncells = (size_t)cl.def[0] * cl.def[1] * cl.def[2];
    if((res = upload(ctx, cl.lwc, ncells, &S.lwc))) return res;
    const uint64_t nblocks = (uint64_t)svo.dims[0][0] * svo.dims[0][1] * svo.dims[0][2];
    {
      float2* bounds = nullptr;
      uint32_t* lmap = nullptr;
      if((res = device_alloc(ctx, svo.nnodes, &bounds))) return res;
      if((res = device_alloc(ctx, nblocks, &lmap))) return res;
      if(ctx->cache_loaded) {
        if(cache.lmap.size() != nblocks || cache.bounds.size() != svo.nnodes
        || cache.nlevels != svo.nlevels) {
          ctx->cache_err = std::string("the cache `") + ctx->cache_path + "' does not fit the cloud grid";
          return HTGX_BAD_ARG;
        }
        HIP_OK(hipMemcpy(lmap, cache.lmap.data(), nblocks * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(bounds, cache.bounds.data(), svo.nnodes * sizeof(float2),
                         hipMemcpyHostToDevice));
        svo.nleaves = cache.nleaves;
      } else {
        if((res = svo_build_device(svo, S.lwc, cl.def, lsc.optical_thickness, bounds, lmap,
                                   &svo.nleaves)) != HTGX_OK)
          return res;
        if(ctx->cache_path) {
          cache.leaf_block = svo.B;
          cache.nlevels = svo.nlevels;
          cache.optical_thickness = lsc.optical_thickness;
          cache.nleaves = svo.nleaves;
          cache.lmap.resize(nblocks);
          cache.bounds.resize(svo.nnodes);
          HIP_OK(hipMemcpy(cache.lmap.data(), lmap, nblocks * sizeof(uint32_t), hipMemcpyDeviceToHost));
          HIP_OK(hipMemcpy(cache.bounds.data(), bounds, svo.nnodes * sizeof(float2),
                           hipMemcpyDeviceToHost));
          if((res = svo_cache_write(ctx->cache_path, digest, cache, ctx->cache_err))) return res;
        }
      }
      S.bounds = bounds;
      S.lmap = lmap;
    }
    if((res = upload(ctx, cl.cext, nb, &S.cext))) return res;
    if((res = upload(ctx, cl.ssa, nb, &S.ssa))) return res;
    if((res = upload(ctx, cl.hg_g, nb, &S.hg_g))) return res;
    ctx->nlevels = svo.nlevels;
    ctx->nnodes = svo.nnodes;
    ctx->nleaves = svo.nleaves;
    ctx->svo_bytes = nblocks * sizeof(uint32_t) + svo.nnodes * sizeof(float2);
  }

  /* Ground */
  S.ground_type = sc.ground.type;
  if(sc.ground.type == HTGX_GROUND_PLANE || sc.ground.type == HTGX_GROUND_MESH) {
    const htgx_ground_desc& g = sc.ground;
    S.plane_z = g.plane_z;
    S.plane_mtl = g.plane_mtl;
    if(g.mtl_albedo)
      if((res = upload(ctx, g.mtl_albedo, g.nmtls, &S.mtl_albedo))) return res;
    if(g.mtl_temperature)
      if((res = upload(ctx, g.mtl_temperature, g.nmtls, &S.mtl_T))) return res;
    ctx->nmtls = g.nmtls;
    S.gmat = g.mtl_brdf_first || g.interfaces;
    if(S.gmat) {
      /* the general model: every material a BRDF list (a grey albedo is a
       * one-entry Lambertian list), every surface an interface (material m
       * alone is {front m, back the sky}, the sky a material with no BRDF
       * appended after the others) */
      std::vector<DevMtl> mtls(g.nmtls + 1);
      std::vector<DevBrdf> brdfs;
      for(uint32_t m = 0; m < g.nmtls; ++m) {
        DevMtl& d = mtls[m];
        d.first = (uint32_t)brdfs.size();
        d.T = g.mtl_temperature ? g.mtl_temperature[m] : 0.0;
        if(g.mtl_brdf_first) {
          d.list = g.mtl_brdf_list[m];
          for(uint32_t i = g.mtl_brdf_first[m]; i < g.mtl_brdf_first[m + 1]; ++i) {
            const htgx_brdf& e = g.brdfs[i];
            brdfs.push_back(DevBrdf{e.wlen[0], e.wlen[1], e.reflectivity, e.type, 0});
          }
        } else {
          d.list = HTGX_BRDF_LIST_WAVELENGTHS;
          brdfs.push_back(DevBrdf{0.0, 0.0, g.mtl_albedo[m], HTGX_BRDF_LAMBERTIAN, 0});
        }
        d.count = (uint32_t)brdfs.size() - d.first;
      }
      mtls[g.nmtls] = DevMtl{(uint32_t)brdfs.size(), 0, HTGX_BRDF_LIST_WAVELENGTHS, 0, 0.0};
      std::vector<DevIfc> ifcs;
      if(g.interfaces) {
        for(uint32_t i = 0; i < g.ninterfaces; ++i)
          ifcs.push_back(DevIfc{g.interfaces[3 * i], g.interfaces[3 * i + 1], g.interfaces[3 * i + 2], 0});
        S.sky_mtl = g.sky_mtl;
        S.explicit_ifc = 1;
      } else {
        for(uint32_t m = 0; m < g.nmtls; ++m) ifcs.push_back(DevIfc{(int32_t)m, (int32_t)g.nmtls, -1, 0});
        S.sky_mtl = (int32_t)g.nmtls;
        S.explicit_ifc = 0;
      }
      if(brdfs.empty()) brdfs.push_back(DevBrdf{0.0, 0.0, 0.0, 0, 0}); /* no zero-size upload */
      if((res = upload(ctx, mtls.data(), mtls.size(), &S.mtls))) return res;
      if((res = upload(ctx, brdfs.data(), brdfs.size(), &S.brdfs))) return res;
      if((res = upload(ctx, ifcs.data(), ifcs.size(), &S.ifcs))) return res;
    }
  }
  if(sc.ground.type == HTGX_GROUND_MESH) {
    BvhHost bvh;
    if((res = bvh_build(sc.ground.verts, sc.ground.nverts, sc.ground.tris, sc.ground.tri_mtl,
                        sc.ground.ntris, bvh)) != HTGX_OK)
      return res;
    S.nbvh = (uint32_t)bvh.nodes.size();
    S.ground_repeat = sc.repeat_ground ? 1 : 0;
    for(int a = 0; a < 3; ++a) { S.glo[a] = bvh.lo[a]; S.ghi[a] = bvh.hi[a]; }
    if((res = upload(ctx, bvh.nodes.data(), bvh.nodes.size(), &S.bvh))) return res;
    if((res = upload(ctx, bvh.tris.data(), bvh.tris.size(), &S.tris))) return res;
    if((res = upload(ctx, bvh.mtl.data(), bvh.mtl.size(), &S.tri_mtl))) return res;
    ctx->bvh_depth = bvh.depth;
  }

  /* Sun: src/atmosphere/htrdr_atmosphere.c:89-111 + ..._sun.c:84-87 */
  {
    const double az = sc.sun_azimuth * kPi / 180.0;
    const double el = sc.sun_elevation * kPi / 180.0;
    S.sun_dir[0] = std::cos(el) * std::cos(az);
    S.sun_dir[1] = std::cos(el) * std::sin(az);
    S.sun_dir[2] = std::sin(el);
    S.sun_cos_half = std::cos(4.6524e-3);
    S.sun_solid_angle = 2.0 * kPi * (1.0 - S.sun_cos_half);
  }

  HIP_OK(hipMalloc(&ctx->d_counters, HTGX_NSTATS * sizeof(unsigned long long)));
  HIP_OK(hipMemset(ctx->d_counters, 0, HTGX_NSTATS * sizeof(unsigned long long)));
  return HTGX_OK;
}

extern "C" int
htgx_ctx_destroy(htgx_ctx* ctx)
{
  if(!ctx) return HTGX_BAD_ARG;
  for(void* p : ctx->bufs) (void)hipFree(p);
  if(ctx->d_partials) (void)hipFree(ctx->d_partials);
  if(ctx->d_wcdf) (void)hipFree(ctx->d_wcdf);
  if(ctx->d_wpdf) (void)hipFree(ctx->d_wpdf);
  if(ctx->d_cie_cdf) (void)hipFree(ctx->d_cie_cdf);
  if(ctx->d_counters) (void)hipFree(ctx->d_counters);
  htgx_wavefront_release(ctx);
  delete ctx;
  return HTGX_OK;
}

extern "C" int
htgx_ctx_get_svo_info(const htgx_ctx* ctx, uint32_t* nlevels, uint64_t* nnodes,
                      uint64_t* nleaves)
{
  if(!ctx) return HTGX_BAD_ARG;
  if(nlevels) *nlevels = ctx->nlevels;
  if(nnodes) *nnodes = ctx->nnodes;
  if(nleaves) *nleaves = ctx->nleaves;
  return HTGX_OK;
}

/*******************************************************************************
 * Map setup
 ******************************************************************************/
static void cross(double r[3], const double a[3], const double b[3])
{
  const double t0 = a[1]*b[2] - a[2]*b[1];
  const double t1 = a[2]*b[0] - a[0]*b[2];
  const double t2 = a[0]*b[1] - a[1]*b[0];
  r[0] = t0; r[1] = t1; r[2] = t2;
}

static double normalize(double r[3], const double a[3])
{
  const double len = std::sqrt(a[0]*a[0] + a[1]*a[1] + a[2]*a[2]);
  r[0] = a[0] / len; r[1] = a[1] / len; r[2] = a[2] / len;
  return len;
}

int
htgx_map_setup(htgx_ctx* ctx, const htgx_map_desc* d, DevMap* M, bool may_allocate)
{
  if(!ctx || !d || !M) return HTGX_BAD_ARG;
  /* check_draw_map_args, src/core/htrdr_draw_map.c:66-79 */
  if(!d->width || !d->height || !d->spp) return HTGX_BAD_ARG;
  if(d->width > 16384 || d->height > 16384) return HTGX_BAD_ARG; /* htrdr_args.c:61 */
  if(d->sensor != HTGX_SENSOR_CAMERA && d->sensor != HTGX_SENSOR_RECTANGLE) return HTGX_BAD_ARG;
  /* CIE XYZ flux maps are invalid (src/atmosphere/htrdr_atmosphere.c:189-194) */
  if(d->spectral_type == HTGX_SPECTRAL_SW_CIE_XYZ && d->sensor != HTGX_SENSOR_CAMERA)
    return HTGX_BAD_ARG;
  if(d->spectral_type != HTGX_SPECTRAL_SW && d->spectral_type != HTGX_SPECTRAL_LW
  && d->spectral_type != HTGX_SPECTRAL_SW_CIE_XYZ)
    return HTGX_BAD_ARG;
  if(d->spectral_type == HTGX_SPECTRAL_LW && !ctx->has_temperature) return HTGX_BAD_ARG;
  if(!(d->wlen_range[1] >= d->wlen_range[0])) return HTGX_BAD_ARG;
  if(d->solver != HTGX_SOLVER_PLAIN_FP64 && d->solver != HTGX_SOLVER_WAVEFRONT_FP32)
    return HTGX_BAD_ARG;

  std::memset(M, 0, sizeof(*M));
  M->width = d->width;
  M->height = d->height;
  M->spp = d->spp;
  M->spp_offset = d->spp_offset;
  M->seed = d->seed;
  M->sensor = d->sensor;
  M->spectral_type = d->spectral_type;
  M->us_per_tick = ctx->us_per_tick;

  double t[3];
  for(int a = 0; a < 3; ++a) { M->pos[a] = d->pos[a]; t[a] = d->tgt[a] - d->pos[a]; }
  if(!(normalize(M->z, t) > 0)) return HTGX_BAD_ARG;
  cross(t, M->z, d->up);
  if(!(normalize(M->x, t) > 0)) return HTGX_BAD_ARG;
  if(d->sensor == HTGX_SENSOR_CAMERA) {
    cross(M->y, M->x, M->z);
    M->aspect = (double)d->width / (double)d->height;
    if(d->projection == HTGX_PROJECTION_ORTHOGRAPHIC) {
      /* image plane height > 0 (htrdr_args_camera_orthographic_check) */
      if(!(d->ortho_height > 0)) return HTGX_BAD_ARG;
      M->ortho = 1;
      M->ortho_half_h = 0.5 * d->ortho_height;
    } else if(d->projection == HTGX_PROJECTION_PERSPECTIVE) {
      /* fov in ]0, 180[, lens radius >= 0, focal distance > 0
       * (htrdr_args_camera_perspective_check, src/core/htrdr_args.c) */
      if(!(d->fov_y > 0 && d->fov_y < 180)) return HTGX_BAD_ARG;
      if(!(d->lens_radius >= 0)) return HTGX_BAD_ARG;
      if(d->lens_radius > 0 && !(d->focal_dst > 0)) return HTGX_BAD_ARG;
      M->tan_half = std::tan(d->fov_y * kPi / 180.0 * 0.5);
      M->lens_radius = d->lens_radius;
      M->focal_dst = d->focal_dst;
    } else {
      return HTGX_BAD_ARG;
    }
  } else {
    /* src/core/htrdr_rectangle.c:93-131 */
    if(!(d->rect_size[0] > 0 && d->rect_size[1] > 0)) return HTGX_BAD_ARG;
    cross(t, M->z, M->x);
    if(!(normalize(M->y, t) > 0)) return HTGX_BAD_ARG;
    for(int a = 0; a < 3; ++a) {
      M->axis_x[a] = M->x[a] * d->rect_size[0] * 0.5;
      M->axis_y[a] = M->y[a] * d->rect_size[1] * 0.5;
    }
  }

  /* Wavelength sampler: htrdr_atmosphere_create, src/atmosphere/
   * htrdr_atmosphere.c:384-415, clamping the range to the sky's bands */
  double range[2] = {d->wlen_range[0], d->wlen_range[1]};
  range[0] = std::max(range[0], ctx->band_lo.front());
  range[1] = std::min(range[1], ctx->band_hi.back());
  if(!(range[1] >= range[0])) return HTGX_BAD_ARG;
  double Tref = d->ref_temperature;
  if(Tref <= 0) Tref = d->spectral_type == HTGX_SPECTRAL_LW ? 290.0 : 5778.0;
  const size_t nwb = (size_t)std::rint(range[1] - range[0]);
  M->range[0] = range[0];
  M->range[1] = range[1];
  M->Tref = Tref;
  if(d->spectral_type == HTGX_SPECTRAL_SW_CIE_XYZ) {
    /* htrdr_ran_wlen_cie_xyz_create (src/core/htrdr_ran_wlen_cie_xyz.c:
     * 288-329): 1-nm bands within [380, 780] nm */
    if(range[0] < 380.0 || range[1] > 780.0 || !nwb) return HTGX_BAD_ARG;
    M->spectral_type = HTGX_SPECTRAL_SW;
    M->wmode = 3;
    M->band_len = (range[1] - range[0]) / (double)nwb;
    M->nwbands = (uint32_t)nwb;
    const bool cached = ctx->d_cie_cdf && ctx->cie_nbands == nwb
      && ctx->cie_range[0] == range[0] && ctx->cie_range[1] == range[1];
    if(!cached) {
      if(!may_allocate) return HTGX_BAD_OP;
      /* setup_cie_xyz, :183-270: per band the trapezoidal integral of each
       * fit over 2 sub-intervals per nm (trapezoidal_integration, :55-78) */
      std::vector<double> cdf(3 * nwb);
      for(int c = 0; c < 3; ++c) {
        double* f = cdf.data() + (size_t)c * nwb;
        double sum = 0;
        for(size_t i = 0; i < nwb; ++i) {
          const double lo = range[0] + (double)i * M->band_len;
          const double hi = std::min(lo + M->band_len, range[1]);
          const size_t n = (size_t)(hi - lo) + 1;
          const double dl = (hi - lo) / (double)n;
          double integral = 0;
          for(size_t k = 0; k < n; ++k) {
            const double f1 = cie_fit(c, lo + dl * (double)k);
            const double f2 = cie_fit(c, lo + dl * (double)(k + 1));
            integral += (f1 + f2) * dl * 0.5;
          }
          f[i] = integral;
          sum += integral;
        }
        ctx->cie_rcp[c] = 1.0 / sum;
        for(size_t i = 0; i < nwb; ++i) {
          f[i] /= sum;
          if(i) f[i] += f[i - 1];
        }
        f[nwb - 1] = 1.0;
      }
      if(ctx->d_cie_cdf) { (void)hipFree(ctx->d_cie_cdf); ctx->d_cie_cdf = nullptr; }
      ctx->cie_nbands = 0;
      HIP_OK(hipMalloc(&ctx->d_cie_cdf, 3 * nwb * sizeof(double)));
      HIP_OK(hipMemcpy(ctx->d_cie_cdf, cdf.data(), 3 * nwb * sizeof(double), hipMemcpyHostToDevice));
      ctx->cie_nbands = nwb;
      ctx->cie_range[0] = range[0];
      ctx->cie_range[1] = range[1];
    }
    *M = htgx_cie_channel_map(ctx, *M, 0);
    M->seed = d->seed; /* channel maps derive their seeds from this one */
    return HTGX_OK;
  }
  if(nwb) {
    M->wmode = 1;
    const bool cached = ctx->d_wcdf && ctx->w_nbands == nwb && ctx->w_range[0] == range[0]
      && ctx->w_range[1] == range[1] && ctx->w_Tref == Tref;
    M->band_len = (range[1] - range[0]) / (double)nwb;
    if(!cached) {
      if(!may_allocate) return HTGX_BAD_OP;
      /* setup_wlen_planck_ran_cdf, src/core/htrdr_ran_wlen_planck.c:55-134 */
      std::vector<double> pdf(nwb), cdf(nwb);
      double sum = 0;
      for(size_t i = 0; i < nwb; ++i) {
        const double lo = range[0] + (double)i * M->band_len;
        const double hi = std::min(lo + M->band_len, range[1]);
        pdf[i] = blackbody_fraction(lo * 1e-9, hi * 1e-9, Tref);
        sum += pdf[i];
      }
      for(size_t i = 0; i < nwb; ++i) {
        pdf[i] /= sum;
        cdf[i] = i ? pdf[i] + cdf[i - 1] : pdf[i];
      }
      cdf[nwb - 1] = 1.0;
      if(ctx->d_wcdf) { (void)hipFree(ctx->d_wcdf); ctx->d_wcdf = nullptr; }
      if(ctx->d_wpdf) { (void)hipFree(ctx->d_wpdf); ctx->d_wpdf = nullptr; }
      HIP_OK(hipMalloc(&ctx->d_wcdf, nwb * sizeof(double)));
      HIP_OK(hipMalloc(&ctx->d_wpdf, nwb * sizeof(double)));
      HIP_OK(hipMemcpy(ctx->d_wcdf, cdf.data(), nwb * sizeof(double), hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(ctx->d_wpdf, pdf.data(), nwb * sizeof(double), hipMemcpyHostToDevice));
      ctx->w_nbands = nwb;
      ctx->w_range[0] = range[0];
      ctx->w_range[1] = range[1];
      ctx->w_Tref = Tref;
    }
    M->nwbands = (uint32_t)nwb;
    M->wcdf = ctx->d_wcdf;
    M->wpdf = ctx->d_wpdf;
  } else if(std::fabs(range[0] - range[1]) <= 1e-6) {
    M->wmode = 0;
  } else {
    M->wmode = 2;
  }
  return HTGX_OK;
}

DevMap htgx_cie_channel_map(const htgx_ctx* ctx, const DevMap& M, int c)
{
  DevMap C = M;
  C.cie_channel = c;
  C.cie_rcp = ctx->cie_rcp[c];
  C.wcdf = ctx->d_cie_cdf + (size_t)c * ctx->cie_nbands;
  C.wpdf = nullptr;
  C.seed = M.seed ^ ((uint64_t)(c + 1) << 56);
  return C;
}

/*******************************************************************************
 * Rendering
 ******************************************************************************/
static int
render_plain(htgx_ctx* ctx, const DevMap& M, htgx_mc_accum* accum, hipStream_t stream)
{
  const uint64_t npix = (uint64_t)M.width * M.height;
  const uint32_t ntx = (M.width + 7) / 8, nty = (M.height + 7) / 8;
  const uint64_t items_per_slice = (uint64_t)ntx * nty * 64;
  /* Enough threads to fill the 256 CUs several times over, each running as
   * many consecutive realisations of its pixel as possible. */
  const uint64_t target = 1ull << 21;
  uint32_t nslices = (uint32_t)std::min<uint64_t>(M.spp,
    std::max<uint64_t>(1, (target + items_per_slice - 1) / items_per_slice));
  const uint32_t spt = (M.spp + nslices - 1) / nslices;
  nslices = (M.spp + spt - 1) / spt;
  if(items_per_slice * nslices > 0xFFFFFFFFull) return HTGX_BAD_ARG;

  const size_t need = (size_t)nslices * npix * sizeof(htgx_mc_accum);
  if(need > ctx->partials_cap) {
    if(ctx->d_partials) { HIP_OK(hipFree(ctx->d_partials)); ctx->d_partials = nullptr; }
    HIP_OK(hipMalloc(&ctx->d_partials, need));
    ctx->partials_cap = need;
  }
  const uint64_t nthreads = items_per_slice * nslices;
  const uint32_t nblocks = (uint32_t)((nthreads + 255) / 256);
  if(ctx->counting) {
    HIP_OK(hipMemsetAsync(ctx->d_counters, 0, HTGX_NSTATS * sizeof(unsigned long long), stream));
    plain_kernel<true><<<nblocks, 256, 0, stream>>>(ctx->S, M, ntx, (uint32_t)items_per_slice,
      nslices, spt, (htgx_mc_accum*)ctx->d_partials, ctx->d_counters);
  } else {
    plain_kernel<false><<<nblocks, 256, 0, stream>>>(ctx->S, M, ntx, (uint32_t)items_per_slice,
      nslices, spt, (htgx_mc_accum*)ctx->d_partials, ctx->d_counters);
  }
  HIP_OK(hipGetLastError());
  merge_kernel<<<(uint32_t)((npix + 255) / 256), 256, 0, stream>>>(
    (const htgx_mc_accum*)ctx->d_partials, nslices, npix, accum);
  HIP_OK(hipGetLastError());
  ctx->last_nrealisations = npix * M.spp;
  return HTGX_OK;
}

extern "C" int
htgx_render_accum(htgx_ctx* ctx, const htgx_map_desc* desc, htgx_mc_accum* accum, void* stream)
{
  DevMap M;
  int res;
  if(!ctx || !desc || !accum) return HTGX_BAD_ARG;
  if((res = htgx_map_setup(ctx, desc, &M)) != HTGX_OK) return res;
  /* a CIE XYZ image is 3 maps, one per channel, into 3 accumulator planes
   * (draw_pixel_image, src/atmosphere/htrdr_atmosphere_draw_map.c:134-199) */
  const int nch = desc->spectral_type == HTGX_SPECTRAL_SW_CIE_XYZ ? 3 : 1;
  const size_t npix = (size_t)desc->width * desc->height;
  uint64_t nreal = 0;
  for(int c = 0; c < nch; ++c) {
    const DevMap Mc = nch == 3 ? htgx_cie_channel_map(ctx, M, c) : M;
    htgx_mc_accum* a = accum + (size_t)c * npix;
    res = desc->solver == HTGX_SOLVER_WAVEFRONT_FP32
      ? htgx_render_wavefront(ctx, Mc, a, (hipStream_t)stream)
      : render_plain(ctx, Mc, a, (hipStream_t)stream);
    if(res != HTGX_OK) return res;
    nreal += ctx->last_nrealisations;
  }
  ctx->last_nrealisations = nreal;
  return HTGX_OK;
}

extern "C" int
htgx_render_accum_noalloc(htgx_ctx* ctx, const htgx_map_desc* desc, htgx_mc_accum* accum,
                          void* stream)
{
  DevMap M;
  int res;
  if(!ctx || !desc || !accum) return HTGX_BAD_ARG;
  if(desc->solver != HTGX_SOLVER_WAVEFRONT_FP32) return HTGX_BAD_ARG;
  if((res = htgx_map_setup(ctx, desc, &M, false)) != HTGX_OK) return res;
  const int nch = desc->spectral_type == HTGX_SPECTRAL_SW_CIE_XYZ ? 3 : 1;
  const size_t npix = (size_t)desc->width * desc->height;
  uint64_t nreal = 0;
  for(int c = 0; c < nch; ++c) {
    const DevMap Mc = nch == 3 ? htgx_cie_channel_map(ctx, M, c) : M;
    if((res = htgx_render_wavefront_noalloc(ctx, Mc, accum + (size_t)c * npix,
                                            (hipStream_t)stream)) != HTGX_OK)
      return res;
    nreal += ctx->last_nrealisations;
  }
  ctx->last_nrealisations = nreal;
  return HTGX_OK;
}

extern "C" int
htgx_get_wavefront_params(htgx_ctx* ctx, htgx_wf_params* params)
{
  if(!ctx || !params) return HTGX_BAD_ARG;
  return htgx_wavefront_params(ctx, params, nullptr);
}

extern "C" int
htgx_set_wavefront_params(htgx_ctx* ctx, const htgx_wf_params* params)
{
  if(!ctx || !params) return HTGX_BAD_ARG;
  return htgx_wavefront_params(ctx, nullptr, params);
}

extern "C" int
htgx_autotune(htgx_ctx* ctx, const htgx_map_desc* desc, double budget_ms, htgx_wf_params* best,
              double* rate, uint32_t* ntrials)
{
  if(!ctx || !desc || !(budget_ms >= 0)) return HTGX_BAD_ARG;
  DevMap M;
  int res;
  if((res = htgx_map_setup(ctx, desc, &M)) != HTGX_OK) return res;
  return htgx_wavefront_autotune(ctx, M, budget_ms, best, rate, ntrials);
}

/* Checks of a batch and its device maps */
static int
batch_setup(const htgx_batch_item* items, uint32_t n, std::vector<DevMap>& Ms,
            std::vector<htgx_ctx*>& ctxs, std::vector<htgx_mc_accum*>& accums)
{
  if(!items || !n) return HTGX_BAD_ARG;
  int dev = -1, res;
  HIP_OK(hipGetDevice(&dev));
  Ms.resize(n);
  ctxs.resize(n);
  accums.resize(n);
  for(uint32_t i = 0; i < n; ++i) {
    const htgx_batch_item& it = items[i];
    if(!it.ctx || !it.desc || !it.accum || it.ctx->device != dev) return HTGX_BAD_ARG;
    for(uint32_t k = 0; k < i; ++k) if(items[k].ctx == it.ctx) return HTGX_BAD_ARG;
    if(it.desc->sensor != items[0].desc->sensor
    || it.desc->spectral_type != items[0].desc->spectral_type)
      return HTGX_BAD_ARG;
    /* a CIE XYZ image is 3 maps: batch its channels as separate items */
    if(it.desc->spectral_type == HTGX_SPECTRAL_SW_CIE_XYZ) return HTGX_BAD_OP;
    if((res = htgx_map_setup(it.ctx, it.desc, &Ms[i])) != HTGX_OK) return res;
    ctxs[i] = it.ctx;
    accums[i] = it.accum;
  }
  return HTGX_OK;
}

extern "C" int
htgx_render_batch(const htgx_batch_item* items, uint32_t n, void* stream, double* scene_ms)
{
  std::vector<DevMap> Ms;
  std::vector<htgx_ctx*> ctxs;
  std::vector<htgx_mc_accum*> accums;
  int res;
  if((res = batch_setup(items, n, Ms, ctxs, accums)) != HTGX_OK) return res;
  return htgx_render_wavefront_batch(ctxs.data(), Ms.data(), accums.data(), n,
                                     (hipStream_t)stream, scene_ms);
}

extern "C" int
htgx_batch_plan_create(const htgx_batch_item* items, uint32_t n, int use_graph,
                       htgx_batch_plan** plan)
{
  if(!plan) return HTGX_BAD_ARG;
  *plan = nullptr;
  std::vector<DevMap> Ms;
  std::vector<htgx_ctx*> ctxs;
  std::vector<htgx_mc_accum*> accums;
  int res;
  if((res = batch_setup(items, n, Ms, ctxs, accums)) != HTGX_OK) return res;
  return wf_plan_build(ctxs.data(), Ms.data(), accums.data(), n, true, use_graph != 0, plan);
}

extern "C" int
htgx_batch_autotune(const htgx_batch_item* items, uint32_t n, uint32_t spp_div, double budget_ms,
                    htgx_wf_params* best, double* rate, uint32_t* ntrials)
{
  if(!(budget_ms >= 0) || !spp_div) return HTGX_BAD_ARG;
  std::vector<DevMap> Ms;
  std::vector<htgx_ctx*> ctxs;
  std::vector<htgx_mc_accum*> accums;
  int res;
  if((res = batch_setup(items, n, Ms, ctxs, accums)) != HTGX_OK) return res;
  return htgx_wavefront_batch_autotune(ctxs.data(), Ms.data(), n, spp_div, budget_ms, best, rate,
                                       ntrials);
}

extern "C" int
htgx_batch_plan_step(htgx_batch_plan* plan, void* stream)
{
  if(!plan) return HTGX_BAD_ARG;
  return wf_plan_enqueue(plan, (hipStream_t)stream);
}

extern "C" int
htgx_batch_plan_rewind(htgx_batch_plan* plan)
{
  if(!plan) return HTGX_BAD_ARG;
  return wf_plan_rewind(plan);
}

extern "C" int
htgx_batch_plan_scene_ms(htgx_batch_plan* plan, double* scene_ms)
{
  if(!plan || !scene_ms) return HTGX_BAD_ARG;
  return wf_plan_times(plan, scene_ms);
}

extern "C" int
htgx_batch_plan_destroy(htgx_batch_plan* plan)
{
  if(!plan) return HTGX_BAD_ARG;
  (void)hipDeviceSynchronize();
  wf_plan_free(plan);
  return HTGX_OK;
}

extern "C" int
htgx_set_counting(htgx_ctx* ctx, int enable)
{
  if(!ctx) return HTGX_BAD_ARG;
  ctx->counting = enable ? 1 : 0;
  return HTGX_OK;
}

extern "C" int
htgx_get_counters(htgx_ctx* ctx, htgx_counters* c)
{
  if(!ctx || !c) return HTGX_BAD_ARG;
  unsigned long long h[HTGX_NSTATS];
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipMemcpy(h, ctx->d_counters, sizeof(h), hipMemcpyDeviceToHost));
  c->nrealisations = ctx->last_nrealisations;
  c->nleaf = h[1];
  c->ncoll = h[3];
  c->nbvh = h[6]; /* counted by the FP32 solver */
  c->ntri = h[7];
  c->nevents = h[5];
  c->nseg = h[2];
  c->nlwc = h[4];
  for(int i = 0; i < 16; ++i) c->diag[i] = h[8 + i];
  uint64_t tr[2];
  int res;
  if((res = htgx_wavefront_truncations(ctx, tr)) != HTGX_OK) return res;
  c->ntrunc_coll = tr[0];
  c->ntrunc_iter = tr[1];
  return HTGX_OK;
}

/* Estimate of an accumulator (src/core/htrdr_accum.h:48-68) */
static void
estimate(double S, double S2, uint64_t n, double* E, double* SE)
{
  if(!n) { *E = 0; *SE = 0; return; }
  const double N = (double)n;
  const double e = S / N;
  const double V = std::max(S2 / N - e * e, 0.0);
  *E = e;
  *SE = std::sqrt(V / N);
}

/* htrdr_brightness_temperature / htrdr_radiance_temperature,
 * src/core/htrdr_spectral.c:31-121 */
static double
radiance_temperature(double lmin, double lmax, double radiance)
{
  const size_t MAX_ITER = 100;
  double avg = radiance;
  if(lmin != lmax) avg /= (lmax - lmin);
  const double eps_T = 1e-4, eps_B = avg * 1e-8;
  double T = 0, T0 = 0, T1 = 0, T2 = 200, B, B0 = 0;
  size_t i;
  for(i = 0; i < MAX_ITER; ++i) {
    if(planck_range(lmin, lmax, T2) >= avg) break;
    T2 *= 2;
  }
  if(i >= MAX_ITER) return 0;
  for(i = 0; i < MAX_ITER; ++i) {
    T = (T1 + T2) * 0.5;
    B = planck_range(lmin, lmax, T);
    if(B < avg) T1 = T; else T2 = T;
    if(std::fabs(T - T0) < eps_T || std::fabs(B - B0) < eps_B) break;
    T0 = T;
    B0 = B;
  }
  if(i >= MAX_ITER) return 0;
  return T;
}

/* Size of the reference pixel struct of a map (atmosphere_get_pixel_format,
 * src/atmosphere/htrdr_atmosphere.c:482-512) */
static size_t pixel_size(const htgx_map_desc* d)
{
  if(d->sensor == HTGX_SENSOR_RECTANGLE) return sizeof(htgx_pixel_flux);
  if(d->spectral_type == HTGX_SPECTRAL_SW_CIE_XYZ) return sizeof(htgx_pixel_image);
  return sizeof(htgx_pixel_xwave);
}

extern "C" int
htgx_finalize_pixels(const htgx_map_desc* d, const htgx_mc_accum* acc, void* pixels, size_t pitch)
{
  if(!d || !acc || !pixels || !d->width || !d->height) return HTGX_BAD_ARG;
  const size_t elmt = pixel_size(d);
  if(pitch < elmt * d->width) return HTGX_BAD_ARG;
  if(d->spectral_type == HTGX_SPECTRAL_SW_CIE_XYZ) {
    /* draw_pixel_image, src/atmosphere/htrdr_atmosphere_draw_map.c:201-206:
     * one estimate per channel, the time over all realisations */
    const size_t npix = (size_t)d->width * d->height;
    for(uint32_t y = 0; y < d->height; ++y) {
      char* row = (char*)pixels + (size_t)y * pitch;
      for(uint32_t x = 0; x < d->width; ++x) {
        htgx_pixel_image* p = (htgx_pixel_image*)(row + x * elmt);
        htgx_estimate* est[3] = {&p->X, &p->Y, &p->Z};
        p->time.sum_weights = p->time.sum_weights_sqr = 0;
        p->time.nweights = p->time.nfailures = 0;
        for(int c = 0; c < 3; ++c) {
          const htgx_mc_accum& a = acc[(size_t)c * npix + (size_t)y * d->width + x];
          estimate(a.sum + a.sum_c, a.sum2 + a.sum2_c, a.n, &est[c]->E, &est[c]->SE);
          p->time.sum_weights += a.tsum;
          p->time.sum_weights_sqr += a.tsum2;
          p->time.nweights += (size_t)a.n;
        }
      }
    }
    return HTGX_OK;
  }
  for(uint32_t y = 0; y < d->height; ++y) {
    char* row = (char*)pixels + (size_t)y * pitch;
    for(uint32_t x = 0; x < d->width; ++x) {
      const htgx_mc_accum& a = acc[(size_t)y * d->width + x];
      const double S = a.sum + a.sum_c, S2 = a.sum2 + a.sum2_c;
      htgx_accum time;
      time.sum_weights = a.tsum;
      time.sum_weights_sqr = a.tsum2;
      time.nweights = (size_t)a.n;
      time.nfailures = 0;
      if(d->sensor == HTGX_SENSOR_RECTANGLE) {
        htgx_pixel_flux* p = (htgx_pixel_flux*)(row + x * elmt);
        p->flux.sum_weights = S;
        p->flux.sum_weights_sqr = S2;
        p->flux.nweights = (size_t)a.n;
        p->flux.nfailures = 0;
        p->time = time;
      } else {
        htgx_pixel_xwave* p = (htgx_pixel_xwave*)(row + x * elmt);
        estimate(S, S2, a.n, &p->radiance.E, &p->radiance.SE);
        p->radiance_temperature.E = 0;
        p->radiance_temperature.SE = 0;
        if(d->spectral_type == HTGX_SPECTRAL_LW) {
          /* draw_pixel_xwave, src/atmosphere/htrdr_atmosphere_draw_map.c:425-435 */
          const double l0 = d->wlen_range[0] * 1e-9, l1 = d->wlen_range[1] * 1e-9;
          p->radiance_temperature.E = radiance_temperature(l0, l1, p->radiance.E);
          const double tmin = radiance_temperature(l0, l1, p->radiance.E - p->radiance.SE);
          const double tmax = radiance_temperature(l0, l1, p->radiance.E + p->radiance.SE);
          p->radiance_temperature.SE = tmax - tmin;
        }
        p->time = time;
      }
    }
  }
  return HTGX_OK;
}

extern "C" int
htgx_draw_map(htgx_ctx* ctx, const htgx_map_desc* desc, void* pixels, size_t pitch)
{
  if(!ctx || !desc || !pixels) return HTGX_BAD_ARG;
  if(!desc->width || !desc->height || !desc->spp) return HTGX_BAD_ARG;
  const size_t elmt = pixel_size(desc);
  if(pitch < elmt * desc->width) return HTGX_BAD_ARG;
  const size_t npix = (size_t)desc->width * desc->height
    * (desc->spectral_type == HTGX_SPECTRAL_SW_CIE_XYZ ? 3 : 1);
  htgx_mc_accum* d_acc = nullptr;
  HIP_OK(hipMalloc(&d_acc, npix * sizeof(htgx_mc_accum)));
  int res = HTGX_OK;
  std::vector<htgx_mc_accum> h(npix);
  if(hipMemset(d_acc, 0, npix * sizeof(htgx_mc_accum)) != hipSuccess) res = HTGX_UNKNOWN_ERR;
  if(res == HTGX_OK) res = htgx_render_accum(ctx, desc, d_acc, nullptr);
  if(res == HTGX_OK && hipMemcpy(h.data(), d_acc, npix * sizeof(htgx_mc_accum),
                                 hipMemcpyDeviceToHost) != hipSuccess)
    res = HTGX_UNKNOWN_ERR;
  (void)hipFree(d_acc);
  if(res == HTGX_OK) res = htgx_finalize_pixels(desc, h.data(), pixels, pitch);
  return res;
}

/*******************************************************************************
 * Debug entry points
 ******************************************************************************/
/* Device array owned for the duration of a debug call: freed on every
 * return path */
template<class T>
struct DevArray {
  T* p = nullptr;
  ~DevArray() { if(p) (void)hipFree(p); }
  int alloc(size_t n)
  {
    HIP_OK(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)));
    return HTGX_OK;
  }
  int upload(const T* h, size_t n)
  {
    int res;
    if((res = alloc(n))) return res;
    HIP_OK(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice));
    return HTGX_OK;
  }
  int download(T* h, size_t n) const
  {
    HIP_OK(hipMemcpy(h, p, n * sizeof(T), hipMemcpyDeviceToHost));
    return HTGX_OK;
  }
};

extern "C" int
htgx_debug_trace_segments(htgx_ctx* ctx, const double* rays, uint32_t nrays, int prop,
                          uint32_t iband, uint32_t iquad, htgx_segment* segs,
                          uint32_t max_segs, uint32_t* nsegs)
{
  if(!ctx || !rays || !segs || !nsegs || !nrays || !max_segs) return HTGX_BAD_ARG;
  if(prop < HTGX_Ka || prop > HTGX_Kext || iband >= ctx->S.nbands) return HTGX_BAD_ARG;
  if(iquad >= ctx->nquads[iband]) return HTGX_BAD_ARG;
  DevArray<double> d_rays;
  DevArray<htgx_segment> d_segs;
  DevArray<uint32_t> d_n;
  int res;
  if((res = d_rays.upload(rays, (size_t)nrays * 8))) return res;
  if((res = d_segs.alloc((size_t)nrays * max_segs))) return res;
  if((res = d_n.alloc(nrays))) return res;
  trace_segments_kernel<<<(nrays + 127) / 128, 128>>>(ctx->S, d_rays.p, nrays, prop, iband,
                                                       iquad, d_segs.p, max_segs, d_n.p);
  HIP_OK(hipGetLastError());
  if((res = d_segs.download(segs, (size_t)nrays * max_segs))) return res;
  return d_n.download(nsegs, nrays);
}

extern "C" int
htgx_debug_trace_segments_fp32(htgx_ctx* ctx, const double* rays, uint32_t nrays, int prop,
                               uint32_t iband, uint32_t iquad, htgx_segment* segs,
                               uint32_t max_segs, uint32_t* nsegs)
{
  if(!ctx || !rays || !segs || !nsegs || !nrays || !max_segs) return HTGX_BAD_ARG;
  if(prop < HTGX_Ka || prop > HTGX_Kext || iband >= ctx->S.nbands) return HTGX_BAD_ARG;
  if(iquad >= ctx->nquads[iband]) return HTGX_BAD_ARG;
  return htgx_debug_trace_segments_fp32_impl(ctx, rays, nrays, prop, iband, iquad, segs, max_segs,
                                             nsegs);
}

extern "C" int
htgx_debug_realisations(htgx_ctx* ctx, const htgx_map_desc* desc, const uint64_t* pix,
                        const uint64_t* real, uint32_t n, double* weights)
{
  if(!ctx || !desc || !pix || !real || !weights || !n) return HTGX_BAD_ARG;
  const uint64_t npix = (uint64_t)desc->width * desc->height;
  for(uint32_t i = 0; i < n; ++i) if(pix[i] >= npix) return HTGX_BAD_ARG;
  if(desc->solver == HTGX_SOLVER_WAVEFRONT_FP32)
    return htgx_debug_realisations_fp32(ctx, desc, pix, real, n, weights);
  DevMap M;
  int res;
  if((res = htgx_map_setup(ctx, desc, &M)) != HTGX_OK) return res;
  const bool cie = desc->spectral_type == HTGX_SPECTRAL_SW_CIE_XYZ;
  for(uint32_t i = 0; i < n; ++i)
    if((real[i] >> 32) > (cie ? 2u : 0u)) return HTGX_BAD_ARG;
  /* one launch per channel over the realisations of that channel */
  for(int c = 0; c < (cie ? 3 : 1); ++c) {
    std::vector<uint64_t> p, r;
    std::vector<uint32_t> idx;
    for(uint32_t i = 0; i < n; ++i)
      if((int)(real[i] >> 32) == c) {
        idx.push_back(i);
        p.push_back(pix[i]);
        r.push_back(real[i] & 0xFFFFFFFFull);
      }
    if(idx.empty()) continue;
    const uint32_t m = (uint32_t)idx.size();
    const DevMap Mc = cie ? htgx_cie_channel_map(ctx, M, c) : M;
    DevArray<uint64_t> d_p, d_r;
    DevArray<double> d_w;
    if((res = d_p.upload(p.data(), m))) return res;
    if((res = d_r.upload(r.data(), m))) return res;
    if((res = d_w.alloc(m))) return res;
    realisations_kernel<<<(m + 63) / 64, 64>>>(ctx->S, Mc, d_p.p, d_r.p, m, d_w.p);
    HIP_OK(hipGetLastError());
    std::vector<double> w(m);
    if((res = d_w.download(w.data(), m))) return res;
    for(uint32_t k = 0; k < m; ++k) weights[idx[k]] = w[k];
  }
  return HTGX_OK;
}

extern "C" int
htgx_debug_threefry(const uint64_t* ctr, const uint64_t* key, uint32_t n, uint64_t* out)
{
  if(!ctr || !key || !out || !n) return HTGX_BAD_ARG;
  DevArray<uint64_t> d_c, d_k, d_o;
  int res;
  if((res = d_c.upload(ctr, 4 * (size_t)n))) return res;
  if((res = d_k.upload(key, 4 * (size_t)n))) return res;
  if((res = d_o.alloc(4 * (size_t)n))) return res;
  threefry_kernel<<<(n + 255) / 256, 256>>>(d_c.p, d_k.p, n, d_o.p);
  HIP_OK(hipGetLastError());
  return d_o.download(out, 4 * (size_t)n);
}

extern "C" int
htgx_debug_threefry2x32(const uint32_t* ctr, const uint32_t* key, uint32_t n, uint32_t* out)
{
  if(!ctr || !key || !out || !n) return HTGX_BAD_ARG;
  DevArray<uint32_t> d_c, d_k, d_o;
  int res;
  if((res = d_c.upload(ctr, 2 * (size_t)n))) return res;
  if((res = d_k.upload(key, 2 * (size_t)n))) return res;
  if((res = d_o.alloc(2 * (size_t)n))) return res;
  threefry2x32_kernel<<<(n + 255) / 256, 256>>>(d_c.p, d_k.p, n, d_o.p);
  HIP_OK(hipGetLastError());
  return d_o.download(out, 2 * (size_t)n);
}

extern "C" int
htgx_debug_sample_dirs(htgx_ctx* ctx, int kind, double param, const double dir_in[3],
                       const double* u, uint32_t n, double* out)
{
  if(!ctx || !dir_in || !u || !out || !n || kind < 0 || kind > 3) return HTGX_BAD_ARG;
  DevArray<double> d_u, d_o;
  int res;
  if((res = d_u.upload(u, 2 * (size_t)n))) return res;
  if((res = d_o.alloc(3 * (size_t)n))) return res;
  sample_dirs_kernel<<<(n + 255) / 256, 256>>>(ctx->S, kind, param, dir_in[0], dir_in[1],
                                               dir_in[2], d_u.p, n, d_o.p);
  HIP_OK(hipGetLastError());
  return d_o.download(out, 3 * (size_t)n);
}

extern "C" int
htgx_debug_mesh_trace(htgx_ctx* ctx, const htgx_scene_desc* scene, const double* rays,
                      const int64_t* prev, const int32_t* prev_edge, uint32_t n, int any,
                      double* t, int64_t* prim, uint32_t* nvisited)
{
  if(!scene || !rays || !prev || !prev_edge || !t || !prim || !nvisited || !n) return HTGX_BAD_ARG;
  const htgx_ground_desc& g = scene->ground;
  if(g.type != HTGX_GROUND_MESH) return HTGX_BAD_ARG;
  if(ctx && ctx->S.ground_type != HTGX_GROUND_MESH) return HTGX_BAD_ARG;
  BvhHost bvh;
  int res;
  if((res = bvh_build(g.verts, g.nverts, g.tris, g.tri_mtl, g.ntris, bvh)) != HTGX_OK) return res;
  if(ctx && ctx->S.nbvh != bvh.nodes.size()) return HTGX_BAD_ARG; /* not this scene's mesh */
  /* input triangle index <-> leaf order */
  std::vector<int32_t> leaf_of(g.ntris);
  for(uint32_t k = 0; k < (uint32_t)bvh.prim.size(); ++k) leaf_of[bvh.prim[k]] = (int32_t)k;
  std::vector<int32_t> lprev(n), ledge(n), ltri(n);
  for(uint32_t i = 0; i < n; ++i) {
    if(prev[i] >= (int64_t)g.ntris) return HTGX_BAD_ARG;
    lprev[i] = prev[i] < 0 ? -1 : leaf_of[prev[i]];
    ledge[i] = prev_edge[i] != 0;
  }
  if(!ctx) {
    for(uint32_t i = 0; i < n; ++i)
      mesh_trace_ray(bvh.nodes.data(), (uint32_t)bvh.nodes.size(), bvh.tris.data(), rays + 8 * (size_t)i,
                     lprev[i], ledge[i] != 0, any != 0, t + i, &ltri[i], nvisited + i);
  } else {
    DevArray<double> d_rays, d_t;
    DevArray<int32_t> d_prev, d_edge, d_tri;
    DevArray<uint32_t> d_nv;
    if((res = d_rays.upload(rays, 8 * (size_t)n))) return res;
    if((res = d_prev.upload(lprev.data(), n))) return res;
    if((res = d_edge.upload(ledge.data(), n))) return res;
    if((res = d_t.alloc(n)) || (res = d_tri.alloc(n)) || (res = d_nv.alloc(n))) return res;
    mesh_trace_kernel<<<(n + 255) / 256, 256>>>(ctx->S, d_rays.p, d_prev.p, d_edge.p, n, any,
                                                d_t.p, d_tri.p, d_nv.p);
    HIP_OK(hipGetLastError());
    if((res = d_t.download(t, n)) || (res = d_tri.download(ltri.data(), n))) return res;
    if((res = d_nv.download(nvisited, n))) return res;
  }
  for(uint32_t i = 0; i < n; ++i) prim[i] = ltri[i] < 0 ? -1 : (int64_t)bvh.prim[ltri[i]];
  return HTGX_OK;
}

extern "C" int
htgx_debug_copy(void* dst, const void* src, size_t nbytes, void* stream)
{
  if(!dst || !src || !nbytes || (nbytes % 16) || ((uintptr_t)dst % 16) || ((uintptr_t)src % 16))
    return HTGX_BAD_ARG;
  const size_t n = nbytes / 16;
  const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 256 * 64);
  copy_kernel<<<blocks, 256, 0, (hipStream_t)stream>>>((const uint4*)src, (uint4*)dst, n);
  HIP_OK(hipGetLastError());
  return HTGX_OK;
}
