/*
 * FP32 batched wavefront solver: a persistent kernel whose lanes advance
 * their own path by one micro-step per loop iteration.
 *
 * Structure (replaces the OpenMP tile loop of src/core/htrdr_draw_map.c:
 * 324-491 and the per-pixel loops of ..._draw_map.c:210-436):
 *  - work queue: chunks of K consecutive realisations of one pixel, handed
 *    out by atomic counters, one counter per XCD over a contiguous range of
 *    8x8 tiles (pixels of a tile in Morton order, src/core/htrdr_draw_map.c:
 *    301-319) so that the waves of an XCD trace neighbouring pixels and share
 *    that XCD's L2; an XCD whose range is empty takes chunks from the others;
 *  - micro-steps: in one iteration a lane either fetches the next octree x
 *    gas-layer segment of its current query (SEG), tests one tentative
 *    collision (COLL) -- the null-collision loops of ..._sw.c:94-143,
 *    174-215 and ..._lw.c:83-133 unrolled one collision at a time -- or
 *    handles an event (EVENT: scattering, surface hit, escape, the start and
 *    end of realisations). Lanes never wait for a slower lane's whole query;
 *  - event regrouping: the EVENT branch is the expensive one (phase and sun
 *    sampling, camera rays, spectral sampling); it runs only when at least
 *    `ev_thresh` lanes of the wavefront need it, or when no lane can do
 *    anything else, so that it executes with most of the wavefront active;
 *  - precision: positions, directions, distances, coefficients and weights
 *    are FP32; the k_min optical depth Tmin of the transmissivity estimator
 *    and the per-pixel accumulators are FP64 (Neumaier compensated);
 *  - results: every chunk writes its own partial accumulator; a merge kernel
 *    folds them per pixel in a fixed order, so a map does not depend on the
 *    schedule.
 * Random numbers: Threefry-2x32-20 blocks keyed by (pixel, realisation), with
 * the block index and the seed as counter; each block gives two FP32
 * canonicals and every draw site computes its own block.
 */
/* Multiply-adds fuse within a source expression only: the arithmetic, hence
 * every result, does not depend on the kernel the code is inlined into (the
 * single-map and batched launches give bit-identical maps) */
#pragma clang fp contract(on)

#include "htgx_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

using namespace htgx;

#define HIP_OK(expr) do {                                                     \
  hipError_t e_ = (expr);                                                    \
  if(e_ != hipSuccess) {                                                     \
    fprintf(stderr, "htgx: %s failed: %s (%s:%d)\n", #expr,                  \
      hipGetErrorString(e_), __FILE__, __LINE__);                            \
    return e_ == hipErrorOutOfMemory ? HTGX_MEM_ERR : HTGX_UNKNOWN_ERR;       \
  }                                                                           \
} while(0)

namespace {

constexpr float kPif = 3.14159265358979323846f;
constexpr int kXcds = 8;

/* FP32 copy of the scene constants (device arrays shared with the FP64
 * solver where the layout is the same) */
struct SceneF {
  uint32_t nbands;
  const double* band_hi;
  const uint32_t* quad_off;
  const double* quad_cdf;
  uint32_t nlayers;
  float zb, dzl, inv_dzl;
  const float* gas_ka; /* [#quads][nlayers] */
  const float* gas_ks; /* [nbands][nlayers] */
  /* vertical optical depth of gas_ka / gas_ks from the bottom of the gas to
   * the bottom of each layer, same layout (FP64 sums of the FP32 values: a
   * run's depth is a difference of two of them, which FP32 would round to
   * about 1e-7 of the whole column) */
  const double* cum_ka;
  const double* cum_ks;
  const float* gas_T;
  int has_clouds;
  uint32_t B;
  uint32_t def[3];
  uint32_t dims0[3];
  float lower[3], bs[3], inv_bs[3], inv_cs[3];
  int repeat;        /* clouds repeated along X and Y (-r) */
  float period[2];   /* repetition period: dims0 * bs */
  const float* lwc;
  /* per finest block: x = level-map word (level << 28 | leaf id), y = the
   * leaf's LWC min (low half) and max (high half) as FP16 multiples of
   * lwc_scale, rounded outwards */
  const uint2* leafrec;
  float lwc_scale;
  const float* cka; /* [nbands] cext*(1-ssa) */
  const float* cks; /* [nbands] cext*ssa */
  const float* hg_g;
  int ground_type;
  float plane_z;
  uint32_t plane_mtl;
  const BvhNode* bvh;   /* mesh ground (shared with the FP64 solver) */
  uint32_t nbvh;
  const BvhTri* tris;
  const uint32_t* tri_mtl;
  int ground_repeat;         /* -R: the mesh repeats along X, Y */
  float glo[3], ghi[3];      /* the mesh's bounding box */
  const float* mtl_albedo;
  const float* mtl_T;
  /* general materials (the FP64 solver's tables): tri_mtl and plane_mtl
   * then index interfaces */
  int gmat;
  int explicit_ifc;
  int32_t sky_mtl;
  const DevMtl* mtls;
  const DevBrdf* brdfs;
  const DevIfc* ifcs;
  float sun_dir[3];
  float sun_cos_half;
  float sun_solid_angle;
};

struct Launch {
  uint32_t width, height, ntx;
  uint32_t K;            /* realisations per chunk */
  uint32_t nsub;         /* chunks per pixel in this launch */
  uint32_t ntiles;
  uint32_t xcd_tiles;    /* tiles per XCD range */
  uint32_t spp_offset;
  uint32_t ev_thresh;    /* min #lanes needing the EVENT branch */
  uint32_t seg_thresh;   /* min #lanes needing the SEG step while others can collide */
  uint32_t max_steps;    /* micro-step budget of one query */
  unsigned int* counters; /* [kXcds] next chunk of each XCD range */
  htgx_mc_accum* partials; /* [ntiles*nsub*64] */
  unsigned long long* stats; /* event counters */
  /* batched launches only (null otherwise): set once the chunk queues are
   * empty; chunks completed, out of nvalid; wall clock of the last one */
  unsigned int* exhausted;
  unsigned int* done_count;
  unsigned long long* done_tick;
  uint32_t nvalid;
  /* queries cut short, counted whether or not counting is on (null: not
   * counted): [0] by the tentative-collision budget, [1] by the segment
   * iterator's bound on passes without a segment */
  unsigned long long* trunc;
  /* list launches only (null otherwise): nvalid (pixel, realisation) pairs,
   * one realisation each, whose accumulators go to partials[0..nvalid) */
  const uint32_t* list;
  /* single-map plans replayed from a HIP graph (null otherwise): the
   * realisation offset of the step, read at each chunk fetch instead of
   * spp_offset and advanced on the device after each step */
  const uint32_t* spp_offset_dev;
};

/* The arguments of the single-map kernels, one struct: the first argument
 * lies at offset 0 of the kernel argument segment, so its fields can be
 * read from there at any point of the kernel */
struct KArgs {
  SceneF S;
  DevMap M;
  Launch P;
};

enum { LS_SEG = 0, LS_COLL = 1, LS_EVENT = 2, LS_DONE = 3 };
enum { MODE_TR = 0, MODE_SC = 1, MODE_LW = 2 };
enum { EV_NONE = 0, EV_HIT = 1, EV_ABS = 2, EV_SCA = 3 };
enum {
  /* ST_ABS: the scattering event proper, entered from ST_SCATTER */
  ST_DIRECT = 0, ST_SCATTER = 1, ST_ABS = 2, ST_SUN = 3, ST_LW = 4,
  ST_WALK_DONE = 5, ST_REAL_DONE = 6, ST_BEGIN = 7, ST_FETCH = 8
};

/*******************************************************************************
 * Lane state
 ******************************************************************************/
/* Per-lane state the tracking steps use every iteration; it stays in
 * registers */
struct Lane {
  /* work */
  uint32_t pixel, r;    /* key of the realisation's random stream */
  uint32_t pair;        /* next Threefry-2x32 block of the realisation */
  /* realisation */
  uint32_t gka_off, gks_off; /* rows of gas ka (band, k-point) and ks (band) */
  float cka, cks;       /* particle absorption / scattering per LWC of the band */
  /* flags, one bit each in one word (bools would each hold a 64-bit lane
   * mask in scalar registers across the loop) */
  uint32_t surface : 1, rayleigh : 1, gh_hit : 1, pending : 1; /* walk */
  uint32_t prev_edge : 1, gh_edge : 1;
  uint32_t nq : 1;        /* a query is requested */
  uint32_t lazy : 1;      /* the range ends at the mesh's top, not yet cast */
  uint32_t itdone : 1, adv : 1, adv_g : 1, adv_c : 1, adv_end : 1; /* segment iterator */
  uint32_t s_cloud : 1;   /* the current segment is in the cloud grid */
  /* small states, bit fields of the same word */
  uint32_t lstate : 2;    /* LS_* */
  uint32_t mode : 2, prop : 2; /* the query's MODE_* and property */
  uint32_t nq_ground : 2; /* its range comes from a ground cast (1), or the
                             query resumes after a late cast (2) */
  uint32_t ev : 2;        /* EV_* of the query's end */
  uint32_t cstate : 3;    /* iterator: 0 before grid, 1 in leaf, 2 outside,
                             3 locate at tit, 4 step out of leaf */
  uint32_t exit_axis : 2; /* the current leaf's exit face */
  /* walk */
  float pos[3], dir[3];
  /* query */
  float qr;             /* end of the query range */
  float inv_dir[3];
  float tc, Ts, maj, inv_maj;   /* next tentative collision (INFINITY: none in the segment) */
  /* MODE_SC: the absorption transmissivity of the same path (k_min/k_max on
   * Ka, ..._sw.c:396-400) tracked alongside the scattering query: Tmin holds
   * its k_min optical depth, Ts_a the residual optical thickness left to its
   * next tentative collision, tc_a that collision in the current segment
   * (INFINITY: none), maj_a the segment's residual majorant */
  float Ts_a, tc_a, maj_a;
  double Tmin;
  /* two counts of the current query: tentative collisions in the low
   * kCollBits bits, and above them the passes of the segment iterator that
   * emitted nothing since its last segment */
  uint32_t nsteps;
  /* segment iterator */
  float tit, tnext;    /* iterator position; grid entry or leaf exit */
  uint32_t lohi[3];    /* leaf block range lo | hi << 16 */
  float lmn, lmx;       /* LWC bounds of the current leaf */
  int32_t j;
  /* current segment */
  float s_t1, s_kmin, s_kmax, s_gka, s_gks; /* gas ka / ks of the segment's layer */
  int32_t s_layer;
  /* counters */
  uint32_t c_leaf, c_seg, c_coll, c_lwc, c_ev, c_bvh, c_tri;
};

/* Per-lane state used only by the event step (realisation constants, walk
 * weights, ground-hit records, chunk sums), kept in LDS as arrays over the
 * block's lanes so that it takes no registers across the tracking steps.
 * Lane t touches element t of each array: no bank conflicts. */
constexpr int kBlock = 256;
struct LaneCold {
  double sum[kBlock], sum2[kBlock]; /* plain FP64 sums over the few realisations of a chunk */
  float pdf[kBlock], g[kBlock], L_sun[kBlock], wlen[kBlock];
  float org[3][kBlock], dir0[3][kBlock], cosN[kBlock], Ld[kBlock]; /* flux SW */
  uint32_t t0[kBlock];  /* realisation start, low bits of the wall clock */
  int stage[kBlock], mask[kBlock], phase[kBlock];
  float dir_next[3][kBlock];
  float ksi[kBlock], w[kBlock], Tr_abs[kBlock], R[kBlock], refl[kBlock], gas_ratio[kBlock],
        ev_dst[kBlock];
  int32_t ev_layer[kBlock];
  int32_t prev_prim[kBlock], gh_prim[kBlock]; /* previous surface hit (-1: none), current hit */
  uint32_t gh_mtl[kBlock];
  float tsum[kBlock], tsum2[kBlock];
  uint32_t n[kBlock], nfail[kBlock], chunk[kBlock], r_end[kBlock];
  float shift[2][kBlock]; /* -r: world X/Y of the origin of the lane's frame */
  uint32_t ntiles[kBlock]; /* -r: copies of the grid the query has crossed */
};
__shared__ LaneCold g_cold;
/* Cap on the copies of a repeated cloud grid one query crosses */
constexpr uint32_t kMaxTilesF = 1u << 16;
/* Lane::nsteps: tentative collisions of the query (at most the budget
 * Launch::max_steps = 2^20, then one more) in the low bits; passes of the
 * segment iterator without a segment above them, bounded by kMaxEmptyPasses
 * (a pass either emits a segment or crosses a boundary, so the bound only
 * stops an iterator that makes no progress) */
constexpr uint32_t kCollBits = 21;
constexpr uint32_t kCollMask = (1u << kCollBits) - 1u;
constexpr uint32_t kMaxEmptyPasses = 2000;
#define LC(f) (g_cold.f[threadIdx.x])
#define LCA(f, a) (g_cold.f[a][threadIdx.x])

/* Next pair of FP32 canonicals of the realisation: one Threefry-2x32-20 block
 * with counter (pair index, seed) under the key (pixel, realisation). The
 * numbers of a realisation depend only on its key and on how many pairs it
 * drew before, not on the lanes it shares a wavefront with. */
__device__ __forceinline__ float2 draw2(Lane& L, uint32_t key)
{
  const uint32_t ctr[2] = {L.pair, key};
  const uint32_t k[2] = {L.pixel, L.r};
  uint32_t out[2];
  threefry2x32_20(ctr, k, out);
  L.pair += 1;
  return make_float2((float)(out[0] >> 8) * 0x1.0p-24f, (float)(out[1] >> 8) * 0x1.0p-24f);
}

/* Exp(1) sample from a canonical, ssp_ran_exp(rng, 1) */
__device__ __forceinline__ float exp1f(float u)
{
  return -__logf(1.0f - u);
}

__device__ __forceinline__ void nsum(double& sum, double& c, double x)
{
  const double t = sum + x;
  if(fabs(sum) >= fabs(x)) c += (sum - t) + x;
  else c += (x - t) + sum;
  sum = t;
}

/* Planck's law in W/m^2/sr/m (src/core/htrdr_spectral.h:102-116) in FP32 */
__device__ __forceinline__ float planckf(float wlen_nm, float T)
{
  const float c1 = 1.191042972e-16f; /* 2 h c^2, W m^2 / sr */
  const float c2 = 1.438776877e-2f;  /* h c / k, m K */
  const float l = wlen_nm * 1e-9f;
  const float l2 = l * l;
  return c1 / (l2 * l2 * l) / (__expf(c2 / (l * T)) - 1.f);
}

/*******************************************************************************
 * FP32 geometry helpers
 ******************************************************************************/
__device__ __forceinline__ float dot3f(const float a[3], const float b[3])
{
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ void local_to_worldf(const float n[3], float cos_t, float phi, float out[3])
{
  const float sign = n[2] >= 0 ? 1.f : -1.f;
  const float a = -1.f / (sign + n[2]);
  const float b = n[0] * n[1] * a;
  const float b1x = 1.f + sign * n[0] * n[0] * a, b1y = sign * b, b1z = -sign * n[0];
  const float b2x = b, b2y = sign + n[1] * n[1] * a, b2z = -n[1];
  const float sin_t = sqrtf(fmaxf(0.f, 1.f - cos_t * cos_t));
  float sp, cp;
  __sincosf(phi, &sp, &cp);
  const float cx = cp * sin_t, cy = sp * sin_t;
  out[0] = b1x * cx + b2x * cy + n[0] * cos_t;
  out[1] = b1y * cx + b2y * cy + n[1] * cos_t;
  out[2] = b1z * cx + b2z * cy + n[2] * cos_t;
}

__device__ __forceinline__ float hg_cosf(float g, float u)
{
  if(fabsf(g) < 1e-6f) return 1.f - 2.f * u;
  const float t = (1.f - g * g) / (1.f - g + 2.f * g * u);
  const float c = (1.f + g * g - t * t) / (2.f * g);
  return fminf(1.f, fmaxf(-1.f, c));
}

__device__ __forceinline__ float hg_evalf(float g, float mu)
{
  const float d = 1.f + g * g - 2.f * g * mu;
  return (1.f - g * g) / (4.f * kPif * d * sqrtf(d));
}

__device__ __forceinline__ float rayleigh_cosf(float u)
{
  const float q = 4.f * u - 2.f;
  const float A = cbrtf(q + sqrtf(q * q + 1.f));
  return fminf(1.f, fmaxf(-1.f, A - 1.f / A));
}

__device__ __forceinline__ float rayleigh_evalf(float mu)
{
  return 3.f / (16.f * kPif) * (1.f + mu * mu);
}

/* Ground hit for the FP32 solver: the plane (primitive 0, a quad repeated to
 * infinity) or the mesh BVH, with the self-hit rules of
 * src/core/htrdr_geometry.c:116-150 relative to the previous surface hit */
struct GroundF {
  bool hit, edge;
  int32_t prim;
  uint32_t mtl;
  float t;
  uint32_t nodes, ntris; /* BVH work of the trace (counting build) */
};

/* World position of a point of the lane's frame (-r moves the frame with
 * the copies of the cloud grid); only a mesh ground needs it */
__device__ __forceinline__ void world_posf(const SceneF& S, const float p[3], float w[3])
{
  w[0] = p[0];
  w[1] = p[1];
  w[2] = p[2];
  if(S.repeat && S.ground_type == HTGX_GROUND_MESH) {
    w[0] += LCA(shift, 0);
    w[1] += LCA(shift, 1);
  }
}

__device__ __forceinline__ GroundF ground_tracef(const SceneF& S, const float o[3], const float d[3],
                                                 float r1, int32_t prev_prim, bool prev_edge,
                                                 bool any)
{
  GroundF g;
  g.hit = g.edge = false;
  g.prim = -1;
  g.mtl = 0;
  g.t = FLT_MAX;
  g.nodes = g.ntris = 0;
  if(S.ground_type == HTGX_GROUND_PLANE) {
    if(prev_prim != 0 && d[2] != 0.f) {
      const float tt = (S.plane_z - o[2]) / d[2];
      if(tt > 0.f && tt < r1) { g.hit = true; g.t = tt; g.prim = 0; g.mtl = S.plane_mtl; }
    }
  } else if(S.ground_type == HTGX_GROUND_MESH && !(o[2] > S.ghi[2] && d[2] >= 0.f)) {
    /* (a ray from above the mesh's top going up or level meets none of it) */
    const MeshHit m = mesh_trace_copies<float>(S.bvh, S.nbvh, S.tris, S.ground_repeat != 0, S.glo,
                                               S.ghi, o, d, 0.f, r1, prev_prim, prev_edge, any);
    g.nodes = m.nodes;
    g.ntris = m.ntris;
    if(m.tri >= 0) {
      g.hit = true;
      g.t = m.t;
      g.prim = m.tri;
      g.mtl = S.tri_mtl[m.tri];
      g.edge = mesh_on_edge(m.u, m.v);
    }
  }
  return g;
}

/* Unit geometric normal of a ground primitive: +z for the plane,
 * (v1 - v0) x (v2 - v0) for a mesh triangle; its side is the sky side */
__device__ __forceinline__ void ground_normalf(const SceneF& S, int32_t prim, float N[3])
{
  if(S.ground_type == HTGX_GROUND_MESH) {
    const BvhTri& T = S.tris[prim];
    N[0] = T.e1[1] * T.e2[2] - T.e1[2] * T.e2[1];
    N[1] = T.e1[2] * T.e2[0] - T.e1[0] * T.e2[2];
    N[2] = T.e1[0] * T.e2[1] - T.e1[1] * T.e2[0];
    const float inv = rsqrtf(dot3f(N, N));
#pragma unroll
    for(int a = 0; a < 3; ++a) N[a] *= inv;
  } else {
    N[0] = N[1] = 0.f;
    N[2] = 1.f;
  }
}

/* Mirror direction of dir about the unit normal N */
__device__ __forceinline__ void reflectf(const float dir[3], const float N[3], float out[3])
{
  const float c = 2.f * dot3f(dir, N);
#pragma unroll
  for(int a = 0; a < 3; ++a) out[a] = dir[a] - c * N[a];
}

/* The normal turned to face the incoming direction's origin side
 * (N . wo >= 0 with wo = -dir, ..._sw.c:402-420) */
__device__ __forceinline__ void face_forwardf(float N[3], const float dir[3])
{
  if(dot3f(N, dir) > 0.f) {
#pragma unroll
    for(int a = 0; a < 3; ++a) N[a] = -N[a];
  }
}

/* What the tracking steps read of the scene at every step (the cloud grid's
 * addressing, the leaf records, the random-stream key, the collision
 * budget). With VREG they are held in vector registers for the whole launch
 * (the empty asm statements make them lane values): in scalar registers
 * they stay live across the whole loop, where the kernel runs out of scalar
 * registers, and are reloaded from their spill slots at every step. The
 * kernels with the rare features have no vector registers to spare and keep
 * them in scalar ones. The mesh kernels keep the copy: the shortwave flux
 * one then spills 8 bytes per lane and still runs faster. */
struct TrackRegs {
  float lower[3], inv_cs[3];
  uint32_t def0, def1, B, key, max_steps;
  const float* lwc;
  const uint2* leafrec; /* the leaf records and their LWC scale */
  float lwc_scale;
};

template<bool VREG>
__device__ __forceinline__ TrackRegs track_regs(const SceneF& S, uint32_t key, uint32_t max_steps)
{
  TrackRegs g;
#pragma unroll
  for(int a = 0; a < 3; ++a) {
    g.lower[a] = S.lower[a];
    g.inv_cs[a] = S.inv_cs[a];
  }
  g.def0 = S.def[0];
  g.def1 = S.def[1];
  g.B = S.B;
  g.key = key;
  g.max_steps = max_steps;
  g.lwc = S.lwc;
  g.leafrec = S.leafrec;
  g.lwc_scale = S.lwc_scale;
  if(VREG) {
    asm volatile("" : "+v"(g.lower[0]), "+v"(g.lower[1]), "+v"(g.lower[2]), "+v"(g.inv_cs[0]),
                      "+v"(g.inv_cs[1]), "+v"(g.inv_cs[2]));
    asm volatile("" : "+v"(g.def0), "+v"(g.def1), "+v"(g.B), "+v"(g.key), "+v"(g.max_steps),
                      "+v"(g.lwc), "+v"(g.leafrec), "+v"(g.lwc_scale));
  }
  return g;
}

/*******************************************************************************
 * Segment iterator (FP32 version of SegIter in htgx_device.h)
 ******************************************************************************/
__device__ __forceinline__ float half_bits_to_float(uint32_t b)
{
  return (float)__builtin_bit_cast(_Float16, (uint16_t)b);
}

__device__ __forceinline__ float planef(const SceneF& S, int a, uint32_t k)
{
  return S.lower[a] + (float)k * S.bs[a];
}

template<bool COUNT>
__device__ __forceinline__ void load_leaf(const SceneF& S, const TrackRegs& G, Lane& L, uint32_t cx,
                                          uint32_t cy, uint32_t cz)
{
  /* one 8-byte load: the leaf's level and its LWC bounds */
  const uint2 rec = G.leafrec[(cz * S.dims0[1] + cy) * S.dims0[0] + cx];
  const uint32_t lvl = rec.x >> kLeafIdBits;
  L.lmn = half_bits_to_float(rec.y & 0xFFFFu) * G.lwc_scale;
  L.lmx = half_bits_to_float(rec.y >> 16) * G.lwc_scale;
  const uint32_t c[3] = {cx, cy, cz};
  float best = INFINITY;
  int ax = 0;
#pragma unroll
  for(int a = 0; a < 3; ++a) {
    const uint32_t lo = (c[a] >> lvl) << lvl;
    const uint32_t hi = min(lo + (1u << lvl), S.dims0[a]);
    L.lohi[a] = lo | (hi << 16);
    const float ta0 = (planef(S, a, L.dir[a] > 0.f ? hi : lo) - L.pos[a]) * L.inv_dir[a];
    const float ta = L.dir[a] != 0.f ? ta0 : INFINITY;
    const bool better = ta < best;
    best = better ? ta : best;
    ax = better ? a : ax;
  }
  L.tnext = best;
  L.exit_axis = ax;
  if(COUNT) L.c_leaf++;
}

/* Cell to load when the iterator is in state 3 (locate the block holding the
 * point at tit, clamped into the grid) or 4 (step across the exit face of
 * the current leaf, other axes clamped into its block range); false when the
 * step leaves the grid. One code path for both, so that a wavefront pays for
 * a single leaf load per pass whatever mix of states its lanes are in. */
__device__ __forceinline__ bool next_cell(const SceneF& S, Lane& L, uint32_t c[3])
{
  const bool step = L.cstate == 4;
  bool out = false;
  int oa = 0;
#pragma unroll
  for(int a = 0; a < 3; ++a) {
    const uint32_t lo = step ? L.lohi[a] & 0xFFFFu : 0u;
    const uint32_t hi = step ? L.lohi[a] >> 16 : S.dims0[a];
    const bool pos_dir = L.dir[a] > 0.f;
    const bool across = step && a == L.exit_axis;
    const bool out_a = across && (pos_dir ? hi >= S.dims0[a] : lo == 0);
    const float p = L.pos[a] + L.tit * L.dir[a];
    const float f = floorf((p - S.lower[a]) * S.inv_bs[a]);
    const uint32_t cl = f < (float)lo ? lo : f > (float)(hi - 1) ? hi - 1 : (uint32_t)f;
    const uint32_t cs = pos_dir ? hi : lo - 1;
    c[a] = across ? cs : cl;
    out = out || out_a;
    oa = out_a ? a : oa;
  }
  if(out) {
    /* -r: into the next copy of the grid along X or Y; the lane's frame
     * moves with it (a bounded number of copies per query) */
    if(oa == 2 || !S.repeat || ++LC(ntiles) > kMaxTilesF) return false;
#pragma unroll
    for(int a = 0; a < 2; ++a) {
      if(a == oa) {
        const bool pos_dir = L.dir[a] > 0.f;
        const float dp = pos_dir ? S.period[a] : -S.period[a];
        L.pos[a] -= dp;
        LCA(shift, a) += dp;
        c[a] = pos_dir ? 0u : S.dims0[a] - 1;
      }
    }
  }
  return true;
}

/* Start the iterator of the query over [0, qr] along (pos, dir); no memory
 * access here */
__device__ __forceinline__ void iter_init(const SceneF& S, Lane& L)
{
  const float r1 = L.qr;
  L.tit = 0.f;
  L.itdone = false;
  L.adv = false;
  L.cstate = 2;
  L.nsteps = 0;
  if(S.repeat) LC(ntiles) = 0;
#pragma unroll
  for(int a = 0; a < 3; ++a) L.inv_dir[a] = 1.f / L.dir[a];
  if(S.has_clouds) {
    float tn = -INFINITY, tf = INFINITY;
    bool miss = false;
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      if(S.repeat && a < 2) continue; /* -r: an infinite layer along X, Y */
      const float plo = planef(S, a, 0), phi = planef(S, a, S.dims0[a]);
      if(L.dir[a] != 0.f) {
        const float ta = (plo - L.pos[a]) * L.inv_dir[a];
        const float tb = (phi - L.pos[a]) * L.inv_dir[a];
        tn = fmaxf(tn, fminf(ta, tb));
        tf = fminf(tf, fmaxf(ta, tb));
      } else if(L.pos[a] < plo || L.pos[a] > phi) {
        miss = true;
      }
    }
    if(!miss && fmaxf(tn, 0.f) < fminf(tf, r1)) {
      if(S.repeat) {
        /* the frame of the copy of the grid holding the entry point */
        const float te = fmaxf(tn, 0.f);
#pragma unroll
        for(int a = 0; a < 2; ++a) {
          const float k = floorf((L.pos[a] + te * L.dir[a] - S.lower[a]) / S.period[a]);
          if(k != 0.f) {
            L.pos[a] -= k * S.period[a];
            LCA(shift, a) += k * S.period[a];
          }
        }
      }
      if(tn > 0.f) { L.cstate = 0; L.tnext = tn; }
      else L.cstate = 3;
    }
  }
  const float fj = floorf((L.pos[2] - S.zb) * S.inv_dzl);
  L.j = fj < -1.f ? -1 : fj > (float)S.nlayers ? (int32_t)S.nlayers : (int32_t)fj;
}

/* Move the iterator past a boundary at t1 (gas layer if adv_g, leaf or grid
 * entry if adv_c). No memory access: a leaf boundary only marks the leaf to
 * load next (state 3 or 4). */
__device__ __forceinline__ void iter_advance(Lane& L, float t1, bool adv_g, bool adv_c,
                                             bool at_end)
{
  if(t1 >= L.qr) { L.itdone = true; return; }
  L.tit = t1;
  if(adv_g) L.j += L.dir[2] > 0.f ? 1 : -1;
  if(adv_c) {
    if(L.cstate == 0) L.cstate = 3;
    else if(L.cstate == 1) L.cstate = 4;
  }
  if(at_end) L.itdone = true;
}

/* Vertical optical depth of the gas between heights za and zb (clamped into
 * the gas), for the property prop of the lane's band and k-point: at each
 * height, the cumulative column at its layer's bottom plus the layer's
 * coefficient times the height into it. Only the difference of the two
 * cumulative columns is taken in FP64 (the long sum that FP32 would round);
 * the parts inside the two layers are FP32 */
__device__ __forceinline__ float gas_depth(const SceneF& S, const Lane& L, int prop, float za, float zb)
{
  const float fa = floorf((za - S.zb) * S.inv_dzl), fb = floorf((zb - S.zb) * S.inv_dzl);
  const float top = (float)(S.nlayers - 1);
  const uint32_t ja = (uint32_t)fminf(fmaxf(fa, 0.f), top), jb = (uint32_t)fminf(fmaxf(fb, 0.f), top);
  const float ha = fminf(fmaxf(za - (S.zb + (float)ja * S.dzl), 0.f), S.dzl);
  const float hb = fminf(fmaxf(zb - (S.zb + (float)jb * S.dzl), 0.f), S.dzl);
  float d = 0.f;
  if(prop != HTGX_Ks) {
    const uint32_t o = L.gka_off;
    d += (float)(S.cum_ka[o + jb] - S.cum_ka[o + ja]) + (S.gas_ka[o + jb] * hb - S.gas_ka[o + ja] * ha);
  }
  if(prop != HTGX_Ka) {
    const uint32_t o = L.gks_off;
    d += (float)(S.cum_ks[o + jb] - S.cum_ks[o + ja]) + (S.gas_ks[o + jb] * hb - S.gas_ks[o + ja] * ha);
  }
  return fabsf(d);
}

/* Next segment holding some medium into L.s_*; false at the end of the range.
 * With `runs' (transmissivity queries, whose gas segments have k_min = k_max
 * and so no collisions), a stretch of the ray outside the cloud grid is one
 * segment over all the gas layers it crosses, up to the grid's entry, the
 * range's end or the edge of the gas, whose k_min = k_max is its mean: its
 * length times that is the stretch's optical depth, from the cumulative
 * column of the gas (gas_depth) instead of a sum over its layers.
 * The iterator only moves past an emitted segment at the next call, so that
 * the leaf range used to clamp the cells of its collisions stays the one of
 * the segment. Each pass of the loop loads at most one leaf, at one site. */
template<bool COUNT>
__device__ __forceinline__ bool iter_next(const SceneF& S, const TrackRegs& G, Lane& L, int prop,
                                          float* t0, bool runs = false)
{
  if(L.adv) {
    L.adv = false;
    iter_advance(L, L.s_t1, L.adv_g, L.adv_c, L.adv_end);
  }
  const float tmax = L.qr;
  for(;;) {
    if(L.itdone || !(L.tit < tmax)) return false;
    if(L.cstate >= 3) {
      uint32_t c[3];
      if(next_cell(S, L, c)) { load_leaf<COUNT>(S, G, L, c[0], c[1], c[2]); L.cstate = 1; }
      else L.cstate = 2;
    }
    const float t = L.tit;
    const float tc = L.cstate == 0 ? L.tnext : L.cstate == 1 ? fmaxf(L.tnext, t) : INFINITY;
    if(runs && L.cstate != 1 && L.j >= 0 && L.j < (int32_t)S.nlayers && L.dir[2] != 0.f) {
      /* a gas run: up to the grid's entry, the range's end or the edge of
       * the gas, whichever comes first */
      const float ze = L.dir[2] > 0.f ? S.zb + (float)S.nlayers * S.dzl : S.zb;
      const float tx = fmaxf((ze - L.pos[2]) * L.inv_dir[2], t);
      const float t1 = fminf(fminf(tc, tx), tmax);
      if(t1 > t) {
        const float dcol = gas_depth(S, L, prop, L.pos[2] + t * L.dir[2], L.pos[2] + t1 * L.dir[2]);
        const float k = dcol * fabsf(L.inv_dir[2]) / (t1 - t);
        *t0 = t;
        L.s_t1 = t1;
        L.s_kmin = L.s_kmax = k;
        L.s_gka = L.s_gks = 0.f;
        L.s_cloud = false;
        L.s_layer = L.j;
        /* past the run: the layer the ray is in just beyond t1 (or out of
         * the gas), the grid entered if the run ended there */
        const float zf = (L.pos[2] + t1 * L.dir[2] - S.zb) * S.inv_dzl;
        int32_t j1 = L.dir[2] > 0.f ? (int32_t)floorf(zf) : (int32_t)ceilf(zf) - 1;
        if(t1 == tx) j1 = L.dir[2] > 0.f ? (int32_t)S.nlayers : -1;
        L.j = j1 < -1 ? -1 : j1 > (int32_t)S.nlayers ? (int32_t)S.nlayers : j1;
        L.adv = true;
        L.adv_g = false;
        L.adv_c = tc == t1;
        L.adv_end = tc == INFINITY && t1 == tx;
        L.nsteps &= kCollMask;
        if(COUNT) L.c_seg++;
        return true;
      }
    }
    float tg;
    if(L.dir[2] > 0.f && L.j < (int32_t)S.nlayers) tg = (S.zb + (float)(L.j + 1) * S.dzl - L.pos[2]) * L.inv_dir[2];
    else if(L.dir[2] < 0.f && L.j >= 0) tg = (S.zb + (float)L.j * S.dzl - L.pos[2]) * L.inv_dir[2];
    else tg = INFINITY;
    tg = fmaxf(tg, t);
    const float t1 = fminf(fminf(tc, tg), tmax);
    const bool at_end = tg == INFINITY && tc == INFINITY;
    const bool in_gas = L.j >= 0 && L.j < (int32_t)S.nlayers;
    const bool in_cloud = L.cstate == 1;
    if(t1 > t && (in_gas || in_cloud)) {
      float gka = 0.f, gks = 0.f, cmn = 0.f, cmx = 0.f;
      if(in_gas) {
        gka = S.gas_ka[L.gka_off + (uint32_t)L.j];
        gks = S.gas_ks[L.gks_off + (uint32_t)L.j];
      }
      if(in_cloud) {
        const float c = prop == HTGX_Ka ? L.cka : prop == HTGX_Ks ? L.cks : L.cka + L.cks;
        cmn = c * L.lmn;
        cmx = c * L.lmx;
      }
      const float gk = prop == HTGX_Ka ? gka : prop == HTGX_Ks ? gks : gka + gks;
      *t0 = t;
      L.s_t1 = t1;
      L.s_kmin = gk + cmn;
      L.s_kmax = gk + cmx;
      L.s_gka = gka;
      L.s_gks = gks;
      L.s_cloud = in_cloud;
      L.s_layer = in_gas ? L.j : -1;
      L.adv = true;
      L.adv_g = tg == t1;
      L.adv_c = tc == t1;
      L.adv_end = at_end;
      L.nsteps &= kCollMask;
      if(COUNT) L.c_seg++;
      return true;
    }
    iter_advance(L, t1, tg == t1, tc == t1, at_end);
    if((L.nsteps >> kCollBits) >= kMaxEmptyPasses) { L.itdone = true; return false; }
    L.nsteps += 1u << kCollBits;
  }
}

/* LWC of the cell holding x, clamped into the current segment's leaf */
template<bool COUNT>
__device__ __forceinline__ float seg_lwcf(const TrackRegs& G, Lane& L, const float x[3])
{
  uint32_t ci[3];
#pragma unroll
  for(int a = 0; a < 3; ++a) {
    const uint32_t lo = (L.lohi[a] & 0xFFFFu) * G.B, hi = (L.lohi[a] >> 16) * G.B - 1;
    const float f = floorf((x[a] - G.lower[a]) * G.inv_cs[a]);
    ci[a] = f < (float)lo ? lo : f > (float)hi ? hi : (uint32_t)f;
  }
  if(COUNT) L.c_lwc++;
  return G.lwc[(ci[2] * G.def1 + ci[1]) * G.def0 + ci[0]];
}

/*******************************************************************************
 * Walk logic (EVENT branch). Stage functions only request the next query;
 * the query is set up once at the end of the event branch. `u' holds the
 * canonicals drawn for this event: slots 0-1 film / position / scattering
 * direction, 2-3 hemisphere / sun, 4-6 spectral / phase choice / Russian
 * roulette, 7-8 flux sun, the last one the optical thickness of the query.
 ******************************************************************************/
__device__ __forceinline__ void request(Lane& L, int mode, int prop, float r1)
{
  L.mode = mode;
  L.prop = prop;
  L.qr = r1;
  L.nq = true;
  L.nq_ground = 0;
}

__device__ __forceinline__ void request_ground(Lane& L, int mode, int prop)
{
  L.mode = mode;
  L.prop = prop;
  L.nq = true;
  L.nq_ground = 1;
}

__device__ __forceinline__ void sw_after_direct(Lane& L)
{
  if(!(LC(mask) & HTGX_RADIANCE_DIFFUSE)) { LC(stage) = ST_WALK_DONE; return; }
  LC(stage) = ST_SCATTER;
  request_ground(L, MODE_SC, HTGX_Ks);
}

template<bool COUNT>
__device__ __forceinline__ void count_ground(Lane& L, const GroundF& g)
{
  if(COUNT) { L.c_bvh += g.nodes; L.c_tri += g.ntris; }
}

template<bool COUNT>
__device__ __forceinline__ void sw_begin(const SceneF& S, Lane& L, int mask, const float p[3],
                                         const float d[3])
{
  LC(mask) = mask;
#pragma unroll
  for(int a = 0; a < 3; ++a) { L.pos[a] = p[a]; L.dir[a] = d[a]; }
  LCA(shift, 0) = LCA(shift, 1) = 0.f; /* a walk starts from a world position */
  LC(ksi) = 1.f;
  LC(w) = 0.f;
  LC(prev_prim) = -1;
  L.prev_edge = false;
  if((mask & HTGX_RADIANCE_DIRECT) && dot3f(L.dir, S.sun_dir) >= S.sun_cos_half) {
    float wp[3];
    world_posf(S, L.pos, wp);
    const GroundF g = ground_tracef(S, wp, L.dir, FLT_MAX, -1, false, true);
    count_ground<COUNT>(L, g);
    if(!g.hit) {
      LC(stage) = ST_DIRECT;
      request(L, MODE_TR, HTGX_Kext, FLT_MAX);
      return;
    }
  }
  sw_after_direct(L);
}

/* Next-event estimate and Russian roulette closing a SW event
 * (..._sw.c:494-504); the next scattering query starts from the event */
__device__ __forceinline__ void sw_accumulate(const SceneF& S, Lane& L, float Tr, float u_rr)
{
  LC(ksi) *= LC(Tr_abs);
  LC(w) += LC(ksi) * LC(L_sun) * Tr * LC(R) * S.sun_solid_angle;
  if(L.surface && u_rr >= LC(refl)) { LC(stage) = ST_WALK_DONE; return; }
#pragma unroll
  for(int a = 0; a < 3; ++a) L.dir[a] = LCA(dir_next, a);
  LC(stage) = ST_SCATTER;
  request_ground(L, MODE_SC, HTGX_Ks);
}

__device__ __forceinline__ float query_tr(const Lane& L)
{
  return L.ev == EV_HIT ? 0.f : (L.Tmin != 0.0 ? (float)exp(-L.Tmin) : 1.f);
}

/* Consume the finished query of the current stage */
template<bool COUNT, int SPECTRAL>
__device__ __forceinline__ void consume(const SceneF& S, Lane& L, const float* u)
{
  /* the stages of the other spectral domain never occur: compiled out */
  constexpr bool kLW = SPECTRAL == HTGX_SPECTRAL_LW;
  const int stage = LC(stage);
  if(kLW ? stage != ST_LW : stage == ST_LW) return;
  switch(stage) {
    case ST_DIRECT:
      LC(w) = LC(L_sun) * query_tr(L);
      sw_after_direct(L);
      break;

    case ST_SCATTER: {
      /* the scattering query (..._sw.c:355-375) and, tracked alongside it,
       * the absorption transmissivity up to its event (:396-400): a real
       * absorption collision on the way makes the latter 0 and ends the
       * walk, as an escape does */
      const bool hit = L.ev == EV_HIT;
      if(L.ev == EV_ABS || (!L.gh_hit && !hit)) { LC(stage) = ST_WALK_DONE; break; }
      if(COUNT) L.c_ev++;
      L.surface = !hit;
      const float dst = L.surface ? L.qr : LC(ev_dst);
      /* ..._sw.c:392: the next traces leave this surface hit, if any */
      LC(prev_prim) = L.surface ? LC(gh_prim) : -1;
      L.prev_edge = L.surface && L.gh_edge;
      LC(ev_dst) = dst;
      LC(Tr_abs) = L.Tmin != 0.0 ? (float)exp(-L.Tmin) : 1.f;
    }
      [[fallthrough]];
    case ST_ABS: {
      if(LC(Tr_abs) <= 0.f) { LC(stage) = ST_WALK_DONE; break; }
      const float dst = LC(ev_dst);
      float N[3] = {0.f, 0.f, 1.f};
#pragma unroll
      for(int a = 0; a < 3; ++a) L.pos[a] = L.pos[a] + L.dir[a] * dst;
      bool specular = false;
      if(L.surface) {
        ground_normalf(S, LC(gh_prim), N);
        if(S.gmat) {
          /* the hit side's material and its BRDF at the wavelength,
           * ..._sw.c:403-411 (the canonical of the phase choice) */
          const int32_t m = interface_mtl(S.ifcs[LC(gh_mtl)], S.mtls, dot3f(N, L.dir) > 0.f);
          const DevBrdf& b = brdf_fetch(S.mtls[m], S.brdfs, (double)LC(wlen), (double)u[4]);
          LC(refl) = (float)b.refl;
          specular = b.type == HTGX_BRDF_SPECULAR;
        } else {
          LC(refl) = S.mtl_albedo[LC(gh_mtl)];
        }
        face_forwardf(N, L.dir);
        float dn[3];
        if(specular) {
          reflectf(L.dir, N, dn);
        } else {
          local_to_worldf(N, sqrtf(1.f - u[0]), 2.f * kPif * u[1], dn);
        }
#pragma unroll
        for(int a = 0; a < 3; ++a) LCA(dir_next, a) = dn[a];
      } else {
        L.rayleigh = u[4] < LC(gas_ratio);
        const float mu = L.rayleigh ? rayleigh_cosf(u[0]) : hg_cosf(LC(g), u[0]);
        float dn[3];
        local_to_worldf(L.dir, mu, 2.f * kPif * u[1], dn);
#pragma unroll
        for(int a = 0; a < 3; ++a) LCA(dir_next, a) = dn[a];
        LC(refl) = 1.f;
      }
      float sun_dir[3];
      local_to_worldf(S.sun_dir, 1.f - u[2] * (1.f - S.sun_cos_half), 2.f * kPif * u[3], sun_dir);
      if(specular) {
        /* the mirror direction is the sun direction, of pdf 1 instead of
         * 1 / solid angle (..._sw.c:446-455) */
#pragma unroll
        for(int a = 0; a < 3; ++a) sun_dir[a] = LCA(dir_next, a);
        const bool lit = dot3f(sun_dir, S.sun_dir) >= S.sun_cos_half && dot3f(N, sun_dir) >= 0.f;
        LC(R) = lit ? LC(refl) / S.sun_solid_angle : 0.f;
      } else if(L.surface) {
        const float c = dot3f(N, sun_dir);
        LC(R) = c < 0.f ? 0.f : LC(refl) / kPif * c;
      } else {
        const float mu = dot3f(L.dir, sun_dir);
        LC(R) = L.rayleigh ? rayleigh_evalf(mu) : hg_evalf(LC(g), mu);
      }
      bool shadowed = LC(R) <= 0.f;
      if(!shadowed) {
        float wp[3];
        world_posf(S, L.pos, wp);
        const GroundF g = ground_tracef(S, wp, sun_dir, FLT_MAX, LC(prev_prim), L.prev_edge, true);
        count_ground<COUNT>(L, g);
        shadowed = g.hit;
      }
      if(shadowed) {
        sw_accumulate(S, L, 0.f, u[5]);
        break;
      }
#pragma unroll
      for(int a = 0; a < 3; ++a) L.dir[a] = sun_dir[a];
      LC(stage) = ST_SUN;
      request(L, MODE_TR, HTGX_Kext, FLT_MAX);
    } break;

    case ST_SUN:
      sw_accumulate(S, L, query_tr(L), u[5]);
      break;

    case ST_LW: {
      if(!L.gh_hit && L.ev == EV_NONE) { LC(w) = 0.f; LC(stage) = ST_WALK_DONE; break; }
      if(COUNT) L.c_ev++;
      const float dst = L.ev != EV_NONE ? LC(ev_dst) : L.qr;
#pragma unroll
      for(int a = 0; a < 3; ++a) L.pos[a] = L.pos[a] + L.dir[a] * dst;
      if(L.ev == EV_ABS) {
        const int32_t nl = (int32_t)S.nlayers;
        const int32_t jj = LC(ev_layer) < 0 ? 0 : LC(ev_layer) >= nl ? nl - 1 : LC(ev_layer);
        /* no gas layers: no emission temperature */
        LC(w) = S.gas_T && nl > 0 ? planckf(LC(wlen), S.gas_T[jj]) : 0.f;
        LC(stage) = ST_WALK_DONE;
        break;
      }
      float dn[3];
      if(L.ev == EV_SCA) {
        local_to_worldf(L.dir, hg_cosf(LC(g), u[0]), 2.f * kPif * u[1], dn);
        LC(prev_prim) = -1; /* ..._lw.c:236 */
        L.prev_edge = false;
      } else {
        float N[3];
        ground_normalf(S, LC(gh_prim), N);
        float refl, T;
        bool specular = false;
        if(S.gmat) { /* ..._lw.c:249-252 */
          const int32_t m = interface_mtl(S.ifcs[LC(gh_mtl)], S.mtls, dot3f(N, L.dir) > 0.f);
          const DevBrdf& b = brdf_fetch(S.mtls[m], S.brdfs, (double)LC(wlen), (double)u[4]);
          refl = (float)b.refl;
          specular = b.type == HTGX_BRDF_SPECULAR;
          T = (float)S.mtls[m].T;
        } else {
          refl = S.mtl_albedo[LC(gh_mtl)];
          T = S.mtl_T ? S.mtl_T[LC(gh_mtl)] : 0.f;
        }
        face_forwardf(N, L.dir);
        if(specular) {
          reflectf(L.dir, N, dn);
        } else {
          local_to_worldf(N, sqrtf(1.f - u[0]), 2.f * kPif * u[1], dn);
        }
        if(u[6] >= refl) {
          LC(w) = T > 0.f ? planckf(LC(wlen), T) : 0.f;
          LC(stage) = ST_WALK_DONE;
          break;
        }
        LC(prev_prim) = LC(gh_prim); /* ..._lw.c:272 */
        L.prev_edge = L.gh_edge;
      }
#pragma unroll
      for(int a = 0; a < 3; ++a) L.dir[a] = dn[a];
      LC(stage) = ST_LW;
      request_ground(L, MODE_LW, HTGX_Kext);
    } break;

    default: break;
  }
}

/* Sensor ray, wavelength, band and k-point, then the first walk
 * (..._draw_map.c:348-400 and :233-302) */
template<int SENSOR, int SPECTRAL, bool COUNT, bool RARE>
__device__ __forceinline__ void real_begin(const SceneF& S, const DevMap& M, Lane& L, const float* u)
{
  const uint32_t px = L.pixel % M.width, py = L.pixel / M.width;
  float org[3], dir[3];
  if(SENSOR == HTGX_SENSOR_CAMERA) {
    /* film sample u[0..1], lens sample u[2..3] (the camera_ray models of
     * htgx_device.h) */
    const float fx = ((float)px + u[0]) / (float)M.width;
    const float fy = ((float)py + u[1]) / (float)M.height;
    if(M.ortho) {
      const float sx = (2.f * fx - 1.f) * (float)(M.ortho_half_h * M.aspect);
      const float sy = (1.f - 2.f * fy) * (float)M.ortho_half_h;
#pragma unroll
      for(int a = 0; a < 3; ++a) {
        org[a] = (float)M.pos[a] + (float)M.x[a] * sx + (float)M.y[a] * sy;
        dir[a] = (float)M.z[a];
      }
    } else {
      const float sx = (2.f * fx - 1.f) * (float)(M.tan_half * M.aspect);
      const float sy = (1.f - 2.f * fy) * (float)M.tan_half;
#pragma unroll
      for(int a = 0; a < 3; ++a) {
        dir[a] = (float)M.z[a] + (float)M.x[a] * sx + (float)M.y[a] * sy;
        org[a] = (float)M.pos[a];
      }
      if(M.lens_radius > 0) {
        const float r = (float)M.lens_radius * sqrtf(u[2]), phi = 2.f * kPif * u[3];
        const float lx = r * cosf(phi), ly = r * sinf(phi);
#pragma unroll
        for(int a = 0; a < 3; ++a) {
          const float focus = org[a] + dir[a] * (float)M.focal_dst;
          org[a] += (float)M.x[a] * lx + (float)M.y[a] * ly;
          dir[a] = focus - org[a];
        }
      }
      const float inv = rsqrtf(dot3f(dir, dir));
#pragma unroll
      for(int a = 0; a < 3; ++a) dir[a] *= inv;
    }
  } else {
    const float sx = ((float)px + u[0]) / (float)M.width;
    const float sy = ((float)py + u[1]) / (float)M.height;
#pragma unroll
    for(int a = 0; a < 3; ++a)
      org[a] = (float)M.pos[a] + (float)M.axis_x[a] * (sx * 2.f - 1.f)
             + (float)M.axis_y[a] * (sy * 2.f - 1.f);
    const float up[3] = {0.f, 0.f, 1.f};
    /* a sample under the ground (the up ray meets the ground's sky side from
     * below, ..._draw_map.c:75-100) is rejected */
    const GroundF gu = ground_tracef(S, org, up, INFINITY, -1, false, false);
    count_ground<COUNT>(L, gu);
    if(gu.hit) {
      float N[3];
      ground_normalf(S, gu.prim, N);
      /* with interfaces: the material of the side the up ray travels in
       * (front if cos(up, -N) < 0) must be the sky's; else the ray must not
       * meet the side N points to from below */
      bool under;
      if(S.explicit_ifc) {
        const DevIfc& f = S.ifcs[gu.mtl];
        under = (N[2] > 0.f ? f.front : f.back) != S.sky_mtl;
      } else {
        under = N[2] >= 0.f;
      }
      if(under) { LC(stage) = ST_REAL_DONE; LC(nfail)++; LC(w) = -1.f; return; }
    }
    const float z[3] = {(float)M.z[0], (float)M.z[1], (float)M.z[2]};
    local_to_worldf(z, sqrtf(1.f - u[2]), 2.f * kPif * u[3], dir);
  }

  double pdf;
  const double wlen = sample_wavelength<RARE>(M, (double)u[4], (double)u[5], &pdf);
  LC(pdf) = (float)(pdf * 1e9);
  LC(wlen) = (float)wlen;
  uint32_t band;
  {
    uint32_t lo = 0, hi = S.nbands - 1;
    while(lo < hi) { const uint32_t mid = (lo + hi) >> 1; if(S.band_hi[mid] > wlen) hi = mid; else lo = mid + 1; }
    band = lo;
    const uint32_t off = S.quad_off[lo], nq = S.quad_off[lo + 1] - off;
    uint32_t q = 0;
    for(; q + 1 < nq; ++q) if(S.quad_cdf[off + q] > (double)u[6]) break;
    L.gka_off = (off + q) * S.nlayers;
    L.gks_off = lo * S.nlayers;
  }
  LC(g) = S.has_clouds ? S.hg_g[band] : 0.f;
  L.cka = S.has_clouds ? S.cka[band] : 0.f;
  L.cks = S.has_clouds ? S.cks[band] : 0.f;
  LC(L_sun) = planckf(LC(wlen), 5778.f);

  if(SPECTRAL == HTGX_SPECTRAL_LW) {
    LC(phase) = 0;
#pragma unroll
    for(int a = 0; a < 3; ++a) { L.pos[a] = org[a]; L.dir[a] = dir[a]; }
    LC(prev_prim) = -1;
    L.prev_edge = false;
    LC(w) = 0.f;
    LC(stage) = ST_LW;
    request_ground(L, MODE_LW, HTGX_Kext);
  } else if(SENSOR == HTGX_SENSOR_CAMERA) {
    LC(phase) = 0;
    sw_begin<COUNT>(S, L, HTGX_RADIANCE_ALL, org, dir);
  } else {
    float sun_dir[3];
    local_to_worldf(S.sun_dir, 1.f - u[7] * (1.f - S.sun_cos_half), 2.f * kPif * u[8], sun_dir);
    const float z[3] = {(float)M.z[0], (float)M.z[1], (float)M.z[2]};
    LC(cosN) = dot3f(z, sun_dir);
    LC(Ld) = 0.f;
#pragma unroll
    for(int a = 0; a < 3; ++a) { LCA(org, a) = org[a]; LCA(dir0, a) = dir[a]; }
    if(LC(cosN) > 0.f) { LC(phase) = 1; sw_begin<COUNT>(S, L, HTGX_RADIANCE_DIRECT, org, sun_dir); }
    else { LC(phase) = 2; sw_begin<COUNT>(S, L, HTGX_RADIANCE_DIFFUSE, org, dir); }
  }
}

template<int SENSOR, int SPECTRAL, bool COUNT>
__device__ __forceinline__ void walk_done(const SceneF& S, Lane& L)
{
  if(SENSOR == HTGX_SENSOR_CAMERA) {
    LC(stage) = ST_REAL_DONE;
    LC(w) = LC(w) / LC(pdf);
  } else if(SPECTRAL == HTGX_SPECTRAL_LW) {
    LC(stage) = ST_REAL_DONE;
    LC(w) = LC(w) * kPif / LC(pdf);
  } else if(LC(phase) == 1) {
    LC(Ld) = LC(w);
    LC(phase) = 2;
    float o[3], d[3];
#pragma unroll
    for(int a = 0; a < 3; ++a) { o[a] = LCA(org, a); d[a] = LCA(dir0, a); }
    sw_begin<COUNT>(S, L, HTGX_RADIANCE_DIFFUSE, o, d);
  } else {
    LC(stage) = ST_REAL_DONE;
    LC(w) = (LC(cosN) * S.sun_solid_angle * LC(Ld) + kPif * LC(w)) / LC(pdf);
  }
}

/* Next chunk from this XCD's tile range, else from the others */
__device__ __forceinline__ bool fetch_chunk(const Launch& P, uint32_t xcd, uint32_t* chunk)
{
  const uint32_t per_tile = P.nsub * 64u;
  for(uint32_t k = 0; k < (uint32_t)kXcds; ++k) {
    const uint32_t x = (xcd + k) % (uint32_t)kXcds;
    const uint32_t t0 = x * P.xcd_tiles;
    const uint32_t t1 = min(t0 + P.xcd_tiles, P.ntiles);
    if(t0 >= t1) continue;
    const uint32_t n = (t1 - t0) * per_tile;
    const uint32_t c = atomicAdd(P.counters + x, 1u);
    if(c < n) { *chunk = t0 * per_tile + c; return true; }
  }
  return false;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v)
{
  for(int off = 32; off > 0; off >>= 1) v += (uint64_t)__shfl_xor((unsigned long long)v, off, 64);
  return v;
}

/* Kernel features: without kFeatRare, the features that few maps use (cloud
 * and ground repetition, spectral BRDFs and interfaces, orthographic and
 * thin-lens cameras, CIE XYZ wavelength sampling, Planck wavelength sampling
 * over a continuous range) are compiled out; without kFeatMesh, the mesh
 * ground (its BVH casts, its deferred casts). Their flags become constants,
 * so that their code and registers do not weigh on the kernels of the maps
 * that do without them. */
enum { kFeatRare = 1, kFeatMesh = 2 };

/* The persistent loop of one wavefront over the chunks of one map */
/* A pointer to constant memory: the kernel argument segment, or a table in
 * device memory that no kernel writes while it is read */
template<class T>
using KPtr = const __attribute__((address_space(4))) T*;

/* The same address, opaque to the optimiser (an empty asm statement in a
 * scalar register): loads through it stay where they are written, in the
 * loop, instead of being hoisted out of it */
template<class T>
__device__ __forceinline__ KPtr<T> launder(KPtr<T> p)
{
  uint64_t a = (uint64_t)p;
  asm volatile("" : "+s"(a));
  return (KPtr<T>)a;
}

/* A copy of a constant-memory struct (scalar loads; the host compilation
 * pass, which only needs the declaration, gets an empty value) */
template<class T>
__device__ __forceinline__ T kload(KPtr<T> p)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return *p;
#else
  (void)p;
  return T{};
#endif
}

/* The scene and map of a kernel instance: the features it was not compiled
 * for (FEAT) read as constant zeros, so that their code folds away */
template<int FEAT>
__device__ __forceinline__ void fold_features(SceneF& S, DevMap& M)
{
  /* no mesh: the ground type is NONE or PLANE (0 or 1), so that every test
   * for HTGX_GROUND_MESH (2) folds to false */
  if(!(FEAT & kFeatMesh)) S.ground_type &= 1;
  if(!(FEAT & kFeatRare)) {
    S.repeat = 0;
    S.ground_repeat = 0;
    S.gmat = 0;
    S.explicit_ifc = 0;
    M.ortho = 0;
    M.lens_radius = 0.0;
  }
}

template<int SENSOR, int SPECTRAL, bool COUNT, int FEAT, bool LIST = false>
__device__ __forceinline__ void run_scene(KPtr<SceneF> pS, KPtr<DevMap> pM, KPtr<Launch> pP)
{
  constexpr bool RARE = (FEAT & kFeatRare) != 0;
  constexpr bool kLW = SPECTRAL == HTGX_SPECTRAL_LW;
  SceneF S = kload(pS);
  DevMap M = kload(pM);
  const Launch P = kload(pP);
  fold_features<FEAT>(S, M);
  /* canonicals per event: 8, plus 2 for the sun of a SW flux map */
  constexpr int NP = (SENSOR == HTGX_SENSOR_RECTANGLE && SPECTRAL == HTGX_SPECTRAL_SW) ? 5 : 4;
  Lane L;
  L.lstate = LS_EVENT;
  LC(stage) = ST_FETCH;
  L.pending = false;
  L.nq = false;
  L.lazy = false;
  L.c_leaf = L.c_seg = L.c_coll = L.c_lwc = L.c_ev = L.c_bvh = L.c_tri = 0;
  LC(n) = LC(nfail) = 0;
  LC(sum) = LC(sum2) = 0.0;
  LC(tsum) = LC(tsum2) = 0.f;
  L.pixel = L.r = 0;
  const uint32_t xcd = blockIdx.x % (uint32_t)kXcds;
  const TrackRegs G =
      track_regs<!RARE>(S, (uint32_t)M.seed ^ (uint32_t)(M.seed >> 32), P.max_steps);
  const uint32_t key = G.key;
  uint64_t d_it = 0, d_ce = 0, d_cl = 0, d_se = 0, d_sl = 0, d_ee = 0, d_el = 0;
  /* shader clock cycles per phase (counting build only) */
  uint64_t cy[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t tk = 0;
#define WF_TICK(i)                                                             \
  if(COUNT) { const uint64_t now = clock64(); cy[i] += now - tk; tk = now; }

  for(;;) {
    const uint64_t want0 = __ballot(L.lstate == LS_EVENT);
    const uint64_t busy0 = __ballot(L.lstate == LS_SEG || L.lstate == LS_COLL);
    if(!(busy0 | want0)) break;
    if(COUNT) tk = clock64();

    /* next segment of the query, and the first tentative collision in it:
     * while some lanes can take a collision step, the segment step waits
     * until seg_thresh lanes need it */
    const uint64_t bs = __ballot(L.lstate == LS_SEG);
    const bool run_seg = bs && (__popcll(bs) >= P.seg_thresh || !__ballot(L.lstate == LS_COLL));
    if(COUNT) {
      d_it++;
      d_se += run_seg;
      d_sl += run_seg ? __popcll(bs) : 0;
    }
    if(run_seg && L.lstate == LS_SEG) {
      float t0;
      /* longwave kernels run MODE_LW queries only, shortwave ones MODE_SC
       * and MODE_TR: the other modes' code is compiled out */
      const int sprop = kLW ? HTGX_Kext : L.mode == MODE_SC ? HTGX_Ks : L.prop;
      if(iter_next<COUNT>(S, G, L, sprop, &t0, !kLW && L.mode == MODE_TR)) {
        const float len = L.s_t1 - t0;
        if(!kLW && L.mode == MODE_TR) L.Tmin += (double)len * (double)L.s_kmin;
        if(!kLW && L.mode == MODE_SC) {
          /* the absorption transmissivity over the segment: its k_min part
           * (taken back beyond a scattering event) and the first tentative
           * collision of its residual tracking */
          const float amin = L.s_gka + (L.s_cloud ? L.cka * L.lmn : 0.f);
          const float maj_a = L.s_cloud ? L.cka * (L.lmx - L.lmn) : 0.f;
          L.Tmin += (double)len * (double)amin;
          const float Ta = len * maj_a;
          if(L.Ts_a > Ta || !(maj_a > 0.f)) {
            L.Ts_a -= Ta;
          } else {
            L.maj_a = maj_a;
            L.tc_a = t0 + L.Ts_a / maj_a;
            L.lstate = LS_COLL;
          }
        }
        const float maj = !kLW && L.mode == MODE_TR ? L.s_kmax - L.s_kmin : L.s_kmax;
        const float T = len * maj;
        if(L.Ts > T || !(maj > 0.f)) {
          L.Ts -= T; /* no collision in this segment */
        } else {
          L.maj = maj;
          L.inv_maj = 1.f / maj;
          L.tc = t0 + L.Ts * L.inv_maj;
          L.lstate = LS_COLL;
        }
      } else if(L.lazy) {
        /* the query reached the top of the mesh ground without an event:
         * cast the ray on the ground now, then go on up to the hit */
        L.lazy = false;
        L.nq = true;
        L.nq_ground = 2;
        L.lstate = LS_EVENT;
      } else {
        if(P.trunc && (L.nsteps >> kCollBits) >= kMaxEmptyPasses) atomicAdd(P.trunc + 1, 1ull);
        L.ev = EV_NONE;
        L.pending = true;
        L.lstate = LS_EVENT;
      }
    }
    WF_TICK(0)

    /* one tentative collision */
    if(COUNT) {
      const uint64_t bc = __ballot(L.lstate == LS_COLL);
      d_ce += bc != 0;
      d_cl += __popcll(bc);
    }
    if(L.lstate == LS_COLL) {
      /* ..._sw.c:123: a scattering candidate beyond 1e9 m is dropped, the
       * query moving on to the next segment with the same optical thickness */
      if(!kLW && L.mode == MODE_SC && L.tc > 1e9f) L.tc = INFINITY;
      /* the nearer of the scattering and the absorption candidates */
      const bool is_a = !kLW && L.tc_a < L.tc;
      const float tx = is_a ? L.tc_a : L.tc;
      if(tx < INFINITY) {
        if(COUNT) L.c_coll++;
        float x[3];
#pragma unroll
        for(int a = 0; a < 3; ++a) x[a] = L.pos[a] + tx * L.dir[a];
        const float v = L.s_cloud ? seg_lwcf<COUNT>(G, L, x) : 0.f;
        /* the draw does not depend on the LWC: written before the LWC's
         * first use, its Threefry rounds run while the load is in flight */
        const float2 uu = draw2(L, key);
        const float gka = L.s_gka, gks = L.s_gks;
        const float ka = gka + L.cka * v;
        const float ks = gks + L.cks * v;
        if(is_a) {
          /* residual tracking of the absorption transmissivity: a real
           * collision makes it 0 and ends the walk (..._sw.c:396-400,
           * 206-209) */
          const float amin = gka + (L.s_cloud ? L.cka * L.lmn : 0.f);
          L.ev = uu.x < (ka - amin) / L.maj_a ? EV_ABS : EV_NONE;
        } else if(kLW) {
          const float pa = ka * L.inv_maj, ps = ks * L.inv_maj;
          L.ev = uu.x < pa ? EV_ABS : uu.x < pa + ps ? EV_SCA : EV_NONE;
        } else {
          const float k = L.prop == HTGX_Ka ? ka : L.prop == HTGX_Ks ? ks : ka + ks;
          const float proba = (L.mode == MODE_TR ? k - L.s_kmin : k) * L.inv_maj;
          L.ev = uu.x < proba ? EV_HIT : EV_NONE;
          if(!kLW && L.mode == MODE_SC && L.ev != EV_NONE) {
            /* gas vs particle scattering at the event, ..._sw.c:427-438 */
            LC(gas_ratio) = gks / ks;
            /* the absorption optical depth stops at the event */
            const float amin = gka + (L.s_cloud ? L.cka * L.lmn : 0.f);
            L.Tmin -= (double)(L.s_t1 - tx) * (double)amin;
          }
        }
        if(L.ev != EV_NONE) {
          LC(ev_dst) = tx;
          LC(ev_layer) = L.s_layer;
          L.pending = true;
          L.lstate = LS_EVENT;
        } else if((++L.nsteps & kCollMask) > G.max_steps) {
          /* tentative-collision budget of a query exhausted: end it without
           * event, and count it */
          if(P.trunc) atomicAdd(P.trunc, 1ull);
          L.pending = true;
          L.lstate = LS_EVENT;
        } else if(is_a) {
          L.Ts_a = exp1f(uu.y);
          const float Ta = (L.s_t1 - L.tc_a) * L.maj_a;
          if(L.Ts_a > Ta) { L.Ts_a -= Ta; L.tc_a = INFINITY; }
          else L.tc_a += L.Ts_a / L.maj_a;
        } else {
          L.Ts = exp1f(uu.y);
          const float T = (L.s_t1 - L.tc) * L.maj;
          if(L.Ts > T) { L.Ts -= T; L.tc = INFINITY; }
          else L.tc += L.Ts * L.inv_maj;
        }
      }
      /* no candidate left in the segment: on to the next one */
      if(L.lstate == LS_COLL && L.tc == INFINITY && L.tc_a == INFINITY) L.lstate = LS_SEG;
    }
    WF_TICK(1)

    /* events, when enough lanes need them or nothing else is left */
    const uint64_t want = __ballot(L.lstate == LS_EVENT);
    const uint64_t other = __ballot(L.lstate == LS_SEG || L.lstate == LS_COLL);
    if(want && (__popcll(want) >= P.ev_thresh || !other)) {
      /* the shortwave event step reads its constants afresh: the many
       * fields it uses then hold scalar registers only while it runs, not
       * across the tracking steps (fewer scalar spills there). The
       * longwave one, which runs twice as often per realisation, keeps the
       * values computed before the loop: recomputing them at each event
       * measured slower there */
      constexpr bool kReload = SPECTRAL != HTGX_SPECTRAL_LW, kReloadP = kReload;
      const SceneF& S_h = S;
      const DevMap& M_h = M;
      const Launch& P_h = P;
      SceneF S = kReload ? kload(launder(pS)) : S_h;
      DevMap M = kReload ? kload(launder(pM)) : M_h;
      const Launch P = kReloadP ? kload(launder(pP)) : P_h;
      if(kReload) fold_features<FEAT>(S, M);
      if(COUNT) { d_ee++; d_el += __popcll(want); }
      if(L.lstate == LS_EVENT) {
        if(LIST && LC(stage) == ST_FETCH) {
          /* the next (pixel, realisation) pair of the list, one realisation */
          const uint32_t c = atomicAdd(P.counters, 1u);
          if(c >= P.nvalid) {
            L.lstate = LS_DONE;
          } else {
            LC(chunk) = c;
            L.pixel = P.list[2 * c];
            L.r = P.list[2 * c + 1];
            LC(r_end) = L.r + 1u;
            LC(n) = LC(nfail) = 0;
            LC(sum) = LC(sum2) = 0.0;
            LC(tsum) = LC(tsum2) = 0.f;
            LC(stage) = ST_BEGIN;
          }
        } else if(LC(stage) == ST_FETCH) {
          uint32_t c;
          if(!fetch_chunk(P, xcd, &c)) {
            L.lstate = LS_DONE;
            if(P.exhausted) atomicExch(P.exhausted, 1u);
          } else {
            LC(chunk) = c;
            const uint32_t m = c & 63u, rest = c >> 6;
            const uint32_t s = rest % P.nsub, t = rest / P.nsub;
            const uint32_t x = (t % P.ntx) * 8u + ((m & 1u) | ((m >> 1) & 2u) | ((m >> 2) & 4u));
            const uint32_t y = (t / P.ntx) * 8u + (((m >> 1) & 1u) | ((m >> 2) & 2u) | ((m >> 3) & 4u));
            if(x < P.width && y < P.height) {
              L.pixel = y * P.width + x;
              L.r = (P.spp_offset_dev ? *P.spp_offset_dev : P.spp_offset) + s * P.K;
              LC(r_end) = L.r + P.K;
              LC(n) = LC(nfail) = 0;
              LC(sum) = LC(sum2) = 0.0;
              LC(tsum) = LC(tsum2) = 0.f;
              LC(stage) = ST_BEGIN;
            } /* else: outside the image, fetch again next time */
          }
        }
      }
      WF_TICK(2)
      if(L.lstate == LS_EVENT && LC(stage) != ST_FETCH) {
        const bool begin = LC(stage) == ST_BEGIN;
        if(begin) {
          L.pair = 0;
          LC(t0) = (uint32_t)wall_clock64();
          LCA(shift, 0) = LCA(shift, 1) = 0.f; /* sensor rays start in world coordinates */
        }
        /* one set of canonicals per event, drawn at a single site */
        float u[2 * NP];
#pragma unroll
        for(int i = 0; i < NP; ++i) { const float2 v = draw2(L, key); u[2 * i] = v.x; u[2 * i + 1] = v.y; }
        WF_TICK(3)
        if(begin) {
          real_begin<SENSOR, SPECTRAL, COUNT, RARE>(S, M, L, u);
        }
        WF_TICK(4)
        if(!begin && L.pending) {
          L.pending = false;
          consume<COUNT, SPECTRAL>(S, L, u);
        }
        WF_TICK(5)
        while(LC(stage) == ST_WALK_DONE) walk_done<SENSOR, SPECTRAL, COUNT>(S, L);
        WF_TICK(6)
        if(LC(stage) == ST_REAL_DONE) {
          const uint64_t t1 = wall_clock64();
          if(LC(w) >= 0.f) {
            const double wd = (double)LC(w);
            const float us = (float)((double)((uint32_t)t1 - LC(t0)) * M.us_per_tick);
            LC(sum) += wd;
            LC(sum2) += wd * wd;
            LC(tsum) += us;
            LC(tsum2) += us * us;
            LC(n)++;
          }
          L.r++;
          if(L.r != LC(r_end)) {
            LC(stage) = ST_BEGIN;
          } else {
            htgx_mc_accum a;
            a.sum = LC(sum); a.sum_c = 0.0; a.sum2 = LC(sum2); a.sum2_c = 0.0;
            a.tsum = LC(tsum); a.tsum2 = LC(tsum2); a.n = LC(n); a.nfail = LC(nfail);
            P.partials[LC(chunk)] = a;
            LC(stage) = ST_FETCH;
            if(P.done_count && atomicAdd(P.done_count, 1u) + 1u == P.nvalid)
              *P.done_tick = wall_clock64();
          }
        }
        WF_TICK(7)
        /* set-up of the requested query (single site) */
        if(L.nq) {
          L.nq = false;
          const bool resume = L.nq_ground == 2;
          L.lazy = false; /* a query that ended in the medium leaves no cast pending */
          if(L.nq_ground) {
            float wp[3];
            world_posf(S, L.pos, wp);
            if(S.ground_type == HTGX_GROUND_MESH && !resume && wp[2] > S.ghi[2] && L.dir[2] < 0.f) {
              /* from above the mesh, downwards: the tracking runs up to the
               * mesh's top first, and most queries end in the medium before
               * it; the cast waits until the tracking gets there */
              L.lazy = true;
              L.gh_hit = false;
              L.qr = (S.ghi[2] - wp[2]) / L.dir[2];
            } else {
              const GroundF gh = ground_tracef(S, wp, L.dir, INFINITY, LC(prev_prim), L.prev_edge,
                                               false);
              count_ground<COUNT>(L, gh);
              L.gh_hit = gh.hit;
              LC(gh_mtl) = gh.mtl;
              LC(gh_prim) = gh.prim;
              L.gh_edge = gh.edge;
              /* resuming: the tracking goes on from the mesh's top, the
               * iterator's leaf and layer being those at that point */
              if(resume) L.tit = L.qr;
              L.qr = gh.t;
            }
          }
          if(resume) {
            L.itdone = false;
          } else {
            L.Ts = exp1f(u[2 * NP - 1]);
            L.Tmin = 0.0;
            L.pending = false;
            if(!kLW && L.mode == MODE_SC) L.Ts_a = exp1f(draw2(L, key).x);
            iter_init(S, L);
          }
          L.tc = L.tc_a = INFINITY;
          L.lstate = LS_SEG;
        }
        WF_TICK(8)
      }
    }
  }
#undef WF_TICK

  if(COUNT) {
    const uint64_t a = wave_sum64(L.c_leaf), b = wave_sum64(L.c_seg), c = wave_sum64(L.c_coll),
                   d = wave_sum64(L.c_lwc), e = wave_sum64(L.c_ev), f = wave_sum64(L.c_bvh),
                   g = wave_sum64(L.c_tri);
    if((threadIdx.x & 63u) == 0) {
      atomicAdd(P.stats + 8, (unsigned long long)d_it);
      atomicAdd(P.stats + 9, (unsigned long long)d_ce);
      atomicAdd(P.stats + 10, (unsigned long long)d_cl);
      atomicAdd(P.stats + 11, (unsigned long long)d_se);
      atomicAdd(P.stats + 12, (unsigned long long)d_sl);
      atomicAdd(P.stats + 13, (unsigned long long)d_ee);
      atomicAdd(P.stats + 14, (unsigned long long)d_el);
#pragma unroll
      for(int i = 0; i < 9; ++i) atomicAdd(P.stats + 15 + i, (unsigned long long)cy[i]);
      atomicAdd(P.stats + 1, (unsigned long long)a);
      atomicAdd(P.stats + 2, (unsigned long long)b);
      atomicAdd(P.stats + 3, (unsigned long long)c);
      atomicAdd(P.stats + 4, (unsigned long long)d);
      atomicAdd(P.stats + 5, (unsigned long long)e);
      atomicAdd(P.stats + 6, (unsigned long long)f);
      atomicAdd(P.stats + 7, (unsigned long long)g);
    }
  }
}

/* 3 waves per SIMD: the register cap (168) spills a little but hides more
 * latency than 2 waves without spills (measured 14% faster at 1024^2) */
template<int SENSOR, int SPECTRAL, bool COUNT, int FEAT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
wavefront_kernel(KArgs A)
{
  (void)A;
  const KPtr<KArgs> k = (KPtr<KArgs>)__builtin_amdgcn_kernarg_segment_ptr();
  run_scene<SENSOR, SPECTRAL, COUNT, FEAT>(&k->S, &k->M, &k->P);
}

/* Given (pixel, realisation) pairs of one map, one realisation each (the
 * debug entry of single FP32 realisations): the loop of the production
 * kernels with the rare and mesh features compiled in, which give the same
 * arithmetic as the leaner instances */
template<int SENSOR, int SPECTRAL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
wavefront_list_kernel(KArgs A)
{
  (void)A;
  const KPtr<KArgs> k = (KPtr<KArgs>)__builtin_amdgcn_kernarg_segment_ptr();
  run_scene<SENSOR, SPECTRAL, false, kFeatRare | kFeatMesh, true>(&k->S, &k->M, &k->P);
}

/* The FP32 segment iterator on given rays over [0, range[1]] (a debug
 * entry): the segments and their property bounds as the production kernels
 * see them, with the leaf id read back from the leaf record */
__global__ void __launch_bounds__(256)
wf_trace_segments_kernel(SceneF S, const double* rays, uint32_t nrays, int prop, uint32_t iband,
                         uint32_t iquad, htgx_segment* segs, uint32_t max_segs, uint32_t* nsegs)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= nrays) return;
  const double* r = rays + (size_t)i * 8;
  Lane L;
#pragma unroll
  for(int a = 0; a < 3; ++a) { L.pos[a] = (float)r[a]; L.dir[a] = (float)r[3 + a]; }
  L.qr = (float)r[7];
  L.gka_off = (S.quad_off[iband] + iquad) * S.nlayers;
  L.gks_off = iband * S.nlayers;
  L.cka = S.has_clouds ? S.cka[iband] : 0.f;
  L.cks = S.has_clouds ? S.cks[iband] : 0.f;
  LCA(shift, 0) = LCA(shift, 1) = 0.f;
  iter_init(S, L);
  uint32_t n = 0;
  float t0;
  const TrackRegs G = track_regs<false>(S, 0u, 0u);
  while(iter_next<false>(S, G, L, prop, &t0)) {
    if(n < max_segs) {
      htgx_segment& o = segs[(size_t)i * max_segs + n];
      o.t0 = t0;
      o.t1 = L.s_t1;
      o.kmin = L.s_kmin;
      o.kmax = L.s_kmax;
      o.leaf = -1;
      if(L.s_cloud) {
        const uint32_t c[3] = {L.lohi[0] & 0xFFFFu, L.lohi[1] & 0xFFFFu, L.lohi[2] & 0xFFFFu};
        o.leaf = S.leafrec[(c[2] * S.dims0[1] + c[1]) * S.dims0[0] + c[0]].x & kLeafIdMask;
      }
      o.layer = L.s_layer;
      o.pad = 0;
    }
    n++;
  }
  nsegs[i] = n;
}

/* One map of a batched launch */
struct BatchScene {
  SceneF S;
  DevMap M;
  Launch P;
  htgx_mc_accum* accum; /* the map's pixel accumulators */
};

/* Several maps of distinct scenes in one persistent launch. Every wavefront
 * walks the scenes in table order (largest first): it drains a scene's chunk
 * queues with the per-scene loop, then moves on, so the tails of the scenes
 * overlap instead of idling the device between launches. The scene is
 * uniform over a wavefront, so its constants stay scalar. */
template<int SENSOR, int SPECTRAL, int FEAT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
wavefront_batch_kernel(const BatchScene* __restrict__ scenes, uint32_t nscenes)
{
  for(uint32_t s = 0; s < nscenes; ++s) {
    const BatchScene& B = scenes[s];
    const unsigned int ex = __hip_atomic_load(B.P.exhausted, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
    if(__builtin_amdgcn_readfirstlane(ex)) continue;
    run_scene<SENSOR, SPECTRAL, false, FEAT>((KPtr<SceneF>)&B.S, (KPtr<DevMap>)&B.M, (KPtr<Launch>)&B.P);
  }
}

/* Start of a batched step: ticks[0] = now, and ticks[1] = now on the first
 * step (*step == 0) */
__global__ void
batch_tick_kernel(unsigned long long* ticks, const unsigned int* step)
{
  const unsigned long long now = wall_clock64();
  ticks[0] = now;
  if(*step == 0u) ticks[1] = now;
}

/* End of a batched step: the next step draws the next spp realisations of
 * every pixel */
__global__ void
batch_advance_kernel(BatchScene* scenes, uint32_t n, unsigned int* step)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) {
    scenes[i].P.spp_offset += scenes[i].M.spp;
    scenes[i].M.spp_offset += scenes[i].M.spp;
  }
  if(i == 0) *step += 1u;
}

/* Fold the chunks of pixel p, in chunk order, into its accumulator */
__device__ __forceinline__ void
merge_pixel(const htgx_mc_accum* partials, uint32_t width, uint32_t ntx, uint32_t nsub,
            htgx_mc_accum* accum, uint32_t p)
{
  const uint32_t x = p % width, y = p / width;
  const uint32_t t = (y / 8u) * ntx + (x / 8u);
  const uint32_t ix = x & 7u, iy = y & 7u;
  const uint32_t m = (ix & 1u) | ((iy & 1u) << 1) | ((ix & 2u) << 1) | ((iy & 2u) << 2)
                   | ((ix & 4u) << 2) | ((iy & 4u) << 3);
  htgx_mc_accum a = accum[p];
  for(uint32_t s = 0; s < nsub; ++s) {
    const htgx_mc_accum q = partials[((size_t)t * nsub + s) * 64u + m];
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

__global__ void __launch_bounds__(256)
wf_merge_kernel(const htgx_mc_accum* partials, uint32_t width, uint32_t height, uint32_t ntx,
                uint32_t nsub, htgx_mc_accum* accum)
{
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if(p < width * height) merge_pixel(partials, width, ntx, nsub, accum, p);
}

/* The merges of every map of a batch: blockIdx.y is the table slot */
__global__ void __launch_bounds__(256)
batch_merge_kernel(const BatchScene* scenes)
{
  const BatchScene& B = scenes[blockIdx.y];
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if(p < B.M.width * B.M.height)
    merge_pixel(B.P.partials, B.M.width, B.P.ntx, B.P.nsub, B.accum, p);
}

__global__ void
zero_u32_kernel(unsigned int* p, uint32_t n)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = 0u;
}

__global__ void
zero_u64_kernel(unsigned long long* p, uint32_t n)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = 0ull;
}

/* FP16 bits of a bound of v >= 0 in units of scale >= v, rounded down (lower
 * bound) or up (upper bound) so that the decoded value, FP32
 * half_bits_to_float(b) * scale as load_leaf computes it, never lies inside
 * the true bounds */
__device__ uint32_t bound_half_bits(float v, float scale, bool up)
{
  uint32_t b = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)(v / scale));
  for(int i = 0; i < 4; ++i) {
    const float d = half_bits_to_float(b) * scale;
    if(up ? d >= v : d <= v) return b;
    b = up ? b + 1u : b - 1u;
  }
  return up ? 0x3C00u /* 1.0: scale >= v */ : 0u;
}

__global__ void
leafrec_kernel(const uint32_t* __restrict__ lmap, const float2* __restrict__ bounds, uint32_t n,
               float scale, uint2* __restrict__ out)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const uint32_t word = lmap[i];
  const float2 b = bounds[word & kLeafIdMask];
  out[i] = make_uint2(word, bound_half_bits(b.x, scale, false)
                          | (bound_half_bits(b.y, scale, true) << 16));
}

template<class F>
int launch_variant(F kernel, const SceneF& S, const DevMap& M, const Launch& P, uint32_t grid,
                   hipStream_t stream)
{
  kernel<<<grid, 256, 0, stream>>>(KArgs{S, M, P});
  HIP_OK(hipGetLastError());
  return HTGX_OK;
}

/* The kernel features a map needs (kFeatRare, kFeatMesh) */
int features(const SceneF& S, const DevMap& M)
{
  const bool rare = S.repeat || S.ground_repeat || S.gmat || M.ortho || M.lens_radius > 0 ||
                    M.wmode == 3 || M.wmode == 2;
  return (rare ? kFeatRare : 0) | (S.ground_type == HTGX_GROUND_MESH ? kFeatMesh : 0);
}

template<bool COUNT, int FEAT>
int launch_full(const SceneF& S, const DevMap& M, const Launch& P, uint32_t grid, hipStream_t stream)
{
  if(M.sensor == HTGX_SENSOR_CAMERA) {
    if(M.spectral_type == HTGX_SPECTRAL_LW)
      return launch_variant(wavefront_kernel<HTGX_SENSOR_CAMERA, HTGX_SPECTRAL_LW, COUNT, FEAT>, S, M, P, grid, stream);
    return launch_variant(wavefront_kernel<HTGX_SENSOR_CAMERA, HTGX_SPECTRAL_SW, COUNT, FEAT>, S, M, P, grid, stream);
  }
  if(M.spectral_type == HTGX_SPECTRAL_LW)
    return launch_variant(wavefront_kernel<HTGX_SENSOR_RECTANGLE, HTGX_SPECTRAL_LW, COUNT, FEAT>, S, M, P, grid, stream);
  return launch_variant(wavefront_kernel<HTGX_SENSOR_RECTANGLE, HTGX_SPECTRAL_SW, COUNT, FEAT>, S, M, P, grid, stream);
}

template<bool COUNT>
int launch(const SceneF& S, const DevMap& M, const Launch& P, uint32_t grid, hipStream_t stream)
{
  switch(features(S, M)) {
    case 0: return launch_full<COUNT, 0>(S, M, P, grid, stream);
    case kFeatRare: return launch_full<COUNT, kFeatRare>(S, M, P, grid, stream);
    case kFeatMesh: return launch_full<COUNT, kFeatMesh>(S, M, P, grid, stream);
    default: return launch_full<COUNT, kFeatRare | kFeatMesh>(S, M, P, grid, stream);
  }
}

} /* namespace */

/*******************************************************************************
 * Scratch of htgx_render_accum_noalloc
 ******************************************************************************/
/* The chunk queue counters (the first kScratchHead bytes) and the chunk
 * partial sums of one FP32 map launched by htgx_render_accum_noalloc, in
 * device memory reserved with the code object on each GPU the library runs
 * on and held until the process ends, instead of memory the call allocates.
 * Every call on a GPU shares that GPU's copy: calls that overlap on one GPU
 * race. */
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr size_t kScratchHead = 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

/* Device address of each GPU's copy, looked up once per device (at the FP32
 * set-up of its first context, with that device current). The array's name
 * in host code is a stand-in, not a device address: only these pointers
 * reach kernels. */
constexpr int kMaxScratchDevices = 64;
static void* g_scratch[kMaxScratchDevices];
static std::mutex g_scratch_lock;

static int scratch_lookup()
{
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if(dev < 0 || dev >= kMaxScratchDevices) return HTGX_BAD_OP;
  std::lock_guard<std::mutex> guard(g_scratch_lock);
  if(g_scratch[dev]) return HTGX_OK;
  void* p = nullptr;
  HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch[dev] = p;
  return HTGX_OK;
}

static void* scratch_of(int dev)
{
  if(dev < 0 || dev >= kMaxScratchDevices) return nullptr;
  std::lock_guard<std::mutex> guard(g_scratch_lock);
  return g_scratch[dev];
}

/*******************************************************************************
 * Host side
 ******************************************************************************/
struct htgx_wavefront_state {
  SceneF S;
  std::vector<void*> bufs;
  unsigned int* d_counters = nullptr;
  unsigned long long* d_trunc = nullptr; /* [2] queries cut short, Launch::trunc */
  void* d_partials = nullptr;
  size_t partials_cap = 0;
  uint32_t grid = 0;
  uint32_t grid_max = 0; /* resident blocks of the device */
  uint32_t K = 4;
  uint32_t ev_thresh = 24;
  uint32_t seg_thresh = 0;
};

/* Device array freed with its owner */
template<class T>
struct DevArrayLite {
  T* p = nullptr;
  DevArrayLite() = default;
  DevArrayLite(const DevArrayLite&) = delete;
  DevArrayLite(DevArrayLite&& o) noexcept : p(o.p) { o.p = nullptr; }
  ~DevArrayLite() { if(p) (void)hipFree(p); }
  int alloc(size_t n)
  {
    HIP_OK(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)));
    return HTGX_OK;
  }
};

template<class T>
static int upload_f(htgx_wavefront_state* st, const std::vector<T>& v, const T** out)
{
  void* d = nullptr;
  if(v.empty()) { *out = nullptr; return HTGX_OK; }
  HIP_OK(hipMalloc(&d, v.size() * sizeof(T)));
  st->bufs.push_back(d);
  HIP_OK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = (const T*)d;
  return HTGX_OK;
}

/* FP32 copy of a device FP64 table, uploaded as a new device table */
template<class T>
static int upload_as_float(htgx_wavefront_state* st, const T* dptr, size_t n, const float** out)
{
  std::vector<T> h(n);
  std::vector<float> f(n);
  if(n) HIP_OK(hipMemcpy(h.data(), dptr, n * sizeof(T), hipMemcpyDeviceToHost));
  for(size_t i = 0; i < n; ++i) f[i] = (float)h[i];
  return upload_f(st, f, out);
}

static void
wavefront_free(htgx_wavefront_state* st)
{
  for(void* p : st->bufs) (void)hipFree(p);
  if(st->d_counters) (void)hipFree(st->d_counters);
  if(st->d_trunc) (void)hipFree(st->d_trunc);
  if(st->d_partials) (void)hipFree(st->d_partials);
  delete st;
}

static int wavefront_build(htgx_ctx* ctx, htgx_wavefront_state* st);

static int
wavefront_prepare(htgx_ctx* ctx)
{
  if(ctx->wf) return HTGX_OK;
  htgx_wavefront_state* st = new(std::nothrow) htgx_wavefront_state();
  if(!st) return HTGX_MEM_ERR;
  const int res = wavefront_build(ctx, st);
  if(res != HTGX_OK) {
    wavefront_free(st);
    return res;
  }
  ctx->wf = st; /* published only once complete */
  return HTGX_OK;
}

static int
wavefront_build(htgx_ctx* ctx, htgx_wavefront_state* st)
{
  const DevScene& D = ctx->S;
  SceneF& S = st->S;
  std::memset(&S, 0, sizeof(S));
  int res;
  S.nbands = D.nbands;
  S.band_hi = D.band_hi;
  S.quad_off = D.quad_off;
  S.quad_cdf = D.quad_cdf;
  S.nlayers = D.nlayers;
  S.zb = (float)D.zb;
  S.dzl = (float)D.dzl;
  S.inv_dzl = (float)D.inv_dzl;
  std::vector<uint32_t> qoff(D.nbands + 1);
  HIP_OK(hipMemcpy(qoff.data(), D.quad_off, qoff.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  const uint32_t nq = qoff[D.nbands];
  if((res = upload_as_float(st, D.gas_ka, (size_t)nq * D.nlayers, &S.gas_ka))) return res;
  if((res = upload_as_float(st, D.gas_ks, (size_t)D.nbands * D.nlayers, &S.gas_ks))) return res;
  {
    /* cumulative vertical optical depths of the FP32 coefficients, for the
     * gas runs of transmissivity queries */
    const uint32_t nl = D.nlayers;
    std::vector<double> ka((size_t)nq * nl), ks((size_t)D.nbands * nl);
    if(nl) {
      HIP_OK(hipMemcpy(ka.data(), D.gas_ka, ka.size() * sizeof(double), hipMemcpyDeviceToHost));
      HIP_OK(hipMemcpy(ks.data(), D.gas_ks, ks.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    const double dz = (double)S.dzl;
    auto cumulate = [&](std::vector<double>& v, size_t nrows) {
      for(size_t r = 0; r < nrows; ++r) {
        double acc = 0.0;
        for(uint32_t j = 0; j < nl; ++j) {
          const double k = (double)(float)v[r * nl + j];
          v[r * nl + j] = acc;
          acc += k * dz;
        }
      }
    };
    cumulate(ka, nq);
    cumulate(ks, D.nbands);
    if((res = upload_f(st, ka, &S.cum_ka))) return res;
    if((res = upload_f(st, ks, &S.cum_ks))) return res;
  }
  if(D.gas_T && (res = upload_as_float(st, D.gas_T, D.nlayers, &S.gas_T))) return res;
  S.has_clouds = D.has_clouds;
  if(D.has_clouds) {
    S.B = D.B;
    for(int a = 0; a < 3; ++a) {
      S.def[a] = D.def[a];
      S.dims0[a] = D.dims0[a];
      S.lower[a] = (float)D.lower[a];
      S.bs[a] = (float)D.bs[a];
      S.inv_bs[a] = (float)D.inv_bs[a];
      S.inv_cs[a] = (float)D.inv_cs[a];
      if(D.dims0[a] > 0xFFFFu) return HTGX_BAD_ARG;
    }
    S.repeat = D.repeat;
    S.period[0] = (float)D.period[0];
    S.period[1] = (float)D.period[1];
    S.lwc = D.lwc;
    /* fused leaf records; the scale is the grid's largest LWC (the root's
     * upper bound, last node of the level-ordered bounds) */
    float2 root;
    HIP_OK(hipMemcpy(&root, D.bounds + (ctx->nnodes - 1), sizeof(float2), hipMemcpyDeviceToHost));
    S.lwc_scale = root.y > 0.f ? root.y : 1.f;
    const uint32_t nblocks = D.dims0[0] * D.dims0[1] * D.dims0[2];
    uint2* rec = nullptr;
    HIP_OK(hipMalloc(&rec, (size_t)nblocks * sizeof(uint2)));
    st->bufs.push_back(rec);
    leafrec_kernel<<<(nblocks + 255) / 256, 256>>>(D.lmap, D.bounds, nblocks, S.lwc_scale, rec);
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    S.leafrec = rec;
    std::vector<double> ce(D.nbands), ssa(D.nbands), g(D.nbands);
    HIP_OK(hipMemcpy(ce.data(), D.cext, D.nbands * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(ssa.data(), D.ssa, D.nbands * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(g.data(), D.hg_g, D.nbands * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<float> ca(D.nbands), cs(D.nbands), gf(D.nbands);
    for(uint32_t b = 0; b < D.nbands; ++b) {
      ca[b] = (float)(ce[b] * (1.0 - ssa[b]));
      cs[b] = (float)(ce[b] * ssa[b]);
      gf[b] = (float)g[b];
    }
    if((res = upload_f(st, ca, &S.cka))) return res;
    if((res = upload_f(st, cs, &S.cks))) return res;
    if((res = upload_f(st, gf, &S.hg_g))) return res;
  }
  S.ground_type = D.ground_type;
  S.plane_z = (float)D.plane_z;
  S.plane_mtl = D.plane_mtl;
  S.bvh = D.bvh;
  S.nbvh = D.nbvh;
  S.tris = D.tris;
  S.tri_mtl = D.tri_mtl;
  S.ground_repeat = D.ground_repeat;
  for(int a = 0; a < 3; ++a) { S.glo[a] = (float)D.glo[a]; S.ghi[a] = (float)D.ghi[a]; }
  if(D.ground_type == HTGX_GROUND_PLANE || D.ground_type == HTGX_GROUND_MESH) {
    const uint32_t nm = ctx->nmtls;
    if(D.mtl_albedo && (res = upload_as_float(st, D.mtl_albedo, nm, &S.mtl_albedo))) return res;
    if(D.mtl_T && (res = upload_as_float(st, D.mtl_T, nm, &S.mtl_T))) return res;
  }
  S.gmat = D.gmat;
  S.explicit_ifc = D.explicit_ifc;
  S.sky_mtl = D.sky_mtl;
  S.mtls = D.mtls;
  S.brdfs = D.brdfs;
  S.ifcs = D.ifcs;
  for(int a = 0; a < 3; ++a) S.sun_dir[a] = (float)D.sun_dir[a];
  S.sun_cos_half = (float)D.sun_cos_half;
  S.sun_solid_angle = (float)D.sun_solid_angle;

  HIP_OK(hipMalloc(&st->d_counters, kXcds * sizeof(unsigned int)));
  HIP_OK(hipMalloc(&st->d_trunc, 2 * sizeof(unsigned long long)));
  if((res = scratch_lookup()) != HTGX_OK) return res;
  HIP_OK(hipMemset(st->d_trunc, 0, 2 * sizeof(unsigned long long)));
  /* persistent grid: as many resident blocks as the device holds */
  int dev = 0, ncu = 0, per_cu = 0;
  HIP_OK(hipGetDevice(&dev));
  HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu,
    wavefront_kernel<HTGX_SENSOR_CAMERA, HTGX_SPECTRAL_SW, false, 0>, 256, 0));
  st->grid = st->grid_max = (uint32_t)std::max(1, ncu * std::max(1, per_cu));
  if(const char* e = std::getenv("HTGX_WF_GRID")) st->grid = (uint32_t)std::max(1, atoi(e));
  if(const char* e = std::getenv("HTGX_WF_EVTHRESH")) st->ev_thresh = (uint32_t)std::max(1, atoi(e));
  if(const char* e = std::getenv("HTGX_WF_CHUNK")) st->K = (uint32_t)std::min(16, std::max(1, atoi(e)));
  if(const char* e = std::getenv("HTGX_WF_SEGTHRESH")) st->seg_thresh = (uint32_t)std::min(64, std::max(0, atoi(e)));
  return HTGX_OK;
}

/* Launch parameters of one map, its partials and counters aside */
static int
fill_launch(htgx_ctx* ctx, const DevMap& M, Launch* P)
{
  htgx_wavefront_state* st = ctx->wf;
  std::memset(P, 0, sizeof(*P));
  P->width = M.width;
  P->height = M.height;
  P->ntx = (M.width + 7) / 8;
  P->ntiles = P->ntx * ((M.height + 7) / 8);
  /* chunk size: the largest divisor of spp not above K */
  uint32_t K = std::min(st->K, M.spp);
  while(M.spp % K) --K;
  P->K = K;
  P->nsub = M.spp / K;
  P->xcd_tiles = (P->ntiles + kXcds - 1) / kXcds;
  P->spp_offset = M.spp_offset;
  P->ev_thresh = st->ev_thresh;
  P->seg_thresh = st->seg_thresh;
  P->max_steps = 1u << 20;
  const uint64_t nchunks = (uint64_t)P->ntiles * P->nsub * 64u;
  if(nchunks > 0xFFFFFFFFull) return HTGX_BAD_ARG;
  P->nvalid = (uint32_t)((uint64_t)M.width * M.height * P->nsub);
  P->stats = ctx->d_counters;
  P->trunc = st->d_trunc;
  return HTGX_OK;
}

/* Launch parameters of one map; sizes the chunk partials of the context */
static int
setup_launch(htgx_ctx* ctx, const DevMap& M, Launch* P)
{
  htgx_wavefront_state* st = ctx->wf;
  int res;
  if((res = fill_launch(ctx, M, P)) != HTGX_OK) return res;
  const size_t need = (size_t)P->ntiles * P->nsub * 64u * sizeof(htgx_mc_accum);
  if(need > st->partials_cap) {
    if(st->d_partials) { HIP_OK(hipFree(st->d_partials)); st->d_partials = nullptr; }
    HIP_OK(hipMalloc(&st->d_partials, need));
    st->partials_cap = need;
  }
  P->counters = st->d_counters;
  P->partials = (htgx_mc_accum*)st->d_partials;
  return HTGX_OK;
}

int htgx_wavefront_truncations(htgx_ctx* ctx, uint64_t out[2])
{
  out[0] = out[1] = 0;
  if(!ctx->wf) return HTGX_OK;
  unsigned long long h[2];
  HIP_OK(hipMemcpy(h, ctx->wf->d_trunc, sizeof(h), hipMemcpyDeviceToHost));
  out[0] = h[0];
  out[1] = h[1];
  return HTGX_OK;
}

int htgx_render_wavefront(htgx_ctx* ctx, const DevMap& M, htgx_mc_accum* accum, hipStream_t stream)
{
  int res;
  if((res = wavefront_prepare(ctx)) != HTGX_OK) return res;
  htgx_wavefront_state* st = ctx->wf;
  Launch P;
  if((res = setup_launch(ctx, M, &P)) != HTGX_OK) return res;
  zero_u32_kernel<<<1, 64, 0, stream>>>(st->d_counters, kXcds);
  HIP_OK(hipGetLastError());
  if(ctx->counting) {
    zero_u64_kernel<<<1, 64, 0, stream>>>(ctx->d_counters, HTGX_NSTATS);
    HIP_OK(hipGetLastError());
    if((res = launch<true>(st->S, M, P, st->grid, stream))) return res;
  } else {
    if((res = launch<false>(st->S, M, P, st->grid, stream))) return res;
  }
  const uint32_t npix = M.width * M.height;
  wf_merge_kernel<<<(npix + 255) / 256, 256, 0, stream>>>(
    (const htgx_mc_accum*)st->d_partials, M.width, M.height, P.ntx, P.nsub, accum);
  HIP_OK(hipGetLastError());
  ctx->last_nrealisations = (uint64_t)npix * M.spp;
  return HTGX_OK;
}

int htgx_render_wavefront_noalloc(htgx_ctx* ctx, const DevMap& M, htgx_mc_accum* accum,
                                  hipStream_t stream)
{
  htgx_wavefront_state* st = ctx->wf;
  if(!st) return HTGX_BAD_OP; /* the FP32 tables would have to be built */
  int dev = 0, res;
  HIP_OK(hipGetDevice(&dev));
  if(dev != ctx->device) return HTGX_BAD_ARG;
  unsigned char* base = (unsigned char*)scratch_of(dev);
  if(!base) return HTGX_BAD_OP;
  Launch P;
  if((res = fill_launch(ctx, M, &P)) != HTGX_OK) return res;
  const uint64_t need = (uint64_t)P.ntiles * P.nsub * 64u * sizeof(htgx_mc_accum);
  if(need > kScratchBytes - kScratchHead) {
    fprintf(stderr, "htgx: htgx_render_accum_noalloc: %llu chunk partials of %zu B exceed the "
            "%zu MiB scratch (fewer pixels, fewer realisations per call, or a larger chunk)\n",
            (unsigned long long)(need / sizeof(htgx_mc_accum)), sizeof(htgx_mc_accum),
            kScratchBytes >> 20);
    return HTGX_BAD_ARG;
  }
  P.counters = (unsigned int*)base;
  P.partials = (htgx_mc_accum*)(base + kScratchHead);
  /* fill kernels, no memset: the call may be captured in a graph */
  zero_u32_kernel<<<1, 64, 0, stream>>>(P.counters, kXcds);
  HIP_OK(hipGetLastError());
  if(ctx->counting) {
    zero_u64_kernel<<<1, 64, 0, stream>>>(ctx->d_counters, HTGX_NSTATS);
    HIP_OK(hipGetLastError());
    if((res = launch<true>(st->S, M, P, st->grid, stream))) return res;
  } else {
    if((res = launch<false>(st->S, M, P, st->grid, stream))) return res;
  }
  const uint32_t npix = M.width * M.height;
  wf_merge_kernel<<<(npix + 255) / 256, 256, 0, stream>>>(P.partials, M.width, M.height, P.ntx,
                                                           P.nsub, accum);
  HIP_OK(hipGetLastError());
  ctx->last_nrealisations = (uint64_t)npix * M.spp;
  return HTGX_OK;
}

int htgx_wavefront_params(htgx_ctx* ctx, htgx_wf_params* get, const htgx_wf_params* set)
{
  int res;
  if((res = wavefront_prepare(ctx)) != HTGX_OK) return res;
  htgx_wavefront_state* st = ctx->wf;
  if(set) {
    if(set->ev_thresh > 64 || set->chunk > 16 || set->grid > st->grid_max || set->seg_thresh > 64)
      return HTGX_BAD_ARG;
    st->ev_thresh = set->ev_thresh ? set->ev_thresh : 24;
    st->seg_thresh = set->seg_thresh;
    st->K = set->chunk ? set->chunk : 4;
    st->grid = set->grid ? set->grid : st->grid_max;
  }
  if(get) {
    get->ev_thresh = st->ev_thresh;
    get->chunk = st->K;
    get->grid = st->grid;
    get->grid_max = st->grid_max;
    get->seg_thresh = st->seg_thresh;
  }
  return HTGX_OK;
}

/* A batch of maps prepared for repeated steps (htgx_batch_plan_*): its own
 * device table, chunk queues, flags and clock ticks, and optionally its own
 * chunk partials and a HIP graph of one step. */
struct htgx_batch_plan {
  int device = 0;
  uint32_t n = 0;
  /* one map: its own single-map kernel instead of the batch kernel (the
   * same realisations, without the scene table's scalar reloads) */
  bool single = false;
  int sensor = 0, spectral = 0;
  int feat = 0;       /* kernel features the maps need, together */
  uint32_t grid = 0;
  uint32_t maxpix = 0;
  std::vector<uint32_t> order;      /* table slot -> item */
  std::vector<BatchScene> table_h;  /* the table at step 0 */
  std::vector<double> us_per_tick;  /* per table slot */
  std::vector<void*> bufs;          /* owned device memory */
  BatchScene* table = nullptr;
  unsigned long long* ticks = nullptr; /* [n] done ticks, step start, first start */
  unsigned int* flags = nullptr;       /* [n] exhausted, [n] done counts, step counter */
  unsigned int* queues = nullptr;      /* [n][kXcds] chunk counters */
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

void wf_plan_free(htgx_batch_plan* p)
{
  if(!p) return;
  if(p->exec) (void)hipGraphExecDestroy(p->exec);
  if(p->graph) (void)hipGraphDestroy(p->graph);
  if(p->cap) (void)hipStreamDestroy(p->cap);
  for(void* b : p->bufs) (void)hipFree(b);
  delete p;
}

template<class T>
static int plan_alloc(htgx_batch_plan* p, size_t n, T** out)
{
  void* d = nullptr;
  HIP_OK(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
  p->bufs.push_back(d);
  *out = (T*)d;
  return HTGX_OK;
}

template<int FEAT>
static void launch_batch(const htgx_batch_plan* p, hipStream_t stream)
{
  const uint32_t n = p->n;
  if(p->sensor == HTGX_SENSOR_CAMERA) {
    if(p->spectral == HTGX_SPECTRAL_LW)
      wavefront_batch_kernel<HTGX_SENSOR_CAMERA, HTGX_SPECTRAL_LW, FEAT><<<p->grid, 256, 0, stream>>>(p->table, n);
    else
      wavefront_batch_kernel<HTGX_SENSOR_CAMERA, HTGX_SPECTRAL_SW, FEAT><<<p->grid, 256, 0, stream>>>(p->table, n);
  } else {
    if(p->spectral == HTGX_SPECTRAL_LW)
      wavefront_batch_kernel<HTGX_SENSOR_RECTANGLE, HTGX_SPECTRAL_LW, FEAT><<<p->grid, 256, 0, stream>>>(p->table, n);
    else
      wavefront_batch_kernel<HTGX_SENSOR_RECTANGLE, HTGX_SPECTRAL_SW, FEAT><<<p->grid, 256, 0, stream>>>(p->table, n);
  }
}

/* Launches of one step on `stream': no allocation, no host copy, no memset
 * (so that a HIP graph can record them) */
static int plan_launch_step(htgx_batch_plan* p, hipStream_t stream)
{
  const uint32_t n = p->n;
  const uint32_t nq = n * kXcds;
  zero_u32_kernel<<<(nq + 255) / 256, 256, 0, stream>>>(p->queues, nq);
  zero_u32_kernel<<<(2 * n + 255) / 256, 256, 0, stream>>>(p->flags, 2 * n);
  batch_tick_kernel<<<1, 1, 0, stream>>>(p->ticks + n, p->flags + 2 * n);
  HIP_OK(hipGetLastError());
  if(p->single) {
    const BatchScene& e = p->table_h[0];
    int res;
    if((res = launch<false>(e.S, e.M, e.P, p->grid, stream)) != HTGX_OK) return res;
  } else switch(p->feat) {
    case 0: launch_batch<0>(p, stream); break;
    case kFeatRare: launch_batch<kFeatRare>(p, stream); break;
    case kFeatMesh: launch_batch<kFeatMesh>(p, stream); break;
    default: launch_batch<kFeatRare | kFeatMesh>(p, stream); break;
  }
  HIP_OK(hipGetLastError());
  batch_merge_kernel<<<dim3((p->maxpix + 255) / 256, n), 256, 0, stream>>>(p->table);
  batch_advance_kernel<<<(n + 255) / 256, 256, 0, stream>>>(p->table, n, p->flags + 2 * n);
  HIP_OK(hipGetLastError());
  return HTGX_OK;
}

int wf_plan_build(htgx_ctx* const* ctxs, const DevMap* Ms, htgx_mc_accum* const* accums,
                  uint32_t n, bool own_partials, bool use_graph, htgx_batch_plan** out)
{
  int res;
  *out = nullptr;
  htgx_batch_plan* p = new htgx_batch_plan;
  struct Guard {
    htgx_batch_plan*& p;
    ~Guard() { wf_plan_free(p); }
  } guard{p};
  HIP_OK(hipGetDevice(&p->device));
  p->n = n;
  p->sensor = Ms[0].sensor;
  p->spectral = Ms[0].spectral_type;
  p->order.resize(n);
  for(uint32_t i = 0; i < n; ++i) {
    if((res = wavefront_prepare(ctxs[i])) != HTGX_OK) return res;
    p->grid = std::max(p->grid, ctxs[i]->wf->grid);
    p->order[i] = i;
  }
  /* largest maps first: the small ones fill the tail */
  std::stable_sort(p->order.begin(), p->order.end(), [&](uint32_t a, uint32_t b) {
    return (uint64_t)Ms[a].width * Ms[a].height * Ms[a].spp
         > (uint64_t)Ms[b].width * Ms[b].height * Ms[b].spp;
  });
  if((res = plan_alloc(p, n, &p->table))) return res;
  if((res = plan_alloc(p, n + 2, &p->ticks))) return res;
  if((res = plan_alloc(p, 2 * n + 1, &p->flags))) return res;
  if((res = plan_alloc(p, (size_t)n * kXcds, &p->queues))) return res;
  p->table_h.resize(n);
  p->us_per_tick.resize(n);
  for(uint32_t k = 0; k < n; ++k) {
    const uint32_t i = p->order[k];
    BatchScene& e = p->table_h[k];
    e.S = ctxs[i]->wf->S;
    e.M = Ms[i];
    e.accum = accums[i];
    if((res = setup_launch(ctxs[i], Ms[i], &e.P)) != HTGX_OK) return res;
    e.P.stats = nullptr;
    e.P.counters = p->queues + (size_t)k * kXcds;
    e.P.exhausted = p->flags + k;
    e.P.done_count = p->flags + n + k;
    e.P.done_tick = p->ticks + k;
    if(own_partials) {
      /* the context's partials and wavelength tables may be reallocated by
       * its next map: the plan keeps copies of its own */
      const size_t np = (size_t)e.P.ntiles * e.P.nsub * 64u;
      if((res = plan_alloc(p, np, &e.P.partials))) return res;
      if(e.M.nwbands) {
        double *cdf = nullptr, *pdf = nullptr;
        if((res = plan_alloc(p, e.M.nwbands, &cdf)) || (res = plan_alloc(p, e.M.nwbands, &pdf)))
          return res;
        HIP_OK(hipMemcpy(cdf, e.M.wcdf, e.M.nwbands * sizeof(double), hipMemcpyDeviceToDevice));
        HIP_OK(hipMemcpy(pdf, e.M.wpdf, e.M.nwbands * sizeof(double), hipMemcpyDeviceToDevice));
        e.M.wcdf = cdf;
        e.M.wpdf = pdf;
      }
    }
    p->maxpix = std::max(p->maxpix, Ms[i].width * Ms[i].height);
    p->feat |= features(e.S, e.M);
    p->us_per_tick[k] = ctxs[i]->us_per_tick;
  }
  if(n == 1) {
    /* the single-map kernel takes its arguments by value, so a graph replays
     * them unchanged: the step's realisation offset is read from the table,
     * where batch_advance_kernel moves it on after each step */
    p->single = true;
    BatchScene& e = p->table_h[0];
    e.P.spp_offset_dev = (const uint32_t*)((const char*)p->table + offsetof(BatchScene, P) +
                                           offsetof(Launch, spp_offset));
  }
  HIP_OK(hipMemcpy(p->table, p->table_h.data(), n * sizeof(BatchScene), hipMemcpyHostToDevice));
  HIP_OK(hipMemset(p->flags, 0, (2 * n + 1) * sizeof(unsigned int)));
  HIP_OK(hipMemset(p->ticks, 0, (n + 2) * sizeof(unsigned long long)));
  if(use_graph) {
    HIP_OK(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking));
    HIP_OK(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
    res = plan_launch_step(p, p->cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(p->cap, &g);
    p->graph = g;
    if(res) return res;
    if(e != hipSuccess || !g) return HTGX_UNKNOWN_ERR;
    HIP_OK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  }
  *out = p;
  p = nullptr; /* handed to the caller */
  return HTGX_OK;
}

int wf_plan_enqueue(htgx_batch_plan* p, hipStream_t stream)
{
  int dev = -1;
  HIP_OK(hipGetDevice(&dev));
  if(dev != p->device) return HTGX_BAD_ARG;
  if(p->exec) {
    HIP_OK(hipGraphLaunch(p->exec, stream));
    return HTGX_OK;
  }
  return plan_launch_step(p, stream);
}

int wf_plan_rewind(htgx_batch_plan* p)
{
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipMemcpy(p->table, p->table_h.data(), p->n * sizeof(BatchScene), hipMemcpyHostToDevice));
  HIP_OK(hipMemset(p->flags + 2 * p->n, 0, sizeof(unsigned int)));
  return HTGX_OK;
}

int wf_plan_times(htgx_batch_plan* p, double* scene_ms)
{
  const uint32_t n = p->n;
  std::vector<unsigned long long> t(n + 2);
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipMemcpy(t.data(), p->ticks, (n + 2) * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for(uint32_t k = 0; k < n; ++k)
    scene_ms[p->order[k]] = (double)(t[k] - t[n + 1]) * p->us_per_tick[k] * 1e-3;
  return HTGX_OK;
}

int htgx_render_wavefront_batch(htgx_ctx* const* ctxs, const DevMap* Ms, htgx_mc_accum* const* accums,
                                uint32_t n, hipStream_t stream, double* scene_ms)
{
  htgx_batch_plan* p = nullptr;
  int res;
  if((res = wf_plan_build(ctxs, Ms, accums, n, false, false, &p)) != HTGX_OK) return res;
  struct Guard {
    htgx_batch_plan* p;
    ~Guard() { wf_plan_free(p); }
  } guard{p};
  if((res = plan_launch_step(p, stream)) != HTGX_OK) return res;
  for(uint32_t i = 0; i < n; ++i)
    ctxs[i]->last_nrealisations = (uint64_t)Ms[i].width * Ms[i].height * Ms[i].spp;
  /* the plan's buffers are freed on return: wait for the launch */
  HIP_OK(hipStreamSynchronize(stream));
  if(scene_ms) return wf_plan_times(p, scene_ms);
  return HTGX_OK;
}

/*******************************************************************************
 * Autotuner of the launch parameters
 ******************************************************************************/
/* A private stream and two events for timed trials */
struct TrialClock {
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~TrialClock()
  {
    if(stream) (void)hipStreamSynchronize(stream);
    if(e0) (void)hipEventDestroy(e0);
    if(e1) (void)hipEventDestroy(e1);
    if(stream) (void)hipStreamDestroy(stream);
  }
  int init()
  {
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    return HTGX_OK;
  }
  /* device time of the launches enqueued by f() on `stream' */
  template<class F>
  int time(F&& f, double* ms)
  {
    HIP_OK(hipEventRecord(e0, stream));
    const int res = f();
    if(res != HTGX_OK) return res;
    HIP_OK(hipEventRecord(e1, stream));
    HIP_OK(hipEventSynchronize(e1));
    float t = 0;
    HIP_OK(hipEventElapsedTime(&t, e0, e1));
    *ms = t;
    return HTGX_OK;
  }
};

/* Coordinate descent over the chunk size, then the event threshold, then the
 * segment threshold, then the grid. The two thresholds interact: when the
 * segment pass moves, the event threshold is searched again, and then the
 * segment threshold if that moved. trial(p, &ms) runs one timed launch with
 * parameters p; each candidate is timed twice and keeps its faster time
 * (single launches vary by 1-2%), and replaces the current parameters only
 * when more than 1% faster. The first launch is an untimed warm-up; the
 * search stops before a trial that would exceed the budget, estimated from
 * the last one. */
template<class Trial>
static int tune_search(Trial&& trial, double budget_ms, uint32_t grid_max, htgx_wf_params* cur,
                       double* cur_ms, uint32_t* ntrials)
{
  static const uint32_t thresholds[] = {8, 16, 24, 32, 40, 48};
  static const uint32_t seg_thresholds[] = {0, 16, 24};
  static const uint32_t chunks[] = {2, 4, 8};
  const uint32_t grids[] = {grid_max, std::max(1u, grid_max * 3 / 4), std::max(1u, grid_max / 2)};
  double ms = 0, ms2 = 0;
  int res;
  if((res = trial(*cur, &ms)) != HTGX_OK) return res;
  if((res = trial(*cur, &ms)) != HTGX_OK) return res;
  if((res = trial(*cur, &ms2)) != HTGX_OK) return res;
  ms = std::min(ms, ms2);
  double spent = 2 * ms, last = 2 * ms;
  uint32_t n = 1;
  *cur_ms = ms;
  /* passes: chunk, event threshold, segment threshold, the event and
   * segment thresholds again (each only if the pass before it moved), grid */
  static const int order[] = {0, 1, 2, 1, 2, 3};
  bool moved = false, stop = false;
  for(int k = 0; k < 6 && !stop; ++k) {
    const int dim = order[k];
    if((k == 3 || k == 4) && !moved) continue;
    moved = false;
    const uint32_t* vals = dim == 0 ? chunks : dim == 1 ? thresholds : dim == 2 ? seg_thresholds : grids;
    const int nv = dim == 1 ? 6 : 3;
    const htgx_wf_params base = *cur;
    for(int i = 0; i < nv; ++i) {
      htgx_wf_params p = base;
      if(dim == 0) p.chunk = vals[i];
      else if(dim == 1) p.ev_thresh = vals[i];
      else if(dim == 2) p.seg_thresh = vals[i];
      else p.grid = vals[i];
      if(p.ev_thresh == base.ev_thresh && p.chunk == base.chunk && p.grid == base.grid &&
         p.seg_thresh == base.seg_thresh)
        continue;
      if(spent + last > budget_ms) { stop = true; break; }
      if((res = trial(p, &ms)) != HTGX_OK) return res;
      if((res = trial(p, &ms2)) != HTGX_OK) return res;
      spent += ms + ms2;
      last = ms + ms2;
      ms = std::min(ms, ms2);
      ++n;
      if(ms < 0.99 * *cur_ms) { *cur = p; *cur_ms = ms; moved = true; }
    }
  }
  *ntrials = n;
  return HTGX_OK;
}

static void set_params(htgx_wavefront_state* st, const htgx_wf_params& p)
{
  st->ev_thresh = p.ev_thresh;
  st->seg_thresh = p.seg_thresh;
  st->K = p.chunk;
  st->grid = p.grid;
}

int htgx_wavefront_autotune(htgx_ctx* ctx, const DevMap& M, double budget_ms,
                            htgx_wf_params* best, double* best_rate, uint32_t* ntrials)
{
  int res;
  if((res = wavefront_prepare(ctx)) != HTGX_OK) return res;
  htgx_wavefront_state* st = ctx->wf;
  const size_t npix = (size_t)M.width * M.height;
  TrialClock clk;
  DevArrayLite<htgx_mc_accum> acc;
  if((res = clk.init()) || (res = acc.alloc(npix))) return res;
  const int counting = ctx->counting;
  ctx->counting = 0;
  htgx_wf_params cur = {st->ev_thresh, st->K, st->grid, st->grid_max, st->seg_thresh};
  double ms = 0;
  uint32_t n = 0;
  res = tune_search([&](const htgx_wf_params& p, double* t) {
    set_params(st, p);
    return clk.time([&] { return htgx_render_wavefront(ctx, M, acc.p, clk.stream); }, t);
  }, budget_ms, st->grid_max, &cur, &ms, &n);
  ctx->counting = counting;
  set_params(st, cur); /* kept for the maps that follow */
  if(res != HTGX_OK) return res;
  if(best) *best = cur;
  if(best_rate) *best_rate = (double)npix * M.spp / (ms * 1e-3);
  if(ntrials) *ntrials = n;
  return HTGX_OK;
}

int htgx_wavefront_batch_autotune(htgx_ctx* const* ctxs, const DevMap* Ms, uint32_t n,
                                  uint32_t spp_div, double budget_ms, htgx_wf_params* best,
                                  double* best_rate, uint32_t* ntrials)
{
  int res;
  uint32_t grid_max = 0;
  std::vector<DevMap> T(Ms, Ms + n);
  std::vector<DevArrayLite<htgx_mc_accum>> acc(n);
  std::vector<htgx_mc_accum*> accp(n);
  double nreal = 0;
  for(uint32_t i = 0; i < n; ++i) {
    if((res = wavefront_prepare(ctxs[i])) != HTGX_OK) return res;
    grid_max = std::max(grid_max, ctxs[i]->wf->grid_max);
    /* trial maps: 1/spp_div of the realisations per pixel, at least 4 */
    T[i].spp = std::max(4u, T[i].spp / spp_div / 4u * 4u);
    T[i].spp_offset = 0;
    if((res = acc[i].alloc((size_t)T[i].width * T[i].height))) return res;
    accp[i] = acc[i].p;
    nreal += (double)T[i].width * T[i].height * T[i].spp;
  }
  TrialClock clk;
  if((res = clk.init())) return res;
  htgx_wavefront_state* st0 = ctxs[0]->wf;
  htgx_wf_params cur = {st0->ev_thresh, st0->K, st0->grid, grid_max, st0->seg_thresh};
  double ms = 0;
  uint32_t nt = 0;
  res = tune_search([&](const htgx_wf_params& p, double* t) {
    for(uint32_t i = 0; i < n; ++i) set_params(ctxs[i]->wf, p);
    htgx_batch_plan* plan = nullptr;
    int r = wf_plan_build(ctxs, T.data(), accp.data(), n, false, false, &plan);
    if(r != HTGX_OK) return r;
    r = clk.time([&] { return wf_plan_enqueue(plan, clk.stream); }, t);
    wf_plan_free(plan);
    return r;
  }, budget_ms, grid_max, &cur, &ms, &nt);
  for(uint32_t i = 0; i < n; ++i) set_params(ctxs[i]->wf, cur);
  if(res != HTGX_OK) return res;
  if(best) *best = cur;
  if(best_rate) *best_rate = nreal / (ms * 1e-3);
  if(ntrials) *ntrials = nt;
  return HTGX_OK;
}

int htgx_debug_realisations_fp32(htgx_ctx* ctx, const htgx_map_desc* desc, const uint64_t* pix,
                                 const uint64_t* real, uint32_t n, double* weights)
{
  int res;
  if((res = wavefront_prepare(ctx)) != HTGX_OK) return res;
  htgx_wavefront_state* st = ctx->wf;
  DevMap M;
  if((res = htgx_map_setup(ctx, desc, &M)) != HTGX_OK) return res;
  const bool cie = desc->spectral_type == HTGX_SPECTRAL_SW_CIE_XYZ;
  for(uint32_t i = 0; i < n; ++i)
    if((real[i] >> 32) > (cie ? 2u : 0u)) return HTGX_BAD_ARG;
  /* one list launch per CIE channel over the realisations of that channel */
  for(int c = 0; c < (cie ? 3 : 1); ++c) {
    std::vector<uint32_t> list;
    std::vector<uint32_t> idx;
    for(uint32_t i = 0; i < n; ++i)
      if((int)(real[i] >> 32) == c) {
        idx.push_back(i);
        list.push_back((uint32_t)pix[i]);
        list.push_back((uint32_t)(real[i] & 0xFFFFFFFFull));
      }
    if(idx.empty()) continue;
    const uint32_t m = (uint32_t)idx.size();
    const DevMap Mc = cie ? htgx_cie_channel_map(ctx, M, c) : M;
    DevArrayLite<uint32_t> d_list;
    DevArrayLite<htgx_mc_accum> d_acc;
    if((res = d_list.alloc(list.size())) || (res = d_acc.alloc(m))) return res;
    HIP_OK(hipMemcpy(d_list.p, list.data(), list.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    Launch P;
    std::memset(&P, 0, sizeof(P));
    P.width = Mc.width;
    P.height = Mc.height;
    P.K = 1;
    P.nsub = 1;
    P.ev_thresh = st->ev_thresh;
    P.seg_thresh = st->seg_thresh;
    P.max_steps = 1u << 20;
    P.counters = st->d_counters;
    P.partials = d_acc.p;
    P.nvalid = m;
    P.trunc = st->d_trunc;
    P.list = d_list.p;
    zero_u32_kernel<<<1, 64>>>(st->d_counters, kXcds);
    const uint32_t grid = std::min(st->grid_max, (m + 255) / 256);
    if(Mc.sensor == HTGX_SENSOR_CAMERA) {
      if(Mc.spectral_type == HTGX_SPECTRAL_LW)
        wavefront_list_kernel<HTGX_SENSOR_CAMERA, HTGX_SPECTRAL_LW><<<grid, 256>>>(KArgs{st->S, Mc, P});
      else
        wavefront_list_kernel<HTGX_SENSOR_CAMERA, HTGX_SPECTRAL_SW><<<grid, 256>>>(KArgs{st->S, Mc, P});
    } else {
      if(Mc.spectral_type == HTGX_SPECTRAL_LW)
        wavefront_list_kernel<HTGX_SENSOR_RECTANGLE, HTGX_SPECTRAL_LW><<<grid, 256>>>(KArgs{st->S, Mc, P});
      else
        wavefront_list_kernel<HTGX_SENSOR_RECTANGLE, HTGX_SPECTRAL_SW><<<grid, 256>>>(KArgs{st->S, Mc, P});
    }
    HIP_OK(hipGetLastError());
    std::vector<htgx_mc_accum> acc(m);
    HIP_OK(hipMemcpy(acc.data(), d_acc.p, m * sizeof(htgx_mc_accum), hipMemcpyDeviceToHost));
    /* a rejected realisation (a flux sample under the ground) is NaN, as in
     * the plain solver */
    for(uint32_t k = 0; k < m; ++k)
      weights[idx[k]] = acc[k].n ? acc[k].sum : std::nan("");
  }
  return HTGX_OK;
}

int htgx_debug_trace_segments_fp32_impl(htgx_ctx* ctx, const double* rays, uint32_t nrays, int prop,
                                        uint32_t iband, uint32_t iquad, htgx_segment* segs,
                                        uint32_t max_segs, uint32_t* nsegs)
{
  int res;
  if((res = wavefront_prepare(ctx)) != HTGX_OK) return res;
  for(uint32_t i = 0; i < nrays; ++i) {
    const double* r = rays + (size_t)i * 8;
    if(r[6] != 0.0 || !(r[7] >= 0.0)) return HTGX_BAD_ARG; /* queries start at their origin */
  }
  DevArrayLite<double> d_rays;
  DevArrayLite<htgx_segment> d_segs;
  DevArrayLite<uint32_t> d_n;
  if((res = d_rays.alloc((size_t)nrays * 8)) || (res = d_segs.alloc((size_t)nrays * max_segs)) ||
     (res = d_n.alloc(nrays)))
    return res;
  HIP_OK(hipMemcpy(d_rays.p, rays, (size_t)nrays * 8 * sizeof(double), hipMemcpyHostToDevice));
  wf_trace_segments_kernel<<<(nrays + 255) / 256, 256>>>(ctx->wf->S, d_rays.p, nrays, prop, iband,
                                                         iquad, d_segs.p, max_segs, d_n.p);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(segs, d_segs.p, (size_t)nrays * max_segs * sizeof(htgx_segment),
                   hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(nsegs, d_n.p, (size_t)nrays * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return HTGX_OK;
}

void htgx_wavefront_release(htgx_ctx* ctx)
{
  if(!ctx || !ctx->wf) return;
  wavefront_free(ctx->wf);
  ctx->wf = nullptr;
}
