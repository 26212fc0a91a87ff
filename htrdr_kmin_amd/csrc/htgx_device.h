/*
 * Device-side FP64 building blocks of the null-collision path tracer.
 *
 * Scene layout in HBM (built by svo.cpp, uploaded by api.hip):
 *  - lwc    : FP32 liquid water content per cloud cell, X fastest;
 *  - lmap   : one 32-bit word per octree leaf block (BxBxB cells): the
 *             global id of the octree leaf that contains the block (bits
 *             0-27) and its level (bits 28-31). Finding the leaf of a point is
 *             one load, not a root-to-leaf descent;
 *  - bounds : FP32 (LWC_min, LWC_max) per octree node, all levels, numbered
 *             level by level (level 0 first), X fastest within a level. The
 *             leaf id of lmap indexes it directly;
 *  - gas    : FP64 absorption / scattering per (band, quad, layer) and
 *             (band, layer), temperature per layer.
 * The property bounds of a leaf for a band are c(prop, band) * LWC_min/max
 * plus the gas value of the layer, with c = cext*(1-ssa), cext*ssa or cext.
 *
 * All arithmetic here is FP64 with floating-point contraction disabled at
 * build time, so that segment distances, voxel indices and bounds are
 * bit-identical to the CPU oracle's for identical rays.
 */
#ifndef HTGX_DEVICE_H
#define HTGX_DEVICE_H

#include "htgx_mesh.h"
#include "htgx_rng.h"
#include "htgx.h"

#include <float.h>
#include <stdint.h>

namespace htgx {

#define HTGX_D __device__ __forceinline__

constexpr double kPi = 3.14159265358979323846;
constexpr uint32_t kLeafIdBits = 28;
constexpr uint32_t kLeafIdMask = (1u << kLeafIdBits) - 1u;

/* Ground materials in the general model (spectral BRDF lists and two-sided
 * interfaces, as htrdr-materials, mrumtl and htrdr-obj files give them) */
struct DevMtl {
  uint32_t first, count; /* BRDF entries; none: semi-transparent (the sky) */
  int32_t list;          /* enum htgx_brdf_list */
  int32_t pad;
  double T;              /* emission temperature, K */
};

struct DevBrdf {
  double lo, hi;         /* nm: [w, w] or the band */
  double refl;
  int32_t type, pad;     /* enum htgx_brdf_type */
};

struct DevIfc {
  int32_t front, back, thin, pad; /* material indices; thin -1: none */
};

struct DevScene {
  /* spectral */
  uint32_t nbands;
  const double* band_hi;     /* [nbands] */
  const uint32_t* quad_off;  /* [nbands + 1] */
  const double* quad_cdf;    /* [#quads] */
  /* gas */
  uint32_t nlayers;
  double zb, dzl, inv_dzl;
  const double* gas_ka;      /* [#quads][nlayers] */
  const double* gas_ks;      /* [nbands][nlayers] */
  const double* gas_T;       /* [nlayers] */
  /* clouds */
  int has_clouds;
  uint32_t B;
  uint32_t def[3];
  uint32_t dims0[3];
  double lower[3];
  double bs[3], inv_bs[3], inv_cs[3];
  int repeat;        /* clouds repeated along X and Y (-r) */
  double period[2];  /* repetition period: dims0 * bs */
  const float* lwc;
  const uint32_t* lmap;
  const float2* bounds;
  const double* cext;
  const double* ssa;
  const double* hg_g;
  /* ground */
  int ground_type;
  double plane_z;
  uint32_t plane_mtl;
  const double* mtl_albedo;
  const double* mtl_T;
  /* mesh ground: BVH nodes, triangles and their materials in leaf order */
  const BvhNode* bvh;
  uint32_t nbvh;
  const BvhTri* tris;
  const uint32_t* tri_mtl;
  int ground_repeat;         /* -R: the mesh repeats along X, Y */
  double glo[3], ghi[3];     /* the mesh's bounding box */
  /* general materials: when set, tri_mtl and plane_mtl index ifcs */
  int gmat;
  const DevMtl* mtls;
  const DevBrdf* brdfs;
  const DevIfc* ifcs;
  int32_t sky_mtl;           /* the sky's material (-1: none) */
  int explicit_ifc;          /* interfaces given, not made from materials */
  /* sun */
  double sun_dir[3];
  double sun_cos_half;
  double sun_solid_angle;
};

struct DevMap {
  uint32_t width, height, spp, spp_offset;
  uint64_t seed;
  int sensor;
  int spectral_type;
  double pos[3], x[3], y[3], z[3];
  double tan_half, aspect;
  int ortho;               /* camera: orthographic projection */
  double ortho_half_h;     /* orthographic: half height of the image plane */
  double lens_radius;      /* perspective: thin-lens radius (0: pinhole) */
  double focal_dst;        /* thin lens: distance of the plane in focus */
  double axis_x[3], axis_y[3];
  /* wavelength sampler (src/core/htrdr_ran_wlen_planck.c) */
  int wmode; /* 0 monochromatic, 1 discrete CDF, 2 continuous bisection,
                3 CIE 1931 channel (piecewise-linear CDF inversion) */
  double range[2];
  double band_len;
  uint32_t nwbands;
  const double* wcdf;
  const double* wpdf;
  double Tref;
  int cie_channel;  /* wmode 3: 0 X, 1 Y, 2 Z */
  double cie_rcp;   /* wmode 3: 1 / integral of the channel's fit, nm^-1 */
  double us_per_tick; /* wall clock period in microseconds */
};

/* Event counters of one thread (roofline byte model) */
struct Counters {
  uint64_t nleaf; /* cloud-leaf lookups: 4 B leaf word + 8 B bounds */
  uint64_t nseg;  /* segments handed to the filters */
  uint64_t ncoll; /* tentative collisions */
  uint64_t ncell; /* LWC fetches (4 B) */
  uint64_t nev;   /* scattering / surface events */
};

/*******************************************************************************
 * Small helpers
 ******************************************************************************/
HTGX_D double dot3(const double a[3], const double b[3])
{
  return a[0]*b[0] + a[1]*b[1] + a[2]*b[2];
}

HTGX_D void normalize3(double v[3])
{
  const double len = sqrt(dot3(v, v));
  v[0] = v[0] / len; v[1] = v[1] / len; v[2] = v[2] / len;
}

/* Orthonormal basis completing unit n (Duff et al., JCGT 6(1) 2017) */
HTGX_D void basis(const double n[3], double b1[3], double b2[3])
{
  const double sign = n[2] >= 0 ? 1.0 : -1.0;
  const double a = -1.0 / (sign + n[2]);
  const double b = n[0] * n[1] * a;
  b1[0] = 1.0 + sign * n[0] * n[0] * a;
  b1[1] = sign * b;
  b1[2] = -sign * n[0];
  b2[0] = b;
  b2[1] = sign + n[1] * n[1] * a;
  b2[2] = -n[1];
}

HTGX_D void local_to_world(const double n[3], double cos_t, double phi, double out[3])
{
  double b1[3], b2[3];
  const double sin_t = sqrt(fmax(0.0, 1.0 - cos_t * cos_t));
  const double cx = cos(phi) * sin_t;
  const double cy = sin(phi) * sin_t;
  basis(n, b1, b2);
  out[0] = b1[0]*cx + b2[0]*cy + n[0]*cos_t;
  out[1] = b1[1]*cx + b2[1]*cy + n[1]*cos_t;
  out[2] = b1[2]*cx + b2[2]*cy + n[2]*cos_t;
}

/* Exp(1) sample, ssp_ran_exp(rng, 1) */
HTGX_D double ran_exp(Stream& rng) { return -log(1.0 - rng.canonical()); }

/*******************************************************************************
 * Spectral functions (src/core/htrdr_spectral.h:56-148)
 ******************************************************************************/
HTGX_HD double planck_mono(double lambda, double T)
{
  const double c = 299792458.0;
  const double h = 6.62607015e-34;
  const double k = 1.380649e-23;
  const double l2 = lambda * lambda;
  const double l5 = l2 * l2 * lambda;
  return ((2.0 * h * c * c) / l5) / (exp(h * c / (lambda * k * T)) - 1.0);
}

HTGX_HD double wiebelt(double v)
{
  const double f = 15.0 / (kPi * kPi * kPi * kPi);
  const double z0 = 1.0 / 3.0;
  const double z1 = 1.0 / 8.0;
  const double z2 = 1.0 / 60.0;
  const double z4 = 1.0 / 5040.0;
  const double z6 = 1.0 / 272160.0;
  const double z8 = 1.0 / 13305600.0;
  double w;
  if(v >= 2.0) {
    w = 0;
    for(int m = 1; m < 6; ++m)
      w += exp(-m * v) / (m * m * m * m) * (((m * v + 3) * m * v + 6) * m * v + 6);
    w = w * f;
  } else {
    const double v2 = v * v;
    const double v4 = v2 * v2;
    w = z0 - z1 * v + z2 * v2 - z4 * v2 * v2 + z6 * v4 * v2 - z8 * v4 * v4;
    w = 1. - f * v2 * v * w;
  }
  return w;
}

HTGX_HD double blackbody_fraction(double l0, double l1, double T)
{
  const double C2 = 1.43877735e-2;
  return wiebelt(C2 / l1 / T) - wiebelt(C2 / l0 / T);
}

HTGX_HD double planck_range(double lmin, double lmax, double T)
{
  if(lmin == lmax) return planck_mono(lmin, T);
  const double T2 = T * T;
  const double T4 = T2 * T2;
  return blackbody_fraction(lmin, lmax, T) * 5.6696e-8 * T4 / (kPi * (lmax - lmin));
}

/* The CIE 1931 colour matching functions as the sums of piecewise Gaussians
 * of C. Wyman, P.-P. Sloan and P. Shirley, "Simple Analytic Approximations
 * to the CIE XYZ Color Matching Functions", JCGT 2(2), 2013 -- the fits
 * src/core/htrdr_ran_wlen_cie_xyz.c:80-107 integrates. A lobe has its peak
 * wavelength and one inverse width on each side of it. */
HTGX_HD double cie_lobe(double l, double peak, double inv_w_lo, double inv_w_hi)
{
  const double t = (l - peak) * (l < peak ? inv_w_lo : inv_w_hi);
  return exp(-0.5 * t * t);
}

HTGX_HD double cie_fit(int channel, double l)
{
  if(channel == 0)
    return 0.362 * cie_lobe(l, 442.0, 0.0624, 0.0374) + 1.056 * cie_lobe(l, 599.8, 0.0264, 0.0323)
         - 0.065 * cie_lobe(l, 501.1, 0.0490, 0.0382);
  if(channel == 1)
    return 0.821 * cie_lobe(l, 568.8, 0.0213, 0.0247) + 0.286 * cie_lobe(l, 530.9, 0.0613, 0.0322);
  return 1.217 * cie_lobe(l, 437.0, 0.0845, 0.0278) + 0.681 * cie_lobe(l, 459.0, 0.0385, 0.0725);
}

/* Wavelength within band [lmin, lmin + len[ whose density is linear between
 * the channel's values at the band ends: root of the quadratic CDF equal to
 * r1 (sample_cie_xyz, src/core/htrdr_ran_wlen_cie_xyz.c:109-181; the band
 * middle when no root falls in the band) */
HTGX_HD double cie_band_sample(int channel, double lmin, double len, double r1)
{
  const double lmax = lmin + len;
  const double fmin = cie_fit(channel, lmin), fmax = cie_fit(channel, lmax);
  const double a = 0.5 * (fmax - fmin) / len;
  const double b = (lmax * fmin - lmin * fmax) / len;
  const double c = -lmin * fmin + lmin * lmin * a;
  const double d = 0.5 * (fmax + fmin) * len;
  double delta = b * b - 4.0 * a * (c - d * r1);
  if(delta < 0 && fabs(delta) <= 1e-6) delta = 0;
  const double sq = sqrt(delta);
  const double l1 = (-b - sq) / (2.0 * a), l2 = (-b + sq) / (2.0 * a);
  if(lmin <= l1 && l1 < lmax) return l1;
  if(lmin <= l2 && l2 < lmax) return l2;
  return (lmin + lmax) * 0.5;
}

/* htrdr_ran_wlen_planck_sample, src/core/htrdr_ran_wlen_planck.c:359-376,
 * and htrdr_ran_wlen_cie_xyz_sample_X/Y/Z, htrdr_ran_wlen_cie_xyz.c:346-397.
 * The pdf is in nm^-1; for a CIE channel it is the constant 1 / integral of
 * the channel, as the reference returns it. Without RARE, the modes that few
 * maps use (3, CIE channels, and 2, Planck over a continuous range) are
 * compiled out: the caller guarantees the mode is 0 or 1. */
template<bool RARE = true>
HTGX_D double sample_wavelength(const DevMap& M, double r0, double r1, double* pdf)
{
  if(RARE && M.wmode == 3) {
    uint32_t lo = 0, hi = M.nwbands - 1;
    while(lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if(M.wcdf[mid] > r0) hi = mid; else lo = mid + 1;
    }
    *pdf = M.cie_rcp;
    return cie_band_sample(M.cie_channel, M.range[0] + (double)lo * M.band_len, M.band_len, r1);
  }
  if(M.wmode == 1) {
    uint32_t lo = 0, hi = M.nwbands - 1;
    while(lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if(M.wcdf[mid] > r0) hi = mid; else lo = mid + 1;
    }
    *pdf = M.wpdf[lo] * (1.0 / M.band_len);
    return M.range[0] + (double)lo * M.band_len + M.band_len * r1;
  } else if(!RARE || M.wmode == 0) {
    *pdf = 1.0;
    return M.range[0];
  } else {
    const double rm0 = M.range[0] * 1e-9, rm1 = M.range[1] * 1e-9;
    double lmin = rm0, lmax = rm1, l = 0, lprev = 0, bf, bfprev = 0;
    const double bfmm = blackbody_fraction(rm0, rm1, M.Tref);
    for(int i = 0; i < 100; ++i) {
      l = (lmin + lmax) * 0.5;
      bf = blackbody_fraction(rm0, l, M.Tref);
      if(bf / bfmm < r0) lmin = l; else lmax = l;
      if(fabs(lprev - l) < 1e-15 || fabs(bfprev - bf) < 1e-6) break;
      lprev = l;
      bfprev = bf;
    }
    *pdf = planck_range(l, l, M.Tref) / (planck_range(rm0, rm1, M.Tref) * (rm1 - rm0)) * 1e-9;
    return l * 1e9;
  }
}

/* htsky_find_spectral_band: first band whose upper bound exceeds wlen */
HTGX_D uint32_t find_band(const DevScene& S, double wlen)
{
  uint32_t lo = 0, hi = S.nbands - 1;
  while(lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if(S.band_hi[mid] > wlen) hi = mid; else lo = mid + 1;
  }
  return lo;
}

/* htsky_spectral_band_sample_quadrature: first k-point whose cumulative
 * weight exceeds r */
HTGX_D uint32_t sample_quad(const DevScene& S, uint32_t iband, double r)
{
  const uint32_t off = S.quad_off[iband];
  const uint32_t nq = S.quad_off[iband + 1] - off;
  uint32_t q = 0;
  for(; q + 1 < nq; ++q) if(S.quad_cdf[off + q] > r) break;
  return q;
}

/*******************************************************************************
 * Octree x gas-layer segments
 ******************************************************************************/
struct Segment {
  double t0, t1, kmin, kmax;
  int in_cloud;
  int32_t layer;
  uint32_t leaf;        /* global leaf id */
  uint32_t lo[3], hi[3]; /* leaf range, in blocks */
  double shift[2];      /* X/Y offset of the leaf's copy of the grid (-r) */
};

/* Cap on the copies of a repeated cloud grid one ray crosses (a horizontal
 * ray inside the cloud layer never leaves it) */
constexpr int64_t kMaxTiles = 1 << 16;

HTGX_D double gas_k(const DevScene& S, int prop, uint32_t iband, uint32_t iquad, int32_t j)
{
  const uint32_t nl = S.nlayers;
  const double ka = S.gas_ka[(size_t)(S.quad_off[iband] + iquad) * nl + (uint32_t)j];
  const double ks = S.gas_ks[(size_t)iband * nl + (uint32_t)j];
  return prop == HTGX_Ka ? ka : prop == HTGX_Ks ? ks : ka + ks;
}

HTGX_D double cloud_c(const DevScene& S, int prop, uint32_t iband)
{
  const double ce = S.cext[iband];
  const double w = S.ssa[iband];
  return prop == HTGX_Ka ? ce * (1.0 - w) : prop == HTGX_Ks ? ce * w : ce;
}

/* Front-to-back iterator over the segments of a ray range: pieces of the ray
 * that lie in one octree leaf (or outside the cloud grid) and one gas layer.
 * This is what the svx traversal hands to the hit filters at
 * ..._sw.c:243-244,369-370 and ..._lw.c:205-206, with the leaf bounds of
 * ..._sw.c:87-88,163-166 folded in. The leaf following the current one is
 * found by integer stepping: the exit face gives the neighbour block along
 * its axis, the exit point gives (clamped into the current leaf) the block
 * along the two other axes, and the leaf word of that block gives the leaf. */
struct SegIter {
  double o[3], d[3], inv[3];
  double t, tmax;
  int cstate;  /* 0 before the cloud grid, 1 inside, 2 after / no clouds */
  double tbox;
  uint32_t lo[3], hi[3]; /* current leaf range in blocks */
  uint32_t leaf;
  float lmn, lmx;
  double tleaf; /* exit distance of the current leaf */
  int exit_axis;
  int32_t j;    /* gas layer, -1 below, nlayers above */
  int done;
  /* repeated clouds: o[0..1] is the origin relative to the copy k of the
   * grid, shift = k * period, ow the world origin */
  int64_t k[2];
  double shift[2], ow[2];
  int64_t ntiles;

  /* Enter copy k of the grid along axis a (0 or 1) */
  HTGX_D void set_tile(const DevScene& S, int a, int64_t ka)
  {
    k[a] = ka;
    shift[a] = (double)ka * S.period[a];
    o[a] = ow[a] - shift[a];
  }

  HTGX_D double plane(const DevScene& S, int a, uint32_t k) const
  {
    return S.lower[a] + (double)k * S.bs[a];
  }

  /* Fetch the leaf containing block c and compute its exit distance */
  HTGX_D void load_leaf(const DevScene& S, uint32_t cx, uint32_t cy, uint32_t cz, Counters* cnt)
  {
    const uint32_t word = S.lmap[((size_t)cz * S.dims0[1] + cy) * S.dims0[0] + cx];
    const uint32_t L = word >> kLeafIdBits;
    leaf = word & kLeafIdMask;
    const float2 b = S.bounds[leaf];
    lmn = b.x;
    lmx = b.y;
    const uint32_t c[3] = {cx, cy, cz};
    double best = INFINITY;
    exit_axis = 0;
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      lo[a] = (c[a] >> L) << L;
      hi[a] = min(lo[a] + (1u << L), S.dims0[a]);
      if(d[a] != 0) {
        const double ta = (plane(S, a, d[a] > 0 ? hi[a] : lo[a]) - o[a]) * inv[a];
        if(ta < best) { best = ta; exit_axis = a; }
      }
    }
    tleaf = best;
    if(cnt) cnt->nleaf++;
  }

  HTGX_D void locate(const DevScene& S, double at, Counters* cnt)
  {
    uint32_t c[3];
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      const double p = o[a] + at * d[a];
      const double f = floor((p - S.lower[a]) * S.inv_bs[a]);
      const double fmaxc = (double)(S.dims0[a] - 1);
      c[a] = f < 0.0 ? 0u : f > fmaxc ? S.dims0[a] - 1 : (uint32_t)f;
    }
    load_leaf(S, c[0], c[1], c[2], cnt);
  }

  /* Move to the leaf across the exit face of the current one */
  HTGX_D void step_leaf(const DevScene& S, Counters* cnt)
  {
    uint32_t c[3];
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      if(a == exit_axis) {
        const bool out = d[a] > 0 ? hi[a] >= S.dims0[a] : lo[a] == 0;
        if(out && (!S.repeat || a == 2 || ++ntiles > kMaxTiles)) { cstate = 2; return; }
        if(out) {
          /* into the next copy of the grid */
          set_tile(S, a, k[a] + (d[a] > 0 ? 1 : -1));
          c[a] = d[a] > 0 ? 0u : S.dims0[a] - 1;
        } else {
          c[a] = d[a] > 0 ? hi[a] : lo[a] - 1;
        }
      } else {
        const double p = o[a] + t * d[a];
        const double f = floor((p - S.lower[a]) * S.inv_bs[a]);
        c[a] = f < (double)lo[a] ? lo[a] : f > (double)(hi[a] - 1) ? hi[a] - 1 : (uint32_t)f;
      }
    }
    load_leaf(S, c[0], c[1], c[2], cnt);
  }

  HTGX_D void init(const DevScene& S, const double org[3], const double dir[3],
                   double r0, double r1, Counters* cnt)
  {
#pragma unroll
    for(int a = 0; a < 3; ++a) { o[a] = org[a]; d[a] = dir[a]; inv[a] = 1.0 / dir[a]; }
    t = r0;
    tmax = r1;
    done = 0;
    cstate = 2;
    k[0] = k[1] = 0;
    shift[0] = shift[1] = 0.0;
    ow[0] = o[0];
    ow[1] = o[1];
    ntiles = 0;
    if(S.has_clouds) {
      double tn = -INFINITY, tf = INFINITY;
      bool miss = false;
      /* repeated clouds: an infinite layer along X and Y */
      const int a0 = S.repeat ? 2 : 0;
      for(int a = a0; a < 3; ++a) {
        const double plo = plane(S, a, 0), phi = plane(S, a, S.dims0[a]);
        if(d[a] != 0) {
          const double ta = (plo - o[a]) * inv[a];
          const double tb = (phi - o[a]) * inv[a];
          const double n = ta < tb ? ta : tb;
          const double f = ta < tb ? tb : ta;
          if(n > tn) tn = n;
          if(f < tf) tf = f;
        } else if(o[a] < plo || o[a] > phi) {
          miss = true;
        }
      }
      if(!miss && fmax(tn, t) < fmin(tf, tmax)) {
        if(S.repeat) {
          /* the copy of the grid holding the entry point */
          const double te = fmax(tn, t);
          for(int a = 0; a < 2; ++a)
            set_tile(S, a, (int64_t)floor((ow[a] + te * d[a] - S.lower[a]) / S.period[a]));
        }
        if(tn > t) { cstate = 0; tbox = tn; }
        else { cstate = 1; locate(S, t, cnt); }
      }
    }
    const double z = o[2] + t * d[2];
    const double fj = floor((z - S.zb) * S.inv_dzl);
    j = fj < -1.0 ? -1 : fj > (double)S.nlayers ? (int32_t)S.nlayers : (int32_t)fj;
  }

  /* Next segment holding some medium; false once the range is exhausted */
  HTGX_D bool next(const DevScene& S, int prop, uint32_t iband, uint32_t iquad,
                   Segment& seg, Counters* cnt)
  {
    while(!done && t < tmax) {
      const double tc = cstate == 0 ? tbox : cstate == 1 ? fmax(tleaf, t) : INFINITY;
      double tg;
      if(d[2] > 0 && j < (int32_t)S.nlayers) {
        tg = (S.zb + (double)(j + 1) * S.dzl - o[2]) * inv[2];
      } else if(d[2] < 0 && j >= 0) {
        tg = (S.zb + (double)j * S.dzl - o[2]) * inv[2];
      } else {
        tg = INFINITY;
      }
      tg = fmax(tg, t);
      const double t1 = fmin(fmin(tc, tg), tmax);
      bool emit = false;
      if(t1 > t) {
        const bool in_gas = j >= 0 && j < (int32_t)S.nlayers;
        const bool in_cloud = cstate == 1;
        if(in_gas || in_cloud) {
          const double g = in_gas ? gas_k(S, prop, iband, iquad, j) : 0.0;
          const double c = in_cloud ? cloud_c(S, prop, iband) : 0.0;
          seg.t0 = t;
          seg.t1 = t1;
          seg.in_cloud = in_cloud;
          seg.layer = in_gas ? j : -1;
          if(in_cloud) {
            seg.kmin = g + c * (double)lmn;
            seg.kmax = g + c * (double)lmx;
            seg.leaf = leaf;
#pragma unroll
            for(int a = 0; a < 3; ++a) { seg.lo[a] = lo[a]; seg.hi[a] = hi[a]; }
            seg.shift[0] = shift[0];
            seg.shift[1] = shift[1];
          } else {
            seg.kmin = g;
            seg.kmax = g;
            seg.leaf = 0xFFFFFFFFu;
            seg.shift[0] = seg.shift[1] = 0.0;
          }
          emit = true;
        }
      }
      if(t1 >= tmax) {
        done = 1;
      } else {
        t = t1;
        if(tg == t1) j += d[2] > 0 ? 1 : -1;
        if(tc == t1) {
          if(cstate == 0) { cstate = 1; locate(S, t, cnt); }
          else if(cstate == 1) step_leaf(S, cnt);
        }
        if(tg == INFINITY && tc == INFINITY) done = 1;
      }
      if(emit) {
        if(cnt) cnt->nseg++;
        return true;
      }
    }
    return false;
  }
};

/* LWC of the cloud cell holding x (world), clamped into the segment's leaf
 * of the segment's copy of the grid */
HTGX_D float seg_lwc(const DevScene& S, const Segment& seg, const double x[3], Counters* cnt)
{
  uint32_t ci[3];
#pragma unroll
  for(int a = 0; a < 3; ++a) {
    const double xa = a < 2 ? x[a] - seg.shift[a] : x[a];
    const double f = floor((xa - S.lower[a]) * S.inv_cs[a]);
    const double lo = (double)(seg.lo[a] * S.B);
    const double hi = (double)(seg.hi[a] * S.B - 1);
    ci[a] = f < lo ? seg.lo[a] * S.B : f > hi ? seg.hi[a] * S.B - 1 : (uint32_t)f;
  }
  if(cnt) cnt->ncell++;
  return S.lwc[((size_t)ci[2] * S.def[1] + ci[1]) * S.def[0] + ci[0]];
}

/* Property at a point of a segment (htsky_fetch_raw_property with the
 * [k_min, k_max] hint of ..._sw.c:202-204): the gas layer is the segment's,
 * and the cloud cell is clamped into the segment's leaf. cpnt: 1 gas,
 * 2 particles, 3 both. */
HTGX_D double raw_in_seg(const DevScene& S, const Segment& seg, int prop, int cpnt,
                         uint32_t iband, uint32_t iquad, const double x[3], Counters* cnt)
{
  double g = 0.0, c = 0.0;
  if((cpnt & 1) && seg.layer >= 0) g = gas_k(S, prop, iband, iquad, seg.layer);
  if((cpnt & 2) && seg.in_cloud) c = cloud_c(S, prop, iband) * (double)seg_lwc(S, seg, x, cnt);
  return g + c;
}

/*******************************************************************************
 * Null-collision tracking along one ray range. One routine serves the three
 * hit filters of the reference:
 *   MODE_TR  k_min/k_max transmissivity, ..._sw.c:147-251 (returns Tr);
 *   MODE_SC  delta tracking on the Ks majorant, ..._sw.c:73-145;
 *   MODE_LW  Kext majorant with absorption / scattering / null collisions,
 *            ..._lw.c:64-135.
 * A walk calls it from a single site, so its code and registers exist once.
 ******************************************************************************/
enum { MODE_TR = 0, MODE_SC = 1, MODE_LW = 2 };
enum { EV_NONE = 0, EV_HIT = 1, EV_ABS = 2, EV_SCA = 3 };

struct TrackOut {
  int event;   /* MODE_TR: EV_HIT = real collision (Tr = 0) */
  double dst;  /* distance of the event */
  double Tr;   /* MODE_TR transmissivity */
  Segment seg; /* segment holding the event */
};

HTGX_D void track(const DevScene& S, Stream& rng, int mode, int prop, uint32_t iband,
                  uint32_t iquad, const double o[3], const double d[3], double r1,
                  TrackOut& out, Counters* cnt)
{
  const int sprop = mode == MODE_SC ? HTGX_Ks : mode == MODE_LW ? HTGX_Kext : prop;
  SegIter it;
  double Ts = ran_exp(rng);
  double Tmin = 0;
  out.event = EV_NONE;
  out.dst = 0;
  it.init(S, o, d, 0.0, r1, cnt);
  while(it.next(S, sprop, iband, iquad, out.seg, cnt)) {
    const Segment& seg = out.seg;
    const double maj = mode == MODE_TR ? seg.kmax - seg.kmin : seg.kmax;
    if(mode == MODE_TR) Tmin += (seg.t1 - seg.t0) * seg.kmin;
    double t = seg.t0;
    for(;;) {
      const double T = (seg.t1 - t) * maj;
      if(Ts > T || !(maj > 0)) { Ts -= T; break; }
      t += Ts / maj;
      if(mode == MODE_SC && t > 1e9) break; /* ..._sw.c:123 */
      double x[3];
      x[0] = o[0] + t * d[0];
      x[1] = o[1] + t * d[1];
      x[2] = o[2] + t * d[2];
      if(cnt) cnt->ncoll++;
      const float v = seg.in_cloud ? seg_lwc(S, seg, x, cnt) : 0.f;
      if(mode == MODE_LW) {
        const double ga = seg.layer >= 0 ? gas_k(S, HTGX_Ka, iband, iquad, seg.layer) : 0.0;
        const double gs = seg.layer >= 0 ? gas_k(S, HTGX_Ks, iband, iquad, seg.layer) : 0.0;
        const double ca = seg.in_cloud ? cloud_c(S, HTGX_Ka, iband) * (double)v : 0.0;
        const double cs = seg.in_cloud ? cloud_c(S, HTGX_Ks, iband) * (double)v : 0.0;
        const double ka = ga + ca, ks = gs + cs;
        const double r = rng.canonical();
        const double pa = ka / maj;
        const double ps = ks / maj;
        if(r < pa) { out.event = EV_ABS; out.dst = t; return; }
        if(r < pa + ps) { out.event = EV_SCA; out.dst = t; return; }
      } else {
        const double g = seg.layer >= 0 ? gas_k(S, sprop, iband, iquad, seg.layer) : 0.0;
        const double c = seg.in_cloud ? cloud_c(S, sprop, iband) * (double)v : 0.0;
        const double k = g + c;
        const double proba = mode == MODE_TR ? (k - seg.kmin) / maj : k / maj;
        if(rng.canonical() < proba) {
          out.event = EV_HIT;
          out.dst = t;
          out.Tr = 0;
          return;
        }
      }
      Ts = ran_exp(rng);
    }
  }
  out.Tr = Tmin != 0 ? exp(-Tmin) : 1.0;
}

/*******************************************************************************
 * Ground, sun, phase functions
 ******************************************************************************/
struct GroundHit {
  bool hit;
  double t;
  double N[3];    /* geometric normal of the hit (unit) */
  uint32_t mtl;
  int32_t prim;   /* hit primitive: 0 the plane, else the mesh triangle; -1 none */
  bool edge;      /* the hit lies on an edge or a vertex of its triangle */
};

/* Closest ground hit in ]r0, r1[, htrdr_atmosphere_ground_trace_ray
 * (..._ground.c:170-217) with the self-hit rules of
 * src/core/htrdr_geometry.c:116-150 relative to the previous surface hit
 * (prev_prim, prev_edge; prev_prim = -1 after a volume event). The plane is
 * a flat quad repeated to infinity (-R): one primitive, id 0. A mesh is
 * traced in FP32 as the reference does. With `any', any accepted hit will do
 * (occlusion queries). */
HTGX_D GroundHit ground_trace(const DevScene& S, const double o[3], const double d[3],
                              double r0, double r1, int32_t prev_prim, bool prev_edge,
                              bool any = false)
{
  GroundHit h;
  h.hit = false;
  h.t = FLT_MAX; /* distance of a missed ray */
  h.N[0] = h.N[1] = h.N[2] = 0;
  h.mtl = 0;
  h.prim = -1;
  h.edge = false;
  if(S.ground_type == HTGX_GROUND_PLANE && prev_prim != 0 && d[2] != 0) {
    const double t = (S.plane_z - o[2]) / d[2];
    if(t > r0 && t < r1) {
      h.hit = true;
      h.t = t;
      h.N[2] = 1;
      h.mtl = S.plane_mtl;
      h.prim = 0;
    }
  } else if(S.ground_type == HTGX_GROUND_MESH) {
    const MeshHit m = mesh_trace_copies<double>(S.bvh, S.nbvh, S.tris, S.ground_repeat != 0, S.glo,
                                                S.ghi, o, d, r0, r1, prev_prim, prev_edge, any);
    if(m.tri >= 0) {
      const BvhTri& T = S.tris[m.tri];
      const float n0 = T.e1[1] * T.e2[2] - T.e1[2] * T.e2[1];
      const float n1 = T.e1[2] * T.e2[0] - T.e1[0] * T.e2[2];
      const float n2 = T.e1[0] * T.e2[1] - T.e1[1] * T.e2[0];
      h.hit = true;
      h.t = (double)m.t;
      h.N[0] = n0; h.N[1] = n1; h.N[2] = n2;
      normalize3(h.N);
      h.mtl = S.tri_mtl[m.tri];
      h.prim = m.tri;
      h.edge = mesh_on_edge(m.u, m.v);
    }
  }
  return h;
}

/* Material reflecting a hit of interface f by a ray along dir, front_hit
 * telling that the ray meets the interface's front side, i.e. that
 * N . dir > 0 for the normal N = (v1 - v0) x (v2 - v0) (the front normal of
 * htrdr-obj(5) is -N): the thin material if it has a BRDF, else the material
 * of the other side, else the hit side's (htrdr_interface_fetch_hit_mtl,
 * src/core/htrdr_interface.h:45-87) */
HTGX_D int32_t interface_mtl(const DevIfc& f, const DevMtl* mtls, bool front_hit)
{
  if(f.thin >= 0 && mtls[f.thin].count) return f.thin;
  const int32_t m = front_hit ? f.back : f.front;
  return mtls[m].count ? m : (front_hit ? f.front : f.back);
}

/* BRDF entry of material m at wavelength wlen (nm) with the canonical r, as
 * mrumtl_fetch_brdf selects it (called at src/core/htrdr_materials.c:430-433):
 * the band holding wlen, or one of the two entries around it, the upper one
 * with probability u = wlen's linear position between them (r < u); the
 * first or last entry outside the list */
HTGX_D const DevBrdf& brdf_fetch(const DevMtl& m, const DevBrdf* brdfs, double wlen, double r)
{
  const DevBrdf* e = brdfs + m.first;
  const uint32_t n = m.count;
  const bool bands = m.list == HTGX_BRDF_LIST_BANDS;
  /* first entry whose upper bound is >= wlen */
  uint32_t lo = 0, hi = n;
  while(lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if(e[mid].hi < wlen) lo = mid + 1; else hi = mid;
  }
  if(lo == n) return e[n - 1];
  if(lo == 0 || (bands && e[lo].lo <= wlen)) return e[lo];
  const double below = bands ? e[lo - 1].hi : e[lo - 1].lo;
  const double u = (wlen - below) / (e[lo].lo - below);
  return r < u ? e[lo] : e[lo - 1];
}

/* Mirror direction of dir about the unit normal N */
HTGX_D void reflect3(const double dir[3], const double N[3], double out[3])
{
  const double c = 2.0 * (dir[0] * N[0] + dir[1] * N[1] + dir[2] * N[2]);
  out[0] = dir[0] - c * N[0];
  out[1] = dir[1] - c * N[1];
  out[2] = dir[2] - c * N[2];
}

/* A rectangle sensor's sample whose up ray hits the ground with `h' lies
 * inside the ground if the side the ray travels in is not the sky's
 * (..._draw_map.c:75-100): with interfaces, the front (cos(up, -N) < 0) or
 * back material is compared with the sky's; else the ray must not meet the
 * side N points to from below (N . up >= 0) */
HTGX_D bool sensor_under_ground(const DevScene& S, const GroundHit& h)
{
  if(!S.explicit_ifc) return h.N[2] >= 0;
  const DevIfc& f = S.ifcs[h.mtl];
  return (h.N[2] > 0 ? f.front : f.back) != S.sky_mtl;
}

HTGX_D bool in_solar_cone(const DevScene& S, const double dir[3])
{
  return dot3(dir, S.sun_dir) >= S.sun_cos_half;
}

HTGX_D double sun_sample(const DevScene& S, Stream& rng, double out[3])
{
  const double u0 = rng.canonical(), u1 = rng.canonical();
  const double cos_t = 1.0 - u0 * (1.0 - S.sun_cos_half);
  local_to_world(S.sun_dir, cos_t, 2.0 * kPi * u1, out);
  return 1.0 / S.sun_solid_angle;
}

HTGX_D double hg_cos_sample(double g, double u)
{
  if(fabs(g) < 1e-6) return 1.0 - 2.0 * u;
  const double t = (1.0 - g * g) / (1.0 - g + 2.0 * g * u);
  const double c = (1.0 + g * g - t * t) / (2.0 * g);
  return c < -1 ? -1 : c > 1 ? 1 : c;
}

HTGX_D double hg_eval(double g, double mu)
{
  const double dd = 1.0 + g * g - 2.0 * g * mu;
  return (1.0 - g * g) / (4.0 * kPi * dd * sqrt(dd));
}

HTGX_D double rayleigh_cos_sample(double u)
{
  const double q = 4.0 * u - 2.0;
  const double A = cbrt(q + sqrt(q * q + 1.0));
  const double c = A - 1.0 / A;
  return c < -1 ? -1 : c > 1 ? 1 : c;
}

HTGX_D double rayleigh_eval(double mu) { return 3.0 / (16.0 * kPi) * (1.0 + mu * mu); }

HTGX_D void cos_hemisphere(const double n[3], double u0, double u1, double out[3])
{
  local_to_world(n, sqrt(1.0 - u0), 2.0 * kPi * u1, out);
}

HTGX_D double layer_temperature(const DevScene& S, int32_t j)
{
  const int32_t nl = (int32_t)S.nlayers;
  if(!S.gas_T || nl <= 0) return 0.0; /* no gas layers: no emission */
  j = j < 0 ? 0 : j >= nl ? nl - 1 : j;
  return S.gas_T[j];
}

/*******************************************************************************
 * One realisation as a state machine. Each stage that needs the medium hands
 * one tracking query to track(); the rest of the walk logic runs between
 * queries. Stages follow atmosphere_compute_radiance_sw (..._sw.c:256-511)
 * and _lw (..._lw.c:140-279); the per-realisation set-up follows
 * draw_pixel_xwave (..._draw_map.c:348-400) and draw_pixel_flux
 * (..._draw_map.c:233-302).
 ******************************************************************************/
enum Stage {
  ST_DIRECT = 0,   /* Kext transmissivity of the direct beam, ..._sw.c:324-337 */
  ST_SCATTER = 1,  /* next scattering or surface event, ..._sw.c:353-375 */
  ST_ABS = 2,      /* Ka transmissivity up to the event, ..._sw.c:394-399 */
  ST_SUN = 3,      /* Kext transmissivity towards the sun, ..._sw.c:468-485 */
  ST_LW = 4,       /* longwave tracking, ..._lw.c:186-206 */
  ST_WALK_DONE = 5,
  ST_REAL_DONE = 6
};

struct Path {
  /* realisation */
  double wlen, pdf, g, L_sun;
  uint32_t iband, iquad;
  int phase; /* 0: one walk; flux SW: 1 direct walk, 2 diffuse walk */
  double org[3], dir0[3];
  double cosN, Ld;
  double weight;
  bool rejected;
  /* walk */
  int stage, mask;
  bool surface, rayleigh, gh_hit, specular;
  bool prev_edge, gh_edge;
  int32_t prev_prim, gh_prim; /* previous surface hit (-1: none) and current hit */
  uint32_t gh_mtl;
  double pos[3], dir[3], pos_next[3], dir_next[3], sun_dir[3], N[3];
  double gh_t, dst, ksi, w, Tr_abs, R, sun_pdf, refl, albedo, gas_ratio;
};

HTGX_D void ground_setup(const DevScene& S, Path& P)
{
  const GroundHit gh = ground_trace(S, P.pos, P.dir, 0.0, DBL_MAX, P.prev_prim, P.prev_edge);
  P.gh_hit = gh.hit;
  P.gh_t = gh.t;
  P.gh_mtl = gh.mtl;
  P.gh_prim = gh.prim;
  P.gh_edge = gh.edge;
  P.N[0] = gh.N[0]; P.N[1] = gh.N[1]; P.N[2] = gh.N[2];
}

HTGX_D void sw_after_direct(const DevScene& S, Path& P)
{
  if(!(P.mask & HTGX_RADIANCE_DIFFUSE)) { P.stage = ST_WALK_DONE; return; }
  ground_setup(S, P);
  P.stage = ST_SCATTER;
}

HTGX_D void sw_begin(const DevScene& S, Path& P, int mask, const double pos[3], const double dir[3])
{
  P.mask = mask;
  P.pos[0] = pos[0]; P.pos[1] = pos[1]; P.pos[2] = pos[2];
  P.dir[0] = dir[0]; P.dir[1] = dir[1]; P.dir[2] = dir[2];
  P.ksi = 1;
  P.w = 0;
  P.prev_prim = -1;
  P.prev_edge = false;
  if((mask & HTGX_RADIANCE_DIRECT) && in_solar_cone(S, P.dir)) {
    const GroundHit h = ground_trace(S, P.pos, P.dir, 0.0, FLT_MAX, -1, false, true);
    if(!h.hit) { P.stage = ST_DIRECT; return; }
  }
  sw_after_direct(S, P);
}

HTGX_D void lw_begin(const DevScene& S, Path& P, const double pos[3], const double dir[3])
{
  P.pos[0] = pos[0]; P.pos[1] = pos[1]; P.pos[2] = pos[2];
  P.dir[0] = dir[0]; P.dir[1] = dir[1]; P.dir[2] = dir[2];
  P.w = 0;
  P.prev_prim = -1;
  P.prev_edge = false;
  ground_setup(S, P);
  P.stage = ST_LW;
}

/* Next-event estimate and Russian roulette closing a SW event, ..._sw.c:494-504 */
HTGX_D void sw_accumulate(const DevScene& S, Path& P, Stream& rng, double Tr)
{
  P.ksi *= P.Tr_abs;
  P.w += P.ksi * P.L_sun * Tr * P.R / P.sun_pdf;
  if(P.surface && rng.canonical() >= P.refl) { P.stage = ST_WALK_DONE; return; }
  P.pos[0] = P.pos_next[0]; P.pos[1] = P.pos_next[1]; P.pos[2] = P.pos_next[2];
  P.dir[0] = P.dir_next[0]; P.dir[1] = P.dir_next[1]; P.dir[2] = P.dir_next[2];
  ground_setup(S, P);
  P.stage = ST_SCATTER;
}

/* Tracking query of the current stage. Values are copied with selects (no
 * pointer into the path state) so that the state stays in registers. */
HTGX_D void stage_query(const Path& P, int* mode, int* prop, double o[3], double d[3], double* r1)
{
  const int st = P.stage;
  const bool sun = st == ST_SUN;
  *mode = st == ST_SCATTER ? MODE_SC : st == ST_LW ? MODE_LW : MODE_TR;
  *prop = st == ST_SCATTER ? HTGX_Ks : st == ST_ABS ? HTGX_Ka : HTGX_Kext;
  *r1 = st == ST_ABS ? P.dst : (st == ST_SCATTER || st == ST_LW) ? P.gh_t : (double)FLT_MAX;
#pragma unroll
  for(int a = 0; a < 3; ++a) {
    o[a] = sun ? P.pos_next[a] : P.pos[a];
    d[a] = sun ? P.sun_dir[a] : P.dir[a];
  }
}

/* Consume the result of the current stage's query */
HTGX_D void stage_consume(const DevScene& S, Path& P, Stream& rng, const TrackOut& out, Counters* cnt)
{
  switch(P.stage) {
    case ST_DIRECT:
      P.w = P.L_sun * out.Tr;
      sw_after_direct(S, P);
      break;

    case ST_SCATTER: {
      const bool hit = out.event == EV_HIT;
      if(!P.gh_hit && !hit) { P.stage = ST_WALK_DONE; break; }
      if(cnt) cnt->nev++;
      P.surface = !hit;
      P.dst = P.surface ? P.gh_t : out.dst;
      P.pos_next[0] = P.pos[0] + P.dir[0] * P.dst;
      P.pos_next[1] = P.pos[1] + P.dir[1] * P.dst;
      P.pos_next[2] = P.pos[2] + P.dir[2] * P.dst;
      /* ..._sw.c:392: the next traces leave this surface hit, if any */
      P.prev_prim = P.surface ? P.gh_prim : -1;
      P.prev_edge = P.surface && P.gh_edge;
      if(!P.surface) {
        /* gas vs particle scattering, ..._sw.c:427-438 */
        const double ks_gas = raw_in_seg(S, out.seg, HTGX_Ks, 1, P.iband, P.iquad, P.pos_next, nullptr);
        const double ks_par = raw_in_seg(S, out.seg, HTGX_Ks, 2, P.iband, P.iquad, P.pos_next, nullptr);
        const double ks = ks_par + ks_gas;
        P.gas_ratio = ks_gas / ks;
      }
      P.stage = ST_ABS;
    } break;

    case ST_ABS: {
      P.Tr_abs = out.Tr;
      if(P.Tr_abs <= 0) { P.stage = ST_WALK_DONE; break; }
      P.specular = false;
      if(P.surface) {
        if(S.gmat) {
          /* the hit side's material and its BRDF at the wavelength,
           * ..._sw.c:403-411 */
          const int32_t m = interface_mtl(S.ifcs[P.gh_mtl], S.mtls, dot3(P.N, P.dir) > 0);
          const DevBrdf& b = brdf_fetch(S.mtls[m], S.brdfs, P.wlen, rng.canonical());
          P.albedo = b.refl;
          P.specular = b.type == HTGX_BRDF_SPECULAR;
        } else {
          P.albedo = S.mtl_albedo[P.gh_mtl];
        }
        if(-(P.dir[0]*P.N[0] + P.dir[1]*P.N[1] + P.dir[2]*P.N[2]) < 0) {
          P.N[0] = -P.N[0]; P.N[1] = -P.N[1]; P.N[2] = -P.N[2];
        }
        if(P.specular) {
          reflect3(P.dir, P.N, P.dir_next);
        } else {
          const double u0 = rng.canonical(), u1 = rng.canonical();
          cos_hemisphere(P.N, u0, u1, P.dir_next);
        }
        P.refl = P.albedo;
      } else {
        P.rayleigh = rng.canonical() < P.gas_ratio;
        const double u0 = rng.canonical(), u1 = rng.canonical();
        const double mu = P.rayleigh ? rayleigh_cos_sample(u0) : hg_cos_sample(P.g, u0);
        local_to_world(P.dir, mu, 2.0 * kPi * u1, P.dir_next);
        P.refl = 1;
      }
      if(P.specular) {
        /* the mirror direction is the sun direction, ..._sw.c:446-455 */
        P.sun_pdf = 1;
        P.sun_dir[0] = P.dir_next[0]; P.sun_dir[1] = P.dir_next[1]; P.sun_dir[2] = P.dir_next[2];
        P.R = !in_solar_cone(S, P.sun_dir) || dot3(P.N, P.sun_dir) < 0 ? 0 : P.albedo;
      } else if(P.surface) {
        P.sun_pdf = sun_sample(S, rng, P.sun_dir);
        const double c = dot3(P.N, P.sun_dir);
        P.R = c < 0 ? 0 : P.albedo / kPi * c;
      } else {
        P.sun_pdf = sun_sample(S, rng, P.sun_dir);
        const double mu = dot3(P.dir, P.sun_dir);
        P.R = P.rayleigh ? rayleigh_eval(mu) : hg_eval(P.g, mu);
      }
      if(P.R <= 0) { sw_accumulate(S, P, rng, 0.0); break; }
      const GroundHit h = ground_trace(S, P.pos_next, P.sun_dir, 0.0, FLT_MAX, P.prev_prim,
                                       P.prev_edge, true);
      if(h.hit) { sw_accumulate(S, P, rng, 0.0); break; }
      P.stage = ST_SUN;
    } break;

    case ST_SUN:
      sw_accumulate(S, P, rng, out.Tr);
      break;

    case ST_LW: {
      if(!P.gh_hit && out.event == EV_NONE) { P.w = 0; P.stage = ST_WALK_DONE; break; }
      if(cnt) cnt->nev++;
      const double dst = out.event != EV_NONE ? out.dst : P.gh_t;
      P.pos_next[0] = P.pos[0] + P.dir[0] * dst;
      P.pos_next[1] = P.pos[1] + P.dir[1] * dst;
      P.pos_next[2] = P.pos[2] + P.dir[2] * dst;
      if(out.event == EV_ABS) {
        P.w = planck_mono(P.wlen * 1e-9, layer_temperature(S, out.seg.layer));
        P.stage = ST_WALK_DONE;
        break;
      }
      if(out.event == EV_SCA) {
        const double u0 = rng.canonical(), u1 = rng.canonical();
        local_to_world(P.dir, hg_cos_sample(P.g, u0), 2.0 * kPi * u1, P.dir_next);
        P.prev_prim = -1; /* ..._lw.c:236 */
        P.prev_edge = false;
      } else {
        double refl, T;
        bool specular = false;
        if(S.gmat) { /* ..._lw.c:249-252 */
          const int32_t m = interface_mtl(S.ifcs[P.gh_mtl], S.mtls, dot3(P.N, P.dir) > 0);
          const DevBrdf& b = brdf_fetch(S.mtls[m], S.brdfs, P.wlen, rng.canonical());
          refl = b.refl;
          specular = b.type == HTGX_BRDF_SPECULAR;
          T = S.mtls[m].T;
        } else {
          refl = S.mtl_albedo[P.gh_mtl];
          T = S.mtl_T ? S.mtl_T[P.gh_mtl] : 0.0;
        }
        if(-(P.dir[0]*P.N[0] + P.dir[1]*P.N[1] + P.dir[2]*P.N[2]) < 0) {
          P.N[0] = -P.N[0]; P.N[1] = -P.N[1]; P.N[2] = -P.N[2];
        }
        if(specular) {
          reflect3(P.dir, P.N, P.dir_next);
        } else {
          const double u0 = rng.canonical(), u1 = rng.canonical();
          cos_hemisphere(P.N, u0, u1, P.dir_next);
        }
        if(rng.canonical() >= refl) {
          P.w = T > 0 ? planck_mono(P.wlen * 1e-9, T) : 0;
          P.stage = ST_WALK_DONE;
          break;
        }
        P.prev_prim = P.gh_prim; /* ..._lw.c:272 */
        P.prev_edge = P.gh_edge;
      }
      P.pos[0] = P.pos_next[0]; P.pos[1] = P.pos_next[1]; P.pos[2] = P.pos_next[2];
      P.dir[0] = P.dir_next[0]; P.dir[1] = P.dir_next[1]; P.dir[2] = P.dir_next[2];
      ground_setup(S, P);
      P.stage = ST_LW;
    } break;

    default: break;
  }
}

/* Camera ray through the film point (fx, fy) in [0,1]^2 (fy = 0 at the top
 * row) with the lens sample (l0, l1), as scam_generate_ray does for the
 * cameras of src/atmosphere/htrdr_atmosphere.c:112-165:
 *  - orthographic: from the image plane of height 2 ortho_half_h centred on
 *    the position, along the view axis;
 *  - pinhole: from the position through the film point;
 *  - thin lens (the model of Pharr, Jakob and Humphreys, PBRT 3rd ed.,
 *    6.2.3): from a point of the lens disk, r = R sqrt(l0), phi = 2 pi l1,
 *    towards where the pinhole ray meets the plane in focus at focal_dst
 *    along the view axis. */
HTGX_D void camera_ray(const DevMap& M, double fx, double fy, double l0, double l1,
                       double org[3], double dir[3])
{
  const double sx = 2.0 * fx - 1.0, sy = 1.0 - 2.0 * fy;
  if(M.ortho) {
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      org[a] = M.pos[a] + M.x[a] * (sx * M.ortho_half_h * M.aspect) + M.y[a] * (sy * M.ortho_half_h);
      dir[a] = M.z[a];
    }
    return;
  }
#pragma unroll
  for(int a = 0; a < 3; ++a) {
    dir[a] = M.z[a] + M.x[a] * (sx * M.tan_half * M.aspect) + M.y[a] * (sy * M.tan_half);
    org[a] = M.pos[a];
  }
  if(M.lens_radius > 0) {
    const double r = M.lens_radius * sqrt(l0), phi = 2.0 * kPi * l1;
    const double lx = r * cos(phi), ly = r * sin(phi);
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      const double focus = M.pos[a] + dir[a] * M.focal_dst;
      org[a] = M.pos[a] + M.x[a] * lx + M.y[a] * ly;
      dir[a] = focus - org[a];
    }
  }
  normalize3(dir);
}

/* Start a realisation: sensor ray, wavelength, band, k-point, first walk */
HTGX_D void real_begin(const DevScene& S, const DevMap& M, Path& P, Stream& rng,
                       uint64_t pixel, uint64_t ireal)
{
  rng.init(M.seed, pixel, ireal);
  P.rejected = false;
  P.weight = 0;
  const uint32_t px = (uint32_t)(pixel % M.width), py = (uint32_t)(pixel / M.width);
  const double psx = 1.0 / (double)M.width, psy = 1.0 / (double)M.height;

  if(M.sensor == HTGX_SENSOR_CAMERA) {
    const double fx = ((double)px + rng.canonical()) * psx;
    const double fy = ((double)py + rng.canonical()) * psy;
    const double l0 = rng.canonical(), l1 = rng.canonical(); /* lens sample */
    camera_ray(M, fx, fy, l0, l1, P.org, P.dir0);
  } else {
    /* sample_rectangle_ray, ..._draw_map.c:52-107 */
    const double sx = ((double)px + rng.canonical()) * psx;
    const double sy = ((double)py + rng.canonical()) * psy;
#pragma unroll
    for(int a = 0; a < 3; ++a)
      P.org[a] = M.pos[a] + M.axis_x[a] * (sx * 2.0 - 1.0) + M.axis_y[a] * (sy * 2.0 - 1.0);
    const double up[3] = {0, 0, 1};
    const GroundHit h = ground_trace(S, P.org, up, 0.0, DBL_MAX, -1, false);
    if(h.hit && sensor_under_ground(S, h)) { P.rejected = true; P.stage = ST_REAL_DONE; return; }
    const double u0 = rng.canonical(), u1 = rng.canonical();
    cos_hemisphere(M.z, u0, u1, P.dir0);
  }

  const double r0 = rng.canonical(), r1 = rng.canonical(), r2 = rng.canonical();
  P.wlen = sample_wavelength(M, r0, r1, &P.pdf);
  P.pdf *= 1e9;
  P.iband = find_band(S, P.wlen);
  P.iquad = sample_quad(S, P.iband, r2);
  P.g = S.has_clouds ? S.hg_g[P.iband] : 0.0;
  P.L_sun = planck_mono(P.wlen * 1e-9, 5778.0);

  if(M.spectral_type == HTGX_SPECTRAL_LW) {
    P.phase = 0;
    lw_begin(S, P, P.org, P.dir0);
  } else if(M.sensor == HTGX_SENSOR_CAMERA) {
    P.phase = 0;
    sw_begin(S, P, HTGX_RADIANCE_ALL, P.org, P.dir0);
  } else {
    double sun_dir[3];
    (void)sun_sample(S, rng, sun_dir);
    P.cosN = dot3(M.z, sun_dir);
    P.Ld = 0;
    if(P.cosN > 0) { P.phase = 1; sw_begin(S, P, HTGX_RADIANCE_DIRECT, P.org, sun_dir); }
    else { P.phase = 2; sw_begin(S, P, HTGX_RADIANCE_DIFFUSE, P.org, P.dir0); }
  }
}

/* A walk ended: weight of the realisation or start of the next walk */
HTGX_D void walk_done(const DevScene& S, const DevMap& M, Path& P)
{
  if(M.sensor == HTGX_SENSOR_CAMERA) {
    P.weight = P.w / P.pdf;
    P.stage = ST_REAL_DONE;
  } else if(M.spectral_type == HTGX_SPECTRAL_LW) {
    P.weight = P.w * kPi / P.pdf;
    P.stage = ST_REAL_DONE;
  } else if(P.phase == 1) {
    P.Ld = P.w;
    P.phase = 2;
    sw_begin(S, P, HTGX_RADIANCE_DIFFUSE, P.org, P.dir0);
  } else {
    P.weight = (P.cosN * S.sun_solid_angle * P.Ld + kPi * P.w) / P.pdf;
    P.stage = ST_REAL_DONE;
  }
}

/* Run queries until the realisation ends */
HTGX_D void run_until_done(const DevScene& S, const DevMap& M, Path& P, Stream& rng,
                           TrackOut& out, Counters* cnt)
{
  for(;;) {
    while(P.stage == ST_WALK_DONE) walk_done(S, M, P);
    if(P.stage == ST_REAL_DONE) return;
    int mode, prop;
    double o[3], d[3], r1;
    stage_query(P, &mode, &prop, o, d, &r1);
    track(S, rng, mode, prop, P.iband, P.iquad, o, d, r1, out, cnt);
    stage_consume(S, P, rng, out, cnt);
  }
}

/* Neumaier compensated accumulation */
HTGX_D void nsum(double& sum, double& c, double x)
{
  const double t = sum + x;
  if(fabs(sum) >= fabs(x)) c += (sum - t) + x;
  else c += (x - t) + sum;
  sum = t;
}

} /* namespace htgx */

#endif /* HTGX_DEVICE_H */
