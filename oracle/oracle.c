/*
 * CPU FP64 oracle of the htrdr-atmosphere null-collision core.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU reference of libhtgx.so's tests. It is never linked
 * into, loaded by or called from the product; only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
 *
 * Design: written for clarity, not speed. The octree is a pointer tree built
 * recursively and traversed recursively front to back (a different algorithm
 * from the GPU's level-map stepping), the random streams are a from-spec
 * Threefry-4x64-20, and the walks follow the reference functions line by line
 * (citations are file:line under the htrdr tree).
 */
#define _POSIX_C_SOURCE 200809L
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979323846

/*******************************************************************************
 * Threefry-4x64-20 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11, section 3 and table 4). Replaces star-sp's
 * Threefry proxy of src/core/htrdr_draw_map.c:433-442.
 ******************************************************************************/
static uint64_t rotl64(uint64_t x, unsigned r) { return (x << r) | (x >> (64 - r)); }

void
oracle_threefry4x64_20(const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4])
{
  /* Rotation constants R_{4x64} of the paper's table 4, rows d mod 8 */
  static const unsigned R[8][2] = {
    {14,16},{52,57},{23,40},{5,37},{25,33},{46,12},{58,22},{32,32}};
  uint64_t ks[5];
  uint64_t x[4];
  int r, i;
  ks[4] = 0x1BD11BDAA9FC1A22ULL; /* key-schedule parity constant C240 */
  for(i = 0; i < 4; ++i) { ks[i] = key[i]; ks[4] ^= key[i]; x[i] = ctr[i] + ks[i]; }
  for(r = 0; r < 20; ++r) {
    const unsigned* rc = R[r % 8];
    if((r % 2) == 0) {
      x[0] += x[1]; x[1] = rotl64(x[1], rc[0]); x[1] ^= x[0];
      x[2] += x[3]; x[3] = rotl64(x[3], rc[1]); x[3] ^= x[2];
    } else {
      x[0] += x[3]; x[3] = rotl64(x[3], rc[0]); x[3] ^= x[0];
      x[2] += x[1]; x[1] = rotl64(x[1], rc[1]); x[1] ^= x[2];
    }
    if((r % 4) == 3) { /* key injection s after every 4 rounds */
      const uint64_t s = (uint64_t)(r / 4 + 1);
      for(i = 0; i < 4; ++i) x[i] += ks[(s + (uint64_t)i) % 5];
      x[3] += s;
    }
  }
  for(i = 0; i < 4; ++i) out[i] = x[i];
}

/* One random stream per realisation: key = (seed, pixel, realisation, 0),
 * counter = (block, 0, 0, 0); the 4 words of a block are used in order and
 * mapped to [0,1) with their 53 high bits. */
struct orng {
  uint64_t key[4];
  uint64_t block;
  uint64_t buf[4];
  int idx;
};

static void
orng_init(struct orng* rng, uint64_t seed, uint64_t pixel, uint64_t realisation)
{
  rng->key[0] = seed; rng->key[1] = pixel; rng->key[2] = realisation;
  rng->key[3] = 0;
  rng->block = 0;
  rng->idx = 4;
}

static double
canonical(struct orng* rng)
{
  if(rng->idx == 4) {
    const uint64_t ctr[4] = {rng->block, 0, 0, 0};
    oracle_threefry4x64_20(ctr, rng->key, rng->buf);
    rng->block += 1;
    rng->idx = 0;
  }
  return (double)(rng->buf[rng->idx++] >> 11) * 0x1.0p-53;
}

/* Exp(1) sample, ssp_ran_exp(rng, 1) (called at ..._sw.c:140,212,239,359):
 * inverse CDF on 1-u so that the result is finite */
static double ran_exp(struct orng* rng) { return -log(1.0 - canonical(rng)); }

void
oracle_stream_canonicals
  (uint64_t seed, uint64_t pixel, uint64_t realisation, uint32_t n, double* out)
{
  struct orng rng;
  uint32_t i;
  orng_init(&rng, seed, pixel, realisation);
  for(i = 0; i < n; ++i) out[i] = canonical(&rng);
}

/*******************************************************************************
 * Small vector helpers
 ******************************************************************************/
static double dot3(const double a[3], const double b[3])
{ return a[0]*b[0] + a[1]*b[1] + a[2]*b[2]; }
static void cross3(double r[3], const double a[3], const double b[3])
{
  double t[3];
  t[0] = a[1]*b[2] - a[2]*b[1];
  t[1] = a[2]*b[0] - a[0]*b[2];
  t[2] = a[0]*b[1] - a[1]*b[0];
  r[0] = t[0]; r[1] = t[1]; r[2] = t[2];
}
static double normalize3(double r[3], const double a[3])
{
  const double len = sqrt(dot3(a, a));
  r[0] = a[0]/len; r[1] = a[1]/len; r[2] = a[2]/len;
  return len;
}

/* Orthonormal basis (b1, b2) completing a unit vector n (Duff, Burgess,
 * Christensen, Hery, Kensler, Liani, Villemin, "Building an orthonormal
 * basis, revisited", JCGT 6(1), 2017, listing 3). */
static void
basis(const double n[3], double b1[3], double b2[3])
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

/* World direction from a local (cos theta, phi) around the unit vector n */
static void
local_to_world(const double n[3], double cos_t, double phi, double out[3])
{
  double b1[3], b2[3];
  const double sin_t = sqrt(fmax(0.0, 1.0 - cos_t*cos_t));
  const double cx = cos(phi) * sin_t;
  const double cy = sin(phi) * sin_t;
  basis(n, b1, b2);
  out[0] = b1[0]*cx + b2[0]*cy + n[0]*cos_t;
  out[1] = b1[1]*cx + b2[1]*cy + n[1]*cos_t;
  out[2] = b1[2]*cx + b2[2]*cy + n[2]*cos_t;
}

/*******************************************************************************
 * Spectral functions: src/core/htrdr_spectral.h:56-148, htrdr_spectral.c
 ******************************************************************************/
/* Planck's law in W/m^2/sr/m, src/core/htrdr_spectral.h:102-116 */
double
oracle_planck_monochromatic(double lambda, double temperature)
{
  const double c = 299792458.0;
  const double h = 6.62607015e-34;
  const double k = 1.380649e-23;
  const double l2 = lambda * lambda;
  const double l5 = l2 * l2 * lambda;
  return ((2.0 * h * c * c) / l5) / (exp(h * c / (lambda * k * temperature)) - 1.0);
}

/* Fraction of black-body emission beyond v = C2/(lambda T), series of
 * src/core/htrdr_spectral.h:56-83 (5 exponential terms for v >= 2, a
 * polynomial in v below) */
double
oracle_wiebelt(double v)
{
  const double f = 15.0 / (PI*PI*PI*PI);
  const double z0 = 1.0/3.0;
  const double z1 = 1.0/8.0;
  const double z2 = 1.0/60.0;
  const double z4 = 1.0/5040.0;
  const double z6 = 1.0/272160.0;
  const double z8 = 1.0/13305600.0;
  double w, v2, v4;
  int m;
  if(v >= 2.0) {
    w = 0;
    for(m = 1; m < 6; ++m)
      w += exp(-m*v)/(m*m*m*m) * (((m*v+3)*m*v+6)*m*v+6);
    w = w * f;
  } else {
    v2 = v*v;
    v4 = v2*v2;
    w = z0 - z1*v + z2*v2 - z4*v2*v2 + z6*v4*v2 - z8*v4*v4;
    w = 1. - f*v2*v*w;
  }
  return w;
}

/* src/core/htrdr_spectral.h:85-99 */
double
oracle_blackbody_fraction(double lambda0, double lambda1, double T)
{
  const double C2 = 1.43877735e-2; /* m.K */
  return oracle_wiebelt(C2 / lambda1 / T) - oracle_wiebelt(C2 / lambda0 / T);
}

/* src/core/htrdr_spectral.h:118-148 */
double
oracle_planck(double lmin, double lmax, double T)
{
  const double T2 = T*T;
  const double T4 = T2*T2;
  if(lmin == lmax) return oracle_planck_monochromatic(lmin, T);
  return oracle_blackbody_fraction(lmin, lmax, T) * 5.6696e-8 * T4 / (PI * (lmax - lmin));
}

/* Doubling then bisection, src/core/htrdr_spectral.c:31-89 */
int
oracle_brightness_temperature(double lmin, double lmax, double radiance, double* out)
{
  const size_t MAX_ITER = 100;
  const double eps_T = 1e-4;
  const double eps_B = radiance * 1e-8;
  double T = 0, T0 = 0, T1 = 0, T2 = 200, B, B0 = 0;
  size_t i;
  for(i = 0; i < MAX_ITER; ++i) {
    if(oracle_planck(lmin, lmax, T2) >= radiance) break;
    T2 *= 2;
  }
  if(i >= MAX_ITER) return HTGX_BAD_OP;
  for(i = 0; i < MAX_ITER; ++i) {
    T = (T1 + T2) * 0.5;
    B = oracle_planck(lmin, lmax, T);
    if(B < radiance) T1 = T; else T2 = T;
    if(fabs(T - T0) < eps_T || fabs(B - B0) < eps_B) break;
    T0 = T;
    B0 = B;
  }
  if(i >= MAX_ITER) return HTGX_BAD_OP;
  *out = T;
  return HTGX_OK;
}

/* src/core/htrdr_spectral.c:91-121 */
double
oracle_radiance_temperature(double lmin, double lmax, double radiance)
{
  double T = 0;
  double avg = radiance;
  if(lmin != lmax) avg /= (lmax - lmin);
  if(oracle_brightness_temperature(lmin, lmax, avg, &T) != HTGX_OK) T = 0;
  return T;
}

/*******************************************************************************
 * Planck wavelength sampler: src/core/htrdr_ran_wlen_planck.c
 ******************************************************************************/
struct oracle_wlen_planck {
  double range[2];
  double band_len;
  double Tref;
  size_t nbands;
  double* pdf;
  double* cdf;
  int cie;        /* -1: Planck; 0, 1, 2: CIE 1931 X, Y, Z channel */
  double cie_rcp; /* 1 / integral of the CIE channel, nm^-1 */
};

/* src/atmosphere/htrdr_atmosphere.c:53-70 */
size_t
oracle_spectral_bands_count(const double range_nm[2])
{
  return (size_t)rint(range_nm[1] - range_nm[0]);
}

/* src/core/htrdr_ran_wlen_planck.c:287-343 (creation) and :55-134 (CDF) */
int
oracle_wlen_planck_create
  (const double range[2], size_t nbands, double Tref, struct oracle_wlen_planck** out)
{
  struct oracle_wlen_planck* p = calloc(1, sizeof(*p));
  size_t i;
  double sum = 0;
  if(!p) return HTGX_MEM_ERR;
  p->range[0] = range[0];
  p->range[1] = range[1];
  p->Tref = Tref;
  p->nbands = nbands;
  p->cie = -1;
  if(nbands) {
    p->band_len = (range[1] - range[0]) / (double)nbands;
    p->pdf = malloc(nbands * sizeof(double));
    p->cdf = malloc(nbands * sizeof(double));
    if(!p->pdf || !p->cdf) { oracle_wlen_planck_destroy(p); return HTGX_MEM_ERR; }
    for(i = 0; i < nbands; ++i) {
      double lo = range[0] + (double)i * p->band_len;
      double hi = fmin(lo + p->band_len, range[1]);
      p->pdf[i] = oracle_blackbody_fraction(lo*1e-9, hi*1e-9, Tref);
      sum += p->pdf[i];
    }
    for(i = 0; i < nbands; ++i) {
      p->pdf[i] /= sum;
      p->cdf[i] = i ? p->pdf[i] + p->cdf[i-1] : p->pdf[i];
    }
    p->cdf[nbands-1] = 1.0;
  }
  *out = p;
  return HTGX_OK;
}

void
oracle_wlen_planck_destroy(struct oracle_wlen_planck* p)
{
  if(!p) return;
  free(p->pdf);
  free(p->cdf);
  free(p);
}

/*******************************************************************************
 * CIE 1931 XYZ wavelength sampler: src/core/htrdr_ran_wlen_cie_xyz.c
 ******************************************************************************/
/* The colour matching functions as fitted by Wyman, Sloan and Shirley
 * (JCGT 2(2), 2013), :80-107 */
static double
cie_fit(int channel, double l)
{
  double a, b, c;
  switch(channel) {
    case 0:
      a = (l - 442.0) * (l < 442.0 ? 0.0624 : 0.0374);
      b = (l - 599.8) * (l < 599.8 ? 0.0264 : 0.0323);
      c = (l - 501.1) * (l < 501.1 ? 0.0490 : 0.0382);
      return 0.362*exp(-0.5*a*a) + 1.056*exp(-0.5*b*b) - 0.065*exp(-0.5*c*c);
    case 1:
      a = (l - 568.8) * (l < 568.8 ? 0.0213 : 0.0247);
      b = (l - 530.9) * (l < 530.9 ? 0.0613 : 0.0322);
      return 0.821*exp(-0.5*a*a) + 0.286*exp(-0.5*b*b);
    default:
      a = (l - 437.0) * (l < 437.0 ? 0.0845 : 0.0278);
      b = (l - 459.0) * (l < 459.0 ? 0.0385 : 0.0725);
      return 1.217*exp(-0.5*a*a) + 0.681*exp(-0.5*b*b);
  }
}

double
oracle_cie_fit(int channel, double lambda_nm)
{
  return cie_fit(channel, lambda_nm);
}

/* trapezoidal_integration, :55-78: (hi - lo) + 1 sub-intervals */
static double
cie_integral(int channel, double lo, double hi)
{
  const size_t n = (size_t)(hi - lo) + 1;
  const double dl = (hi - lo) / (double)n;
  double sum = 0;
  size_t i;
  for(i = 0; i < n; ++i) {
    const double f1 = cie_fit(channel, lo + dl*(double)i);
    const double f2 = cie_fit(channel, lo + dl*(double)(i+1));
    sum += (f1 + f2)*dl*0.5;
  }
  return sum;
}

/* htrdr_ran_wlen_cie_xyz_create + setup_cie_xyz, :183-329, for one channel */
int
oracle_wlen_cie_create
  (const double range[2], size_t nbands, int channel, struct oracle_wlen_planck** out)
{
  struct oracle_wlen_planck* p;
  double sum = 0;
  size_t i;
  if(!nbands || channel < 0 || channel > 2 || range[0] < 380 || range[1] > 780
  || !(range[0] < range[1]))
    return HTGX_BAD_ARG;
  p = calloc(1, sizeof(*p));
  if(!p) return HTGX_MEM_ERR;
  p->range[0] = range[0];
  p->range[1] = range[1];
  p->nbands = nbands;
  p->cie = channel;
  p->band_len = (range[1] - range[0]) / (double)nbands;
  p->pdf = malloc(nbands * sizeof(double));
  p->cdf = malloc(nbands * sizeof(double));
  if(!p->pdf || !p->cdf) { oracle_wlen_planck_destroy(p); return HTGX_MEM_ERR; }
  for(i = 0; i < nbands; ++i) {
    const double lo = range[0] + (double)i * p->band_len;
    const double hi = fmin(lo + p->band_len, range[1]);
    p->pdf[i] = cie_integral(channel, lo, hi);
    sum += p->pdf[i];
  }
  p->cie_rcp = 1.0 / sum;
  for(i = 0; i < nbands; ++i) {
    p->pdf[i] /= sum;
    p->cdf[i] = i ? p->pdf[i] + p->cdf[i-1] : p->pdf[i];
  }
  p->cdf[nbands-1] = 1.0;
  *out = p;
  return HTGX_OK;
}

/* sample_cie_xyz, :109-181: the band whose cumulative first exceeds r0, then
 * the root in the band of the quadratic CDF of the linear density through
 * the fit's values at the band ends; the band middle if no root is in it */
static double
cie_sample(const struct oracle_wlen_planck* p, double r0, double r1)
{
  size_t lo = 0, hi = p->nbands - 1;
  double lmin, lmax, fmin, fmax, a, b, c, d, delta, sq, l1, l2;
  while(lo < hi) {
    const size_t mid = (lo + hi) / 2;
    if(p->cdf[mid] > r0) hi = mid; else lo = mid + 1;
  }
  lmin = p->range[0] + p->band_len * (double)lo;
  lmax = lmin + p->band_len;
  fmin = cie_fit(p->cie, lmin);
  fmax = cie_fit(p->cie, lmax);
  a = 0.5 * (fmax - fmin) / p->band_len;
  b = (lmax * fmin - lmin * fmax) / p->band_len;
  c = -lmin * fmin + lmin*lmin * a;
  d = 0.5 * (fmax + fmin) * p->band_len;
  delta = b*b - 4*a*(c - d*r1);
  if(delta < 0 && fabs(delta) <= 1.e-6) delta = 0;
  sq = sqrt(delta);
  l1 = (-b - sq) / (2*a);
  l2 = (-b + sq) / (2*a);
  if(lmin <= l1 && l1 < lmax) return l1;
  if(lmin <= l2 && l2 < lmax) return l2;
  return (lmin + lmax) * 0.5;
}

/* Bisection on the black-body fraction, src/core/htrdr_ran_wlen_planck.c:136-216 */
static double
wlen_sample_continue(const struct oracle_wlen_planck* p, double r, double* pdf)
{
  const double rm0 = p->range[0] * 1e-9, rm1 = p->range[1] * 1e-9;
  double lmin = rm0, lmax = rm1, l = 0, lprev = 0, bf = 0, bfprev = 0;
  const double bfmm = oracle_blackbody_fraction(rm0, rm1, p->Tref);
  size_t i;
  for(i = 0; i < 100; ++i) {
    l = (lmin + lmax) * 0.5;
    bf = oracle_blackbody_fraction(rm0, l, p->Tref);
    if(bf / bfmm < r) lmin = l; else lmax = l;
    if(fabs(lprev - l) < 1e-15 || fabs(bfprev - bf) < 1e-6) break;
    lprev = l;
    bfprev = bf;
  }
  if(pdf) {
    const double Bl = oracle_planck(l, l, p->Tref);
    const double Bm = oracle_planck(rm0, rm1, p->Tref);
    *pdf = Bl / (Bm * (rm1 - rm0)) * 1e-9;
  }
  return l * 1e9;
}

/* src/core/htrdr_ran_wlen_planck.c:218-268 and :359-376 */
double
oracle_wlen_planck_sample
  (const struct oracle_wlen_planck* p, double r0, double r1, double* pdf)
{
  if(p->cie >= 0) {
    if(pdf) *pdf = p->cie_rcp;
    return cie_sample(p, r0, r1);
  }
  if(p->nbands) {
    /* first band whose cumulative is strictly greater than r0 */
    size_t lo = 0, hi = p->nbands - 1;
    double blo;
    while(lo < hi) {
      const size_t mid = (lo + hi) / 2;
      if(p->cdf[mid] > r0) hi = mid; else lo = mid + 1;
    }
    blo = p->range[0] + (double)lo * p->band_len;
    if(pdf) *pdf = p->pdf[lo] * (1.0 / p->band_len);
    return blo + p->band_len * r1;
  } else if(fabs(p->range[0] - p->range[1]) <= 1e-6) {
    if(pdf) *pdf = 1;
    return p->range[0];
  } else {
    return wlen_sample_continue(p, r0, pdf);
  }
}

/*******************************************************************************
 * Sky: gas layers + cloud octree (the property model and acceleration
 * structure htsky builds at src/atmosphere/htrdr_atmosphere.c:366-382).
 ******************************************************************************/
struct onode {
  float mn, mx;         /* exact FP32 min/max of the LWC over the node */
  uint8_t level;
  uint8_t leaf;
  uint32_t x, y, z;     /* node coordinates at its level */
  struct onode* child[8];
};

struct obvh;

struct oracle_sky {
  struct htgx_scene_desc scene; /* shallow copy (arrays owned by caller) */
  uint32_t* quad_offset;       /* [nbands] first quadrature point of a band */
  double* quad_cdf;            /* [sum nquads] */
  /* clouds */
  int has_clouds;
  uint32_t B;                  /* leaf block */
  uint32_t nlevels;
  uint32_t dims[32][3];        /* #nodes per axis per level */
  uint64_t level_offset[32];
  double cs[3], inv_cs[3];     /* cell size and inverse */
  double bs[3], inv_bs[3];     /* block size and inverse */
  double cext_max;
  double period[2];            /* -r repetition period along X, Y */
  double glo[3], ghi[3];       /* bounding box of the mesh ground (-R period) */
  struct onode* root;
  uint64_t nnodes_alloc, nleaves;
  /* optional BVH over the mesh ground, for timing runs only (the casts it
   * gives are those of the brute force, bit for bit) */
  struct obvh* bvh;
  uint32_t nbvh;
  uint32_t* bvh_tris;
  /* gas */
  double inv_dz_layer;
  /* sun */
  double sun_dir[3];
  double sun_cos_half;
  double sun_solid_angle;
  /* spectral sampler per (range) is created per draw_map call */
};

static int
cell_exists(const struct oracle_sky* s, int64_t x, int64_t y, int64_t z)
{
  return x >= 0 && y >= 0 && z >= 0
      && x < s->scene.clouds.def[0] && y < s->scene.clouds.def[1]
      && z < s->scene.clouds.def[2];
}

/* Recursive octree build. A level-0 node is a leaf block of BxBxB cells; a
 * node of a higher level merges into a leaf when all its children are leaves
 * and cext_max * (LWC_max - LWC_min) * edge <= T, with edge the largest
 * nominal edge of the node. This is the k_min/k_max merge of htsky driven by
 * the optical thickness threshold -T (doc/htrdr-atmosphere.1.in:226-228,
 * default src/../config.mk:52). */
static struct onode*
build_node(struct oracle_sky* s, uint32_t level, uint32_t x, uint32_t y, uint32_t z)
{
  struct onode* n = calloc(1, sizeof(*n));
  if(!n) return NULL;
  n->level = (uint8_t)level;
  n->x = x; n->y = y; n->z = z;
  s->nnodes_alloc++;
  if(level == 0) {
    const uint32_t B = s->B;
    uint32_t i, j, k;
    float mn = FLT_MAX, mx = -FLT_MAX;
    for(k = z*B; k < (z+1)*B; ++k)
    for(j = y*B; j < (y+1)*B; ++j)
    for(i = x*B; i < (x+1)*B; ++i) {
      const float v = s->scene.clouds.lwc
        [((size_t)k*s->scene.clouds.def[1] + j)*s->scene.clouds.def[0] + i];
      if(v < mn) mn = v;
      if(v > mx) mx = v;
    }
    n->mn = mn; n->mx = mx;
    n->leaf = 1;
  } else {
    int c, all_leaves = 1;
    float mn = FLT_MAX, mx = -FLT_MAX;
    double edge = 0;
    for(c = 0; c < 8; ++c) {
      const uint32_t cx = 2*x + (uint32_t)(c & 1);
      const uint32_t cy = 2*y + (uint32_t)((c >> 1) & 1);
      const uint32_t cz = 2*z + (uint32_t)((c >> 2) & 1);
      if(cx >= s->dims[level-1][0] || cy >= s->dims[level-1][1]
      || cz >= s->dims[level-1][2]) continue;
      n->child[c] = build_node(s, level-1, cx, cy, cz);
      if(!n->child[c]) return NULL;
      if(n->child[c]->mn < mn) mn = n->child[c]->mn;
      if(n->child[c]->mx > mx) mx = n->child[c]->mx;
      all_leaves &= n->child[c]->leaf;
    }
    n->mn = mn; n->mx = mx;
    {
      int a;
      for(a = 0; a < 3; ++a) {
        const double e = s->bs[a] * (double)(1u << level);
        if(e > edge) edge = e;
      }
    }
    if(all_leaves
    && s->cext_max * ((double)mx - (double)mn) * edge
       <= s->scene.optical_thickness) {
      n->leaf = 1;
      for(c = 0; c < 8; ++c) { free(n->child[c]); n->child[c] = NULL; }
    }
  }
  return n;
}

static void
free_node(struct onode* n)
{
  int c;
  if(!n) return;
  for(c = 0; c < 8; ++c) free_node(n->child[c]);
  free(n);
}

static void
count_leaves(const struct onode* n, uint64_t* nleaves)
{
  int c;
  if(n->leaf) { (*nleaves)++; return; }
  for(c = 0; c < 8; ++c) if(n->child[c]) count_leaves(n->child[c], nleaves);
}

static void
sun_setup(struct oracle_sky* s)
{
  /* spherical_to_cartesian_dir, src/atmosphere/htrdr_atmosphere.c:89-111;
   * cone constants, src/atmosphere/htrdr_atmosphere_sun.c:84-87 */
  const double az = s->scene.sun_azimuth * PI / 180.0;
  const double el = s->scene.sun_elevation * PI / 180.0;
  s->sun_dir[0] = cos(el) * cos(az);
  s->sun_dir[1] = cos(el) * sin(az);
  s->sun_dir[2] = sin(el);
  s->sun_cos_half = cos(4.6524e-3);
  s->sun_solid_angle = 2.0 * PI * (1.0 - s->sun_cos_half);
}

int
oracle_sky_create(const struct htgx_scene_desc* scene, struct oracle_sky** out)
{
  struct oracle_sky* s;
  uint32_t b, nq = 0;
  if(!scene || !out) return HTGX_BAD_ARG;
  s = calloc(1, sizeof(*s));
  if(!s) return HTGX_MEM_ERR;
  s->scene = *scene;
  s->quad_offset = calloc(scene->spectral.nbands + 1, sizeof(uint32_t));
  for(b = 0; b < scene->spectral.nbands; ++b) {
    s->quad_offset[b] = nq;
    nq += scene->spectral.nquads[b];
  }
  s->quad_offset[scene->spectral.nbands] = nq;
  s->quad_cdf = calloc(nq ? nq : 1, sizeof(double));
  for(b = 0; b < scene->spectral.nbands; ++b) {
    uint32_t q;
    double acc = 0;
    for(q = 0; q < scene->spectral.nquads[b]; ++q) {
      acc += scene->spectral.quad_weights[s->quad_offset[b] + q];
      s->quad_cdf[s->quad_offset[b] + q] = acc;
    }
    s->quad_cdf[s->quad_offset[b] + scene->spectral.nquads[b] - 1] = 1.0;
  }
  s->inv_dz_layer = 1.0 / scene->gas.layer_thickness;
  sun_setup(s);
  if(scene->ground.type == HTGX_GROUND_MESH) {
    /* FP32 bounds of the vertices of the triangles */
    uint32_t t;
    int a, k;
    for(a = 0; a < 3; ++a) { s->glo[a] = FLT_MAX; s->ghi[a] = -FLT_MAX; }
    for(t = 0; t < scene->ground.ntris; ++t)
      for(k = 0; k < 3; ++k)
        for(a = 0; a < 3; ++a) {
          const double v = scene->ground.verts[3 * (size_t)scene->ground.tris[3 * (size_t)t + k] + a];
          if(v < s->glo[a]) s->glo[a] = v;
          if(v > s->ghi[a]) s->ghi[a] = v;
        }
  }

  s->has_clouds = scene->clouds.def[0] && scene->clouds.def[1]
               && scene->clouds.def[2] && scene->clouds.lwc;
  if(s->has_clouds) {
    int a;
    uint32_t l;
    s->B = scene->leaf_block ? scene->leaf_block : 1;
    for(a = 0; a < 3; ++a) {
      if(scene->clouds.def[a] % s->B) { oracle_sky_destroy(s); return HTGX_BAD_ARG; }
      s->cs[a] = (scene->clouds.upper[a] - scene->clouds.lower[a])
               / (double)scene->clouds.def[a];
      s->inv_cs[a] = 1.0 / s->cs[a];
      s->bs[a] = s->cs[a] * (double)s->B;
      s->inv_bs[a] = 1.0 / s->bs[a];
      s->dims[0][a] = scene->clouds.def[a] / s->B;
    }
    s->nlevels = 1;
    s->level_offset[0] = 0;
    while(s->dims[s->nlevels-1][0] > 1 || s->dims[s->nlevels-1][1] > 1
       || s->dims[s->nlevels-1][2] > 1) {
      l = s->nlevels;
      for(a = 0; a < 3; ++a) s->dims[l][a] = (s->dims[l-1][a] + 1) / 2;
      s->level_offset[l] = s->level_offset[l-1]
        + (uint64_t)s->dims[l-1][0] * s->dims[l-1][1] * s->dims[l-1][2];
      s->nlevels++;
    }
    s->period[0] = (double)s->dims[0][0] * s->bs[0];
    s->period[1] = (double)s->dims[0][1] * s->bs[1];
    s->cext_max = 0;
    for(b = 0; b < scene->spectral.nbands; ++b)
      if(scene->clouds.cext[b] > s->cext_max) s->cext_max = scene->clouds.cext[b];
    s->root = build_node(s, s->nlevels - 1, 0, 0, 0);
    if(!s->root) { oracle_sky_destroy(s); return HTGX_MEM_ERR; }
    count_leaves(s->root, &s->nleaves);
  }
  *out = s;
  return HTGX_OK;
}

void
oracle_sky_destroy(struct oracle_sky* s)
{
  if(!s) return;
  free(s->bvh);
  free(s->bvh_tris);
  free_node(s->root);
  free(s->quad_offset);
  free(s->quad_cdf);
  free(s);
}

int
oracle_sky_get_svo_info
  (const struct oracle_sky* s, uint32_t* nlevels, uint64_t* nnodes, uint64_t* nleaves)
{
  if(!s) return HTGX_BAD_ARG;
  if(nlevels) *nlevels = s->has_clouds ? s->nlevels : 0;
  if(nnodes) *nnodes = s->has_clouds
    ? s->level_offset[s->nlevels-1] + 1 : 0;
  if(nleaves) *nleaves = s->nleaves;
  return HTGX_OK;
}

static uint64_t
leaf_id(const struct oracle_sky* s, const struct onode* n)
{
  const uint32_t* d = s->dims[n->level];
  return s->level_offset[n->level]
       + ((uint64_t)n->z * d[1] + n->y) * d[0] + n->x;
}

int
oracle_leaf_of_cell
  (const struct oracle_sky* s, const uint32_t cell[3], int64_t* leaf,
   float* mn, float* mx)
{
  const struct onode* n;
  uint32_t blk[3];
  int a;
  if(!s || !s->has_clouds) return HTGX_BAD_ARG;
  if(!cell_exists(s, cell[0], cell[1], cell[2])) return HTGX_BAD_ARG;
  for(a = 0; a < 3; ++a) blk[a] = cell[a] / s->B;
  n = s->root;
  while(!n->leaf) {
    const uint32_t l = (uint32_t)n->level - 1;
    const int c = (int)(((blk[0] >> l) & 1) | (((blk[1] >> l) & 1) << 1)
                | (((blk[2] >> l) & 1) << 2));
    n = n->child[c];
  }
  *leaf = (int64_t)leaf_id(s, n);
  *mn = n->mn;
  *mx = n->mx;
  return HTGX_OK;
}

/* Node box in block units along each axis */
static void
node_range(const struct oracle_sky* s, const struct onode* n, uint32_t lo[3], uint32_t hi[3])
{
  int a;
  const uint32_t c[3] = {n->x, n->y, n->z};
  for(a = 0; a < 3; ++a) {
    const uint32_t sz = 1u << n->level;
    lo[a] = c[a] * sz;
    hi[a] = (c[a] + 1) * sz;
    if(hi[a] > s->dims[0][a]) hi[a] = s->dims[0][a];
  }
}

/* World coordinate of the block plane k along axis a */
static double
plane(const struct oracle_sky* s, int a, uint32_t k)
{
  return s->scene.clouds.lower[a] + (double)k * s->bs[a];
}

struct ray {
  double o[3], d[3], inv[3];
  double range[2];
};

/* Cap on the copies of a repeated cloud grid one ray crosses, as the GPU
 * iterator's (a horizontal ray inside the cloud layer never leaves it) */
#define ORACLE_MAX_TILES (1 << 16)

/* Slab test of a box given by block planes */
static void
slab(const struct oracle_sky* s, const struct ray* r, const uint32_t lo[3],
     const uint32_t hi[3], double* tn, double* tf)
{
  int a;
  *tn = -INFINITY;
  *tf = INFINITY;
  for(a = 0; a < 3; ++a) {
    const double plo = plane(s, a, lo[a]);
    const double phi = plane(s, a, hi[a]);
    if(r->d[a] != 0) {
      const double t0 = (plo - r->o[a]) * r->inv[a];
      const double t1 = (phi - r->o[a]) * r->inv[a];
      const double n = t0 < t1 ? t0 : t1;
      const double f = t0 < t1 ? t1 : t0;
      if(n > *tn) *tn = n;
      if(f < *tf) *tf = f;
    } else if(r->o[a] < plo || r->o[a] > phi) {
      *tn = INFINITY;
      *tf = -INFINITY;
    }
  }
}

struct leaf_hit {
  double tf;        /* leaf exit distance (unclipped) */
  int64_t id;
  float mn, mx;
  uint32_t lo[3], hi[3]; /* block range */
  double shift[2];  /* X/Y offset of the leaf's copy of the grid (-r) */
};

struct leaf_list {
  struct leaf_hit* v;
  size_t n, cap;
};

static void
leaf_list_push(struct leaf_list* l, const struct leaf_hit* h)
{
  if(l->n == l->cap) {
    l->cap = l->cap ? l->cap * 2 : 64;
    l->v = realloc(l->v, l->cap * sizeof(*l->v));
  }
  l->v[l->n++] = *h;
}

/* Recursive front-to-back traversal: the leaves the ray crosses with a
 * positive length inside its range (svx_octree_trace_ray semantics as used by
 * htsky_trace_ray, call sites ..._sw.c:243-244,369-370, ..._lw.c:205-206). */
static void
traverse(const struct oracle_sky* s, const struct onode* n, const struct ray* r,
         struct leaf_list* out)
{
  uint32_t lo[3], hi[3];
  double tn, tf;
  node_range(s, n, lo, hi);
  slab(s, r, lo, hi, &tn, &tf);
  if(fmax(tn, r->range[0]) >= fmin(tf, r->range[1])) return;
  if(n->leaf) {
    struct leaf_hit h;
    int a;
    h.shift[0] = h.shift[1] = 0;
    h.tf = tf;
    h.id = (int64_t)leaf_id(s, n);
    h.mn = n->mn;
    h.mx = n->mx;
    for(a = 0; a < 3; ++a) { h.lo[a] = lo[a]; h.hi[a] = hi[a]; }
    leaf_list_push(out, &h);
  } else {
    const struct onode* kids[8];
    double ktn[8];
    int nk = 0, c, i, j;
    for(c = 0; c < 8; ++c) {
      uint32_t clo[3], chi[3];
      double ctn, ctf;
      if(!n->child[c]) continue;
      node_range(s, n->child[c], clo, chi);
      slab(s, r, clo, chi, &ctn, &ctf);
      if(fmax(ctn, r->range[0]) >= fmin(ctf, r->range[1])) continue;
      kids[nk] = n->child[c];
      ktn[nk] = ctn;
      nk++;
    }
    /* insertion sort on the entry distance */
    for(i = 1; i < nk; ++i) {
      for(j = i; j > 0 && ktn[j-1] > ktn[j]; --j) {
        const double t = ktn[j]; const struct onode* k = kids[j];
        ktn[j] = ktn[j-1]; kids[j] = kids[j-1];
        ktn[j-1] = t; kids[j-1] = k;
      }
    }
    for(i = 0; i < nk; ++i) traverse(s, kids[i], r, out);
  }
}

/* Gas and particle coefficients of a property (htsky_fetch_raw_property and
 * htsky_fetch_svx_voxel_property, call sites ..._sw.c:87-88,130-131,163-166,
 * 202-203,427-430, ..._lw.c:78-79,113-116). The cloud part is
 * c(prop, band) * LWC with c = cext*(1-ssa), cext*ssa or cext. */
static double
gas_k(const struct oracle_sky* s, int prop, uint32_t iband, uint32_t iquad, int32_t j)
{
  const uint32_t nl = s->scene.gas.nlayers;
  const double ka = s->scene.gas.ka[(size_t)(s->quad_offset[iband] + iquad) * nl + (uint32_t)j];
  const double ks = s->scene.gas.ks[(size_t)iband * nl + (uint32_t)j];
  switch(prop) {
    case HTGX_Ka: return ka;
    case HTGX_Ks: return ks;
    default: return ka + ks;
  }
}

static double
cloud_c(const struct oracle_sky* s, int prop, uint32_t iband)
{
  const double ce = s->scene.clouds.cext[iband];
  const double w = s->scene.clouds.ssa[iband];
  switch(prop) {
    case HTGX_Ka: return ce * (1.0 - w);
    case HTGX_Ks: return ce * w;
    default: return ce;
  }
}

/* A segment of the overlay of the octree leaves with the gas layers */
struct oseg {
  double t0, t1, kmin, kmax;
  int in_cloud;
  struct leaf_hit leaf;
  int32_t layer;  /* -1 if no gas */
};

typedef int (*seg_cb)(const struct oseg* seg, void* ctx); /* 0 = stop */

/* Enumerate the segments of [range[0], range[1]] along the ray, front to
 * back, calling cb on each segment of positive length that holds some medium.
 * Returns 1 if cb stopped the traversal. */
static int
for_each_segment
  (const struct oracle_sky* s, const double org[3], const double dir[3],
   const double range[2], int prop, uint32_t iband, uint32_t iquad,
   seg_cb cb, void* ctx)
{
  struct ray r;
  struct leaf_list leaves = {NULL, 0, 0};
  double tbox = INFINITY; /* entry distance into the cloud grid */
  size_t ileaf = 0;
  int cloud_state = 2; /* 0 before the grid, 1 inside, 2 after / no clouds */
  const uint32_t nl = s->scene.gas.nlayers;
  const double zb = s->scene.gas.z_bottom, dzl = s->scene.gas.layer_thickness;
  int32_t j;
  double t, tmax;
  int a, stopped = 0;

  for(a = 0; a < 3; ++a) {
    r.o[a] = org[a];
    r.d[a] = dir[a];
    r.inv[a] = 1.0 / dir[a];
  }
  r.range[0] = range[0];
  r.range[1] = range[1];
  t = range[0];
  tmax = range[1];

  if(s->has_clouds && !s->scene.repeat_clouds) {
    uint32_t lo[3] = {0,0,0};
    double tn, tf;
    slab(s, &r, lo, s->dims[0], &tn, &tf);
    if(fmax(tn, t) < fmin(tf, tmax)) {
      traverse(s, s->root, &r, &leaves);
      if(leaves.n) {
        if(tn > t) { cloud_state = 0; tbox = tn; } else { cloud_state = 1; }
      }
    }
  } else if(s->has_clouds) {
    /* -r: the grid repeats along X and Y with the period of its extent, an
     * infinite layer of copies crossed one after the other, each traversed
     * with the ray origin relative to it (the slab walk of
     * src/core/htrdr_slab.c:35-137, here stepping to the next copy across
     * the face the ray leaves the current one by) */
    double tn = -INFINITY, tf = INFINITY;
    const double plo = plane(s, 2, 0), phi = plane(s, 2, s->dims[0][2]);
    if(dir[2] != 0) {
      const double ta = (plo - org[2]) * r.inv[2], tb = (phi - org[2]) * r.inv[2];
      tn = ta < tb ? ta : tb;
      tf = ta < tb ? tb : ta;
    } else if(org[2] < plo || org[2] > phi) {
      tn = INFINITY;
      tf = -INFINITY;
    }
    if(fmax(tn, t) < fmin(tf, tmax)) {
      const double te = fmax(tn, t);
      int64_t k[2], ntiles = 0;
      for(a = 0; a < 2; ++a)
        k[a] = (int64_t)floor((org[a] + te * dir[a] - s->scene.clouds.lower[a]) / s->period[a]);
      for(;;) {
        struct ray rl = r;
        const size_t first = leaves.n;
        double sh[2], tx[3];
        size_t i;
        int ax;
        for(a = 0; a < 2; ++a) {
          sh[a] = (double)k[a] * s->period[a];
          rl.o[a] = org[a] - sh[a];
        }
        traverse(s, s->root, &rl, &leaves);
        for(i = first; i < leaves.n; ++i) {
          leaves.v[i].shift[0] = sh[0];
          leaves.v[i].shift[1] = sh[1];
        }
        /* the face the ray leaves this copy by */
        for(a = 0; a < 3; ++a) {
          tx[a] = dir[a] == 0 ? INFINITY
            : (plane(s, a, dir[a] > 0 ? s->dims[0][a] : 0) - rl.o[a]) * r.inv[a];
        }
        ax = tx[0] < tx[1] ? (tx[0] < tx[2] ? 0 : 2) : (tx[1] < tx[2] ? 1 : 2);
        if(ax == 2 || tx[ax] >= tmax || ++ntiles > ORACLE_MAX_TILES) break;
        k[ax] += dir[ax] > 0 ? 1 : -1;
      }
      if(leaves.n) {
        if(tn > t) { cloud_state = 0; tbox = tn; } else { cloud_state = 1; }
      }
    }
  }

  /* gas layer of the starting point */
  {
    const double z = org[2] + t * dir[2];
    const double fj = floor((z - zb) * s->inv_dz_layer);
    j = fj < -1.0 ? -1 : fj > (double)nl ? (int32_t)nl : (int32_t)fj;
  }

  while(t < tmax) {
    double tc, tg, t1;
    /* next cloud boundary */
    if(cloud_state == 0) tc = tbox;
    else if(cloud_state == 1) tc = fmax(leaves.v[ileaf].tf, t);
    else tc = INFINITY;
    /* next gas boundary */
    if(dir[2] > 0 && j < (int32_t)nl) {
      tg = (zb + (double)(j + 1) * dzl - org[2]) * r.inv[2];
    } else if(dir[2] < 0 && j >= 0) {
      tg = (zb + (double)j * dzl - org[2]) * r.inv[2];
    } else {
      tg = INFINITY;
    }
    tg = fmax(tg, t);
    t1 = fmin(fmin(tc, tg), tmax);

    if(t1 > t) {
      const int in_gas = j >= 0 && j < (int32_t)nl;
      const int in_cloud = cloud_state == 1;
      if(in_gas || in_cloud) {
        struct oseg seg;
        const double g = in_gas ? gas_k(s, prop, iband, iquad, j) : 0.0;
        const double c = in_cloud ? cloud_c(s, prop, iband) : 0.0;
        seg.t0 = t;
        seg.t1 = t1;
        seg.in_cloud = in_cloud;
        seg.layer = in_gas ? j : -1;
        if(in_cloud) {
          seg.leaf = leaves.v[ileaf];
          seg.kmin = g + c * (double)seg.leaf.mn;
          seg.kmax = g + c * (double)seg.leaf.mx;
        } else {
          memset(&seg.leaf, 0, sizeof(seg.leaf));
          seg.leaf.id = -1;
          seg.kmin = g;
          seg.kmax = g;
        }
        if(!cb(&seg, ctx)) { stopped = 1; break; }
      }
    }
    if(t1 >= tmax) break;
    t = t1;
    if(tg == t1) j += dir[2] > 0 ? 1 : -1;
    if(tc == t1) {
      if(cloud_state == 0) cloud_state = 1;
      else if(cloud_state == 1) {
        ileaf++;
        if(ileaf >= leaves.n) cloud_state = 2;
      }
    }
    if(tg == INFINITY && tc == INFINITY) break;
  }
  free(leaves.v);
  return stopped;
}

struct collect_ctx {
  struct htgx_segment* segs;
  uint32_t max, n;
};

static int
collect_cb(const struct oseg* seg, void* p)
{
  struct collect_ctx* c = p;
  if(c->n < c->max) {
    struct htgx_segment* o = c->segs + c->n;
    o->t0 = seg->t0;
    o->t1 = seg->t1;
    o->kmin = seg->kmin;
    o->kmax = seg->kmax;
    o->leaf = seg->in_cloud ? seg->leaf.id : -1;
    o->layer = seg->layer;
    o->pad = 0;
  }
  c->n++;
  return 1;
}

int
oracle_trace_segments
  (const struct oracle_sky* s, const double ray[8], int prop, uint32_t iband,
   uint32_t iquad, struct htgx_segment* segs, uint32_t max_segs, uint32_t* nsegs)
{
  struct collect_ctx c;
  c.segs = segs;
  c.max = max_segs;
  c.n = 0;
  for_each_segment(s, ray, ray + 3, ray + 6, prop, iband, iquad, collect_cb, &c);
  *nsegs = c.n;
  return HTGX_OK;
}

/* Property at a point inside a segment. The gas layer is the segment's and the
 * cloud cell is clamped into the segment's leaf so that k lies in
 * [kmin, kmax] (the hint of htsky_fetch_raw_property, ..._sw.c:202-204). */
static double
raw_in_seg(const struct oracle_sky* s, const struct oseg* seg, int prop,
           int cpnt, uint32_t iband, uint32_t iquad, const double x[3])
{
  double g = 0.0, c = 0.0;
  if((cpnt & 1) && seg->layer >= 0) g = gas_k(s, prop, iband, iquad, seg->layer);
  if((cpnt & 2) && seg->in_cloud) {
    int a;
    int64_t ci[3];
    float v;
    for(a = 0; a < 3; ++a) {
      /* relative to the segment's copy of the grid (-r) */
      const double xa = a < 2 ? x[a] - seg->leaf.shift[a] : x[a];
      const double f = floor((xa - s->scene.clouds.lower[a]) * s->inv_cs[a]);
      const int64_t lo = (int64_t)seg->leaf.lo[a] * s->B;
      const int64_t hi = (int64_t)seg->leaf.hi[a] * s->B - 1;
      int64_t i = f < (double)lo ? lo : f > (double)hi ? hi : (int64_t)f;
      ci[a] = i;
    }
    v = s->scene.clouds.lwc
      [((size_t)ci[2]*s->scene.clouds.def[1] + (size_t)ci[1])*s->scene.clouds.def[0]
       + (size_t)ci[0]];
    c = cloud_c(s, prop, iband) * (double)v;
  }
  return g + c;
}

double
oracle_fetch_raw
  (const struct oracle_sky* s, int prop, int cpnt, uint32_t iband, uint32_t iquad,
   const double pos[3])
{
  /* Standalone lookup (no segment hint): layer and cell from the position */
  double g = 0.0, c = 0.0;
  const double fj = floor((pos[2] - s->scene.gas.z_bottom) * s->inv_dz_layer);
  if((cpnt & 1) && fj >= 0 && fj < (double)s->scene.gas.nlayers)
    g = gas_k(s, prop, iband, iquad, (int32_t)fj);
  if((cpnt & 2) && s->has_clouds) {
    int a, inside = 1;
    int64_t ci[3];
    for(a = 0; a < 3; ++a) {
      const double f = floor((pos[a] - s->scene.clouds.lower[a]) * s->inv_cs[a]);
      if(f < 0 || f >= (double)s->scene.clouds.def[a]) inside = 0;
      ci[a] = (int64_t)f;
    }
    if(inside) {
      const float v = s->scene.clouds.lwc
        [((size_t)ci[2]*s->scene.clouds.def[1] + (size_t)ci[1])
         * s->scene.clouds.def[0] + (size_t)ci[0]];
      c = cloud_c(s, prop, iband) * (double)v;
    }
  }
  return g + c;
}

/*******************************************************************************
 * Null-collision hit filters
 ******************************************************************************/
/* k_min/k_max transmissivity, ..._sw.c:147-251 */
struct tr_ctx {
  const struct oracle_sky* s;
  struct orng* rng;
  const double* o; const double* d;
  int prop; uint32_t iband, iquad;
  double Ts, Tmin;
  int collided;
};

static int
tr_cb(const struct oseg* seg, void* p)
{
  struct tr_ctx* c = p;
  const double kmin = seg->kmin, kmax = seg->kmax;
  double t = seg->t0;
  c->Tmin += (seg->t1 - seg->t0) * kmin;               /* ..._sw.c:169 */
  for(;;) {
    const double Tdif = (seg->t1 - t) * (kmax - kmin);  /* ..._sw.c:175-176 */
    if(c->Ts > Tdif || !(kmax > kmin)) {                /* ..._sw.c:179-183 */
      c->Ts -= Tdif;
      return 1;
    } else {
      double x[3], k, proba;
      t += c->Ts / (kmax - kmin);                       /* ..._sw.c:190-193 */
      x[0] = c->o[0] + t * c->d[0];
      x[1] = c->o[1] + t * c->d[1];
      x[2] = c->o[2] + t * c->d[2];
      k = raw_in_seg(c->s, seg, c->prop, 3, c->iband, c->iquad, x);
      proba = (k - kmin) / (kmax - kmin);               /* ..._sw.c:206 */
      if(canonical(c->rng) < proba) { c->collided = 1; return 0; }
      c->Ts = ran_exp(c->rng);                          /* ..._sw.c:212 */
    }
  }
}

static double
transmissivity
  (const struct oracle_sky* s, struct orng* rng, int prop, uint32_t iband,
   uint32_t iquad, const double pos[3], const double dir[3], const double range[2])
{
  struct tr_ctx c;
  c.s = s; c.rng = rng; c.o = pos; c.d = dir;
  c.prop = prop; c.iband = iband; c.iquad = iquad;
  c.Ts = ran_exp(rng);                                  /* ..._sw.c:239 */
  c.Tmin = 0;
  c.collided = 0;
  for_each_segment(s, pos, dir, range, prop, iband, iquad, tr_cb, &c);
  if(c.collided) return 0;                              /* ..._sw.c:246-250 */
  return c.Tmin ? exp(-c.Tmin) : 1.0;
}

double
oracle_transmissivity
  (const struct oracle_sky* s, uint64_t seed, uint64_t pixel, uint64_t realisation,
   int prop, uint32_t iband, uint32_t iquad, const double pos[3],
   const double dir[3], const double range[2])
{
  struct orng rng;
  orng_init(&rng, seed, pixel, realisation);
  return transmissivity(s, &rng, prop, iband, iquad, pos, dir, range);
}

/* Delta tracking on the Ks majorant, ..._sw.c:73-145 */
struct sc_ctx {
  const struct oracle_sky* s;
  struct orng* rng;
  const double* o; const double* d;
  uint32_t iband, iquad;
  double Ts;
  double dst;
  int hit;
  struct oseg seg; /* segment of the collision */
};

static int
sc_cb(const struct oseg* seg, void* p)
{
  struct sc_ctx* c = p;
  const double ks_max = seg->kmax;
  double t = seg->t0;
  for(;;) {
    const double T = (seg->t1 - t) * ks_max;           /* ..._sw.c:96-97 */
    if(c->Ts > T || !(ks_max > 0)) {                   /* ..._sw.c:100-104 */
      c->Ts -= T;
      return 1;
    } else {
      double x[3], ks;
      t += c->Ts / ks_max;                             /* ..._sw.c:111-114 */
      if(t > 1e9) return 1;                            /* ..._sw.c:123 */
      x[0] = c->o[0] + t * c->d[0];
      x[1] = c->o[1] + t * c->d[1];
      x[2] = c->o[2] + t * c->d[2];
      ks = raw_in_seg(c->s, seg, HTGX_Ks, 3, c->iband, c->iquad, x);
      if(canonical(c->rng) < ks / ks_max) {            /* ..._sw.c:134-138 */
        c->hit = 1;
        c->dst = t;
        c->seg = *seg;
        return 0;
      }
      c->Ts = ran_exp(c->rng);                         /* ..._sw.c:140 */
    }
  }
}

/* Kext majorant with absorption / scattering / null events, ..._lw.c:64-135 */
enum { EV_ABS, EV_SCA, EV_NONE };
struct lw_ctx {
  const struct oracle_sky* s;
  struct orng* rng;
  const double* o; const double* d;
  uint32_t iband, iquad;
  double Ts, dst;
  int event;
  struct oseg seg;
};

static int
lw_cb(const struct oseg* seg, void* p)
{
  struct lw_ctx* c = p;
  const double kext_max = seg->kmax;
  double t = seg->t0;
  for(;;) {
    const double T = (seg->t1 - t) * kext_max;          /* ..._lw.c:84-85 */
    if(c->Ts > T || !(kext_max > 0)) {
      c->Ts -= T;
      return 1;
    } else {
      double x[3], ka, ks, r, pa, ps;
      t += c->Ts / kext_max;                            /* ..._lw.c:96-104 */
      x[0] = c->o[0] + t * c->d[0];
      x[1] = c->o[1] + t * c->d[1];
      x[2] = c->o[2] + t * c->d[2];
      ka = raw_in_seg(c->s, seg, HTGX_Ka, 3, c->iband, c->iquad, x);
      ks = raw_in_seg(c->s, seg, HTGX_Ks, 3, c->iband, c->iquad, x);
      r = canonical(c->rng);                            /* ..._lw.c:118-131 */
      pa = ka / kext_max;
      ps = ks / kext_max;
      if(r < pa) { c->event = EV_ABS; c->dst = t; c->seg = *seg; return 0; }
      if(r < pa + ps) { c->event = EV_SCA; c->dst = t; c->seg = *seg; return 0; }
      c->Ts = ran_exp(c->rng);
    }
  }
}

/*******************************************************************************
 * Ground: the plane repeated to infinity (-R over a flat quad,
 * src/atmosphere/htrdr_atmosphere_ground.c:170-217 + src/core/htrdr_slab.c)
 * with the self-hit rules of src/core/htrdr_geometry.c:116-150.
 ******************************************************************************/
struct ghit {
  int hit;
  double t;
  double N[3];
  uint32_t mtl;
  int64_t prim;  /* 0 the plane, else the mesh triangle (input order); -1 none */
  int edge;      /* the hit lies on an edge or vertex of its triangle */
};

static const struct ghit GHIT_NONE = {0, FLT_MAX, {0, 0, 0}, 0, -1, 0};

/* Self-hit filter of src/core/htrdr_geometry.c:96-150 */
static int
on_edge_f(float u, float v)
{
  const float eps = 1e-4f;
  const float w = 1.f - u - v;
  return fabsf(u) <= eps || fabsf(u - 1.f) <= eps || fabsf(v) <= eps
      || fabsf(v - 1.f) <= eps || fabsf(w) <= eps || fabsf(w - 1.f) <= eps;
}

/* Moller-Trumbore in FP32 on (v0, e1 = v1 - v0, e2 = v2 - v0), double-sided */
static int
tri_hit_f(const float v0[3], const float e1[3], const float e2[3], const float o[3],
          const float d[3], float* t, float* u, float* v)
{
  const float p0 = d[1]*e2[2] - d[2]*e2[1];
  const float p1 = d[2]*e2[0] - d[0]*e2[2];
  const float p2 = d[0]*e2[1] - d[1]*e2[0];
  const float det = e1[0]*p0 + e1[1]*p1 + e1[2]*p2;
  float inv, s0, s1, s2, uu, q0, q1, q2, vv;
  if(det == 0.f) return 0;
  inv = 1.f / det;
  s0 = o[0] - v0[0]; s1 = o[1] - v0[1]; s2 = o[2] - v0[2];
  uu = (s0*p0 + s1*p1 + s2*p2) * inv;
  if(!(uu >= 0.f && uu <= 1.f)) return 0;
  q0 = s1*e1[2] - s2*e1[1];
  q1 = s2*e1[0] - s0*e1[2];
  q2 = s0*e1[1] - s1*e1[0];
  vv = (d[0]*q0 + d[1]*q1 + d[2]*q2) * inv;
  if(!(vv >= 0.f && uu + vv <= 1.f)) return 0;
  *t = (e2[0]*q0 + e2[1]*q1 + e2[2]*q2) * inv;
  *u = uu;
  *v = vv;
  return 1;
}

/* Closest ground hit in ]range[0], range[1][ leaving the previous surface
 * hit `prev' (htrdr_atmosphere_ground_trace_ray, ..._ground.c:170-217). The
 * plane is the quad repeated to infinity with -R, a single primitive; a mesh
 * is cast against every triangle in FP32, as the reference converts the ray
 * (src/core/htrdr_geometry.c:698-700). */
static struct ghit mesh_trace_once
  (const struct oracle_sky* s, const double o[3], const double d[3], const double range[2],
   int64_t pprim, int pedge);

static struct ghit
ground_trace(const struct oracle_sky* s, const double o[3], const double d[3],
             const double range[2], const struct ghit* prev)
{
  struct ghit h = GHIT_NONE;
  const int64_t pprim = prev ? prev->prim : -1;
  const int pedge = prev ? prev->edge : 0;
  const struct htgx_ground_desc* g = &s->scene.ground;
  if(g->type == HTGX_GROUND_PLANE && pprim != 0 && d[2] != 0) {
    const double t = (g->plane_z - o[2]) / d[2];
    if(t > range[0] && t < range[1]) {
      h.hit = 1;
      h.t = t;
      h.N[0] = 0; h.N[1] = 0; h.N[2] = 1;
      h.mtl = g->plane_mtl;
      h.prim = 0;
    }
  } else if(g->type == HTGX_GROUND_MESH && s->scene.repeat_ground) {
    /* -R: the copies of the mesh along X and Y, as htrdr_slab_trace_ray
     * walks them (src/core/htrdr_slab.c:35-137; called at
     * src/atmosphere/htrdr_atmosphere_ground.c:197-210 with 32 steps) */
    const double* lo = s->glo;
    const double* hi = s->ghi;
    double tz0 = (lo[2] - o[2]) / d[2], tz1 = (hi[2] - o[2]) / d[2];
    double sz[2], tmax[3], tdelta[2];
    int64_t xy[2];
    int incr[2], a, step;
    if(tz0 > tz1) { const double tt = tz0; tz0 = tz1; tz1 = tt; }
    tz0 = fmax(tz0, range[0]);
    tz1 = fmin(tz1, range[1]);
    if(tz0 > tz1) return h;
    tmax[2] = tz1;
    for(a = 0; a < 2; ++a) {
      double low;
      sz[a] = hi[a] - lo[a];
      xy[a] = (int64_t)floor((o[a] + tz0 * d[a] - lo[a]) / sz[a]);
      incr[a] = d[a] < 0 ? -1 : 1;
      low = lo[a] + (double)xy[a] * sz[a];
      tmax[a] = ((d[a] < 0 ? low : low + sz[a]) - o[a]) / d[a];
      tdelta[a] = (d[a] < 0 ? -sz[a] : sz[a]) / d[a];
    }
    for(step = 0; step < 32; ++step) {
      const double oc[3] = {o[0] - (double)xy[0] * sz[0], o[1] - (double)xy[1] * sz[1], o[2]};
      struct ghit m = mesh_trace_once(s, oc, d, range, pprim, pedge);
      if(m.hit) return m;
      a = tmax[0] < tmax[1] ? (tmax[0] < tmax[2] ? 0 : 2) : (tmax[1] < tmax[2] ? 1 : 2);
      if(a == 2 || tmax[a] >= range[1]) break;
      tmax[a] += tdelta[a];
      xy[a] += incr[a];
    }
  } else if(g->type == HTGX_GROUND_MESH) {
    return mesh_trace_once(s, o, d, range, pprim, pedge);
  }
  return h;
}

/* The best hit so far of a mesh cast: the smallest distance, the lowest
 * triangle index among equal distances (the brute force's order) */
struct cand {
  float best, bu, bv;
  int64_t bt;
};

/* Triangle t as a candidate of the cast (of, df) in FP32, with the range
 * and the self-hit rules of src/core/htrdr_geometry.c:116-150 */
static void
tri_candidate(const struct htgx_ground_desc* g, uint32_t t, const float of[3], const float df[3],
              float r0, float r1, int64_t pprim, int pedge, struct cand* c)
{
  const float* p0 = g->verts + 3*(size_t)g->tris[3*(size_t)t + 0];
  const float* p1 = g->verts + 3*(size_t)g->tris[3*(size_t)t + 1];
  const float* p2 = g->verts + 3*(size_t)g->tris[3*(size_t)t + 2];
  float e1[3], e2[3], tt, u, v;
  int a;
  for(a = 0; a < 3; ++a) { e1[a] = p1[a] - p0[a]; e2[a] = p2[a] - p0[a]; }
  if(!tri_hit_f(p0, e1, e2, of, df, &tt, &u, &v)) return;
  if(tt > c->best || (tt == c->best && (c->bt < 0 || (int64_t)t > c->bt))) return;
  if(tt <= r0 || tt >= r1) return;
  if((int64_t)t == pprim) return;
  if(pprim >= 0 && fabsf(tt) <= 1e-1f && pedge && on_edge_f(u, v)) return;
  c->best = tt; c->bu = u; c->bv = v; c->bt = (int64_t)t;
}

/* A bounding volume hierarchy over the mesh, median split on the longest
 * axis, leaves of up to 4 triangles, boxes in FP64 padded so that no FP32
 * hit can fall outside them. Visiting its leaves instead of every triangle
 * finds the same candidate (tri_candidate keeps the smallest distance, then
 * the lowest index), so the casts equal the brute force's. */
struct obvh {
  double lo[3], hi[3];
  uint32_t a, b;  /* inner: the two children; leaf: first triangle slot, count */
  int leaf;
};

static void
tri_box(const struct htgx_ground_desc* g, uint32_t t, double lo[3], double hi[3])
{
  int k, a;
  for(a = 0; a < 3; ++a) { lo[a] = INFINITY; hi[a] = -INFINITY; }
  for(k = 0; k < 3; ++k) {
    const float* p = g->verts + 3*(size_t)g->tris[3*(size_t)t + k];
    for(a = 0; a < 3; ++a) {
      if(p[a] < lo[a]) lo[a] = p[a];
      if(p[a] > hi[a]) hi[a] = p[a];
    }
  }
}

/* centroid keys of the triangles being sorted (qsort has no context) */
static const double* g_sort_key;

static int
cmp_key(const void* pa, const void* pb)
{
  const uint32_t ta = *(const uint32_t*)pa, tb = *(const uint32_t*)pb;
  const double ca = g_sort_key[ta], cb = g_sort_key[tb];
  return ca < cb ? -1 : ca > cb ? 1 : (ta < tb ? -1 : ta > tb);
}

static uint32_t
bvh_node(struct oracle_sky* s, double* key, uint32_t first, uint32_t n)
{
  const struct htgx_ground_desc* g = &s->scene.ground;
  const uint32_t id = s->nbvh++;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double ext = -1;
  uint32_t i;
  int a, axis = 0;
  for(i = first; i < first + n; ++i) {
    double tl[3], th[3];
    tri_box(g, s->bvh_tris[i], tl, th);
    for(a = 0; a < 3; ++a) {
      if(tl[a] < lo[a]) lo[a] = tl[a];
      if(th[a] > hi[a]) hi[a] = th[a];
    }
  }
  for(a = 0; a < 3; ++a) {
    const double pad = 1e-4 * (fabs(lo[a]) + fabs(hi[a])) + 1e-3;
    s->bvh[id].lo[a] = lo[a] - pad;
    s->bvh[id].hi[a] = hi[a] + pad;
    if(hi[a] - lo[a] > ext) { ext = hi[a] - lo[a]; axis = a; }
  }
  if(n <= 4) {
    s->bvh[id].leaf = 1;
    s->bvh[id].a = first;
    s->bvh[id].b = n;
    return id;
  }
  for(i = first; i < first + n; ++i) {
    double tl[3], th[3];
    tri_box(g, s->bvh_tris[i], tl, th);
    key[s->bvh_tris[i]] = tl[axis] + th[axis];
  }
  g_sort_key = key;
  qsort(s->bvh_tris + first, n, sizeof(uint32_t), cmp_key);
  s->bvh[id].leaf = 0;
  s->bvh[id].a = bvh_node(s, key, first, n / 2);
  s->bvh[id].b = bvh_node(s, key, first + n / 2, n - n / 2);
  return id;
}

int
oracle_sky_use_bvh(struct oracle_sky* s, int on)
{
  const struct htgx_ground_desc* g;
  double* key;
  uint32_t t;
  if(!s) return HTGX_BAD_ARG;
  free(s->bvh);
  free(s->bvh_tris);
  s->bvh = NULL;
  s->bvh_tris = NULL;
  s->nbvh = 0;
  g = &s->scene.ground;
  if(!on || g->type != HTGX_GROUND_MESH || !g->ntris) return HTGX_OK;
  s->bvh = malloc(2 * (size_t)g->ntris * sizeof(struct obvh));
  s->bvh_tris = malloc((size_t)g->ntris * sizeof(uint32_t));
  key = malloc((size_t)g->ntris * sizeof(double));
  if(!s->bvh || !s->bvh_tris || !key) {
    free(key);
    oracle_sky_use_bvh(s, 0);
    return HTGX_MEM_ERR;
  }
  for(t = 0; t < g->ntris; ++t) s->bvh_tris[t] = t;
  bvh_node(s, key, 0, g->ntris);
  free(key);
  return HTGX_OK;
}

/* Entry distance of the ray into a box within [0, r1], or INFINITY */
static double
box_entry(const struct obvh* nd, const double o[3], const double inv[3], double r1)
{
  double t0 = 0, t1 = r1;
  int a;
  for(a = 0; a < 3; ++a) {
    double ta = (nd->lo[a] - o[a]) * inv[a], tb = (nd->hi[a] - o[a]) * inv[a];
    if(ta > tb) { const double x = ta; ta = tb; tb = x; }
    if(ta != ta) ta = -INFINITY; /* 0 * inf: the ray runs in the slab's boundary */
    if(tb != tb) tb = INFINITY;
    if(ta > t0) t0 = ta;
    if(tb < t1) t1 = tb;
  }
  return t0 <= t1 ? t0 : INFINITY;
}

static void
bvh_trace(const struct oracle_sky* s, const double o[3], const double d[3], const float of[3],
          const float df[3], float r0, float r1, int64_t pprim, int pedge, struct cand* c)
{
  uint32_t stack[256];
  int top = 0, a;
  double inv[3];
  for(a = 0; a < 3; ++a) inv[a] = 1.0 / d[a];
  stack[top++] = 0;
  while(top > 0) {
    const struct obvh* nd = s->bvh + stack[--top];
    /* boxes beyond the best distance so far (with a margin for the FP32
     * distance) hold no better candidate */
    const double lim = c->best == INFINITY ? INFINITY : (double)c->best * (1 + 1e-5) + 1e-2;
    if(box_entry(nd, o, inv, lim) == INFINITY) continue;
    if(nd->leaf) {
      uint32_t i;
      for(i = nd->a; i < nd->a + nd->b; ++i)
        tri_candidate(&s->scene.ground, s->bvh_tris[i], of, df, r0, r1, pprim, pedge, c);
    } else if(top + 2 > 256) {
      /* deeper than any median split of 2^32 triangles: cast by brute force */
      uint32_t t;
      for(t = 0; t < s->scene.ground.ntris; ++t)
        tri_candidate(&s->scene.ground, t, of, df, r0, r1, pprim, pedge, c);
      return;
    } else {
      /* the nearer child on top of the stack */
      const double ea = box_entry(s->bvh + nd->a, o, inv, lim);
      const double eb = box_entry(s->bvh + nd->b, o, inv, lim);
      if(ea <= eb) { stack[top++] = nd->b; stack[top++] = nd->a; }
      else { stack[top++] = nd->a; stack[top++] = nd->b; }
    }
  }
}

/* Closest mesh hit by brute force over every triangle, the ray in FP32 */
static struct ghit
mesh_trace_once(const struct oracle_sky* s, const double o[3], const double d[3],
                const double range[2], int64_t pprim, int pedge)
{
  struct ghit h = GHIT_NONE;
  const struct htgx_ground_desc* g = &s->scene.ground;
  {
    const float of[3] = {(float)o[0], (float)o[1], (float)o[2]};
    const float df[3] = {(float)d[0], (float)d[1], (float)d[2]};
    const float r0 = (float)range[0];
    const float r1 = range[1] > (double)FLT_MAX ? INFINITY : (float)range[1];
    struct cand c;
    uint32_t t;
    c.best = r1; c.bu = 0; c.bv = 0; c.bt = -1;
    if(s->bvh) {
      bvh_trace(s, o, d, of, df, r0, r1, pprim, pedge, &c);
    } else {
      for(t = 0; t < g->ntris; ++t) tri_candidate(g, t, of, df, r0, r1, pprim, pedge, &c);
    }
    const float best = c.best, bu = c.bu, bv = c.bv;
    const int64_t bt = c.bt;
    if(bt >= 0) {
      const float* p0 = g->verts + 3*(size_t)g->tris[3*(size_t)bt + 0];
      const float* p1 = g->verts + 3*(size_t)g->tris[3*(size_t)bt + 1];
      const float* p2 = g->verts + 3*(size_t)g->tris[3*(size_t)bt + 2];
      float e1[3], e2[3];
      double n[3];
      int a;
      for(a = 0; a < 3; ++a) { e1[a] = p1[a] - p0[a]; e2[a] = p2[a] - p0[a]; }
      n[0] = (float)(e1[1]*e2[2] - e1[2]*e2[1]);
      n[1] = (float)(e1[2]*e2[0] - e1[0]*e2[2]);
      n[2] = (float)(e1[0]*e2[1] - e1[1]*e2[0]);
      normalize3(h.N, n);
      h.hit = 1;
      h.t = (double)best;
      h.mtl = g->tri_mtl[bt];
      h.prim = bt;
      h.edge = on_edge_f(bu, bv);
    }
  }
  return h;
}

int
oracle_ground_trace(const struct oracle_sky* s, const double ray[8], int64_t prev_prim,
                    int prev_edge, double* t, int64_t* prim, int* edge)
{
  struct ghit prev = GHIT_NONE;
  struct ghit h;
  prev.prim = prev_prim;
  prev.edge = prev_edge;
  h = ground_trace(s, ray, ray + 3, ray + 6, prev_prim >= 0 ? &prev : NULL);
  *t = h.hit ? h.t : INFINITY;
  *prim = h.hit ? h.prim : -1;
  *edge = h.edge;
  return h.hit;
}

/*******************************************************************************
 * Ground materials: spectral BRDF lists and two-sided interfaces
 ******************************************************************************/
static int
general_materials(const struct htgx_ground_desc* g)
{
  return g->mtl_brdf_first != NULL || g->interfaces != NULL;
}

/* BRDF entries of material m; m = -1 is the sky side of a surface given by a
 * lone material (an interface {m, sky}) */
static uint32_t
brdf_count(const struct htgx_ground_desc* g, int32_t m)
{
  if(m < 0) return 0;
  if(!g->mtl_brdf_first) return 1;
  return g->mtl_brdf_first[m + 1] - g->mtl_brdf_first[m];
}

static void
interface_of(const struct htgx_ground_desc* g, uint32_t i, int32_t f[3])
{
  if(g->interfaces) {
    f[0] = g->interfaces[3 * (size_t)i];
    f[1] = g->interfaces[3 * (size_t)i + 1];
    f[2] = g->interfaces[3 * (size_t)i + 2];
  } else {
    f[0] = (int32_t)i; f[1] = -1; f[2] = -1;
  }
}

/* src/core/htrdr_interface.h:45-87. N is the hit normal (v1 - v0) x (v2 - v0);
 * the front normal of the s3d hit, towards the side from which the vertices
 * turn clockwise (htrdr-obj(5)), is -N. */
static int32_t
hit_material(const struct htgx_ground_desc* g, uint32_t ifc, const double N[3], const double dir[3])
{
  const double Nf[3] = {-N[0], -N[1], -N[2]};
  int32_t f[3], m;
  int back_hit;
  interface_of(g, ifc, f);
  if(f[2] >= 0 && brdf_count(g, f[2])) return f[2];   /* thin material */
  back_hit = !(dot3(Nf, dir) < 0);                     /* :67 */
  m = back_hit ? f[0] : f[1];         /* the material on the other side */
  if(!brdf_count(g, m)) m = back_hit ? f[1] : f[0];    /* :75-84 */
  return m;
}

/* mrumtl_fetch_brdf(mrumtl, wavelength, r), src/core/htrdr_materials.c:430-433
 * (mrumtl 0.2, not vendored; restated from its documented behaviour): in a
 * band list, the band holding wlen; otherwise, between the entries i-1 and i
 * around wlen (wavelengths, or the upper end of band i-1 and the lower end
 * of band i), entry i if r < u = wlen's linear position between them, else
 * i-1; before the first entry the first, after the last the last. A grey
 * albedo is one Lambertian entry. */
static void
brdf_at(const struct htgx_ground_desc* g, int32_t m, double wlen, double r, int* type,
        double* refl)
{
  const struct htgx_brdf* e;
  uint32_t n, i;
  if(!g->mtl_brdf_first) { *type = HTGX_BRDF_LAMBERTIAN; *refl = g->mtl_albedo[m]; return; }
  e = g->brdfs + g->mtl_brdf_first[m];
  n = brdf_count(g, m);
  if(g->mtl_brdf_list[m] == HTGX_BRDF_LIST_BANDS) {
    for(i = 0; i < n && e[i].wlen[1] < wlen; ++i) {}
    if(i == n) {
      i = n - 1;
    } else if(i > 0 && wlen < e[i].wlen[0]) {
      const double u = (wlen - e[i - 1].wlen[1]) / (e[i].wlen[0] - e[i - 1].wlen[1]);
      if(!(r < u)) --i;
    }
  } else {
    for(i = 0; i < n && e[i].wlen[0] < wlen; ++i) {}
    if(i == n) {
      i = n - 1;
    } else if(i > 0) {
      const double u = (wlen - e[i - 1].wlen[0]) / (e[i].wlen[0] - e[i - 1].wlen[0]);
      if(!(r < u)) --i;
    }
  }
  *type = e[i].type;
  *refl = e[i].reflectivity;
}

static double
mtl_temperature(const struct htgx_ground_desc* g, int32_t m)
{
  return g->mtl_temperature ? g->mtl_temperature[m] : 0.0;
}

int
oracle_brdf_fetch(const struct htgx_ground_desc* g, int32_t m, double wlen, double r,
                  double* refl)
{
  int type;
  brdf_at(g, m, wlen, r, &type, refl);
  return type;
}

int32_t
oracle_hit_material(const struct htgx_ground_desc* g, uint32_t ifc, const double N[3],
                    const double dir[3])
{
  return hit_material(g, ifc, N, dir);
}

/* specular reflection of the incoming direction about N (ssf_specular_reflection:
 * wi = 2 (wo . N) N - wo with wo = -dir) */
static void
reflect_dir(const double dir[3], const double N[3], double out[3])
{
  const double c = 2.0 * dot3(dir, N);
  out[0] = dir[0] - c * N[0];
  out[1] = dir[1] - c * N[1];
  out[2] = dir[2] - c * N[2];
}

/*******************************************************************************
 * Sun, phase functions, BRDF
 ******************************************************************************/
static int
in_solar_cone(const struct oracle_sky* s, const double dir[3])
{
  return dot3(dir, s->sun_dir) >= s->sun_cos_half; /* ..._sun.c:154-166 */
}

/* ..._sun.c:126-136: uniform direction in the cone, pdf = 1/solid angle */
static double
sun_sample(const struct oracle_sky* s, struct orng* rng, double out[3])
{
  const double u0 = canonical(rng), u1 = canonical(rng);
  const double cos_t = 1.0 - u0 * (1.0 - s->sun_cos_half);
  local_to_world(s->sun_dir, cos_t, 2.0 * PI * u1, out);
  return 1.0 / s->sun_solid_angle;
}

static double
hg_cos_sample(double g, double u)
{
  if(fabs(g) < 1e-6) return 1.0 - 2.0 * u;
  {
    const double t = (1.0 - g*g) / (1.0 - g + 2.0*g*u);
    double c = (1.0 + g*g - t*t) / (2.0*g);
    return c < -1 ? -1 : c > 1 ? 1 : c;
  }
}

static double
hg_eval(double g, double mu)
{
  const double d = 1.0 + g*g - 2.0*g*mu;
  return (1.0 - g*g) / (4.0 * PI * d * sqrt(d));
}

/* p(mu) = 3/(16 pi)(1+mu^2); inverse of the CDF (mu^3+3mu+4)/8 by Cardano */
static double
rayleigh_cos_sample(double u)
{
  const double q = 4.0 * u - 2.0;  /* -q/2 of the depressed cubic */
  const double A = cbrt(q + sqrt(q*q + 1.0));
  const double c = A - 1.0 / A;
  return c < -1 ? -1 : c > 1 ? 1 : c;
}

static double
rayleigh_eval(double mu) { return 3.0 / (16.0 * PI) * (1.0 + mu*mu); }

double
oracle_phase_eval(int kind, double param, double mu)
{
  return kind == 0 ? hg_eval(param, mu) : rayleigh_eval(mu);
}

/* Cosine-weighted hemisphere around n (ssp_ran_hemisphere_cos, ..._draw_map.c:104;
 * Lambertian ssf_bsdf_sample) */
static void
cos_hemisphere(const double n[3], double u0, double u1, double out[3])
{
  local_to_world(n, sqrt(1.0 - u0), 2.0 * PI * u1, out);
}

int
oracle_sample_dirs
  (const struct oracle_sky* s, int kind, double param, const double dir_in[3],
   const double* u, uint32_t n, double* out)
{
  uint32_t i;
  for(i = 0; i < n; ++i) {
    const double u0 = u[2*i], u1 = u[2*i+1];
    double* o = out + 3*i;
    switch(kind) {
      case 0: local_to_world(dir_in, hg_cos_sample(param, u0), 2.0*PI*u1, o); break;
      case 1: local_to_world(dir_in, rayleigh_cos_sample(u0), 2.0*PI*u1, o); break;
      case 2: {
        const double cos_t = 1.0 - u0 * (1.0 - s->sun_cos_half);
        local_to_world(s->sun_dir, cos_t, 2.0*PI*u1, o);
      } break;
      case 3: cos_hemisphere(dir_in, u0, u1, o); break;
      default: return HTGX_BAD_ARG;
    }
  }
  return HTGX_OK;
}

/*******************************************************************************
 * Random walks
 ******************************************************************************/
/* ..._sw.c:256-511 */
static double
radiance_sw
  (const struct oracle_sky* s, struct orng* rng, int cpnt_mask,
   const double pos_in[3], const double dir_in[3], double wlen,
   uint32_t iband, uint32_t iquad)
{
  const double g = s->scene.clouds.hg_g ? s->scene.clouds.hg_g[iband] : 0.0;
  const double L_sun = oracle_planck_monochromatic(wlen * 1e-9, 5778.0);
  double pos[3], dir[3];
  double ksi = 1, w = 0;
  struct ghit hit_prev = GHIT_NONE; /* ..._sw.c:270 */
  memcpy(pos, pos_in, sizeof(pos));
  memcpy(dir, dir_in, sizeof(dir));

  if((cpnt_mask & HTGX_RADIANCE_DIRECT) && in_solar_cone(s, dir)) { /* :324-337 */
    const double range[2] = {0, FLT_MAX};
    const struct ghit h = ground_trace(s, pos, dir, range, NULL);
    if(!h.hit) w = L_sun * transmissivity(s, rng, HTGX_Kext, iband, iquad, pos, dir, range);
  }
  if(!(cpnt_mask & HTGX_RADIANCE_DIFFUSE)) return w;

  for(;;) {
    struct sc_ctx sc;
    struct ghit gh;
    double range[2], pos_next[3], dir_next[3], sun_dir[3], N[3] = {0,0,0};
    double Tr_abs, Tr, R, sun_pdf, refl = 1, dst, albedo = 0;
    int surface, phase_kind = 0, specular = 0;
    const double r_inf[2] = {0, DBL_MAX};

    gh = ground_trace(s, pos, dir, r_inf, &hit_prev);     /* :355-356 */
    sc.s = s; sc.rng = rng; sc.o = pos; sc.d = dir;
    sc.iband = iband; sc.iquad = iquad;
    sc.Ts = ran_exp(rng);                                 /* :359 */
    sc.hit = 0;
    sc.dst = 0;
    range[0] = 0; range[1] = gh.t;                        /* :368 */
    for_each_segment(s, pos, dir, range, HTGX_Ks, iband, iquad, sc_cb, &sc);
    if(!gh.hit && !sc.hit) break;                         /* :373-375 */

    surface = !sc.hit;                                    /* :384 */
    dst = surface ? gh.t : sc.dst;
    pos_next[0] = pos[0] + dir[0] * dst;                  /* :387-389 */
    pos_next[1] = pos[1] + dir[1] * dst;
    pos_next[2] = pos[2] + dir[2] * dst;
    hit_prev = surface ? gh : GHIT_NONE;                  /* :392 */

    range[0] = 0; range[1] = dst;                         /* :396-399 */
    Tr_abs = transmissivity(s, rng, HTGX_Ka, iband, iquad, pos, dir, range);
    if(Tr_abs <= 0) break;

    if(surface) {                                         /* :402-420 */
      const struct htgx_ground_desc* gd = &s->scene.ground;
      const double wo[3] = {-dir[0], -dir[1], -dir[2]};
      N[0] = gh.N[0]; N[1] = gh.N[1]; N[2] = gh.N[2];
      if(general_materials(gd)) {                         /* :405-407 */
        int type;
        const int32_t m = hit_material(gd, gh.mtl, gh.N, dir);
        brdf_at(gd, m, wlen, canonical(rng), &type, &albedo);
        specular = type == HTGX_BRDF_SPECULAR;
      } else {
        albedo = gd->mtl_albedo[gh.mtl];
      }
      if(dot3(N, wo) < 0) { N[0] = -N[0]; N[1] = -N[1]; N[2] = -N[2]; }
      if(specular) {
        reflect_dir(dir, N, dir_next);
      } else {
        const double u0 = canonical(rng), u1 = canonical(rng);
        cos_hemisphere(N, u0, u1, dir_next);
      }
      refl = albedo;
    } else {                                              /* :422-443 */
      const double ks_gas = raw_in_seg(s, &sc.seg, HTGX_Ks, 1, iband, iquad, pos_next);
      const double ks_par = raw_in_seg(s, &sc.seg, HTGX_Ks, 2, iband, iquad, pos_next);
      const double ks = ks_par + ks_gas;
      double u0, u1, mu;
      phase_kind = canonical(rng) < ks_gas / ks ? 1 : 0; /* 1 Rayleigh, 0 HG */
      u0 = canonical(rng); u1 = canonical(rng);
      mu = phase_kind ? rayleigh_cos_sample(u0) : hg_cos_sample(g, u0);
      local_to_world(dir, mu, 2.0 * PI * u1, dir_next);
    }

    if(specular) {                                        /* :446-455 */
      memcpy(sun_dir, dir_next, sizeof(sun_dir));
      R = !in_solar_cone(s, sun_dir) ? 0 : dot3(N, sun_dir) < 0 ? 0 : albedo;
      sun_pdf = 1.0;
    } else if(surface) {                                  /* :457-466 */
      double c;
      sun_pdf = sun_sample(s, rng, sun_dir);
      c = dot3(N, sun_dir);
      R = c < 0 ? 0 : albedo / PI * c;
    } else {
      sun_pdf = sun_sample(s, rng, sun_dir);
      const double mu = dot3(dir, sun_dir);
      R = phase_kind ? rayleigh_eval(mu) : hg_eval(g, mu);
    }

    if(R <= 0) {                                          /* :470-485 */
      Tr = 0;
    } else {
      const double r_sun[2] = {0, FLT_MAX};
      const struct ghit h = ground_trace(s, pos_next, sun_dir, r_sun, &hit_prev);
      Tr = h.hit ? 0
        : transmissivity(s, rng, HTGX_Kext, iband, iquad, pos_next, sun_dir, r_sun);
    }

    ksi *= Tr_abs;                                        /* :495-496 */
    w += ksi * L_sun * Tr * R / sun_pdf;

    if(surface && canonical(rng) >= refl) break;          /* :499-500 */

    memcpy(pos, pos_next, sizeof(pos));
    memcpy(dir, dir_next, sizeof(dir));
  }
  return w;
}

static double
layer_temperature(const struct oracle_sky* s, int32_t j)
{
  const int32_t nl = (int32_t)s->scene.gas.nlayers;
  if(!s->scene.gas.temperature || nl <= 0) return 0.0; /* no gas layers */
  if(j < 0) j = 0;
  if(j >= nl) j = nl - 1;
  return s->scene.gas.temperature[j];
}

/* ..._lw.c:140-279 */
static double
radiance_lw
  (const struct oracle_sky* s, struct orng* rng, const double pos_in[3],
   const double dir_in[3], double wlen, uint32_t iband, uint32_t iquad)
{
  const double g = s->scene.clouds.hg_g ? s->scene.clouds.hg_g[iband] : 0.0;
  const double wlen_m = wlen * 1e-9;
  double pos[3], dir[3], w = 0;
  struct ghit hit_prev = GHIT_NONE; /* ..._lw.c:152 */
  memcpy(pos, pos_in, sizeof(pos));
  memcpy(dir, dir_in, sizeof(dir));

  for(;;) {
    struct lw_ctx c;
    struct ghit gh;
    double range[2], pos_next[3], dir_next[3], dst;
    const double r_inf[2] = {0, DBL_MAX};

    gh = ground_trace(s, pos, dir, r_inf, &hit_prev);     /* :188-190 */
    c.s = s; c.rng = rng; c.o = pos; c.d = dir;
    c.iband = iband; c.iquad = iquad;
    c.Ts = ran_exp(rng);                                  /* :193 */
    c.event = EV_NONE;
    c.dst = 0;
    range[0] = 0; range[1] = gh.t;                        /* :202 */
    for_each_segment(s, pos, dir, range, HTGX_Kext, iband, iquad, lw_cb, &c);

    if(!gh.hit && c.event == EV_NONE) { w = 0; break; }   /* :210-213 */

    dst = c.event != EV_NONE ? c.dst : gh.t;
    pos_next[0] = pos[0] + dir[0] * dst;                  /* :216-218 */
    pos_next[1] = pos[1] + dir[1] * dst;
    pos_next[2] = pos[2] + dir[2] * dst;

    if(c.event == EV_ABS) {                               /* :221-227 */
      w = oracle_planck_monochromatic(wlen_m, layer_temperature(s, c.seg.layer));
      break;
    }
    if(c.event == EV_SCA) {                               /* :233-236 */
      const double u0 = canonical(rng), u1 = canonical(rng);
      local_to_world(dir, hg_cos_sample(g, u0), 2.0 * PI * u1, dir_next);
      hit_prev = GHIT_NONE;                               /* :236 */
    } else {                                              /* :239-273 */
      const struct htgx_ground_desc* gd = &s->scene.ground;
      const double wo[3] = {-dir[0], -dir[1], -dir[2]};
      double N[3], refl, T;
      int specular = 0;
      N[0] = gh.N[0]; N[1] = gh.N[1]; N[2] = gh.N[2];
      if(general_materials(gd)) {                         /* :250-252 */
        int type;
        const int32_t m = hit_material(gd, gh.mtl, gh.N, dir);
        brdf_at(gd, m, wlen, canonical(rng), &type, &refl);
        specular = type == HTGX_BRDF_SPECULAR;
        T = mtl_temperature(gd, m);
      } else {
        refl = gd->mtl_albedo[gh.mtl];
        T = mtl_temperature(gd, (int32_t)gh.mtl);
      }
      if(dot3(N, wo) < 0) { N[0] = -N[0]; N[1] = -N[1]; N[2] = -N[2]; }
      if(specular) {
        reflect_dir(dir, N, dir_next);
      } else {
        const double u0 = canonical(rng), u1 = canonical(rng);
        cos_hemisphere(N, u0, u1, dir_next);
      }
      if(canonical(rng) >= refl) {                        /* :266-270 */
        w = T > 0 ? oracle_planck_monochromatic(wlen_m, T) : 0;
        break;
      }
      hit_prev = gh;                                      /* :272 */
    }
    memcpy(pos, pos_next, sizeof(pos));
    memcpy(dir, dir_next, sizeof(dir));
  }
  return w;
}

double
oracle_radiance
  (const struct oracle_sky* s, uint64_t seed, uint64_t pixel, uint64_t realisation,
   int spectral_type, int cpnt_mask, const double pos[3], const double dir[3],
   double wlen, uint32_t iband, uint32_t iquad)
{
  struct orng rng;
  orng_init(&rng, seed, pixel, realisation);
  return spectral_type == HTGX_SPECTRAL_LW
    ? radiance_lw(s, &rng, pos, dir, wlen, iband, iquad)
    : radiance_sw(s, &rng, cpnt_mask, pos, dir, wlen, iband, iquad);
}

void
oracle_radiance_many
  (const struct oracle_sky* s, uint64_t seed, uint64_t pixel, uint64_t realisation0,
   uint64_t n, int spectral_type, int cpnt_mask, const double pos[3],
   const double dir[3], double wlen, uint32_t iband, uint32_t iquad, double* out,
   int nthreads)
{
  int64_t i;
  if(nthreads > 0) omp_set_num_threads(nthreads);
  #pragma omp parallel for schedule(static, 4096)
  for(i = 0; i < (int64_t)n; ++i) {
    out[i] = oracle_radiance(s, seed, pixel, realisation0 + (uint64_t)i, spectral_type,
                             cpnt_mask, pos, dir, wlen, iband, iquad);
  }
}

/*******************************************************************************
 * Per-pixel realisations: ..._draw_map.c:52-107, 210-322, 325-436
 ******************************************************************************/
struct sensor {
  double pos[3];
  double x[3], y[3], z[3]; /* camera: right, up, forward; rectangle: frame */
  double tan_half, aspect;
  int ortho;               /* orthographic camera (-P) */
  double half_h;           /* its image plane half height */
  double lens_radius, focal_dst; /* thin-lens perspective camera */
  double axis_x[3], axis_y[3];
};

static int
sensor_setup(const struct htgx_map_desc* d, struct sensor* c)
{
  double t[3];
  int a;
  for(a = 0; a < 3; ++a) { c->pos[a] = d->pos[a]; t[a] = d->tgt[a] - d->pos[a]; }
  if(normalize3(c->z, t) <= 0) return HTGX_BAD_ARG;
  cross3(t, c->z, d->up);
  if(normalize3(c->x, t) <= 0) return HTGX_BAD_ARG;
  if(d->sensor == HTGX_SENSOR_CAMERA) {
    /* Pinhole camera: x = image right, y = image up (film v = 0 is the top
     * row, the first row of an htrdr-image, doc/htrdr-image.5) */
    cross3(c->y, c->x, c->z);
    c->tan_half = tan(d->fov_y * PI / 180.0 * 0.5);
    c->aspect = (double)d->width / (double)d->height;
    if(d->projection != HTGX_PROJECTION_PERSPECTIVE && d->projection != HTGX_PROJECTION_ORTHOGRAPHIC)
      return HTGX_BAD_ARG;
    c->ortho = d->projection == HTGX_PROJECTION_ORTHOGRAPHIC;
    c->half_h = 0.5 * d->ortho_height;
    c->lens_radius = d->lens_radius;
    c->focal_dst = d->focal_dst;
    if(c->ortho ? !(d->ortho_height > 0)
                : !(d->fov_y > 0 && d->fov_y < 180) || !(d->lens_radius >= 0)
                  || (d->lens_radius > 0 && !(d->focal_dst > 0)))
      return HTGX_BAD_ARG;
  } else {
    /* Rectangle frame, src/core/htrdr_rectangle.c:101-131 */
    cross3(t, c->z, c->x);
    normalize3(c->y, t);
    for(a = 0; a < 3; ++a) {
      c->axis_x[a] = c->x[a] * d->rect_size[0] * 0.5;
      c->axis_y[a] = c->y[a] * d->rect_size[1] * 0.5;
    }
  }
  return HTGX_OK;
}

struct spectral_ctx {
  struct oracle_wlen_planck* planck;
};

static uint32_t
find_band(const struct oracle_sky* s, double wlen)
{
  const uint32_t nb = s->scene.spectral.nbands;
  uint32_t lo = 0, hi = nb - 1;
  while(lo < hi) {
    const uint32_t mid = (lo + hi) / 2;
    if(s->scene.spectral.band_hi[mid] > wlen) hi = mid; else lo = mid + 1;
  }
  return lo;
}

static uint32_t
sample_quad(const struct oracle_sky* s, uint32_t iband, double r)
{
  const uint32_t off = s->quad_offset[iband];
  const uint32_t nq = s->scene.spectral.nquads[iband];
  uint32_t q;
  for(q = 0; q + 1 < nq; ++q) if(s->quad_cdf[off + q] > r) break;
  return q;
}

static double
realisation
  (const struct oracle_sky* s, const struct htgx_map_desc* d,
   const struct sensor* sn, const struct oracle_wlen_planck* planck,
   uint64_t pixel, uint64_t ireal, int* rejected)
{
  struct orng rng;
  const uint64_t px = pixel % d->width, py = pixel / d->width;
  const double psx = 1.0 / (double)d->width, psy = 1.0 / (double)d->height;
  double org[3], dir[3], r0, r1, r2, wlen, pdf, weight;
  uint32_t iband, iquad;
  int a;

  orng_init(&rng, d->seed, pixel, ireal);
  *rejected = 0;

  if(d->sensor == HTGX_SENSOR_CAMERA) {
    const double fx = ((double)px + canonical(&rng)) * psx;   /* :364-369 */
    const double fy = ((double)py + canonical(&rng)) * psy;
    const double l0 = canonical(&rng), l1 = canonical(&rng); /* lens sample */
    double v[3];
    if(sn->ortho) {
      /* orthographic: from the image plane, along the view axis */
      for(a = 0; a < 3; ++a) {
        org[a] = sn->pos[a] + sn->x[a] * ((2.0*fx - 1.0) * sn->half_h * sn->aspect)
                            + sn->y[a] * ((1.0 - 2.0*fy) * sn->half_h);
        dir[a] = sn->z[a];
      }
    } else {
      for(a = 0; a < 3; ++a) {
        v[a] = sn->z[a] + sn->x[a] * ((2.0*fx - 1.0) * sn->tan_half * sn->aspect)
                        + sn->y[a] * ((1.0 - 2.0*fy) * sn->tan_half);
        org[a] = sn->pos[a];
      }
      if(sn->lens_radius > 0) {
        /* thin lens: a point of the lens disk towards the point of the
         * pinhole ray on the plane in focus */
        const double r = sn->lens_radius * sqrt(l0), phi = 2.0 * PI * l1;
        const double lx = r * cos(phi), ly = r * sin(phi);
        for(a = 0; a < 3; ++a) {
          const double focus = sn->pos[a] + v[a] * sn->focal_dst;
          org[a] = sn->pos[a] + sn->x[a] * lx + sn->y[a] * ly;
          v[a] = focus - org[a];
        }
      }
      normalize3(dir, v);
    }
  } else {
    const double sx = ((double)px + canonical(&rng)) * psx;   /* :69-73 */
    const double sy = ((double)py + canonical(&rng)) * psy;
    const double up[3] = {0, 0, 1};
    const double r_inf[2] = {0, DBL_MAX};
    struct ghit h;
    for(a = 0; a < 3; ++a)
      org[a] = sn->pos[a] + sn->axis_x[a] * (sx*2.0 - 1.0) + sn->axis_y[a] * (sy*2.0 - 1.0);
    h = ground_trace(s, org, up, r_inf, NULL);                /* :75-100 */
    if(h.hit) {
      const struct htgx_ground_desc* gd = &s->scene.ground;
      int under;
      if(gd->interfaces) {
        /* the side the up ray travels in: front if cos(up, -N) < 0 (:93);
         * rejected unless its material is the sky's (:96) */
        int32_t f[3];
        interface_of(gd, h.mtl, f);
        under = (-h.N[2] < 0 ? f[0] : f[1]) != gd->sky_mtl;
      } else {
        under = dot3(up, h.N) >= 0; /* the up ray meets the sky side from below */
      }
      if(under) { *rejected = 1; return 0; }
    }
    {
      const double u0 = canonical(&rng), u1 = canonical(&rng);
      cos_hemisphere(sn->z, u0, u1, dir);                     /* :103-104 */
    }
  }

  r0 = canonical(&rng); r1 = canonical(&rng); r2 = canonical(&rng);
  wlen = oracle_wlen_planck_sample(planck, r0, r1, &pdf);
  pdf *= 1e9;
  iband = find_band(s, wlen);
  iquad = sample_quad(s, iband, r2);

  if(d->sensor == HTGX_SENSOR_CAMERA) {
    weight = d->spectral_type == HTGX_SPECTRAL_LW
      ? radiance_lw(s, &rng, org, dir, wlen, iband, iquad)
      : radiance_sw(s, &rng, HTGX_RADIANCE_ALL, org, dir, wlen, iband, iquad);
    weight /= pdf;
  } else if(d->spectral_type == HTGX_SPECTRAL_LW) {           /* :265-268 */
    weight = radiance_lw(s, &rng, org, dir, wlen, iband, iquad) * PI / pdf;
  } else {                                                    /* :269-302 */
    double sun_dir[3], Ld = 0, Lf, cosN;
    (void)sun_sample(s, &rng, sun_dir);
    cosN = dot3(sn->z, sun_dir);
    if(cosN > 0)
      Ld = radiance_sw(s, &rng, HTGX_RADIANCE_DIRECT, org, sun_dir, wlen, iband, iquad);
    Lf = radiance_sw(s, &rng, HTGX_RADIANCE_DIFFUSE, org, dir, wlen, iband, iquad);
    weight = (cosN * s->sun_solid_angle * Ld + PI * Lf) / pdf;
  }
  return weight;
}

static int
make_planck(const struct oracle_sky* s, const struct htgx_map_desc* d,
            struct oracle_wlen_planck** p);

/* The desc and sampler of channel c of a CIE XYZ image: the channel's CDF
 * and its own random streams, keyed by seed ^ (c+1) << 56 */
static int
cie_channel(const struct oracle_sky* s, const struct htgx_map_desc* d, int c,
            struct htgx_map_desc* dc, struct oracle_wlen_planck** p)
{
  const uint32_t nb = s->scene.spectral.nbands;
  double range[2];
  range[0] = fmax(d->wlen_range[0], s->scene.spectral.band_lo[0]);
  range[1] = fmin(d->wlen_range[1], s->scene.spectral.band_hi[nb - 1]);
  *dc = *d;
  dc->seed = d->seed ^ ((uint64_t)(c + 1) << 56);
  dc->spectral_type = HTGX_SPECTRAL_SW;
  return oracle_wlen_cie_create(range, oracle_spectral_bands_count(range), c, p);
}

static int
make_planck(const struct oracle_sky* s, const struct htgx_map_desc* d,
            struct oracle_wlen_planck** p)
{
  /* src/atmosphere/htrdr_atmosphere.c:384-415: the range is clamped to the
   * sky's spectral data; default reference temperatures of
   * src/atmosphere/htrdr_atmosphere_args.c:216-230 */
  const uint32_t nb = s->scene.spectral.nbands;
  double range[2];
  double Tref = d->ref_temperature;
  size_t n;
  range[0] = fmax(d->wlen_range[0], s->scene.spectral.band_lo[0]);
  range[1] = fmin(d->wlen_range[1], s->scene.spectral.band_hi[nb - 1]);
  if(!(range[1] >= range[0])) return HTGX_BAD_ARG;
  if(Tref <= 0) Tref = d->spectral_type == HTGX_SPECTRAL_LW ? 290.0 : 5778.0;
  n = oracle_spectral_bands_count(range);
  return oracle_wlen_planck_create(range, n, Tref, p);
}

double
oracle_realisation
  (const struct oracle_sky* s, const struct htgx_map_desc* d, uint64_t pixel,
   uint64_t ireal)
{
  struct sensor sn;
  struct oracle_wlen_planck* planck = NULL;
  struct htgx_map_desc dc = *d;
  int rejected;
  double w;
  if(sensor_setup(d, &sn) != HTGX_OK) return NAN;
  if(d->spectral_type == HTGX_SPECTRAL_SW_CIE_XYZ) {
    /* bits 32-33 of the realisation index select the channel */
    if((ireal >> 32) > 2) return NAN;
    if(cie_channel(s, d, (int)(ireal >> 32), &dc, &planck) != HTGX_OK) return NAN;
    ireal &= 0xFFFFFFFFull;
  } else if(make_planck(s, d, &planck) != HTGX_OK) {
    return NAN;
  }
  w = realisation(s, &dc, &sn, planck, pixel, ireal, &rejected);
  oracle_wlen_planck_destroy(planck);
  return rejected ? NAN : w;
}

/* Neumaier compensated summation */
static void
nsum(double* sum, double* c, double x)
{
  const double t = *sum + x;
  if(fabs(*sum) >= fabs(x)) *c += (*sum - t) + x;
  else *c += (x - t) + *sum;
  *sum = t;
}

static int draw_map_channel
  (const struct oracle_sky* s, const struct htgx_map_desc* d, const struct sensor* sn,
   const struct oracle_wlen_planck* planck, struct htgx_mc_accum* accum, int nthreads);

int
oracle_draw_map
  (const struct oracle_sky* s, const struct htgx_map_desc* d,
   struct htgx_mc_accum* accum, int nthreads)
{
  struct sensor sn;
  struct oracle_wlen_planck* planck = NULL;
  int res, c;
  if(!s || !d || !accum || !d->width || !d->height || !d->spp) return HTGX_BAD_ARG;
  if((res = sensor_setup(d, &sn)) != HTGX_OK) return res;
  if(d->spectral_type != HTGX_SPECTRAL_SW_CIE_XYZ) {
    if((res = make_planck(s, d, &planck)) != HTGX_OK) return res;
    res = draw_map_channel(s, d, &sn, planck, accum, nthreads);
    oracle_wlen_planck_destroy(planck);
    return res;
  }
  /* draw_pixel_image, ..._draw_map.c:134-199: spp realisations per channel,
   * into 3 accumulator planes */
  for(c = 0; c < 3; ++c) {
    struct htgx_map_desc dc;
    if((res = cie_channel(s, d, c, &dc, &planck)) != HTGX_OK) return res;
    res = draw_map_channel(s, &dc, &sn, planck,
                           accum + (size_t)c * d->width * d->height, nthreads);
    oracle_wlen_planck_destroy(planck);
    if(res != HTGX_OK) return res;
  }
  return HTGX_OK;
}

static int
draw_map_channel
  (const struct oracle_sky* s, const struct htgx_map_desc* d, const struct sensor* snp,
   const struct oracle_wlen_planck* planck, struct htgx_mc_accum* accum, int nthreads)
{
  const struct sensor sn = *snp;
  int64_t npix, ip;
  npix = (int64_t)d->width * d->height;
#ifdef _OPENMP
  if(nthreads > 0) omp_set_num_threads(nthreads);
  #pragma omp parallel for schedule(dynamic, 16)
#endif
  for(ip = 0; ip < npix; ++ip) {
    struct htgx_mc_accum a;
    uint32_t r;
    memset(&a, 0, sizeof(a));
    for(r = 0; r < d->spp; ++r) {
      struct timespec t0, t1;
      int rejected;
      double w, usec;
      clock_gettime(CLOCK_MONOTONIC, &t0);
      w = realisation(s, d, &sn, planck, (uint64_t)ip, (uint64_t)(d->spp_offset + r), &rejected);
      clock_gettime(CLOCK_MONOTONIC, &t1);
      if(rejected) { a.nfail++; continue; }
      usec = (double)(t1.tv_sec - t0.tv_sec) * 1e6 + (double)(t1.tv_nsec - t0.tv_nsec) * 1e-3;
      nsum(&a.sum, &a.sum_c, w);
      nsum(&a.sum2, &a.sum2_c, w * w);
      a.tsum += usec;
      a.tsum2 += usec * usec;
      a.n++;
    }
    accum[ip] = a;
  }
  (void)nthreads;
  return HTGX_OK;
}

static int
draw_pixels_channel
  (const struct oracle_sky* s, const struct htgx_map_desc* d, const struct sensor* sn,
   const struct oracle_wlen_planck* planck, const uint64_t* pixels, uint64_t npixels,
   struct htgx_mc_accum* accum, int nthreads)
{
  int64_t i;
#ifdef _OPENMP
  if(nthreads > 0) omp_set_num_threads(nthreads);
  #pragma omp parallel for schedule(dynamic, 4)
#endif
  for(i = 0; i < (int64_t)npixels; ++i) {
    struct htgx_mc_accum a;
    uint32_t r;
    memset(&a, 0, sizeof(a));
    for(r = 0; r < d->spp; ++r) {
      int rejected;
      const double w = realisation(s, d, sn, planck, pixels[i],
        (uint64_t)(d->spp_offset + r), &rejected);
      if(rejected) { a.nfail++; continue; }
      nsum(&a.sum, &a.sum_c, w);
      nsum(&a.sum2, &a.sum2_c, w * w);
      a.n++;
    }
    accum[i] = a;
  }
  (void)nthreads;
  return HTGX_OK;
}

int
oracle_draw_pixels
  (const struct oracle_sky* s, const struct htgx_map_desc* d,
   const uint64_t* pixels, uint64_t npixels, struct htgx_mc_accum* accum, int nthreads)
{
  struct sensor sn;
  struct oracle_wlen_planck* planck = NULL;
  int res, c;
  if(!s || !d || !pixels || !accum || !d->width || !d->height || !d->spp) return HTGX_BAD_ARG;
  if((res = sensor_setup(d, &sn)) != HTGX_OK) return res;
  if(d->spectral_type != HTGX_SPECTRAL_SW_CIE_XYZ) {
    if((res = make_planck(s, d, &planck)) != HTGX_OK) return res;
    res = draw_pixels_channel(s, d, &sn, planck, pixels, npixels, accum, nthreads);
    oracle_wlen_planck_destroy(planck);
    return res;
  }
  for(c = 0; c < 3; ++c) {
    struct htgx_map_desc dc;
    if((res = cie_channel(s, d, c, &dc, &planck)) != HTGX_OK) return res;
    res = draw_pixels_channel(s, &dc, &sn, planck, pixels, npixels,
                              accum + (size_t)c * npixels, nthreads);
    oracle_wlen_planck_destroy(planck);
    if(res != HTGX_OK) return res;
  }
  return HTGX_OK;
}

/* src/core/htrdr_accum.h:48-68 */
void
oracle_accum_estimate(const struct htgx_mc_accum* acc, double* E, double* SE)
{
  if(!acc->n) { *E = 0; *SE = 0; return; }
  {
    const double N = (double)acc->n;
    const double S = acc->sum + acc->sum_c;
    const double S2 = acc->sum2 + acc->sum2_c;
    const double e = S / N;
    const double V = fmax(S2 / N - e*e, 0);
    *E = e;
    *SE = sqrt(V / N);
  }
}
