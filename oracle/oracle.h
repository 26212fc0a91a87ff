/*
 * CPU FP64 oracle of the htrdr-atmosphere null-collision core.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the CPU reference the GPU
 * library (libhtgx.so) is compared against in the tests; nothing in the product links, loads or
 * calls it. Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline
 * leg may use it.
 *
 * It restates, in plain C with recursion and no GPU-oriented layout, the
 * algorithms of the reference path (citations are file:line under the htrdr
 * tree) and the semantics the GPU library fixes for the un-vendored libraries
 * (htsky property model and octree, svx traversal, star-sp samplers, star-sf
 * phase functions, star-camera pinhole); see oracle.c for each function.
 *
 * Parity pins (what checks each function, tests/ file):
 *   threefry4x64_20 ........ Random123's published known answers for the zero
 *                            and the all-ones counter and key (the latter
 *                            runs the key schedule's parity word through every
 *                            injection); every counter and key word changes
 *                            every output word (test_oracle_rng.py)
 *   octree / segments ...... brute-force DDA over the finest cells on tiny
 *                            grids, exact FP32 min/max (test_oracle_svo.py)
 *   planck / wiebelt ....... Stefan-Boltzmann integral, Wien's law, numerical
 *                            quadrature of Planck (test_oracle_spectral.py)
 *   brightness temperature . closed-form monochromatic inverse, round trips
 *   wavelength sampler ..... brute-force CDF on small band counts, mono case
 *   phase / cone / lambert . Kolmogorov-Smirnov against analytic CDFs, moments
 *                            (test_oracle_samplers.py)
 *   transmissivity ......... k_min = k_max gives exp(-tau) with zero variance;
 *                            absorbing cloud against an exact cell-by-cell
 *                            integral at three octree thresholds
 *                            (test_oracle_walks.py)
 *   SW walk ................ vacuum + Lambertian ground; direct sun through
 *                            Kext; volume scattering by the first-order closed
 *                            form of an optically thin slab (HG particles,
 *                            Rayleigh gas, a 3:1 mix; 20 and 150 degrees from
 *                            the sun), which a flipped scattering angle or
 *                            swapped gas/particle odds fail by many standard
 *                            errors (test_oracle_walks.py, closed_forms.py)
 *   LW walk ................ isothermal enclosures, absorbing slab; emission at
 *                            the layer temperature under a 9 K-per-layer
 *                            gradient (gas and cloud absorption, looking up
 *                            and down from three heights), which a layer index
 *                            off by one or the start layer's temperature fail
 *                            (test_oracle_walks.py, closed_forms.py)
 *   flux maps .............. SW vacuum cos Omega L_sun; LW thick isothermal pi B
 *   mesh ground casts ...... flat mesh = plane, heightfield rejection area,
 *                            no leaks (test_oracle_ground.py); the optional
 *                            BVH = the brute force bit for bit
 *   CIE 1931 sampler ....... CIE table, equal-energy integrals, exact in-band
 *                            inversion, KS, vacuum closed form (test_oracle_cie.py)
 *   cameras ................ ground-footprint closed forms (test_oracle_cameras.py)
 *   cloud / ground copies .. explicit tiling, infinite-layer and tiled-plane
 *                            closed forms (test_oracle_repeat.py)
 *   ground materials ....... list selection, mean = linear interpolation,
 *                            interface sides, mirror and two-sided closed forms
 *                            (test_oracle_materials.py)
 * "Parity unpinned": the bits of the reference's own random streams, which
 * counter-based Threefry replaces by design (per-pixel statistical parity
 * only); the thin-lens sample mapping of star-camera (not vendored; the
 * expectation per pixel is pinned, not the mapping of the lens samples); and
 * which of two mrumtl BRDF entries around a wavelength a canonical selects
 * (mrumtl not vendored; the mean reflectivity is pinned, not the choice).
 */
#ifndef HTGX_ORACLE_H
#define HTGX_ORACLE_H

#include "htgx.h"

#ifdef __cplusplus
extern "C" {
#endif

struct oracle_sky;

/* Random streams */
void oracle_threefry4x64_20
  (const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4]);
/* n canonical numbers of the stream keyed by (seed, pixel, realisation) */
void oracle_stream_canonicals
  (uint64_t seed, uint64_t pixel, uint64_t realisation, uint32_t n, double* out);

/* Scene */
int oracle_sky_create(const struct htgx_scene_desc* scene, struct oracle_sky** sky);
void oracle_sky_destroy(struct oracle_sky* sky);
int oracle_sky_get_svo_info
  (const struct oracle_sky* sky, uint32_t* nlevels, uint64_t* nnodes,
   uint64_t* nleaves);

/* Leaves/layer segments along a ray, in front-to-back order */
int oracle_trace_segments
  (const struct oracle_sky* sky,
   const double ray[8], /* org[3], dir[3], range[2] */
   int prop, uint32_t iband, uint32_t iquad,
   struct htgx_segment* segs, uint32_t max_segs, uint32_t* nsegs);

/* Leaf id and bounds (LWC min/max, g/m^3) of the octree leaf containing the
 * cloud cell (ix,iy,iz) */
int oracle_leaf_of_cell
  (const struct oracle_sky* sky, const uint32_t cell[3], int64_t* leaf,
   float* lwc_min, float* lwc_max);

/* Property at a point (htsky_fetch_raw_property) */
double oracle_fetch_raw
  (const struct oracle_sky* sky, int prop, int cpnt /*1 gas,2 particles,3 all*/,
   uint32_t iband, uint32_t iquad, const double pos[3]);

/* Transmissivity estimator of ..._sw.c:219-251 for one realisation whose
 * random stream is keyed by (seed, pixel, realisation). */
double oracle_transmissivity
  (const struct oracle_sky* sky, uint64_t seed, uint64_t pixel,
   uint64_t realisation, int prop, uint32_t iband, uint32_t iquad,
   const double pos[3], const double dir[3], const double range[2]);

/* One realisation weight, same semantics as htgx_debug_realisations */
double oracle_realisation
  (const struct oracle_sky* sky, const struct htgx_map_desc* desc,
   uint64_t pixel, uint64_t realisation);

/* Radiance of one walk from a given ray (W/m^2/sr/m), wavelength in nm */
double oracle_radiance
  (const struct oracle_sky* sky, uint64_t seed, uint64_t pixel,
   uint64_t realisation, int spectral_type, int cpnt_mask,
   const double pos[3], const double dir[3], double wlen,
   uint32_t iband, uint32_t iquad);

/* oracle_radiance of realisations realisation0 .. realisation0 + n - 1 of one
 * ray into out[n], nthreads OpenMP threads (0 = all) */
void oracle_radiance_many
  (const struct oracle_sky* sky, uint64_t seed, uint64_t pixel, uint64_t realisation0,
   uint64_t n, int spectral_type, int cpnt_mask, const double pos[3],
   const double dir[3], double wlen, uint32_t iband, uint32_t iquad, double* out,
   int nthreads);

/* The whole map into per-pixel accumulators (row-major; for a CIE XYZ
 * image 3 planes, X, Y, Z), nthreads OpenMP threads (0 = all). CIE channel c
 * draws from the streams keyed by seed ^ (c+1) << 56; oracle_realisation
 * takes the channel from bits 32-33 of the realisation index. */
int oracle_draw_map
  (const struct oracle_sky* sky, const struct htgx_map_desc* desc,
   struct htgx_mc_accum* accum, int nthreads);

/* A subset of the pixels of a map (pixel ids y*W+x) into accum[npixels]
 * (for a CIE XYZ image accum[3][npixels]) */
int oracle_draw_pixels
  (const struct oracle_sky* sky, const struct htgx_map_desc* desc,
   const uint64_t* pixels, uint64_t npixels, struct htgx_mc_accum* accum,
   int nthreads);

/* Spectral helpers (src/core/htrdr_spectral.h, htrdr_spectral.c) */
double oracle_planck_monochromatic(double lambda_m, double temperature);
double oracle_wiebelt(double v);
double oracle_blackbody_fraction(double lambda0_m, double lambda1_m, double T);
double oracle_planck(double lambda_min_m, double lambda_max_m, double T);
int oracle_brightness_temperature
  (double lambda_min_m, double lambda_max_m, double radiance, double* T);
double oracle_radiance_temperature
  (double lambda_min_m, double lambda_max_m, double radiance);

/* Planck wavelength sampler (src/core/htrdr_ran_wlen_planck.c) */
struct oracle_wlen_planck;
int oracle_wlen_planck_create
  (const double range_nm[2], size_t nbands, double ref_temperature,
   struct oracle_wlen_planck** out);
void oracle_wlen_planck_destroy(struct oracle_wlen_planck* p);
double oracle_wlen_planck_sample
  (const struct oracle_wlen_planck* p, double r0, double r1, double* pdf_nm);
/* #bands of the discrete CDF as htrdr_atmosphere_create computes it
 * (src/atmosphere/htrdr_atmosphere.c:53-70) */
size_t oracle_spectral_bands_count(const double range_nm[2]);

/* Direction samplers, kind as htgx_debug_sample_dirs; the sun of `sky' */
int oracle_sample_dirs
  (const struct oracle_sky* sky, int kind, double param, const double dir_in[3],
   const double* u, uint32_t n, double* out);
double oracle_phase_eval(int kind, double param, double cos_theta);

/* Closest ground hit of ray (org, dir, range) leaving triangle prev_prim
 * (-1: none; prev_edge: that hit was on an edge), by brute force over every
 * mesh triangle (htrdr_atmosphere_ground_trace_ray, ..._ground.c:170-217).
 * Returns whether it hits; t = INFINITY and prim = -1 when not. */
int oracle_ground_trace
  (const struct oracle_sky* sky, const double ray[8], int64_t prev_prim,
   int prev_edge, double* t, int64_t* prim, int* edge);

/* Cast mesh-ground rays through a BVH of the oracle's own (on) or against
 * every triangle (off, the default). Both give the same casts bit for bit;
 * the BVH only makes timing runs on large terrains affordable. */
int oracle_sky_use_bvh(struct oracle_sky* sky, int on);

/* Ground materials of a scene's ground description: the BRDF entry of
 * material m at wavelength wlen (nm) for the canonical r, returning its type
 * (enum htgx_brdf_type) and reflectivity (mrumtl_fetch_brdf, called at
 * src/core/htrdr_materials.c:430-433); the material reflecting a hit of
 * interface ifc by a ray along dir, N = (v1 - v0) x (v2 - v0) being the hit
 * triangle's normal (src/core/htrdr_interface.h:45-87) */
int oracle_brdf_fetch
  (const struct htgx_ground_desc* ground, int32_t m, double wlen, double r, double* refl);
int32_t oracle_hit_material
  (const struct htgx_ground_desc* ground, uint32_t ifc, const double N[3], const double dir[3]);

/* CIE 1931 XYZ sampler of one channel (0 X, 1 Y, 2 Z), range within
 * [380, 780] nm in nbands bands; sampled with oracle_wlen_planck_sample (pdf
 * = 1 / integral of the channel, nm^-1) and freed with
 * oracle_wlen_planck_destroy (src/core/htrdr_ran_wlen_cie_xyz.c) */
int oracle_wlen_cie_create
  (const double range[2], size_t nbands, int channel, struct oracle_wlen_planck** out);
double oracle_cie_fit(int channel, double lambda_nm);

/* Per-pixel estimate (src/core/htrdr_accum.h:48-68) */
void oracle_accum_estimate
  (const struct htgx_mc_accum* acc, double* E, double* SE);

#ifdef __cplusplus
}
#endif

#endif /* HTGX_ORACLE_H */
