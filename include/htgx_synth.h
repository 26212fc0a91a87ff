/*
 * Seeded synthetic scenes for libhtgx.so.
 *
 * The reference reads its atmosphere from htgop (-a), htcp (-c) and htmie
 * (-m) files (doc/htrdr-atmosphere.1.in:144-149,274-277) through htsky; none
 * of those readers nor their netCDF dependency exist here, so the benchmark
 * and test scenes are generated in memory from a parameter block. The
 * generated arrays fill a struct htgx_scene_desc that htgx_ctx_create
 * consumes exactly as it would consume arrays read from files.
 *
 * Cloud field: LWC = exp(mu + sigma*N(0,1)) g/m^3 per cell, masked by
 * thresholding an independent N(0,1) field smoothed with a periodic Gaussian
 * of `smooth_sigma' cells at its (1 - cover) quantile. Gas: uniform layers,
 * ka = ka0*exp(-z/H), ks = ks0*exp(-z/H) at the layer middle,
 * T = T_ground - lapse*z. Longwave bands: per (band, k-point), ka0 is
 * log-uniform in [lw_ka_min, lw_ka_max] and the k-point weights are
 * Dirichlet(1). Shortwave bands with several k-points: ka0 is log-uniform
 * in [ka0/10, 10 ka0] per k-point, weights Dirichlet(1).
 *
 * Mesh ground (ground_type = HTGX_GROUND_MESH): a heightfield of hf_res^2
 * vertices over the cloud grid's XY extent, fractional Brownian motion of
 * hf_octaves octaves of periodic value noise with amplitude ratio
 * 2^-hurst per octave, rescaled to [0, hf_relief] m. Each grid quad gives
 * two triangles wound counter-clockwise seen from above, so that
 * (v1 - v0) x (v2 - v0), the side of the sky, points up. Triangles take one
 * of 4 materials by the height band of their centroid (albedo 0.1, 0.25,
 * 0.4, 0.6, from low to high).
 */
#ifndef HTGX_SYNTH_H
#define HTGX_SYNTH_H

#include "htgx.h"

#ifdef __cplusplus
extern "C" {
#endif

struct htgx_synth_params {
  int spectral_type;        /* enum htgx_spectral_type */
  /* clouds */
  uint32_t cloud_def[3];    /* 0 => clear sky */
  double cloud_lower[3];
  double cloud_upper[3];
  double lwc_mu, lwc_sigma; /* lognormal LWC, g/m^3 */
  double cover;             /* cloud fraction in ]0, 1] */
  double smooth_sigma;      /* cells */
  double cext;              /* m^2/g */
  double ssa;
  double hg_g;
  /* gas */
  uint32_t nlayers;
  double layer_thickness;   /* m */
  double ka0, ks0;          /* m^-1 at z = 0 (shortwave) */
  double scale_height;      /* m */
  double T_ground;          /* K */
  double lapse_rate;        /* K/m */
  /* spectral */
  uint32_t nbands, nquads;  /* per band */
  double wlen_lo, wlen_hi;  /* nm, whole band range */
  double lw_ka_min, lw_ka_max;
  /* ground */
  int ground_type;          /* enum htgx_ground_type (plane only here) */
  double ground_albedo;
  double ground_temperature;
  /* sun */
  double sun_azimuth, sun_elevation;
  /* octree */
  double optical_thickness;
  uint32_t leaf_block;
  uint64_t seed;
  /* mesh ground */
  uint32_t hf_res;          /* vertices per side */
  uint32_t hf_octaves;
  double hf_relief;         /* m */
  double hf_hurst;
};

struct htgx_synth;

/* Parameters of a benchmark configuration: 0 = 64x64x32 clouds, SW 550 nm;
 * 1 = 256x256x64 clouds, SW; 2 = 256x256x64 clouds, LW 8 bands x 16 k;
 * 4 = 512x512x128 clouds over 51.2 km, SW, 1024^2-vertex fBm ground */
int htgx_synth_params_preset(int config, struct htgx_synth_params* params);

/* Sensor/spectral settings of a configuration (image or flux map) */
int htgx_synth_map_preset(int config, struct htgx_map_desc* desc);

int htgx_synth_create(const struct htgx_synth_params* params, struct htgx_synth** out);
const struct htgx_scene_desc* htgx_synth_get_scene(const struct htgx_synth* synth);
int htgx_synth_destroy(struct htgx_synth* synth);

#ifdef __cplusplus
}
#endif

#endif /* HTGX_SYNTH_H */
