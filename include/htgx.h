/*
 * htgx -- MI355X (gfx950) null-collision path-tracing core for htrdr-atmosphere.
 *
 * C-ABI of the GPU library libhtgx.so. Plain C types only: pointers, sizes,
 * fixed-width integers and doubles. Every entry point returns an `int' whose
 * values are those of rsys `res_T' (RES_OK = 0, RES_BAD_ARG, RES_MEM_ERR,
 * RES_IO_ERR, RES_UNKNOWN_ERR, RES_BAD_OP), so a caller written against
 * htrdr's core can test them exactly as it tests `htrdr_draw_map'.
 *
 * What the library replaces in the reference (file:line under the htrdr tree):
 *  - htrdr_draw_map + draw_map + draw_tile, the OpenMP/MPI tile dispatcher
 *    (src/core/htrdr_draw_map.c:271-621, src/core/htrdr_draw_map.h:89-93),
 *    together with the per-pixel callbacks it drives for the atmosphere
 *    command: draw_pixel_xwave and draw_pixel_flux
 *    (src/atmosphere/htrdr_atmosphere_draw_map.c:210-436);
 *  - the radiative random walks they call,
 *    atmosphere_compute_radiance_sw / _lw
 *    (src/atmosphere/htrdr_atmosphere_compute_radiance_sw.c:257-511,
 *     src/atmosphere/htrdr_atmosphere_compute_radiance_lw.c:141-279);
 *  - the star-vx octree traversal and htsky property fetches they rely on
 *    (call sites ..._sw.c:87-88,130-131,163-166,202-203,243-244,369-370);
 *  - the per-tile star-sp Threefry proxies (src/core/htrdr_draw_map.c:427-449),
 *    replaced by counter-based Threefry-4x64-20 streams keyed by
 *    (seed, pixel, realisation).
 *
 * Pixel layouts written by htgx_draw_map are bit-compatible with
 * struct atmosphere_pixel_xwave / atmosphere_pixel_flux
 * (src/atmosphere/htrdr_atmosphere_c.h:46-68) and struct htrdr_accum /
 * htrdr_estimate (src/core/htrdr_accum.h:31-44), so the reference's
 * dump_buffer (src/atmosphere/htrdr_atmosphere_draw_map.c:546-587) can write
 * them unchanged.
 */
#ifndef HTGX_H
#define HTGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Result codes: same values as rsys res_T. */
enum htgx_res {
  HTGX_OK = 0,
  HTGX_BAD_ARG = 1,
  HTGX_MEM_ERR = 2,
  HTGX_IO_ERR = 3,
  HTGX_UNKNOWN_ERR = 4,
  HTGX_BAD_OP = 5
};

/* Same order as enum htrdr_spectral_type (src/core/htrdr_spectral.h:35-39) */
enum htgx_spectral_type {
  HTGX_SPECTRAL_LW = 0,
  HTGX_SPECTRAL_SW = 1,
  HTGX_SPECTRAL_SW_CIE_XYZ = 2 /* Shortwave CIE 1931 XYZ image (camera only: a flux
                                  map returns HTGX_BAD_ARG, as in the reference) */
};

/* Radiance components, enum atmosphere_radiance_cpnt_flag
 * (src/atmosphere/htrdr_atmosphere_c.h:38-44) */
enum htgx_radiance_cpnt {
  HTGX_RADIANCE_DIRECT = 1,
  HTGX_RADIANCE_DIFFUSE = 2,
  HTGX_RADIANCE_ALL = 3
};

/* Optical properties, enum htsky_property as used at the call sites
 * (..._sw.c:87,163,397,483; ..._lw.c:78,113) */
enum htgx_property {
  HTGX_Ka = 0,   /* absorption coefficient */
  HTGX_Ks = 1,   /* scattering coefficient */
  HTGX_Kext = 2  /* extinction coefficient */
};

enum htgx_sensor_type {
  HTGX_SENSOR_CAMERA = 0,    /* pinhole perspective camera, option -C */
  HTGX_SENSOR_RECTANGLE = 1  /* flux-density map, option -p */
};

/* Camera projections (star-camera, set up at src/atmosphere/
 * htrdr_atmosphere.c:112-165): -C is perspective (a pinhole when the lens
 * radius is 0, else a thin lens), -P orthographic */
enum htgx_projection {
  HTGX_PROJECTION_PERSPECTIVE = 0,
  HTGX_PROJECTION_ORTHOGRAPHIC = 1
};

enum htgx_solver {
  HTGX_SOLVER_PLAIN_FP64 = 0,    /* one path per thread, FP64 everywhere */
  HTGX_SOLVER_WAVEFRONT_FP32 = 1 /* FP32 transport, FP64 Tmin + accumulators */
};

enum htgx_ground_type {
  HTGX_GROUND_NONE = 0,  /* no ground geometry */
  HTGX_GROUND_PLANE = 1, /* horizontal plane z = plane_z, i.e. a flat quad
                            repeated to infinity with -R */
  HTGX_GROUND_MESH = 2   /* triangle mesh traced through a GPU BVH */
};

/*******************************************************************************
 * Output pixel formats (bit-compatible with the reference)
 ******************************************************************************/
/* struct htrdr_accum, src/core/htrdr_accum.h:31-36 */
struct htgx_accum {
  double sum_weights;
  double sum_weights_sqr;
  size_t nweights;
  size_t nfailures;
};

/* struct htrdr_estimate, src/core/htrdr_accum.h:41-44 */
struct htgx_estimate {
  double E;
  double SE;
};

/* struct atmosphere_pixel_xwave, src/atmosphere/htrdr_atmosphere_c.h:46-50 */
struct htgx_pixel_xwave {
  struct htgx_estimate radiance; /* W/m^2/sr (W/m^2/sr/nm if monochromatic) */
  struct htgx_estimate radiance_temperature; /* K, longwave only */
  struct htgx_accum time; /* microseconds per realisation */
};

/* struct atmosphere_pixel_flux, src/atmosphere/htrdr_atmosphere_c.h:59-62 */
struct htgx_pixel_flux {
  struct htgx_accum flux; /* W/m^2 */
  struct htgx_accum time; /* microseconds per realisation */
};

/* struct atmosphere_pixel_image, src/atmosphere/htrdr_atmosphere_c.h:70-75:
 * the pixel of a CIE XYZ image */
struct htgx_pixel_image {
  struct htgx_estimate X; /* W/m^2/sr */
  struct htgx_estimate Y; /* W/m^2/sr */
  struct htgx_estimate Z; /* W/m^2/sr */
  struct htgx_accum time; /* microseconds per realisation, all 3 channels */
};

/* Device-side per-pixel Monte Carlo accumulator. The weight sums are FP64
 * with Neumaier compensation; they are folded into struct htgx_accum when the
 * map is finalised. 64 bytes. */
struct htgx_mc_accum {
  double sum, sum_c;   /* sum of weights and its compensation term */
  double sum2, sum2_c; /* sum of squared weights and its compensation */
  double tsum, tsum2;  /* per-realisation time (us) and its square */
  uint64_t n;          /* accepted realisations */
  uint64_t nfail;      /* rejected samples (flux map, inside geometry) */
};

/*******************************************************************************
 * Scene description (what htsky + the ground + the sun provide in the
 * reference; src/atmosphere/htrdr_atmosphere.c:288-446)
 ******************************************************************************/
/* Spectral bands and correlated-k quadrature: htsky_find_spectral_band and
 * htsky_spectral_band_sample_quadrature (call site ..._draw_map.c:383-384) */
struct htgx_spectral_desc {
  uint32_t nbands;
  const double* band_lo;      /* [nbands] nm, sorted and contiguous */
  const double* band_hi;      /* [nbands] nm */
  const uint32_t* nquads;     /* [nbands] #quadrature points per band */
  const double* quad_weights; /* [sum nquads] weights, summing to 1 per band */
};

/* Gas: horizontally infinite, piecewise constant in uniform layers along Z */
struct htgx_gas_desc {
  uint32_t nlayers;
  double z_bottom;        /* m */
  double layer_thickness; /* m */
  const double* ka;       /* [sum nquads][nlayers] m^-1, per (band, quad) */
  const double* ks;       /* [nbands][nlayers] m^-1 */
  const double* temperature; /* [nlayers] K; may be NULL for shortwave */
};

/* Clouds: liquid water content on a regular grid. Per band, the particle
 * extinction is kext = cext * LWC with single scattering albedo ssa and a
 * Henyey-Greenstein phase function of asymmetry g. */
struct htgx_cloud_desc {
  uint32_t def[3];   /* #cells along X, Y, Z; 0 => no clouds */
  double lower[3];   /* m */
  double upper[3];   /* m */
  const float* lwc;  /* [def[2]][def[1]][def[0]] g/m^3, X fastest */
  const double* cext; /* [nbands] m^2/g */
  const double* ssa;  /* [nbands] */
  const double* hg_g; /* [nbands] */
};

/* One entry of a material's spectral BRDF list (an mrumtl file: a BRDF per
 * wavelength or per band, src/core/htrdr_materials.c:415-463) */
enum htgx_brdf_type {
  HTGX_BRDF_LAMBERTIAN = 0, /* ssf_lambertian_reflection */
  HTGX_BRDF_SPECULAR = 1    /* ssf_specular_reflection, constant Fresnel term */
};

enum htgx_brdf_list {
  HTGX_BRDF_LIST_WAVELENGTHS = 0, /* entries at wavelengths, strictly increasing */
  HTGX_BRDF_LIST_BANDS = 1        /* entries over [lo, hi] bands, increasing,
                                     not overlapping (they may touch) */
};

struct htgx_brdf {
  double wlen[2];      /* nm: [w, w] in a wavelength list, [lo, hi] in a band list */
  int type;            /* enum htgx_brdf_type */
  double reflectivity; /* in [0, 1] */
};

/* Ground. The plane ground is one material; a mesh carries one material index
 * per triangle. By default materials are grey Lambertian reflectors with an
 * albedo and a temperature (the emission temperature used by the longwave
 * walk, ..._lw.c:266-270), seen the same from both sides of a surface.
 *
 * Materials as htrdr-materials, mrumtl and htrdr-obj files give them
 * (src/core/htrdr_materials.c, src/core/htrdr_geometry.c:177-566) use the
 * optional fields at the end, each left zero for the default:
 *  - spectral BRDFs: with mtl_brdf_first set, material m reflects with the
 *    entries brdfs[mtl_brdf_first[m] .. mtl_brdf_first[m + 1]), a list of
 *    kind mtl_brdf_list[m], and mtl_albedo is ignored. At wavelength lambda
 *    the list gives the entry of the band holding lambda, or one of the two
 *    entries around lambda, the upper one with the probability of lambda's
 *    linear position between them (mrumtl_fetch_brdf with a canonical,
 *    src/core/htrdr_materials.c:430-433); the first or last entry outside
 *    the list. A material with no entry is semi-transparent (`none', e.g.
 *    the air);
 *  - interfaces: with `interfaces' set, tri_mtl[t] and plane_mtl index
 *    interfaces, not materials. Interface i is the material of its front
 *    side interfaces[3i], of its back side interfaces[3i+1], and a thin
 *    material interfaces[3i+2] (-1: none). The front side of triangle
 *    (v0, v1, v2) is the side from which its vertices turn clockwise, the
 *    one (v2 - v0) x (v1 - v0) points to (htrdr-obj(5)); the plane's front
 *    side is below it. A hit from one side reflects with the material of
 *    the other side, or with the thin material if it has a BRDF; if the
 *    material of the other side has none, with the material of the hit side
 *    (src/core/htrdr_interface.h:45-87). At least one of the three must
 *    have a BRDF. Without `interfaces', material m is the interface whose
 *    front is m and whose back is the sky;
 *  - sky_mtl: with `interfaces' set, the material of the sky (option -n).
 *    A rectangle sensor's sample whose up ray meets the ground from a side
 *    whose material is not the sky is rejected (..._draw_map.c:75-100). */
struct htgx_ground_desc {
  int type;              /* enum htgx_ground_type */
  double plane_z;        /* HTGX_GROUND_PLANE */
  uint32_t plane_mtl;    /* HTGX_GROUND_PLANE material (or interface) index */
  uint32_t nverts;       /* HTGX_GROUND_MESH */
  uint32_t ntris;
  const float* verts;    /* [nverts][3] m */
  const uint32_t* tris;  /* [ntris][3] */
  const uint32_t* tri_mtl; /* [ntris] material (or interface) index */
  uint32_t nmtls;
  const double* mtl_albedo;      /* [nmtls]; may be NULL with mtl_brdf_first */
  const double* mtl_temperature; /* [nmtls] K; may be NULL (0 K) */
  /* spectral BRDFs */
  const uint32_t* mtl_brdf_first; /* [nmtls + 1], mtl_brdf_first[0] = 0 */
  const int32_t* mtl_brdf_list;   /* [nmtls] enum htgx_brdf_list */
  const struct htgx_brdf* brdfs;  /* [mtl_brdf_first[nmtls]] */
  /* interfaces */
  uint32_t ninterfaces;
  const int32_t* interfaces;      /* [ninterfaces][3] front, back, thin */
  int32_t sky_mtl;                /* -1: no material is the sky */
};

struct htgx_scene_desc {
  struct htgx_spectral_desc spectral;
  struct htgx_gas_desc gas;
  struct htgx_cloud_desc clouds;
  struct htgx_ground_desc ground;
  double sun_azimuth;   /* degrees, option -D */
  double sun_elevation; /* degrees, option -D */
  double optical_thickness; /* octree merge threshold, option -T */
  uint32_t leaf_block;  /* edge (in cells) of the finest octree leaf block:
                           1, 2, 4, ...; the cloud definition must be a
                           multiple of it */
  int repeat_clouds;    /* option -r: the cloud grid repeats along X and Y
                           with the period of its extent (htsky's
                           repeat_clouds, src/atmosphere/htrdr_atmosphere.c:375) */
  int repeat_ground;    /* option -R: a mesh ground repeats along X and Y with
                           the period of its bounding box, traced copy after
                           copy as htrdr_atmosphere_ground_trace_ray does
                           (src/atmosphere/htrdr_atmosphere_ground.c:197-210);
                           the plane ground is infinite already */
};

/*******************************************************************************
 * Map description: what htrdr_draw_map_args + the atmosphere command state
 * carry for one image or flux map (src/core/htrdr_draw_map.h:58-63,
 * src/atmosphere/htrdr_atmosphere_c.h:95-126)
 ******************************************************************************/
struct htgx_map_desc {
  uint32_t width, height; /* image definition, option -i def=WxH */
  uint32_t spp;           /* realisations per pixel in this call */
  uint32_t spp_offset;    /* index of the first realisation of this call */
  int sensor;             /* enum htgx_sensor_type */
  double pos[3], tgt[3], up[3]; /* camera (-C) or rectangle (-p) frame */
  double fov_y;           /* camera vertical field of view, degrees */
  double rect_size[2];    /* rectangle size, m */
  int spectral_type;      /* enum htgx_spectral_type */
  double wlen_range[2];   /* nm, option -s sw=/lw= */
  double ref_temperature; /* K, option -s Tref= */
  uint64_t seed;          /* key of the Threefry streams */
  int solver;             /* enum htgx_solver */
  /* camera projection (zero-initialised: a pinhole perspective camera) */
  int projection;         /* enum htgx_projection */
  double lens_radius;     /* perspective: lens radius, m; 0 = pinhole (-C lens-radius) */
  double focal_dst;       /* perspective with a lens: distance in focus, m (-C focal-dst) */
  double ortho_height;    /* orthographic: height of the image plane, m (-P height) */
};

/*******************************************************************************
 * Context: device copy of the scene (SVO, LWC grid, gas tables, spectral
 * tables, ground) on the current HIP device. One context per GPU; not
 * re-entrant.
 ******************************************************************************/
struct htgx_ctx;

/* Upload the scene to the current device and build the octree there.
 * Replaces the htsky_create + ground/sun setup of htrdr_atmosphere_create
 * (src/atmosphere/htrdr_atmosphere.c:348-382). */
int htgx_ctx_create(const struct htgx_scene_desc* scene, struct htgx_ctx** ctx);

/* htgx_ctx_create with the octree cache of option -O (htsky_args
 * cache_filename, src/atmosphere/htrdr_atmosphere.c:369;
 * doc/htrdr-atmosphere.1.in:283-309). If cache_path does not exist, the
 * octree is built and written there. If it exists and was built from the
 * same atmosphere, cloud and Mie data (spectral bands and quadrature, gas
 * layers, cloud grid and LWC, cloud optics), the octree is read from it
 * instead, with the merge threshold and leaf block it was built with
 * (scene->optical_thickness and leaf_block are then ignored, as the reference
 * ignores -T and -V). A cache built from other data, or not a cache, is
 * refused with HTGX_BAD_ARG. *loaded (may be NULL) says whether the cache was
 * read; err (may be NULL) receives a message on failure. A scene without
 * clouds neither reads nor writes the file. */
int htgx_ctx_create_cached
  (const struct htgx_scene_desc* scene, const char* cache_path, int* loaded,
   char* err, size_t err_len, struct htgx_ctx** ctx);
int htgx_ctx_destroy(struct htgx_ctx* ctx);

/* Number of nodes, leaves and octree levels (diagnostics). */
int htgx_ctx_get_svo_info
  (const struct htgx_ctx* ctx,
   uint32_t* nlevels,
   uint64_t* nnodes,
   uint64_t* nleaves);

/* Option -d of htrdr-atmosphere (dump_volumetric_acceleration_structure,
 * src/atmosphere/htrdr_atmosphere.c:228-255, calling htsky_dump_cloud_vtk per
 * band and k-point): writes to the file descriptor fd, for every band and
 * k-point in order, the octree over the clouds as a legacy VTK file, each
 * followed by a line `---'. A tree is the unstructured grid of its leaves
 * (VTK_VOXEL cells, corners in metres) with the cell scalars kext_min and
 * kext_max (m^-1): the cloud extinction cext * LWC bounds of the leaf plus
 * the bounds of the gas extinction over the layers the leaf overlaps. This
 * build keeps one tree of LWC bounds for all bands; the trees written are
 * its per-band images. htsky 0.3 is not in the reference, so this is the
 * content the option documents (doc/htrdr-atmosphere.1.in:222-230), not
 * htsky's byte layout. Host only (no device involved); HTGX_BAD_OP for a
 * scene without clouds, HTGX_IO_ERR if a write fails. */
int htgx_dump_octrees_vtk(const struct htgx_scene_desc* scene, int fd);

/* Replaces htrdr_draw_map (src/core/htrdr_draw_map.c:497-621) as called by
 * atmosphere_draw_map (src/atmosphere/htrdr_atmosphere_draw_map.c:593-616):
 * renders desc->spp realisations per pixel and writes one pixel struct per
 * pixel, row-major, into the caller's HOST memory `pixels' whose rows are
 * `pitch' bytes apart. The pixel struct is htgx_pixel_image for a CIE XYZ
 * camera (spp realisations per channel, draw_pixel_image, ..._draw_map.c:
 * 110-207), htgx_pixel_xwave for another camera and htgx_pixel_flux for a
 * rectangle. Returns HTGX_BAD_ARG on the conditions
 * of check_draw_map_args (src/core/htrdr_draw_map.c:66-79): null spp,
 * null definition, pitch smaller than a row. */
int htgx_draw_map
  (struct htgx_ctx* ctx,
   const struct htgx_map_desc* desc,
   void* pixels,
   size_t pitch);

/* Device-resident variant for callers that keep the accumulators in HBM and
 * render in several calls (realisation slices [spp_offset, spp_offset+spp)).
 * `accum' is a DEVICE pointer to width*height struct htgx_mc_accum, row-major
 * (for a CIE XYZ image: 3 such planes, X then Y then Z; channel c draws from
 * streams keyed by seed ^ (c+1) << 56); `stream' is a hipStream_t (NULL =
 * default stream). Asynchronous. */
int htgx_render_accum
  (struct htgx_ctx* ctx,
   const struct htgx_map_desc* desc,
   struct htgx_mc_accum* accum,
   void* stream);

/* htgx_render_accum on the FP32 wavefront solver without any allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Its chunk counters and partial sums live
 * in a static device array of 64 MiB that each GPU holds from the first FP32
 * set-up of a context on it until the process ends; a map needing more
 * chunk partials than it holds ((64 MiB - 256 B) / 64 B, one per pixel and
 * chunk of the context's chunk-size launch parameter) returns HTGX_BAD_ARG.
 * The context's FP32 tables must exist (built by its first FP32 map or by
 * htgx_get_wavefront_params), and a band or CIE map must use the wavelength
 * tables of the context's last map (same range and reference temperature):
 * otherwise HTGX_BAD_OP and nothing is allocated. desc->solver must be
 * HTGX_SOLVER_WAVEFRONT_FP32. Same arguments and, with the same launch
 * parameters, the same accumulators bit for bit as htgx_render_accum. Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU race
 * (serialise them, e.g. on one stream). */
int htgx_render_accum_noalloc
  (struct htgx_ctx* ctx,
   const struct htgx_map_desc* desc,
   struct htgx_mc_accum* accum,
   void* stream);

/* Launch parameters of the FP32 wavefront solver of a context. No
 * counterpart in the reference, whose tile dispatch has no such knobs. They
 * change the schedule only: the realisations, and so the estimates, are the
 * same for any values; the chunk size changes how partial sums are grouped,
 * hence the last bits of the accumulators. */
struct htgx_wf_params {
  uint32_t ev_thresh; /* lanes of a 64-lane wavefront that must wait for the
                         event step before it runs (1-64; 0 = default 24) */
  uint32_t chunk;     /* realisations per work-queue chunk (1-16; 0 = default 4) */
  uint32_t grid;      /* blocks of 256 lanes of the persistent launch
                         (1-grid_max; 0 = grid_max) */
  uint32_t grid_max;  /* read only: blocks resident on the device at once */
  uint32_t seg_thresh; /* lanes of a 64-lane wavefront that must wait for the
                          segment step before it runs while other lanes can
                          take a collision step (0-64; 0 = run it whenever a
                          lane needs it) */
};

int htgx_get_wavefront_params(struct htgx_ctx* ctx, struct htgx_wf_params* params);
int htgx_set_wavefront_params(struct htgx_ctx* ctx, const struct htgx_wf_params* params);

/* Autotuner of the parameters above for a map: trial launches of the map as
 * described (desc->spp realisations per pixel: pass the launch size that
 * will be used), timed on the device, by coordinate descent (chunk size,
 * then event threshold, then segment threshold, then both thresholds again
 * while they move, then grid) within about budget_ms of device time;
 * each candidate is timed twice, and must be more than 1% faster to be
 * taken. The best
 * parameters are kept by the context for the maps that follow; best, rate
 * (realisations/s of the best trial) and ntrials are optional outputs.
 * Synchronous. */
int htgx_autotune
  (struct htgx_ctx* ctx,
   const struct htgx_map_desc* desc,
   double budget_ms,
   struct htgx_wf_params* best,
   double* rate,
   uint32_t* ntrials);

/* Batched rendering. No counterpart in the reference, which draws one map
 * per process: several maps of distinct scenes in one persistent FP32
 * launch, so that maps of uneven sizes share the device and their tails
 * overlap. Each item renders exactly what htgx_render_accum renders for it
 * (same random streams, same chunk order: bit-identical accumulators). */
struct htgx_batch_item {
  struct htgx_ctx* ctx;              /* distinct contexts on the current device */
  const struct htgx_map_desc* desc;  /* same sensor and spectral type for all, not
                                        CIE XYZ (HTGX_BAD_OP); the solver field
                                        is ignored (FP32 wavefront) */
  struct htgx_mc_accum* accum;       /* DEVICE accumulators of the map, as in
                                        htgx_render_accum */
};

/* Render n items in one launch on `stream' (a hipStream_t, NULL = default);
 * returns once they are done. scene_ms (optional, HOST array of n): time from
 * the start of the launch to the completion of each item, in ms. */
int htgx_render_batch
  (const struct htgx_batch_item* items,
   uint32_t n,
   void* stream,
   double* scene_ms);

/* A batch prepared once and rendered in repeated steps (no counterpart in the
 * reference). Step s renders, for each item, the realisations
 * [desc->spp_offset + s*spp, desc->spp_offset + (s+1)*spp) of every pixel
 * and accumulates them into the item's accumulators: bit-identical to
 * htgx_render_accum with that spp_offset. A plan of one item runs the
 * single-map kernel of htgx_render_accum (it reads the step's offset from
 * the plan's table), more items the batch kernel. A step is six launches on
 * `stream', with no allocation, host copy or synchronisation; with use_graph
 * they are recorded once, at creation, as a HIP graph that each step
 * replays. The plan keeps its own copy of the descriptors, chunk queues and
 * partial sums: the contexts may draw other maps between steps, but must
 * outlive the plan; steps of one plan must not overlap (same stream, or
 * synchronised). */
struct htgx_batch_plan;

int htgx_batch_plan_create
  (const struct htgx_batch_item* items,
   uint32_t n,
   int use_graph,
   struct htgx_batch_plan** plan);

/* The autotuner for a batch: the same search, each trial one batched launch
 * of all items at 1/spp_div of their realisations per pixel (at least 4).
 * Small trials (large spp_div) time mostly the tails of the maps, not their
 * steady state: 8 represents a launch of whole maps. The best parameters are
 * set on every item's context; call it before htgx_batch_plan_create, whose
 * plan keeps the parameters it was built with. rate counts the trial's
 * realisations. Synchronous. */
int htgx_batch_autotune
  (const struct htgx_batch_item* items,
   uint32_t n,
   uint32_t spp_div,
   double budget_ms,
   struct htgx_wf_params* best,
   double* rate,
   uint32_t* ntrials);

/* Enqueue the next step on `stream' (asynchronous), on the plan's device */
int htgx_batch_plan_step(struct htgx_batch_plan* plan, void* stream);

/* Back to step 0 (same realisations again; completion times restart).
 * Synchronises the device. */
int htgx_batch_plan_rewind(struct htgx_batch_plan* plan);

/* Completion time of each item in its last step, in ms after the start of
 * the first step since creation or rewind (HOST array of n, in item order).
 * Synchronises the device. */
int htgx_batch_plan_scene_ms(struct htgx_batch_plan* plan, double* scene_ms);

/* Waits for the device, then frees the plan */
int htgx_batch_plan_destroy(struct htgx_batch_plan* plan);

/* Fold device accumulators into reference pixel structs (host memory). For
 * a longwave camera, the brightness temperature and its error are computed as
 * in draw_pixel_xwave (src/atmosphere/htrdr_atmosphere_draw_map.c:423-435);
 * a CIE XYZ image takes its 3 accumulator planes. */
int htgx_finalize_pixels
  (const struct htgx_map_desc* desc,
   const struct htgx_mc_accum* accum_host, /* HOST copy of the accumulators */
   void* pixels,
   size_t pitch);

/* Counters of the last htgx_render_accum call on this context (summed over
 * all realisations): leaf visits, tentative collisions, ground-BVH nodes and
 * triangles tested, and the number of realisations. Used for the algorithmic
 * byte model of the roofline. Only maintained when counting is enabled. */
struct htgx_counters {
  uint64_t nrealisations;
  uint64_t nleaf;   /* octree leaf x gas layer segments visited */
  uint64_t ncoll;   /* tentative collisions (property fetches) */
  uint64_t nbvh;    /* BVH nodes visited (FP32 solver only) */
  uint64_t ntri;    /* triangles tested (FP32 solver only) */
  uint64_t nevents; /* scattering / surface events */
  uint64_t nseg;    /* segments handed to the null-collision loops */
  uint64_t nlwc;    /* cloud LWC fetches (4 B each) */
  /* Wavefront solver only: [0] loop iterations summed over wavefronts,
   * [1..6] for its collision, segment and event branches the number of
   * wavefront executions and the active lanes summed over them, [7..15]
   * shader clock cycles summed over wavefronts spent in: segment step,
   * collision step, chunk fetch, event draws, realisation start, event
   * consumption, end of walk, end of realisation, query set-up. */
  uint64_t diag[16];
  /* FP32 solver, maintained whether or not counting is enabled and summed
   * over every map of the context since its creation: queries cut short by
   * the budget of 2^20 tentative collisions (a majorant optical depth above
   * about 10^6), and by the segment iterator's bound on consecutive passes
   * that emit no segment (an iterator making no progress). Either ends the
   * query as if it had left the medium, which would bias the estimate: a
   * non-zero count says so. */
  uint64_t ntrunc_coll;
  uint64_t ntrunc_iter;
};
int htgx_set_counting(struct htgx_ctx* ctx, int enable);
int htgx_get_counters(struct htgx_ctx* ctx, struct htgx_counters* counters);

/*******************************************************************************
 * Test entry points (device execution of single building blocks so that they
 * can be compared with the CPU oracle on identical inputs)
 ******************************************************************************/
/* One segment of the front-to-back traversal: the intersection of an octree
 * leaf (or the cloud-free region) with a gas layer, as handed to the hit
 * filters of ..._sw.c:74,148 and ..._lw.c:65 (svx_hit.distance[0..1] plus the
 * voxel bounds fetched at ..._sw.c:163-166). */
struct htgx_segment {
  double t0, t1;     /* distances along the ray */
  double kmin, kmax; /* bounds of the requested property on the segment */
  int64_t leaf;      /* global id of the octree leaf, -1 outside the clouds */
  int32_t layer;     /* gas layer index, -1 outside the gas */
  int32_t pad;
};

/* Trace nrays rays (org/dir/range packed as 8 doubles per ray) and write at
 * most max_segs segments per ray to segs[iray*max_segs ...] and the count to
 * nsegs[iray] (may exceed max_segs: then only the first max_segs are
 * written). All pointers are HOST memory. */
int htgx_debug_trace_segments
  (struct htgx_ctx* ctx,
   const double* rays, /* [nrays][8] = org[3], dir[3], range[2] */
   uint32_t nrays,
   int prop, uint32_t iband, uint32_t iquad,
   struct htgx_segment* segs,
   uint32_t max_segs,
   uint32_t* nsegs);

/* The same segments as the FP32 wavefront solver's iterator sees them:
 * FP32 origin, direction and distances, the leaf bounds of its fused leaf
 * records (FP16 multiples of the grid's largest LWC, rounded outwards, so
 * that kmin never exceeds and kmax never falls below the exact bounds up to
 * the FP32 rounding of the gas term). Queries start at their origin:
 * range[0] must be 0. Same arguments as htgx_debug_trace_segments. */
int htgx_debug_trace_segments_fp32
  (struct htgx_ctx* ctx,
   const double* rays,
   uint32_t nrays,
   int prop, uint32_t iband, uint32_t iquad,
   struct htgx_segment* segs,
   uint32_t max_segs,
   uint32_t* nsegs);

/* Weight of individual realisations (pixel index y*W+x, realisation index),
 * computed by the solver desc->solver with the random streams of
 * htgx_draw_map: the plain FP64 solver in one thread per realisation, the
 * FP32 wavefront solver in its persistent loop fed from the list (for the
 * FP32 solver the realisation index is taken modulo 2^32, as its streams are
 * keyed). For a CIE XYZ image, bits 32-33 of the realisation index select
 * the channel (0 X, 1 Y, 2 Z). A rejected realisation (a flux sample under
 * the ground) is NaN. HOST arrays. */
int htgx_debug_realisations
  (struct htgx_ctx* ctx,
   const struct htgx_map_desc* desc,
   const uint64_t* pixel_ids,
   const uint64_t* realisation_ids,
   uint32_t n,
   double* weights);

/* The octree cache of option -O prepared on the host, without a device:
 * if path holds a cache it is checked against the scene as
 * htgx_ctx_create_cached checks it (same inputs; then its structure and its
 * bounds against the scene's LWC) and *loaded = 1; if path does not exist
 * the octree is built on the host (the same tree as the device build, bit
 * for bit) and written there, *loaded = 0. HTGX_BAD_ARG with a message in
 * err (errlen bytes, may be NULL) for a cache of other inputs, a damaged or
 * hand-made one whose tree is invalid, or a cache that does not fit the
 * cloud grid. A scene without clouds neither reads nor writes the file.
 * No counterpart in the reference, where htsky reads its cache inside
 * htsky_create (src/atmosphere/htrdr_atmosphere.c:369): it lets a cache be
 * made or checked on a machine without the GPU. */
int htgx_octree_cache_prepare
  (const struct htgx_scene_desc* scene, const char* path, int* loaded, char* err, size_t errlen);

/* Leaf block for option -V x,y,z, the maximum definition of the
 * acceleration structure (src/atmosphere/htrdr_atmosphere_args.c:49-78,
 * passed to htsky at src/atmosphere/htrdr_atmosphere.c:370-372; "by default
 * the finest definition is that of clouds", doc/htrdr-atmosphere.1.in:
 * 414-416). The octree's finest level is the grid of B x B x B blocks of
 * cloud cells, so the finest definition along an axis is def / B; this
 * returns the smallest B that divides the cloud definition def along every
 * axis and keeps def / B <= max_def. HTGX_BAD_ARG if a definition is 0 or
 * if no such block exists. */
int htgx_leaf_block_for_definition
  (const uint32_t def[3], const uint32_t max_def[3], uint32_t* leaf_block);

/* The k_min/k_max octree of a scene's clouds built on the host (svo_build)
 * or on the current device (the build htgx_ctx_create uses), for comparing
 * the two: *nnodes node bounds (LWC min, max as 2 floats, level by level)
 * into bounds and *nblocks level-map words (level << 28 | leaf id, one per
 * finest block) into lmap; *nleaves leaves; *build_ms the build's time
 * (device: kernels only, the LWC grid already uploaded). With lmap or bounds
 * NULL only the sizes are returned. HTGX_BAD_OP for a scene without clouds.
 * HOST arrays. */
int htgx_debug_octree
  (const struct htgx_scene_desc* scene, int on_device, uint32_t* lmap, float* bounds,
   uint64_t* nblocks, uint64_t* nnodes, uint64_t* nleaves, double* build_ms);

/* Threefry-4x64-20 blocks computed on the device: out[i][0..3] =
 * threefry(ctr[i][0..3], key[i][0..3]). HOST arrays. */
int htgx_debug_threefry
  (const uint64_t* ctr, const uint64_t* key, uint32_t n, uint64_t* out);

/* Threefry-2x32-20 blocks computed on the device (the generator of the FP32
 * solver): out[i][0..1] = threefry(ctr[i][0..1], key[i][0..1]). HOST arrays. */
int htgx_debug_threefry2x32
  (const uint32_t* ctr, const uint32_t* key, uint32_t n, uint32_t* out);

/* Direction samplers on the device (phase functions, sun cone, cosine
 * hemisphere). kind: 0 = HG(param) relative to dir_in, 1 = Rayleigh relative
 * to dir_in, 2 = sun cone (param unused), 3 = cosine hemisphere around dir_in.
 * u = [n][2] canonical numbers, out = [n][3]. For the sun cone, the scene sun
 * of ctx is used. HOST arrays. */
int htgx_debug_sample_dirs
  (struct htgx_ctx* ctx,
   int kind, double param, const double dir_in[3],
   const double* u, uint32_t n, double* out);

/* Closest (or, with `any', first accepted) hits of n rays on the ground mesh
 * of `scene' through its BVH, with the self-hit filter of
 * src/core/htrdr_geometry.c:96-150 (replaces htrdr_geometry_trace_ray,
 * :686-718). rays = [n][8] org, dir, range; prev[i] = input index of the
 * triangle the ray leaves (-1: none) and prev_edge[i] whether that hit was on
 * an edge. With ctx NULL the BVH is built and walked on the host by the same
 * code; else ctx must hold this scene and its device BVH is walked on the
 * GPU. Outputs: t (INFINITY: no hit), prim (input triangle index or -1) and
 * the number of BVH nodes visited. HOST arrays. */
int htgx_debug_mesh_trace
  (struct htgx_ctx* ctx, const struct htgx_scene_desc* scene,
   const double* rays, const int64_t* prev, const int32_t* prev_edge,
   uint32_t n, int any, double* t, int64_t* prim, uint32_t* nvisited);

/* Copy nbytes (multiple of 16, 16-byte aligned DEVICE pointers) with a plain
 * streaming kernel on `stream': the known-size reference used to calibrate
 * the HBM traffic counters of the profiler. Asynchronous. */
int htgx_debug_copy(void* dst, const void* src, size_t nbytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HTGX_H */
