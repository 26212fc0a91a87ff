/* Internal state of a libhtgx.so context (not part of the C-ABI). */
#ifndef HTGX_INTERNAL_H
#define HTGX_INTERNAL_H

#include "htgx.h"
#include "htgx_device.h"

#include <hip/hip_runtime.h>
#include <string>
#include <vector>

struct htgx_wavefront_state;

/* Device counter slots: [1..5] walk counters, [6..7] BVH nodes and triangles,
 * [8..23] wavefront diagnostics */
#define HTGX_NSTATS 32

struct htgx_ctx {
  int device = 0;
  htgx::DevScene S;
  std::vector<void*> bufs;          /* scene buffers owned by the context */
  std::vector<double> band_lo, band_hi;
  std::vector<uint32_t> nquads;     /* k-points of each band */
  bool has_temperature = false;
  uint32_t nlevels = 0;
  uint64_t nnodes = 0, nleaves = 0;
  uint32_t nmtls = 0;     /* ground materials */
  uint32_t bvh_depth = 0; /* mesh ground BVH depth */
  size_t svo_bytes = 0;
  const char* cache_path = nullptr; /* option -O, during set-up only */
  bool cache_loaded = false;
  std::string cache_err;
  double us_per_tick = 0.01;
  /* per-slice partial accumulators of the plain solver */
  void* d_partials = nullptr;
  size_t partials_cap = 0;
  /* wavelength CDF of the last map (cached by range and Tref) */
  double* d_wcdf = nullptr;
  double* d_wpdf = nullptr;
  size_t w_nbands = 0;
  double w_range[2] = {0, 0};
  double w_Tref = 0;
  /* CIE XYZ channel CDFs of the last image ([3][cie_nbands], cached by range) */
  double* d_cie_cdf = nullptr;
  size_t cie_nbands = 0;
  double cie_range[2] = {0, 0};
  double cie_rcp[3] = {0, 0, 0};
  /* event counters */
  int counting = 0;
  unsigned long long* d_counters = nullptr; /* HTGX_NSTATS slots */
  uint64_t last_nrealisations = 0, last_ncell = 0, last_nseg = 0;
  /* FP32 wavefront solver state */
  htgx_wavefront_state* wf = nullptr;
};

int htgx_ctx_init(htgx_ctx* ctx, const htgx_scene_desc& scene);
/* may_allocate false: a map whose wavelength tables are not those the
 * context holds from its last map returns HTGX_BAD_OP instead of building
 * and uploading them */
int htgx_map_setup(htgx_ctx* ctx, const htgx_map_desc* desc, htgx::DevMap* map,
                   bool may_allocate = true);
/* The map of channel c (0 X, 1 Y, 2 Z) of a CIE XYZ image set up by
 * htgx_map_setup: the channel's wavelength CDF and its own random streams */
htgx::DevMap htgx_cie_channel_map(const htgx_ctx* ctx, const htgx::DevMap& M, int c);
int htgx_render_wavefront(htgx_ctx* ctx, const htgx::DevMap& map, htgx_mc_accum* accum,
                          hipStream_t stream);
/* Several maps of distinct contexts (same device, same sensor and spectral
 * type) in one persistent FP32 launch; synchronous. scene_ms[i] (optional):
 * completion time of map i after the launch started. */
int htgx_render_wavefront_batch(htgx_ctx* const* ctxs, const htgx::DevMap* Ms, htgx_mc_accum* const* accums,
                                uint32_t n, hipStream_t stream, double* scene_ms);
/* Launch parameters of the FP32 solver of a context: read (get) and/or
 * replace (set; zero fields take the defaults) */
int htgx_wavefront_params(htgx_ctx* ctx, htgx_wf_params* get, const htgx_wf_params* set);
int htgx_wavefront_autotune(htgx_ctx* ctx, const htgx::DevMap& M, double budget_ms,
                            htgx_wf_params* best, double* best_rate, uint32_t* ntrials);
int htgx_wavefront_batch_autotune(htgx_ctx* const* ctxs, const htgx::DevMap* Ms, uint32_t n,
                                  uint32_t spp_div, double budget_ms, htgx_wf_params* best,
                                  double* best_rate, uint32_t* ntrials);
/* Batch plans (htgx_batch_plan_* of htgx.h). own_partials: the plan holds
 * its own chunk partials and wavelength tables, so that the contexts may draw
 * other maps between its steps. use_graph: one step is recorded as a HIP
 * graph at build time and replayed by wf_plan_enqueue. */
int wf_plan_build(htgx_ctx* const* ctxs, const htgx::DevMap* Ms, htgx_mc_accum* const* accums,
                  uint32_t n, bool own_partials, bool use_graph, htgx_batch_plan** out);
int wf_plan_enqueue(htgx_batch_plan* p, hipStream_t stream);
int wf_plan_rewind(htgx_batch_plan* p);
int wf_plan_times(htgx_batch_plan* p, double* scene_ms);
void wf_plan_free(htgx_batch_plan* p);
int htgx_debug_realisations_fp32(htgx_ctx* ctx, const htgx_map_desc* desc,
                                 const uint64_t* pix, const uint64_t* real, uint32_t n,
                                 double* weights);
int htgx_debug_trace_segments_fp32_impl(htgx_ctx* ctx, const double* rays, uint32_t nrays, int prop,
                                        uint32_t iband, uint32_t iquad, htgx_segment* segs,
                                        uint32_t max_segs, uint32_t* nsegs);
/* Queries of the FP32 solver cut short since the context was created:
 * [0] by the tentative-collision budget, [1] by the segment iterator's bound
 * on passes without a segment */
int htgx_wavefront_truncations(htgx_ctx* ctx, uint64_t out[2]);
/* htgx_render_accum_noalloc of one map (channel) on the FP32 solver: the
 * chunk counters and partials in the library's static scratch array */
int htgx_render_wavefront_noalloc(htgx_ctx* ctx, const htgx::DevMap& map, htgx_mc_accum* accum,
                                  hipStream_t stream);
void htgx_wavefront_release(htgx_ctx* ctx);

#endif /* HTGX_INTERNAL_H */
