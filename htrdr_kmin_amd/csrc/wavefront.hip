/* FP32 wavefront solver (work in progress: not yet available). */
#include "htgx_internal.h"

int htgx_render_wavefront(htgx_ctx*, const htgx::DevMap&, htgx_mc_accum*, hipStream_t)
{
  return HTGX_BAD_OP;
}

int htgx_debug_realisations_fp32(htgx_ctx*, const htgx_map_desc*, const uint64_t*,
                                 const uint64_t*, uint32_t, double*)
{
  return HTGX_BAD_OP;
}

void htgx_wavefront_release(htgx_ctx*) {}
