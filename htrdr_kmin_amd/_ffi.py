"""ctypes mirror of include/htgx.h and include/htgx_synth.h.

The product is the C-ABI library ``libhtgx.so`` built in-tree by
``__graft_entry__.build()``; this module only declares its types and entry
points. Loading fails loudly when the library is missing: there is no CPU
fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhtgx.so")

# enum htgx_res (rsys res_T values)
OK, BAD_ARG, MEM_ERR, IO_ERR, UNKNOWN_ERR, BAD_OP = range(6)
RES_NAMES = {0: "OK", 1: "BAD_ARG", 2: "MEM_ERR", 3: "IO_ERR", 4: "UNKNOWN_ERR", 5: "BAD_OP"}

SPECTRAL_LW, SPECTRAL_SW, SPECTRAL_SW_CIE_XYZ = 0, 1, 2
PROJECTION_PERSPECTIVE, PROJECTION_ORTHOGRAPHIC = 0, 1
RADIANCE_DIRECT, RADIANCE_DIFFUSE, RADIANCE_ALL = 1, 2, 3
Ka, Ks, Kext = 0, 1, 2
SENSOR_CAMERA, SENSOR_RECTANGLE = 0, 1
SOLVER_PLAIN_FP64, SOLVER_WAVEFRONT_FP32 = 0, 1
GROUND_NONE, GROUND_PLANE, GROUND_MESH = 0, 1, 2
BRDF_LAMBERTIAN, BRDF_SPECULAR = 0, 1
BRDF_LIST_WAVELENGTHS, BRDF_LIST_BANDS = 0, 1

c_double_p = C.POINTER(C.c_double)
c_uint32_p = C.POINTER(C.c_uint32)
c_uint64_p = C.POINTER(C.c_uint64)
c_float_p = C.POINTER(C.c_float)


class Accum(C.Structure):  # struct htrdr_accum
    _fields_ = [("sum_weights", C.c_double), ("sum_weights_sqr", C.c_double),
                ("nweights", C.c_size_t), ("nfailures", C.c_size_t)]


class Estimate(C.Structure):  # struct htrdr_estimate
    _fields_ = [("E", C.c_double), ("SE", C.c_double)]


class PixelXwave(C.Structure):  # struct atmosphere_pixel_xwave
    _fields_ = [("radiance", Estimate), ("radiance_temperature", Estimate), ("time", Accum)]


class PixelFlux(C.Structure):  # struct atmosphere_pixel_flux
    _fields_ = [("flux", Accum), ("time", Accum)]


class PixelImage(C.Structure):  # struct atmosphere_pixel_image (CIE XYZ)
    _fields_ = [("X", Estimate), ("Y", Estimate), ("Z", Estimate), ("time", Accum)]


class McAccum(C.Structure):
    _fields_ = [("sum", C.c_double), ("sum_c", C.c_double), ("sum2", C.c_double),
                ("sum2_c", C.c_double), ("tsum", C.c_double), ("tsum2", C.c_double),
                ("n", C.c_uint64), ("nfail", C.c_uint64)]


class SpectralDesc(C.Structure):
    _fields_ = [("nbands", C.c_uint32), ("band_lo", c_double_p), ("band_hi", c_double_p),
                ("nquads", c_uint32_p), ("quad_weights", c_double_p)]


class GasDesc(C.Structure):
    _fields_ = [("nlayers", C.c_uint32), ("z_bottom", C.c_double),
                ("layer_thickness", C.c_double), ("ka", c_double_p), ("ks", c_double_p),
                ("temperature", c_double_p)]


class CloudDesc(C.Structure):
    _fields_ = [("def_", C.c_uint32 * 3), ("lower", C.c_double * 3), ("upper", C.c_double * 3),
                ("lwc", c_float_p), ("cext", c_double_p), ("ssa", c_double_p),
                ("hg_g", c_double_p)]


class Brdf(C.Structure):
    _fields_ = [("wlen", C.c_double * 2), ("type", C.c_int), ("reflectivity", C.c_double)]


class GroundDesc(C.Structure):
    _fields_ = [("type", C.c_int), ("plane_z", C.c_double), ("plane_mtl", C.c_uint32),
                ("nverts", C.c_uint32), ("ntris", C.c_uint32), ("verts", c_float_p),
                ("tris", c_uint32_p), ("tri_mtl", c_uint32_p), ("nmtls", C.c_uint32),
                ("mtl_albedo", c_double_p), ("mtl_temperature", c_double_p),
                ("mtl_brdf_first", c_uint32_p), ("mtl_brdf_list", C.POINTER(C.c_int32)),
                ("brdfs", C.POINTER(Brdf)), ("ninterfaces", C.c_uint32),
                ("interfaces", C.POINTER(C.c_int32)), ("sky_mtl", C.c_int32)]


class SceneDesc(C.Structure):
    _fields_ = [("spectral", SpectralDesc), ("gas", GasDesc), ("clouds", CloudDesc),
                ("ground", GroundDesc), ("sun_azimuth", C.c_double),
                ("sun_elevation", C.c_double), ("optical_thickness", C.c_double),
                ("leaf_block", C.c_uint32), ("repeat_clouds", C.c_int),
                ("repeat_ground", C.c_int)]


class MapDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("spp_offset", C.c_uint32), ("sensor", C.c_int), ("pos", C.c_double * 3),
                ("tgt", C.c_double * 3), ("up", C.c_double * 3), ("fov_y", C.c_double),
                ("rect_size", C.c_double * 2), ("spectral_type", C.c_int),
                ("wlen_range", C.c_double * 2), ("ref_temperature", C.c_double),
                ("seed", C.c_uint64), ("solver", C.c_int), ("projection", C.c_int),
                ("lens_radius", C.c_double), ("focal_dst", C.c_double),
                ("ortho_height", C.c_double)]


class WfParams(C.Structure):
    _fields_ = [("ev_thresh", C.c_uint32), ("chunk", C.c_uint32), ("grid", C.c_uint32),
                ("grid_max", C.c_uint32), ("seg_thresh", C.c_uint32)]


class BatchItem(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("desc", C.POINTER(MapDesc)), ("accum", C.c_void_p)]


class Segment(C.Structure):
    _fields_ = [("t0", C.c_double), ("t1", C.c_double), ("kmin", C.c_double),
                ("kmax", C.c_double), ("leaf", C.c_int64), ("layer", C.c_int32),
                ("pad", C.c_int32)]


class Counters(C.Structure):
    _fields_ = [("nrealisations", C.c_uint64), ("nleaf", C.c_uint64), ("ncoll", C.c_uint64),
                ("nbvh", C.c_uint64), ("ntri", C.c_uint64), ("nevents", C.c_uint64),
                ("nseg", C.c_uint64), ("nlwc", C.c_uint64), ("diag", C.c_uint64 * 16),
                ("ntrunc_coll", C.c_uint64), ("ntrunc_iter", C.c_uint64)]


class SynthParams(C.Structure):
    _fields_ = [("spectral_type", C.c_int), ("cloud_def", C.c_uint32 * 3),
                ("cloud_lower", C.c_double * 3), ("cloud_upper", C.c_double * 3),
                ("lwc_mu", C.c_double), ("lwc_sigma", C.c_double), ("cover", C.c_double),
                ("smooth_sigma", C.c_double), ("cext", C.c_double), ("ssa", C.c_double),
                ("hg_g", C.c_double), ("nlayers", C.c_uint32), ("layer_thickness", C.c_double),
                ("ka0", C.c_double), ("ks0", C.c_double), ("scale_height", C.c_double),
                ("T_ground", C.c_double), ("lapse_rate", C.c_double), ("nbands", C.c_uint32),
                ("nquads", C.c_uint32), ("wlen_lo", C.c_double), ("wlen_hi", C.c_double),
                ("lw_ka_min", C.c_double), ("lw_ka_max", C.c_double), ("ground_type", C.c_int),
                ("ground_albedo", C.c_double), ("ground_temperature", C.c_double),
                ("sun_azimuth", C.c_double), ("sun_elevation", C.c_double),
                ("optical_thickness", C.c_double), ("leaf_block", C.c_uint32),
                ("seed", C.c_uint64), ("hf_res", C.c_uint32), ("hf_octaves", C.c_uint32),
                ("hf_relief", C.c_double), ("hf_hurst", C.c_double)]


# Entry points: name -> (restype, argtypes)
SIGNATURES = {
    "htgx_ctx_create": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    "htgx_ctx_create_cached": (C.c_int, [C.POINTER(SceneDesc), C.c_char_p, C.POINTER(C.c_int),
                                         C.c_char_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "htgx_ctx_destroy": (C.c_int, [C.c_void_p]),
    "htgx_ctx_get_svo_info": (C.c_int, [C.c_void_p, c_uint32_p, c_uint64_p, c_uint64_p]),
    "htgx_draw_map": (C.c_int, [C.c_void_p, C.POINTER(MapDesc), C.c_void_p, C.c_size_t]),
    "htgx_render_accum": (C.c_int, [C.c_void_p, C.POINTER(MapDesc), C.c_void_p, C.c_void_p]),
    "htgx_render_accum_noalloc": (C.c_int, [C.c_void_p, C.POINTER(MapDesc), C.c_void_p, C.c_void_p]),
    "htgx_finalize_pixels": (C.c_int, [C.POINTER(MapDesc), C.c_void_p, C.c_void_p, C.c_size_t]),
    "htgx_set_counting": (C.c_int, [C.c_void_p, C.c_int]),
    "htgx_get_counters": (C.c_int, [C.c_void_p, C.POINTER(Counters)]),
    "htgx_debug_trace_segments": (C.c_int, [C.c_void_p, c_double_p, C.c_uint32, C.c_int,
                                            C.c_uint32, C.c_uint32, C.POINTER(Segment),
                                            C.c_uint32, c_uint32_p]),
    "htgx_debug_trace_segments_fp32": (C.c_int, [C.c_void_p, c_double_p, C.c_uint32, C.c_int,
                                            C.c_uint32, C.c_uint32, C.POINTER(Segment),
                                            C.c_uint32, c_uint32_p]),
    "htgx_debug_realisations": (C.c_int, [C.c_void_p, C.POINTER(MapDesc), c_uint64_p,
                                          c_uint64_p, C.c_uint32, c_double_p]),
    "htgx_render_batch": (C.c_int, [C.POINTER(BatchItem), C.c_uint32, C.c_void_p, c_double_p]),
    "htgx_get_wavefront_params": (C.c_int, [C.c_void_p, C.POINTER(WfParams)]),
    "htgx_set_wavefront_params": (C.c_int, [C.c_void_p, C.POINTER(WfParams)]),
    "htgx_autotune": (C.c_int, [C.c_void_p, C.POINTER(MapDesc), C.c_double, C.POINTER(WfParams),
                                c_double_p, c_uint32_p]),
    "htgx_batch_autotune": (C.c_int, [C.POINTER(BatchItem), C.c_uint32, C.c_uint32, C.c_double,
                                      C.POINTER(WfParams), c_double_p, c_uint32_p]),
    "htgx_batch_plan_create": (C.c_int, [C.POINTER(BatchItem), C.c_uint32, C.c_int,
                                         C.POINTER(C.c_void_p)]),
    "htgx_batch_plan_step": (C.c_int, [C.c_void_p, C.c_void_p]),
    "htgx_batch_plan_rewind": (C.c_int, [C.c_void_p]),
    "htgx_batch_plan_scene_ms": (C.c_int, [C.c_void_p, c_double_p]),
    "htgx_batch_plan_destroy": (C.c_int, [C.c_void_p]),
    "htgx_debug_threefry": (C.c_int, [c_uint64_p, c_uint64_p, C.c_uint32, c_uint64_p]),
    "htgx_debug_threefry2x32": (C.c_int, [c_uint32_p, c_uint32_p, C.c_uint32, c_uint32_p]),
    "htgx_debug_sample_dirs": (C.c_int, [C.c_void_p, C.c_int, C.c_double, c_double_p,
                                         c_double_p, C.c_uint32, c_double_p]),
    "htgx_debug_mesh_trace": (C.c_int, [C.c_void_p, C.POINTER(SceneDesc), c_double_p,
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_uint32,
                                        C.c_int, c_double_p, C.POINTER(C.c_int64), c_uint32_p]),
    "htgx_debug_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "htgx_synth_params_preset": (C.c_int, [C.c_int, C.POINTER(SynthParams)]),
    "htgx_synth_map_preset": (C.c_int, [C.c_int, C.POINTER(MapDesc)]),
    "htgx_synth_create": (C.c_int, [C.POINTER(SynthParams), C.POINTER(C.c_void_p)]),
    "htgx_synth_get_scene": (C.POINTER(SceneDesc), [C.c_void_p]),
    "htgx_synth_destroy": (C.c_int, [C.c_void_p]),
    "htgx_ground_files_load": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "htgx_ground_files_get_desc": (C.POINTER(GroundDesc), [C.c_void_p]),
    "htgx_ground_files_material_name": (C.c_char_p, [C.c_void_p, C.c_uint32]),
    "htgx_ground_files_interface_name": (C.c_char_p, [C.c_void_p, C.c_uint32]),
    "htgx_ground_files_destroy": (C.c_int, [C.c_void_p]),
    "htgx_files_last_error": (C.c_char_p, []),
    "htgx_dump_octrees_vtk": (C.c_int, [C.POINTER(SceneDesc), C.c_int]),
    "htgx_debug_octree": (C.c_int, [C.POINTER(SceneDesc), C.c_int, c_uint32_p, c_float_p, c_uint64_p,
                                    c_uint64_p, c_uint64_p, c_double_p]),
    "htgx_octree_cache_prepare": (C.c_int, [C.POINTER(SceneDesc), C.c_char_p, C.POINTER(C.c_int),
                                            C.c_char_p, C.c_size_t]),
    "htgx_leaf_block_for_definition": (C.c_int, [c_uint32_p, c_uint32_p, c_uint32_p]),
}

_lib = None


def load():
    """Load libhtgx.so (built in-tree). Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
            "g.build()'` (there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class HtgxError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__(f"{fn} failed: {RES_NAMES.get(code, code)}")
        self.code = code


def check(fn, code):
    if code != OK:
        raise HtgxError(fn, code)
    return code
