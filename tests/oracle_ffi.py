"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: the oracle is the CPU reference the GPU library
is compared against.
"""
import ctypes as C
import os

import numpy as np

from htrdr_kmin_amd import _ffi as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_SIG = {
    "oracle_threefry4x64_20": (None, [F.c_uint64_p, F.c_uint64_p, F.c_uint64_p]),
    "oracle_stream_canonicals": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, F.c_double_p]),
    "oracle_sky_create": (C.c_int, [C.POINTER(F.SceneDesc), C.POINTER(C.c_void_p)]),
    "oracle_sky_destroy": (None, [C.c_void_p]),
    "oracle_sky_get_svo_info": (C.c_int, [C.c_void_p, F.c_uint32_p, F.c_uint64_p, F.c_uint64_p]),
    "oracle_trace_segments": (C.c_int, [C.c_void_p, F.c_double_p, C.c_int, C.c_uint32, C.c_uint32,
                                        C.POINTER(F.Segment), C.c_uint32, F.c_uint32_p]),
    "oracle_leaf_of_cell": (C.c_int, [C.c_void_p, F.c_uint32_p, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "oracle_fetch_raw": (C.c_double, [C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint32, F.c_double_p]),
    "oracle_transmissivity": (C.c_double, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int,
                                           C.c_uint32, C.c_uint32, F.c_double_p, F.c_double_p,
                                           F.c_double_p]),
    "oracle_realisation": (C.c_double, [C.c_void_p, C.POINTER(F.MapDesc), C.c_uint64, C.c_uint64]),
    "oracle_radiance": (C.c_double, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                     F.c_double_p, F.c_double_p, C.c_double, C.c_uint32, C.c_uint32]),
    "oracle_radiance_many": (None, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int,
                                    C.c_int, F.c_double_p, F.c_double_p, C.c_double, C.c_uint32,
                                    C.c_uint32, F.c_double_p, C.c_int]),
    "oracle_draw_map": (C.c_int, [C.c_void_p, C.POINTER(F.MapDesc), C.c_void_p, C.c_int]),
    "oracle_draw_pixels": (C.c_int, [C.c_void_p, C.POINTER(F.MapDesc), F.c_uint64_p, C.c_uint64,
                                     C.c_void_p, C.c_int]),
    "oracle_planck_monochromatic": (C.c_double, [C.c_double, C.c_double]),
    "oracle_wiebelt": (C.c_double, [C.c_double]),
    "oracle_blackbody_fraction": (C.c_double, [C.c_double, C.c_double, C.c_double]),
    "oracle_planck": (C.c_double, [C.c_double, C.c_double, C.c_double]),
    "oracle_brightness_temperature": (C.c_int, [C.c_double, C.c_double, C.c_double, F.c_double_p]),
    "oracle_radiance_temperature": (C.c_double, [C.c_double, C.c_double, C.c_double]),
    "oracle_wlen_planck_create": (C.c_int, [F.c_double_p, C.c_size_t, C.c_double, C.POINTER(C.c_void_p)]),
    "oracle_wlen_planck_destroy": (None, [C.c_void_p]),
    "oracle_wlen_planck_sample": (C.c_double, [C.c_void_p, C.c_double, C.c_double, F.c_double_p]),
    "oracle_spectral_bands_count": (C.c_size_t, [F.c_double_p]),
    "oracle_sample_dirs": (C.c_int, [C.c_void_p, C.c_int, C.c_double, F.c_double_p, F.c_double_p,
                                     C.c_uint32, F.c_double_p]),
    "oracle_phase_eval": (C.c_double, [C.c_int, C.c_double, C.c_double]),
    "oracle_accum_estimate": (None, [C.c_void_p, F.c_double_p, F.c_double_p]),
    "oracle_wlen_cie_create": (C.c_int, [F.c_double_p, C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]),
    "oracle_cie_fit": (C.c_double, [C.c_int, C.c_double]),
    "oracle_ground_trace": (C.c_int, [C.c_void_p, F.c_double_p, C.c_int64, C.c_int, F.c_double_p,
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "oracle_sky_use_bvh": (C.c_int, [C.c_void_p, C.c_int]),
    "oracle_brdf_fetch": (C.c_int, [C.POINTER(F.GroundDesc), C.c_int32, C.c_double, C.c_double,
                                    F.c_double_p]),
    "oracle_hit_material": (C.c_int32, [C.POINTER(F.GroundDesc), C.c_uint32, F.c_double_p,
                                        F.c_double_p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: run __graft_entry__.build()")
        L = C.CDLL(ORACLE_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(F.c_double_p)


class Sky:
    def __init__(self, scene_ptr):
        self._h = C.c_void_p()
        r = lib().oracle_sky_create(scene_ptr, C.byref(self._h))
        if r != 0:
            raise RuntimeError(f"oracle_sky_create: {r}")

    @property
    def h(self):
        return self._h

    def svo_info(self):
        nl, nn, nf = C.c_uint32(), C.c_uint64(), C.c_uint64()
        lib().oracle_sky_get_svo_info(self._h, C.byref(nl), C.byref(nn), C.byref(nf))
        return nl.value, nn.value, nf.value

    def trace_segments(self, rays, prop, iband=0, iquad=0, max_segs=256):
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 8)
        n = rays.shape[0]
        out = np.zeros((n, max_segs), F.Segment)
        cnt = np.zeros(n, np.uint32)
        seg_t = F.Segment * max_segs
        for i in range(n):
            buf = seg_t()
            c = C.c_uint32()
            lib().oracle_trace_segments(self._h, dptr(rays[i]), prop, iband, iquad, buf, max_segs,
                                        C.byref(c))
            out[i] = np.ctypeslib.as_array(buf)
            cnt[i] = c.value
        return out, cnt

    def leaf_of_cell(self, cell):
        c = (C.c_uint32 * 3)(*cell)
        leaf, mn, mx = C.c_int64(), C.c_float(), C.c_float()
        r = lib().oracle_leaf_of_cell(self._h, c, C.byref(leaf), C.byref(mn), C.byref(mx))
        if r != 0:
            raise RuntimeError("oracle_leaf_of_cell")
        return leaf.value, mn.value, mx.value

    def realisation(self, desc, pixel, ireal):
        return lib().oracle_realisation(self._h, C.byref(desc), pixel, ireal)

    def draw_map(self, desc, nthreads=0):
        """Accumulators (height, width), or (3, height, width) for CIE XYZ"""
        from htrdr_kmin_amd.atmosphere import ACCUM_DTYPE, accum_planes
        nc = accum_planes(desc)
        out = np.zeros(nc * desc.width * desc.height, ACCUM_DTYPE)
        r = lib().oracle_draw_map(self._h, C.byref(desc), out.ctypes.data_as(C.c_void_p), nthreads)
        if r != 0:
            raise RuntimeError(f"oracle_draw_map: {r}")
        shape = (desc.height, desc.width) if nc == 1 else (nc, desc.height, desc.width)
        return out.reshape(shape)

    def draw_pixels(self, desc, pixels, nthreads=0):
        """Accumulators (npixels,), or (3, npixels) for CIE XYZ"""
        from htrdr_kmin_amd.atmosphere import ACCUM_DTYPE, accum_planes
        pixels = np.ascontiguousarray(pixels, np.uint64)
        nc = accum_planes(desc)
        out = np.zeros(nc * pixels.size, ACCUM_DTYPE)
        r = lib().oracle_draw_pixels(self._h, C.byref(desc), pixels.ctypes.data_as(F.c_uint64_p),
                                     pixels.size, out.ctypes.data_as(C.c_void_p), nthreads)
        if r != 0:
            raise RuntimeError(f"oracle_draw_pixels: {r}")
        return out if nc == 1 else out.reshape(nc, pixels.size)

    def close(self):
        if self._h:
            lib().oracle_sky_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
