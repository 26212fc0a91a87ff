"""Host-side mirror of htrdr-atmosphere's draw-map path over libhtgx.so.

This is the Python face of the drop-in boundary: the same steps as
``atmosphere_draw_map`` (src/atmosphere/htrdr_atmosphere_draw_map.c:593-641)
-- render the map through ``htgx_draw_map`` (the replacement of
``htrdr_draw_map``, src/core/htrdr_draw_map.c:497-621), then write the
htrdr-image text with ``dump_buffer`` (..._draw_map.c:477-587) and log the
per-path time and, for flux maps, the overall flux. Errors are the library's
rsys-compatible result codes, raised as ``HtgxError``.
"""
import ctypes as C
import os
import math

import numpy as np

from . import _ffi as F

XWAVE_DTYPE = np.dtype([("radiance_E", "<f8"), ("radiance_SE", "<f8"),
                        ("temperature_E", "<f8"), ("temperature_SE", "<f8"),
                        ("time_sum", "<f8"), ("time_sum2", "<f8"),
                        ("time_n", "<u8"), ("time_nfail", "<u8")])
FLUX_DTYPE = np.dtype([("flux_sum", "<f8"), ("flux_sum2", "<f8"),
                       ("flux_n", "<u8"), ("flux_nfail", "<u8"),
                       ("time_sum", "<f8"), ("time_sum2", "<f8"),
                       ("time_n", "<u8"), ("time_nfail", "<u8")])
IMAGE_DTYPE = np.dtype([("X_E", "<f8"), ("X_SE", "<f8"), ("Y_E", "<f8"), ("Y_SE", "<f8"),
                        ("Z_E", "<f8"), ("Z_SE", "<f8"),
                        ("time_sum", "<f8"), ("time_sum2", "<f8"),
                        ("time_n", "<u8"), ("time_nfail", "<u8")])
ACCUM_DTYPE = np.dtype([("sum", "<f8"), ("sum_c", "<f8"), ("sum2", "<f8"),
                        ("sum2_c", "<f8"), ("tsum", "<f8"), ("tsum2", "<f8"),
                        ("n", "<u8"), ("nfail", "<u8")])
assert XWAVE_DTYPE.itemsize == C.sizeof(F.PixelXwave) == 64
assert FLUX_DTYPE.itemsize == C.sizeof(F.PixelFlux) == 64
assert IMAGE_DTYPE.itemsize == C.sizeof(F.PixelImage) == 80


def pixel_dtype(desc):
    """Reference pixel struct of a map (atmosphere_get_pixel_format)."""
    if desc.sensor == F.SENSOR_RECTANGLE:
        return FLUX_DTYPE
    return IMAGE_DTYPE if desc.spectral_type == F.SPECTRAL_SW_CIE_XYZ else XWAVE_DTYPE


def accum_planes(desc):
    """Accumulator planes of a map: 3 for a CIE XYZ image (X, Y, Z), else 1."""
    return 3 if desc.spectral_type == F.SPECTRAL_SW_CIE_XYZ else 1
assert ACCUM_DTYPE.itemsize == C.sizeof(F.McAccum) == 64


def dump_octrees_vtk(scene_ptr, out):
    """Option -d: the octrees of every band and k-point as legacy VTK, each
    followed by a line `---' (htgx_dump_octrees_vtk), into the open file
    ``out`` (anything with fileno()) or the file at path ``out``."""
    lib = F.load()
    if hasattr(out, "fileno"):
        out.flush()
        F.check("htgx_dump_octrees_vtk", lib.htgx_dump_octrees_vtk(scene_ptr, out.fileno()))
        return
    with open(out, "w") as f:
        F.check("htgx_dump_octrees_vtk", lib.htgx_dump_octrees_vtk(scene_ptr, f.fileno()))


def octree_cache_prepare(scene_ptr, path):
    """Option -O's cache prepared on the host (htgx_octree_cache_prepare):
    checks the cache at ``path`` against the scene and returns True, or
    builds the octree and writes it there and returns False; HtgxError
    with the message for a cache that is refused."""
    lib = F.load()
    loaded, err = C.c_int(0), C.create_string_buffer(512)
    res = lib.htgx_octree_cache_prepare(scene_ptr, os.fsencode(path), C.byref(loaded), err, len(err))
    if res != F.OK:
        raise F.HtgxError(f"htgx_octree_cache_prepare ({err.value.decode()})", res)
    return bool(loaded.value)


def leaf_block_for_definition(definition, max_definition):
    """Leaf block for option -V (htgx_leaf_block_for_definition)"""
    lib = F.load()
    d = np.ascontiguousarray(definition, np.uint32)
    m = np.ascontiguousarray(max_definition, np.uint32)
    B = C.c_uint32(0)
    F.check("htgx_leaf_block_for_definition", lib.htgx_leaf_block_for_definition(
        d.ctypes.data_as(F.c_uint32_p), m.ctypes.data_as(F.c_uint32_p), C.byref(B)))
    return B.value


class GroundFiles:
    """The ground of option -g (an htrdr-obj file) with the materials of
    option -M (an htrdr-materials file and its mrumtl files) and the sky
    material of option -n, read by the library (include/htgx_files.h).

    ``desc`` is an htgx_ground_desc owned by this object; ``with_scene``
    returns a copy of a scene description with this ground.
    """

    def __init__(self, obj_path, materials_path, sky_mtl="air"):
        lib = F.load()
        self._lib = lib
        self._h = C.c_void_p()
        res = lib.htgx_ground_files_load(str(obj_path).encode(), str(materials_path).encode(),
                                         None if sky_mtl is None else sky_mtl.encode(), C.byref(self._h))
        if res != F.OK:
            raise F.HtgxError(f"htgx_ground_files_load ({lib.htgx_files_last_error().decode()})", res)
        self.desc = lib.htgx_ground_files_get_desc(self._h).contents
        self.material_names = [lib.htgx_ground_files_material_name(self._h, m).decode()
                               for m in range(self.desc.nmtls)]
        self.interface_names = [lib.htgx_ground_files_interface_name(self._h, i).decode()
                                for i in range(self.desc.ninterfaces)]

    def with_scene(self, scene):
        """A copy of the htgx_scene_desc ``scene`` whose ground is this one
        (valid while this object is open)"""
        d = F.SceneDesc.from_buffer_copy(scene)
        d.ground = self.desc
        return d

    def close(self):
        if self._h:
            self._lib.htgx_ground_files_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Synth:
    """Seeded synthetic scene (benchmark configuration ``config``)."""

    def __init__(self, config=0, seed=0, **overrides):
        lib = F.load()
        self.params = F.SynthParams()
        F.check("htgx_synth_params_preset", lib.htgx_synth_params_preset(config, C.byref(self.params)))
        self.params.seed = seed
        for k, v in overrides.items():
            field = getattr(self.params, k)
            if isinstance(field, C.Array):
                for i, x in enumerate(v):
                    field[i] = x
            else:
                setattr(self.params, k, v)
        self._h = C.c_void_p()
        F.check("htgx_synth_create", lib.htgx_synth_create(C.byref(self.params), C.byref(self._h)))
        self.scene_ptr = lib.htgx_synth_get_scene(self._h)
        self._lib = lib

    @property
    def scene(self):
        return self.scene_ptr.contents

    def lwc(self):
        sc = self.scene.clouds
        nx, ny, nz = sc.def_[0], sc.def_[1], sc.def_[2]
        if not nx:
            return np.zeros((0, 0, 0), np.float32)
        return np.ctypeslib.as_array(sc.lwc, shape=(nz, ny, nx)).copy()

    def close(self):
        if self._h:
            self._lib.htgx_synth_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def map_preset(config, **overrides):
    """Sensor and spectral settings of benchmark configuration ``config``."""
    lib = F.load()
    d = F.MapDesc()
    F.check("htgx_synth_map_preset", lib.htgx_synth_map_preset(config, C.byref(d)))
    for k, v in overrides.items():
        field = getattr(d, k)
        if isinstance(field, C.Array):
            for i, x in enumerate(v):
                field[i] = x
        else:
            setattr(d, k, v)
    return d


class Context:
    """A libhtgx context: the scene resident on the current HIP device.

    cache: path of the octree cache (option -O of htrdr-atmosphere,
    htgx_ctx_create_cached): read when it holds the octree of the same inputs,
    written when it does not exist, refused (HtgxError with the message)
    when it was built from other inputs. `cache_loaded' says which."""

    def __init__(self, scene_ptr, cache=None):
        self._lib = F.load()
        self._h = C.c_void_p()
        self.cache_loaded = False
        if cache is None:
            F.check("htgx_ctx_create", self._lib.htgx_ctx_create(scene_ptr, C.byref(self._h)))
            return
        loaded, err = C.c_int(0), C.create_string_buffer(512)
        res = self._lib.htgx_ctx_create_cached(scene_ptr, os.fsencode(cache), C.byref(loaded),
                                               err, len(err), C.byref(self._h))
        if res != F.OK:
            raise F.HtgxError(f"htgx_ctx_create_cached ({err.value.decode()})", res)
        self.cache_loaded = bool(loaded.value)

    @property
    def handle(self):
        return self._h

    def svo_info(self):
        nl, nn, nf = C.c_uint32(), C.c_uint64(), C.c_uint64()
        F.check("htgx_ctx_get_svo_info",
                self._lib.htgx_ctx_get_svo_info(self._h, C.byref(nl), C.byref(nn), C.byref(nf)))
        return nl.value, nn.value, nf.value

    def draw_map(self, desc):
        """Pixels of the map as a structured array of shape (height, width)."""
        dt = pixel_dtype(desc)
        out = np.zeros((desc.height, desc.width), dt)
        F.check("htgx_draw_map", self._lib.htgx_draw_map(
            self._h, C.byref(desc), out.ctypes.data_as(C.c_void_p), desc.width * dt.itemsize))
        return out

    def render_accum(self, desc, accum_ptr, stream=None):
        F.check("htgx_render_accum", self._lib.htgx_render_accum(
            self._h, C.byref(desc), C.c_void_p(accum_ptr), C.c_void_p(stream or 0)))

    def wavefront_params(self):
        p = F.WfParams()
        F.check("htgx_get_wavefront_params", self._lib.htgx_get_wavefront_params(self._h, C.byref(p)))
        return {k: getattr(p, k) for k, _ in F.WfParams._fields_}

    def set_wavefront_params(self, ev_thresh=0, chunk=0, grid=0, seg_thresh=0):
        p = F.WfParams(ev_thresh, chunk, grid, 0, seg_thresh)
        F.check("htgx_set_wavefront_params", self._lib.htgx_set_wavefront_params(self._h, C.byref(p)))

    def autotune(self, desc, budget_ms=2000.0):
        """Tune the FP32 solver's launch parameters for `desc' (htgx_autotune);
        returns (params, realisations/s of the best trial, #trials)."""
        p, rate, n = F.WfParams(), C.c_double(), C.c_uint32()
        F.check("htgx_autotune", self._lib.htgx_autotune(self._h, C.byref(desc), budget_ms, C.byref(p),
                                                         C.byref(rate), C.byref(n)))
        return {k: getattr(p, k) for k, _ in F.WfParams._fields_}, rate.value, n.value

    def set_counting(self, on):
        F.check("htgx_set_counting", self._lib.htgx_set_counting(self._h, 1 if on else 0))

    def counters(self):
        c = F.Counters()
        F.check("htgx_get_counters", self._lib.htgx_get_counters(self._h, C.byref(c)))
        out = {k: getattr(c, k) for k, _ in F.Counters._fields_ if k != "diag"}
        out["diag"] = list(c.diag)
        return out

    def trace_segments(self, rays, prop, iband=0, iquad=0, max_segs=256, fp32=False):
        """Segments of rays [org, dir, range] as one solver's iterator sees
        them: the FP64 solver's (htgx_debug_trace_segments) or, with fp32,
        the FP32 wavefront solver's (htgx_debug_trace_segments_fp32)"""
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 8)
        n = rays.shape[0]
        segs = (F.Segment * (n * max_segs))()
        nsegs = np.zeros(n, np.uint32)
        name = "htgx_debug_trace_segments_fp32" if fp32 else "htgx_debug_trace_segments"
        F.check(name, getattr(self._lib, name)(
            self._h, rays.ctypes.data_as(F.c_double_p), n, prop, iband, iquad, segs, max_segs,
            nsegs.ctypes.data_as(F.c_uint32_p)))
        arr = np.ctypeslib.as_array(segs).reshape(n, max_segs)
        return arr, nsegs

    def realisations(self, desc, pixels, reals):
        pixels = np.ascontiguousarray(pixels, np.uint64)
        reals = np.ascontiguousarray(reals, np.uint64)
        out = np.zeros(pixels.size, np.float64)
        F.check("htgx_debug_realisations", self._lib.htgx_debug_realisations(
            self._h, C.byref(desc), pixels.ctypes.data_as(F.c_uint64_p),
            reals.ctypes.data_as(F.c_uint64_p), pixels.size, out.ctypes.data_as(F.c_double_p)))
        return out

    def sample_dirs(self, kind, param, dir_in, u):
        u = np.ascontiguousarray(u, np.float64).reshape(-1, 2)
        d = np.ascontiguousarray(dir_in, np.float64)
        out = np.zeros((u.shape[0], 3), np.float64)
        F.check("htgx_debug_sample_dirs", self._lib.htgx_debug_sample_dirs(
            self._h, kind, param, d.ctypes.data_as(F.c_double_p), u.ctypes.data_as(F.c_double_p),
            u.shape[0], out.ctypes.data_as(F.c_double_p)))
        return out

    def close(self):
        if self._h:
            self._lib.htgx_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def uneven_batch(seed=0, n=64, spp_scale=1.0):
    """The uneven batched workload: n seeded SW scenes and camera images.

    Per scene: cloud grid uniform in [32, 256]^2 x [32, 96] cells (even),
    100 m cells with the benchmark scenes' LWC law; image N^2 with N uniform in
    [64, 512] (multiple of 8); realisations per pixel log-uniform in
    [16, 1024] (times spp_scale, a multiple of 16: 4 steps of a multiple of
    4 each); 1-16 k-points. The camera
    looks at the grid from its south edge like the small scene's camera.
    Returns [(Synth, MapDesc)]."""
    rng = np.random.default_rng(0x0BA7C4 + seed)
    out = []
    for i in range(n):
        nx, ny = (int(rng.integers(16, 129)) * 2 for _ in range(2))
        nz = int(rng.integers(16, 49)) * 2
        npx = int(rng.integers(8, 65)) * 8
        spp = max(16, int(round(2.0 ** rng.uniform(4, 10) * spp_scale / 16)) * 16)
        nq = int(rng.integers(1, 17))
        ext = (nx * 100.0, ny * 100.0, 1000.0 + nz * 100.0)
        s = Synth(0, seed=seed * 1000 + i, cloud_def=(nx, ny, nz), cloud_upper=ext, nquads=nq)
        tilt = math.radians(60.0)
        pos = (ext[0] / 2, -2000.0, 1000.0 + nz * 100.0 + 1800.0)
        d = map_preset(0, width=npx, height=npx, spp=spp, pos=pos,
                       tgt=(pos[0], pos[1] + math.cos(tilt), pos[2] - math.sin(tilt)),
                       solver=F.SOLVER_WAVEFRONT_FP32, seed=i)
        out.append((s, d))
    return out


def render_batch(items, stream=None):
    """Several maps in one persistent launch (htgx_render_batch).

    items: [(Context, MapDesc, device accumulator address)]. Returns the
    completion time of each map after the launch started, in ms."""
    lib = F.load()
    n = len(items)
    arr = _batch_items(items)
    ms = np.zeros(n)
    F.check("htgx_render_batch", lib.htgx_render_batch(
        arr, n, C.c_void_p(stream or 0), ms.ctypes.data_as(F.c_double_p)))
    return ms


def _batch_items(items):
    arr = (F.BatchItem * len(items))()
    for i, (ctx, desc, accum) in enumerate(items):
        arr[i].ctx = ctx.handle
        arr[i].desc = C.pointer(desc)
        arr[i].accum = C.c_void_p(accum)
    return arr


def batch_autotune(items, budget_ms=10000.0, spp_div=8):
    """Tune the launch parameters of a batch (htgx_batch_autotune) and set
    them on every context; returns (params, realisations/s of the best trial,
    #trials). items: [(Context, MapDesc, device accumulator address)]."""
    lib = F.load()
    p, rate, n = F.WfParams(), C.c_double(), C.c_uint32()
    F.check("htgx_batch_autotune", lib.htgx_batch_autotune(
        _batch_items(items), len(items), spp_div, budget_ms, C.byref(p), C.byref(rate), C.byref(n)))
    return {k: getattr(p, k) for k, _ in F.WfParams._fields_}, rate.value, n.value


class BatchPlan:
    """A batch rendered in repeated steps (htgx_batch_plan_*), each step
    replaying one HIP graph when use_graph.

    items: [(Context, MapDesc, device accumulator address)]."""

    def __init__(self, items, use_graph=True):
        self._lib = F.load()
        self.n = len(items)
        self._h = C.c_void_p()
        F.check("htgx_batch_plan_create", self._lib.htgx_batch_plan_create(
            _batch_items(items), self.n, int(use_graph), C.byref(self._h)))

    def step(self, stream=None):
        F.check("htgx_batch_plan_step", self._lib.htgx_batch_plan_step(self._h, C.c_void_p(stream or 0)))

    def rewind(self):
        F.check("htgx_batch_plan_rewind", self._lib.htgx_batch_plan_rewind(self._h))

    def scene_ms(self):
        ms = np.zeros(self.n)
        F.check("htgx_batch_plan_scene_ms", self._lib.htgx_batch_plan_scene_ms(
            self._h, ms.ctypes.data_as(F.c_double_p)))
        return ms

    def close(self):
        if self._h:
            self._lib.htgx_batch_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def threefry(ctr, key):
    """Threefry-4x64-20 blocks computed on the device."""
    lib = F.load()
    ctr = np.ascontiguousarray(ctr, np.uint64).reshape(-1, 4)
    key = np.ascontiguousarray(key, np.uint64).reshape(-1, 4)
    out = np.zeros_like(ctr)
    F.check("htgx_debug_threefry", lib.htgx_debug_threefry(
        ctr.ctypes.data_as(F.c_uint64_p), key.ctypes.data_as(F.c_uint64_p), ctr.shape[0],
        out.ctypes.data_as(F.c_uint64_p)))
    return out


def threefry2x32(ctr, key):
    """Threefry-2x32-20 blocks computed on the device."""
    lib = F.load()
    ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 2)
    key = np.ascontiguousarray(key, np.uint32).reshape(-1, 2)
    out = np.zeros_like(ctr)
    F.check("htgx_debug_threefry2x32", lib.htgx_debug_threefry2x32(
        ctr.ctypes.data_as(F.c_uint32_p), key.ctypes.data_as(F.c_uint32_p), ctr.shape[0],
        out.ctypes.data_as(F.c_uint32_p)))
    return out


def mesh_trace(scene_ptr, rays, prev=None, prev_edge=None, any_hit=False, ctx=None):
    """Ground-mesh hits of rays [n][8] (org, dir, range) through the BVH: on
    the host when ctx is None, else on the GPU of `ctx' (which must hold the
    same scene). Returns (t, prim, nodes visited)."""
    lib = F.load()
    rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 8)
    n = rays.shape[0]
    prev = np.full(n, -1, np.int64) if prev is None else np.ascontiguousarray(prev, np.int64)
    prev_edge = (np.zeros(n, np.int32) if prev_edge is None
                 else np.ascontiguousarray(prev_edge, np.int32))
    t = np.zeros(n, np.float64)
    prim = np.zeros(n, np.int64)
    nv = np.zeros(n, np.uint32)
    F.check("htgx_debug_mesh_trace", lib.htgx_debug_mesh_trace(
        ctx.handle if ctx is not None else None, scene_ptr, rays.ctypes.data_as(F.c_double_p),
        prev.ctypes.data_as(C.POINTER(C.c_int64)), prev_edge.ctypes.data_as(C.POINTER(C.c_int32)),
        n, int(any_hit), t.ctypes.data_as(F.c_double_p), prim.ctypes.data_as(C.POINTER(C.c_int64)),
        nv.ctypes.data_as(F.c_uint32_p)))
    return t, prim, nv


def finalize(desc, accum):
    """Reference pixel structs from a host copy of the device accumulators."""
    lib = F.load()
    accum = np.ascontiguousarray(accum, ACCUM_DTYPE)
    dt = pixel_dtype(desc)
    out = np.zeros((desc.height, desc.width), dt)
    F.check("htgx_finalize_pixels", lib.htgx_finalize_pixels(
        C.byref(desc), accum.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
        desc.width * dt.itemsize))
    return out


def estimate(s, s2, n):
    """htrdr_accum_get_estimation (src/core/htrdr_accum.h:48-68), vectorised."""
    s = np.asarray(s, np.float64)
    s2 = np.asarray(s2, np.float64)
    n = np.asarray(n, np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        E = np.where(n > 0, s / np.maximum(n, 1), 0.0)
        V = np.maximum(s2 / np.maximum(n, 1) - E * E, 0.0)
        SE = np.where(n > 0, np.sqrt(V / np.maximum(n, 1)), 0.0)
    return E, SE


def accum_estimate(acc):
    return estimate(acc["sum"] + acc["sum_c"], acc["sum2"] + acc["sum2_c"], acc["n"])


def _fmt_g(x):
    return "%g" % x


def dump_buffer(pixels, desc, stream):
    """Write an htrdr-image (doc/htrdr-image.5) as dump_buffer does
    (src/atmosphere/htrdr_atmosphere_draw_map.c:477-587). Returns the
    per-path time accumulator and, for flux maps, the flux accumulator
    (sum, sum2, n) summed over the pixels with nweights > 0."""
    h, w = pixels.shape
    stream.write("%lu %lu\n" % (w, h))
    tacc = [0.0, 0.0, 0]
    facc = [0.0, 0.0, 0]

    def dump_accum(s, s2, n, acc):
        if n == 0:
            stream.write("0 0 ")
            return
        E, SE = estimate(s, s2, n)
        stream.write("%s %s " % (_fmt_g(float(E)), _fmt_g(float(SE))))
        if acc is not None:
            acc[0] += s
            acc[1] += s2
            acc[2] += n

    flux = desc.sensor == F.SENSOR_RECTANGLE
    cie = not flux and desc.spectral_type == F.SPECTRAL_SW_CIE_XYZ
    for y in range(h):
        for x in range(w):
            p = pixels[y, x]
            if flux:
                dump_accum(p["flux_sum"], p["flux_sum2"], int(p["flux_n"]), facc)
                stream.write("0 0 0 0 ")
            elif cie:
                # dump_pixel_image, ..._draw_map.c:516-527
                for c in "XYZ":
                    stream.write("%s %s " % (_fmt_g(float(p[c + "_E"])), _fmt_g(float(p[c + "_SE"]))))
            else:
                # ..._draw_map.c:536-540: "%g %g %f %f 0 0 "; the temperature
                # columns are written 0 0 for shortwave maps (tile memory is
                # left uninitialised there in the reference).
                stream.write("%s %s %f %f 0 0 " % (
                    _fmt_g(float(p["temperature_E"])), _fmt_g(float(p["temperature_SE"])),
                    float(p["radiance_E"]), float(p["radiance_SE"])))
            dump_accum(p["time_sum"], p["time_sum2"], int(p["time_n"]), tacc)
            stream.write("\n")
        stream.write("\n")
    return tacc, facc
