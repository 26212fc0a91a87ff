"""htgx_render_accum_noalloc: htgx_render_accum on the FP32 solver with its
chunk counters and partials in the library's static scratch array instead of
allocated memory. Through the C-ABI, two calls in a row (the second on a
smaller map) give the same accumulators bit for bit as htgx_render_accum
with the same launch parameters; a map beyond the scratch, a plain-solver
map, a band map whose wavelength tables the context does not hold and a
context without FP32 tables are refused without allocating."""
import ctypes as C

import numpy as np
import pytest

from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from htrdr_kmin_amd import hip as H

pytestmark = pytest.mark.gpu

FIELDS = ("sum", "sum_c", "sum2", "sum2_c", "n", "nfail")


def noalloc(ctx, d, address):
    return F.load().htgx_render_accum_noalloc(ctx.handle, C.byref(d), C.c_void_p(address), None)


def both(ctx, d):
    npx = d.width * d.height * (3 if d.spectral_type == F.SPECTRAL_SW_CIE_XYZ else 1)
    ref = H.DeviceBuffer(npx * A.ACCUM_DTYPE.itemsize)
    got = H.DeviceBuffer(npx * A.ACCUM_DTYPE.itemsize)
    ref.zero()
    got.zero()
    ctx.render_accum(d, ref.address)
    H.synchronize()
    assert noalloc(ctx, d, got.address) == F.OK
    H.synchronize()
    r = ref.download(np.zeros(npx, A.ACCUM_DTYPE))
    g = got.download(np.zeros(npx, A.ACCUM_DTYPE))
    ref.free()
    got.free()
    return r, g


def test_noalloc_equals_render_accum_two_calls():
    s = A.Synth(0, seed=2)
    with A.Context(s.scene_ptr) as ctx:
        big = A.map_preset(0, width=96, height=64, spp=16, solver=F.SOLVER_WAVEFRONT_FP32)
        small = A.map_preset(0, width=40, height=24, spp=8, spp_offset=16,
                             solver=F.SOLVER_WAVEFRONT_FP32)
        for d in (big, small):
            r, g = both(ctx, d)
            assert np.all(r["n"] == d.spp)
            for f in FIELDS:
                assert np.array_equal(r[f], g[f]), f
    s.close()


def test_noalloc_band_and_cie_maps_after_their_tables_exist():
    """a band map (Planck wavelength CDF) and a CIE XYZ image (3 planes):
    refused while the context holds other wavelength tables, equal to
    htgx_render_accum once it holds theirs"""
    s = A.Synth(0, seed=3)
    with A.Context(s.scene_ptr) as ctx:
        band = A.map_preset(0, width=24, height=16, spp=8, wlen_range=(500.0, 600.0),
                            solver=F.SOLVER_WAVEFRONT_FP32)
        cie = A.map_preset(0, width=16, height=16, spp=4, spectral_type=F.SPECTRAL_SW_CIE_XYZ,
                           wlen_range=(380.0, 780.0), solver=F.SOLVER_WAVEFRONT_FP32)
        ctx.wavefront_params()  # builds the FP32 tables
        buf = H.DeviceBuffer(3 * 24 * 16 * A.ACCUM_DTYPE.itemsize)
        assert noalloc(ctx, band, buf.address) == F.BAD_OP
        assert noalloc(ctx, cie, buf.address) == F.BAD_OP
        buf.free()
        for d in (band, cie):
            r, g = both(ctx, d)
            for f in FIELDS:
                assert np.array_equal(r[f], g[f]), f
    s.close()


def test_noalloc_refusals():
    s = A.Synth(0, seed=4)
    with A.Context(s.scene_ptr) as ctx:
        d = A.map_preset(0, width=16, height=16, spp=4, solver=F.SOLVER_WAVEFRONT_FP32)
        buf = H.DeviceBuffer(512 * 512 * A.ACCUM_DTYPE.itemsize)
        # no FP32 tables yet: building them would allocate
        assert noalloc(ctx, d, buf.address) == F.BAD_OP
        ctx.wavefront_params()
        assert noalloc(ctx, d, buf.address) == F.OK
        H.synchronize()
        # the plain solver has no allocation-free path
        d64 = A.map_preset(0, width=16, height=16, spp=4, solver=F.SOLVER_PLAIN_FP64)
        assert noalloc(ctx, d64, buf.address) == F.BAD_ARG
        # more chunk partials than the 64 MiB scratch holds: 512^2 pixels x
        # 256 realisations in chunks of at most 16
        huge = A.map_preset(0, width=512, height=512, spp=256, solver=F.SOLVER_WAVEFRONT_FP32)
        assert noalloc(ctx, huge, buf.address) == F.BAD_ARG
        assert F.load().htgx_render_accum_noalloc(ctx.handle, C.byref(d), None, None) == F.BAD_ARG
        buf.free()
    s.close()
