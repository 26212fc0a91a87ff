"""Edge cases of the map dispatch (htrdr_draw_map's 8x8 Morton tiles,
src/core/htrdr_draw_map.c:271-621, replaced by the plain kernel's tiles and
the persistent kernel's chunk queues): ragged images whose last tiles hang
over the edge (down to 1x1), realisation counts that are not multiples of
the chunk size, slices at odd offsets, and a caller's row pitch wider than a
row (htrdr_buffer_layout, src/core/htrdr_buffer.h:68-94). Every pixel must get
exactly spp realisations, FP64 maps must agree with the oracle (pixels within
3 combined sigma >= 99.5%), FP32 slices must add up to one launch (the same
realisations; the sums differ only by their grouping, 1e-12) and the bytes
between rows must stay untouched."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from htrdr_kmin_amd import hip as H
from test_gpu_parity_maps import pixel_parity

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (13, 7), (9, 17), (65, 3)]


@pytest.fixture(scope="module")
def small():
    s = A.Synth(0, seed=4)
    ctx = A.Context(s.scene_ptr)
    sky = O.Sky(s.scene_ptr)
    yield s, ctx, sky
    sky.close()
    ctx.close()
    s.close()


def desc(w, h, spp, solver, **kw):
    return A.map_preset(0, width=w, height=h, spp=spp, solver=solver, **kw)


@pytest.mark.parametrize("w,h", SIZES)
def test_ragged_images(small, w, h):
    _, ctx, sky = small
    for solver in (F.SOLVER_PLAIN_FP64, F.SOLVER_WAVEFRONT_FP32):
        for spp in (1, 7):
            p = ctx.draw_map(desc(w, h, spp, solver))
            assert p.shape == (h, w)
            assert np.all(p["time_n"] == spp), (solver, spp)
    d = desc(w, h, 64, F.SOLVER_PLAIN_FP64)
    p = ctx.draw_map(d)
    Eo, SEo = A.accum_estimate(sky.draw_map(d))
    assert pixel_parity(p["radiance_E"].ravel(), p["radiance_SE"].ravel(), Eo.ravel(), SEo.ravel()) >= 0.995


@pytest.mark.parametrize("chunk", [1, 2, 4, 16])
def test_fp32_odd_slices_add_up(small, chunk):
    """spp 3 + 4 + 6 at offsets 0, 3, 7 into one accumulator = spp 13 in one
    launch, whatever the chunk size (13 is a multiple of none)"""
    _, ctx, _ = small
    ctx.set_wavefront_params(24, chunk, 0)
    w, h = 13, 7
    npix = w * h
    one = H.DeviceBuffer(npix * A.ACCUM_DTYPE.itemsize)
    one.zero()
    ctx.render_accum(desc(w, h, 13, F.SOLVER_WAVEFRONT_FP32), one.address)
    parts = H.DeviceBuffer(npix * A.ACCUM_DTYPE.itemsize)
    parts.zero()
    for off, n in ((0, 3), (3, 4), (7, 6)):
        ctx.render_accum(desc(w, h, n, F.SOLVER_WAVEFRONT_FP32, spp_offset=off), parts.address)
    H.synchronize()
    a = one.download(np.zeros(npix, A.ACCUM_DTYPE))
    b = parts.download(np.zeros(npix, A.ACCUM_DTYPE))
    ctx.set_wavefront_params(0, 0, 0)
    assert np.all(a["n"] == 13) and np.all(b["n"] == 13)
    Ea, _ = A.accum_estimate(a)
    Eb, _ = A.accum_estimate(b)
    np.testing.assert_allclose(Ea, Eb, rtol=1e-12, atol=0)


def test_row_pitch(small):
    """rows written `pitch' bytes apart; the padding between them keeps the
    caller's bytes"""
    _, ctx, _ = small
    lib = F.load()
    w, h = 13, 7
    for solver in (F.SOLVER_PLAIN_FP64, F.SOLVER_WAVEFRONT_FP32):
        d = desc(w, h, 4, solver)
        dt = A.pixel_dtype(d)
        pitch = w * dt.itemsize + 40
        buf = np.full(pitch * h, 0xA5, np.uint8)
        assert lib.htgx_draw_map(ctx.handle, C.byref(d), buf.ctypes.data_as(C.c_void_p), pitch) == F.OK
        ref = ctx.draw_map(d)
        rows = buf.reshape(h, pitch)
        assert np.all(rows[:, w * dt.itemsize:] == 0xA5)
        px = np.frombuffer(rows[:, :w * dt.itemsize].tobytes(), dt).reshape(h, w)
        assert np.array_equal(px["radiance_E"], ref["radiance_E"])
        assert np.all(px["time_n"] == 4)
        # a pitch narrower than a row is refused
        assert lib.htgx_draw_map(ctx.handle, C.byref(d), buf.ctypes.data_as(C.c_void_p),
                                 w * dt.itemsize - 8) == F.BAD_ARG
