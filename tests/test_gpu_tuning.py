"""Launch parameters of the FP32 wavefront solver and their autotuner
(htgx_get/set_wavefront_params, htgx_autotune).

The parameters change the schedule only: every realisation draws from its
own counter-based stream, so the event threshold and the grid leave the
accumulators bit-identical, and the chunk size only regroups the partial
sums (same counts, sums equal to 1e-12 relative)."""
import ctypes as C

import numpy as np
import pytest

from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from htrdr_kmin_amd import hip as H

pytestmark = pytest.mark.gpu

EXACT = ("sum", "sum_c", "sum2", "sum2_c", "n", "nfail")


@pytest.fixture(scope="module")
def scene():
    s = A.Synth(0, seed=3)
    ctx = A.Context(s.scene_ptr)
    yield s, ctx
    ctx.close()
    s.close()


def render(ctx, desc):
    n = desc.width * desc.height
    b = H.DeviceBuffer(n * A.ACCUM_DTYPE.itemsize)
    b.zero()
    ctx.render_accum(desc, b.address)
    H.synchronize()
    return b.download(np.zeros(n, A.ACCUM_DTYPE))


def test_defaults_and_bad_values(scene):
    _, ctx = scene
    ctx.set_wavefront_params()
    p = ctx.wavefront_params()
    assert p["ev_thresh"] == 24 and p["chunk"] == 4 and p["grid"] == p["grid_max"] > 0
    assert p["seg_thresh"] == 0
    lib = F.load()
    for bad in (F.WfParams(65, 0, 0, 0), F.WfParams(0, 17, 0, 0), F.WfParams(0, 0, p["grid_max"] + 1, 0),
                F.WfParams(0, 0, 0, 0, 65)):
        assert lib.htgx_set_wavefront_params(ctx.handle, C.byref(bad)) == F.BAD_ARG
    assert ctx.wavefront_params() == p


def test_schedule_does_not_change_results(scene):
    _, ctx = scene
    d = A.map_preset(0, width=96, height=80, spp=16, solver=F.SOLVER_WAVEFRONT_FP32)
    ctx.set_wavefront_params()
    ref = render(ctx, d)
    gmax = ctx.wavefront_params()["grid_max"]
    for ev, grid, seg in ((1, 0, 0), (64, 7, 0), (8, gmax // 2, 0), (24, 0, 16), (1, 0, 64)):
        ctx.set_wavefront_params(ev_thresh=ev, grid=grid, seg_thresh=seg)
        got = render(ctx, d)
        for k in EXACT:
            assert np.array_equal(got[k], ref[k]), (ev, grid, seg, k)
    ctx.set_wavefront_params(chunk=16)
    got = render(ctx, d)
    assert np.array_equal(got["n"], ref["n"]) and np.array_equal(got["nfail"], ref["nfail"])
    np.testing.assert_allclose(got["sum"] + got["sum_c"], ref["sum"] + ref["sum_c"], rtol=1e-12,
                               atol=1e-300)
    np.testing.assert_allclose(got["sum2"] + got["sum2_c"], ref["sum2"] + ref["sum2_c"], rtol=1e-12,
                               atol=1e-300)
    ctx.set_wavefront_params()


def test_autotune(scene):
    _, ctx = scene
    d = A.map_preset(0, width=128, height=128, spp=32, solver=F.SOLVER_WAVEFRONT_FP32)
    ctx.set_wavefront_params()
    # no budget: the current parameters are timed once and kept
    p0, rate0, n0 = ctx.autotune(d, budget_ms=0.0)
    assert n0 == 1 and rate0 > 0
    assert (p0["ev_thresh"], p0["chunk"]) == (24, 4)
    best, rate, n = ctx.autotune(d, budget_ms=5000.0)
    assert n >= 2 and rate > 0
    assert best["ev_thresh"] in (8, 16, 24, 32, 40, 48) and best["chunk"] in (2, 4, 8)
    assert best["seg_thresh"] in (0, 16, 24)
    assert 1 <= best["grid"] <= best["grid_max"]
    assert ctx.wavefront_params() == best  # kept for the maps that follow
    # the tuned context renders the same realisations
    pix = render(ctx, d)
    assert np.all(pix["n"] == 32)
    ctx.set_wavefront_params()


def test_batch_autotune():
    items = A.uneven_batch(seed=4, n=4, spp_scale=0.03)
    ctxs = [A.Context(s.scene_ptr) for s, _ in items]
    bufs = [H.DeviceBuffer(d.width * d.height * A.ACCUM_DTYPE.itemsize) for _, d in items]
    try:
        best, rate, n = A.batch_autotune([(c, d, b.address) for c, (_, d), b in zip(ctxs, items, bufs)],
                                         budget_ms=3000.0, spp_div=1)
        assert n >= 2 and rate > 0
        for c in ctxs:
            assert c.wavefront_params() == best
        # a batch with the tuned parameters still renders every realisation
        for b in bufs:
            b.zero()
        A.render_batch([(c, d, b.address) for c, (_, d), b in zip(ctxs, items, bufs)])
        for (_, d), b in zip(items, bufs):
            assert np.all(b.download(np.zeros(d.width * d.height, A.ACCUM_DTYPE))["n"] == d.spp)
    finally:
        for c in ctxs:
            c.close()
