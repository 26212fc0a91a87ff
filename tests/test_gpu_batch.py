"""Batched rendering (htgx_render_batch, htgx_batch_plan_*): several maps of
distinct scenes in one persistent launch give bit-identical accumulators to
separate htgx_render_accum calls (same random streams, same chunk order),
except the wall-clock time sums; so does every step of a batch plan, replayed
from a HIP graph or launched directly."""
import ctypes as C

import numpy as np
import pytest

from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from htrdr_kmin_amd import hip as H

pytestmark = pytest.mark.gpu

EXACT = ("sum", "sum_c", "sum2", "sum2_c", "n", "nfail")


@pytest.fixture(scope="module")
def batch():
    items = A.uneven_batch(seed=1, n=6, spp_scale=0.03)
    ctxs = [A.Context(s.scene_ptr) for s, _ in items]
    yield items, ctxs
    for c in ctxs:
        c.close()


def test_batch_equals_separate_launches(batch):
    items, ctxs = batch
    bufs = [H.DeviceBuffer(d.width * d.height * A.ACCUM_DTYPE.itemsize) for _, d in items]
    for b in bufs:
        b.zero()
    ms = A.render_batch([(c, d, b.address) for c, (_, d), b in zip(ctxs, items, bufs)])
    got = [b.download(np.zeros(d.width * d.height, A.ACCUM_DTYPE)) for (_, d), b in zip(items, bufs)]
    assert np.all(ms > 0)
    for c, (_, d), b, g in zip(ctxs, items, bufs, got):
        b.zero()
        c.render_accum(d, b.address)
        H.synchronize()
        ref = b.download(np.zeros(d.width * d.height, A.ACCUM_DTYPE))
        assert np.all(ref["n"] == d.spp)
        for k in EXACT:
            assert np.array_equal(g[k], ref[k]), k


def test_batch_bad_arguments(batch):
    items, ctxs = batch
    lib = F.load()
    buf = H.DeviceBuffer(512 * 512 * A.ACCUM_DTYPE.itemsize)
    (s0, d0), (s1, d1) = items[0], items[1]
    arr = (F.BatchItem * 2)()
    # the same context twice
    for i in range(2):
        arr[i].ctx, arr[i].desc, arr[i].accum = ctxs[0].handle, C.pointer(d0), C.c_void_p(buf.address)
    assert lib.htgx_render_batch(arr, 2, None, None) == F.BAD_ARG
    # mixed sensor types
    d_flux = A.map_preset(0, width=16, height=16, spp=4, sensor=F.SENSOR_RECTANGLE,
                          rect_size=(100.0, 100.0))
    arr[1].ctx, arr[1].desc = ctxs[1].handle, C.pointer(d_flux)
    assert lib.htgx_render_batch(arr, 2, None, None) == F.BAD_ARG
    assert lib.htgx_render_batch(arr, 0, None, None) == F.BAD_ARG


@pytest.mark.parametrize("nmaps", [6, 1])
@pytest.mark.parametrize("use_graph", [True, False])
def test_plan_steps_equal_sliced_launches(batch, use_graph, nmaps):
    """Each step of a plan (a replayed HIP graph, or direct launches) equals
    htgx_render_accum of the next spp slice, after every step, with other
    launches of the same contexts between the steps. A plan of one map runs
    the single-map kernel, whose graph reads each step's realisation offset
    from the plan's table."""
    items, ctxs = batch
    items, ctxs = items[:nmaps], ctxs[:nmaps]
    npx = [d.width * d.height for _, d in items]
    plan_bufs = [H.DeviceBuffer(n * A.ACCUM_DTYPE.itemsize) for n in npx]
    ref_bufs = [H.DeviceBuffer(n * A.ACCUM_DTYPE.itemsize) for n in npx]
    for b in plan_bufs + ref_bufs:
        b.zero()
    plan = A.BatchPlan([(c, d, b.address) for c, (_, d), b in zip(ctxs, items, plan_bufs)],
                       use_graph=use_graph)
    try:
        for step in range(3):
            plan.step()
            H.synchronize()
            for c, (_, d), b in zip(ctxs, items, ref_bufs):
                d.spp_offset = step * d.spp
                c.render_accum(d, b.address)
                d.spp_offset = 0
            H.synchronize()
            if nmaps == 1:
                # eager work of other contexts between two replays
                for c, (_, d) in zip(batch[1][1:3], batch[0][1:3]):
                    tmp = H.DeviceBuffer(d.width * d.height * A.ACCUM_DTYPE.itemsize)
                    tmp.zero()
                    c.render_accum(d, tmp.address)
                H.synchronize()
            for (_, d), n, pb, rb in zip(items, npx, plan_bufs, ref_bufs):
                got = pb.download(np.zeros(n, A.ACCUM_DTYPE))
                ref = rb.download(np.zeros(n, A.ACCUM_DTYPE))
                assert np.all(ref["n"] == (step + 1) * d.spp)
                for k in EXACT:
                    assert np.array_equal(got[k], ref[k]), (step, k)
        ms = plan.scene_ms()
        assert np.all(ms > 0)
        # rewind: step 0 again, from zeroed accumulators
        plan.rewind()
        for b in plan_bufs:
            b.zero()
        plan.step()
        ms0 = plan.scene_ms()
        assert np.all(ms0 > 0)
        for c, (_, d), n, pb in zip(ctxs, items, npx, plan_bufs):
            rb = H.DeviceBuffer(n * A.ACCUM_DTYPE.itemsize)
            rb.zero()
            c.render_accum(d, rb.address)
            H.synchronize()
            got = pb.download(np.zeros(n, A.ACCUM_DTYPE))
            ref = rb.download(np.zeros(n, A.ACCUM_DTYPE))
            for k in EXACT:
                assert np.array_equal(got[k], ref[k]), k
    finally:
        plan.close()


def test_plan_bad_arguments(batch):
    items, ctxs = batch
    lib = F.load()
    buf = H.DeviceBuffer(512 * 512 * A.ACCUM_DTYPE.itemsize)
    (_, d0) = items[0]
    arr = (F.BatchItem * 2)()
    for i in range(2):
        arr[i].ctx, arr[i].desc, arr[i].accum = ctxs[0].handle, C.pointer(d0), C.c_void_p(buf.address)
    h = C.c_void_p()
    assert lib.htgx_batch_plan_create(arr, 2, 1, C.byref(h)) == F.BAD_ARG and not h.value
    assert lib.htgx_batch_plan_create(arr, 0, 1, C.byref(h)) == F.BAD_ARG
    assert lib.htgx_batch_plan_step(None, None) == F.BAD_ARG
    assert lib.htgx_batch_plan_destroy(None) == F.BAD_ARG


def test_mixed_plane_and_mesh_batch():
    """A batch of a plane-ground map, a mesh-ground map and a map with a rare
    feature (an orthographic camera) runs in the kernel that carries all of
    them; each map still equals its own launch (in its own, leaner kernel)
    bit for bit."""
    plane = A.Synth(0, seed=3)
    mesh = A.Synth(0, seed=4, ground_type=F.GROUND_MESH, hf_res=33, hf_octaves=5, hf_relief=300.0,
                   hf_hurst=0.7)
    ortho = A.Synth(0, seed=5)
    descs = [A.map_preset(0, width=24, height=16, spp=8, solver=F.SOLVER_WAVEFRONT_FP32),
             A.map_preset(0, width=16, height=16, spp=8, solver=F.SOLVER_WAVEFRONT_FP32,
                          pos=(3200.0, -1500.0, 1800.0), tgt=(3200.0, -1499.0, 1799.4)),
             A.map_preset(0, width=16, height=8, spp=8, solver=F.SOLVER_WAVEFRONT_FP32,
                          projection=F.PROJECTION_ORTHOGRAPHIC, ortho_height=3000.0)]
    synths = [plane, mesh, ortho]
    ctxs = [A.Context(s.scene_ptr) for s in synths]
    bufs = [H.DeviceBuffer(d.width * d.height * A.ACCUM_DTYPE.itemsize) for d in descs]
    for b in bufs:
        b.zero()
    A.render_batch([(c, d, b.address) for c, d, b in zip(ctxs, descs, bufs)])
    got = [b.download(np.zeros(d.width * d.height, A.ACCUM_DTYPE)) for d, b in zip(descs, bufs)]
    for c, d, b, g in zip(ctxs, descs, bufs, got):
        b.zero()
        c.render_accum(d, b.address)
        H.synchronize()
        ref = b.download(np.zeros(d.width * d.height, A.ACCUM_DTYPE))
        assert np.all(ref["n"] == d.spp)
        for k in EXACT:
            assert np.array_equal(g[k], ref[k]), k
    for c in ctxs:
        c.close()
    for s in synths:
        s.close()


def test_uneven_batch_full_size():
    """The uneven batched workload at full size: 64 seeded scenes (cloud
    grids 32^3 to 256x256x96, images 64^2 to 512^2, 16-1024 realisations per
    pixel, 1-16 k-points), 1.44 G realisations in one graph-replayed
    persistent step. Every pixel of every map gets exactly its realisations,
    with finite non-negative sums and no failed realisation; every map
    reports a completion time within the step; the three smallest maps equal
    their own launches bit for bit; and 128 pixels of each of the two
    smallest maps agree with the CPU oracle at the same realisation counts
    (>= 97 % within 3 combined sigma, pooled z mean within +-0.35)."""
    import time

    import oracle_ffi as O
    items = A.uneven_batch(seed=0, n=64)
    ctxs = [A.Context(s.scene_ptr) for s, _ in items]
    npx = [d.width * d.height for _, d in items]
    bufs = [H.DeviceBuffer(n * A.ACCUM_DTYPE.itemsize) for n in npx]
    for b in bufs:
        b.zero()
    plan = A.BatchPlan([(c, d, b.address) for c, (_, d), b in zip(ctxs, items, bufs)], use_graph=True)
    try:
        H.synchronize()
        t0 = time.perf_counter()
        plan.step()
        H.synchronize()
        wall = time.perf_counter() - t0
        ms = plan.scene_ms()
    finally:
        plan.close()
    total = sum(n * d.spp for n, (_, d) in zip(npx, items))
    print(f"64 scenes, {total / 1e9:.3f} G realisations in {wall:.2f} s "
          f"({total / wall / 1e6:.1f} Mreal/s); completion {ms.min() / 1e3:.2f}-{ms.max() / 1e3:.2f} s")
    assert np.all(ms > 0) and ms.max() <= wall * 1e3 * 1.01
    got = [b.download(np.zeros(n, A.ACCUM_DTYPE)) for n, b in zip(npx, bufs)]
    for (_, d), g in zip(items, got):
        assert np.all(g["n"] == d.spp) and np.all(g["nfail"] == 0)
        assert np.all(np.isfinite(g["sum"])) and np.all(g["sum"] >= 0) and np.all(g["sum2"] >= 0)
    order = np.argsort([n * d.spp for n, (_, d) in zip(npx, items)])
    for i in order[:3]:
        (_, d), b = items[i], bufs[i]
        b.zero()
        ctxs[i].render_accum(d, b.address)
        H.synchronize()
        ref = b.download(np.zeros(npx[i], A.ACCUM_DTYPE))
        for k in EXACT:
            assert np.array_equal(got[i][k], ref[k]), (i, k)
    rng = np.random.default_rng(7)
    for i in order[:2]:
        s, d = items[i]
        pix = np.sort(rng.choice(npx[i], 128, replace=False)).astype(np.uint64)
        sky = O.Sky(s.scene_ptr)
        acc = sky.draw_pixels(d, pix)
        sky.close()
        E, SE = A.accum_estimate(acc)
        Eg, SEg = A.accum_estimate(got[i][pix.astype(np.int64)])
        sig = np.sqrt(SE ** 2 + SEg ** 2)
        m = sig > 0
        z = (Eg[m] - E[m]) / sig[m]
        frac = ((np.abs(z) <= 3).sum() + (~m).sum()) / pix.size
        print(f"scene {i}: {d.width}x{d.height} x {d.spp}: {frac:.3f} of 128 pixels within 3 sigma, "
              f"z mean {z.mean():+.3f}")
        assert frac >= 0.97 and abs(z.mean()) < 0.35
    for c in ctxs:
        c.close()
    for s, _ in items:
        s.close()
