"""The FP32 wavefront solver's building blocks, checked directly.

- Its segment iterator (htgx_debug_trace_segments_fp32) against the CPU
  oracle's: on rays whose segments are all longer than 5 cm (so that FP32
  and FP64 boundary distances cannot swap two crossings), the same leaves
  and gas layers in the same order, distances to FP32 precision, and
  property bounds that contain the oracle's exact ones (FP16 leaf bounds
  rounded outwards: kmin never above, kmax never below, up to the FP32
  rounding of the gas term) and exceed them by at most one FP16 step.
- Its single realisations (htgx_debug_realisations with the FP32 solver):
  they sum to the map's accumulators, do not depend on the list they come
  in, and equal the closed form where every realisation is deterministic.
- Its truncation counters: zero on the benchmark scenes, and counting a
  query that does exhaust the tentative-collision budget.
"""
import numpy as np
import pytest

from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from htrdr_kmin_amd import hip as H
import oracle_ffi as O
from scenes import Scene, camera, planck

pytestmark = pytest.mark.gpu

FP16_STEP = 2.0 ** -10   # relative step of a normal FP16 value
FP16_SUB = 2.0 ** -24    # step of the subnormal FP16 values (times the scale)
F32_REL = 4e-7           # FP32 rounding of kmin/kmax (a few half-ulps)


def fp32_rays(rng, n, lo, hi):
    """Random rays [org, dir, 0, r1] with FP32-representable origins and
    directions (the FP32 solver rounds them; the oracle then traces the same
    ray exactly)"""
    org = rng.uniform(lo, hi, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d[: n // 8, 2] = 0.0  # some horizontal rays
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    org = org.astype(np.float32).astype(np.float64)
    d = d.astype(np.float32).astype(np.float64)
    r1 = np.where(rng.random(n) < 0.5, 3.4028234663852886e38, rng.uniform(0, 20000, n))
    r1 = r1.astype(np.float32).astype(np.float64)
    return np.concatenate([org, d, np.zeros((n, 1)), r1[:, None]], axis=1)


def cloud_factor(scene, prop, band=0):
    c = scene.clouds
    ce, w = c.cext[band], c.ssa[band]
    return {F.Ka: ce * (1 - w), F.Ks: ce * w, F.Kext: ce}[prop]


@pytest.fixture(scope="module")
def fine():
    """The small scene's sky with an octree of single blocks (threshold 0,
    leaf block 1): many short leaves along every ray"""
    s = A.Synth(0, seed=1, optical_thickness=0.0, leaf_block=1)
    sky = O.Sky(s.scene_ptr)
    with A.Context(s.scene_ptr) as ctx:
        yield s, sky, ctx
    sky.close()
    s.close()


@pytest.mark.parametrize("prop", [F.Ka, F.Ks, F.Kext])
@pytest.mark.parametrize("tree", ["merged", "fine"])
def test_fp32_segments_match_oracle(tree, prop, request):
    if tree == "merged":
        synth, sky, ctx = request.getfixturevalue("synth0"), request.getfixturevalue("sky0"), \
            request.getfixturevalue("ctx0")
    else:
        synth, sky, ctx = request.getfixturevalue("fine")
    rays = fp32_rays(np.random.default_rng(10 + prop), 3000, (-3000, -3000, -500), (9400, 9400, 22000))
    g, gn = ctx.trace_segments(rays, prop, max_segs=512, fp32=True)
    o, on = sky.trace_segments(rays, prop, max_segs=512)
    assert on.max() < 512 and gn.max() < 512
    c = cloud_factor(synth.scene, prop)
    scale = float(synth.lwc().max())
    checked = nseg = 0
    for i in range(len(rays)):
        a, b = g[i, : gn[i]], o[i, : on[i]]
        if on[i] and np.min(b["t1"] - b["t0"]) < 0.05:
            continue  # a crossing within FP32 reach of another one
        checked += 1
        nseg += on[i]
        assert gn[i] == on[i], i
        assert np.array_equal(a["leaf"], b["leaf"]) and np.array_equal(a["layer"], b["layer"]), i
        # FP32 distances: a few ulps of the distance or of the coordinates
        assert np.all(np.abs(a["t0"] - b["t0"]) <= 4e-7 * (np.abs(b["t0"]) + 3e4)), i
        fin = b["t1"] < 1e30
        assert np.all(np.abs(a["t1"] - b["t1"])[fin] <= 4e-7 * (np.abs(b["t1"][fin]) + 3e4)), i
        # outward bounds, at most one FP16 step looser
        cl = b["leaf"] >= 0
        slack = np.where(cl, FP16_STEP * b["kmax"] + c * scale * FP16_SUB, 0.0)
        assert np.all(a["kmin"] <= b["kmin"] * (1 + F32_REL)), i
        assert np.all(a["kmax"] >= b["kmax"] * (1 - F32_REL)), i
        assert np.all(a["kmax"] <= b["kmax"] * (1 + F32_REL) + slack), i
        assert np.all(a["kmin"] >= b["kmin"] * (1 - F32_REL) - slack), i
    print(f"{checked} of {len(rays)} rays compared, {nseg} segments")
    assert checked > 0.8 * len(rays) and nseg > 10 * checked


def test_fp32_trace_rejects_a_range_not_starting_at_the_origin(ctx0):
    rays = np.array([[0, 0, 1000, 0, 1, 0, 5.0, 100.0]])
    with pytest.raises(F.HtgxError):
        ctx0.trace_segments(rays, F.Ks, fp32=True)
    with pytest.raises(F.HtgxError):
        ctx0.trace_segments(np.array([[0, 0, 1000, 0, 1, 0, 0.0, 100.0]]), F.Ks, iquad=1, fp32=True)


def test_fp32_realisations_sum_to_the_map(ctx0):
    """Each pixel's FP32 realisations 0..7 (one list launch) sum to the
    accumulators of the 8-spp map (its chunked launch): the same
    realisations, to the rounding of the summation order (1e-12)"""
    d = A.map_preset(0, width=24, height=16, spp=8, solver=F.SOLVER_WAVEFRONT_FP32)
    buf = H.DeviceBuffer(24 * 16 * A.ACCUM_DTYPE.itemsize)
    buf.zero()
    ctx0.render_accum(d, buf.address)
    H.synchronize()
    acc = buf.download(np.zeros(24 * 16, A.ACCUM_DTYPE))
    buf.free()
    pix = np.repeat(np.arange(24 * 16, dtype=np.uint64), 8)
    real = np.tile(np.arange(8, dtype=np.uint64), 24 * 16)
    w = ctx0.realisations(d, pix, real).reshape(-1, 8)
    assert np.all(np.isfinite(w)) and np.all(w >= 0)
    assert np.all(acc["n"] == 8)
    assert np.allclose(w.sum(1), acc["sum"] + acc["sum_c"], rtol=1e-12, atol=0)
    assert np.allclose((w ** 2).sum(1), acc["sum2"] + acc["sum2_c"], rtol=1e-12, atol=0)
    # the same pairs shuffled, and a subset: the same weights
    rng = np.random.default_rng(3)
    perm = rng.permutation(pix.size)
    w2 = ctx0.realisations(d, pix[perm], real[perm])
    assert np.array_equal(w2, w.reshape(-1)[perm])
    sub = perm[:100]
    assert np.array_equal(ctx0.realisations(d, pix[sub], real[sub]), w.reshape(-1)[sub])


def test_fp32_realisations_cie_channels(ctx0):
    """CIE XYZ: bits 32-33 of the realisation index pick the channel; each
    channel's realisations sum to its accumulator plane"""
    d = A.map_preset(0, width=16, height=8, spp=4, solver=F.SOLVER_WAVEFRONT_FP32,
                     spectral_type=F.SPECTRAL_SW_CIE_XYZ, wlen_range=(380.0, 780.0))
    npx = 16 * 8
    buf = H.DeviceBuffer(3 * npx * A.ACCUM_DTYPE.itemsize)
    buf.zero()
    ctx0.render_accum(d, buf.address)
    H.synchronize()
    acc = buf.download(np.zeros(3 * npx, A.ACCUM_DTYPE)).reshape(3, npx)
    buf.free()
    for c in range(3):
        pix = np.repeat(np.arange(npx, dtype=np.uint64), 4)
        real = np.tile(np.arange(4, dtype=np.uint64), npx) | np.uint64(c << 32)
        w = ctx0.realisations(d, pix, real).reshape(-1, 4)
        assert np.allclose(w.sum(1), acc[c]["sum"] + acc[c]["sum_c"], rtol=1e-12, atol=0), c


def test_fp32_realisations_isothermal_enclosure_closed_form():
    """Every realisation in an isothermal enclosure with a black ground is
    B_lambda(T) / pdf, pdf = 1e9 m^-1 for a monochromatic map: each FP32
    realisation within 2e-5 of it (FP32 Planck)"""
    T = 280.0
    sc = Scene(bands=((9500.0, 10500.0),), nlayers=50, dz=400.0, ka=np.full((1, 50), 2e-5),
               T=np.full(50, T), ground=("plane", 0.0), albedo=0.0, ground_T=T)
    d = camera(8, 8, 16, pos=(0, 0, 5000), tgt=(0.2, -0.1, 4999), fov=50.0, wlen=(10000.0, 10000.0),
               spectral=F.SPECTRAL_LW, Tref=T, solver=F.SOLVER_WAVEFRONT_FP32)
    with A.Context(sc.ptr) as ctx:
        rng = np.random.default_rng(5)
        pix = rng.integers(0, 64, 500).astype(np.uint64)
        real = rng.integers(0, 2 ** 32, 500).astype(np.uint64)
        w = ctx.realisations(d, pix, real)
    assert np.allclose(w, planck(1e-5, T) / 1e9, rtol=2e-5, atol=0)


def test_truncation_counters_zero_on_the_benchmark_scenes(ctx0):
    """No query of the small scene's image or of its LW flux map is cut
    short by either bound"""
    before = ctx0.counters()
    ctx0.draw_map(A.map_preset(0, width=32, height=32, spp=16, solver=F.SOLVER_WAVEFRONT_FP32))
    after = ctx0.counters()
    assert after["ntrunc_coll"] == before["ntrunc_coll"] == 0
    assert after["ntrunc_iter"] == before["ntrunc_iter"] == 0


def test_truncation_counter_counts_an_exhausted_budget():
    """A 2 km cloud grid forming one octree leaf whose majorant comes from a
    single dense cell (1e3 m^-1) while the rays cross only empty cells: a
    scattering query meets about 2e6 tentative collisions, all null, beyond
    the budget of 2^20. Such queries end early and are counted."""
    lwc = np.zeros((4, 4, 4), np.float32)
    lwc[0, 0, 0] = 1e4
    sc = Scene(lwc=lwc, lower=(0, 0, 1000), upper=(2000, 2000, 3000), cext=0.1, ssa=1.0, g=0.85,
               optical_thickness=1e30, leaf_block=1, sun=(0.0, 60.0))
    # looking along +y at x, z inside the empty cells: 2 km of path in the grid
    d = camera(2, 2, 2, pos=(1500.0, -10.0, 2500.0), tgt=(1500.0, 1.0, 2500.0), fov=1e-3,
               solver=F.SOLVER_WAVEFRONT_FP32)
    with A.Context(sc.ptr) as ctx:
        assert ctx.svo_info()[2] == 1
        p = ctx.draw_map(d)
        c = ctx.counters()
    assert np.all(p["time_n"] == 2)
    assert c["ntrunc_coll"] > 0 and c["ntrunc_iter"] == 0
