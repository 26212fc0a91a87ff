"""Cloud repetition (option -r, htsky's repeat_clouds,
src/atmosphere/htrdr_atmosphere.c:375) in the oracle, pinned by

* an explicit tiling: the segments of a repeated 4x3x2 grid equal those of
  the same grid copied 3x3 times into one larger grid (no leaf merging), on
  rays that stay within the 3x3 copies;
* a closed form: the direct SW flux under an infinite, uniform, purely
  absorbing cloud layer is L_sun Omega <cos> exp(-tau / mu) wherever the
  sensor is, also where the sun's rays would miss a single copy of the grid.
"""
import math

import numpy as np

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd.atmosphere import accum_estimate
from scenes import Scene, planck, rectangle, sun_solid_angle
from test_oracle_walks import cone_mean_cos

L_SUN_550 = planck(550e-9, 5778.0)
NX, NY, NZ = 4, 3, 2
LOW, UP = (0.0, 0.0, 1000.0), (400.0, 300.0, 1200.0)


def small_grid(seed=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.1, 2.0, (NZ, NY, NX)).astype(np.float32)


def test_repeated_segments_equal_explicit_copies():
    lwc = small_grid()
    rep = Scene(lwc=lwc, lower=LOW, upper=UP, cext=0.01, optical_thickness=0.0, leaf_block=1,
                repeat_clouds=True)
    big = Scene(lwc=np.tile(lwc, (1, 3, 3)), lower=(-400.0, -300.0, 1000.0), upper=(800.0, 600.0, 1200.0),
                cext=0.01, optical_thickness=0.0, leaf_block=1)
    s_rep, s_big = O.Sky(rep.ptr), O.Sky(big.ptr)
    rng = np.random.default_rng(1)
    rays = []
    while len(rays) < 150:
        o = np.array([rng.uniform(-390, 790), rng.uniform(-290, 590), rng.uniform(900, 1300)])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        # stay within the 3x3 copies along X and Y
        tx = ((800.0 if d[0] > 0 else -400.0) - o[0]) / d[0]
        ty = ((600.0 if d[1] > 0 else -300.0) - o[1]) / d[1]
        r1 = 0.999 * min(tx, ty)
        if r1 > 50:
            rays.append(np.concatenate([o, d, [0.0, r1]]))
    rays = np.array(rays)
    a, na = s_rep.trace_segments(rays, F.Kext)
    b, nb = s_big.trace_segments(rays, F.Kext)
    crossings = 0
    for r in range(len(rays)):
        assert na[r] == nb[r], r
        sa, sb = a[r, :na[r]], b[r, :nb[r]]
        np.testing.assert_allclose(sa["t0"], sb["t0"], rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(sa["t1"], sb["t1"], rtol=1e-12, atol=1e-9)
        assert np.array_equal(sa["kmin"], sb["kmin"]) and np.array_equal(sa["kmax"], sb["kmax"])
        # copies of the grid the ray crosses (cells of the explicit grid)
        ids = sb["leaf"][sb["leaf"] >= 0]
        tiles = {((i % (3 * NX)) // NX, (i // (3 * NX) % (3 * NY)) // NY) for i in ids}
        crossings += len(tiles) > 1
    assert crossings > 40
    s_rep.close()
    s_big.close()


def test_infinite_absorbing_layer_direct_flux():
    el, tau = 30.0, 0.2
    lwc = np.full((NZ, NY, NX), 1.0, np.float32)
    k = tau / (UP[2] - LOW[2])  # cext * LWC
    expect = sun_solid_angle() * L_SUN_550 * cone_mean_cos(el) / 1e9 * math.exp(
        -tau / math.sin(math.radians(el)))
    for repeat in (True, False):
        sc = Scene(lwc=lwc, lower=LOW, upper=UP, cext=k, ssa=0.0, sun=(0.0, el), repeat_clouds=repeat)
        sky = O.Sky(sc.ptr)
        # sensor under and far beside the grid: the sun ray to its points
        # crosses only other copies
        d = rectangle(4, 4, 64, pos=(5000.0, 150.0, 0.0), size=(400.0, 300.0))
        E, SE = accum_estimate(sky.draw_map(d, nthreads=4))
        if repeat:
            assert np.all(np.abs(E - expect) <= 4 * SE + 2e-4 * expect), (E, expect)
        else:
            assert np.all(E > expect * 1.1)  # unshaded: the only copy is far away
        sky.close()
