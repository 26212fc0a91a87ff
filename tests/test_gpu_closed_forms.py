"""Both HIP solvers on scenes with closed-form answers (the scenes of
test_oracle_walks.py, seen through the map interface).

Tolerances: deterministic cases 1e-9 relative for the FP64 solver and 2e-5
for the FP32 one (FP32 geometry and Planck; Tmin stays FP64); the narrow
camera's film jitter (fov 1e-4 deg) moves the optical path by < 1e-6
relative. Estimators: 5 standard errors per pixel, 4 on the image mean.
"""
import math

import numpy as np
import pytest

import closed_forms as CF
from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from scenes import Scene, camera, cells_along, planck, rectangle, sun_dir, sun_solid_angle, SUN_HALF_ANGLE

pytestmark = pytest.mark.gpu

SOLVERS = [F.SOLVER_PLAIN_FP64, F.SOLVER_WAVEFRONT_FP32]
DET_RTOL = {F.SOLVER_PLAIN_FP64: 1e-9, F.SOLVER_WAVEFRONT_FP32: 2e-5}
L_SUN_550 = planck(550e-9, 5778.0)
LW_WLEN = 10000.0
LW_BANDS = ((9500.0, 10500.0),)


def cone_mean_cos(elevation_deg):
    return math.sin(math.radians(elevation_deg)) * (1 + math.cos(SUN_HALF_ANGLE)) / 2


def render(scene, desc):
    with A.Context(scene.ptr) as ctx:
        return ctx.draw_map(desc)


def check_estimates(E, SE, expect, rtol=1e-12):
    assert np.all(np.abs(E - expect) <= 5 * SE + rtol * abs(expect))
    assert abs(E.mean() - expect) <= 4 * math.sqrt((SE ** 2).sum()) / E.size + rtol * abs(expect)


def sun_camera(org, s, n, spp, solver, **kw):
    return camera(n, n, spp, pos=tuple(org), tgt=tuple(np.asarray(org) + s), up=(0, 0, 1),
                  fov=1e-4, solver=solver, **kw)


@pytest.mark.parametrize("solver", SOLVERS)
def test_vacuum_ground(solver):
    sc = Scene(ground=("plane", 0.0), albedo=0.3, sun=(0.0, 60.0))
    p = render(sc, camera(16, 12, 32, pos=(0, 0, 3000), tgt=(0, 1, 2999), fov=40.0, solver=solver))
    expect = L_SUN_550 * sun_solid_angle() * 0.3 / math.pi * cone_mean_cos(60.0) / 1e9
    assert np.all(p["time_n"] == 32)
    check_estimates(p["radiance_E"], p["radiance_SE"], expect)


@pytest.mark.parametrize("solver", SOLVERS)
def test_absorbing_slab_direct_sun(solver):
    zmid = (np.arange(50) + 0.5) * 400.0
    ka = 3e-5 * np.exp(-zmid / 8000.0)
    sc = Scene(nlayers=50, dz=400.0, ka=ka[None, :], sun=(30.0, 50.0))
    s = sun_dir(30.0, 50.0)
    p = render(sc, sun_camera((100.0, -50.0, 0.0), s, 4, 8, solver))
    expect = L_SUN_550 * math.exp(-ka.sum() * 400.0 / s[2]) / 1e9
    assert np.allclose(p["radiance_E"], expect, rtol=max(DET_RTOL[solver], 2e-6), atol=0)


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("threshold", [1.0, 1e3])
def test_absorbing_cloud_transmissivity(solver, threshold):
    """At threshold 1 the leaves are single cells (k_min = k_max, nearly
    deterministic); at 1e3 the whole grid is one leaf and the residual
    tracking carries all the randomness. The pixels' rays sit up to 1e-6 rad
    off the closed form's ray and cross cell boundaries elsewhere, which
    moves the optical depth by up to ~1e-4 relative: rtol 2e-4. Checked on
    the pooled image mean only: with tight bounds a real residual collision
    (a zero) is a rare event, and a pixel that drew none reports an SE blind
    to it (the FP32 solver's FP16 leaf bounds are rounded outwards, so its
    k_min < k_max even for single-cell leaves)."""
    rng = np.random.default_rng(1)
    lwc = np.exp(rng.normal(-1.0, 0.8, (4, 8, 8))).astype(np.float32)
    lwc[rng.random(lwc.shape) < 0.3] = 0.0
    sc = Scene(lwc=lwc, lower=(0, 0, 1000), upper=(800, 800, 1400), cext=0.02, ssa=0.0,
               sun=(20.0, 55.0), optical_thickness=threshold, leaf_block=1)
    s = sun_dir(20.0, 55.0)
    org = np.array([-400.0, 100.0, 0.0])
    tau = sum((tb - ta) * 0.02 * float(lwc[iz, iy, ix])
              for ta, tb, (ix, iy, iz) in cells_along(lwc, sc.lower, sc.upper, org, s, 0.0, 1e5))
    p = render(sc, sun_camera(org, s, 4, 1024, solver))
    expect = L_SUN_550 * math.exp(-tau) / 1e9
    E, SE = p["radiance_E"], p["radiance_SE"]
    assert abs(E.mean() - expect) <= 4 * math.sqrt((SE ** 2).sum()) / E.size + 2e-4 * expect


@pytest.mark.parametrize("solver", SOLVERS)
def test_lw_isothermal_enclosure(solver):
    T = 280.0
    sc = Scene(bands=LW_BANDS, nlayers=50, dz=400.0, ka=np.full((1, 50), 2e-5), T=np.full(50, T),
               ground=("plane", 0.0), albedo=0.0, ground_T=T)
    d = camera(8, 8, 16, pos=(0, 0, 5000), tgt=(0.2, -0.1, 4999), fov=50.0,
               wlen=(LW_WLEN, LW_WLEN), spectral=F.SPECTRAL_LW, Tref=T, solver=solver)
    p = render(sc, d)
    assert np.allclose(p["radiance_E"], planck(LW_WLEN * 1e-9, T) / 1e9, rtol=DET_RTOL[solver], atol=0)
    # Monochromatic map: the reference inverts the per-nm estimate (pdf 1 nm^-1
    # = 1e9 m^-1, ..._draw_map.c:379,399) as a per-metre radiance
    # (:425-435), so T_b is the temperature whose B_lambda is B_lambda(T)/1e9.
    lam = LW_WLEN * 1e-9
    E = p["radiance_E"]
    Tq = 6.62607015e-34 * 299792458.0 / (lam * 1.380649e-23 * np.log1p(
        2 * 6.62607015e-34 * 299792458.0 ** 2 / (lam ** 5 * E)))
    assert np.allclose(p["temperature_E"], Tq, atol=2e-3)
    # Band map with T_ref = T: each realisation is B_lambda(T) / pdf(lambda),
    # nearly the band integral, and T_b is T (bisection to 1e-4 K; the
    # 1-nm discrete pdf and the reference's sigma leave < 0.05 K)
    d = camera(8, 8, 16, pos=(0, 0, 5000), tgt=(0.2, -0.1, 4999), fov=50.0,
               wlen=LW_BANDS[0], spectral=F.SPECTRAL_LW, Tref=T, solver=solver)
    p = render(sc, d)
    assert np.allclose(p["temperature_E"], T, atol=0.05)


@pytest.mark.parametrize("solver", SOLVERS)
def test_lw_thick_scattering_enclosure(solver):
    T = 265.0
    rng = np.random.default_rng(3)
    lwc = np.exp(rng.normal(-1.0, 0.8, (4, 8, 8))).astype(np.float32)
    lwc[rng.random(lwc.shape) < 0.3] = 0.0
    sc = Scene(bands=LW_BANDS, nlayers=50, dz=400.0, ka=np.full((1, 50), 1e-2), T=np.full(50, T),
               lwc=lwc, lower=(0, 0, 1000), upper=(800, 800, 1400), cext=0.3, ssa=0.6, g=0.8,
               ground=("plane", 0.0), albedo=0.0, ground_T=T)
    d = camera(8, 8, 16, pos=(300, 300, 1200), tgt=(300.1, 300.3, 1201), fov=70.0,
               wlen=(LW_WLEN, LW_WLEN), spectral=F.SPECTRAL_LW, Tref=T, solver=solver)
    p = render(sc, d)
    assert np.allclose(p["radiance_E"], planck(LW_WLEN * 1e-9, T) / 1e9, rtol=DET_RTOL[solver], atol=0)


@pytest.mark.parametrize("solver", SOLVERS)
def test_lw_absorbing_slab_upward(solver):
    T = 250.0
    ka = np.full((1, 50), 1.5e-5)
    sc = Scene(bands=LW_BANDS, nlayers=50, dz=400.0, ka=ka, T=np.full(50, T))
    d = np.array([0.3, 0.0, 1.0])
    d /= np.linalg.norm(d)
    desc = camera(4, 4, 2048, pos=(0, 0, 1.0), tgt=tuple(np.array([0, 0, 1.0]) + d), fov=1e-4,
                  wlen=(LW_WLEN, LW_WLEN), spectral=F.SPECTRAL_LW, Tref=T, solver=solver)
    p = render(sc, desc)
    B = planck(LW_WLEN * 1e-9, T)
    tau = ka.sum() * 400.0 - 1.5e-5
    check_estimates(p["radiance_E"], p["radiance_SE"], B * (1 - math.exp(-tau / d[2])) / 1e9)


@pytest.mark.parametrize("solver", SOLVERS)
def test_flux_sw_vacuum(solver):
    sc = Scene(sun=(45.0, 35.0))
    p = render(sc, rectangle(8, 8, 64, pos=(0, 0, 2000.0), size=(1000.0, 1000.0), solver=solver))
    E, SE = A.estimate(p["flux_sum"], p["flux_sum2"], p["flux_n"])
    assert np.all(p["flux_n"] == 64)
    check_estimates(E, SE, sun_solid_angle() * L_SUN_550 * cone_mean_cos(35.0) / 1e9)


@pytest.mark.parametrize("solver", SOLVERS)
def test_flux_lw_thick_isothermal(solver):
    T = 275.0
    sc = Scene(bands=LW_BANDS, nlayers=50, dz=400.0, ka=np.full((1, 50), 1e-2), T=np.full(50, T))
    p = render(sc, rectangle(8, 8, 16, pos=(0, 0, 2000.0), wlen=(LW_WLEN, LW_WLEN),
                             spectral=F.SPECTRAL_LW, Tref=T, solver=solver))
    E, SE = A.estimate(p["flux_sum"], p["flux_sum2"], p["flux_n"])
    assert np.allclose(E, math.pi * planck(LW_WLEN * 1e-9, T) / 1e9, rtol=DET_RTOL[solver], atol=0)


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("config", [0, 2])
def test_looser_majorants_leave_the_estimate_unchanged(solver, config):
    """Null-collision estimators are unbiased for any majorant: the octree
    built with threshold 0 (per-block bounds) and with threshold 1e9 (one
    leaf: the whole grid's bounds) give the same images (full SW walks with
    delta tracking on Ks, LW walks on Kext), while the loose tree draws many
    more tentative collisions."""
    out = []
    for T in (0.0, 1e9):
        s = A.Synth(config, seed=2, optical_thickness=T, cloud_def=(32, 32, 16),
                    cloud_upper=(3200.0, 3200.0, 2600.0))
        if config == 0:
            d = A.map_preset(0, width=16, height=16, spp=256, solver=solver, pos=(1600.0, -1000.0, 3000.0),
                             tgt=(1600.0, -999.5, 2999.134))
        else:
            d = A.map_preset(2, width=16, height=16, spp=128, solver=solver, pos=(1600.0, 1600.0, 2000.0),
                             tgt=(1600.0, 1600.0, 2001.0), rect_size=(3200.0, 3200.0))
        with A.Context(s.scene_ptr) as ctx:
            ctx.set_counting(True)
            p = ctx.draw_map(d)
            out.append((p, ctx.counters()["ncoll"], ctx.svo_info()[2]))
        s.close()
    (a, ca, la), (b, cb, lb) = out
    assert lb == 1 and la > 100 and cb > 1.5 * ca
    if config == 0:
        E1, S1, E2, S2 = a["radiance_E"], a["radiance_SE"], b["radiance_E"], b["radiance_SE"]
    else:
        (E1, S1), (E2, S2) = A.estimate(a["flux_sum"], a["flux_sum2"], a["flux_n"]), A.estimate(
            b["flux_sum"], b["flux_sum2"], b["flux_n"])
    z = (E1 - E2) / np.sqrt(S1 ** 2 + S2 ** 2)
    assert np.mean(np.abs(z) <= 3) >= 0.98 and abs(z.mean()) < 0.2


def pooled(E, SE):
    """Mean of equally sampled pixel estimates and its standard error"""
    return float(E.mean()), float(math.sqrt((SE ** 2).sum()) / E.size)


def narrow_camera(pos, d, n, spp, solver, **kw):
    """n x n pixels all looking along d (fov 1e-4 deg); straight-down views
    take +y as the camera's up"""
    d = np.asarray(d, np.float64) / np.linalg.norm(d)
    up = (0, 1, 0) if abs(d[2]) > 0.99 else (0, 0, 1)
    return camera(n, n, spp, pos=tuple(pos), tgt=tuple(np.asarray(pos) + d), up=up, fov=1e-4,
                  solver=solver, **kw)


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("view", sorted(CF.THIN_VIEWS))
@pytest.mark.parametrize("medium", sorted(CF.THIN_MEDIA))
def test_single_scattering_thin_slab(medium, view, solver):
    """SW volume scattering on the device: the first-order closed form of the
    thin slab (HG particles, Rayleigh gas, 3:1 mix; 20 and 150 degrees from
    the sun), 4.2 M walks. Tolerance: 4 pooled standard errors + 0.5 % for
    second-order scattering; a flipped scattering angle or swapped
    gas/particle odds lie more than 5 tolerances away."""
    ks_gas, ks_par = CF.THIN_MEDIA[medium]
    pos, d, _ = CF.THIN_VIEWS[view]
    p = render(CF.thin_slab_scene(ks_gas, ks_par), narrow_camera(pos, d, 16, 16384, solver))
    assert np.all(p["time_n"] == 16384)
    m, se = pooled(p["radiance_E"] * 1e9, p["radiance_SE"] * 1e9)  # per nm -> per m (pdf 1e9 m^-1)
    expect, alternatives = CF.thin_slab_expectations(L_SUN_550, ks_gas, ks_par, view)
    tol = 4 * se + CF.THIN_MS_RTOL * expect
    assert abs(m - expect) <= tol, (m, expect, se)
    for name, alt in alternatives.items():
        if abs(alt - expect) > 0.2 * expect:
            assert abs(m - alt) > 5 * tol, name


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("view", sorted(CF.LW_VIEWS))
@pytest.mark.parametrize("medium", ["gas", "cloud"])
def test_lw_emission_under_a_temperature_gradient(medium, view, solver):
    """LW emission at the absorption point's layer temperature on the device:
    ten layers 9 K apart, absorbing gas or an absorbing cloud through six of
    them, seen up and down from three heights, 262 k walks. Tolerance: 4
    pooled standard errors + 2e-5 (FP32 Planck); a layer index off by one or
    the starting layer's temperature lie more than 10 standard errors away."""
    pos, d, _ = CF.LW_VIEWS[view]
    desc = narrow_camera(pos, d, 8, 4096, solver, wlen=(CF.LW_WLEN_NM, CF.LW_WLEN_NM),
                         spectral=F.SPECTRAL_LW, Tref=280.0)
    p = render(CF.lw_layers_scene(medium), desc)
    m, se = pooled(p["radiance_E"] * 1e9, p["radiance_SE"] * 1e9)
    expect = CF.lw_layers_expectations(medium, view)
    assert abs(m - expect) <= 4 * se + 2e-5 * expect, (m, expect, se)
    for name, alt in CF.lw_layers_alternatives(medium, view).items():
        assert abs(m - alt) > 10 * se, name
