"""The oracle's SW / LW walks and transmissivity estimator on scenes with
closed-form answers.

Each case is chosen so that a plausible mistake fails it: a dropped cosine
or solid angle (ground), a wrong layer or path length (slab), a biased
k_min/k_max residual (cloud, against an exact cell-by-cell integral), a wrong
emission temperature or pdf (isothermal enclosure), a wrong flux weight
(rectangle sensor). Tolerances are stated per test: deterministic cases are
compared to 1e-12 relative, estimators to 4 standard errors.
"""
import math

import numpy as np
import pytest

import closed_forms as CF
import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from scenes import (Scene, camera, cells_along, planck, rectangle, sun_dir, sun_solid_angle,
                    SUN_HALF_ANGLE)

L_SUN_550 = planck(550e-9, 5778.0)  # W/m^2/sr/m


def walk(sky, pos, d, n, wlen=550.0, spectral=F.SPECTRAL_SW, mask=F.RADIANCE_ALL, pixel=0):
    pos = np.asarray(pos, np.float64)
    d = np.asarray(d, np.float64)
    d = d / np.linalg.norm(d)
    return np.array([O.lib().oracle_radiance(sky.h, 0, pixel, r, spectral, mask, O.dptr(pos),
                                             O.dptr(d), wlen, 0, 0) for r in range(n)])


def mean_se(x):
    return x.mean(), x.std(ddof=1) / math.sqrt(x.size) if x.size > 1 else 0.0


def cone_mean_cos(elevation_deg):
    """E[n . s] for s uniform over the sun cap and n = +z: the cap axis times
    (1 + cos half-angle)/2."""
    return math.sin(math.radians(elevation_deg)) * (1 + math.cos(SUN_HALF_ANGLE)) / 2


def gas_slab(nlayers=50, dz=400.0, ka0=1e-5, ks0=0.0, H=8000.0):
    zmid = (np.arange(nlayers) + 0.5) * dz
    return ka0 * np.exp(-zmid / H), ks0 * np.exp(-zmid / H)


def test_vacuum_lambertian_ground():
    """No medium, Lambertian ground of albedo 0.3, sun at 60 deg: each walk
    scores L_sun * Omega * albedo/pi * cos(sun sample) once, then Russian
    roulette sends it to space. Mean = L_sun Omega (albedo/pi) E[cos]."""
    sc = Scene(ground=("plane", 0.0), albedo=0.3, sun=(0.0, 60.0))
    sky = O.Sky(sc.ptr)
    w = walk(sky, (0, 0, 1000), (0.3, 0.2, -1.0), 4000)
    expect = L_SUN_550 * sun_solid_angle() * 0.3 / math.pi * cone_mean_cos(60.0)
    m, se = mean_se(w)
    # cos(N, s) spreads by cot(60 deg) * 4.65e-3 ~ 2.7e-3 relative over the cone
    assert abs(m - expect) <= 4 * se
    assert se < 1e-4 * expect
    lo = L_SUN_550 * sun_solid_angle() * 0.3 / math.pi * math.sin(math.radians(60) - SUN_HALF_ANGLE)
    assert w.min() >= lo * (1 - 1e-12) and w.max() <= expect * 1.01


def test_absorbing_slab_direct_sun():
    """Gas-only slab, looking straight at the sun: k_min = k_max in every
    segment, so the transmissivity is deterministic, exp(-tau/mu) with tau the
    sum of the layer coefficients times their thickness."""
    ka, ks = gas_slab(ka0=3e-5, ks0=0.0)
    sc = Scene(nlayers=50, dz=400.0, ka=ka[None, :], ks=ks[None, :], sun=(30.0, 50.0))
    sky = O.Sky(sc.ptr)
    s = sun_dir(30.0, 50.0)
    w = walk(sky, (100.0, -50.0, 0.0), s, 8)
    tau = ka.sum() * 400.0
    expect = L_SUN_550 * math.exp(-tau / s[2])
    assert np.allclose(w, expect, rtol=1e-12, atol=0)


def test_direct_component_uses_extinction():
    """With scattering gas and only the direct component requested, the sun
    term is attenuated by ka + ks (Kext, ..._sw.c:334) and nothing else is
    added."""
    ka, ks = gas_slab(ka0=1e-5, ks0=2e-5)
    sc = Scene(nlayers=50, dz=400.0, ka=ka[None, :], ks=ks[None, :], sun=(0.0, 70.0))
    sky = O.Sky(sc.ptr)
    s = sun_dir(0.0, 70.0)
    w = walk(sky, (0, 0, 10.0), s, 8, mask=F.RADIANCE_DIRECT)
    # the walk starts 10 m above the bottom of the first layer
    expect = L_SUN_550 * math.exp(-((ka + ks).sum() * 400.0 - (ka[0] + ks[0]) * 10.0) / s[2])
    assert np.allclose(w, expect, rtol=1e-12, atol=0)
    # away from the sun and without scattering, there is nothing to see
    sc2 = Scene(nlayers=50, dz=400.0, ka=ka[None, :], sun=(0.0, 70.0))
    sky2 = O.Sky(sc2.ptr)
    assert np.all(walk(sky2, (0, 0, 10.0), (0.0, 0.6, 0.8), 16) == 0.0)


def random_cloud(seed, shape=(4, 8, 8)):
    rng = np.random.default_rng(seed)
    lwc = np.exp(rng.normal(-1.0, 0.8, shape)).astype(np.float32)
    lwc[rng.random(shape) < 0.3] = 0.0
    return lwc


@pytest.mark.parametrize("threshold", [1e-6, 1.0, 1e3])
def test_absorbing_cloud_transmissivity(threshold):
    """Purely absorbing cloud (ssa = 0) seen against the sun: the k_min/k_max
    estimator's mean is exp(-tau/mu) with tau integrated exactly over the
    cells the ray crosses (brute-force grid crossing, no octree). The octree
    threshold changes k_min/k_max (finest leaves vs whole-grid bounds) but
    not the mean."""
    lwc = random_cloud(1)
    lower, upper = (0.0, 0.0, 1000.0), (800.0, 800.0, 1400.0)
    sc = Scene(lwc=lwc, lower=lower, upper=upper, cext=0.02, ssa=0.0, g=0.85,
               sun=(20.0, 55.0), optical_thickness=threshold, leaf_block=1)
    sky = O.Sky(sc.ptr)
    s = sun_dir(20.0, 55.0)
    org = np.array([-400.0, 100.0, 0.0])  # crosses the grid at x 260-520, y 340-435
    tau = sum((tb - ta) * 0.02 * float(lwc[iz, iy, ix])
              for ta, tb, (ix, iy, iz) in cells_along(lwc, sc.lower, sc.upper, org, s, 0.0, 1e5))
    assert tau > 0.5
    w = walk(sky, org, s, 20000, mask=F.RADIANCE_DIRECT)
    expect = L_SUN_550 * math.exp(-tau)
    m, se = mean_se(w)
    assert abs(m - expect) <= 4 * se
    # each realisation is 0 or exp(-Tmin) <= 1
    assert np.all((w == 0) | ((w >= expect * (1 - 1e-12)) & (w <= L_SUN_550)))


def test_uniform_cloud_zero_variance():
    """k_min = k_max (uniform LWC): the estimator returns exp(-tau) exactly
    in every realisation."""
    lwc = np.full((4, 4, 4), 0.7, np.float32)
    sc = Scene(lwc=lwc, lower=(0, 0, 1000), upper=(400, 400, 1400), cext=0.01, ssa=0.0,
               sun=(0.0, 80.0))
    sky = O.Sky(sc.ptr)
    s = sun_dir(0.0, 80.0)
    org = np.array([200.0, 200.0, 500.0])
    tau = 400.0 / s[2] * 0.01 * float(np.float32(0.7))  # LWC is stored in FP32
    w = walk(sky, org, s, 16, mask=F.RADIANCE_DIRECT)
    assert np.allclose(w, L_SUN_550 * math.exp(-tau), rtol=1e-12, atol=0)


LW_WLEN = 10000.0  # nm
LW_BANDS = ((9500.0, 10500.0),)


def test_lw_isothermal_enclosure_is_exact():
    """Isothermal gas over a black ground at the same temperature: every
    downward LW walk ends in emission at T, so each realisation is exactly
    B_lambda(T)."""
    T = 280.0
    sc = Scene(nlayers=50, dz=400.0, ka=np.full((1, 50), 2e-5), T=np.full(50, T), bands=LW_BANDS,
               ground=("plane", 0.0), albedo=0.0, ground_T=T)
    sky = O.Sky(sc.ptr)
    w = walk(sky, (0, 0, 5000.0), (0.2, -0.1, -1.0), 64, wlen=LW_WLEN, spectral=F.SPECTRAL_LW)
    assert np.allclose(w, planck(LW_WLEN * 1e-9, T), rtol=1e-12, atol=0)


def test_lw_thick_scattering_enclosure_is_exact():
    """Optically thick isothermal gas with scattering clouds and a black
    ground at T: absorption, ground emission and cloud scattering chains all
    end in B_lambda(T), whatever the path."""
    T = 265.0
    lwc = random_cloud(3)
    sc = Scene(nlayers=50, dz=400.0, ka=np.full((1, 50), 1e-2), T=np.full(50, T), bands=LW_BANDS,
               lwc=lwc, lower=(0, 0, 1000), upper=(800, 800, 1400), cext=0.3, ssa=0.6, g=0.8,
               ground=("plane", 0.0), albedo=0.0, ground_T=T)
    sky = O.Sky(sc.ptr)
    w = walk(sky, (300.0, 300.0, 1200.0), (0.1, 0.3, 1.0), 64, wlen=LW_WLEN, spectral=F.SPECTRAL_LW)
    assert np.allclose(w, planck(LW_WLEN * 1e-9, T), rtol=1e-12, atol=0)


def test_lw_absorbing_slab_upward():
    """Looking up through a purely absorbing isothermal slab with nothing
    beyond: emission with probability 1 - exp(-tau/mu), else 0."""
    T = 250.0
    ka = np.full((1, 50), 1.5e-5)
    sc = Scene(nlayers=50, dz=400.0, ka=ka, T=np.full(50, T), bands=LW_BANDS)
    sky = O.Sky(sc.ptr)
    d = np.array([0.3, 0.0, 1.0])
    d /= np.linalg.norm(d)
    w = walk(sky, (0, 0, 1.0), d, 20000, wlen=LW_WLEN, spectral=F.SPECTRAL_LW)
    B = planck(LW_WLEN * 1e-9, T)
    tau = ka.sum() * 400.0 - 1.5e-5 * 1.0  # the walk starts 1 m above the bottom
    expect = B * (1 - math.exp(-tau / d[2]))
    m, se = mean_se(w)
    assert abs(m - expect) <= 4 * se
    assert np.all((w == 0) | np.isclose(w, B, rtol=1e-14, atol=0))


def test_camera_map_vacuum_ground():
    """The same ground case through the pinhole camera and the pixel
    estimator: every pixel estimates L_sun Omega (albedo/pi) E[cos] / pdf,
    pdf = 1e9 m^-1 for a monochromatic map (..._draw_map.c:380,400)."""
    sc = Scene(ground=("plane", 0.0), albedo=0.3, sun=(0.0, 60.0))
    sky = O.Sky(sc.ptr)
    d = camera(8, 6, 16, pos=(0, 0, 3000), tgt=(0, 1, 2999), fov=40.0)
    acc = sky.draw_map(d, nthreads=4)
    from htrdr_kmin_amd.atmosphere import accum_estimate
    E, SE = accum_estimate(acc)
    expect = L_SUN_550 * sun_solid_angle() * 0.3 / math.pi * cone_mean_cos(60.0) / 1e9
    assert np.all(acc["n"] == 16)
    assert np.all(np.abs(E - expect) <= 5 * SE)
    assert abs(E.mean() - expect) <= 4 * math.sqrt((SE ** 2).sum()) / E.size


def test_flux_map_sw_vacuum():
    """SW flux through an upward rectangle in vacuum: cos(N, s) Omega L_sun
    from the sun leg, nothing diffuse (..._draw_map.c:279-301)."""
    sc = Scene(sun=(45.0, 35.0))
    sky = O.Sky(sc.ptr)
    d = rectangle(4, 4, 32, pos=(0, 0, 2000.0), size=(1000.0, 1000.0))
    acc = sky.draw_map(d, nthreads=4)
    from htrdr_kmin_amd.atmosphere import accum_estimate
    E, SE = accum_estimate(acc)
    expect = sun_solid_angle() * L_SUN_550 * cone_mean_cos(35.0) / 1e9
    assert np.all(np.abs(E - expect) <= 5 * SE)
    assert abs(E.mean() - expect) <= 4 * math.sqrt((SE ** 2).sum()) / E.size


def test_flux_map_lw_thick_isothermal():
    """LW flux through an upward rectangle under an optically thick
    isothermal sky: pi B(T) exactly (w pi / pdf, ..._draw_map.c:266-268)."""
    T = 275.0
    sc = Scene(nlayers=50, dz=400.0, ka=np.full((1, 50), 1e-2), T=np.full(50, T), bands=LW_BANDS)
    sky = O.Sky(sc.ptr)
    d = rectangle(4, 4, 16, pos=(0, 0, 2000.0), wlen=(LW_WLEN, LW_WLEN), spectral=F.SPECTRAL_LW,
                  Tref=T)
    acc = sky.draw_map(d, nthreads=4)
    from htrdr_kmin_amd.atmosphere import accum_estimate
    E, SE = accum_estimate(acc)
    assert np.allclose(E, math.pi * planck(LW_WLEN * 1e-9, T) / 1e9, rtol=1e-12, atol=0)
    assert np.all(SE <= 1e-12 * E)


def walk_many(sky, pos, d, n, wlen=550.0, spectral=F.SPECTRAL_SW, mask=F.RADIANCE_ALL):
    pos = np.asarray(pos, np.float64)
    d = np.asarray(d, np.float64)
    d = d / np.linalg.norm(d)
    out = np.zeros(n)
    O.lib().oracle_radiance_many(sky.h, 0, 0, 0, n, spectral, mask, O.dptr(pos), O.dptr(d), wlen, 0, 0,
                                 O.dptr(out), 0)
    return out


@pytest.mark.parametrize("view", sorted(CF.THIN_VIEWS))
@pytest.mark.parametrize("medium", sorted(CF.THIN_MEDIA))
def test_single_scattering_thin_slab(medium, view):
    """SW volume scattering (..._sw.c:422-466) in an optically thin slab:
    the mean radiance is the first-order closed form L_sun Omega
    sum_c ks_c <p_c(dir . sun)> * path integral, for HG particles, Rayleigh
    gas and a 3:1 gas/particle mix, 20 and 150 degrees from the sun. The
    same data reject what a flipped scattering-angle convention (100x off
    for HG) or swapped gas/particle odds would give by many standard errors.
    Tolerance: 4 standard errors + 0.5 % for second-order scattering."""
    ks_gas, ks_par = CF.THIN_MEDIA[medium]
    sky = O.Sky(CF.thin_slab_scene(ks_gas, ks_par).ptr)
    pos, d, _ = CF.THIN_VIEWS[view]
    w = walk_many(sky, pos, d, 4_000_000)
    m, se = mean_se(w)
    expect, alternatives = CF.thin_slab_expectations(L_SUN_550, ks_gas, ks_par, view)
    tol = 4 * se + CF.THIN_MS_RTOL * expect
    assert abs(m - expect) <= tol, (m, expect, se)
    assert se < 0.03 * expect
    for name, alt in alternatives.items():
        if abs(alt - expect) > 0.2 * expect:  # Rayleigh alone is even in mu
            assert abs(m - alt) > 5 * tol, name
    # a walk scores at most once here (single scattering dominates): every
    # non-zero weight is one next-event term L_sun Omega p Tr_abs Tr
    assert np.mean(w > 0) < 3e-3


@pytest.mark.parametrize("view", sorted(CF.LW_VIEWS))
@pytest.mark.parametrize("medium", ["gas", "cloud"])
def test_lw_emission_under_a_temperature_gradient(medium, view):
    """LW absorption emits B_lambda(T) at the temperature of the layer
    holding the absorption point (..._lw.c:221-227): ten layers 9 K apart,
    absorbing gas or an absorbing cloud across six of them, seen up and
    down, from the ground, the top and mid-layer, over a black ground at
    305 K. Expected: sum_i B(T_i)(1 - exp(-tau_i/mu)) times the
    transmissivity of the layers before it. The same data reject a layer
    index off by one either way and the starting layer's temperature.
    Tolerance: 4 standard errors."""
    sky = O.Sky(CF.lw_layers_scene(medium).ptr)
    pos, d, _ = CF.LW_VIEWS[view]
    w = walk_many(sky, pos, d, 200_000, wlen=CF.LW_WLEN_NM, spectral=F.SPECTRAL_LW)
    m, se = mean_se(w)
    expect = CF.lw_layers_expectations(medium, view)
    assert abs(m - expect) <= 4 * se, (m, expect, se)
    for name, alt in CF.lw_layers_alternatives(medium, view).items():
        assert abs(m - alt) > 10 * se, name
    # each realisation is one emission: B of a layer's or the ground's temperature, or 0
    Bs = np.r_[planck(CF.LW_WLEN_NM * 1e-9, CF.LW_T), planck(CF.LW_WLEN_NM * 1e-9, CF.LW_TG), 0.0]
    assert np.all(np.min(np.abs(w[:, None] - Bs[None, :]) / Bs.max(), axis=1) < 1e-13)
