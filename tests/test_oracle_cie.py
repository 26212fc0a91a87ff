"""The oracle's CIE 1931 XYZ wavelength sampler (src/core/htrdr_ran_wlen_cie_xyz.c)
and CIE XYZ images (draw_pixel_image, src/atmosphere/htrdr_atmosphere_draw_map.c:
110-207), pinned by

* the colour matching functions against the CIE 1931 2-degree observer table
  (the analytic fits of Wyman, Sloan and Shirley, JCGT 2013, are within 0.012
  of it) and their integrals (about 106.86 nm each, the equal-energy white);
* the sampler: the band a sample falls in is the first whose cumulative
  (trapezoidal integrals of the fit, computed here independently) exceeds r0,
  the sample solves the band's linear-density CDF for r1 exactly, the pdf is
  1 / integral, and the samples follow the fit (Kolmogorov-Smirnov);
* the image: a vacuum over a Lambertian ground gives X, Y, Z in closed form
  (integral of the fit times the reflected sun's spectrum).
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from scenes import Scene, camera, planck, sun_solid_angle, SUN_HALF_ANGLE

L = O.lib()

# CIE 1931 2-degree standard observer (x, y, z bar)
CIE_1931 = {420: (0.1344, 0.0040, 0.6456), 450: (0.3362, 0.0380, 1.7721),
            500: (0.0049, 0.3230, 0.2720), 550: (0.4334, 0.9950, 0.0087),
            555: (0.5121, 1.0000, 0.0057), 600: (1.0622, 0.6310, 0.0008),
            650: (0.2835, 0.1070, 0.0000), 700: (0.0114, 0.0041, 0.0000)}


def fit(c, lam):
    return L.oracle_cie_fit(c, float(lam))


@pytest.mark.parametrize("lam", sorted(CIE_1931))
def test_fits_match_cie_1931_table(lam):
    for c in range(3):
        assert abs(fit(c, lam) - CIE_1931[lam][c]) < 0.012, (lam, c)


def test_fit_integrals_are_equal_energy_white():
    for c in range(3):
        I = integrate.quad(lambda x: fit(c, x), 380, 780, limit=200)[0]
        assert abs(I - 106.86) < 0.3, (c, I)


class Sampler:
    def __init__(self, rng, channel):
        self.h = C.c_void_p()
        r = np.array(rng, np.float64)
        n = int(round(rng[1] - rng[0]))
        assert L.oracle_wlen_cie_create(O.dptr(r), n, channel, C.byref(self.h)) == 0
        self.range, self.n, self.c = rng, n, channel

    def sample(self, r0, r1):
        pdf = C.c_double()
        lam = L.oracle_wlen_planck_sample(self.h, float(r0), float(r1), C.byref(pdf))
        return lam, pdf.value

    def close(self):
        L.oracle_wlen_planck_destroy(self.h)


def trapezoid_cdf(c, lo, hi, n):
    """Cumulative of the per-band trapezoidal integrals (2 sub-intervals per
    1-nm band), normalised"""
    edges = lo + (hi - lo) / n * np.arange(n + 1)
    mids = 0.5 * (edges[:-1] + edges[1:])
    f = lambda x: np.array([fit(c, v) for v in x])
    I = 0.25 * (edges[1:] - edges[:-1]) * (f(edges[:-1]) + 2 * f(mids) + f(edges[1:]))
    return np.cumsum(I) / I.sum(), I.sum()


@pytest.mark.parametrize("channel", [0, 1, 2])
def test_sampler_band_and_in_band_inversion(channel):
    lo, hi = 380.0, 780.0
    s = Sampler((lo, hi), channel)
    cdf, total = trapezoid_cdf(channel, lo, hi, s.n)
    rng = np.random.default_rng(channel)
    for r0, r1 in rng.random((400, 2)):
        lam, pdf = s.sample(r0, r1)
        assert math.isclose(pdf, 1.0 / total, rel_tol=1e-12)
        band = int(np.searchsorted(cdf, r0, side="right"))
        assert band == int(math.floor(lam - lo)), (r0, lam)
        # the linear density between the fit's values at the band ends
        a, b = lo + band, lo + band + 1.0
        fa, fb = fit(channel, a), fit(channel, b)
        x = lam - a
        F_lin = (fa * x + 0.5 * (fb - fa) * x * x) / (0.5 * (fa + fb))
        assert abs(F_lin - r1) < 1e-9, (r1, F_lin)
    s.close()


@pytest.mark.parametrize("channel", [0, 1, 2])
def test_sampler_follows_the_fit(channel):
    s = Sampler((380.0, 780.0), channel)
    n = 20000
    u = (np.arange(n) + 0.5) / n
    r1 = np.random.default_rng(7).random(n)
    lam = np.sort([s.sample(a, b)[0] for a, b in zip(u, r1)])
    s.close()
    total = integrate.quad(lambda x: fit(channel, x), 380, 780, limit=200)[0]
    grid = np.linspace(380, 780, 81)
    cdf = np.array([integrate.quad(lambda x: fit(channel, x), 380, g, limit=200)[0] for g in grid]) / total
    emp = np.searchsorted(lam, grid, side="right") / n
    assert np.max(np.abs(emp - cdf)) < 2e-3


def test_sampler_rejects_ranges_outside_the_colour_space():
    h = C.c_void_p()
    for rng in ((370.0, 700.0), (400.0, 790.0)):
        r = np.array(rng, np.float64)
        assert L.oracle_wlen_cie_create(O.dptr(r), int(rng[1] - rng[0]), 0, C.byref(h)) != 0


def vacuum_xyz(albedo, elevation):
    """X, Y, Z of a camera looking at a Lambertian ground in vacuum, W/m^2/sr:
    integral over lambda of fit * rho/pi * L_sun * Omega_sun * sin(el) *
    (1 + cos(half angle)) / 2 (the sun cone's mean cosine)."""
    k = albedo / math.pi * sun_solid_angle() * math.sin(math.radians(elevation)) * (
        1 + math.cos(SUN_HALF_ANGLE)) / 2
    return [k * 1e-9 * integrate.quad(lambda x: fit(c, x) * planck(x * 1e-9, 5778.0), 380, 780,
                                       limit=400)[0] for c in range(3)]


def test_cie_image_vacuum_closed_form():
    sc = Scene(bands=((380.0, 780.0),), ground=("plane", 0.0), albedo=0.4, sun=(0.0, 50.0))
    sky = O.Sky(sc.ptr)
    # looking down: every pixel sees the ground
    d = camera(4, 4, 1500, (10.0, 20.0, 1000.0), (13.0, 22.0, 0.0), wlen=(380.0, 780.0),
               spectral=F.SPECTRAL_SW_CIE_XYZ)
    acc = sky.draw_map(d)
    assert acc.shape == (3, 4, 4)
    expect = vacuum_xyz(0.4, 50.0)
    for c in range(3):
        a = acc[c].ravel()
        assert np.all(a["n"] == 1500)
        w_sum, w2_sum, n = a["sum"].sum(), a["sum2"].sum(), a["n"].sum()
        m = w_sum / n
        se = math.sqrt(max(w2_sum / n - m * m, 0) / n)
        assert abs(m - expect[c]) <= 4 * se and se < 0.02 * expect[c], (c, m, expect[c], se)
    sky.close()


def test_cie_channels_draw_their_own_streams():
    """The plane of channel c sums the realisations whose index carries c in
    bits 32-33 (streams keyed by seed ^ (c+1) << 56); the three channels
    differ; there is no fourth channel."""
    sc = Scene(bands=((380.0, 780.0),), ground=("plane", 0.0), albedo=0.4)
    sky = O.Sky(sc.ptr)
    d = camera(4, 4, 8, (10.0, 20.0, 1000.0), (13.0, 22.0, 0.0), wlen=(380.0, 780.0),
               spectral=F.SPECTRAL_SW_CIE_XYZ, seed=5)
    acc = sky.draw_map(d)
    w = [[L.oracle_realisation(sky.h, C.byref(d), 6, (c << 32) | r) for r in range(8)] for c in range(3)]
    for c in range(3):
        a = acc[c].ravel()[6]
        assert math.isclose(a["sum"] + a["sum_c"], math.fsum(w[c]), rel_tol=1e-12)
    assert len({round(x[0], 12) for x in w}) == 3
    assert math.isnan(L.oracle_realisation(sky.h, C.byref(d), 6, 3 << 32))
    sky.close()
