"""The oracle's spectral functions (src/core/htrdr_spectral.h, .c) and Planck
wavelength sampler (src/core/htrdr_ran_wlen_planck.c) against closed forms,
numerical quadrature and brute force."""
import math

import numpy as np
import pytest
from scipy import constants as K
from scipy import integrate

import oracle_ffi as O

L = O.lib()
SIGMA_REF = 5.6696e-8  # the reference's Stefan-Boltzmann constant (htrdr_spectral.h:141)


def B(lam, T):
    """Planck's law from scipy.constants (W/m^2/sr/m)."""
    return 2 * K.h * K.c ** 2 / lam ** 5 / math.expm1(K.h * K.c / (lam * K.k * T))


def quad_B(l0, l1, T):
    return integrate.quad(lambda x: B(x, T), l0, l1, epsabs=0, epsrel=1e-12, limit=400)[0]


@pytest.mark.parametrize("lam,T", [(550e-9, 5778.0), (10e-6, 288.0), (4e-6, 220.0), (20e-6, 310.0),
                                   (300e-9, 6000.0)])
def test_planck_monochromatic(lam, T):
    assert math.isclose(L.oracle_planck_monochromatic(lam, T), B(lam, T), rel_tol=1e-13)


@pytest.mark.parametrize("T", [250.0, 288.0, 5778.0])
def test_planck_peak_is_wien(T):
    lam = np.linspace(0.5, 1.5, 20001) * K.Wien / T
    v = np.array([L.oracle_planck_monochromatic(x, T) for x in lam])
    assert abs(lam[v.argmax()] * T / K.Wien - 1) < 1e-4


@pytest.mark.parametrize("T", [220.0, 288.0, 5778.0])
def test_blackbody_fraction_is_integral(T):
    """Fraction of sigma T^4 / pi in [l0, l1], against quadrature of Planck's
    law; the reference's series truncation and its constants (C2 =
    1.43877735e-2) hold it to ~1e-6 absolute."""
    total = K.Stefan_Boltzmann * T ** 4 / math.pi
    for l0, l1 in [(0.2e-6, 0.7e-6), (4e-6, 20e-6), (8e-6, 12e-6), (1e-6, 100e-6), (3e-6, 3.5e-6)]:
        f = L.oracle_blackbody_fraction(l0, l1, T)
        assert abs(f - quad_B(l0, l1, T) / total) < 2e-6
    # the whole spectrum holds the whole emission
    assert abs(L.oracle_blackbody_fraction(1e-9, 1.0, T) - 1) < 1e-6


def test_wiebelt_is_fraction_beyond_v():
    """Both branches (polynomial for v < 2, exponential series above) against
    15/pi^4 * integral_v^inf x^3/(e^x - 1) dx: each within 2e-6, the largest
    error sitting on either side of the switch at v = 2."""
    def exact(v):
        return 15 / math.pi ** 4 * integrate.quad(lambda x: x ** 3 / math.expm1(x), v, 200,
                                                  epsrel=1e-13)[0]
    for v in (0.01, 0.5, 1.0, 1.5, 1.999999, 2.0, 2.5, 4.0, 8.0, 20.0):
        assert abs(L.oracle_wiebelt(v) - exact(v)) < 2e-6
    assert L.oracle_wiebelt(0.0) == 1.0
    assert L.oracle_wiebelt(60.0) < 1e-20


@pytest.mark.parametrize("band,T", [((8e-6, 12e-6), 288.0), ((4e-6, 20e-6), 250.0),
                                    ((0.5e-6, 0.6e-6), 5778.0)])
def test_band_planck_is_band_mean(band, T):
    """oracle_planck over a band = its mean spectral radiance, up to the
    reference's sigma (5.6696e-8 vs CODATA 5.670374e-8, 1.4e-4 relative)."""
    l0, l1 = band
    mean = quad_B(l0, l1, T) / (l1 - l0)
    v = L.oracle_planck(l0, l1, T)
    assert math.isclose(v, mean * SIGMA_REF / K.Stefan_Boltzmann, rel_tol=3e-6)
    assert L.oracle_planck(l0, l0, T) == L.oracle_planck_monochromatic(l0, T)


@pytest.mark.parametrize("lam", [4e-6, 10e-6, 15e-6])
@pytest.mark.parametrize("T", [180.0, 255.5, 300.0, 340.0])
def test_brightness_temperature_monochromatic_inverse(lam, T):
    """Mono case: the closed-form inverse of Planck's law,
    T = hc / (lambda k ln(1 + 2 h c^2 / (lambda^5 B)))."""
    rad = L.oracle_planck_monochromatic(lam, T)
    Tb = K.h * K.c / (lam * K.k * math.log1p(2 * K.h * K.c ** 2 / (lam ** 5 * rad)))
    assert abs(Tb - T) < 1e-9
    assert abs(L.oracle_radiance_temperature(lam, lam, rad) - T) < 2e-4


@pytest.mark.parametrize("band", [(4e-6, 20e-6), (8e-6, 12e-6)])
@pytest.mark.parametrize("T", [200.0, 288.0, 330.0])
def test_brightness_temperature_round_trip(band, T):
    """Band case: radiance_temperature inverts the band-integrated radiance to
    the bisection's 1e-4 K (htrdr_spectral.c:31-89)."""
    l0, l1 = band
    integrated = L.oracle_planck(l0, l1, T) * (l1 - l0)
    assert abs(L.oracle_radiance_temperature(l0, l1, integrated) - T) < 2e-4


def test_bands_count():
    assert L.oracle_spectral_bands_count(O.dptr(np.array([4000.0, 20000.0]))) == 16000
    assert L.oracle_spectral_bands_count(O.dptr(np.array([550.0, 550.0]))) == 0
    assert L.oracle_spectral_bands_count(O.dptr(np.array([500.0, 600.4]))) == 100


class Sampler:
    def __init__(self, lo, hi, nbands, T):
        import ctypes as C
        self.h = C.c_void_p()
        rng = np.array([lo, hi])
        assert L.oracle_wlen_planck_create(O.dptr(rng), nbands, T, C.byref(self.h)) == 0

    def sample(self, r0, r1):
        import ctypes as C
        pdf = C.c_double()
        lam = L.oracle_wlen_planck_sample(self.h, r0, r1, C.byref(pdf))
        return lam, pdf.value

    def __del__(self):
        L.oracle_wlen_planck_destroy(self.h)


def test_wlen_sampler_mono():
    s = Sampler(550.0, 550.0, 0, 5778.0)
    for r0 in (0.0, 0.3, 0.999):
        assert s.sample(r0, 0.7) == (550.0, 1.0)


def test_wlen_sampler_discrete_brute_force():
    """5 one-nm bands: band probabilities are the normalised Planck integrals
    (quadrature), r0 picks the first band whose cumulative exceeds it, the
    wavelength is uniform in the band and pdf = p_band / band width."""
    lo, hi, T = 500.0, 505.0, 5778.0
    s = Sampler(lo, hi, 5, T)
    w = np.array([quad_B((lo + i) * 1e-9, (lo + i + 1) * 1e-9, T) for i in range(5)])
    p = w / w.sum()
    cdf = np.cumsum(p)
    for i in range(5):
        r0 = (cdf[i - 1] if i else 0.0) + 0.5 * p[i]
        lam, pdf = s.sample(r0, 0.25)
        assert lam == pytest.approx(lo + i + 0.25, abs=1e-12)
        assert pdf == pytest.approx(p[i], rel=1e-6)
    # just above / below a cumulative boundary (the series CDF is within 1e-6)
    assert s.sample(cdf[1] + 1e-5, 0.0)[0] == pytest.approx(lo + 2, abs=1e-9)
    assert s.sample(cdf[1] - 1e-5, 0.0)[0] == pytest.approx(lo + 1, abs=1e-9)
    # the pdf integrates to 1 over the range
    tot = sum(s.sample((cdf[i - 1] if i else 0.0) + 0.5 * p[i], 0.5)[1] * 1.0 for i in range(5))
    assert tot == pytest.approx(1.0, rel=1e-9)


def test_wlen_sampler_continuous_inverts_cdf():
    """No discrete bands (the continuous branch): the sampled wavelength
    solves CDF(lambda) = r0 with the Planck CDF over the range, and the pdf is
    B(lambda) / integral of B."""
    lo, hi, T = 8000.0, 12000.0, 288.0
    s = Sampler(lo, hi, 0, T)
    tot = quad_B(lo * 1e-9, hi * 1e-9, T)
    for r0 in (0.05, 0.3, 0.5, 0.77, 0.95):
        lam, pdf = s.sample(r0, 0.0)
        cdf = quad_B(lo * 1e-9, lam * 1e-9, T) / tot
        assert abs(cdf - r0) < 2e-5
        assert pdf == pytest.approx(B(lam * 1e-9, T) / tot * 1e-9, rel=2e-4)
