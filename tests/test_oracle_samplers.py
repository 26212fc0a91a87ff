"""The oracle's direction samplers and phase functions against their analytic
distributions: Kolmogorov-Smirnov on the polar cosine and on the azimuth
(p > 1e-3 at fixed seeds), normalisation and mean cosine of the phase
functions."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate, stats

import oracle_ffi as O
from scenes import Scene, SUN_HALF_ANGLE, sun_dir

L = O.lib()
N = 20000


def sample(sky, kind, param, d, seed):
    u = np.random.default_rng(seed).random((N, 2))
    out = np.zeros((N, 3))
    d = np.asarray(d, np.float64) / np.linalg.norm(d)
    assert L.oracle_sample_dirs(sky.h, kind, param, O.dptr(d), O.dptr(u), N, O.dptr(out)) == 0
    return out, d


def azimuths(out, d):
    a = np.array([1.0, 0, 0]) if abs(d[0]) < 0.9 else np.array([0, 1.0, 0])
    e1 = np.cross(d, a)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(d, e1)
    return (np.arctan2(out @ e2, out @ e1) / (2 * math.pi)) % 1.0


def hg_cdf(g):
    if abs(g) < 1e-6:
        return lambda mu: (mu + 1) / 2
    return lambda mu: (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * mu) - 1 / (1 + g))


@pytest.fixture(scope="module")
def sky():
    s = O.Sky(Scene(sun=(40.0, 25.0)).ptr)
    yield s
    s.close()


@pytest.mark.parametrize("g", [0.85, 0.3, 0.0, -0.5])
def test_hg_sampling(sky, g):
    out, d = sample(sky, 0, g, (0.3, -0.2, 0.9), 1)
    assert np.allclose(np.linalg.norm(out, axis=1), 1, atol=1e-12)
    mu = out @ d
    assert stats.kstest(mu, hg_cdf(g)).pvalue > 1e-3
    assert stats.kstest(azimuths(out, d), "uniform").pvalue > 1e-3
    assert abs(mu.mean() - g) < 4 * mu.std() / math.sqrt(N)


def test_rayleigh_sampling(sky):
    out, d = sample(sky, 1, 0.0, (0.0, 0.0, -1.0), 2)
    mu = out @ d
    assert stats.kstest(mu, lambda m: 3 / 8 * (m + m ** 3 / 3) + 0.5).pvalue > 1e-3
    assert stats.kstest(azimuths(out, d), "uniform").pvalue > 1e-3


def test_sun_cone_sampling(sky):
    """Uniform over the cap of half-angle 4.6524e-3 rad around the sun
    direction: the cosine to the axis is uniform on [cos h, 1]."""
    out, _ = sample(sky, 2, 0.0, (0, 0, 1), 3)
    s = sun_dir(40.0, 25.0)
    mu = out @ s
    ch = math.cos(SUN_HALF_ANGLE)
    assert mu.min() >= ch - 1e-15
    assert stats.kstest((mu - ch) / (1 - ch), "uniform").pvalue > 1e-3
    assert stats.kstest(azimuths(out, s), "uniform").pvalue > 1e-3


def test_cosine_hemisphere_sampling(sky):
    out, d = sample(sky, 3, 0.0, (0.2, 0.5, 0.8), 4)
    mu = out @ d
    assert mu.min() >= 0
    assert stats.kstest(mu ** 2, "uniform").pvalue > 1e-3
    assert stats.kstest(azimuths(out, d), "uniform").pvalue > 1e-3


@pytest.mark.parametrize("kind,param", [(0, 0.85), (0, -0.4), (0, 0.0), (1, 0.0)])
def test_phase_functions_normalised(kind, param):
    """2 pi * integral of the phase function over mu in [-1, 1] is 1, and the
    HG mean cosine is g."""
    f = lambda mu: L.oracle_phase_eval(kind, param, mu)
    assert abs(2 * math.pi * integrate.quad(f, -1, 1, epsrel=1e-12)[0] - 1) < 1e-9
    if kind == 0:
        assert abs(2 * math.pi * integrate.quad(lambda m: m * f(m), -1, 1, epsrel=1e-12)[0] - param) < 1e-9
