"""Closed forms shared by the oracle tests and the device tests.

Single scattering in an optically thin homogeneous slab lit by the sun cone,
and LW emission through horizontal slabs of different temperatures. Both
are written from the radiative transfer equation, independently of any
walk: they pin the SW scattering branch (the gas/particle choice by ks_gas /
ks, the phase function and its argument cos = dir . sun_dir, the next-event
weight R / pdf = p * Omega_sun, ..._sw.c:422-466) and the LW emission branch
(B_lambda(T) at the temperature of the absorption point, ..._lw.c:221-227).

Units: metres, m^-1, W/m^2/sr/m.
"""
import math

import numpy as np

from scenes import SUN_HALF_ANGLE, sun_solid_angle


def hg(g, mu):
    """Henyey-Greenstein phase function, normalised over the sphere (sr^-1)"""
    d = 1.0 + g * g - 2.0 * g * np.asarray(mu, np.float64)
    return (1.0 - g * g) / (4.0 * math.pi * d * np.sqrt(d))


def rayleigh(mu):
    """Rayleigh phase function, normalised over the sphere (sr^-1)"""
    mu = np.asarray(mu, np.float64)
    return 3.0 / (16.0 * math.pi) * (1.0 + mu * mu)


def sun_cone_mean(f, cos_theta0, half=SUN_HALF_ANGLE, nc=16, nphi=64):
    """Mean of f(d . s) over sun directions s uniform on the cap of half-angle
    `half` around an axis at angle theta0 from d: Gauss-Legendre in the cap's
    cosine (uniform in solid angle), periodic trapezoid in azimuth."""
    x, wgt = np.polynomial.legendre.leggauss(nc)
    c = 1.0 - (1.0 - math.cos(half)) * 0.5 * (x + 1.0)  # cos of the angle to the cap axis
    wgt = 0.5 * wgt
    phi = 2.0 * math.pi * (np.arange(nphi) + 0.5) / nphi
    s0 = math.sqrt(max(0.0, 1.0 - cos_theta0 ** 2))
    mu = c[:, None] * cos_theta0 + np.sqrt(1.0 - c[:, None] ** 2) * s0 * np.cos(phi)[None, :]
    return float((wgt[:, None] * f(mu)).sum() / nphi)


def single_scatter_slab(L_sun, H, kext, scatterers, d, s, looking_up):
    """First-order scattered radiance seen through a homogeneous slab of
    thickness H and extinction kext, sun cap around s, view direction d
    (unit vectors; from below the slab looking up, or from above looking
    down). scatterers: [(ks_c, phase_c)] with phase_c(mu) normalised over
    the sphere; mu = d . s_sample is the cosine of the scattering angle.

        L = L_sun Omega * sum_c ks_c <p_c> * int exp(-kext l - kext l_sun(l)) dl

    over the path length l inside the slab, l_sun(l) being the sun leg's
    length from the scattering point to the slab's top."""
    d = np.asarray(d, np.float64)
    s = np.asarray(s, np.float64)
    mu_v, mu_s = abs(d[2]), s[2]
    cos0 = float(d @ s)
    pk = sum(ks * sun_cone_mean(p, cos0) for ks, p in scatterers)
    L = H / mu_v  # path length inside the slab
    if looking_up:
        # scattering at height h = l mu_v above the base; sun leg (H - h)/mu_s
        a = kext * (1.0 - mu_v / mu_s)
        pre = math.exp(-kext * H / mu_s)
    else:
        # scattering at depth l mu_v below the top; sun leg l mu_v / mu_s
        a = kext * (1.0 + mu_v / mu_s)
        pre = 1.0
    integral = L if a == 0 else -math.expm1(-a * L) / a
    return L_sun * sun_solid_angle() * pk * pre * integral


def emission_through_slabs(pieces, z0, mu, looking_up, ground_B=0.0):
    """LW radiance seen from height z0 along a direction of |cos| = mu through
    horizontal pieces [(z_lo, z_hi, k_abs, B)] of purely absorbing medium,
    each at one temperature (B = B_lambda(T) of the piece), plus a black
    ground below them all, of emission ground_B, when looking down:

        L = sum_i B_i (1 - exp(-k_i l_i)) exp(-sum_{j before i} k_j l_j)
            + ground_B exp(-sum_j k_j l_j)

    with l_i the path length in piece i (Beer-Lambert per piece)."""
    segs = []
    for zlo, zhi, k, B in pieces:
        if looking_up:
            lo, hi = max(zlo, z0), zhi
            order = lo
        else:
            lo, hi = zlo, min(zhi, z0)
            order = -hi
        if hi > lo:
            segs.append((order, (hi - lo) / mu, k, B))
    segs.sort(key=lambda t: t[0])
    L, tau = 0.0, 0.0
    for _, length, k, B in segs:
        L += B * -math.expm1(-k * length) * math.exp(-tau)
        tau += k * length
    if not looking_up and ground_B:
        L += ground_B * math.exp(-tau)
    return L


def split_by_layers(z_lo, z_hi, k, layer_edges, layer_B):
    """The slab [z_lo, z_hi] of extinction k cut at the layer edges, each
    piece carrying the emission of the layer it lies in"""
    out = []
    for j in range(len(layer_edges) - 1):
        lo, hi = max(z_lo, layer_edges[j]), min(z_hi, layer_edges[j + 1])
        if hi > lo:
            out.append((lo, hi, k, layer_B[j]))
    return out


# --- Scenes and views of the thin-slab single-scattering cases -------------
# A slab z in [1000, 1100] m of vertical optical depth 1e-3: scattering gas
# (one Rayleigh layer), non-absorbing particles (a uniform cloud grid, HG
# g = 0.85, single-scattering albedo 1), or both; no ground, sun at 60 deg
# elevation. Second-order scattering adds < 0.5 % here (the oracle at 4e7
# walks per case agrees with the first-order form within its 0.4-0.6 %
# standard error), the allowance THIN_MS_RTOL.
THIN_Z0, THIN_H = 1000.0, 100.0
THIN_SUN = (0.0, 60.0)
THIN_G = 0.85
THIN_MS_RTOL = 5e-3
# forward: from the ground, 40 deg elevation towards the sun's azimuth, 20 deg
# from the sun; backward: from 2 km straight down, 150 deg from the sun
THIN_VIEWS = {
    "forward": ((0.0, 0.0, 0.0), (math.cos(math.radians(40.0)), 0.0, math.sin(math.radians(40.0))), True),
    "backward": ((0.0, 0.0, 2000.0), (0.0, 0.0, -1.0), False),
}
# (ks_gas, ks_particles) m^-1
THIN_MEDIA = {"particles": (0.0, 1e-5), "gas": (1e-5, 0.0), "mixed": (0.75e-5, 0.25e-5)}


def thin_slab_scene(ks_gas, ks_par):
    from scenes import Scene
    kw = {}
    if ks_gas:
        kw.update(nlayers=1, z_bottom=THIN_Z0, dz=THIN_H, ka=np.zeros((1, 1)), ks=np.full((1, 1), ks_gas))
    if ks_par:
        kw.update(lwc=np.ones((2, 4, 4), np.float32), lower=(-20000.0, -20000.0, THIN_Z0),
                  upper=(20000.0, 20000.0, THIN_Z0 + THIN_H), cext=ks_par, ssa=1.0, g=THIN_G)
    return Scene(sun=THIN_SUN, **kw)


def thin_slab_expectations(L_sun, ks_gas, ks_par, view):
    """(expected, {name: what a plausible mistake would give}) for one view:
    the phase functions evaluated at -mu (a flipped scattering-angle
    convention) and the gas/particle odds swapped"""
    from scenes import sun_dir
    _, d, up = THIN_VIEWS[view]
    s = sun_dir(*THIN_SUN)
    kext = ks_gas + ks_par
    hgp = lambda mu: hg(THIN_G, mu)  # noqa: E731

    def form(sc):
        return single_scatter_slab(L_sun, THIN_H, kext, [x for x in sc if x[0] > 0], d, s, up)

    expect = form([(ks_gas, rayleigh), (ks_par, hgp)])
    alt = {
        "flipped mu": form([(ks_gas, lambda mu: rayleigh(-mu)), (ks_par, lambda mu: hgp(-mu))]),
        "swapped odds": form([(ks_par, rayleigh), (ks_gas, hgp)]),
    }
    return expect, alt


# --- LW emission through layers of different temperatures -----------------
# 10 gas layers of 500 m, T = 300 - 9 K per layer, ka growing upwards (layer
# optical depths 0.05 to 0.275), a black ground at 305 K; or the same
# temperatures with no gas absorption and a purely absorbing uniform cloud
# over z in [1200, 3700] m (4e-4 m^-1), which cuts through six layers.
LW_NL, LW_DZ = 10, 500.0
LW_T = 300.0 - 9.0 * np.arange(LW_NL)
LW_KA = 1e-4 * (1.0 + 0.5 * np.arange(LW_NL))
LW_TG = 305.0
LW_CLOUD = (1200.0, 3700.0, 4e-4)
LW_WLEN_NM = 10000.0
LW_BAND = ((9500.0, 10500.0),)
# (position, direction, looking up)
LW_VIEWS = {
    "up_from_ground": ((0.0, 0.0, 0.0), (0.3, 0.2, 1.0), True),
    "up_from_mid_layer": ((0.0, 0.0, 2250.0), (0.0, 0.5, 1.0), True),
    "down_from_top": ((0.0, 0.0, 5000.0), (0.4, 0.0, -1.0), False),
    "down_from_mid_layer": ((0.0, 0.0, 2250.0), (0.0, -0.3, -1.0), False),
}


def lw_layers_scene(medium):
    from scenes import Scene
    common = dict(bands=LW_BAND, nlayers=LW_NL, dz=LW_DZ, T=LW_T, ground=("plane", 0.0), albedo=0.0,
                  ground_T=LW_TG)
    if medium == "gas":
        return Scene(ka=LW_KA[None, :], **common)
    z0, z1, k = LW_CLOUD
    return Scene(ka=np.zeros((1, LW_NL)), lwc=np.ones((4, 4, 4), np.float32), lower=(-20000.0, -20000.0, z0),
                 upper=(20000.0, 20000.0, z1), cext=k, ssa=0.0, g=0.85, **common)


def lw_layers_expectations(medium, view, temperatures=None):
    """Expected radiance for one medium and view; `temperatures` replaces
    the layers' temperatures (to price a mistaken temperature lookup)"""
    from scenes import planck
    T = LW_T if temperatures is None else np.asarray(temperatures)
    B = planck(LW_WLEN_NM * 1e-9, T)
    edges = np.arange(LW_NL + 1) * LW_DZ
    if medium == "gas":
        pieces = [(edges[j], edges[j + 1], LW_KA[j], B[j]) for j in range(LW_NL)]
    else:
        pieces = split_by_layers(LW_CLOUD[0], LW_CLOUD[1], LW_CLOUD[2], edges, B)
    pos, d, up = LW_VIEWS[view]
    d = np.asarray(d, np.float64) / np.linalg.norm(d)
    return emission_through_slabs(pieces, pos[2], abs(d[2]), up,
                                  ground_B=float(planck(LW_WLEN_NM * 1e-9, LW_TG)))


def lw_layers_alternatives(medium, view):
    """What a layer index off by one either way, or the temperature of the
    walk's starting layer used for every emission, would give"""
    pos = LW_VIEWS[view][0]
    j0 = min(LW_NL - 1, int(pos[2] // LW_DZ))
    return {
        "layer above": lw_layers_expectations(medium, view, np.r_[LW_T[1:], LW_T[-1]]),
        "layer below": lw_layers_expectations(medium, view, np.r_[LW_T[0], LW_T[:-1]]),
        "start layer": lw_layers_expectations(medium, view, np.full(LW_NL, LW_T[j0])),
    }
