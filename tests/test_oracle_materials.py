"""Ground materials in the oracle: spectral BRDF lists (mrumtl files, fetched
at src/core/htrdr_materials.c:415-463), Lambertian and specular BRDFs, and
two-sided interfaces (src/core/htrdr_interface.h:45-87, htrdr-obj(5)), pinned
by

* the BRDF lists: the entry inside a band, the first/last outside the list,
  and between two entries the upper one with probability u, wlen's linear
  position between them, so that the mean reflectivity over the canonical is
  the linear interpolation (exact on a midpoint grid). mrumtl is not in the
  reference; which of the two entries r < u selects cannot be checked
  against it (parity unpinned for that convention; the mean is pinned);
* the interface rule of htrdr_interface_fetch_hit_mtl on both sides, the thin
  material, and the fallback of a side with no BRDF;
* closed forms: a mirror seen along the sun's mirror direction by an
  orthographic camera is rho L_sun exactly, and black everywhere else; a
  Lambertian plane with a wavelength-interpolated albedo; black bodies of
  two temperatures on the two sides of one quad in the longwave;
* the rectangle sensor's rejection by the sky material (..._draw_map.c:75-100).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd.atmosphere import accum_estimate
from scenes import (Scene, air, camera, lambertian, planck, rectangle, sun_solid_angle,
                    with_materials, SUN_HALF_ANGLE)
from test_oracle_ground import walk, with_mesh

L = O.lib()
LAM, SPEC = F.BRDF_LAMBERTIAN, F.BRDF_SPECULAR
WL, BANDS = F.BRDF_LIST_WAVELENGTHS, F.BRDF_LIST_BANDS
L_SUN_550 = planck(550e-9, 5778.0)


def ground_of(mtls, interfaces=None, sky_mtl=-1):
    sc = with_materials(Scene(ground=("plane", 0.0)), mtls, interfaces, sky_mtl)
    return sc, C.byref(sc.desc.ground)


def fetch(g, m, wlen, r):
    refl = C.c_double()
    t = L.oracle_brdf_fetch(g, m, wlen, r, C.byref(refl))
    return t, refl.value


WLIST = (WL, [(400.0, 400.0, LAM, 0.1), (500.0, 500.0, SPEC, 0.5), (600.0, 600.0, LAM, 0.9)], 0.0)
BLIST = (BANDS, [(400.0, 450.0, LAM, 0.2), (500.0, 550.0, SPEC, 0.6), (550.0, 600.0, LAM, 0.4)], 0.0)


def test_wavelength_list_selection():
    sc, g = ground_of([WLIST])
    for r in (0.0, 0.3, 0.999):
        assert fetch(g, 0, 350.0, r) == (LAM, 0.1)   # before the list
        assert fetch(g, 0, 400.0, r) == (LAM, 0.1)
        assert fetch(g, 0, 500.0, r) == (SPEC, 0.5)  # on an entry: u = 1
        assert fetch(g, 0, 700.0, r) == (LAM, 0.9)   # after the list
    # 450 nm: halfway between 400 and 500
    assert fetch(g, 0, 450.0, 0.49) == (SPEC, 0.5)
    assert fetch(g, 0, 450.0, 0.51) == (LAM, 0.1)


@pytest.mark.parametrize("mtl,wlen,expect", [
    (WLIST, 575.0, 0.25 * 0.5 + 0.75 * 0.9),
    (WLIST, 420.0, 0.8 * 0.1 + 0.2 * 0.5),
    (BLIST, 475.0, 0.5 * 0.2 + 0.5 * 0.6),
    (BLIST, 460.0, 0.8 * 0.2 + 0.2 * 0.6),
])
def test_mean_reflectivity_is_the_linear_interpolation(mtl, wlen, expect):
    sc, g = ground_of([mtl])
    n = 1000
    r = (np.arange(n) + 0.5) / n
    mean = np.mean([fetch(g, 0, wlen, x)[1] for x in r])
    assert abs(mean - expect) < 1e-12


def test_band_list_selection():
    sc, g = ground_of([BLIST])
    for r in (0.0, 0.5, 0.999):
        assert fetch(g, 0, 380.0, r) == (LAM, 0.2)
        assert fetch(g, 0, 420.0, r) == (LAM, 0.2)
        assert fetch(g, 0, 450.0, r) == (LAM, 0.2)
        assert fetch(g, 0, 525.0, r) == (SPEC, 0.6)
        assert fetch(g, 0, 550.0, r) == (SPEC, 0.6)  # touching bands: the first
        assert fetch(g, 0, 551.0, r) == (LAM, 0.4)
        assert fetch(g, 0, 900.0, r) == (LAM, 0.4)


def test_grey_albedo_is_one_lambertian_entry():
    sc = Scene(ground=("plane", 0.0), albedo=0.37)
    g = C.byref(sc.desc.ground)
    for wlen in (300.0, 550.0, 20000.0):
        assert fetch(g, 0, wlen, 0.7) == (LAM, 0.37)


def hit_mtl(g, ifc, N, d):
    N = np.asarray(N, np.float64)
    d = np.asarray(d, np.float64)
    return L.oracle_hit_material(g, ifc, O.dptr(N), O.dptr(d))


def test_interface_sides():
    """N = (v1 - v0) x (v2 - v0) = +z: the vertices turn counterclockwise
    seen from above, so the front side (clockwise) is below. A ray from
    above meets the back side and reflects with the front material; from
    below, with the back material."""
    mtls = [lambertian(0.1), lambertian(0.2), lambertian(0.3), air()]
    sc, g = ground_of(mtls, [(0, 1, -1), (0, 1, 2), (0, 3, -1), (3, 1, -1), (0, 1, 3)])
    up, down = (0.0, 0.0, 1.0), (0.0, 0.0, -1.0)
    N = (0.0, 0.0, 1.0)
    assert hit_mtl(g, 0, N, down) == 0 and hit_mtl(g, 0, N, up) == 1
    assert hit_mtl(g, 1, N, down) == 2 and hit_mtl(g, 1, N, up) == 2    # thin
    assert hit_mtl(g, 2, N, down) == 0 and hit_mtl(g, 2, N, up) == 0    # back: air
    assert hit_mtl(g, 3, N, down) == 1 and hit_mtl(g, 3, N, up) == 1    # front: air
    assert hit_mtl(g, 4, N, down) == 0 and hit_mtl(g, 4, N, up) == 1    # thin: air, ignored
    # the normal's orientation decides, not the world's up
    assert hit_mtl(g, 0, (0.0, 0.0, -1.0), down) == 1
    # without interfaces, material m is {front m, back the sky}
    sc2 = Scene(ground=("plane", 0.0), albedo=0.3)
    g2 = C.byref(sc2.desc.ground)
    assert hit_mtl(g2, 0, N, down) == 0 and hit_mtl(g2, 0, N, up) == 0


def mirror_scene(el, refl=0.8, kind=SPEC):
    sc = Scene(ground=("plane", 0.0), sun=(0.0, el))
    return with_materials(sc, [(WL, [(550.0, 550.0, kind, refl)], 0.0)])


def ortho_down(d, spp, size=8):
    """An orthographic camera at 1 km looking along d (downwards)"""
    pos = (0.0, 0.0, 1000.0)
    up = (0.0, 1.0, 0.0) if abs(d[2]) > 0.999 else (0.0, 0.0, 1.0)
    cam = camera(size, size, spp, pos, tuple(p + 100.0 * x for p, x in zip(pos, d)), up=up)
    cam.projection = F.PROJECTION_ORTHOGRAPHIC
    cam.ortho_height = 200.0
    return cam


@pytest.mark.parametrize("el", [30.0, 60.0, 90.0])
def test_mirror_along_the_sun_mirror_direction(el):
    """An orthographic camera looking along the sun's mirror direction: every
    path reflects into the centre of the sun disk, w = rho L_sun exactly"""
    e = math.radians(el)
    d = (math.cos(e), 0.0, -math.sin(e)) if el < 90 else (0.0, 0.0, -1.0)
    sc = mirror_scene(el)
    sky = O.Sky(sc.ptr)
    acc = sky.draw_map(ortho_down(d, 16))
    E, SE = accum_estimate(acc)
    expect = 0.8 * L_SUN_550 / 1e9
    np.testing.assert_allclose(E, expect, rtol=1e-12)
    assert np.all(SE <= 1e-12 * expect)
    sky.close()


def test_mirror_elsewhere_is_black_and_lambertian_is_not():
    """Straight down with the sun at 60 degrees: the mirror reflects the
    zenith, black; a Lambertian reflector of the same albedo is lit"""
    sc = mirror_scene(60.0)
    sky = O.Sky(sc.ptr)
    E, SE = accum_estimate(sky.draw_map(ortho_down((0.0, 0.0, -1.0), 16)))
    assert np.all(E == 0)
    sky.close()
    sc = mirror_scene(60.0, kind=LAM)
    sky = O.Sky(sc.ptr)
    E, SE = accum_estimate(sky.draw_map(ortho_down((0.0, 0.0, -1.0), 256)))
    expect = 0.8 / math.pi * L_SUN_550 * sun_solid_angle() * math.sin(math.radians(60.0)) * (
        1 + math.cos(SUN_HALF_ANGLE)) / 2 / 1e9
    assert np.all(np.abs(E - expect) <= 4 * SE + 1e-9 * expect)
    sky.close()


def test_interpolated_albedo_closed_form():
    """Lambertian entries 0.2 at 500 nm and 0.6 at 600 nm, a 550 nm image:
    the plane's radiance is that of albedo 0.4"""
    sc = Scene(ground=("plane", 0.0), sun=(0.0, 50.0))
    with_materials(sc, [(WL, [(500.0, 500.0, LAM, 0.2), (600.0, 600.0, LAM, 0.6)], 0.0)])
    sky = O.Sky(sc.ptr)
    w = walk(sky, (10.0, 20.0, 1000.0), (0.2, 0.1, -1.0), 6000)
    expect = 0.4 / math.pi * L_SUN_550 * sun_solid_angle() * math.sin(math.radians(50.0)) * (
        1 + math.cos(SUN_HALF_ANGLE)) / 2
    m, se = w.mean(), w.std(ddof=1) / math.sqrt(w.size)
    assert abs(m - expect) <= 4 * se and se < 0.02 * expect
    # two values only: 0.2 or 0.6 times the same direct term
    sky.close()


LW_BANDS = ((9000.0, 11000.0),)


def two_sided_quad(front, back, thin=-1, T=(300.0, 250.0, 200.0)):
    """A 2 km flat quad at z = 0, counterclockwise seen from above (its front
    side is below), between black bodies at T[0] (material 0), T[1]
    (material 1), T[2] (material 2), and the air (material 3)"""
    v = [(-1000.0, -1000.0, 0.0), (1000.0, -1000.0, 0.0), (1000.0, 1000.0, 0.0), (-1000.0, 1000.0, 0.0)]
    # one transparent gas layer: a longwave map needs temperatures
    sc = Scene(bands=LW_BANDS, nlayers=1, z_bottom=-2500.0, dz=5000.0, ka=0.0, T=270.0)
    sc = with_mesh(sc, v, [(0, 1, 2), (0, 2, 3)], [0, 0], [0.0], [0.0])
    blk = [(WL, [(10000.0, 10000.0, LAM, 0.0)], t) for t in T]
    return with_materials(sc, blk + [air()], [(front, back, thin)], sky_mtl=3)


@pytest.mark.parametrize("ifc,above,below", [
    ((0, 1, -1), 0, 1), ((0, 1, 2), 2, 2), ((3, 1, -1), 1, 1), ((0, 3, -1), 0, 0)])
def test_longwave_sides_of_an_interface(ifc, above, below):
    T = (300.0, 250.0, 200.0)
    sc = two_sided_quad(*ifc, T=T)
    sky = O.Sky(sc.ptr)
    wlen = 10000.0
    for pos, d, m in (((0.0, 0.0, 500.0), (0.1, 0.0, -1.0), above),
                      ((0.0, 0.0, -500.0), (0.0, 0.1, 1.0), below)):
        w = walk(sky, pos, d, 8, wlen=wlen, spectral=F.SPECTRAL_LW)
        np.testing.assert_allclose(w, planck(wlen * 1e-9, T[m]), rtol=1e-12)
    sky.close()


@pytest.mark.parametrize("ifc,sky_mtl,rejected", [
    ((3, 0, -1), 3, False),  # the sensor below is on the front side: the air
    ((0, 3, -1), 3, True),   # below is the ground material
    ((3, 0, -1), -1, True),  # no material is the sky
])
def test_rectangle_rejection_by_the_sky_material(ifc, sky_mtl, rejected):
    sc = two_sided_quad(*ifc)
    sc.desc.ground.sky_mtl = sky_mtl
    sky = O.Sky(sc.ptr)
    d = rectangle(2, 2, 8, pos=(0.0, 0.0, -100.0), wlen=(10000.0, 10000.0), spectral=F.SPECTRAL_LW,
                  size=(50.0, 50.0))
    acc = sky.draw_map(d).ravel()
    if rejected:
        assert np.all(acc["n"] == 0) and np.all(acc["nfail"] == 8)
    else:
        assert np.all(acc["n"] == 8) and np.all(acc["nfail"] == 0)
    sky.close()


def test_general_model_of_a_grey_ground_has_the_same_mean():
    """A grey albedo given as one-entry lists and an explicit {m, air}
    interface with the air as the sky: the same radiance (the walks draw one
    more canonical per surface event, so the realisations differ)"""
    v = [(-1e5, -1e5, 0.0), (1e5, -1e5, 0.0), (1e5, 1e5, 0.0), (-1e5, 1e5, 0.0)]
    a = with_mesh(Scene(sun=(0.0, 40.0)), v, [(0, 1, 2), (0, 2, 3)], [0, 0], [0.3], [0.0])
    b = with_mesh(Scene(sun=(0.0, 40.0)), v, [(0, 1, 2), (0, 2, 3)], [0, 0], [0.3], [0.0])
    with_materials(b, [lambertian(0.3), air()], [(0, 1, -1)], sky_mtl=1)
    out = []
    for sc in (a, b):
        sky = O.Sky(sc.ptr)
        out.append(walk(sky, (10.0, 20.0, 1000.0), (0.2, 0.1, -1.0), 6000))
        sky.close()
    assert not np.array_equal(out[0], out[1])
    m = [w.mean() for w in out]
    se = [w.std(ddof=1) / math.sqrt(w.size) for w in out]
    assert abs(m[0] - m[1]) <= 4 * math.hypot(*se)
