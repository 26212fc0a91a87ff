"""Orthographic and thin-lens cameras (the star-camera models set up at
src/atmosphere/htrdr_atmosphere.c:112-165) in the oracle, pinned by closed
forms.

Scene: vacuum over a 100 m x 60 m flat mesh quad, sun at the zenith. A camera
looking straight down sees, per pixel, the quad's share of the pixel's
footprint on the ground times the quad's radiance rho/pi L_sun Omega <cos>,
and 0 elsewhere (rays escape to the empty sky). The footprint of a pixel is
a rectangle of the ground: for the orthographic camera the pixel's share of
the image plane of height h (width h * aspect) centred under the camera; for
a pinhole (and a thin lens focused on the ground) the same with
h = 2 d tan(fov / 2), d the height of the camera. A thin lens focused
elsewhere blurs the quad's edges.
"""
import math

import numpy as np
import pytest

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from scenes import Scene, camera, planck, sun_solid_angle, SUN_HALF_ANGLE
from test_oracle_ground import with_mesh

QUAD = (0.0, 100.0, 0.0, 60.0)
ALBEDO = 0.5
CAM = (50.0, 30.0, 500.0)
W, H = 8, 4


def quad_scene():
    v = [(QUAD[0], QUAD[2], 0.0), (QUAD[1], QUAD[2], 0.0), (QUAD[1], QUAD[3], 0.0), (QUAD[0], QUAD[3], 0.0)]
    return with_mesh(Scene(sun=(0.0, 90.0)), v, [(0, 1, 2), (0, 2, 3)], [0, 0], [ALBEDO], [0.0])


def quad_radiance(wlen=550.0):
    return (ALBEDO / math.pi * planck(wlen * 1e-9, 5778.0) * sun_solid_angle()
            * (1 + math.cos(SUN_HALF_ANGLE)) / 2 * 1e-9)


def expected(plane_h):
    """Per pixel (row-major, top row first): quad coverage of the pixel's
    ground footprint times the quad radiance"""
    plane_w = plane_h * W / H
    out = np.zeros((H, W))
    for j in range(H):
        for i in range(W):
            x0 = CAM[0] - plane_w / 2 + plane_w * i / W
            x1 = x0 + plane_w / W
            y1 = CAM[1] + plane_h / 2 - plane_h * j / H  # the top row is +y
            y0 = y1 - plane_h / H
            ox = max(0.0, min(x1, QUAD[1]) - max(x0, QUAD[0]))
            oy = max(0.0, min(y1, QUAD[3]) - max(y0, QUAD[2]))
            out[j, i] = ox * oy / ((x1 - x0) * (y1 - y0))
    return out * quad_radiance()


def down_camera(spp, **kw):
    d = camera(W, H, spp, CAM, (CAM[0], CAM[1], 0.0), up=(0.0, 1.0, 0.0), fov=30.0)
    for k, v in kw.items():
        setattr(d, k, v)
    return d


def mean_se(acc):
    n = acc["n"].astype(float)
    m = (acc["sum"] + acc["sum_c"]) / n
    v = np.maximum((acc["sum2"] + acc["sum2_c"]) / n - m * m, 0.0)
    return m, np.sqrt(v / n)


def check_closed_form(acc, expect):
    m, se = mean_se(acc)
    tol = 4 * se + 1e-9 * expect.max()
    assert np.all(np.abs(m - expect) <= tol), (m, expect, se)


def test_orthographic_footprints():
    sc = quad_scene()
    sky = O.Sky(sc.ptr)
    d = down_camera(400, projection=F.PROJECTION_ORTHOGRAPHIC, ortho_height=120.0)
    acc = sky.draw_map(d)
    e = expected(120.0)
    assert (e > 0).any() and (e == 0).any() and ((e > 0) & (e < e.max())).any()
    check_closed_form(acc, e)
    sky.close()


def test_pinhole_and_focused_thin_lens_footprints():
    sc = quad_scene()
    sky = O.Sky(sc.ptr)
    e = expected(2 * CAM[2] * math.tan(math.radians(15.0)))
    check_closed_form(sky.draw_map(down_camera(400)), e)
    # a 20 m lens focused on the ground: the same footprints
    check_closed_form(sky.draw_map(down_camera(400, lens_radius=20.0, focal_dst=CAM[2])), e)
    sky.close()


def test_defocused_thin_lens_blurs_the_edges():
    sc = quad_scene()
    sky = O.Sky(sc.ptr)
    e = expected(2 * CAM[2] * math.tan(math.radians(15.0)))
    # focused at half the height: a blur disk of the lens radius on the ground
    m, se = mean_se(sky.draw_map(down_camera(1000, lens_radius=100.0, focal_dst=250.0)))
    z = np.abs(m - e) / np.maximum(se, 1e-30)
    inside = e == e.max()
    assert z.max() > 8  # some pixel far from its sharp value
    # blur moves light across edges but keeps the total (the quad is whole
    # in the view either way)
    assert abs(m.sum() - e.sum()) <= 4 * math.sqrt(np.sum(se ** 2)) + 0.05 * e.sum()
    assert np.all(m[inside] <= e.max() * (1 + 1e-9))
    sky.close()


@pytest.mark.parametrize("kw", [dict(projection=F.PROJECTION_ORTHOGRAPHIC, ortho_height=0.0),
                                dict(lens_radius=-1.0), dict(lens_radius=1.0, focal_dst=0.0),
                                dict(projection=7)])
def test_bad_cameras(kw):
    sc = quad_scene()
    sky = O.Sky(sc.ptr)
    with pytest.raises(RuntimeError):
        sky.draw_map(down_camera(1, **kw))
    sky.close()
