"""Orthographic and thin-lens cameras on the device, both solvers, against
the closed forms of test_oracle_cameras.py (per-pixel quad coverage of the
ground footprints) and, for the FP64 solver, realisation by realisation
against the oracle (>= 97% identical to 1e-9). Tolerance of the closed forms:
4 standard errors, plus 2e-4 relative for the FP32 solver."""
import math

import numpy as np
import pytest

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from test_gpu_parity_maps import realisation_parity
from test_oracle_cameras import CAM, down_camera, expected, quad_scene

pytestmark = pytest.mark.gpu

SOLVERS = [F.SOLVER_PLAIN_FP64, F.SOLVER_WAVEFRONT_FP32]
CAMERAS = {
    "ortho": (dict(projection=F.PROJECTION_ORTHOGRAPHIC, ortho_height=120.0), 120.0),
    "pinhole": (dict(), 2 * CAM[2] * math.tan(math.radians(15.0))),
    "lens": (dict(lens_radius=20.0, focal_dst=CAM[2]), 2 * CAM[2] * math.tan(math.radians(15.0))),
}


@pytest.fixture(scope="module")
def quad():
    sc = quad_scene()
    ctx = A.Context(sc.ptr)
    sky = O.Sky(sc.ptr)
    yield sc, ctx, sky
    sky.close()
    ctx.close()


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("cam", sorted(CAMERAS))
def test_camera_footprints(quad, solver, cam):
    _, ctx, _ = quad
    kw, plane_h = CAMERAS[cam]
    p = ctx.draw_map(down_camera(512, solver=solver, **kw))
    e = expected(plane_h)
    tol = 4 * p["radiance_SE"] + (2e-4 if solver == F.SOLVER_WAVEFRONT_FP32 else 1e-9) * e.max()
    assert np.all(np.abs(p["radiance_E"] - e) <= tol), (p["radiance_E"], e)


@pytest.mark.parametrize("cam", ["ortho", "lens"])
def test_camera_realisations(quad, cam):
    _, ctx, sky = quad
    kw, _ = CAMERAS[cam]
    assert realisation_parity(ctx, sky, down_camera(8, **kw), n=600) >= 0.97


def test_defocus_differs_from_pinhole(quad):
    _, ctx, _ = quad
    e = expected(CAMERAS["pinhole"][1])
    p = ctx.draw_map(down_camera(1024, solver=F.SOLVER_WAVEFRONT_FP32, lens_radius=100.0,
                                 focal_dst=250.0))
    z = np.abs(p["radiance_E"] - e) / np.maximum(p["radiance_SE"], 1e-30)
    assert z.max() > 8
