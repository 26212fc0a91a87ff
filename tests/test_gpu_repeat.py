"""Cloud repetition (option -r) on the device: the FP64 iterator's segments
bit for bit against the oracle on rays crossing several copies of the grid,
single FP64 realisations (>= 97% identical to 1e-9), the infinite absorbing
layer closed form for both solvers (4 standard errors, plus 2e-4 relative),
and FP32 vs FP64 per pixel (>= 99% within 3 sigma, |mean z| < 0.1, z std
< 1.2)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from scenes import Scene, rectangle, sun_solid_angle
from test_gpu_parity_maps import realisation_parity, zstats
from test_oracle_repeat import LOW, NX, NY, NZ, UP, L_SUN_550
from test_oracle_walks import cone_mean_cos

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rep():
    s = A.Synth(0, seed=5)
    d = F.SceneDesc.from_buffer_copy(s.scene)
    d.repeat_clouds = 1
    ptr = C.pointer(d)
    ctx = A.Context(ptr)
    sky = O.Sky(ptr)
    yield s, d, ctx, sky
    sky.close()
    ctx.close()
    s.close()


def test_repeated_segments_bitexact(rep):
    _, d, ctx, sky = rep
    rng = np.random.default_rng(8)
    n = 1500
    org = rng.uniform((-8000, -8000, 0), (14000, 14000, 6000), size=(n, 3))
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    r0 = np.where(rng.random(n) < 0.5, 0.0, rng.uniform(0, 3000, n))
    r1 = r0 + rng.uniform(1000, 30000, n)
    rays = np.concatenate([org, dirs, r0[:, None], r1[:, None]], axis=1)
    for prop in (F.Ka, F.Kext):
        g, gn = ctx.trace_segments(rays, prop, max_segs=1024)
        o, on = sky.trace_segments(rays, prop, max_segs=1024)
        assert np.array_equal(gn, on), f"segment counts differ on {np.sum(gn != on)} rays"
        assert gn.max() < 1024 and gn.sum() > 20 * n
        for i in range(n):
            a, b = g[i, :gn[i]], o[i, :gn[i]]
            for f in ("t0", "t1", "kmin", "kmax", "leaf", "layer"):
                assert np.array_equal(a[f], b[f]), (i, f)


def camera_across(solver=F.SOLVER_PLAIN_FP64, spp=16):
    # looking along the cloud layer: the rays cross many copies of the grid
    return A.map_preset(0, width=32, height=24, spp=spp, pos=(3200.0, -2000.0, 2600.0),
                        tgt=(3200.0, 0.0, 2400.0), solver=solver)


def test_repeated_realisations(rep):
    _, _, ctx, sky = rep
    assert realisation_parity(ctx, sky, camera_across(), n=600) >= 0.97


def test_repeated_fp32_vs_fp64(rep):
    _, _, ctx, _ = rep
    p64 = ctx.draw_map(camera_across(F.SOLVER_PLAIN_FP64, spp=32))
    p32 = ctx.draw_map(camera_across(F.SOLVER_WAVEFRONT_FP32, spp=32))
    w, zm, zs = zstats(p32["radiance_E"], p32["radiance_SE"], p64["radiance_E"], p64["radiance_SE"])
    assert w >= 0.99 and abs(zm) < 0.1 and zs < 1.2


@pytest.mark.parametrize("solver", [F.SOLVER_PLAIN_FP64, F.SOLVER_WAVEFRONT_FP32])
def test_infinite_absorbing_layer(solver):
    el, tau = 30.0, 0.2
    lwc = np.full((NZ, NY, NX), 1.0, np.float32)
    sc = Scene(lwc=lwc, lower=LOW, upper=UP, cext=tau / (UP[2] - LOW[2]), ssa=0.0, sun=(0.0, el),
               repeat_clouds=True)
    expect = sun_solid_angle() * L_SUN_550 * cone_mean_cos(el) / 1e9 * math.exp(
        -tau / math.sin(math.radians(el)))
    with A.Context(sc.ptr) as ctx:
        p = ctx.draw_map(rectangle(8, 8, 64, pos=(5000.0, 150.0, 0.0), size=(400.0, 300.0),
                                   solver=solver))
    E, SE = A.estimate(p["flux_sum"], p["flux_sum2"], p["flux_n"])
    rel = 2e-4 if solver == F.SOLVER_PLAIN_FP64 else 5e-4
    assert np.all(np.abs(E - expect) <= 4 * SE + rel * expect), (E.min(), E.max(), expect)


@pytest.mark.parametrize("solver", [F.SOLVER_PLAIN_FP64, F.SOLVER_WAVEFRONT_FP32])
def test_repeated_mesh_tiles_the_plane(solver):
    from scenes import camera
    from test_oracle_repeat import plane_radiance, repeated_quad_scene
    sc = repeated_quad_scene(True)
    d = camera(8, 8, 128, (5030.0, -2970.0, 800.0), (5035.0, -2968.0, 0.0), up=(0.0, 1.0, 0.0),
               solver=solver)
    with A.Context(sc.ptr) as ctx:
        p = ctx.draw_map(d)
    L = plane_radiance(60.0)
    rel = 1e-9 if solver == F.SOLVER_PLAIN_FP64 else 2e-4
    assert np.all(np.abs(p["radiance_E"] - L) <= 4 * p["radiance_SE"] + rel * L)


def test_repeated_terrain_realisations():
    """The small seeded cloud field over a repeated 17x17-vertex fBm terrain (seams
    included), looking across many copies of it"""
    from test_oracle_ground import terrain
    s, _, _, _ = terrain(res=17, relief=300.0, seed=4)
    d = F.SceneDesc.from_buffer_copy(s.scene)
    d.repeat_ground = 1
    ptr = C.pointer(d)
    with A.Context(ptr) as ctx:
        sky = O.Sky(ptr)
        desc = A.map_preset(0, width=24, height=16, spp=4, pos=(3200.0, -9000.0, 1500.0),
                            tgt=(3200.0, -8000.0, 1200.0))
        assert realisation_parity(ctx, sky, desc, n=500) >= 0.97
        sky.close()
    s.close()
