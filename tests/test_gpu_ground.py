"""The mesh ground on the device (the GPU BVH replacing Star-3D/Embree,
src/core/htrdr_geometry.c:686-718) against the oracle's brute-force ray
casting, and both solvers on mesh scenes with closed forms.

Bars: FP64 device vs oracle single realisations >= 97% identical to 1e-9
(BVH and brute force may pick different triangles on exact distance ties);
flux-map rejected samples equal to within 0.2%; pixels within 3 combined
sigma >= 99.5%; FP32 vs FP64 z-scores as in test_gpu_wavefront.py; closed
forms as in test_gpu_closed_forms.py.
"""
import math

import numpy as np
import pytest

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from scenes import Scene, camera, planck, rectangle, sun_solid_angle, SUN_HALF_ANGLE
from test_oracle_ground import flat_quad, terrain, with_mesh
from test_gpu_parity_maps import flux_estimate, pixel_parity, realisation_parity, zstats

pytestmark = pytest.mark.gpu

SOLVERS = [F.SOLVER_PLAIN_FP64, F.SOLVER_WAVEFRONT_FP32]
DET_RTOL = {F.SOLVER_PLAIN_FP64: 1e-9, F.SOLVER_WAVEFRONT_FP32: 2e-5}


@pytest.fixture(scope="module")
def hills():
    """The small seeded cloud field over a 33x33-vertex fBm terrain of 400 m relief"""
    s = A.Synth(0, seed=7, ground_type=F.GROUND_MESH, hf_res=33, hf_octaves=5, hf_relief=400.0,
                hf_hurst=0.7)
    ctx = A.Context(s.scene_ptr)
    sky = O.Sky(s.scene_ptr)
    yield s, ctx, sky
    sky.close()
    ctx.close()
    s.close()


def hills_camera(**kw):
    args = dict(width=32, height=24, spp=16, pos=(3200.0, -1500.0, 1800.0),
                tgt=(3200.0, -1499.0, 1799.4))
    args.update(kw)
    return A.map_preset(0, **args)


def hills_flux(**kw):
    args = dict(width=24, height=24, spp=16, sensor=F.SENSOR_RECTANGLE, pos=(3200.0, 3200.0, 200.0),
                tgt=(3200.0, 3200.0, 201.0), up=(0.0, 1.0, 0.0), rect_size=(6000.0, 6000.0))
    args.update(kw)
    return A.map_preset(0, **args)


def test_mesh_realisations_camera(hills):
    _, ctx, sky = hills
    assert realisation_parity(ctx, sky, hills_camera(), n=800) >= 0.97


def test_mesh_realisations_flux(hills):
    _, ctx, sky = hills
    assert realisation_parity(ctx, sky, hills_flux(), n=800) >= 0.97


def test_mesh_flux_map_vs_oracle(hills):
    _, ctx, sky = hills
    d = hills_flux()
    p = ctx.draw_map(d)
    acc = sky.draw_map(d)
    # rejected samples: under the terrain (..._draw_map.c:75-100, :251)
    nf_gpu = d.width * d.height * d.spp - int(p["flux_n"].sum())
    nf_cpu = int(acc["nfail"].sum())
    assert nf_cpu > 0
    assert abs(nf_gpu - nf_cpu) <= 2e-3 * nf_cpu + 2
    E, SE = flux_estimate(p)
    Eo, SEo = A.accum_estimate(acc)
    m = (p["flux_n"] > 0) & (acc["n"] > 0)
    assert pixel_parity(E[m], SE[m], Eo[m], SEo[m]) >= 0.995


def test_mesh_camera_map_vs_oracle(hills):
    _, ctx, sky = hills
    d = hills_camera()
    p = ctx.draw_map(d)
    Eo, SEo = A.accum_estimate(sky.draw_map(d))
    assert pixel_parity(p["radiance_E"], p["radiance_SE"], Eo, SEo) >= 0.995


def test_mesh_fp32_vs_fp64(hills):
    _, ctx, _ = hills
    for make in (hills_camera, hills_flux):
        d = make(spp=64)
        p64 = ctx.draw_map(d)
        d.solver = F.SOLVER_WAVEFRONT_FP32
        p32 = ctx.draw_map(d)
        if d.sensor == F.SENSOR_RECTANGLE:
            # the same samples are rejected up to FP32 origins within a few mm
            # of the terrain; pixels with few accepted samples have no usable
            # standard error
            n32, n64 = p32["flux_n"].astype(np.int64), p64["flux_n"].astype(np.int64)
            assert abs(n32.sum() - n64.sum()) <= 5e-3 * n64.sum()
            keep = (n32 >= 16) & (n64 >= 16)
            a = tuple(x[keep] for x in flux_estimate(p32))
            b = tuple(x[keep] for x in flux_estimate(p64))
        else:
            a, b = (p32["radiance_E"], p32["radiance_SE"]), (p64["radiance_E"], p64["radiance_SE"])
        within, zm, zs = zstats(*a, *b)
        assert within >= 0.99 and abs(zm) < 0.1 and zs < 1.2


@pytest.mark.parametrize("solver", SOLVERS)
def test_flat_mesh_vacuum_ground(solver):
    v, t = flat_quad()
    sc = with_mesh(Scene(sun=(0.0, 60.0)), v, t, [0, 0], [0.3], [0.0])
    with A.Context(sc.ptr) as ctx:
        p = ctx.draw_map(camera(16, 12, 32, pos=(0, 0, 3000), tgt=(0, 1, 2999), fov=40.0,
                                solver=solver))
    L_sun = planck(550e-9, 5778.0)
    expect = L_sun * sun_solid_angle() * 0.3 / math.pi * math.sin(math.radians(60)) * (
        1 + math.cos(SUN_HALF_ANGLE)) / 2 / 1e9
    E, SE = p["radiance_E"], p["radiance_SE"]
    assert np.all(np.abs(E - expect) <= 5 * SE)
    assert abs(E.mean() - expect) <= 4 * math.sqrt((SE ** 2).sum()) / E.size


@pytest.mark.parametrize("solver", SOLVERS)
def test_lw_isothermal_heightfield(solver):
    """Thick isothermal gas over a reflecting heightfield at the same T: every
    realisation is B(T) (a leak through the mesh would score 0)."""
    T = 270.0
    _, v, t, _ = terrain(res=17, relief=300.0, seed=5)
    sc = Scene(bands=((9500.0, 10500.0),), nlayers=50, dz=400.0, ka=np.full((1, 50), 1e-2),
               T=np.full(50, T))
    sc = with_mesh(sc, v, t, np.zeros(len(t)), [0.6], [T])
    d = camera(16, 16, 16, pos=(3200.0, 3200.0, 450.0), tgt=(3200.3, 3200.2, 449.0), fov=60.0,
               wlen=(10000.0, 10000.0), spectral=F.SPECTRAL_LW, Tref=T, solver=solver)
    with A.Context(sc.ptr) as ctx:
        p = ctx.draw_map(d)
    assert np.allclose(p["radiance_E"], planck(10e-6, T) / 1e9, rtol=DET_RTOL[solver], atol=0)


def test_mesh_counters(hills):
    """The FP32 solver counts BVH nodes and triangles (the 32 B and 36 B
    terms of the byte model); the plane scenes count none."""
    _, ctx, _ = hills
    d = hills_flux(solver=F.SOLVER_WAVEFRONT_FP32)
    ctx.set_counting(True)
    ctx.draw_map(d)
    c = ctx.counters()
    ctx.set_counting(False)
    assert c["nbvh"] > c["nrealisations"] and c["ntri"] > 0


def test_device_bvh_walk_matches_host_and_oracle():
    """The device BVH walk gives the host walk's hits and node counts exactly, and the oracle's brute-force distances bit for bit."""
    from test_host_bvh import continuation_rays, oracle_trace, random_rays
    s, v, _, _ = terrain(res=129, relief=600.0, seed=5)
    ctx = A.Context(s.scene_ptr)
    sky = O.Sky(s.scene_ptr)
    rng = np.random.default_rng(5)
    rays = random_rays(rng, v.min(0), v.max(0), 4000)
    for any_hit in (False, True):
        t_h, p_h, n_h = A.mesh_trace(s.scene_ptr, rays, any_hit=any_hit)
        t_d, p_d, n_d = A.mesh_trace(s.scene_ptr, rays, any_hit=any_hit, ctx=ctx)
        assert np.array_equal(t_d, t_h) and np.array_equal(p_d, p_h) and np.array_equal(n_d, n_h)
    none = np.full(rays.shape[0], -1)
    t_o, p_o, e_o = oracle_trace(sky, rays, none, np.zeros_like(none))
    assert np.array_equal(A.mesh_trace(s.scene_ptr, rays, ctx=ctx)[0], t_o)
    r2, p2, e2 = continuation_rays(rng, rays, t_o, p_o, e_o)
    t2, q2, _ = A.mesh_trace(s.scene_ptr, r2, p2, e2, ctx=ctx)
    t2_o, q2_o, _ = oracle_trace(sky, r2, p2, e2)
    assert np.array_equal(t2, t2_o) and (q2 != q2_o).mean() < 1e-3
    sky.close()
    ctx.close()


@pytest.mark.parametrize("spectral", [F.SPECTRAL_SW, F.SPECTRAL_LW])
def test_deferred_mesh_casts(spectral):
    """The FP32 solver casts a downward query from above the mesh only once
    its tracking reaches the mesh's top, then resumes the query there with
    the same optical-depth state. Scattering and absorbing gas layers
    straddle the top of a 300 m terrain, so that many queries resume inside
    the medium: FP32 vs FP64 (which casts first) per pixel."""
    _, v, t, _ = terrain(res=17, relief=300.0, seed=6)
    if spectral == F.SPECTRAL_SW:
        sc = Scene(nlayers=10, dz=100.0, ka=np.full((1, 10), 3e-4), ks=np.full((1, 10), 2e-3),
                   sun=(30.0, 50.0))
        wlen, Tref = (550.0, 550.0), 0.0
    else:
        sc = Scene(bands=((9500.0, 10500.0),), nlayers=10, dz=100.0, ka=np.full((1, 10), 1e-3),
                   ks=np.full((1, 10), 1e-3), T=np.linspace(288.0, 282.0, 10))
        wlen, Tref = (10000.0, 10000.0), 285.0
    sc = with_mesh(sc, v, t, np.zeros(len(t)), [0.4], [290.0])
    out = {}
    with A.Context(sc.ptr) as ctx:
        for solver, seed in ((F.SOLVER_PLAIN_FP64, 0), (F.SOLVER_WAVEFRONT_FP32, 1)):
            d = camera(24, 24, 64, pos=(3200.0, 3200.0, 900.0), tgt=(3200.3, 3200.2, 0.0), fov=70.0,
                       wlen=wlen, spectral=spectral, Tref=Tref, solver=solver, seed=seed)
            out[solver] = ctx.draw_map(d)
    a, b = out[F.SOLVER_WAVEFRONT_FP32], out[F.SOLVER_PLAIN_FP64]
    w, zm, zs = zstats(a["radiance_E"], a["radiance_SE"], b["radiance_E"], b["radiance_SE"])
    assert w >= 0.99 and abs(zm) < 0.1 and zs < 1.2
