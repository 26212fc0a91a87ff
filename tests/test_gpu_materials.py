"""Ground materials on the device (spectral BRDF lists, Lambertian and
specular BRDFs, two-sided interfaces and the sky material of the rectangle
sensor; src/core/htrdr_materials.c:415-463, src/core/htrdr_interface.h:45-87,
..._draw_map.c:75-100) against the oracle and closed forms, both solvers.

Scene: the small seeded cloud field (64x64x32 cells) over a 33x33-vertex fBm terrain whose four height bands
get general materials: a wavelength list interpolated at 550 nm, a band list
mixing a specular and a Lambertian entry around 550 nm, a mirror, a wide
Lambertian band, and the air; one interface reversed (air in front), one
with a thin material. Bars as in test_gpu_ground.py: FP64 device vs oracle
single realisations >= 97% identical to 1e-9; pixels within 3 combined sigma
>= 99.5%; FP32 vs FP64 >= 99% within 3 sigma, |mean z| < 0.1, z std < 1.2;
closed forms exact for FP64 (1e-9) and to 2e-5 for FP32.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from scenes import Scene, air, camera, lambertian, planck, sun_solid_angle, with_materials, SUN_HALF_ANGLE
from test_gpu_ground import hills_camera, hills_flux
from test_gpu_parity_maps import flux_estimate, pixel_parity, realisation_parity, zstats
from test_oracle_materials import L_SUN_550, mirror_scene, ortho_down, two_sided_quad

pytestmark = pytest.mark.gpu

LAM, SPEC = F.BRDF_LAMBERTIAN, F.BRDF_SPECULAR
WL, BANDS = F.BRDF_LIST_WAVELENGTHS, F.BRDF_LIST_BANDS
SOLVERS = [F.SOLVER_PLAIN_FP64, F.SOLVER_WAVEFRONT_FP32]
DET_RTOL = {F.SOLVER_PLAIN_FP64: 1e-9, F.SOLVER_WAVEFRONT_FP32: 2e-5}

MTLS = [
    (WL, [(450.0, 450.0, LAM, 0.1), (650.0, 650.0, LAM, 0.5), (10000.0, 10000.0, LAM, 0.05)], 280.0),
    (BANDS, [(500.0, 540.0, SPEC, 0.6), (560.0, 600.0, LAM, 0.3), (9000.0, 11000.0, SPEC, 0.2)], 285.0),
    (WL, [(550.0, 550.0, SPEC, 0.9)], 290.0),
    (BANDS, [(300.0, 800.0, LAM, 0.45), (800.0, 20000.0, LAM, 0.1)], 300.0),
    air(),
]
IFCS = [(0, 4, -1), (4, 1, -1), (2, 4, -1), (0, 1, 3)]


class Copy:
    """A scene description copied from a Synth (whose arrays it keeps using),
    with the arrays added here kept alive"""

    def __init__(self, synth):
        self.synth = synth
        self.desc = F.SceneDesc.from_buffer_copy(synth.scene)
        self._keep = []

    @property
    def ptr(self):
        return C.pointer(self.desc)


@pytest.fixture(scope="module")
def gm():
    s = A.Synth(0, seed=7, ground_type=F.GROUND_MESH, hf_res=33, hf_octaves=5, hf_relief=400.0,
                hf_hurst=0.7)
    sc = with_materials(Copy(s), MTLS, IFCS, sky_mtl=4)
    ctx = A.Context(sc.ptr)
    sky = O.Sky(sc.ptr)
    yield sc, ctx, sky
    sky.close()
    ctx.close()
    s.close()


def test_general_realisations_camera(gm):
    _, ctx, sky = gm
    assert realisation_parity(ctx, sky, hills_camera(), n=800) >= 0.97


def test_general_realisations_flux(gm):
    """the rectangle cuts through the terrain: samples under it are rejected
    by the sky material (NaN on both sides)"""
    _, ctx, sky = gm
    d = hills_flux()
    assert realisation_parity(ctx, sky, d, n=800) >= 0.97
    rng = np.random.default_rng(1)
    pix = rng.integers(0, d.width * d.height, size=300).astype(np.uint64)
    real = rng.integers(0, 1 << 20, size=300).astype(np.uint64)
    g = ctx.realisations(d, pix, real)
    assert 0.05 < np.isnan(g).mean() < 0.95


def test_general_maps_vs_oracle(gm):
    _, ctx, sky = gm
    d = hills_camera()
    p = ctx.draw_map(d)
    Eo, SEo = A.accum_estimate(sky.draw_map(d))
    assert pixel_parity(p["radiance_E"].ravel(), p["radiance_SE"].ravel(), Eo.ravel(), SEo.ravel()) >= 0.995
    d = hills_flux()
    p = ctx.draw_map(d)
    acc = sky.draw_map(d)
    E, SE = flux_estimate(p)
    Eo, SEo = A.accum_estimate(acc)
    m = (p["flux_n"] > 0) & (acc["n"] > 0)
    assert pixel_parity(E[m], SE[m], Eo[m], SEo[m]) >= 0.995
    # rejected samples are not counted in the pixel (..._draw_map.c:251)
    nf_gpu = d.width * d.height * d.spp - int(p["flux_n"].sum())
    nf_cpu = int(acc["nfail"].sum())
    assert nf_cpu > 0 and abs(nf_gpu - nf_cpu) <= 2e-3 * nf_cpu + 2


def test_general_fp32_vs_fp64(gm):
    """FP32 vs FP64 per pixel. On the flux map the rectangle cuts through
    the terrain, so a pixel keeps only its samples above the ground: pixels
    are compared where both solvers kept at least 8 of their 32 (a mean of 1
    or 2 samples has no usable standard error: one side at n = 0 against
    n = 2 on the other gives any z)"""
    _, ctx, _ = gm
    for mk in (hills_camera, hills_flux):
        p64 = ctx.draw_map(mk(spp=32, solver=F.SOLVER_PLAIN_FP64))
        p32 = ctx.draw_map(mk(spp=32, solver=F.SOLVER_WAVEFRONT_FP32, seed=1))
        if mk is hills_camera:
            a, b = (p32["radiance_E"], p32["radiance_SE"]), (p64["radiance_E"], p64["radiance_SE"])
            keep = np.ones(a[0].shape, bool)
        else:
            a, b = flux_estimate(p32), flux_estimate(p64)
            keep = (p32["flux_n"] >= 8) & (p64["flux_n"] >= 8)
            assert keep.mean() > 0.5
        w, zm, zs = zstats(a[0][keep], a[1][keep], b[0][keep], b[1][keep])
        assert w >= 0.99 and abs(zm) < 0.1 and zs < 1.2, mk.__name__


def test_general_longwave(gm):
    """the same terrain and materials in the longwave (emission at the
    material temperatures) under absorbing gas layers"""
    sc, _, _ = gm
    lw = Copy(sc.synth)
    lw.desc.ground = sc.desc.ground
    sp = Scene(bands=((9500.0, 10500.0),), nlayers=20, dz=200.0, ka=np.full((1, 20), 2e-4),
               T=np.linspace(288.0, 262.0, 20))
    lw._keep.append(sp)
    lw.desc.spectral, lw.desc.gas = sp.desc.spectral, sp.desc.gas
    lw.desc.clouds = F.CloudDesc()
    d = camera(24, 16, 16, pos=(3200.0, -1500.0, 1800.0), tgt=(3200.0, -1499.0, 1799.4),
               wlen=(10000.0, 10000.0), spectral=F.SPECTRAL_LW, Tref=280.0)
    with A.Context(lw.ptr) as ctx:
        sky = O.Sky(lw.ptr)
        assert realisation_parity(ctx, sky, d, n=500) >= 0.97
        sky.close()
        p64 = ctx.draw_map(d)
        d.solver, d.seed = F.SOLVER_WAVEFRONT_FP32, 3
        p32 = ctx.draw_map(d)
    w, zm, zs = zstats(p32["radiance_E"], p32["radiance_SE"], p64["radiance_E"], p64["radiance_SE"])
    assert w >= 0.99 and abs(zm) < 0.1 and zs < 1.2


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("el", [30.0, 90.0])
def test_mirror_closed_form(solver, el):
    e = math.radians(el)
    dv = (math.cos(e), 0.0, -math.sin(e)) if el < 90 else (0.0, 0.0, -1.0)
    sc = mirror_scene(el)
    d = ortho_down(dv, 16)
    d.solver = solver
    with A.Context(sc.ptr) as ctx:
        p = ctx.draw_map(d)
    np.testing.assert_allclose(p["radiance_E"], 0.8 * L_SUN_550 / 1e9, rtol=DET_RTOL[solver])
    d = ortho_down((0.0, 0.0, -1.0), 16)
    d.solver = solver
    sc = mirror_scene(60.0)
    with A.Context(sc.ptr) as ctx:
        assert np.all(ctx.draw_map(d)["radiance_E"] == 0)


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("ifc,above,below", [((0, 1, -1), 0, 1), ((0, 1, 2), 2, 2), ((3, 1, -1), 1, 1)])
def test_longwave_sides_closed_form(solver, ifc, above, below):
    T = (300.0, 250.0, 200.0)
    sc = two_sided_quad(*ifc, T=T)
    with A.Context(sc.ptr) as ctx:
        for z, m in ((500.0, above), (-500.0, below)):
            d = camera(8, 8, 4, (0.0, 0.0, z), (0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=30.0,
                       wlen=(10000.0, 10000.0), spectral=F.SPECTRAL_LW, Tref=280.0, solver=solver)
            p = ctx.draw_map(d)
            np.testing.assert_allclose(p["radiance_E"], planck(1e-5, T[m]) / 1e9, rtol=DET_RTOL[solver])


@pytest.mark.parametrize("solver", SOLVERS)
def test_interpolated_albedo_closed_form(solver):
    sc = Scene(ground=("plane", 0.0), sun=(0.0, 50.0))
    with_materials(sc, [(WL, [(500.0, 500.0, LAM, 0.2), (600.0, 600.0, LAM, 0.6)], 0.0)])
    d = camera(16, 16, 256, (10.0, 20.0, 1000.0), (13.0, 22.0, 0.0), solver=solver)
    with A.Context(sc.ptr) as ctx:
        p = ctx.draw_map(d)
    expect = 0.4 / math.pi * L_SUN_550 * sun_solid_angle() * math.sin(math.radians(50.0)) * (
        1 + math.cos(SUN_HALF_ANGLE)) / 2 / 1e9
    E, SE = p["radiance_E"].ravel(), p["radiance_SE"].ravel()
    m, se = E.mean(), math.sqrt(np.sum(SE ** 2)) / E.size
    assert abs(m - expect) <= 4 * se + (2e-4 * expect if solver else 0.0)


@pytest.mark.parametrize("solver", SOLVERS)
def test_grey_ground_in_the_general_model(solver):
    """The hills' four grey materials as one-entry lists and {m, air}
    interfaces with the air as the sky: the same maps within their errors"""
    s = A.Synth(0, seed=7, ground_type=F.GROUND_MESH, hf_res=33, hf_octaves=5, hf_relief=400.0,
                hf_hurst=0.7)
    g = s.scene.ground
    alb = np.ctypeslib.as_array(g.mtl_albedo, shape=(g.nmtls,)).copy()
    sc = with_materials(Copy(s), [lambertian(a) for a in alb] + [air()],
                        [(m, len(alb), -1) for m in range(len(alb))], sky_mtl=len(alb))
    out = []
    for ptr in (s.scene_ptr, sc.ptr):
        with A.Context(ptr) as ctx:
            out.append((ctx.draw_map(hills_camera(spp=32, solver=solver)),
                        ctx.draw_map(hills_flux(spp=32, solver=solver))))
    s.close()
    a, b = out
    w, zm, zs = zstats(a[0]["radiance_E"], a[0]["radiance_SE"], b[0]["radiance_E"], b[0]["radiance_SE"])
    assert w >= 0.99 and abs(zm) < 0.1 and zs < 1.2
    # the same samples are rejected (the rejection comes before any draw
    # the two models make differently)
    assert np.array_equal(a[1]["flux_n"], b[1]["flux_n"]) and a[1]["flux_n"].min() < 32
    m = a[1]["flux_n"] > 1
    (E1, S1), (E2, S2) = flux_estimate(a[1]), flux_estimate(b[1])
    w, zm, zs = zstats(E1[m], S1[m], E2[m], S2[m])
    assert w >= 0.99 and abs(zm) < 0.1 and zs < 1.2


def test_ground_files_parity(tmp_path):
    """A terrain read from htrdr-obj/htrdr-materials/mrumtl files under the small
    seeded cloud field: device FP64 vs oracle realisations, FP32 vs FP64 pixels"""
    from test_cli import write_ground
    obj, mats = write_ground(tmp_path)
    s = A.Synth(0, seed=3, ground_type=F.GROUND_NONE)
    with A.GroundFiles(obj, mats) as g:
        scene = g.with_scene(s.scene)
        ptr = C.pointer(scene)
        d = A.map_preset(0, width=24, height=16, spp=16, pos=(3200.0, -2000.0, 2500.0),
                         tgt=(3200.0, -1999.5, 2499.4))
        with A.Context(ptr) as ctx:
            sky = O.Sky(ptr)
            assert realisation_parity(ctx, sky, d, n=600) >= 0.97
            sky.close()
            p64 = ctx.draw_map(A.map_preset(0, width=24, height=16, spp=32, pos=(3200.0, -2000.0, 2500.0),
                                            tgt=(3200.0, -1999.5, 2499.4)))
            p32 = ctx.draw_map(A.map_preset(0, width=24, height=16, spp=32, pos=(3200.0, -2000.0, 2500.0),
                                            tgt=(3200.0, -1999.5, 2499.4),
                                            solver=F.SOLVER_WAVEFRONT_FP32, seed=1))
    s.close()
    w, zm, zs = zstats(p32["radiance_E"], p32["radiance_SE"], p64["radiance_E"], p64["radiance_SE"])
    assert w >= 0.99 and abs(zm) < 0.1 and zs < 1.2


def test_general_materials_cie_image():
    """The hills' spectral BRDF lists seen in a CIE XYZ image: wavelengths
    sampled over 380-780 nm per channel select different entries (bands and
    interpolated wavelengths); FP64 device vs oracle realisations, FP32 vs
    FP64 per channel"""
    s = A.Synth(0, seed=7, ground_type=F.GROUND_MESH, hf_res=33, hf_octaves=5, hf_relief=400.0,
                hf_hurst=0.7, nbands=4, wlen_lo=380.0, wlen_hi=780.0)
    sc = with_materials(Copy(s), MTLS, IFCS, sky_mtl=4)
    d = hills_camera(spectral_type=F.SPECTRAL_SW_CIE_XYZ, wlen_range=(380.0, 780.0))
    with A.Context(sc.ptr) as ctx:
        sky = O.Sky(sc.ptr)
        rng = np.random.default_rng(5)
        n = 600
        pix = rng.integers(0, d.width * d.height, size=n).astype(np.uint64)
        real = (rng.integers(0, 3, size=n).astype(np.uint64) << np.uint64(32)) | rng.integers(
            0, 1 << 20, size=n).astype(np.uint64)
        g = ctx.realisations(d, pix, real)
        o = np.array([sky.realisation(d, int(p), int(r)) for p, r in zip(pix, real)])
        sky.close()
        same = np.isclose(g, o, rtol=1e-9, atol=0) | (np.isnan(g) & np.isnan(o))
        assert same.mean() >= 0.97
        p64 = ctx.draw_map(hills_camera(spp=32, spectral_type=F.SPECTRAL_SW_CIE_XYZ,
                                        wlen_range=(380.0, 780.0)))
        p32 = ctx.draw_map(hills_camera(spp=32, spectral_type=F.SPECTRAL_SW_CIE_XYZ,
                                        wlen_range=(380.0, 780.0), solver=F.SOLVER_WAVEFRONT_FP32,
                                        seed=1))
    s.close()
    for c in "XYZ":
        w, zm, zs = zstats(p32[c + "_E"], p32[c + "_SE"], p64[c + "_E"], p64[c + "_SE"])
        assert w >= 0.99 and abs(zm) < 0.1 and zs < 1.2, c
