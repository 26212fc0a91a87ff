"""CIE 1931 XYZ images on the device (draw_pixel_image,
src/atmosphere/htrdr_atmosphere_draw_map.c:110-207, with the wavelength
sampler of src/core/htrdr_ran_wlen_cie_xyz.c) against the oracle and a
closed form.

Scene: the small seeded cloud field and its gas with 4 shortwave bands over 380-780 nm. Bars as
in test_gpu_parity_maps.py: FP64 device vs oracle single realisations >= 97%
identical to 1e-9; per pixel and channel >= 99.5% within 3 combined sigma;
FP32 vs FP64 >= 99% within 3 sigma, |mean z| < 0.1, z std < 1.2; the vacuum
closed form within 4 sigma (plus 2e-4 relative for FP32).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from htrdr_kmin_amd import hip as H
from scenes import Scene, camera
from test_gpu_parity_maps import pixel_parity, zstats
from test_oracle_cie import vacuum_xyz

pytestmark = pytest.mark.gpu

CH = "XYZ"


@pytest.fixture(scope="module")
def cie():
    s = A.Synth(0, seed=2, nbands=4, wlen_lo=380.0, wlen_hi=780.0)
    ctx = A.Context(s.scene_ptr)
    sky = O.Sky(s.scene_ptr)
    yield s, ctx, sky
    sky.close()
    ctx.close()
    s.close()


def image(solver=F.SOLVER_PLAIN_FP64, spp=16, size=32, seed=0):
    return A.map_preset(0, width=size, height=size, spp=spp, spectral_type=F.SPECTRAL_SW_CIE_XYZ,
                        wlen_range=(380.0, 780.0), solver=solver, seed=seed)


def test_cie_realisations(cie):
    _, ctx, sky = cie
    d = image()
    rng = np.random.default_rng(9)
    n = 900
    pix = rng.integers(0, d.width * d.height, size=n).astype(np.uint64)
    real = (rng.integers(0, 3, size=n).astype(np.uint64) << np.uint64(32)) | rng.integers(
        0, 1 << 20, size=n).astype(np.uint64)
    g = ctx.realisations(d, pix, real)
    o = np.array([sky.realisation(d, int(p), int(r)) for p, r in zip(pix, real)])
    same = np.isclose(g, o, rtol=1e-9, atol=0)
    print("identical realisations", same.mean())
    assert same.mean() >= 0.97
    lib = F.load()
    bad = np.array([3 << 32], np.uint64)
    w = np.zeros(1)
    assert lib.htgx_debug_realisations(ctx.handle, C.byref(d), pix[:1].ctypes.data_as(F.c_uint64_p),
                                       bad.ctypes.data_as(F.c_uint64_p), 1,
                                       w.ctypes.data_as(F.c_double_p)) == F.BAD_ARG


def test_cie_image_vs_oracle(cie):
    _, ctx, sky = cie
    d = image()
    p = ctx.draw_map(d)
    assert p.dtype == A.IMAGE_DTYPE and p.shape == (32, 32)
    assert np.all(p["time_n"] == 3 * d.spp)
    acc = sky.draw_map(d)
    for c in range(3):
        Eo, SEo = A.accum_estimate(acc[c])
        assert pixel_parity(p[CH[c] + "_E"], p[CH[c] + "_SE"], Eo, SEo) >= 0.995, CH[c]
        # sky pixels whose 16 paths all escape unscattered are 0
        assert np.all(p[CH[c] + "_E"] >= 0) and p[CH[c] + "_E"].mean() > 0


def test_cie_render_accum_planes(cie):
    """htgx_render_accum fills 3 planes, X, Y, Z; finalising them gives the
    pixels htgx_draw_map writes."""
    _, ctx, _ = cie
    d = image(spp=8, size=16, solver=F.SOLVER_WAVEFRONT_FP32)
    npix = d.width * d.height
    buf = H.DeviceBuffer(3 * npix * A.ACCUM_DTYPE.itemsize)
    buf.zero()
    ctx.render_accum(d, buf.address)
    H.synchronize()
    acc = buf.download(np.zeros(3 * npix, A.ACCUM_DTYPE))
    assert np.all(acc["n"] == d.spp)
    px = A.finalize(d, acc)
    ref = ctx.draw_map(d)
    for c in range(3):
        E, SE = A.accum_estimate(acc[c * npix:(c + 1) * npix])
        assert np.array_equal(px[CH[c] + "_E"].ravel(), E)
        assert np.array_equal(ref[CH[c] + "_E"], px[CH[c] + "_E"])


def test_cie_fp32_vs_fp64(cie):
    _, ctx, _ = cie
    p64 = ctx.draw_map(image(F.SOLVER_PLAIN_FP64, spp=32))
    p32 = ctx.draw_map(image(F.SOLVER_WAVEFRONT_FP32, spp=32, seed=1))
    for c in CH:
        w, zm, zs = zstats(p32[c + "_E"], p32[c + "_SE"], p64[c + "_E"], p64[c + "_SE"])
        assert w >= 0.99 and abs(zm) < 0.1 and zs < 1.2, c


@pytest.mark.parametrize("solver", [F.SOLVER_PLAIN_FP64, F.SOLVER_WAVEFRONT_FP32])
def test_cie_vacuum_closed_form(solver):
    sc = Scene(bands=((380.0, 780.0),), ground=("plane", 0.0), albedo=0.4, sun=(0.0, 50.0))
    with A.Context(sc.ptr) as ctx:
        d = camera(16, 16, 256, (10.0, 20.0, 1000.0), (13.0, 22.0, 0.0), wlen=(380.0, 780.0),
                   spectral=F.SPECTRAL_SW_CIE_XYZ, solver=solver)
        p = ctx.draw_map(d)
    expect = vacuum_xyz(0.4, 50.0)
    for c in range(3):
        E, SE = p[CH[c] + "_E"].ravel(), p[CH[c] + "_SE"].ravel()
        m, se = E.mean(), math.sqrt(np.sum(SE ** 2)) / E.size
        tol = 4 * se + (2e-4 * expect[c] if solver == F.SOLVER_WAVEFRONT_FP32 else 0.0)
        assert abs(m - expect[c]) <= tol, (CH[c], m, expect[c], se)


def test_cie_not_batched(cie):
    _, ctx, _ = cie
    lib = F.load()
    d = image()
    buf = H.DeviceBuffer(3 * d.width * d.height * A.ACCUM_DTYPE.itemsize)
    arr = (F.BatchItem * 1)()
    arr[0].ctx, arr[0].desc, arr[0].accum = ctx.handle, C.pointer(d), C.c_void_p(buf.address)
    assert lib.htgx_render_batch(arr, 1, None, None) == F.BAD_OP
