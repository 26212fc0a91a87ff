"""The C command line `htrdr-atmosphere` (htrdr_kmin_amd/cli): the
reference's options (src/atmosphere/htrdr_atmosphere_args.c:124-255) and its
htrdr-image output (doc/htrdr-image.5). CPU tests cover argument handling;
the GPU test renders through the CLI and compares its output with the
Python mirror's dump of the same map."""
import ctypes as C
import io
import os
import subprocess

import numpy as np
import pytest

from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "htrdr_kmin_amd", "htrdr-atmosphere")


def run(*args, env=None):
    if not os.path.exists(CLI):
        pytest.skip("htrdr_kmin_amd/htrdr-atmosphere not built")
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120, env=env)


def test_help():
    r = run("-h")
    assert r.returncode == 0 and "Usage: htrdr-atmosphere" in r.stdout


@pytest.mark.parametrize("args,msg", [
    ((), "option '-a'"),
    (("-a", "gas.nc"), "htgop files cannot be read"),
    (("-a", "synth:0", "-c", "clouds.nc"), "Mie's data"),
    (("-a", "synth:0", "-c", "clouds.nc", "-m", "mie.nc"), "htcp files cannot be read"),
    (("-a", "synth:0", "-D", "400,10"), "invalid option argument"),
    (("-a", "synth:0", "-D", "10,95"), "invalid option argument"),
    (("-a", "synth:0", "-i", "def=0x4"), "invalid option argument"),
    (("-a", "synth:0", "-C", "fov=180"), "invalid option argument"),
    (("-a", "synth:0", "-s", "sw=600,500"), "invalid option argument"),
    (("-a", "synth:0", "-T", "-1"), "invalid option argument"),
    (("-a", "synth:0", "-d"), "no clouds"),
    (("-a", "synth:0", "-P", "height=0"), "invalid option argument"),
    (("-a", "synth:0", "-C", "focal-length=0.05"), "focal-length is not supported"),
    (("-a", "synth:0", "-C", "lens-radius=-1"), "invalid option argument"),
    (("-a", "synth:9", "-s", "sw=550,550"), "unknown synthetic configuration"),
    (("-a", "synth:0", "-g", "ground.obj"), "option '-M'"),
    (("-a", "synth:0", "-g", "/nonexistent/ground.obj", "-M", "/nonexistent/materials"),
     "cannot open the materials file"),
])
def test_bad_arguments(args, msg):
    r = run(*args)
    assert r.returncode != 0
    assert msg in r.stderr


def test_no_overwrite_without_force(tmp_path):
    out = tmp_path / "img.txt"
    out.write_text("keep")
    r = run("-a", "synth:0", "-s", "sw=550,550", "-o", str(out))
    assert r.returncode != 0 and "use -f" in r.stderr
    assert out.read_text() == "keep"


def parse_image(text):
    lines = text.split("\n")
    w, h = map(int, lines[0].split())
    rows = []
    for y in range(h):
        base = 1 + y * (w + 1)
        rows.append([list(map(float, lines[base + x].split())) for x in range(w)])
        assert lines[base + w] == ""
    return w, h, np.array(rows)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["image", "flux", "cie", "ortho", "lens", "repeat"])
def test_cli_matches_library(tmp_path, kind):
    """Same scene and map through the CLI and through Context.draw_map +
    dump_buffer: identical columns 1-6 (the time columns are wall clocks).
    Without -s, a camera renders a CIE XYZ image over [380, 780] nm (the
    reference's default, src/core/htrdr_args.h.in:121-125), clamped to the
    sky's bands."""
    out = tmp_path / "out.txt"
    common = ["-a", "synth:0,3", "-c", "synth", "-g", "plane", "-R", "-D", "0,60", "-f",
              "-o", str(out), "-i", "def=24x16:spp=8"]
    cam = ["-C", "pos=3200,-2000,5000:tgt=3200,-1999.5,4999.133974596:up=0,0,1:fov=60"]
    if kind == "image":
        extra = ["-s", "sw=550,550"] + cam
    elif kind == "cie":
        extra = cam
    elif kind == "ortho":
        extra = ["-s", "sw=550,550", "-P",
                 "pos=3200,-2000,5000:tgt=3200,-1999.5,4999.133974596:up=0,0,1:height=3000"]
    elif kind == "repeat":
        extra = ["-r", "-s", "sw=550,550"] + cam
    elif kind == "lens":
        extra = ["-s", "sw=550,550", "-C", "pos=3200,-2000,5000:tgt=3200,-1999.5,4999.133974596:"
                 "up=0,0,1:fov=60:lens-radius=50:focal-dst=4000"]
    else:
        extra = ["-s", "sw=550,550", "-p", "pos=3200,3200,2000:tgt=3200,3200,2001:up=0,1,0:sz=6400,6400"]
    r = run(*(common + extra))
    assert r.returncode == 0, r.stderr
    assert "Time per radiative path" in r.stderr
    w, h, cli = parse_image(out.read_text())
    assert (w, h) == (24, 16)

    s = A.Synth(0, seed=3, sun_azimuth=0.0, sun_elevation=60.0)
    d = A.map_preset(0, width=24, height=16, spp=8, solver=F.SOLVER_WAVEFRONT_FP32)
    if kind == "cie":
        d.spectral_type = F.SPECTRAL_SW_CIE_XYZ
        d.wlen_range[:] = (380.0, 780.0)
    if kind == "ortho":
        d.projection = F.PROJECTION_ORTHOGRAPHIC
        d.ortho_height = 3000.0
    if kind == "lens":
        d.lens_radius, d.focal_dst = 50.0, 4000.0
    if kind in ("image", "cie", "ortho", "lens", "repeat"):
        d.pos[:] = (3200, -2000, 5000)
        d.tgt[:] = (3200, -1999.5, 4999.133974596)
        d.fov_y = 60.0
    else:
        d.sensor = F.SENSOR_RECTANGLE
        d.pos[:] = (3200, 3200, 2000)
        d.tgt[:] = (3200, 3200, 2001)
        d.up[:] = (0, 1, 0)
        d.rect_size[:] = (6400.0, 6400.0)
    scene = F.SceneDesc.from_buffer_copy(s.scene)
    scene.repeat_clouds = int(kind == "repeat")
    with A.Context(C.pointer(scene)) as ctx:
        px = ctx.draw_map(d)
    buf = io.StringIO()
    A.dump_buffer(px, d, buf)
    _, _, lib = parse_image(buf.getvalue())
    assert np.array_equal(cli[..., :6], lib[..., :6])


def write_ground(tmp_path):
    """A 9x9-vertex terrain over the cloud grid, its south half `air:grass'
    (a wavelength list), its north half `air:ice' (a band list mixing a
    specular and a Lambertian entry around 550 nm)"""
    (tmp_path / "grass.mrumtl").write_text("wavelengths 2\n500 lambertian 0.1\n600 lambertian 0.3\n")
    (tmp_path / "ice.mrumtl").write_text("bands 2\n400 540 specular 0.7\n560 800 lambertian 0.8\n")
    (tmp_path / "materials").write_text(
        f"grass {tmp_path / 'grass.mrumtl'} 290\nice {tmp_path / 'ice.mrumtl'} 260\nair none\n")
    rng = np.random.default_rng(0)
    n = 9
    lines = []
    for j in range(n):
        for i in range(n):
            lines.append(f"v {6400.0 * i / (n - 1)} {6400.0 * j / (n - 1)} {rng.uniform(0, 300):.3f}")
    for half, mtl in ((range(0, (n - 1) // 2), "air:grass"), (range((n - 1) // 2, n - 1), "air:ice")):
        lines.append(f"usemtl {mtl}")
        for j in half:
            for i in range(n - 1):
                a, b, c, d = j * n + i + 1, j * n + i + 2, (j + 1) * n + i + 2, (j + 1) * n + i + 1
                # clockwise seen from above: the front side, the air, is up
                lines += [f"f {a} {c} {b}", f"f {a} {d} {c}"]
    (tmp_path / "ground.obj").write_text("\n".join(lines) + "\n")
    return tmp_path / "ground.obj", tmp_path / "materials"


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["image", "flux"])
def test_cli_ground_files(tmp_path, kind):
    """-g ground.obj -M materials -n air: the CLI's map equals the library's
    on the same scene with the ground read by GroundFiles"""
    obj, mats = write_ground(tmp_path)
    out = tmp_path / "out.txt"
    common = ["-a", "synth:0,3", "-c", "synth", "-g", str(obj), "-M", str(mats), "-n", "air",
              "-D", "0,60", "-f", "-o", str(out), "-i", "def=24x16:spp=8", "-s", "sw=550,550"]
    if kind == "image":
        extra = ["-C", "pos=3200,-2000,2500:tgt=3200,-1999.5,2499.4:up=0,0,1:fov=60"]
    else:
        extra = ["-p", "pos=3200,3200,150:tgt=3200,3200,151:up=0,1,0:sz=6000,6000"]
    r = run(*(common + extra))
    assert r.returncode == 0, r.stderr
    _, _, cli = parse_image(out.read_text())

    s = A.Synth(0, seed=3, sun_azimuth=0.0, sun_elevation=60.0, ground_type=F.GROUND_NONE)
    d = A.map_preset(0, width=24, height=16, spp=8, solver=F.SOLVER_WAVEFRONT_FP32)
    if kind == "image":
        d.pos[:] = (3200, -2000, 2500)
        d.tgt[:] = (3200, -1999.5, 2499.4)
        d.fov_y = 60.0
    else:
        d.sensor = F.SENSOR_RECTANGLE
        d.pos[:] = (3200, 3200, 150)
        d.tgt[:] = (3200, 3200, 151)
        d.up[:] = (0, 1, 0)
        d.rect_size[:] = (6000.0, 6000.0)
    with A.GroundFiles(obj, mats, "air") as g:
        scene = g.with_scene(s.scene)
        with A.Context(C.pointer(scene)) as ctx:
            px = ctx.draw_map(d)
    buf = io.StringIO()
    A.dump_buffer(px, d, buf)
    _, _, lib = parse_image(buf.getvalue())
    assert np.array_equal(cli[..., :6], lib[..., :6])
    if kind == "flux":  # the rectangle cuts through the terrain
        assert 0 < px["flux_n"].sum() < 24 * 16 * 8


def columns(path, n=6):
    return [l.split()[:n] for l in path.read_text().splitlines() if l.strip()]


@pytest.mark.gpu
def test_cli_octree_cache(tmp_path):
    """-O: the first run writes the cache, the second reads it and prints the
    same estimates; a run with other cloud data and the same cache aborts
    (doc/htrdr-atmosphere.1.in:283-309)"""
    cache = tmp_path / "octree.cache"
    args = ["-a", "synth:0,3", "-c", "synth", "-g", "plane", "-D", "0,60", "-f", "-s", "sw=550,550",
            "-i", "def=16x8:spp=8", "-O", str(cache),
            "-C", "pos=3200,-2000,5000:tgt=3200,-1999.5,4999.133974596:up=0,0,1:fov=60"]
    r1 = run(*args, "-o", str(tmp_path / "a.txt"))
    assert r1.returncode == 0, r1.stderr
    assert cache.exists()
    data = cache.read_bytes()
    r2 = run(*args, "-T", "5", "-o", str(tmp_path / "b.txt"))
    assert r2.returncode == 0, r2.stderr
    assert columns(tmp_path / "a.txt") == columns(tmp_path / "b.txt")
    other = [a if a != "synth:0,3" else "synth:0,4" for a in args]
    r3 = run(*other, "-o", str(tmp_path / "c.txt"))
    assert r3.returncode != 0 and "other atmosphere, cloud or Mie inputs" in r3.stderr
    assert cache.read_bytes() == data
