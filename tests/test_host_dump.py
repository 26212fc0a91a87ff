"""Option -d (htgx_dump_octrees_vtk, the counterpart of
dump_volumetric_acceleration_structure, src/atmosphere/htrdr_atmosphere.c:
228-255): one legacy VTK tree per band and k-point, each followed by `---'
(doc/htrdr-atmosphere.1.in:222-230). Pinned by brute force on small grids:
the leaves tile the cloud grid exactly once, and each leaf's kext_min/max
is cext * the min/max LWC of the cells it covers plus the min/max gas
extinction (ka of the k-point + ks of the band) over the layers it overlaps
(0 where it reaches outside them). Host only."""
import os
import subprocess

import numpy as np
import pytest

from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from scenes import Scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def parse(text):
    trees = text.split("---\n")
    assert trees[-1] == ""
    out = []
    for t in trees[:-1]:
        lines = t.split("\n")
        assert lines[0] == "# vtk DataFile Version 2.0" and lines[2] == "ASCII"
        assert lines[3] == "DATASET UNSTRUCTURED_GRID"
        npts = int(lines[4].split()[1])
        pts = np.array([list(map(float, l.split())) for l in lines[5:5 + npts]])
        i = 5 + npts
        ncells = int(lines[i].split()[1])
        cells = np.array([list(map(int, l.split())) for l in lines[i + 1:i + 1 + ncells]])
        i += 1 + ncells
        assert lines[i] == f"CELL_TYPES {ncells}"
        assert all(l == "11" for l in lines[i + 1:i + 1 + ncells])
        i += 1 + ncells
        assert lines[i] == f"CELL_DATA {ncells}"
        assert lines[i + 1] == "SCALARS kext_min double 1" and lines[i + 2] == "LOOKUP_TABLE default"
        kmin = np.array(list(map(float, lines[i + 3:i + 3 + ncells])))
        i += 3 + ncells
        assert lines[i] == "SCALARS kext_max double 1"
        kmax = np.array(list(map(float, lines[i + 2:i + 2 + ncells])))
        assert (cells[:, 0] == 8).all()
        out.append((lines[1], pts, cells[:, 1:], kmin, kmax))
    return out


def dump(sc, tmp_path):
    p = tmp_path / "octrees.vtk"
    A.dump_octrees_vtk(sc.ptr, str(p))
    return p.read_text()


def scene(seed=3, T=0.5, bands=((500.0, 520.0), (520.0, 600.0)), quads=([0.4, 0.6], [1.0])):
    rng = np.random.default_rng(seed)
    lwc = rng.uniform(0.0, 1.0, (6, 8, 12)).astype(np.float32)
    lwc[:, :4, :6] = 0.3  # a uniform region that merges
    nq = sum(len(q) for q in quads)
    ka = rng.uniform(0, 1e-4, (nq, 5))
    ks = rng.uniform(0, 1e-4, (len(bands), 5))
    return Scene(bands=bands, quad_weights=quads, nlayers=5, z_bottom=900.0, dz=100.0, ka=ka, ks=ks,
                 lwc=lwc, lower=(0.0, 0.0, 1000.0), upper=(1200.0, 800.0, 1300.0), cext=[0.2, 0.1],
                 optical_thickness=T), lwc, ka, ks


@pytest.mark.parametrize("T", [0.0, 0.5, 1e9])
def test_leaves_tile_the_grid_with_exact_bounds(tmp_path, T):
    sc, lwc, ka, ks = scene(T=T)
    trees = parse(dump(sc, tmp_path))
    assert len(trees) == 3  # (band 0: 2 k-points) + (band 1: 1)
    cs = np.array([100.0, 100.0, 50.0])
    lower = np.array([0.0, 0.0, 1000.0])
    band_of = [0, 0, 1]
    for it, (title, pts, conn, kmin, kmax) in enumerate(trees):
        b = band_of[it]
        cover = np.zeros(lwc.shape[::-1], int)  # x, y, z
        for c in range(len(conn)):
            p = pts[conn[c]]
            lo, hi = p.min(axis=0), p.max(axis=0)
            # corner order: x fastest, then y, then z
            assert np.allclose(p[0], lo) and np.allclose(p[7], hi) and np.allclose(p[1], [hi[0], lo[1], lo[2]])
            i0 = np.rint((lo - lower) / cs).astype(int)
            i1 = np.rint((hi - lower) / cs).astype(int)
            cover[i0[0]:i1[0], i0[1]:i1[1], i0[2]:i1[2]] += 1
            cells = lwc[i0[2]:i1[2], i0[1]:i1[1], i0[0]:i1[0]]
            cext = [0.2, 0.1][b]
            layers = [j for j in range(5) if 900.0 + 100.0 * (j + 1) > lo[2] and 900.0 + 100.0 * j < hi[2]]
            g = ka[it, layers] + ks[b, layers]
            gmin, gmax = g.min(), g.max()
            if lo[2] < 900.0 or hi[2] > 1400.0:
                gmin, gmax = min(gmin, 0.0), max(gmax, 0.0)
            assert np.isclose(kmin[c], cext * float(cells.min()) + gmin, rtol=1e-7, atol=1e-12)
            assert np.isclose(kmax[c], cext * float(cells.max()) + gmax, rtol=1e-7, atol=1e-12)
        assert (cover == 1).all()
        if T == 0.0:
            assert len(conn) < lwc.size  # the uniform region merged (zero spread)
        if T == 1e9:
            assert len(conn) == 1  # one leaf: the root


def test_no_clouds_is_bad_op(tmp_path):
    sc = Scene(nlayers=2, ka=1e-4)
    with pytest.raises(F.HtgxError) as e:
        A.dump_octrees_vtk(sc.ptr, str(tmp_path / "x.vtk"))
    assert e.value.code == F.BAD_OP


def test_cli_dump(tmp_path):
    cli = os.path.join(ROOT, "htrdr_kmin_amd", "htrdr-atmosphere")
    if not os.path.exists(cli):
        pytest.skip("htrdr-atmosphere not built")
    out = tmp_path / "cli.vtk"
    r = subprocess.run([cli, "-a", "synth:0,3", "-c", "synth", "-T", "2", "-d", "-o", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    s = A.Synth(0, seed=3, optical_thickness=2.0)
    mine = tmp_path / "lib.vtk"
    A.dump_octrees_vtk(s.scene_ptr, str(mine))
    assert out.read_text() == mine.read_text()
    assert len(parse(out.read_text())) >= 1
    r = subprocess.run([cli, "-a", "synth:0", "-d"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "no clouds" in r.stderr
