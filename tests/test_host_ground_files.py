"""The ground-file readers (include/htgx_files.h): htrdr-obj geometry
(doc/htrdr-obj.5), htrdr-materials lists (doc/htrdr-materials.5) and mrumtl
BRDF files, read as htrdr_geometry_create and htrdr_materials_create do
(src/core/htrdr_geometry.c:177-566, src/core/htrdr_materials.c:72-235), and
the oracle on a loaded ground against a closed form.

The htrdr-obj examples are the two of doc/htrdr-obj.5 (the sheet of paper
between air and concrete, the wooden cube against a brick); the materials
list follows the example of doc/htrdr-materials.5.
"""
import ctypes as C
import math

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings, strategies as st

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd.atmosphere import GroundFiles
from scenes import Scene, planck, sun_solid_angle, SUN_HALF_ANGLE
from test_oracle_ground import walk

PAPER_OBJ = """\
v -1.0  -1.0  0
v -1.0   1.0  0
v  1.0  -1.0  0
v  1.0   1.0  0

usemtl air:paper:concrete
f 1 2 3
f 3 2 4
"""

CUBE_OBJ = """\
v 0 0 0
v 1 0 0
v 0 1 0
v 1 1 0
v 0 0 1
v 1 0 1
v 0 1 1
v 1 1 1

usemtl wood:air
f 1 3 2
f 2 3 4
f 1 5 3
f 3 5 7
f 5 6 7
f 7 6 8
f 4 8 2
f 2 8 6

usemtl wood:brick
f 3 7 4
f 4 7 8
f 1 2 5
f 5 2 6
"""


@pytest.fixture
def files(tmp_path):
    def write(name, text):
        p = tmp_path / name
        p.write_text(text)
        return p

    write("paper.mrumtl", "# white paper\nwavelengths 2\n400 lambertian 0.6\n800 lambertian 0.8\n")
    write("concrete.mrumtl", "bands 2\n300 500 lambertian 0.2\n500 20000 specular 0.35\n")
    write("wood.mrumtl", "wavelengths 1\n550 lambertian 0.4 # one entry\n")
    write("brick.mrumtl", "\n\nbands 1\n0 1e5 lambertian 0.25\n")
    mats = write("materials", (
        "# name  brdf  temperature\n"
        f"paper    {tmp_path / 'paper.mrumtl'}    300\n"
        f"concrete {tmp_path / 'concrete.mrumtl'} 310.5\n"
        f"wood\t{tmp_path / 'wood.mrumtl'}\t290\n"
        f"brick {tmp_path / 'brick.mrumtl'} 280   # a comment\n"
        "air     none\n"))
    return tmp_path, write, mats


def arrays(g):
    d = g.desc
    v = np.ctypeslib.as_array(d.verts, shape=(d.nverts, 3))
    t = np.ctypeslib.as_array(d.tris, shape=(d.ntris, 3))
    m = np.ctypeslib.as_array(d.tri_mtl, shape=(d.ntris,))
    ifc = np.ctypeslib.as_array(d.interfaces, shape=(d.ninterfaces, 3))
    return v, t, m, ifc


def test_paper_sheet(files):
    tmp, write, mats = files
    obj = write("paper.obj", PAPER_OBJ)
    with GroundFiles(obj, mats) as g:
        d = g.desc
        assert d.type == F.GROUND_MESH
        assert g.material_names == ["paper", "concrete", "wood", "brick", "air"]
        assert g.interface_names == ["air:paper:concrete"]
        v, t, m, ifc = arrays(g)
        assert np.array_equal(v, [[-1, -1, 0], [-1, 1, 0], [1, -1, 0], [1, 1, 0]])
        assert np.array_equal(t, [[0, 1, 2], [2, 1, 3]]) and np.array_equal(m, [0, 0])
        # front air, back concrete, thin paper
        assert ifc.tolist() == [[4, 1, 0]]
        assert d.sky_mtl == 4
        first = np.ctypeslib.as_array(d.mtl_brdf_first, shape=(d.nmtls + 1,))
        assert first.tolist() == [0, 2, 4, 5, 6, 6]
        kinds = np.ctypeslib.as_array(d.mtl_brdf_list, shape=(d.nmtls,))
        assert kinds.tolist() == [F.BRDF_LIST_WAVELENGTHS, F.BRDF_LIST_BANDS, F.BRDF_LIST_WAVELENGTHS,
                                  F.BRDF_LIST_BANDS, F.BRDF_LIST_WAVELENGTHS]
        b = [d.brdfs[i] for i in range(6)]
        assert [(e.wlen[0], e.wlen[1], e.type, e.reflectivity) for e in b] == [
            (400, 400, F.BRDF_LAMBERTIAN, 0.6), (800, 800, F.BRDF_LAMBERTIAN, 0.8),
            (300, 500, F.BRDF_LAMBERTIAN, 0.2), (500, 20000, F.BRDF_SPECULAR, 0.35),
            (550, 550, F.BRDF_LAMBERTIAN, 0.4), (0, 1e5, F.BRDF_LAMBERTIAN, 0.25)]
        T = np.ctypeslib.as_array(d.mtl_temperature, shape=(d.nmtls,))
        assert T.tolist() == [300, 310.5, 290, 280, 0]
        # the thin paper reflects on both sides
        g_ = C.byref(d)
        for dz in (-1.0, 1.0):
            N, dv = np.array([0.0, 0.0, 1.0]), np.array([0.0, 0.0, dz])
            assert O.lib().oracle_hit_material(g_, 0, O.dptr(N), O.dptr(dv)) == 0


def test_wooden_cube(files):
    tmp, write, mats = files
    obj = write("cube.obj", CUBE_OBJ)
    with GroundFiles(obj, mats, sky_mtl="nothing") as g:
        v, t, m, ifc = arrays(g)
        assert len(v) == 8 and len(t) == 12
        assert m.tolist() == [0] * 8 + [1] * 4
        assert ifc.tolist() == [[2, 4, -1], [2, 3, -1]]  # wood:air, wood:brick
        assert g.desc.sky_mtl == -1                       # no material named so
        # front sides (clockwise vertices) face the cube's inside: the
        # bottom face (1 3 2) has (v2 - v0) x (v1 - v0) = +z
        v0, v1, v2 = v[t[0]]
        assert np.cross(v2 - v0, v1 - v0)[2] > 0


def test_obj_variants(files):
    """v/vt/vn face vertices, negative indices, faces outside a material
    group dropped, a group name used again, other statements ignored"""
    tmp, write, mats = files
    obj = write("var.obj", (
        "mtllib x.mtl\no ground\n"
        "v 0 0 0\nv 1 0 0 0.5 0.5 0.5\nv 1 1 0\nv 0 1 0 1.0\n"
        "vt 0 0\nvn 0 0 1\n"
        "f 1 2 3\n"                       # before any usemtl: not ground
        "usemtl air:brick\ns off\ng a\n"
        "f 1/1 2/1/1 3//1\n"
        "usemtl air:wood\n"
        "f -4 -2 -1\n"
        "usemtl air:brick\n"
        "f 2 3 4\n"))
    with GroundFiles(obj, mats) as g:
        v, t, m, ifc = arrays(g)
        assert t.tolist() == [[0, 1, 2], [0, 2, 3], [1, 2, 3]]
        assert m.tolist() == [0, 1, 0]
        assert ifc.tolist() == [[4, 3, -1], [4, 2, -1]]


BAD = [
    ("quad.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nusemtl air:wood\nf 1 2 3 4\n", F.BAD_ARG,
     "non-triangulated"),
    ("nomtl.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nf 1 2 3\n", F.BAD_ARG, "no material"),
    ("unknown.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nusemtl air:gold\nf 1 2 3\n", F.BAD_ARG,
     "`gold' is unknown"),
    ("noback.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nusemtl wood\nf 1 2 3\n", F.BAD_ARG, "back material"),
    ("index.obj", "v 0 0 0\nv 1 0 0\nusemtl air:wood\nf 1 2 3\n", F.BAD_ARG, "face vertex"),
    ("allair.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nusemtl air:air\nf 1 2 3\n", F.BAD_ARG, "BRDF"),
]


@pytest.mark.parametrize("name,text,code,msg", BAD)
def test_bad_obj(files, name, text, code, msg):
    tmp, write, mats = files
    obj = write(name, text)
    with pytest.raises(F.HtgxError) as e:
        GroundFiles(obj, mats)
    assert e.value.code == code and msg in str(e.value), str(e.value)


@pytest.mark.parametrize("mrumtl,msg", [
    ("wavelengths 2\n400 lambertian 0.5\n", "1 entries instead of the 2"),
    ("wavelengths 1\n400 lambertian 0.5\n500 lambertian 0.5\n", "more entries"),
    ("wavelengths 2\n500 lambertian 0.5\n400 lambertian 0.5\n", "increasing"),
    ("bands 2\n400 500 lambertian 0.5\n450 600 lambertian 0.5\n", "overlapping"),
    ("bands 1\n500 400 lambertian 0.5\n", "invalid band"),
    ("wavelengths 1\n400 glossy 0.5\n", "unknown BRDF"),
    ("wavelengths 1\n400 lambertian 1.5\n", "reflectivity"),
    ("spectrum 1\n400 lambertian 0.5\n", "expected `wavelengths"),
    ("# empty\n", "no `wavelengths'"),
])
def test_bad_mrumtl(files, mrumtl, msg):
    tmp, write, _ = files
    write("bad.mrumtl", mrumtl)
    mats = write("mats_bad", f"ground {tmp / 'bad.mrumtl'} 300\nair none\n")
    obj = write("q.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nusemtl air:ground\nf 1 2 3\n")
    with pytest.raises(F.HtgxError) as e:
        GroundFiles(obj, mats)
    assert e.value.code == F.BAD_ARG and msg in str(e.value), str(e.value)


@pytest.mark.parametrize("line,msg,code", [
    ("ground", "missing the MruMtl file", F.BAD_ARG),
    ("ground {wood}", "missing temperature", F.BAD_ARG),
    ("ground {wood} hot", "error parsing the temperature", F.BAD_ARG),
    ("ground /nonexistent/x.mrumtl 300", "cannot open", F.IO_ERR),
])
def test_bad_materials_list(files, line, msg, code):
    tmp, write, _ = files
    mats = write("mats_bad", line.format(wood=tmp / "wood.mrumtl") + "\nair none\n")
    obj = write("q.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nusemtl air:ground\nf 1 2 3\n")
    with pytest.raises(F.HtgxError) as e:
        GroundFiles(obj, mats)
    assert e.value.code == code and msg in str(e.value), str(e.value)


def test_missing_files(files):
    tmp, write, mats = files
    with pytest.raises(F.HtgxError) as e:
        GroundFiles(tmp / "none.obj", mats)
    assert e.value.code == F.IO_ERR
    with pytest.raises(F.HtgxError) as e:
        GroundFiles(write("p.obj", PAPER_OBJ), tmp / "none.txt")
    assert e.value.code == F.IO_ERR


def test_redefined_material_keeps_the_last(files):
    tmp, write, _ = files
    mats = write("mats2", f"wood {tmp / 'wood.mrumtl'} 290\nair none\nwood {tmp / 'brick.mrumtl'} 250\n")
    obj = write("q.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nusemtl air:wood\nf 1 2 3\n")
    with GroundFiles(obj, mats) as g:
        assert g.material_names == ["wood", "air"]
        assert g.desc.mtl_temperature[0] == 250 and g.desc.brdfs[0].reflectivity == 0.25


def test_loaded_ground_in_the_oracle(files):
    """A 200 km flat quad read from files, `air:ground' with the ground a
    wavelength list (0.2 at 500 nm, 0.6 at 600 nm) seen at 550 nm in vacuum:
    the radiance of a Lambertian plane of albedo 0.4 (its vertices turn
    counterclockwise seen from above, so the air is below: a ray from above
    meets the back side and reflects with the front material, the air, which
    has no BRDF, hence with the ground)"""
    tmp, write, _ = files
    write("g.mrumtl", "wavelengths 2\n500 lambertian 0.2\n600 lambertian 0.6\n")
    mats = write("m", f"ground {tmp / 'g.mrumtl'} 300\nair none\n")
    obj = write("flat.obj", "v -1e5 -1e5 0\nv 1e5 -1e5 0\nv 1e5 1e5 0\nv -1e5 1e5 0\n"
                            "usemtl air:ground\nf 1 2 3\nf 1 3 4\n")
    with GroundFiles(obj, mats) as g:
        sc = Scene(sun=(0.0, 50.0))
        sc.desc.ground = g.desc
        sky = O.Sky(sc.ptr)
        w = walk(sky, (10.0, 20.0, 1000.0), (0.2, 0.1, -1.0), 6000)
        sky.close()
    expect = 0.4 / math.pi * planck(550e-9, 5778.0) * sun_solid_angle() * math.sin(
        math.radians(50.0)) * (1 + math.cos(SUN_HALF_ANGLE)) / 2
    m, se = w.mean(), w.std(ddof=1) / math.sqrt(w.size)
    assert abs(m - expect) <= 4 * se and se < 0.02 * expect


LINE = st.one_of(
    st.text(alphabet=" \t#:/-.0123456789eabcdfilmnorstuvw\r", max_size=40),
    st.sampled_from(["v 0 0 0", "v 1 0 0", "v 1 1 0", "f 1 2 3", "f -1 -2 -3", "f 1/1/1 2//2 3",
                     "usemtl air:ground", "usemtl ground:air", "usemtl air:ground:air", "usemtl :",
                     "f 1 2", "f 0 1 2", "v nan 0 0", "v 1e999 0 0", "f 99999999999999999999 1 2"]))


@settings(max_examples=300, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(obj=st.tuples(st.lists(LINE, max_size=6), st.lists(LINE, max_size=6)).map(
           lambda a: a[0] + ["v 0 0 0", "v 1 0 0", "v 0 1 0", "usemtl air:ground"] + a[1] + ["f 1 3 2"]),
       mtl=st.sampled_from(["wavelengths 2\n400 lambertian 0.5\n600 specular 1\n",
                            "bands 1\n0 1e9 lambertian 0\n", "wavelengths 0\n", "bands 2\n1 2 lambertian 0.5\n",
                            "wavelengths 1\n500 lambertian nan\n", "\x00\x01garbage\n", ""]))
def test_malformed_files_fail_cleanly(files, obj, mtl):
    """Random obj lines and mrumtl texts: the reader returns a description or
    an error carrying a message, and never crashes; a description it returns
    has triangles whose vertex indices all lie in the vertex array"""
    tmp, write, _ = files
    write("f.mrumtl", mtl)
    mats = write("m_fuzz", f"ground {tmp / 'f.mrumtl'} 300\nair none\n")
    o = write("fuzz.obj", "\n".join(obj) + "\n")
    try:
        g = GroundFiles(o, mats)
    except F.HtgxError as e:
        assert e.code in (F.BAD_ARG, F.IO_ERR) and "(" in str(e)
        return
    with g:
        d = g.desc
        assert d.ntris > 0 and d.ninterfaces > 0
        t = np.ctypeslib.as_array(d.tris, shape=(d.ntris, 3))
        assert t.max() < d.nverts
