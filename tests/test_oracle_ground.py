"""The oracle's triangle-mesh ground (htrdr_atmosphere_ground_trace_ray,
src/atmosphere/htrdr_atmosphere_ground.c:170-217, with the self-hit filter of
src/core/htrdr_geometry.c:96-150):

* a flat two-triangle mesh reproduces the plane ground's closed form;
* the flux-map sample rejection under a heightfield matches the fraction of
  the sensor area the terrain covers (brute-force interpolation of the mesh);
* isothermal enclosures over a heightfield stay exact, so no path leaks
  through shared edges or re-hits the triangle it leaves.
"""
import math

import numpy as np
import pytest

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from scenes import Scene, planck, rectangle, sun_solid_angle, SUN_HALF_ANGLE

L_SUN_550 = planck(550e-9, 5778.0)


def with_mesh(sc, verts, tris, mtl, albedo, temperature):
    """Turn a Scene's ground into the given mesh (arrays kept alive)."""
    import ctypes as C
    verts = np.ascontiguousarray(verts, np.float32)
    tris = np.ascontiguousarray(tris, np.uint32)
    mtl = np.ascontiguousarray(mtl, np.uint32)
    alb = np.ascontiguousarray(albedo, np.float64)
    tem = np.ascontiguousarray(temperature, np.float64)
    sc._keep += [verts, tris, mtl, alb, tem]
    g = sc.desc.ground
    g.type = F.GROUND_MESH
    g.nverts, g.ntris, g.nmtls = len(verts), len(tris), len(alb)
    g.verts = verts.ctypes.data_as(F.c_float_p)
    g.tris = tris.ctypes.data_as(F.c_uint32_p)
    g.tri_mtl = mtl.ctypes.data_as(F.c_uint32_p)
    g.mtl_albedo = alb.ctypes.data_as(F.c_double_p)
    g.mtl_temperature = tem.ctypes.data_as(F.c_double_p)
    return sc


def flat_quad(half=1e5, z=0.0):
    v = [(-half, -half, z), (half, -half, z), (half, half, z), (-half, half, z)]
    return v, [(0, 1, 2), (0, 2, 3)]


def walk(sky, pos, d, n, wlen=550.0, spectral=F.SPECTRAL_SW):
    pos = np.asarray(pos, np.float64)
    d = np.asarray(d, np.float64) / np.linalg.norm(d)
    return np.array([O.lib().oracle_radiance(sky.h, 0, 0, r, spectral, F.RADIANCE_ALL, O.dptr(pos),
                                             O.dptr(d), wlen, 0, 0) for r in range(n)])


def test_flat_mesh_matches_plane_closed_form():
    v, t = flat_quad()
    sc = with_mesh(Scene(sun=(0.0, 60.0)), v, t, [0, 0], [0.3], [0.0])
    sky = O.Sky(sc.ptr)
    w = walk(sky, (10.0, 20.0, 1000.0), (0.3, 0.2, -1.0), 4000)
    expect = L_SUN_550 * sun_solid_angle() * 0.3 / math.pi * math.sin(math.radians(60)) * (
        1 + math.cos(SUN_HALF_ANGLE)) / 2
    m, se = w.mean(), w.std(ddof=1) / math.sqrt(w.size)
    assert abs(m - expect) <= 4 * se and se < 1e-4 * expect


def terrain(res=17, relief=300.0, seed=2):
    # the heightfield spans the cloud grid's XY extent, 6400 m
    s = A.Synth(0, seed=seed, ground_type=F.GROUND_MESH, hf_res=res, hf_octaves=4,
                hf_relief=relief, hf_hurst=0.7)
    g = s.scene.ground
    v = np.ctypeslib.as_array(g.verts, shape=(g.nverts, 3)).copy()
    t = np.ctypeslib.as_array(g.tris, shape=(g.ntris, 3)).copy()
    m = np.ctypeslib.as_array(g.tri_mtl, shape=(g.ntris,)).copy()
    return s, v, t, m


def height_at(v, res, x, y):
    """Terrain height by the mesh's own triangles (quad split along v00-v11)."""
    x0, x1, y0, y1 = v[:, 0].min(), v[:, 0].max(), v[:, 1].min(), v[:, 1].max()
    fx = (x - x0) / (x1 - x0) * (res - 1)
    fy = (y - y0) / (y1 - y0) * (res - 1)
    i = np.clip(np.floor(fx).astype(int), 0, res - 2)
    j = np.clip(np.floor(fy).astype(int), 0, res - 2)
    ax, ay = fx - i, fy - j
    z = v[:, 2].reshape(res, res)
    z00, z10, z01, z11 = z[j, i], z[j, i + 1], z[j + 1, i], z[j + 1, i + 1]
    lower = ax >= ay  # triangle (v00, v10, v11)
    return np.where(lower, z00 + ax * (z10 - z00) + ay * (z11 - z10),
                    z00 + ay * (z01 - z00) + ax * (z11 - z01))


def test_flux_rejection_under_heightfield():
    """A flux rectangle cutting through the terrain: a sample is rejected iff
    the terrain is above it (..._draw_map.c:75-100); the rejected fraction is
    the covered fraction of the sensor."""
    res = 17
    s, v, t, m = terrain(res=res, relief=300.0)
    sky = O.Sky(s.scene_ptr)
    zs = 150.0
    d = rectangle(8, 8, 64, pos=(3200.0, 3200.0, zs), size=(6000.0, 6000.0))
    acc = sky.draw_map(d, nthreads=4)
    rejected = acc["nfail"].sum() / (acc["nfail"].sum() + acc["n"].sum())
    rng = np.random.default_rng(0)
    xy = rng.uniform(200.0, 6200.0, (200000, 2))
    covered = (height_at(v, res, xy[:, 0], xy[:, 1]) > zs).mean()
    ntot = (acc["nfail"].sum() + acc["n"].sum())
    assert 0.05 < covered < 0.95
    assert abs(rejected - covered) <= 4 * math.sqrt(covered * (1 - covered) / ntot) + 2e-3


@pytest.mark.parametrize("albedo", [0.0, 0.6])
def test_lw_isothermal_heightfield_is_exact(albedo):
    """Optically thick isothermal gas over a heightfield at the same
    temperature: absorption, emission and reflections all end in B(T). A ray
    leaking through the mesh (into the void below) would score 0."""
    T = 270.0
    res = 17
    s, v, t, m = terrain(res=res, relief=300.0, seed=5)
    # mean free path 100 m: paths started 5-150 m above the terrain's middle
    # reach the ground or are absorbed well inside its 6.4 km footprint
    sc = Scene(bands=((9500.0, 10500.0),), nlayers=50, dz=400.0, ka=np.full((1, 50), 1e-2),
               T=np.full(50, T))
    sc = with_mesh(sc, v, t, np.zeros(len(t)), [albedo], [T])
    sky = O.Sky(sc.ptr)
    rng = np.random.default_rng(3)
    B = planck(10e-6, T)
    for k in range(40):
        x, y = rng.uniform(2000, 4400, 2)
        pos = (x, y, float(height_at(v, res, np.array([x]), np.array([y]))[0]) + rng.uniform(5, 150))
        d = (rng.normal(), rng.normal(), -abs(rng.normal()) - 0.5)
        w = walk(sky, pos, d, 8, wlen=10000.0, spectral=F.SPECTRAL_LW)
        assert np.allclose(w, B, rtol=1e-12, atol=0), (k, w)


def test_mesh_scene_validation():
    """Mesh arrays are checked: out-of-range vertex or material indices."""
    v, t = flat_quad()
    sc = with_mesh(Scene(), v, [(0, 1, 7), (0, 2, 3)], [0, 0], [0.3], [0.0])
    import ctypes as C
    h = C.c_void_p()
    lib = F.load()
    assert lib.htgx_ctx_create(sc.ptr, C.byref(h)) == F.BAD_ARG
    sc = with_mesh(Scene(), v, t, [0, 3], [0.3], [0.0])
    assert lib.htgx_ctx_create(sc.ptr, C.byref(h)) == F.BAD_ARG


def test_oracle_bvh_casts_equal_brute_force():
    """The oracle's optional BVH (for timing runs) gives the brute force's
    casts bit for bit: hit, distance, triangle and edge flag, on a terrain,
    for fresh rays and for rays leaving a previous hit, tied distances
    included (shared edges and vertices)"""
    import ctypes as C
    s, v, t, _ = terrain(res=33, relief=500.0, seed=8)
    sky = O.Sky(s.scene_ptr)
    rng = np.random.default_rng(3)
    rays, prevs = [], []
    for i in range(3000):
        o = np.array([rng.uniform(-500, 6900), rng.uniform(-500, 6900), rng.uniform(-100, 1500)])
        if i % 3 == 0:  # aim at vertices: ties between triangles sharing them
            p = v[rng.integers(len(v))].astype(np.float64)
            d = p - o
        else:
            d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        rays.append(np.concatenate([o, d, [0.0, 1e30]]))
        prevs.append((int(rng.integers(len(t))) if i % 4 == 1 else -1, i % 8 == 1))

    def casts():
        out = []
        for ray, (pp, pe) in zip(rays, prevs):
            tt, prim, edge = C.c_double(), C.c_int64(), C.c_int()
            hit = O.lib().oracle_ground_trace(sky.h, O.dptr(ray), pp, int(pe), C.byref(tt), C.byref(prim),
                                              C.byref(edge))
            out.append((hit, tt.value, prim.value, edge.value))
        return out

    brute = casts()
    assert O.lib().oracle_sky_use_bvh(sky.h, 1) == 0
    fast = casts()
    assert O.lib().oracle_sky_use_bvh(sky.h, 0) == 0
    assert sum(h for h, _, _, _ in brute) > 1000
    assert fast == brute
    sky.close()
    s.close()
