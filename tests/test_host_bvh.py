"""The ground-mesh BVH of the library (build, skip-link layout, the
stackless walk and the self-hit filter of src/core/htrdr_geometry.c:96-150),
walked on the host by htgx_debug_mesh_trace, against the oracle's brute force
over every triangle (htrdr_atmosphere_ground_trace_ray,
src/atmosphere/htrdr_atmosphere_ground.c:170-217).

Both cast the ray in FP32 with the same Moller-Trumbore arithmetic, so hit
distances agree bit for bit; the hit triangle agrees except at exact ties
(a ray through a shared edge), where the distances still agree.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from scenes import Scene
from test_oracle_ground import terrain, with_mesh


def oracle_trace(sky, rays, prev, prev_edge):
    n = rays.shape[0]
    t = np.zeros(n)
    prim = np.zeros(n, np.int64)
    edge = np.zeros(n, np.int32)
    for i in range(n):
        tt, pp, ee = C.c_double(), C.c_int64(), C.c_int()
        O.lib().oracle_ground_trace(sky.h, O.dptr(rays[i]), int(prev[i]), int(prev_edge[i]),
                                    C.byref(tt), C.byref(pp), C.byref(ee))
        t[i], prim[i], edge[i] = tt.value, pp.value, ee.value
    return t, prim, edge


def random_rays(rng, lo, hi, n, vertical=0.2):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    pad = 0.1 * (hi - lo)
    org = rng.uniform(lo - pad, hi + pad, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nv = int(vertical * n)  # flux-map up rays and nadir rays: two zero components
    d[:nv] = 0.0
    d[:nv, 2] = np.where(rng.random(nv) < 0.5, 1.0, -1.0)
    rays = np.zeros((n, 8))
    rays[:, :3], rays[:, 3:6] = org, d
    rays[:, 7] = np.inf
    cut = rng.random(n) < 0.2  # finite ranges, some ending before the hit
    rays[cut, 7] = rng.uniform(0, np.linalg.norm(hi - lo), cut.sum())
    return rays


def continuation_rays(rng, rays, t, prim, edge):
    """Rays leaving the surface they hit, in a new direction (as after a
    reflection), with the previous hit passed on."""
    hit = np.isfinite(t)
    r = rays[hit].copy()
    r[:, :3] = r[:, :3] + t[hit, None] * r[:, 3:6]
    d = rng.normal(size=(r.shape[0], 3))
    r[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 6], r[:, 7] = 0.0, np.inf
    return r, prim[hit], edge[hit]


def check_against_oracle(scene_ptr, sky, rays, prev, prev_edge):
    t, prim, nv = A.mesh_trace(scene_ptr, rays, prev, prev_edge)
    t_o, prim_o, edge_o = oracle_trace(sky, rays, prev, prev_edge)
    assert np.array_equal(np.isfinite(t), np.isfinite(t_o))
    assert np.array_equal(t, t_o)  # bit-exact FP32 distances
    ties = prim != prim_o
    assert ties.mean() < 1e-3, ties.sum()
    # any-hit queries: the same rays are occluded, at an accepted distance
    ta, pa, _ = A.mesh_trace(scene_ptr, rays, prev, prev_edge, any_hit=True)
    assert np.array_equal(np.isfinite(ta), np.isfinite(t_o))
    h = np.isfinite(ta)
    assert np.all(ta[h] >= t_o[h]) and np.all(ta[h] > rays[h, 6]) and np.all(ta[h] < rays[h, 7])
    assert np.all(pa[~h] == -1) and np.all(pa[h] >= 0)
    assert np.all(nv > 0)
    return t_o, prim_o, edge_o, nv


@pytest.mark.parametrize("res,relief", [(33, 300.0), (97, 800.0)])
def test_terrain_walk_matches_brute_force(res, relief):
    s, v, _, _ = terrain(res=res, relief=relief, seed=res)
    sky = O.Sky(s.scene_ptr)
    rng = np.random.default_rng(res)
    rays = random_rays(rng, v.min(0), v.max(0), 3000)
    none = np.full(rays.shape[0], -1)
    t, prim, edge, nv = check_against_oracle(s.scene_ptr, sky, rays, none, np.zeros_like(none))
    assert np.isfinite(t).mean() > 0.3
    # paths leaving the terrain: the self-hit filter sees the same triangles
    r2, p2, e2 = continuation_rays(rng, rays, t, prim, edge)
    t2, _, _, _ = check_against_oracle(s.scene_ptr, sky, r2, p2, e2)
    assert np.isfinite(t2).any()
    sky.close()


def test_triangle_soup_walk_matches_brute_force():
    """Overlapping triangles of every size and orientation: deep, unbalanced
    trees whose boxes overlap."""
    rng = np.random.default_rng(7)
    nt = 1500
    c = rng.uniform(-1000, 1000, (nt, 1, 3))
    size = np.exp(rng.uniform(np.log(5.0), np.log(400.0), (nt, 1, 1)))
    verts = (c + size * rng.normal(size=(nt, 3, 3))).reshape(-1, 3)
    tris = np.arange(3 * nt).reshape(nt, 3)
    sc = with_mesh(Scene(), verts, tris, np.zeros(nt), [0.5], [280.0])
    sky = O.Sky(sc.ptr)
    rays = random_rays(rng, verts.min(0), verts.max(0), 3000)
    none = np.full(rays.shape[0], -1)
    t, prim, edge, _ = check_against_oracle(sc.ptr, sky, rays, none, np.zeros_like(none))
    r2, p2, e2 = continuation_rays(rng, rays, t, prim, edge)
    check_against_oracle(sc.ptr, sky, r2, p2, e2)
    sky.close()


def test_walk_prunes():
    """Boxes beyond the closest hit so far are skipped: nadir rays over a
    terrain visit a small part of the tree, from above and from below."""
    s, v, _, _ = terrain(res=129, relief=500.0, seed=3)
    rng = np.random.default_rng(3)
    n = 500
    rays = np.zeros((n, 8))
    rays[:, :2] = rng.uniform(v[:, :2].min(0) + 10, v[:, :2].max(0) - 10, (n, 2))
    rays[:, 2] = v[:, 2].max() + 100.0
    rays[:, 5] = -1.0
    rays[:, 7] = np.inf
    t, prim, nv = A.mesh_trace(s.scene_ptr, rays)
    assert np.all(prim >= 0)
    ntris = s.scene.ground.ntris
    assert nv.mean() < 0.02 * (2 * ntris)  # the tree has about 2 ntris / 4 nodes
    up = rays.copy()
    up[:, 2] = v[:, 2].min() - 100.0
    up[:, 5] = 1.0
    t_up, prim_up, nv_up = A.mesh_trace(s.scene_ptr, up)
    assert np.all(prim_up >= 0) and nv_up.mean() < 0.02 * (2 * ntris)


def test_bad_arguments():
    lib = F.load()
    sc = Scene()  # plane ground
    r = np.zeros(8)
    p = np.full(1, -1, np.int64)
    e = np.zeros(1, np.int32)
    t, prim, nv = np.zeros(1), np.zeros(1, np.int64), np.zeros(1, np.uint32)
    args = (r.ctypes.data_as(F.c_double_p), p.ctypes.data_as(C.POINTER(C.c_int64)),
            e.ctypes.data_as(C.POINTER(C.c_int32)), 1, 0, t.ctypes.data_as(F.c_double_p),
            prim.ctypes.data_as(C.POINTER(C.c_int64)), nv.ctypes.data_as(F.c_uint32_p))
    assert lib.htgx_debug_mesh_trace(None, sc.ptr, *args) == F.BAD_ARG
    s, _, _, _ = terrain(res=9)
    p[0] = s.scene.ground.ntris  # out of range previous triangle
    assert lib.htgx_debug_mesh_trace(None, s.scene_ptr, *args) == F.BAD_ARG
