"""The k_min/k_max octree built on the device (svo_device.hip, what
htgx_ctx_create uses) against the host build (svo_build) bit for bit: node
bounds, level-map words and leaf counts, over grids whose sizes are not
powers of two, leaf blocks 1, 2 and 4, and merge thresholds from 0 (only
uniform regions merge) to 1e9 (one leaf). The host build is the one the
segment tests pin against the oracle's own octree (test_oracle_svo.py,
test_gpu_parity.py), so this carries that parity over to the device build;
those segment tests also run on contexts whose tree the device built."""
import ctypes as C

import numpy as np
import pytest

from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from scenes import Scene

pytestmark = pytest.mark.gpu


def build(desc_ptr, on_device):
    lib = F.load()
    nb, nn, nl, ms = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_double()
    assert lib.htgx_debug_octree(desc_ptr, on_device, None, None, C.byref(nb), C.byref(nn), C.byref(nl),
                                 C.byref(ms)) == F.OK
    lmap = np.zeros(nb.value, np.uint32)
    bounds = np.zeros((nn.value, 2), np.float32)
    assert lib.htgx_debug_octree(desc_ptr, on_device, lmap.ctypes.data_as(F.c_uint32_p),
                                 bounds.ctypes.data_as(F.c_float_p), C.byref(nb), C.byref(nn),
                                 C.byref(nl), C.byref(ms)) == F.OK
    return lmap, bounds, nl.value, ms.value


def same(desc_ptr):
    lh, bh, nh, _ = build(desc_ptr, 0)
    ld, bd, nd, _ = build(desc_ptr, 1)
    assert np.array_equal(lh, ld)
    assert np.array_equal(bh.view(np.uint32), bd.view(np.uint32))  # bitwise, signed zeros included
    assert nh == nd
    return nh, lh.size


@pytest.mark.parametrize("shape,block", [((6, 10, 14), 1), ((12, 20, 28), 2), ((8, 12, 20), 4),
                                         ((3, 5, 7), 1)])
@pytest.mark.parametrize("T", [0.0, 0.3, 1.0, 1e9])
def test_random_grids(shape, block, T):
    rng = np.random.default_rng(sum(shape) + block)
    lwc = np.exp(rng.normal(-1.0, 1.0, shape)).astype(np.float32)
    lwc[rng.random(shape) < 0.4] = 0.0
    lwc[: shape[0] // 2, : shape[1] // 2, : shape[2] // 2] = 0.5  # a uniform region
    sc = Scene(lwc=lwc, lower=(0.0, 0.0, 1000.0), upper=(50.0 * shape[2], 40.0 * shape[1], 1000.0 + 30.0 * shape[0]),
               cext=0.1, optical_thickness=T, leaf_block=block)
    nleaves, nblocks = same(sc.ptr)
    if T == 1e9:
        assert nleaves == 1
    assert 1 <= nleaves <= nblocks
    if T == 0.0 and block <= 2:
        assert nleaves < nblocks  # the uniform region merged


@pytest.mark.parametrize("config", [0, 1, 4])
def test_benchmark_grids(config):
    """the benchmark scenes' grids, up to 512x512x128 cells; the build times
    are printed (device: kernels only)"""
    s = A.Synth(config, seed=0)
    nleaves, nblocks = same(s.scene_ptr)
    _, _, _, mh = build(s.scene_ptr, 0)
    _, _, _, md = build(s.scene_ptr, 1)
    print(f"config {config}: {nblocks} blocks, {nleaves} leaves, host {mh:.1f} ms, device {md:.2f} ms")
    s.close()


def test_context_tree_is_the_device_tree():
    """a context's octree info (built on the device) = the host build's"""
    s = A.Synth(0, seed=9)
    lh, bh, nh, _ = build(s.scene_ptr, 0)
    with A.Context(s.scene_ptr) as ctx:
        nl, nn, nf = ctx.svo_info()
    assert nn == len(bh) and nf == nh
    s.close()


def test_no_clouds():
    lib = F.load()
    sc = Scene(nlayers=2, ka=1e-4)
    z = C.c_uint64()
    ms = C.c_double()
    assert lib.htgx_debug_octree(sc.ptr, 1, None, None, C.byref(z), C.byref(z), C.byref(z),
                                 C.byref(ms)) == F.BAD_OP
