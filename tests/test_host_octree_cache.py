"""The octree cache of option -O checked on the host
(htgx_octree_cache_prepare, the checks htgx_ctx_create_cached makes before
anything reaches the device), and option -V (htgx_leaf_block_for_definition).

A cache is refused, with a message, when a byte of its body changed (its
payload digest), when it was made by hand with a valid digest but an invalid
tree (bounds that do not hold the cells, leaf words naming the wrong node,
overlapping leaves, a wrong leaf count), when its leaf block does not fit the
cloud grid, and when it is truncated. Host only."""
import os
import subprocess
import struct

import numpy as np
import pytest

from htrdr_kmin_amd import _ffi as F
from htrdr_kmin_amd import atmosphere as A
from scenes import Scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = 64
MASK64 = (1 << 64) - 1


def scene(seed=0, block=2):
    rng = np.random.default_rng(seed)
    shape = (8, 12, 16)
    lwc = np.exp(rng.normal(-1.0, 1.0, shape)).astype(np.float32)
    lwc[rng.random(shape) < 0.3] = 0.0
    return Scene(lwc=lwc, lower=(0.0, 0.0, 1000.0), upper=(1600.0, 1200.0, 1800.0), cext=0.1,
                 nlayers=2, dz=1000.0, ks=1e-6, optical_thickness=1.0, leaf_block=block)


def split(data):
    """header fields, level map and node bounds of a cache file"""
    magic, dig, pdig, block, nlev, thick, nblk, nnod, nleaf = struct.unpack("<8sQQIIdQQQ", data[:HEADER])
    lmap = np.frombuffer(data, np.uint32, nblk, HEADER).copy()
    bounds = np.frombuffer(data, np.float32, 2 * nnod, HEADER + 4 * nblk).reshape(-1, 2).copy()
    return dict(magic=magic, dig=dig, pdig=pdig, block=block, nlev=nlev, thick=thick, nblk=nblk,
                nnod=nnod, nleaf=nleaf), lmap, bounds


def payload_digest(lmap, bounds):
    """The cache's payload checksum, restated to forge caches with a valid
    one: a 64-bit multiply-xorshift over the level map then the node bounds,
    each preceded by its byte length, in little-endian 8-byte words with the
    tail zero-padded"""
    h = 0x243F6A8885A308D3

    def word(w):
        nonlocal h
        h ^= w
        h = (h * 0x9E3779B97F4A7C15) & MASK64
        h ^= h >> 29

    for data in (lmap.astype(np.uint32).tobytes(), bounds.astype(np.float32).tobytes()):
        word(len(data))
        n8 = len(data) // 8 * 8
        for (w,) in struct.iter_unpack("<Q", data[:n8]):
            word(w)
        word(int.from_bytes(data[n8:].ljust(8, b"\0"), "little"))
    return h


def forge(h, lmap, bounds):
    """a cache with these fields and arrays, its payload digest recomputed"""
    head = struct.pack("<8sQQIIdQQQ", h["magic"], h["dig"], payload_digest(lmap, bounds), h["block"],
                       h["nlev"], h["thick"], h["nblk"], h["nnod"], h["nleaf"])
    return head + lmap.astype(np.uint32).tobytes() + bounds.astype(np.float32).tobytes()


def octree_sizes(sc):
    import ctypes as C
    nb, nn, nl, ms = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_double()
    F.check("htgx_debug_octree", F.load().htgx_debug_octree(sc.ptr, 0, None, None, C.byref(nb), C.byref(nn),
                                                            C.byref(nl), C.byref(ms)))
    return nb.value, nn.value


@pytest.fixture
def cache(tmp_path):
    sc = scene()
    path = tmp_path / "octree.cache"
    assert A.octree_cache_prepare(sc.ptr, path) is False  # built and written
    assert A.octree_cache_prepare(sc.ptr, path) is True   # read back and checked
    return sc, path, path.read_bytes()


def refused(sc, path, data, words):
    path.write_bytes(data)
    with pytest.raises(F.HtgxError) as e:
        A.octree_cache_prepare(sc.ptr, path)
    assert e.value.code == F.BAD_ARG
    assert words in str(e.value), str(e.value)


def test_written_tree_is_the_host_build(cache):
    sc, path, data = cache
    h, lmap, bounds = split(data)
    assert h["magic"] == b"HTGXOCT2" and h["block"] == 2
    nblocks, nnodes = octree_sizes(sc)
    assert h["nblk"] == nblocks and h["nnod"] == nnodes
    # the payload digest restated here agrees with the library's
    assert payload_digest(lmap, bounds) == h["pdig"]
    assert forge(h, lmap, bounds) == data


@pytest.mark.parametrize("where", ["lmap", "bounds", "header"])
def test_a_flipped_byte_is_refused(cache, where):
    sc, path, data = cache
    h, lmap, bounds = split(data)
    off = {"lmap": HEADER + 5, "bounds": HEADER + 4 * h["nblk"] + 13, "header": 8 + 8 + 3}[where]
    bad = bytearray(data)
    bad[off] ^= 0x10
    refused(sc, path, bytes(bad), "corrupted")


def test_truncated_and_extended_are_refused(cache):
    sc, path, data = cache
    refused(sc, path, data[:-3], "truncated")
    refused(sc, path, data + b"\0", "trailing data")


def test_forged_invalid_trees_are_refused(cache):
    """Caches with a valid payload digest whose tree is wrong"""
    sc, path, data = cache
    h, lmap, bounds = split(data)
    leaf = int(lmap[7] & 0x0FFFFFFF)
    # bounds that no longer hold the cells (a tighter maximum): a biased majorant
    b = bounds.copy()
    if bounds[leaf, 1] > bounds[leaf, 0]:
        b[leaf, 1] = b[leaf, 0]
    else:  # a uniform leaf: bounds above its cells
        b[leaf] = (bounds[leaf, 1] + 1.0, bounds[leaf, 1] + 2.0)
    refused(sc, path, forge(h, lmap, b), "bounds do not hold")
    # a negative or NaN bound
    b = bounds.copy()
    b[leaf, 0] = -1.0
    refused(sc, path, forge(h, lmap, b), "node bound is invalid")
    b[leaf, 0] = np.nan
    refused(sc, path, forge(h, lmap, b), "node bound is invalid")
    # a leaf word naming another node
    m = lmap.copy()
    m[3] ^= 1
    refused(sc, path, forge(h, m, bounds), "does not name the node")
    # a level beyond the tree
    m = lmap.copy()
    m[3] = (np.uint32(15) << np.uint32(28)) | (m[3] & np.uint32(0x0FFFFFFF))
    refused(sc, path, forge(h, m, bounds), "level is out of range")
    # every block its own level-0 leaf while the header keeps the real count
    nb = h["nblk"]
    m = np.arange(nb, dtype=np.uint32)
    if nb != h["nleaf"]:
        refused(sc, path, forge(h, m, bounds), "leaf count")
    # the root's bounds below the grid's maximum (the FP32 records' scale)
    b = bounds.copy()
    b[-1, 1] = b[-1, 1] * 0.5
    refused(sc, path, forge(h, lmap, b), "root's bounds")


def test_overlapping_leaves_are_refused(cache):
    """one block claims the leaf of a higher level over it while another
    block under that node keeps its own leaf"""
    sc, path, data = cache
    h, lmap, bounds = split(data)
    # level-0 dims of the 16x12x8 grid in blocks of 2: 8 x 6 x 4; level 1: 4 x 3 x 2
    L1 = 8 * 6 * 4
    m = lmap.copy()
    b = bounds.copy()
    # block (0,0,0) and its level-1 node's other blocks: make (0,0,0) claim the
    # level-1 node 0 and leave (1,0,0) as its own level-0 leaf
    m[0] = (np.uint32(1) << np.uint32(28)) | np.uint32(L1)
    m[1] = np.uint32(1)
    for i in (0, 1, 8, 9, 48, 49, 56, 57):  # the 8 blocks under level-1 node 0
        if i not in (0, 1):
            m[i] = (np.uint32(1) << np.uint32(28)) | np.uint32(L1)
    b[L1] = (0.0, float(bounds[:, 1].max()))
    b[1] = (0.0, float(bounds[:, 1].max()))
    refused(sc, path, forge(h, m, b), "overlap")


def test_other_inputs_and_bad_blocks_are_refused(cache, tmp_path):
    sc, path, data = cache
    other = scene(seed=1)
    refused(other, path, data, "other atmosphere, cloud or Mie inputs")
    h, lmap, bounds = split(data)
    h3 = dict(h, block=3)  # 3 does not divide 16 x 12 x 8
    refused(sc, path, forge(h3, lmap, bounds), "does not fit the cloud grid")
    h4 = dict(h, nblk=h["nblk"] + 1)
    refused(sc, path, forge(h4, lmap, bounds), "sizes do not fit")
    junk = tmp_path / "junk"
    refused(sc, junk, b"not an octree cache" * 8, "not an octree cache")


def test_an_accepted_cache_is_left_alone(cache):
    sc, path, data = cache
    assert A.octree_cache_prepare(sc.ptr, path) is True
    assert path.read_bytes() == data


@pytest.mark.parametrize("definition,maxdef,block", [
    ((64, 64, 32), (64, 64, 32), 1),      # the clouds' definition: finest
    ((64, 64, 32), (1000, 1000, 1000), 1),
    ((64, 64, 32), (32, 32, 16), 2),
    ((64, 64, 32), (32, 64, 32), 2),       # cubic blocks: the coarsest axis decides
    ((256, 256, 64), (64, 64, 8), 8),
    ((48, 36, 12), (10, 10, 10), 6),       # 5 and 4 do not divide all three
    ((12, 12, 12), (1, 1, 1), 12),
])
def test_leaf_block_for_definition(definition, maxdef, block):
    assert A.leaf_block_for_definition(definition, maxdef) == block


@pytest.mark.parametrize("definition,maxdef", [((64, 64, 32), (0, 1, 1)), ((0, 4, 4), (4, 4, 4)),
                                              ((7, 5, 3), (2, 2, 2))])
def test_leaf_block_for_definition_refused(definition, maxdef):
    with pytest.raises(F.HtgxError):
        A.leaf_block_for_definition(definition, maxdef)


def test_cli_grid_definition(tmp_path):
    """-V x,y,z bounds the octree's finest definition: -d at -V 16,16,8 on the
    64x64x32 cloud grid writes no leaf smaller than 4x4x4 cells (400 m); at
    the clouds' own definition the finest leaves are single cells; bad
    definitions are refused as the reference refuses them"""
    cli = os.path.join(ROOT, "htrdr_kmin_amd", "htrdr-atmosphere")
    if not os.path.exists(cli):
        pytest.skip("htrdr_kmin_amd/htrdr-atmosphere not built")

    def leaf_edges(vdef):
        out = tmp_path / f"oct_{vdef}.vtk"
        r = subprocess.run([cli, "-a", "synth:0,3", "-c", "synth", "-T", "0", "-d", "-V", vdef, "-f",
                            "-o", str(out)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        tree = out.read_text().split("---\n")[0].split("\n")
        npts = int(tree[4].split()[1])
        pts = np.array([list(map(float, l.split())) for l in tree[5:5 + npts]]).reshape(-1, 8, 3)
        return (pts.max(axis=1) - pts.min(axis=1)).min(axis=0)

    assert np.allclose(leaf_edges("64,64,32"), 100.0)
    assert np.allclose(leaf_edges("16,16,8"), 400.0)
    for bad, msg in (("16,16", "Invalid grid definition"), ("0,4,4", "Invalid null grid definition"),
                     ("1,1,1", "option '-V'")):
        r = subprocess.run([cli, "-a", "synth:0,3", "-c", "synth", "-d", "-V", bad], capture_output=True,
                           text=True, timeout=60)
        assert r.returncode != 0 and msg in r.stderr, (bad, r.stderr)
