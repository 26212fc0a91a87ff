"""The oracle's Threefry-4x64-20 and random stream layout against published
known answers (tests/golden/threefry_kat.json), and the Threefry-2x32-20
fixture against rocRAND's implementation."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_ffi as O
from htrdr_kmin_amd import _ffi as F

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "threefry_kat.json")
KAT_BIN = os.path.join(O.ROOT, "oracle", "_build", "kat_rocrand")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def hexw(v):
    return [int(x, 16) for x in v]


def oracle_block(ctr, key):
    c = np.array(ctr, np.uint64)
    k = np.array(key, np.uint64)
    o = np.zeros(4, np.uint64)
    O.lib().oracle_threefry4x64_20(c.ctypes.data_as(F.c_uint64_p), k.ctypes.data_as(F.c_uint64_p),
                                   o.ctypes.data_as(F.c_uint64_p))
    return [int(x) for x in o]


def test_threefry4x64_published():
    for v in golden()["published"]["threefry4x64_20"]:
        assert oracle_block(hexw(v["ctr"]), hexw(v["key"])) == hexw(v["out"])


def test_threefry4x64_key_schedule_is_used():
    """A key word or counter word flips every output word (the zero vector
    alone cannot see a key-parity mistake)."""
    base = oracle_block([0] * 4, [0] * 4)
    for i in range(4):
        k = [0] * 4
        k[i] = 1
        c = [0] * 4
        c[i] = 1
        for out in (oracle_block([0] * 4, k), oracle_block(c, [0] * 4)):
            assert all(a != b for a, b in zip(out, base))


def test_stream_canonicals_layout():
    """Draw i of the stream keyed (seed, pixel, realisation) is word i % 4 of
    block i // 4, counter (block, 0, 0, 0), key (seed, pixel, realisation, 0),
    mapped to [0, 1) by its 53 high bits."""
    seed, pixel, real = 0x1234, 777, 42
    n = 11
    out = np.zeros(n)
    O.lib().oracle_stream_canonicals(seed, pixel, real, n, O.dptr(out))
    words = []
    for b in range(3):
        words += oracle_block([b, 0, 0, 0], [seed, pixel, real, 0])
    expect = [(w >> 11) * 2.0 ** -53 for w in words[:n]]
    assert list(out) == expect


def test_threefry2x32_fixture_consistent():
    """The published 2x32 answers are what rocRAND computes for them too (the
    fixture's rocRAND vectors come from the same program)."""
    g = golden()
    assert len(g["rocrand_2x32"]) == 64
    if not os.path.exists(KAT_BIN):
        pytest.skip("oracle/_build/kat_rocrand not built")
    lines = ["2x32 " + " ".join(v["ctr"] + v["key"]) for v in g["published"]["threefry2x32_20"]]
    lines += ["2x32 " + " ".join(v["ctr"] + v["key"]) for v in g["rocrand_2x32"]]
    res = subprocess.run([KAT_BIN], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True).stdout.split("\n")
    res = [r.split() for r in res if r.strip()]
    expect = [v["out"] for v in g["published"]["threefry2x32_20"]] + [v["out"] for v in g["rocrand_2x32"]]
    assert res == expect
