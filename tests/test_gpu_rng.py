"""Device Threefry generators against the known answers of
tests/golden/threefry_kat.json: bit-exact."""
import json
import os

import numpy as np
import pytest

from htrdr_kmin_amd import atmosphere as A

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "threefry_kat.json")


def load():
    with open(GOLDEN) as f:
        return json.load(f)


def words(vs, key):
    return np.array([[int(x, 16) for x in v[key]] for v in vs], np.uint64)


def test_device_threefry4x64_published():
    vs = load()["published"]["threefry4x64_20"]
    out = A.threefry(words(vs, "ctr"), words(vs, "key"))
    assert np.array_equal(out, words(vs, "out"))


def test_device_threefry2x32_published_and_rocrand():
    g = load()
    vs = g["published"]["threefry2x32_20"] + g["rocrand_2x32"]
    out = A.threefry2x32(words(vs, "ctr").astype(np.uint32), words(vs, "key").astype(np.uint32))
    assert np.array_equal(out, words(vs, "out").astype(np.uint32))
