"""htgx_ctx_create refuses scenes whose coefficients are not finite or out of
range before it touches the device (the checks run on the host): NaN or
infinite extinctions would leave the tracking loops without progress."""
import ctypes as C

import numpy as np
import pytest

from htrdr_kmin_amd import _ffi as F
from scenes import Scene


def make(**kw):
    lwc = np.full((4, 4, 4), 0.5, np.float32)
    args = dict(lwc=lwc, lower=(0, 0, 1000), upper=(400, 400, 1400), nlayers=2, ks=1e-6)
    args.update(kw)
    return Scene(**args)


@pytest.mark.parametrize("kw", [
    {"lwc": np.full((4, 4, 4), np.nan, np.float32)},
    {"lwc": np.full((4, 4, 4), -1.0, np.float32)},
    {"lwc": np.full((4, 4, 4), np.inf, np.float32)},
    {"cext": np.inf}, {"cext": -0.1}, {"ssa": 1.5}, {"ssa": np.nan}, {"g": 1.0}, {"g": -1.0},
    {"ks": np.inf}, {"ka": -1.0}, {"T": np.nan},
    {"lower": (0, 0, 1400), "upper": (400, 400, 1000)},
])
def test_bad_coefficients_are_refused(kw):
    sc = make(**kw)
    h = C.c_void_p()
    assert F.load().htgx_ctx_create(sc.ptr, C.byref(h)) == F.BAD_ARG
    assert not h.value
