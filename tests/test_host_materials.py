"""Ground-material descriptions the library refuses (no GPU involved: the
scene is checked before anything reaches the device): unsorted or
overlapping BRDF lists, reflectivities out of [0, 1], unknown list kinds,
interfaces naming unknown materials or reflecting with none, surfaces
naming unknown interfaces, an unknown sky material."""
import ctypes as C

import pytest

from htrdr_kmin_amd import _ffi as F
from scenes import Scene, air, lambertian, with_materials

LAM = F.BRDF_LAMBERTIAN
WL, BANDS = F.BRDF_LIST_WAVELENGTHS, F.BRDF_LIST_BANDS


def create(sc):
    lib = F.load()
    h = C.c_void_p()
    return lib.htgx_ctx_create(sc.ptr, C.byref(h))


@pytest.mark.parametrize("case", ["unsorted", "duplicate", "refl", "band_overlap", "band_reversed",
                                  "kind", "type", "ifc_range", "thin_range", "all_air", "tri_ifc",
                                  "sky"])
def test_bad_materials(case):
    mtls = [lambertian(0.3), air()]
    ifcs, sky = [(0, 1, -1)], 1
    if case == "unsorted":
        mtls[0] = (WL, [(600.0, 600.0, LAM, 0.2), (500.0, 500.0, LAM, 0.2)], 0.0)
    elif case == "duplicate":
        mtls[0] = (WL, [(500.0, 500.0, LAM, 0.2), (500.0, 500.0, LAM, 0.3)], 0.0)
    elif case == "refl":
        mtls[0] = (WL, [(550.0, 550.0, LAM, 1.2)], 0.0)
    elif case == "band_overlap":
        mtls[0] = (BANDS, [(400.0, 500.0, LAM, 0.2), (450.0, 600.0, LAM, 0.2)], 0.0)
    elif case == "band_reversed":
        mtls[0] = (BANDS, [(500.0, 400.0, LAM, 0.2)], 0.0)
    elif case == "kind":
        mtls[0] = (7, [(550.0, 550.0, LAM, 0.2)], 0.0)
    elif case == "type":
        mtls[0] = (WL, [(550.0, 550.0, 5, 0.2)], 0.0)
    elif case == "ifc_range":
        ifcs = [(0, 2, -1)]
    elif case == "thin_range":
        ifcs = [(0, 1, 9)]
    elif case == "all_air":
        ifcs = [(1, 1, -1)]
    elif case == "sky":
        sky = 5
    sc = with_materials(Scene(ground=("plane", 0.0)), mtls, ifcs, sky_mtl=sky)
    if case == "tri_ifc":
        sc.desc.ground.plane_mtl = 1  # there is one interface only
    assert create(sc) == F.BAD_ARG


def test_good_materials_pass_the_checks():
    """the same descriptions, valid, get past the checks (the call then needs
    a device: any result but BAD_ARG)"""
    import torch
    if torch.cuda.is_available():
        pytest.skip("would create a context on the device")
    sc = with_materials(Scene(ground=("plane", 0.0)),
                        [(BANDS, [(400.0, 500.0, LAM, 0.2), (500.0, 600.0, LAM, 0.3)], 0.0), air()],
                        [(0, 1, -1)], sky_mtl=1)
    assert create(sc) != F.BAD_ARG
