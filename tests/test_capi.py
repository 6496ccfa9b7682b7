"""CPU-side checks of the C ABI: the HIP library loads, exports every entry point the
header declares, and validates arguments before touching a device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header="ewg_capi.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ewg_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_plugin_surface():
    names = _declared()
    for must in ("ewg_create", "ewg_linearize", "ewg_solve", "ewg_newton", "ewg_run",
                 "ewg_get_pattern", "ewg_set_solution", "ewg_get_solution", "ewg_destroy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import ewg
    path = ewg.library_path()
    assert os.path.exists(path), "build libewg.so first (__graft_entry__.build())"
    lib = C.CDLL(path)
    for header in ("ewg_capi.h", "ewg_cnn.h", "ewg_ensemble.h"):
        names = _declared(header)
        assert names
        missing = [n for n in names if not hasattr(lib, n)]
        assert not missing, (header, missing)


def test_default_params_match_reference_defaults():
    import ewg
    lib = ewg.load()
    p = ewg.Params()
    lib.ewg_default_params.argtypes = [C.POINTER(ewg.Params)]
    lib.ewg_default_params(C.byref(p))
    q = ewg.default_params()
    for name, _ in ewg.Params._fields_:
        assert getattr(p, name) == getattr(q, name), name


@pytest.mark.parametrize("dims", [(2, 3, 3, 3), (3, 2, 4, 4), (3, 4, 2, 4), (4, 5, 5, 5), (2, 8, 8, 2)])
def test_create_rejects_invalid_grids_without_a_device(dims):
    """Grids the structured ILU(0) does not cover, or inconsistent dims, are rejected with
    EWG_INVALID_ARGUMENT before any HIP call."""
    import ewg
    from ewg.capi import FluidDesc, GridDesc, MaterialDesc
    lib = ewg.load()
    dim, nx, ny, nz = dims
    g = GridDesc(dim, nx, ny, nz, 1.0, 1.0, 1.0)
    f = FluidDesc()
    regions = (C.c_double * 4)(1e-4, 2.0, 0.05, 0.0)
    m = MaterialDesc(0, 1, C.cast(regions, C.POINTER(C.c_double)))
    p = ewg.default_params()
    h = C.c_void_p()
    rc = lib.ewg_create(C.byref(g), C.byref(f), C.byref(m), C.byref(p), C.byref(h))
    assert rc == -1 and not h.value


def test_null_arguments_are_rejected():
    import ewg
    lib = ewg.load()
    h = C.c_void_p()
    assert lib.ewg_create(None, None, None, None, C.byref(h)) == -1
    lib.ewg_destroy(None)
    assert lib.ewg_last_error(None) is not None
