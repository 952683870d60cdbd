"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and rejects bad descriptors before touching the device."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qlles.h")


def header_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(qlles_[a-z_]+)\s*\(", txt)))


def test_header_matches_binding_table():
    from l3es_amd import _lib

    assert sorted(_lib.EXPORTS) == header_symbols()


def test_library_exports_every_header_symbol():
    from l3es_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libqlles.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.parametrize("field,value,code", [
    ("dim", 4, 2), ("N", 100, 2), ("N", 512, 2), ("dtype", 7, 1), ("dt", 0.0, 1), ("forcing", 9, 1),
    ("viscous", 3, 1),
])
def test_plan_create_rejects_bad_descriptors(field, value, code):
    from l3es_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libqlles.so not built")
    lib = _lib.load()
    d = _lib.Desc(3, 64, 1, 0.01, 0.01, 0, 0, 0.0, 0.0, 0, 0.0, 0.0, 0)
    setattr(d, field, value)
    h = ctypes.c_void_p()
    rc = lib.qlles_plan_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == code
    assert not h.value
    assert lib.qlles_last_error().decode()


def test_forcing_validation_without_device():
    from l3es_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libqlles.so not built")
    lib = _lib.load()
    h = ctypes.c_void_p()
    # ekin_tot needs 0 < kf < 16 and 3-D
    d = _lib.Desc(3, 64, 1, 0.01, 0.01, 1, 0, 1.0, 0.0, 0, 0.0, 0.0, 0)
    assert lib.qlles_plan_create(ctypes.byref(d), ctypes.byref(h)) == _lib.QLLES_ERR_ARG
    d = _lib.Desc(2, 64, 1, 0.01, 0.01, 1, 3, 1.0, 0.0, 0, 0.0, 0.0, 0)
    assert lib.qlles_plan_create(ctypes.byref(d), ctypes.byref(h)) == _lib.QLLES_ERR_ARG
    # Kolmogorov forcing is 2-D with 0 < k < N/2
    d = _lib.Desc(2, 64, 1, 0.01, 0.01, 3, 0, 0.0, 0.0, 40, 1.0, 0.1, 0)
    assert lib.qlles_plan_create(ctypes.byref(d), ctypes.byref(h)) == _lib.QLLES_ERR_ARG
    assert lib.qlles_interp_create(3, 33, 0, 0, 1e-6, ctypes.byref(h)) == _lib.QLLES_ERR_SHAPE
    assert lib.qlles_interp_create(3, 32, 0, 0, 0.5, ctypes.byref(h)) == _lib.QLLES_ERR_ARG
    assert lib.qlles_step(None, None, None, None, 1, None, None) == _lib.QLLES_ERR_ARG


@pytest.mark.parametrize("dim,n,box,u_ref,code", [
    (4, 32, 6.28, 1.0, 2), (3, 6, 6.28, 1.0, 2), (3, 257, 6.28, 1.0, 2), (2, 4097, 6.28, 1.0, 2),
    (3, 32, 0.0, 1.0, 1), (3, 32, -1.0, 1.0, 1), (2, 32, 6.28, float("nan"), 1),
])
def test_sph_create_rejects_bad_arguments(dim, n, box, u_ref, code):
    from l3es_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libqlles.so not built")
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.qlles_sph_create(dim, n, box, 1, u_ref, ctypes.byref(h))
    assert rc == code
    assert not h.value
    assert lib.qlles_last_error().decode()
