"""CPU-side checks of the C ABI: the library loads, exports exactly what
include/gbg.h declares, and host-side argument logic refuses bad input
without touching a GPU."""

import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "gbg.h")
LIB = os.path.join(REPO, "gbgperiodicfields.jl_amd", "libgbg.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gbg_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib_path():
    if not os.path.exists(LIB):
        pytest.fail("libgbg.so is not built: run __graft_entry__.build()")
    return LIB


def test_exports_every_declared_symbol(lib_path):
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (gbg_[a-z0-9_]+)", out))
    decl = declared_symbols()
    assert len(decl) >= 30
    missing = [s for s in decl if s not in exported]
    assert not missing, missing
    # nothing extra with the public prefix escapes the header
    extra = sorted(s for s in exported if s not in decl)
    assert not extra, extra


def test_ctypes_signatures_match_header(gbg):
    from importlib import import_module  # noqa: F401
    lib = gbg._lib
    assert sorted(lib.SIGNATURES) == declared_symbols()
    L = lib.load()
    assert b"gfx950" in L.gbg_version()


def test_null_and_value_refusals_without_gpu(gbg):
    """Argument checks run before any device call."""
    import ctypes as C
    L = gbg._lib.load()
    assert L.gbg_plan_create(None, None, 8, None) == 1
    n = (C.c_int * 3)(8, 8, 8)
    Lb = (C.c_double * 3)(1.0, 1.0, 1.0)
    h = C.c_void_p()
    assert L.gbg_plan_create(n, Lb, 2, C.byref(h)) == 1          # elem_bytes
    bad = (C.c_int * 3)(0, 8, 8)
    assert L.gbg_plan_create(bad, Lb, 8, C.byref(h)) == 1        # size
    neg = (C.c_double * 3)(-1.0, 1.0, 1.0)
    assert L.gbg_plan_create(n, neg, 8, C.byref(h)) == 1         # length
    assert b"positive" in L.gbg_last_error()
    assert L.gbg_to_spectral(None, None, None, 1, None) == 1     # null plan
    assert L.gbg_uxw(None, None, None, 1.0, None, None, 0, None) == 1
    assert L.gbg_uxw_noalloc(None, None, None, 1.0, None, None, 0, None) == 1
    assert L.gbg_free(None) == 0


def test_python_mirror_refusals_without_gpu(gbg):
    with pytest.raises(gbg.ArgumentError):
        gbg.Grid3D((8, 0, 8))
    with pytest.raises(gbg.ArgumentError):
        gbg.Grid3D((8, 8, 8), (1.0, -1.0, 1.0))
    assert gbg.dealias_cutoff(512) == ((2.0 / 3.0) * 256.0) * 1.0
    assert gbg.spectral_cutoff(None, 3.0) == gbg.spectral_band(None, 0.0, 3.0)
    with pytest.raises(gbg.ArgumentError):
        gbg.spectral_band(None, 2.0, 1.0)
    with pytest.raises(gbg.ArgumentError):
        gbg.spectral_band(None, float("nan"), 1.0)
    assert np.isclose(gbg.dealias_cutoff(64), 64 / 3)


def test_python_constants_match_the_header():
    """Every #define GBG_... integer of include/gbg.h that the Python binding
    mirrors carries the header's value."""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "gbg_lib_consts", os.path.join(REPO, "gbgperiodicfields.jl_amd", "_lib.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    defs = dict(re.findall(r"#define\s+(GBG_[A-Z0-9_]+)\s+(\d+)u?\b", open(HEADER).read()))
    mirrored = [k for k in defs if hasattr(mod, k)]
    assert {"GBG_UXW_UNFUSED", "GBG_UXW_SLABS"} <= set(mirrored)
    for k in mirrored:
        assert getattr(mod, k) == int(defs[k]), k
