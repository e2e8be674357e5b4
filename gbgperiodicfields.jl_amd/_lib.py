"""ctypes binding of libgbg.so (declared in include/gbg.h).

The library is loaded from this directory; a missing library raises at once
-- there is no fallback path.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgbg.so")

GBG_OK = 0
STATUS_NAMES = {0: "GBG_OK", 1: "GBG_E_ARG", 2: "GBG_E_DIM", 3: "GBG_E_UNSUPPORTED",
                4: "GBG_E_ALIAS", 5: "GBG_E_OOM", 6: "GBG_E_DEVICE"}

GBG_UXW_UNFUSED = 1
GBG_UXW_SLABS = 2

class GbgJob(C.Structure):
    """struct gbg_job of include/gbg.h."""

    _fields_ = [("kind", C.c_int), ("n", C.c_int * 3), ("L", C.c_double * 3),
                ("elem_bytes", C.c_int), ("ncomp", C.c_int), ("in0", C.c_void_p),
                ("in1", C.c_void_p), ("out", C.c_void_p), ("kc", C.c_double)]


class GbgUxwCombination(C.Structure):
    """struct gbg_uxw_combination of include/gbg.h."""

    _fields_ = [("dst", C.c_void_p), ("a", C.c_void_p), ("alpha", C.c_double), ("b", C.c_void_p),
                ("beta", C.c_double), ("delta", C.c_double), ("gamma", C.c_double),
                ("coefficient", C.c_double)]


vp = C.c_void_p
st = C.c_int
sz = C.c_size_t
dbl = C.c_double
i32 = C.c_int
dp = C.POINTER(C.c_double)

# name -> (restype, argtypes); the exact set of symbols include/gbg.h declares
SIGNATURES = {
    "gbg_last_error": (C.c_char_p, []),
    "gbg_version": (C.c_char_p, []),
    "gbg_set_device": (st, [i32]),
    "gbg_device_count": (st, [C.POINTER(C.c_int)]),
    "gbg_malloc": (st, [C.POINTER(vp), sz]),
    "gbg_free": (st, [vp]),
    "gbg_memcpy_h2d": (st, [vp, vp, sz, vp]),
    "gbg_memcpy_d2h": (st, [vp, vp, sz, vp]),
    "gbg_memcpy_d2d": (st, [vp, vp, sz, vp]),
    "gbg_memset": (st, [vp, i32, sz, vp]),
    "gbg_stream_create": (st, [C.POINTER(vp)]),
    "gbg_stream_destroy": (st, [vp]),
    "gbg_stream_sync": (st, [vp]),
    "gbg_plan_create": (st, [C.POINTER(C.c_int), C.POINTER(C.c_double), i32, C.POINTER(vp)]),
    "gbg_plan_create_layout": (st, [C.POINTER(C.c_int), C.POINTER(C.c_double), i32,
                                    C.POINTER(C.c_int), C.POINTER(vp)]),
    "gbg_plan_destroy": (st, [vp]),
    "gbg_plan_query": (st, [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)]),
    "gbg_to_spectral": (st, [vp, vp, vp, i32, vp]),
    "gbg_to_physical": (st, [vp, vp, vp, i32, vp, vp]),
    "gbg_gradient": (st, [vp, vp, vp, i32, vp]),
    "gbg_symmetric_gradient": (st, [vp, vp, vp, vp]),
    "gbg_divergence": (st, [vp, vp, vp, i32, i32, dbl, dbl, dbl, dbl, vp]),
    "gbg_curl": (st, [vp, vp, vp, vp]),
    "gbg_inverse_curl": (st, [vp, vp, vp, vp]),
    "gbg_laplacian": (st, [vp, vp, vp, i32, vp]),
    "gbg_inverse_laplacian": (st, [vp, vp, vp, i32, vp]),
    "gbg_helmholtz": (st, [vp, vp, vp, i32, dbl, dbl, dbl, dbl, vp]),
    "gbg_convolve_band": (st, [vp, vp, vp, i32, dbl, dbl, dbl, dbl, vp]),
    "gbg_mean_free": (st, [vp, vp, i32, vp]),
    "gbg_cross": (st, [vp, vp, vp, vp, vp]),
    "gbg_multiply": (st, [vp, vp, vp, vp, vp]),
    "gbg_inner_spectral": (st, [vp, vp, vp, i32, dp, vp]),
    "gbg_inner_physical": (st, [vp, vp, vp, i32, dp, vp]),
    "gbg_radial_shells": (st, [vp, C.POINTER(C.c_int), dp]),
    "gbg_radial_spectrum": (st, [vp, vp, vp, i32, dp, i32, vp]),
    "gbg_exponential_convolve": (st, [vp, vp, vp, i32, dbl, vp, dbl, dbl, dbl, vp]),
    "gbg_exponential_combine": (st, [vp, vp, dbl, vp, dbl, vp, i32, dbl, vp, dbl, vp]),
    "gbg_reality_residual": (st, [vp, vp, i32, dp, vp]),
    "gbg_reality_projection": (st, [vp, vp, vp, i32, vp]),
    "gbg_uxw": (st, [vp, vp, vp, dbl, dp, dp, C.c_uint, vp]),
    "gbg_uxw_combine": (st, [vp, vp, vp, dbl, C.POINTER(GbgUxwCombination), dp, dp, C.c_uint, vp]),
    "gbg_uxw_noalloc": (st, [vp, vp, vp, dbl, dp, dp, C.c_uint, vp]),
    "gbg_dealiased_product": (st, [vp, vp, vp, vp, dbl, vp]),
    "gbg_plan_autotune": (st, [vp, vp, dp]),
    "gbg_batch_create": (st, [i32, i32, C.POINTER(vp)]),
    "gbg_batch_add": (st, [vp, i32, C.POINTER(C.c_int), C.POINTER(C.c_double), i32, vp, vp, vp, dbl]),
    "gbg_batch_instantiate": (st, [vp, vp]),
    "gbg_batch_launch": (st, [vp, vp]),
    "gbg_batch_destroy": (st, [vp]),
    "gbg_server_create": (st, [sz, i32, C.POINTER(vp)]),
    "gbg_server_submit": (st, [vp, vp, C.POINTER(C.c_ulonglong)]),
    "gbg_server_wait": (st, [vp, C.c_ulonglong, dp]),
    "gbg_server_poll": (st, [vp, C.c_ulonglong, C.POINTER(C.c_int), dp]),
    "gbg_server_stats": (st, [vp, dp, dp, dp, C.POINTER(C.c_ulonglong)]),
    "gbg_server_resident": (st, [vp, C.POINTER(sz), C.POINTER(C.c_int)]),
    "gbg_server_destroy": (st, [vp]),
    "gbg_plan_set_profiling": (st, [vp, i32]),
    "gbg_plan_stage_times": (st, [vp, dp, C.POINTER(C.c_int)]),
}

_lib = None


def load(path: str | None = None):
    """Load libgbg.so once and attach the typed signatures."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libgbg.so not built at {p}: run __graft_entry__.build() "
                           "or python gbgperiodicfields.jl_amd/build_lib.py")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class GbgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class ArgumentError(GbgError, ValueError):
    pass


class DimensionMismatch(GbgError, ValueError):
    pass


class UnsupportedError(GbgError, NotImplementedError):
    pass


class AliasError(GbgError, ValueError):
    pass


_EXC = {1: ArgumentError, 2: DimensionMismatch, 3: UnsupportedError, 4: AliasError}


def check(status: int):
    """Raise the Python counterpart of the reference's refusal for a status."""
    if status == GBG_OK:
        return
    msg = load().gbg_last_error().decode(errors="replace")
    raise _EXC.get(status, GbgError)(status, msg)
