"""Host-side mirror of GBGPeriodicFields.jl's spectral hot path over libgbg.

The reference's public calls for this path (``src/GBGPeriodicFields.jl:281-430``)
keep their names and argument meaning; Julia's in-place ``name!`` becomes
``name_`` here, the allocating form keeps the plain name.  Fields live in
device memory owned by this module (``gbg_malloc``), laid out exactly as the
reference's column-major storage:

* physical  ``(Nx, Ny, Nz[, 3[, 3]])``  -> NumPy view ``([3, [3,]] Nz, Ny, Nx)``
* spectral  ``(Nx//2+1, Ny, Nz[, ...])`` -> NumPy view ``([...,] Nz, Ny, Nx//2+1)``

Every operation runs on the GPU through the C ABI; there is no CPU path.
Refusals raise the Python counterparts of the reference's exceptions
(``ArgumentError``, ``DimensionMismatch``, ...).
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (AliasError, ArgumentError, DimensionMismatch, GbgError,  # noqa: F401
                   UnsupportedError, check, load)

TWO_PI = 2.0 * math.pi

__all__ = [
    "Grid3D", "RealFourierLayout", "FieldTransformPlans", "Field", "ScalarFieldPhysical", "VectorFieldPhysical",
    "ScalarFieldSpectral", "VectorFieldSpectral", "TensorFieldSpectral", "Stream",
    "to_spectral_", "to_physical_", "to_spectral", "to_physical", "gradient_", "gradient",
    "symmetric_gradient_", "symmetric_gradient",
    "divergence_", "curl_", "curl", "inverse_curl_", "laplacian_", "inverse_laplacian_",
    "helmholtz_projection_", "spectral_cutoff", "spectral_band", "convolve_",
    "mean_free_projection_", "cross_", "multiply_", "inner", "l2norm", "uxw_",
    "dealiased_product_", "dealias_cutoff", "device_count", "uxw_combine_",
]


def device_count() -> int:
    n = C.c_int(0)
    st = load().gbg_device_count(C.byref(n))
    return n.value if st == 0 else 0


class Stream:
    """A HIP stream owned by the library (``None`` stands for the default)."""

    def __init__(self):
        h = C.c_void_p()
        check(load().gbg_stream_create(C.byref(h)))
        self.handle = h.value

    def synchronize(self):
        check(load().gbg_stream_sync(self.handle))

    def __del__(self):
        try:
            if self.handle:
                load().gbg_stream_destroy(self.handle)
        except Exception:
            pass


def _sh(stream):
    return None if stream is None else (stream.handle if isinstance(stream, Stream) else stream)


@dataclass(frozen=True)
class Grid3D:
    """``Grid3D(grid_size, domain_size)`` (``src/Fields.jl:165-200``)."""

    grid_size: tuple
    domain_size: tuple = (TWO_PI, TWO_PI, TWO_PI)

    def __post_init__(self):
        if len(self.grid_size) != 3 or len(self.domain_size) != 3:
            raise ArgumentError(1, "grid and domain need three axes")
        if any(int(n) < 1 for n in self.grid_size):
            raise ArgumentError(1, "grid sizes must be positive")
        if any(not (float(L) > 0) or not math.isfinite(float(L)) for L in self.domain_size):
            raise ArgumentError(1, "domain lengths must be positive and finite")

    @property
    def nh(self):
        return self.grid_size[0] // 2 + 1

    def physical_shape(self, order=0):
        Nx, Ny, Nz = self.grid_size
        return (3,) * order + (Nz, Ny, Nx)

    def spectral_extents(self, layout=None):
        """``_rfft_shape`` (``src/Fields.jl:239``): the folded axis halved."""
        f = (layout or RealFourierLayout()).fold
        return tuple(n // 2 + 1 if a == f else n for a, n in enumerate(self.grid_size))

    def spectral_shape(self, order=0, layout=None):
        sx, sy, sz = self.spectral_extents(layout)
        return (3,) * order + (sz, sy, sx)


@dataclass(frozen=True)
class RealFourierLayout:
    """``RealFourierLayout(fft_dims)`` (``src/Fields.jl:219-226``): a
    permutation of (1, 2, 3) whose first entry is the folded (halved) axis."""

    fft_dims: tuple = (1, 2, 3)

    def __post_init__(self):
        d = tuple(int(v) for v in self.fft_dims)
        if len(d) != 3 or sorted(d) != [1, 2, 3]:       # _validate_fft_dims, :122
            raise ArgumentError(1, f"fft_dims must be a permutation of (1, 2, 3), got {self.fft_dims}")
        object.__setattr__(self, "fft_dims", d)

    @property
    def fold(self) -> int:
        return self.fft_dims[0] - 1

    @property
    def default(self) -> bool:
        return self.fft_dims == (1, 2, 3)


def dealias_cutoff(n: int, L: float = TWO_PI) -> float:
    """The 2/3-rule cutoff ``(2/3)*(N/2)*(2pi/L)`` of ``examples/HIST_CPU.jl:357``."""
    return ((2.0 / 3.0) * (n / 2.0)) * (TWO_PI / L)


class FieldTransformPlans:
    """Plans for one grid and element type: ``FieldTransformPlans`` plus the
    canonical ``FieldSpectralMultipliers`` (``src/Fields.jl:1381-1405,
    1875-1943``), held by a ``gbg_plan``.  Rank-free: one plan serves scalar,
    vector and tensor fields."""

    def __init__(self, grid: Grid3D, dtype=np.float64, layout: "RealFourierLayout" = None):
        dt = np.dtype(dtype)
        if dt not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise ArgumentError(1, "element type must be Float64 or Float32")
        self.grid = grid
        self.dtype = dt
        self.layout = layout or RealFourierLayout()
        n = (C.c_int * 3)(*[int(v) for v in grid.grid_size])
        L = (C.c_double * 3)(*[float(v) for v in grid.domain_size])
        h = C.c_void_p()
        if self.layout.default:
            check(load().gbg_plan_create(n, L, dt.itemsize, C.byref(h)))
        else:
            dims = (C.c_int * 3)(*self.layout.fft_dims)
            check(load().gbg_plan_create_layout(n, L, dt.itemsize, dims, C.byref(h)))
        self.handle = h.value

    @property
    def cdtype(self):
        return np.complex128 if self.dtype == np.float64 else np.complex64

    def device_bytes(self) -> int:
        a, b, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
        check(load().gbg_plan_query(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return c.value

    def set_profiling(self, on: bool):
        check(load().gbg_plan_set_profiling(self.handle, 1 if on else 0))

    def stage_times(self):
        buf = (C.c_double * 16)()
        n = C.c_int(16)
        check(load().gbg_plan_stage_times(self.handle, buf, C.byref(n)))
        return [buf[i] for i in range(n.value)]

    def close(self):
        if getattr(self, "handle", None):
            load().gbg_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Field:
    """A device-resident field: representation ('physical' | 'spectral'),
    tensor order R (0, 1, 2), element type and grid.  ``data`` is the device
    pointer of the contiguous column-major storage."""

    def __init__(self, grid: Grid3D, representation: str, order: int, dtype=np.float64,
                 layout: "RealFourierLayout" = None):
        if representation not in ("physical", "spectral"):
            raise ArgumentError(1, "representation must be physical or spectral")
        if order not in (0, 1, 2):
            raise DimensionMismatch(2, "tensor order must be 0, 1 or 2")
        self.grid = grid
        self.representation = representation
        self.order = order
        self.dtype = np.dtype(dtype)
        self.layout = layout or RealFourierLayout()
        if self.spectral:
            self.shape = grid.spectral_shape(order, self.layout)
            self.npdtype = np.dtype(np.complex128 if self.dtype == np.float64 else np.complex64)
        else:
            self.shape = grid.physical_shape(order)
            self.npdtype = self.dtype
        self.nbytes = int(np.prod(self.shape)) * self.npdtype.itemsize
        p = C.c_void_p()
        check(load().gbg_malloc(C.byref(p), self.nbytes))
        self.data = p.value

    @property
    def spectral(self):
        return self.representation == "spectral"

    @property
    def ncomp(self):
        return 3 ** self.order

    def upload(self, arr: np.ndarray, stream=None) -> "Field":
        a = np.ascontiguousarray(arr, dtype=self.npdtype)
        if a.shape != self.shape:
            raise DimensionMismatch(2, f"array shape {a.shape} != field shape {self.shape}")
        check(load().gbg_memcpy_h2d(self.data, a.ctypes.data, self.nbytes, _sh(stream)))
        return self

    def download(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.npdtype)
        check(load().gbg_memcpy_d2h(out.ctypes.data, self.data, self.nbytes, _sh(stream)))
        return out

    def copy_from(self, other: "Field", stream=None) -> "Field":
        _same_domain(self, other)
        if self.shape != other.shape:
            raise DimensionMismatch(2, "shape mismatch")
        check(load().gbg_memcpy_d2d(self.data, other.data, self.nbytes, _sh(stream)))
        return self

    def fill_zero(self, stream=None) -> "Field":
        check(load().gbg_memset(self.data, 0, self.nbytes, _sh(stream)))
        return self

    def similar(self, representation=None, order=None) -> "Field":
        return Field(self.grid, representation or self.representation,
                     self.order if order is None else order, self.dtype, self.layout)

    def free(self):
        if getattr(self, "data", None):
            load().gbg_free(self.data)
            self.data = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def ScalarFieldPhysical(grid, dtype=np.float64, layout=None):
    return Field(grid, "physical", 0, dtype, layout)


def VectorFieldPhysical(grid, dtype=np.float64, layout=None):
    return Field(grid, "physical", 1, dtype, layout)


def ScalarFieldSpectral(grid, dtype=np.float64, layout=None):
    return Field(grid, "spectral", 0, dtype, layout)


def VectorFieldSpectral(grid, dtype=np.float64, layout=None):
    return Field(grid, "spectral", 1, dtype, layout)


def TensorFieldSpectral(grid, dtype=np.float64, layout=None):
    return Field(grid, "spectral", 2, dtype, layout)


# ---- argument checks mirroring the reference's refusals ------------------------

def _same_domain(a: Field, b: Field):
    """``_check_field_domain_compatibility`` (``src/Fields.jl:1225-1250``)."""
    if a.grid != b.grid:
        raise ArgumentError(1, "fields live on different grids")
    if a.dtype != b.dtype:
        raise ArgumentError(1, "field scalar types must be equal")
    if a.layout != b.layout:
        raise ArgumentError(1, "fields were built for different Fourier layouts")


def _with_plan(f: Field, plans: FieldTransformPlans):
    """``_check_field_matches_context`` (``src/Fields.jl:1301-1305``)."""
    if f.grid != plans.grid:
        raise ArgumentError(1, "field grid does not match the plans' grid")
    if f.dtype != plans.dtype:
        raise ArgumentError(1, "field and plans scalar types must be equal")
    if f.layout != plans.layout:
        raise ArgumentError(1, "field Fourier layout must be equal to FieldTransformPlans Fourier layout")


def _default_layout(f: Field, what: str):
    """The batch and the server make plans of the default layout."""
    if not f.layout.default:
        raise UnsupportedError(3, f"{what} takes fields of the default layout RealFourierLayout((1, 2, 3))")


def _need(f: Field, rep: str, order=None):
    if f.representation != rep:
        raise UnsupportedError(3, f"no method for a {f.representation} field (needs {rep})")
    if order is not None and f.order != order:
        raise DimensionMismatch(2, f"tensor order {f.order}, expected {order}")


# ---- transforms (src/Fields.jl:1589-1706) ---------------------------------------

def to_spectral_(dst: Field, src: Field, plans: FieldTransformPlans, stream=None) -> Field:
    """``to_spectral!(dst, src, plans)``: unnormalised forward rfft."""
    _need(dst, "spectral")
    _need(src, "physical")
    _with_plan(dst, plans)
    _with_plan(src, plans)
    if dst.order != src.order:
        raise DimensionMismatch(2, "tensor orders differ")
    check(load().gbg_to_spectral(plans.handle, dst.data, src.data, src.ncomp, _sh(stream)))
    return dst


def to_physical_(dst: Field, src: Field, plans: FieldTransformPlans, scratch: Field | None = None,
                 stream=None) -> Field:
    """``to_physical!(dst, src, plans[, scratch])``: inverse c2r with 1/N.  The
    3-argument form consumes ``src``; with ``scratch`` the source survives and
    ``scratch is src`` is refused (``src/Fields.jl:1600-1619``)."""
    _need(dst, "physical")
    _need(src, "spectral")
    _with_plan(dst, plans)
    _with_plan(src, plans)
    if dst.order != src.order:
        raise DimensionMismatch(2, "tensor orders differ")
    sc = None
    if scratch is not None:
        if scratch is src:
            raise AliasError(4, "scratch must not be the spectral source")
        _need(scratch, "spectral", src.order)
        _with_plan(scratch, plans)
        sc = scratch.data
    check(load().gbg_to_physical(plans.handle, dst.data, src.data, src.ncomp, sc, _sh(stream)))
    return dst


def to_spectral(src: Field, plans: FieldTransformPlans, stream=None) -> Field:
    out = Field(src.grid, "spectral", src.order, src.dtype, src.layout)
    return to_spectral_(out, src, plans, stream)


def to_physical(src: Field, plans: FieldTransformPlans, stream=None) -> Field:
    """Allocating form: copies the source first so it is preserved."""
    out = Field(src.grid, "physical", src.order, src.dtype, src.layout)
    tmp = src.similar()
    return to_physical_(out, src, plans, scratch=tmp, stream=stream)


# ---- operators (src/DifferentialOperators.jl) -------------------------------------

def gradient_(dst: Field, src: Field, plans: FieldTransformPlans, stream=None) -> Field:
    """``gradient!`` scalar->vector / vector->tensor (``:812-837``)."""
    _need(src, "spectral")
    _need(dst, "spectral", src.order + 1)
    _with_plan(src, plans)
    _same_domain(dst, src)
    if src.order > 1:
        raise DimensionMismatch(2, "gradient of a tensor field is not defined here")
    check(load().gbg_gradient(plans.handle, dst.data, src.data, src.order, _sh(stream)))
    return dst


def gradient(src: Field, plans: FieldTransformPlans, stream=None) -> Field:
    return gradient_(Field(src.grid, "spectral", src.order + 1, src.dtype, src.layout), src, plans,
                     stream)


def symmetric_gradient_(dst: Field, src: Field, plans: FieldTransformPlans, stream=None) -> Field:
    """``symmetric_gradient!`` vector->symmetric tensor (``:1289-1300``)."""
    _need(src, "spectral", 1)
    _need(dst, "spectral", 2)
    _with_plan(src, plans)
    _same_domain(dst, src)
    check(load().gbg_symmetric_gradient(plans.handle, dst.data, src.data, _sh(stream)))
    return dst


def symmetric_gradient(src: Field, plans: FieldTransformPlans, stream=None) -> Field:
    return symmetric_gradient_(Field(src.grid, "spectral", 2, src.dtype, src.layout), src, plans,
                               stream)


def divergence_(dst: Field, src: Field, plans: FieldTransformPlans, kernel=None, alpha=1.0,
                beta=0.0, stream=None) -> Field:
    """``divergence!`` vector->scalar / tensor->vector (``:948-1037``)."""
    _need(src, "spectral")
    if src.order < 1:
        raise DimensionMismatch(2, "divergence needs a vector or tensor field")
    _need(dst, "spectral", src.order - 1)
    _with_plan(src, plans)
    _same_domain(dst, src)
    hb, lo, hi = _band_args(kernel)
    check(load().gbg_divergence(plans.handle, dst.data, src.data, src.order, hb, lo, hi,
                                float(alpha), float(beta), _sh(stream)))
    return dst


def curl_(dst: Field, src: Field, plans: FieldTransformPlans, stream=None) -> Field:
    """``curl!`` (``:1135-1149``); dst may be src."""
    _need(src, "spectral", 1)
    _need(dst, "spectral", 1)
    _with_plan(src, plans)
    _same_domain(dst, src)
    check(load().gbg_curl(plans.handle, dst.data, src.data, _sh(stream)))
    return dst


def curl(src: Field, plans: FieldTransformPlans, stream=None) -> Field:
    return curl_(src.similar(), src, plans, stream)


def inverse_curl_(dst: Field, src: Field, plans: FieldTransformPlans, stream=None) -> Field:
    """``inverse_curl!`` (``:1212-1224``); dst may be src."""
    _need(src, "spectral", 1)
    _need(dst, "spectral", 1)
    _with_plan(src, plans)
    _same_domain(dst, src)
    check(load().gbg_inverse_curl(plans.handle, dst.data, src.data, _sh(stream)))
    return dst


def laplacian_(dst: Field, src: Field, plans: FieldTransformPlans, stream=None) -> Field:
    """``laplacian!`` (``:1369-1376``)."""
    _need(src, "spectral")
    _need(dst, "spectral", src.order)
    _with_plan(src, plans)
    _same_domain(dst, src)
    check(load().gbg_laplacian(plans.handle, dst.data, src.data, src.ncomp, _sh(stream)))
    return dst


def inverse_laplacian_(dst: Field, src: Field, plans: FieldTransformPlans, stream=None) -> Field:
    """``inverse_laplacian!`` (``:1457-1464``)."""
    _need(src, "spectral")
    _need(dst, "spectral", src.order)
    _with_plan(src, plans)
    _same_domain(dst, src)
    check(load().gbg_inverse_laplacian(plans.handle, dst.data, src.data, src.ncomp, _sh(stream)))
    return dst


def helmholtz_projection_(dst: Field, src: Field, plans: FieldTransformPlans, kernel=None,
                          alpha=1.0, beta=0.0, stream=None) -> Field:
    """``helmholtz_projection!(dst, src, m[, kernel, alpha, beta])`` (``:1574-1593``)."""
    _need(src, "spectral", 1)
    _need(dst, "spectral", 1)
    _with_plan(src, plans)
    _same_domain(dst, src)
    hb, lo, hi = _band_args(kernel)
    check(load().gbg_helmholtz(plans.handle, dst.data, src.data, hb, lo, hi, float(alpha),
                               float(beta), _sh(stream)))
    return dst


@dataclass(frozen=True)
class RadialBand:
    """``spectral_band``/``spectral_cutoff`` as a radial symbol (``src/Convolution.jl:1832-1882``)."""

    k_lo: float
    k_hi: float


def spectral_band(plans: FieldTransformPlans, k_lo: float, k_hi: float) -> RadialBand:
    if not (math.isfinite(k_lo) and math.isfinite(k_hi)):
        raise ArgumentError(1, f"band edges must be finite, got ({k_lo}, {k_hi})")
    if not (0 <= k_lo <= k_hi):
        raise ArgumentError(1, f"band edges must satisfy 0 <= low <= high, got ({k_lo}, {k_hi})")
    return RadialBand(float(k_lo), float(k_hi))


def spectral_cutoff(plans: FieldTransformPlans, kc: float) -> RadialBand:
    return spectral_band(plans, 0.0, kc)


def _band_args(kernel):
    if kernel is None:
        return 0, 0.0, 0.0
    if not isinstance(kernel, RadialBand):
        raise UnsupportedError(3, "only radial band kernels are supported")
    return 1, kernel.k_lo, kernel.k_hi


def convolve_(dst: Field, src: Field, kernel: RadialBand, plans: FieldTransformPlans, alpha=1.0,
              beta=0.0, stream=None) -> Field:
    """``convolve!(dst, src, kernel[, alpha, beta])`` for a radial band kernel
    (``src/Convolution.jl:1035-1065``)."""
    _need(src, "spectral")
    _need(dst, "spectral", src.order)
    _with_plan(src, plans)
    _same_domain(dst, src)
    _, lo, hi = _band_args(kernel)
    check(load().gbg_convolve_band(plans.handle, dst.data, src.data, src.ncomp, lo, hi,
                                   float(alpha), float(beta), _sh(stream)))
    return dst


def mean_free_projection_(f: Field, plans: FieldTransformPlans, stream=None) -> Field:
    """``mean_free_projection!`` (``src/Fields.jl:3124-3128``)."""
    _need(f, "spectral")
    _with_plan(f, plans)
    check(load().gbg_mean_free(plans.handle, f.data, f.ncomp, _sh(stream)))
    return f


# ---- physical products (src/FieldAlgebra.jl) ---------------------------------------

def cross_(dst: Field, u: Field, v: Field, plans: FieldTransformPlans, stream=None) -> Field:
    """``cross!(dst, u, v)`` (``src/FieldAlgebra.jl:1088-1092``)."""
    for f in (dst, u, v):
        _need(f, "physical", 1)
        _with_plan(f, plans)
    check(load().gbg_cross(plans.handle, dst.data, u.data, v.data, _sh(stream)))
    return dst


def multiply_(dst: Field, a: Field, b: Field, plans: FieldTransformPlans, stream=None) -> Field:
    """``multiply!(dst, a, b)`` for scalar fields (``src/FieldAlgebra.jl:778``)."""
    for f in (dst, a, b):
        _need(f, "physical", 0)
        _with_plan(f, plans)
    check(load().gbg_multiply(plans.handle, dst.data, a.data, b.data, _sh(stream)))
    return dst


# ---- reductions (src/Reductions.jl) -------------------------------------------------

def inner(f: Field, g: Field, plans: FieldTransformPlans, stream=None) -> float:
    """``inner(f, g)``: spectral through Parseval (``src/Reductions.jl:883-925``)
    or physical, the cell average of the pointwise contraction (``:838-853``)."""
    if f.representation != g.representation:
        raise ArgumentError(1, "inner needs two fields in the same representation")
    _with_plan(f, plans)
    _same_domain(f, g)
    if f.order != g.order:
        raise DimensionMismatch(2, f"sections have different tensor orders: {f.order} and {g.order}")
    out = C.c_double(0.0)
    fn = load().gbg_inner_spectral if f.spectral else load().gbg_inner_physical
    check(fn(plans.handle, f.data, g.data, f.ncomp, C.byref(out), _sh(stream)))
    return out.value


def l2norm(f: Field, plans: FieldTransformPlans, stream=None) -> float:
    """``l2norm(f) = sqrt(inner(f, f))`` (``src/Reductions.jl:947-948``)."""
    return math.sqrt(inner(f, f, plans, stream))


def radial_spectrum(f: Field, *args, stream=None):
    """``radial_spectrum(f, m)`` / ``radial_spectrum(f, g, m)`` (``src/Reductions.jl:1309-1455``):
    returns ``(shells, spectrum)``, FP64 numpy arrays; the shell abscissa is
    ``m dk`` with ``dk = 2 pi / max L``."""
    if len(args) == 1:
        g, plans = None, args[0]
    elif len(args) == 2:
        g, plans = args
    else:
        raise ArgumentError(1, "radial_spectrum(f, plans) or radial_spectrum(f, g, plans)")
    _need(f, "spectral")
    _with_plan(f, plans)
    if g is not None:
        _need(g, "spectral")
        _same_domain(f, g)
        if f.order != g.order:
            raise ArgumentError(1, f"radial_spectrum(f, g, m) needs equal tensor orders "
                                   f"(got {f.order} and {g.order})")
    n, dk = C.c_int(0), C.c_double(0.0)
    check(load().gbg_radial_shells(plans.handle, C.byref(n), C.byref(dk)))
    spec = np.zeros(n.value, dtype=np.float64)
    check(load().gbg_radial_spectrum(plans.handle, f.data, None if g is None else g.data, f.ncomp,
                                     spec.ctypes.data_as(C.POINTER(C.c_double)), n.value,
                                     _sh(stream)))
    return np.arange(n.value) * dk.value, spec


# ---- exponential kernel (src/Convolution.jl:1197-1795) -----------------------------

def _offset_args(offset, f: Field):
    if offset is None:
        return None
    _need(offset, "spectral", 0)
    _same_domain(offset, f)
    return offset.data


def exponential_convolve_(dst: Field, *args, coefficient: float, offset: Field | None = None,
                          offset_scale: float = 1.0, stream=None) -> Field:
    """``exponential_convolve!(dst, f, m[, alpha, beta]; coefficient, offset, offset_scale)``
    and the two-operand ``exponential_convolve!(dst, alpha, a, beta, b, m; ...)``
    (``src/Convolution.jl:1750``): the kernel ``exp(c |k|^2 + s w)`` evaluated per
    mode, never stored."""
    _need(dst, "spectral")
    if len(args) >= 1 and isinstance(args[0], Field):
        f, plans = args[0], args[1]
        alpha, beta = (args[2], args[3]) if len(args) == 4 else (1.0, 0.0)
        if len(args) not in (2, 4):
            raise ArgumentError(1, "exponential_convolve!(dst, f, m[, alpha, beta])")
        _need(f, "spectral", dst.order)
        _with_plan(f, plans)
        _same_domain(dst, f)
        check(load().gbg_exponential_convolve(plans.handle, dst.data, f.data, f.ncomp,
                                              float(coefficient), _offset_args(offset, f),
                                              float(offset_scale), float(alpha), float(beta),
                                              _sh(stream)))
        return dst
    if len(args) != 5:
        raise ArgumentError(1, "exponential_convolve!(dst, alpha, a, beta, b, m)")
    alpha, a, beta, b, plans = args
    for x in (a, b):
        _need(x, "spectral", dst.order)
        _same_domain(dst, x)
    _with_plan(a, plans)
    check(load().gbg_exponential_combine(plans.handle, dst.data, float(alpha), a.data, float(beta),
                                         b.data, a.ncomp, float(coefficient),
                                         _offset_args(offset, a), float(offset_scale), _sh(stream)))
    return dst


def exponential_convolve(f: Field, plans: FieldTransformPlans, *, coefficient: float,
                         offset: Field | None = None, offset_scale: float = 1.0,
                         stream=None) -> Field:
    """Allocating ``exponential_convolve``."""
    return exponential_convolve_(f.similar(), f, plans, coefficient=coefficient, offset=offset,
                                 offset_scale=offset_scale, stream=stream)


# ---- reality condition (src/Fields.jl:2858-3094) -----------------------------------

def reality_residual(f: Field, plans: FieldTransformPlans, stream=None) -> float:
    """``reality_residual(f)``: the largest ``|f(k) - conj f(-k)|`` on the
    self-conjugate planes (``src/Fields.jl:2918-2928``)."""
    _need(f, "spectral")
    _with_plan(f, plans)
    out = C.c_double(0.0)
    check(load().gbg_reality_residual(plans.handle, f.data, f.ncomp, C.byref(out), _sh(stream)))
    return out.value


def reality_projection_(dst: Field, *args, stream=None) -> Field:
    """``reality_projection!(f, plans)`` in place or ``reality_projection!(dst, src,
    plans)`` (``src/Fields.jl:2940-3094``)."""
    if len(args) == 1:
        src, plans = dst, args[0]
    elif len(args) == 2:
        src, plans = args
    else:
        raise ArgumentError(1, "reality_projection!(f, plans) or (dst, src, plans)")
    _need(src, "spectral")
    _need(dst, "spectral", src.order)
    _with_plan(src, plans)
    _same_domain(dst, src)
    check(load().gbg_reality_projection(plans.handle, dst.data, src.data, src.ncomp, _sh(stream)))
    return dst


# ---- fused pipelines ---------------------------------------------------------------

def _uxw_flags(fused: bool, slabs: bool) -> int:
    return (0 if fused else _lib.GBG_UXW_UNFUSED) | (_lib.GBG_UXW_SLABS if slabs else 0)


def uxw_(n_hat: Field, u_hat: Field, plans: FieldTransformPlans, kc: float | None = None,
         fused: bool = True, reductions: bool = True, slabs: bool = False, stream=None):
    """The dealiased rotational nonlinear term ``N^ = Q M P F(u x w)`` of
    ``examples/HIST_CPU.jl:229-247`` in one call.  Returns ``(E, Z)`` (FP64)
    when ``reductions`` is true, else ``None`` (no synchronisation).
    ``slabs``: the middle passes z-slab by z-slab (same bits, smaller work
    area)."""
    _need(u_hat, "spectral", 1)
    _need(n_hat, "spectral", 1)
    _with_plan(u_hat, plans)
    _same_domain(n_hat, u_hat)
    if kc is None:
        kc = dealias_cutoff(plans.grid.grid_size[0], plans.grid.domain_size[0])
    flags = _uxw_flags(fused, slabs)
    if reductions:
        E, Z = C.c_double(0.0), C.c_double(0.0)
        check(load().gbg_uxw(plans.handle, u_hat.data, n_hat.data, float(kc), C.byref(E), C.byref(Z),
                             flags, _sh(stream)))
        return E.value, Z.value
    check(load().gbg_uxw(plans.handle, u_hat.data, n_hat.data, float(kc), None, None, flags,
                         _sh(stream)))
    return None


def uxw_combine_(dst: Field, u_hat: Field, plans: FieldTransformPlans, kc: float | None = None, *,
                 n_hat: Field | None = None, a: Field | None = None, alpha: float = 1.0,
                 b: Field | None = None, beta: float = 0.0, delta: float = 0.0, gamma: float = 0.0,
                 coefficient: float = 0.0, fused: bool = True, reductions: bool = True,
                 slabs: bool = False, stream=None):
    """u x omega with the integrating-factor step of ``step!``
    (``examples/HIST_CPU.jl:267-282``) folded into its last pass:
    ``dst = exp(c|k|^2) (alpha a + beta b + delta N^) + gamma N^``, and N^ into
    ``n_hat`` when given.  Returns ``(E, Z)`` as ``uxw_``."""
    _need(u_hat, "spectral", 1)
    _need(dst, "spectral", 1)
    _with_plan(u_hat, plans)
    _same_domain(dst, u_hat)
    for f in (n_hat, a, b):
        if f is not None:
            _need(f, "spectral", 1)
            _same_domain(f, u_hat)
    if dst is u_hat or (n_hat is not None and (n_hat is u_hat or n_hat is dst)):
        raise AliasError(4, "dst and n_hat must not be u_hat or each other")
    if kc is None:
        kc = dealias_cutoff(plans.grid.grid_size[0], plans.grid.domain_size[0])
    cmb = _lib.GbgUxwCombination(dst.data, None if a is None else a.data, float(alpha),
                                 None if b is None else b.data, float(beta), float(delta),
                                 float(gamma), float(coefficient))
    flags = _uxw_flags(fused, slabs)
    E, Z = C.c_double(0.0), C.c_double(0.0)
    check(load().gbg_uxw_combine(plans.handle, u_hat.data, None if n_hat is None else n_hat.data,
                                 float(kc), C.byref(cmb), C.byref(E) if reductions else None,
                                 C.byref(Z) if reductions else None, flags, _sh(stream)))
    return (E.value, Z.value) if reductions else None


class Batch:
    """Many independent (field, operation) items run as one graph launch.

    ``add_scalar(a_hat)`` -> (grad a, dealiased a*a); ``add_vector(u_hat)`` ->
    (grad u, dealiased u x w).  Outputs are allocated here and returned; all
    buffers stay put until the batch is closed."""

    SCALAR, VECTOR = 0, 1

    def __init__(self, nstreams: int = 4, autotune: bool = True):
        h = C.c_void_p()
        check(load().gbg_batch_create(int(nstreams), 1 if autotune else 0, C.byref(h)))
        self.handle = h.value
        self.keep = []

    def _add(self, kind, f: Field, kc):
        _default_layout(f, "the batch")
        g = f.grid
        if kc is None:
            kc = min(dealias_cutoff(n, L) for n, L in zip(g.grid_size, g.domain_size))
        out0 = Field(g, "spectral", f.order + 1, f.dtype)
        out1 = Field(g, "spectral", f.order, f.dtype)
        n = (C.c_int * 3)(*g.grid_size)
        L = (C.c_double * 3)(*[float(v) for v in g.domain_size])
        check(load().gbg_batch_add(self.handle, kind, n, L, f.dtype.itemsize, f.data, out0.data,
                                   out1.data, float(kc)))
        self.keep += [f, out0, out1]
        return out0, out1

    def add_scalar(self, a_hat: Field, kc=None):
        _need(a_hat, "spectral", 0)
        return self._add(self.SCALAR, a_hat, kc)

    def add_vector(self, u_hat: Field, kc=None):
        _need(u_hat, "spectral", 1)
        return self._add(self.VECTOR, u_hat, kc)

    def instantiate(self, stream=None):
        check(load().gbg_batch_instantiate(self.handle, _sh(stream)))

    def launch(self, stream=None):
        check(load().gbg_batch_launch(self.handle, _sh(stream)))

    def close(self):
        if getattr(self, "handle", None):
            load().gbg_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop over the C ABI's gbg_server (see
    include/gbg.h).  ``submit`` returns a ticket at once; ``wait`` blocks for
    that job and returns its latency in ms."""

    TO_SPECTRAL, TO_PHYSICAL, CURL, INVERSE_LAPLACIAN, DEALIASED_PRODUCT, UXW = range(6)

    def __init__(self, hbm_budget_bytes: int, slots: int = 4):
        h = C.c_void_p()
        check(load().gbg_server_create(C.c_size_t(int(hbm_budget_bytes)), int(slots), C.byref(h)))
        self.handle = h.value
        self.inflight = {}

    # kind -> (representation of in0, of out, tensor order or None = any, kept)
    _IO = {0: ("physical", "spectral", None), 1: ("spectral", "physical", None),
           2: ("spectral", "spectral", 1), 3: ("spectral", "spectral", None),
           4: ("spectral", "spectral", 0), 5: ("spectral", "spectral", 1)}

    def submit(self, kind: int, out: Field, in0: Field, in1: Field | None = None,
               kc: float | None = None) -> int:
        """Queue one job; the refusals of the matching standalone call apply
        (representation, tensor order, grid and precision of every buffer;
        no output that is an input, except the in-place curl and inverse
        Laplacian)."""
        if kind not in self._IO:
            raise ArgumentError(1, f"unknown job kind {kind}")
        rin, rout, order = self._IO[kind]
        _need(in0, rin, order)
        _default_layout(in0, "the server")
        _need(out, rout, in0.order)
        _same_domain(out, in0)
        if kind == self.DEALIASED_PRODUCT:
            if in1 is None:
                raise ArgumentError(1, "the dealiased product takes two inputs")
            _need(in1, "spectral", 0)
            _same_domain(in1, in0)
            if out is in1:
                raise AliasError(4, "the product's output must not be an input")
        elif in1 is not None:
            raise ArgumentError(1, "this job kind takes one input")
        if out is in0 and kind not in (self.CURL, self.INVERSE_LAPLACIAN):
            raise AliasError(4, "this job kind cannot write over its input")
        g = in0.grid
        if kc is None:
            kc = min(dealias_cutoff(n, L) for n, L in zip(g.grid_size, g.domain_size))
        j = _lib.GbgJob()
        j.kind = int(kind)
        j.n = (C.c_int * 3)(*g.grid_size)
        j.L = (C.c_double * 3)(*[float(v) for v in g.domain_size])
        j.elem_bytes = in0.dtype.itemsize
        j.ncomp = in0.ncomp
        j.in0 = in0.data
        j.in1 = in1.data if in1 is not None else None
        j.out = out.data
        j.kc = float(kc)
        t = C.c_ulonglong(0)
        check(load().gbg_server_submit(self.handle, C.byref(j), C.byref(t)))
        self.inflight[t.value] = (out, in0, in1)   # keep buffers alive
        return t.value

    def wait(self, ticket: int) -> float:
        ms = C.c_double(0.0)
        try:
            check(load().gbg_server_wait(self.handle, C.c_ulonglong(ticket), C.byref(ms)))
        finally:
            self.inflight.pop(ticket, None)
        return ms.value

    def poll(self, ticket: int):
        """The job's latency (ms) once it has finished, else None (non-blocking)."""
        done, ms = C.c_int(0), C.c_double(0.0)
        try:
            check(load().gbg_server_poll(self.handle, C.c_ulonglong(ticket), C.byref(done),
                                         C.byref(ms)))
        except Exception:
            self.inflight.pop(ticket, None)
            raise
        if not done.value:
            return None
        self.inflight.pop(ticket, None)
        return ms.value

    def stats(self):
        p50, p99, tput = C.c_double(), C.c_double(), C.c_double()
        n = C.c_ulonglong()
        check(load().gbg_server_stats(self.handle, C.byref(p50), C.byref(p99), C.byref(tput),
                                      C.byref(n)))
        return {"p50_ms": p50.value, "p99_ms": p99.value, "jobs_per_s": tput.value,
                "completed": n.value}

    def resident(self):
        b, k = C.c_size_t(), C.c_int()
        check(load().gbg_server_resident(self.handle, C.byref(b), C.byref(k)))
        return b.value, k.value

    def close(self):
        if getattr(self, "handle", None):
            load().gbg_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(plans: FieldTransformPlans, stream=None) -> float:
    """Tune the plan's per-axis pass choice; returns the forward + inverse ms."""
    ms = C.c_double(0.0)
    check(load().gbg_plan_autotune(plans.handle, _sh(stream), C.byref(ms)))
    return ms.value


def dealiased_product_(p_hat: Field, a_hat: Field, b_hat: Field, plans: FieldTransformPlans,
                       kc: float | None = None, stream=None) -> Field:
    """``p^ = M F(F^-1(M a^) F^-1(M b^))`` for scalar fields."""
    for f in (p_hat, a_hat, b_hat):
        _need(f, "spectral", 0)
        _with_plan(f, plans)
    if kc is None:
        kc = dealias_cutoff(plans.grid.grid_size[0], plans.grid.domain_size[0])
    check(load().gbg_dealiased_product(plans.handle, p_hat.data, a_hat.data, b_hat.data, float(kc),
                                       _sh(stream)))
    return p_hat
