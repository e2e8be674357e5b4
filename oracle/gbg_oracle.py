"""FP64 CPU oracle for the spectral hot path of GBGPeriodicFields.jl.

TEST INFRASTRUCTURE ONLY.  This module restates the reference's mathematics in
NumPy so that the HIP library can be checked against it.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and only as the checker: nothing in the product path calls it.

Array conventions (a zero-copy view of the reference's Julia column-major
storage, ``src/Fields.jl:239, 369, 388``):

* physical scalar   ``(Nz, Ny, Nx)`` float64, C order, so x is fastest;
* physical vector   ``(3, Nz, Ny, Nx)`` -- component blocks contiguous;
* spectral scalar   ``(Nz, Ny, Nx//2+1)`` complex128 (the folded axis is x,
  ``fft_dims = (1, 2, 3)``);
* spectral vector   ``(3, Nz, Ny, Nx//2+1)``.

Every function cites the reference file:line it follows.  The FFT itself is
the third-party part of the reference (FFTW 3.3.11 through FFTW.jl 1.10 /
AbstractFFTs 1.5, called at ``src/Fields.jl:1475-1481, 1520, 1535``); it is
restated here with NumPy/SciPy pocketfft using the same convention (forward
unnormalised, inverse carrying 1/N, ``docs/FORMALIZATION.md:86-100``) and is
pinned in ``tests/test_oracle.py`` against a brute-force DFT.

Pinning (see ``tests/test_oracle.py``): wavenumber tables vs rfftfreq/fftfreq,
transforms vs direct O(N^2) DFT and round trips, operators vs closed forms
(sin/cos derivatives, curl of a shear, Nyquist split, projector identities),
the 2/3 mask vs an integer predicate, Parseval vs the physical inner, and the
u x omega pipeline vs two closed forms (a Beltrami field whose product
vanishes and a shear pair whose projected product is known analytically).
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

try:  # pocketfft with threads; numpy.fft is the single-threaded fallback
    import scipy.fft as _sfft
except Exception:  # pragma: no cover
    _sfft = None

TWO_PI = 2.0 * math.pi


def _workers() -> int:
    return int(os.environ.get("GBG_ORACLE_WORKERS", os.cpu_count() or 1))


# ---------------------------------------------------------------------------
# Grid, wavenumbers, multipliers
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class Grid3D:
    """Uniform periodic box ``src/Fields.jl:165-200``.  ``n = (Nx, Ny, Nz)``,
    ``L = (Lx, Ly, Lz)``; ``fold`` is the folded (real-FFT, halved) axis,
    0-based: ``RealFourierLayout(fft_dims).fft_dims[1] - 1`` (``:219-226``),
    x by default."""

    n: tuple
    L: tuple = (TWO_PI, TWO_PI, TWO_PI)
    fold: int = 0

    @property
    def nh(self) -> int:                       # stored length of the folded axis
        return self.n[self.fold] // 2 + 1

    @property
    def extents(self):                         # _rfft_shape, src/Fields.jl:239
        return tuple(self.nh if a == self.fold else self.n[a] for a in range(3))

    @property
    def physical_shape(self):                  # field_physical_shape, :369
        return (self.n[2], self.n[1], self.n[0])

    @property
    def spectral_shape(self):                  # field_spectral_shape, :388
        s = self.extents
        return (s[2], s[1], s[0])

    @property
    def npoints(self) -> int:
        return self.n[0] * self.n[1] * self.n[2]

    @property
    def nmodes(self) -> int:
        s = self.extents
        return s[0] * s[1] * s[2]

    def fold_view(self, a: np.ndarray) -> np.ndarray:
        """A 1-D array along the folded axis shaped to broadcast over
        (..., Nz', Ny', Nx')."""
        shape = [1, 1, 1]
        shape[2 - self.fold] = a.shape[0]
        return a.reshape(shape)


def axis_wavenumber_values(n: int, L: float, half: bool):
    """Restates ``_axis_wavenumber_values`` (``src/Fields.jl:2006-2026``).

    Returns ``(derivative, squared)``: ``unit = 2*pi/L``; stored index i ->
    frequency i on the folded axis, else i if i < (n+1)//2 else i-n; ``k =
    unit*freq``; squared = k*k (true k^2 everywhere); the derivative family is
    zero at the Nyquist entry of an even axis."""
    unit = (2.0 * math.pi) / float(L)
    stored = n // 2 + 1 if half else n
    i = np.arange(stored)
    freq = i if half else np.where(i < (n + 1) // 2, i, i - n)
    k = unit * freq.astype(np.float64)
    squared = k * k
    deriv = k.copy()
    if n % 2 == 0:
        nyq = stored - 1 if half else n // 2
        deriv[nyq] = 0.0
    return deriv, squared


@dataclass
class Multipliers:
    """Canonical ``FieldSpectralMultipliers`` (``src/Fields.jl:1875-1943,
    1989-1998``): three derivative and three squared per-axis tables."""

    grid: Grid3D
    k1: np.ndarray
    k2: np.ndarray
    k3: np.ndarray
    s1: np.ndarray
    s2: np.ndarray
    s3: np.ndarray

    @classmethod
    def build(cls, grid: Grid3D) -> "Multipliers":
        k1, s1 = axis_wavenumber_values(grid.n[0], grid.L[0], grid.fold == 0)
        k2, s2 = axis_wavenumber_values(grid.n[1], grid.L[1], grid.fold == 1)
        k3, s3 = axis_wavenumber_values(grid.n[2], grid.L[2], grid.fold == 2)
        return cls(grid, k1, k2, k3, s1, s2, s3)

    # broadcast views onto (Nz, Ny, Nh)
    def a1(self):
        return self.k1[None, None, :]

    def a2(self):
        return self.k2[None, :, None]

    def a3(self):
        return self.k3[:, None, None]

    def ksq(self):
        """``_squared_wavenumber_broadcast`` (``src/Fields.jl:2105-2109``):
        (k1^2 + k2^2) + k3^2, left association."""
        return (self.s1[None, None, :] + self.s2[None, :, None]) + self.s3[:, None, None]


# ---------------------------------------------------------------------------
# Transforms  (src/Fields.jl:1515-1542, 1589-1706)
# ---------------------------------------------------------------------------

def _fft_axes(fold: int):
    """Array axes of an rfftn over (..., z, y, x) whose last entry (the one
    rfftn halves) is the folded spatial axis."""
    axes = [a for a in (-3, -2, -1) if a != -1 - fold]
    return tuple(axes + [-1 - fold])


def to_spectral(phys: np.ndarray, workers: int | None = None, fold: int = 0) -> np.ndarray:
    """Unnormalised forward rfft over the last three axes, halving the
    spatial axis ``fold`` (x by default)."""
    w = _workers() if workers is None else workers
    axes = _fft_axes(fold)
    if _sfft is not None:
        return _sfft.rfftn(phys, axes=axes, workers=w)
    return np.fft.rfftn(phys, axes=axes)


def to_physical(spec: np.ndarray, grid: Grid3D, workers: int | None = None) -> np.ndarray:
    """Inverse c2r carrying 1/N: full inverses along the other two axes, then
    c2r along the folded axis, whose self-conjugate entries have their
    imaginary parts discarded, as FFTW's c2r does."""
    w = _workers() if workers is None else workers
    axes = _fft_axes(grid.fold)
    s = [grid.physical_shape[a] for a in axes]
    if _sfft is not None:
        return _sfft.irfftn(spec, s=s, axes=axes, workers=w)
    return np.fft.irfftn(spec, s=s, axes=axes)


def direct_dft_forward(phys: np.ndarray) -> np.ndarray:
    """Brute-force O(P*M) forward transform f^(k) = sum_x f(x) e^{-i k.x}
    (``docs/FORMALIZATION.md:86-100``), used only to pin ``to_spectral`` on
    tiny grids -- the reference's own DirectDFT idea
    (``test/transfer_spectral.jl:45-51``)."""
    Nz, Ny, Nx = phys.shape
    nh = Nx // 2 + 1
    ex = np.exp(-2j * np.pi * np.outer(np.arange(nh), np.arange(Nx)) / Nx)
    ey = np.exp(-2j * np.pi * np.outer(np.arange(Ny), np.arange(Ny)) / Ny)
    ez = np.exp(-2j * np.pi * np.outer(np.arange(Nz), np.arange(Nz)) / Nz)
    out = np.zeros((Nz, Ny, nh), dtype=np.complex128)
    for kz in range(Nz):
        for ky in range(Ny):
            for kx in range(nh):
                w = ez[kz][:, None, None] * ey[ky][None, :, None] * ex[kx][None, None, :]
                out[kz, ky, kx] = np.sum(phys * w)
    return out


def direct_dft_inverse(spec: np.ndarray, grid: Grid3D) -> np.ndarray:
    """Brute-force inverse of a half spectrum: f(x) = (1/N) Re sum over the
    Hermitian completion, with the self-conjugate entries taken real on the
    folded axis after the full-axis sums."""
    Nx, Ny, Nz = grid.n
    nh = grid.nh
    xs, ys, zs = np.arange(Nx), np.arange(Ny), np.arange(Nz)
    # full-axis inverse sums over kz and ky first (complex), then c2r on x
    ey = np.exp(2j * np.pi * np.outer(np.arange(Ny), ys) / Ny)   # [ky, y]
    ez = np.exp(2j * np.pi * np.outer(np.arange(Nz), zs) / Nz)   # [kz, z]
    g = np.einsum("kjx,kz,jy->zyx", spec, ez, ey)                # (Nz, Ny, nh)
    out = np.zeros((Nz, Ny, Nx))
    for kx in range(nh):
        selfconj = kx == 0 or (Nx % 2 == 0 and kx == Nx // 2)
        phase = np.exp(2j * np.pi * kx * xs / Nx)
        if selfconj:
            out += (g[:, :, kx].real[:, :, None] * phase.real[None, None, :])
        else:
            out += 2.0 * np.real(g[:, :, kx][:, :, None] * phase[None, None, :])
    return out / (Nx * Ny * Nz)


# ---------------------------------------------------------------------------
# Spectral operators  (src/DifferentialOperators.jl)
# ---------------------------------------------------------------------------

def gradient(s: np.ndarray, m: Multipliers) -> np.ndarray:
    """Scalar -> vector, ``_scalar_gradient_kernel!`` (``:357-368``):
    ``(i*a_j) * s`` with the factor formed first."""
    return np.stack([(1j * m.a1()) * s, (1j * m.a2()) * s, (1j * m.a3()) * s])


def symmetric_gradient(u: np.ndarray, m: Multipliers) -> np.ndarray:
    """Vector -> symmetric tensor, ``_symmetric_gradient_kernel!``
    (``src/DifferentialOperators.jl:666-699``, entry ``:1289-1300``):
    ``S_ii = (i*a_i) * u_i`` and ``S_ij = S_ji = (i/2) * (a_j u_i + a_i u_j)``
    with ``i/2`` the complex ``(0, 1/2)``.  Returned as ``(3, 3, ...)`` with
    ``[i, j] = S_ij`` (symmetric, so also the ``[j, i]`` view of the stored
    blocks ``i + 3 j``)."""
    a = (m.a1(), m.a2(), m.a3())
    half_i = complex(0.0, 0.5)
    S = np.empty((3, 3) + u.shape[1:], dtype=np.complex128)
    for i in range(3):
        S[i, i] = (1j * a[i]) * u[i]
        for j in range(i + 1, 3):
            S[i, j] = half_i * (a[j] * u[i] + a[i] * u[j])
            S[j, i] = S[i, j]
    return S


def divergence(u: np.ndarray, m: Multipliers) -> np.ndarray:
    """Vector -> scalar, ``_apply_spectral_divergence!`` (``:120-139``):
    ``i * ((a1 u1 + a2 u2) + a3 u3)``."""
    return 1j * ((m.a1() * u[0] + m.a2() * u[1]) + m.a3() * u[2])


def curl(u: np.ndarray, m: Multipliers) -> np.ndarray:
    """``_curl_kernel!`` (``:502-523``): ``i*(a2 u3 - a3 u2, a3 u1 - a1 u3,
    a1 u2 - a2 u1)``."""
    a1, a2, a3 = m.a1(), m.a2(), m.a3()
    return np.stack([1j * (a2 * u[2] - a3 * u[1]),
                     1j * (a3 * u[0] - a1 * u[2]),
                     1j * (a1 * u[1] - a2 * u[0])])


def _deriv_sq(m: Multipliers):
    a1, a2, a3 = m.a1(), m.a2(), m.a3()
    return (a1 * a1 + a2 * a2) + a3 * a3


def inverse_curl(w: np.ndarray, m: Multipliers) -> np.ndarray:
    """``_inverse_curl_kernel!`` (``:580-605``): ``i a x w / |a|^2`` with the
    derivative-family |a|^2 and zero where it vanishes."""
    sq = _deriv_sq(m)
    with np.errstate(divide="ignore"):
        rec = np.where(sq == 0.0, 0.0, 1.0 / np.where(sq == 0.0, 1.0, sq))
    c = curl(w, m)
    return c * rec[None]


def laplacian(f: np.ndarray, m: Multipliers) -> np.ndarray:
    """``_apply_spectral_laplacian!`` (``:147-150``): ``-|k|^2 * f``."""
    return (-m.ksq()) * f


def inverse_laplacian(f: np.ndarray, m: Multipliers) -> np.ndarray:
    """``_apply_spectral_inverse_laplacian!`` (``:159-164``):
    ``(-f) * (|k|^2 == 0 ? 0 : inv(|k|^2))``."""
    ks = m.ksq()
    inv = np.where(ks == 0.0, 0.0, 1.0 / np.where(ks == 0.0, 1.0, ks))
    return (-f) * inv


def band_mask(m: Multipliers, k_lo: float, k_hi: float) -> np.ndarray:
    """``spectral_band``/``spectral_cutoff`` (``src/Convolution.jl:1832-1882``)
    evaluated through the radial pass (``:743-787``):
    ``lo^2 <= (k1^2 + k2^2) + k3^2 <= hi^2`` with lo^2 = T(lo)^2, hi^2 =
    T(hi)^2 in FP64.  Returns a float64 0/1 array over (Nz, Ny, Nh)."""
    if not (math.isfinite(k_lo) and math.isfinite(k_hi)) or not (0 <= k_lo <= k_hi):
        raise ValueError("band edges must satisfy 0 <= low <= high")
    lo2 = float(k_lo) ** 2
    hi2 = float(k_hi) ** 2
    ks = m.ksq()
    return ((lo2 <= ks) & (ks <= hi2)).astype(np.float64)


def band_mask_float32(grid: Grid3D, k_lo: float, k_hi: float) -> np.ndarray:
    """The band mask as the reference's Float32 path evaluates it: per-axis
    tables in Float32 (``_axis_wavenumber_values``, ``src/Fields.jl:2006-2026``
    with T = Float32: ``unit = 2*T(pi)/L``, ``k = unit*n``, ``k^2``), ``|k|^2 =
    (k1^2 + k2^2) + k3^2`` in Float32 (``src/Fields.jl:2105-2109``) and edges
    ``T(lo)^2``, ``T(hi)^2`` (``src/Convolution.jl:1832-1850``).  The library's
    FP32 form keeps the FP64 predicate of ``band_mask`` (mode handling in
    FP64); this restatement shows where the two select different boundary
    modes.  Returns a bool array over (Nz, Ny, Nh)."""
    T = np.float32

    def sq(n, L, half):
        unit = T(2) * T(math.pi) / T(L)
        stored = n // 2 + 1 if half else n
        i = np.arange(stored)
        freq = i if half else np.where(i < (n + 1) // 2, i, i - n)
        k = unit * freq.astype(T)
        return k * k

    s1 = sq(grid.n[0], grid.L[0], grid.fold == 0)
    s2 = sq(grid.n[1], grid.L[1], grid.fold == 1)
    s3 = sq(grid.n[2], grid.L[2], grid.fold == 2)
    q = (s1[None, None, :] + s2[None, :, None]) + s3[:, None, None]
    return (T(k_lo) ** 2 <= q) & (q <= T(k_hi) ** 2)


def mask_digest(m: Multipliers, kc: float) -> tuple[int, str]:
    """Kept-mode count and SHA-256 of the bit-packed 2/3 mask of
    ``band_mask(m, 0, kc)`` in storage order (kx fastest), computed one kz
    plane at a time so full-size grids fit in memory (every plane holds a
    multiple of 8 modes for the grids this is used on, so the packed planes
    concatenate to the packed whole)."""
    import hashlib
    h = hashlib.sha256()
    kept = 0
    hi2 = float(kc) ** 2
    plane = m.s1[None, :] + m.s2[:, None]          # (Ny, Nh): k1^2 + k2^2
    for kz in range(m.grid.n[2]):
        q = plane + m.s3[kz]
        bits = (0.0 <= q) & (q <= hi2)
        kept += int(bits.sum())
        h.update(np.packbits(bits.ravel()).tobytes())
    return kept, h.hexdigest()


def dealias_cutoff(n: int, L: float = TWO_PI) -> float:
    """The consumer's 2/3 cutoff ``(2/3)*(N/2)*(2pi/L)`` evaluated in Float64
    left to right (``examples/HIST_CPU.jl:357``)."""
    return ((2.0 / 3.0) * (n / 2.0)) * (TWO_PI / L)


def convolve_band(f: np.ndarray, m: Multipliers, k_lo: float, k_hi: float,
                  alpha: float = 1.0, beta: float = 0.0, dst: np.ndarray | None = None):
    """Radial ``convolve!`` (``src/Convolution.jl:1035-1065`` ->
    ``_radial_convolution_pass!`` ``:756-787``): ``alpha*sigma(|k|^2)*f
    (+beta*dst)``."""
    g = band_mask(m, k_lo, k_hi)
    out = (alpha * g) * f
    if beta != 0.0:
        out = out + beta * dst
    return out


def helmholtz_projection(u: np.ndarray, m: Multipliers, kernel: np.ndarray | None = None,
                         alpha: float = 1.0, beta: float = 0.0, dst: np.ndarray | None = None):
    """``_helmholtz_projection_kernel!`` (``src/DifferentialOperators.jl:191-230``):
    ``L = (a.u)/(a.a == 0 ? 1 : a.a)``; ``u_i <- alpha*g*(u_i - a_i L)
    (+beta d_i)`` with the derivative family in numerator and denominator."""
    a1, a2, a3 = m.a1(), m.a2(), m.a3()
    sq = _deriv_sq(m)
    lon = (a1 * u[0] + a2 * u[1] + a3 * u[2]) / np.where(sq == 0.0, 1.0, sq)
    scale = alpha if kernel is None else alpha * kernel
    out = np.stack([scale * (u[0] - a1 * lon), scale * (u[1] - a2 * lon), scale * (u[2] - a3 * lon)])
    if beta != 0.0:
        out = out + beta * dst
    return out


def mean_free(f: np.ndarray) -> np.ndarray:
    """``mean_free_projection!`` (``src/Fields.jl:3124-3128``): zero the k=0
    coefficient of every component."""
    out = f.copy()
    out[..., 0, 0, 0] = 0.0
    return out


# ---------------------------------------------------------------------------
# Physical products  (src/FieldAlgebra.jl)
# ---------------------------------------------------------------------------

def cross(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """``_cross_formula`` via ``cross!`` (``src/FieldAlgebra.jl:1088-1092``,
    ``src/FiberSections.jl:641-652``): ``(u x v)_i = u_j v_k - u_k v_j``."""
    return np.stack([u[1] * v[2] - u[2] * v[1],
                     u[2] * v[0] - u[0] * v[2],
                     u[0] * v[1] - u[1] * v[0]])


def multiply(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Scalar ``multiply!`` (``src/FieldAlgebra.jl:703-724, 778``)."""
    return a * b


# ---------------------------------------------------------------------------
# Reductions  (src/Reductions.jl)
# ---------------------------------------------------------------------------

def inner_spectral(f: np.ndarray, g: np.ndarray, grid: Grid3D) -> float:
    """Parseval ``inner`` (``src/Reductions.jl:883-925``):
    ``Re(2*sum f conj(g) - sum_{k_f=0} - sum_{k_f=Nyq if N_f even}) / N^2``
    over the folded axis f."""
    ax = -1 - grid.fold
    tot = np.sum(f * np.conj(g))
    corr = np.sum(np.take(f, 0, axis=ax) * np.conj(np.take(g, 0, axis=ax)))
    if grid.n[grid.fold] % 2 == 0:
        corr = corr + np.sum(np.take(f, -1, axis=ax) * np.conj(np.take(g, -1, axis=ax)))
    return float(np.real(2.0 * tot - corr)) / float(grid.npoints) ** 2


def inner_physical(f: np.ndarray, g: np.ndarray, grid: Grid3D) -> float:
    """Physical ``inner`` (``src/Reductions.jl:838-853``): cell average of
    the pointwise contraction."""
    return float(np.sum(f * g)) / grid.npoints


def parseval_weight(grid: Grid3D) -> np.ndarray:
    """``_parseval_weight`` (``src/Reductions.jl:956-965``) along the folded
    axis, shaped to broadcast over the spectral array."""
    w = np.full(grid.nh, 2.0)
    w[0] = 1.0
    if grid.n[grid.fold] % 2 == 0:
        w[-1] = 1.0
    return grid.fold_view(w)


# ---------------------------------------------------------------------------
# Pipelines  (examples/HIST_CPU.jl:229-247)
# ---------------------------------------------------------------------------

def uxw(u_hat: np.ndarray, grid: Grid3D, kc: float | None = None, m: Multipliers | None = None,
        workers: int | None = None):
    """Dealiased rotational nonlinear term N^ = Q M P F(F^-1 u^ x F^-1 curl u^),
    following ``to_velocity_vorticity!`` + ``rhs!`` (``examples/HIST_CPU.jl:229-247``).

    The 2/3 mask M (``spectral_cutoff`` at kc) is applied to the input as the
    first step (the consumer keeps u^ dealiased by construction,
    ``examples/HIST_CPU.jl:362-363``, so this is the identity there) and after
    the projection.  E = 1/2 <u,u> and Z = 1/2 <w,w> of the truncated fields
    (``examples/HIST_CPU.jl:170-171``) through the Parseval inner.
    Returns ``(N_hat, E, Z)``."""
    m = m or Multipliers.build(grid)
    if kc is None:
        kc = dealias_cutoff(grid.n[0], grid.L[0])
    mask = band_mask(m, 0.0, kc)
    ut = mask[None] * u_hat
    wt = curl(ut, m)
    E = 0.5 * inner_spectral(ut, ut, grid)
    Z = 0.5 * inner_spectral(wt, wt, grid)
    u = to_physical(ut, grid, workers)
    w = to_physical(wt, grid, workers)
    n = cross(u, w)
    del u, w
    nt = to_spectral(n, workers, grid.fold)
    del n
    nh = helmholtz_projection(nt, m)
    nh = mask[None] * nh
    nh[:, 0, 0, 0] = 0.0
    return nh, E, Z


def dealiased_product(a_hat: np.ndarray, b_hat: np.ndarray, grid: Grid3D, kc: float | None = None,
                      m: Multipliers | None = None, workers: int | None = None) -> np.ndarray:
    """p^ = M . F(F^-1(M a^) * F^-1(M b^)): 2/3 truncation of both inputs,
    physical pointwise product, transform back, truncate (the consumer's
    composition of ``convolve!``/``to_physical!``/``multiply!``/``to_spectral!``)."""
    m = m or Multipliers.build(grid)
    if kc is None:
        kc = dealias_cutoff(grid.n[0], grid.L[0])
    mask = band_mask(m, 0.0, kc)
    a = to_physical(mask * a_hat, grid, workers)
    b = to_physical(mask * b_hat, grid, workers)
    return mask * to_spectral(a * b, workers, grid.fold)


# ---------------------------------------------------------------------------
# Exponential kernel  (src/Convolution.jl:1197-1795)
# ---------------------------------------------------------------------------

def exponential_factor(m: Multipliers, coefficient: float, offset: np.ndarray | None = None,
                       offset_scale: float = 1.0) -> np.ndarray:
    """Per-mode transfer function ``exp(c |k|^2 [+ s w])`` with ``|k|^2 =
    (k1^2 + k2^2) + k3^2`` from the squared family (``_exponential_exponent``,
    ``src/Convolution.jl:1289-1294``; kernel ``:1326-1346``).  A complex offset
    gives a complex exponential, as the reference's broadcast does."""
    e = coefficient * m.ksq()
    if offset is None:
        return np.exp(e)
    return np.exp(e + offset_scale * offset)


def exponential_convolve(f: np.ndarray, m: Multipliers, coefficient: float, alpha: float = 1.0,
                         beta: float = 0.0, dst: np.ndarray | None = None,
                         offset: np.ndarray | None = None, offset_scale: float = 1.0) -> np.ndarray:
    """``exponential_convolve!`` single-operand form (``src/Convolution.jl:1750``;
    ``_exponential_convolution_kernel!`` ``:1326-1346``, scalar broadcast
    ``:1211-1232``): ``(alpha exp(...)) f``, plus ``beta dst`` only when beta != 0."""
    fac = alpha * exponential_factor(m, coefficient, offset, offset_scale)
    out = fac * f
    if beta != 0.0:
        out = out + beta * dst
    return out


def exponential_combine(alpha: float, a: np.ndarray, beta: float, b: np.ndarray, m: Multipliers,
                        coefficient: float, offset: np.ndarray | None = None,
                        offset_scale: float = 1.0) -> np.ndarray:
    """``exponential_convolve!(dst, alpha, a, beta, b, m; ...)`` (the two-operand
    form, ``_exponential_combination_kernel!`` ``src/Convolution.jl:1349-1362``,
    scalar broadcast ``:1237-1255``): ``exp(...) (alpha a + beta b)``."""
    return exponential_factor(m, coefficient, offset, offset_scale) * (alpha * a + beta * b)


# ---------------------------------------------------------------------------
# Radial (shell-binned) spectrum  (src/Reductions.jl:1309-1455)
# ---------------------------------------------------------------------------

def radial_shells(m: Multipliers) -> tuple[int, float]:
    """Shell count and width of ``_radial_spectrum_bin`` (``src/Reductions.jl:1334-1338``):
    ``dk = 2 pi / max L``; ``round(sqrt(max k1^2 + max k2^2 + max k3^2) / dk) + 1``
    shells (``round`` ties to even)."""
    dk = TWO_PI / max(m.grid.L)
    corner = float(m.s1.max()) + float(m.s2.max()) + float(m.s3.max())
    return int(np.rint(math.sqrt(corner) / dk)) + 1, dk


def radial_spectrum(f: np.ndarray, m: Multipliers, g: np.ndarray | None = None):
    """``radial_spectrum(f[, g], m)`` (``_spectral_shell_bin_kernel!``
    ``src/Reductions.jl:1309-1327``, ``_radial_spectrum_bin`` ``:1330-1358``):
    shell ``round(|k| / dk)`` (ties to even), Parseval weight 1 on the kx = 0 and
    even-Nx Nyquist planes and 2 elsewhere, density ``sum_c |f_c|^2`` (or
    ``Re(f_c conj(g_c))``), each shell divided by ``(Nx Ny Nz)^2``.  Returns
    ``(shells, spectrum)``."""
    grid = m.grid
    nsh, dk = radial_shells(m)
    k2 = (m.s1[None, None, :] + m.s2[None, :, None]) + m.s3[:, None, None]
    shell = np.rint(np.sqrt(k2) / dk).astype(np.int64)
    ff = f.reshape((-1,) + grid.spectral_shape)
    if g is None:
        dens = np.sum(np.abs(ff) ** 2, axis=0)
    else:
        gg = g.reshape((-1,) + grid.spectral_shape)
        dens = np.sum((ff * np.conj(gg)).real, axis=0)
    w = parseval_weight(grid)
    spec = np.bincount(shell.ravel(), weights=(w * dens).ravel(), minlength=nsh)[:nsh]
    spec = spec / float(grid.npoints) ** 2
    return np.arange(nsh) * dk, spec


# ---------------------------------------------------------------------------
# Reality condition on the self-conjugate planes  (src/Fields.jl:2858-3094)
# ---------------------------------------------------------------------------

def _reality_partner(plane: np.ndarray) -> np.ndarray:
    """The entry at (-k_a, -k_b) for every (k_a, k_b) of a (..., Na, Nb) plane."""
    na, nb = plane.shape[-2], plane.shape[-1]
    ja = (-np.arange(na)) % na
    jb = (-np.arange(nb)) % nb
    return plane[..., ja, :][..., :, jb]


def _self_conjugate_planes(grid: Grid3D):
    """Indices along the folded axis of its self-conjugate planes: 0 and, for
    an even folded length, the Nyquist plane."""
    return [0] + ([grid.nh - 1] if grid.n[grid.fold] % 2 == 0 else [])


def reality_residual(spec: np.ndarray, grid: Grid3D) -> float:
    """``reality_residual`` (``src/Fields.jl:2858-2869, 2918-2928``): the largest
    ``|f(k) - conj(f(-k))|`` over the k_f = 0 plane of the folded axis f and,
    for even N_f, its Nyquist plane, every component."""
    ax = spec.ndim - 1 - grid.fold
    r = 0.0
    for p in _self_conjugate_planes(grid):
        pl = np.take(spec, p, axis=ax)
        r = max(r, float(np.max(np.abs(pl - np.conj(_reality_partner(pl))))))
    return r


def reality_projection(spec: np.ndarray, grid: Grid3D) -> np.ndarray:
    """``reality_projection!`` (``src/Fields.jl:2940-3094``; ``_project_pair!``
    ``:2949-2958``): every entry of a self-conjugate plane becomes
    ``(f(k) + conj(f(-k))) / 2`` (its partner the conjugate of that, a
    self-paired corner its real part); all other entries are untouched."""
    out = spec.copy()
    ax = spec.ndim - 1 - grid.fold
    for p in _self_conjugate_planes(grid):
        pl = np.take(spec, p, axis=ax)
        idx = [slice(None)] * spec.ndim
        idx[ax] = p
        out[tuple(idx)] = (pl + np.conj(_reality_partner(pl))) / 2.0
    return out


# ---------------------------------------------------------------------------
# Synthetic inputs (seeded; the workload generator)
# ---------------------------------------------------------------------------

def reality_pair_mean(spec: np.ndarray, grid: Grid3D) -> np.ndarray:
    """Project the self-conjugate planes of the folded axis onto the reality
    condition with the pair mean of ``_project_pair!`` (``src/Fields.jl:2949-2958``):
    both members of a pair (k, -k) become (x + conj(y))/2 and its conjugate."""
    return reality_projection(spec, grid)


def synthetic_field(grid: Grid3D, seed: int, vector: bool = False, m: Multipliers | None = None) -> np.ndarray:
    """Seeded spectral field with amplitude |k|^(-5/3) exp(-(|k|/20)^2),
    uniform phases in [0, 2pi) drawn per stored mode in column-major order
    (PCG64(seed)), k=0 set to zero, Hermitian on the self-conjugate planes.
    Vector fields are projected solenoidal and the whole field is scaled by
    one factor so that <u,u> = 3 (a scalar field: <f,f> = 1)."""
    m = m or Multipliers.build(grid)
    rng = np.random.Generator(np.random.PCG64(seed))
    kmag = np.sqrt(m.ksq())
    with np.errstate(divide="ignore"):
        amp = np.where(kmag > 0, kmag ** (-5.0 / 3.0) * np.exp(-(kmag / 20.0) ** 2), 0.0)
    ncomp = 3 if vector else 1
    comps = []
    for _ in range(ncomp):
        ph = rng.uniform(0.0, 2.0 * np.pi, size=grid.nmodes).reshape(grid.spectral_shape)
        comps.append(amp * np.exp(1j * ph))
    f = np.stack(comps) if vector else comps[0]
    f = reality_pair_mean(f, grid)
    if vector:
        f = helmholtz_projection(f, m)
        target = 3.0
    else:
        target = 1.0
    f[..., 0, 0, 0] = 0.0
    e = inner_spectral(f, f, grid)
    if e > 0:
        f = f * math.sqrt(target / e)
    return f


def rel_l2(x: np.ndarray, ref: np.ndarray) -> float:
    """``||x - ref||_2 / ||ref||_2`` over all stored entries."""
    num = np.linalg.norm((np.asarray(x) - np.asarray(ref)).ravel())
    den = np.linalg.norm(np.asarray(ref).ravel())
    return float(num / den) if den > 0 else float(num)
