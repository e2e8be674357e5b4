"""Seeded synthetic spectral fields: the inputs of the benchmark workloads.

A field has random Fourier modes of amplitude |k|^(-5/3) exp(-(|k|/20)^2),
uniform phases in [0, 2pi) drawn from PCG64(seed) per stored mode in
column-major order, the k = 0 mode set to zero, and the self-conjugate planes
of the folded axis made Hermitian with the pair mean of the reference's
reality projection (src/Fields.jl:2949-2958).  Vector fields are projected
solenoidal (src/DifferentialOperators.jl:191-230) and the whole field is
scaled by one factor so that <u,u> = 3 (a scalar field: <f,f> = 1).

This is host-side input generation (NumPy), not part of the GPU path.
"""

from __future__ import annotations

import math

import numpy as np


def _axis(n: int, L: float, half: bool):
    unit = (2.0 * math.pi) / float(L)
    stored = n // 2 + 1 if half else n
    i = np.arange(stored)
    f = i if half else np.where(i < (n + 1) // 2, i, i - n)
    k = unit * f.astype(np.float64)
    d = k.copy()
    if n % 2 == 0:
        d[stored - 1 if half else n // 2] = 0.0
    return d, k * k


def _tables(n, L):
    k1, s1 = _axis(n[0], L[0], True)
    k2, s2 = _axis(n[1], L[1], False)
    k3, s3 = _axis(n[2], L[2], False)
    return (k1, k2, k3), (s1, s2, s3)


def _inner(f, nx, npts):
    tot = np.sum(f.real * f.real + f.imag * f.imag)
    corr = np.sum(f[..., 0].real ** 2 + f[..., 0].imag ** 2)
    if nx % 2 == 0:
        corr += np.sum(f[..., -1].real ** 2 + f[..., -1].imag ** 2)
    return float(2.0 * tot - corr) / float(npts) ** 2


def synthetic_field(n, seed: int, vector: bool = False, L=(2 * math.pi,) * 3,
                    dtype=np.complex128) -> np.ndarray:
    """Return a spectral field of shape ([3,] Nz, Ny, Nx//2+1)."""
    Nx, Ny, Nz = n
    nh = Nx // 2 + 1
    (k1, k2, k3), (s1, s2, s3) = _tables(n, L)
    ksq = (s1[None, None, :] + s2[None, :, None]) + s3[:, None, None]
    kmag = np.sqrt(ksq)
    with np.errstate(divide="ignore"):
        amp = np.where(kmag > 0, kmag ** (-5.0 / 3.0) * np.exp(-(kmag / 20.0) ** 2), 0.0)
    del ksq, kmag
    rng = np.random.Generator(np.random.PCG64(seed))
    ncomp = 3 if vector else 1
    M = nh * Ny * Nz
    f = np.empty((ncomp, Nz, Ny, nh), dtype=np.complex128)
    for c in range(ncomp):
        ph = rng.uniform(0.0, 2.0 * np.pi, size=M).reshape(Nz, Ny, nh)
        f[c] = amp * np.exp(1j * ph)
        del ph
    del amp
    jy = (-np.arange(Ny)) % Ny
    jz = (-np.arange(Nz)) % Nz
    planes = [0] + ([nh - 1] if Nx % 2 == 0 else [])
    for p in planes:
        pl = f[:, :, :, p]
        partner = pl[:, jz, :][:, :, jy]
        f[:, :, :, p] = (pl + np.conj(partner)) / 2.0
    if vector:
        a1, a2, a3 = k1[None, None, :], k2[None, :, None], k3[:, None, None]
        sq = (a1 * a1 + a2 * a2) + a3 * a3
        den = np.where(sq == 0.0, 1.0, sq)
        lon = (a1 * f[0] + a2 * f[1] + a3 * f[2]) / den
        f[0] = f[0] - a1 * lon
        f[1] = f[1] - a2 * lon
        f[2] = f[2] - a3 * lon
        del lon
        target = 3.0
    else:
        target = 1.0
    f[:, 0, 0, 0] = 0.0
    e = _inner(f, Nx, Nx * Ny * Nz)
    if e > 0:
        f *= math.sqrt(target / e)
    out = f if vector else f[0]
    return out.astype(dtype, copy=False)
