"""Closed-form fields of the u x omega tests, as stored modes (c, kz, ky, kx)
of the unnormalised forward rfft (x folded; kz, ky signed).  Pinned against
the oracle in tests/test_oracle_next.py."""

import numpy as np


def shear_modes(P):
    """u = (sin z, sin x, 0) and P(u x w) = (0, -sin z cos x, 0) as stored
    modes (c, kz, ky, kx) of the unnormalised forward rfft (x folded)."""
    u = {(0, 1, 0, 0): -0.5j * P, (0, -1, 0, 0): 0.5j * P, (1, 0, 0, 1): -0.5j * P}
    n = {(1, 1, 0, 1): 0.25j * P, (1, -1, 0, 1): -0.25j * P}
    return u, n


def abc_modes(P, A, B, Cc):
    """ABC flow u = (A sin z + C cos y, B sin x + A cos z, C sin y + B cos x)."""
    h = 0.5 * P
    return {(0, 1, 0, 0): -1j * A * h, (0, -1, 0, 0): 1j * A * h,
            (0, 0, 1, 0): Cc * h, (0, 0, -1, 0): Cc * h,
            (1, 0, 0, 1): -1j * B * h, (1, 1, 0, 0): A * h, (1, -1, 0, 0): A * h,
            (2, 0, 1, 0): -1j * Cc * h, (2, 0, -1, 0): 1j * Cc * h, (2, 0, 0, 1): B * h}


def dense(modes, n):
    f = np.zeros((3, n[2], n[1], n[0] // 2 + 1), dtype=np.complex128)
    for (c, kz, ky, kx), v in modes.items():
        f[c, kz % n[2], ky % n[1], kx] = v
    return f
