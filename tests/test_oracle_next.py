"""Pins the oracle's exponential kernel, radial spectrum and reality projector
(CPU only) against closed forms, the reference's own known answers
(test/convolution.jl:282-330, test/fields.jl:324-470) and brute-force loops
written independently of the vectorised oracle."""

import math

import numpy as np
import pytest

from closed_forms import abc_modes, dense, shear_modes
from oracle import gbg_oracle as O

TWO_PI = 2 * math.pi


def phys_field(grid, fn):
    Nx, Ny, Nz = grid.n
    x = np.arange(Nx) * grid.L[0] / Nx
    y = np.arange(Ny) * grid.L[1] / Ny
    z = np.arange(Nz) * grid.L[2] / Nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return fn(X, Y, Z)


def rand_spec(grid, rng, ncomp=None):
    shape = grid.spectral_shape if ncomp is None else (ncomp,) + grid.spectral_shape
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# --- exponential kernel: test/convolution.jl:282-330 -------------------------

def test_exponential_single_mode_eigenvector():
    # exp(c|k|^2) on cos(x) is exp(c) cos(x); on cos(2y) sin(3z) it is exp(13 c)
    g = O.Grid3D((8, 8, 8))
    m = O.Multipliers.build(g)
    c = -0.037
    f = O.to_spectral(phys_field(g, lambda X, Y, Z: np.cos(X)))
    out = O.to_physical(O.exponential_convolve(f, m, c), g)
    np.testing.assert_allclose(out, math.exp(c) * phys_field(g, lambda X, Y, Z: np.cos(X)), atol=1e-14)
    f2 = O.to_spectral(phys_field(g, lambda X, Y, Z: np.cos(2 * Y) * np.sin(3 * Z)))
    out2 = O.to_physical(O.exponential_convolve(f2, m, c), g)
    ref2 = math.exp(13 * c) * phys_field(g, lambda X, Y, Z: np.cos(2 * Y) * np.sin(3 * Z))
    np.testing.assert_allclose(out2, ref2, atol=1e-14)


def test_exponential_anisotropic_box_uses_true_k2():
    # on L = (2pi, 4pi, 3) the mode cos(2 * 2pi y / 4pi) has |k|^2 = 1; the x
    # Nyquist mode of an even axis keeps its true k^2 (not the zeroed derivative)
    g = O.Grid3D((8, 6, 4), (TWO_PI, 4 * math.pi, 3.0))
    m = O.Multipliers.build(g)
    c = -0.2
    f = O.to_spectral(phys_field(g, lambda X, Y, Z: np.cos(Y)))
    np.testing.assert_allclose(O.to_physical(O.exponential_convolve(f, m, c), g),
                               math.exp(c) * phys_field(g, lambda X, Y, Z: np.cos(Y)), atol=1e-14)
    fn = O.to_spectral(phys_field(g, lambda X, Y, Z: np.cos(4 * X)))
    np.testing.assert_allclose(O.to_physical(O.exponential_convolve(fn, m, c), g),
                               math.exp(16 * c) * phys_field(g, lambda X, Y, Z: np.cos(4 * X)), atol=1e-14)


def test_exponential_semigroup_and_accumulate():
    g = O.Grid3D((8, 7, 6))
    m = O.Multipliers.build(g)
    rng = np.random.default_rng(1)
    f = rand_spec(g, rng, 3)
    d0 = rand_spec(g, rng, 3)
    a = O.exponential_convolve(O.exponential_convolve(f, m, -0.01), m, -0.02)
    np.testing.assert_allclose(a, O.exponential_convolve(f, m, -0.03), rtol=1e-13)
    # alpha/beta accumulate against the hand formula of test/convolution.jl:304-306
    k2 = m.ksq()
    got = O.exponential_convolve(f, m, -0.037, 0.5, -1.5, d0)
    np.testing.assert_allclose(got, 0.5 * (np.exp(-0.037 * k2) * f) + (-1.5) * d0, rtol=1e-14)
    # beta == 0 overwrites: a NaN destination does not leak
    nan = np.full_like(d0, np.nan)
    assert np.all(np.isfinite(O.exponential_convolve(f, m, -0.01, 1.0, 0.0, nan)))


def test_exponential_offset_and_two_operand():
    g = O.Grid3D((8, 8, 6))
    m = O.Multipliers.build(g)
    rng = np.random.default_rng(2)
    f = rand_spec(g, rng)
    b = rand_spec(g, rng)
    c, s = -0.037, 0.011
    # offset w = |k|^2 as a (complex-stored) field: exp(c k2 + s k2) = exp((c + s) k2)
    w = m.ksq().astype(np.complex128)
    np.testing.assert_allclose(O.exponential_convolve(f, m, c, offset=w, offset_scale=s),
                               O.exponential_convolve(f, m, c + s), rtol=1e-13)
    # a complex offset: exp(i s phi) rotates each mode by s phi
    phi = rng.uniform(-1, 1, g.spectral_shape)
    got = O.exponential_convolve(f, m, 0.0, offset=1j * phi, offset_scale=s)
    np.testing.assert_allclose(got, f * (np.cos(s * phi) + 1j * np.sin(s * phi)), rtol=1e-14)
    # the two-operand form == combine then convolve (test/convolution.jl:314-324)
    seq = O.exponential_convolve(0.3 * f + 0.7 * b, m, c)
    np.testing.assert_allclose(O.exponential_combine(0.3, f, 0.7, b, m, c), seq, rtol=1e-15)
    seq_w = O.exponential_convolve(0.3 * f + 0.7 * b, m, c, offset=w, offset_scale=s)
    np.testing.assert_allclose(O.exponential_combine(0.3, f, 0.7, b, m, c, w, s), seq_w, rtol=1e-15)


# --- radial spectrum: src/Reductions.jl:1309-1455 ----------------------------

def brute_radial_spectrum(f, grid, g=None):
    """Loop restatement: every stored mode, every component."""
    Nx, Ny, Nz = grid.n
    dk = TWO_PI / max(grid.L)

    def kk(i, n, L, half):
        j = i if (half or i < (n + 1) // 2) else i - n
        return (TWO_PI / L * j) ** 2

    corner = max(kk(i, Nx, grid.L[0], True) for i in range(grid.nh)) + \
        max(kk(i, Ny, grid.L[1], False) for i in range(Ny)) + \
        max(kk(i, Nz, grid.L[2], False) for i in range(Nz))
    nsh = round(math.sqrt(corner) / dk) + 1          # Python round: ties to even
    out = [0.0] * nsh
    ff = f.reshape((-1,) + grid.spectral_shape)
    gg = None if g is None else g.reshape((-1,) + grid.spectral_shape)
    for iz in range(Nz):
        for iy in range(Ny):
            for ix in range(grid.nh):
                k2 = (kk(ix, Nx, grid.L[0], True) + kk(iy, Ny, grid.L[1], False)) + kk(iz, Nz, grid.L[2], False)
                sh = round(math.sqrt(k2) / dk)
                wgt = 1.0 if (ix == 0 or (Nx % 2 == 0 and ix == grid.nh - 1)) else 2.0
                acc = 0.0
                for c in range(ff.shape[0]):
                    v = ff[c, iz, iy, ix]
                    acc += abs(v) ** 2 if gg is None else (v * np.conj(gg[c, iz, iy, ix])).real
                out[sh] += wgt * acc
    return np.array(out) / float(grid.npoints) ** 2


@pytest.mark.parametrize("n,L", [((8, 8, 8), None), ((7, 8, 6), None), ((8, 6, 5), (TWO_PI, 2.5, 7.0))])
def test_radial_spectrum_brute_force_and_parseval(n, L):
    g = O.Grid3D(n) if L is None else O.Grid3D(n, L)
    m = O.Multipliers.build(g)
    rng = np.random.default_rng(3)
    f = O.reality_projection(rand_spec(g, rng, 3), g)
    h = O.reality_projection(rand_spec(g, rng, 3), g)
    sh, spec = O.radial_spectrum(f, m)
    np.testing.assert_allclose(spec, brute_radial_spectrum(f, g), rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(sh, np.arange(len(spec)) * TWO_PI / max(g.L), rtol=1e-15)
    # every mode is in some shell: the spectrum refines the Parseval inner
    assert math.isclose(spec.sum(), O.inner_spectral(f, f, g), rel_tol=1e-12)
    _, co = O.radial_spectrum(f, m, h)
    np.testing.assert_allclose(co, brute_radial_spectrum(f, g, h), rtol=1e-12, atol=1e-15)
    _, co2 = O.radial_spectrum(h, m, f)
    np.testing.assert_allclose(co, co2, rtol=1e-13, atol=1e-16)
    assert math.isclose(co.sum(), O.inner_spectral(f, h, g), rel_tol=1e-11, abs_tol=1e-14)


def test_radial_spectrum_known_shells():
    # <cos 3x, cos 3x> = 1/2 lands in shell 3; cos 4y in shell 4; cos(3x)cos(4y)
    # (|k| = 5) carries <.,.> = 1/4 into shell 5
    g = O.Grid3D((16, 16, 8))
    m = O.Multipliers.build(g)
    f = O.to_spectral(phys_field(g, lambda X, Y, Z: np.cos(3 * X) + np.cos(4 * Y)
                                 + np.cos(3 * X) * np.cos(4 * Y)))
    _, spec = O.radial_spectrum(f, m)
    want = np.zeros_like(spec)
    want[3], want[4], want[5] = 0.5, 0.5, 0.25
    np.testing.assert_allclose(spec, want, atol=1e-14)


# --- reality condition: test/fields.jl:324-470 -------------------------------

def brute_reality_projection(spec, grid):
    out = spec.copy()
    Nx, Ny, Nz = grid.n
    planes = [0] + ([grid.nh - 1] if Nx % 2 == 0 else [])
    flat = spec.reshape((-1,) + grid.spectral_shape)
    o = out.reshape((-1,) + grid.spectral_shape)
    for c in range(flat.shape[0]):
        for p in planes:
            for iz in range(Nz):
                for iy in range(Ny):
                    pz, py = (-iz) % Nz, (-iy) % Ny
                    o[c, iz, iy, p] = (flat[c, iz, iy, p] + np.conj(flat[c, pz, py, p])) / 2
    return out


@pytest.mark.parametrize("n", [(8, 8, 8), (8, 7, 6), (7, 8, 8)])
def test_reality_projection_and_residual(n):
    g = O.Grid3D(n)
    rng = np.random.default_rng(4)
    for ncomp in (None, 3, 9):
        f = rand_spec(g, rng, ncomp)
        p = O.reality_projection(f, g)
        # the projector against the loop, exact; residual exactly 0; idempotent bitwise
        assert np.array_equal(p, brute_reality_projection(f, g))
        assert O.reality_residual(p, g) == 0.0
        assert np.array_equal(O.reality_projection(p, g), p)
        # off-plane entries untouched
        planes = [0] + ([g.nh - 1] if n[0] % 2 == 0 else [])
        off = [i for i in range(g.nh) if i not in planes]
        assert np.array_equal(p[..., off], f[..., off])
        # self-paired corners are real
        assert np.all(p[..., 0, 0, 0].imag == 0)
        # break one member of a pair by delta: the residual is |delta|
        q = p.copy()
        idx = (0,) * (0 if ncomp is None else 1)
        delta = 0.6 + 0.25j
        q[idx + (3, 3, 0)] += delta
        assert math.isclose(O.reality_residual(q, g), abs(delta), rel_tol=1e-12)
        # fix the partner: back to round-off
        q[idx + ((-3) % n[2], (-3) % n[1], 0)] += np.conj(delta)
        assert O.reality_residual(q, g) < 1e-12
        # a self-paired corner must be real: the residual is twice its imaginary part
        q = p.copy()
        q[idx + (0, 0, 0)] += 0.3j
        assert math.isclose(O.reality_residual(q, g), 0.6, rel_tol=1e-12)
        # linear
        h = rand_spec(g, rng, ncomp)
        np.testing.assert_allclose(O.reality_projection(2 * f - 3 * h, g),
                                   2 * O.reality_projection(f, g) - 3 * O.reality_projection(h, g),
                                   rtol=1e-14, atol=1e-14)


def test_reality_round_trip_is_the_projection():
    # the physical round trip projects too (up to transform round-off); a spectrum
    # from real data has residual at round-off level
    g = O.Grid3D((8, 6, 10))
    rng = np.random.default_rng(5)
    f = rand_spec(g, rng)
    rt = O.to_spectral(O.to_physical(f, g))
    np.testing.assert_allclose(rt, O.reality_projection(f, g), atol=1e-13)
    assert O.reality_residual(O.to_spectral(rng.standard_normal(g.physical_shape)), g) < 1e-12


def test_closed_forms_pinned_by_oracle():
    """The mode lists of the 1024^3 closed-form GPU tests (tests/closed_forms.py)
    are what the oracle gives on 8^3: the input spectra equal the transform of
    the physical fields, the shear pair's projected product and the ABC
    flow's vanishing product come out of the oracle's u x omega, and so do
    the closed-form energy and enstrophy."""
    n = (8, 8, 8)
    g = O.Grid3D(n)
    x = np.arange(8) * (O.TWO_PI / 8)
    Z3, Y3, X3 = np.meshgrid(x, x, x, indexing="ij")
    u = np.stack([np.sin(Z3), np.sin(X3), 0 * X3])
    us, ns = shear_modes(512.0)
    assert np.abs(O.to_spectral(u) - dense(us, n)).max() < 1e-12
    got, E, Z = O.uxw(dense(us, n), g, O.dealias_cutoff(1024))
    assert np.abs(got - dense(ns, n)).max() < 1e-12 and abs(E - 0.5) < 1e-15 and abs(Z - 0.5) < 1e-15
    A, B, Cc = 1.0, 0.7, 0.4
    abc = np.stack([A * np.sin(Z3) + Cc * np.cos(Y3), B * np.sin(X3) + A * np.cos(Z3),
                    Cc * np.sin(Y3) + B * np.cos(X3)])
    assert np.abs(O.to_spectral(abc) - dense(abc_modes(512.0, A, B, Cc), n)).max() < 1e-12
    got, E, Z = O.uxw(dense(abc_modes(512.0, A, B, Cc), n), g, O.dealias_cutoff(1024))
    assert np.abs(got).max() < 1e-12
    assert abs(E - (A * A + B * B + Cc * Cc) / 2) < 1e-14 and abs(Z - E) < 1e-14


def test_curl_formed_after_the_z_transform():
    """The fused u x omega schedule's split of w = curl u^: the z-inverse pass
    writes U1, U2, U3 and the kz terms D1 = Fz(i a3 u1), D2 = Fz(i a3 u2), and
    the y pass forms w1 = i a2 U3 - D2, w2 = D1 - i a1 U3, w3 = i (a1 U2 - a2 U1)
    (a1, a2 are constant along a z pencil, Fz is linear).  On a dense random
    spectrum of an anisotropic grid that equals the z transform of the
    oracle's curl, and a sign or factor slip in any of the three fails it."""
    n = (12, 10, 16)
    g = O.Grid3D(n)
    m = O.Multipliers.build(g)
    rng = np.random.default_rng(7)
    shape = (3,) + g.spectral_shape
    u = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    def fz(f):    # unnormalised inverse transform along z (axis 0 of a component)
        return np.fft.ifft(f, axis=0) * n[2]

    ref = np.stack([fz(c) for c in O.curl(u, m)])
    a1, a2, a3 = m.a1(), m.a2(), m.a3()
    U = [fz(c) for c in u]
    d1, d2 = fz(1j * a3 * u[0]), fz(1j * a3 * u[1])
    w = np.stack([1j * a2 * U[2] - d2, d1 - 1j * a1 * U[2], 1j * (a1 * U[1] - a2 * U[0])])
    scale = np.abs(ref).max()
    assert np.abs(w - ref).max() <= 1e-13 * scale
    # the same split with D2's sign flipped is caught
    bad = 1j * a2 * U[2] + d2
    assert np.abs(bad - ref[0]).max() > 1e-3 * scale
