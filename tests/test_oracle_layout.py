"""Pins the oracle's permuted real-FFT layouts (``RealFourierLayout((2,1,3))``
and ``((3,2,1))``, ``src/Fields.jl:219-239``): the folded axis is y or z.
Each check ties the folded-axis form to something fixed independently: the
full complex FFT, the x-folded form on transposed data (itself pinned against
a brute-force DFT), Parseval against the physical inner, the reference's own
multiplier test (``test/fields.jl:243-270``) and the reality round trip."""

import math

import numpy as np
import pytest

from oracle import gbg_oracle as O

TWO_PI = 2 * math.pi
BOX = (TWO_PI, 2 * TWO_PI, 3 * TWO_PI)


def rnd(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


@pytest.mark.parametrize("fold", [1, 2])
@pytest.mark.parametrize("n", [(8, 7, 6), (6, 8, 7), (5, 6, 9)])
def test_folded_forward_is_the_sliced_full_fft(n, fold):
    g = O.Grid3D(n, BOX, fold)
    f = rnd((2,) + g.physical_shape, 1)
    full = np.fft.fftn(f, axes=(-3, -2, -1))
    idx = [slice(None)] * 4
    idx[3 - fold] = slice(0, g.nh)
    spec = O.to_spectral(f, fold=fold)
    assert spec.shape == (2,) + g.spectral_shape
    np.testing.assert_allclose(spec, full[tuple(idx)], atol=1e-12)
    np.testing.assert_allclose(O.to_physical(spec, g), f, atol=1e-13)


@pytest.mark.parametrize("fold", [1, 2])
def test_folded_inverse_matches_x_folded_on_transposed_data(fold):
    """A non-Hermitian half spectrum folded on axis f, transposed so that f
    becomes x, inverts like the (brute-force pinned) x-folded c2r."""
    n = (8, 7, 6)
    g = O.Grid3D(n, BOX, fold)
    rng = np.random.default_rng(3)
    h = rng.standard_normal(g.spectral_shape) + 1j * rng.standard_normal(g.spectral_shape)
    # array axes (z, y, x); swap x with the folded axis
    sw = (0, 2) if fold == 2 else (1, 2)
    nt = list(n)
    nt[0], nt[fold] = nt[fold], nt[0]
    gt = O.Grid3D(tuple(nt), BOX, 0)
    ht = np.swapaxes(h, *sw)
    want = np.swapaxes(O.direct_dft_inverse(ht, gt), *sw)
    np.testing.assert_allclose(O.to_physical(h, g), want, atol=1e-13)


@pytest.mark.parametrize("fold", [0, 1, 2])
def test_folded_multipliers_follow_rfftfreq_on_the_fold(fold):
    """test/fields.jl:243-270: the derivative family zeroes an even axis's
    Nyquist, the squared family keeps the true k^2, and only the folded axis
    is halved."""
    n = (8, 7, 6)
    g = O.Grid3D(n, BOX, fold)
    m = O.Multipliers.build(g)
    tabs = ((m.k1, m.s1), (m.k2, m.s2), (m.k3, m.s3))
    for a in range(3):
        scale = TWO_PI / BOX[a]
        full = scale * (np.fft.rfftfreq(n[a], 1.0 / n[a]) if a == fold else np.fft.fftfreq(n[a], 1.0 / n[a]))
        deriv = full.copy()
        if n[a] % 2 == 0:
            deriv[n[a] // 2] = 0.0
        np.testing.assert_allclose(tabs[a][0], deriv, atol=1e-14)
        np.testing.assert_allclose(tabs[a][1], full ** 2, atol=1e-12)


@pytest.mark.parametrize("fold", [1, 2])
@pytest.mark.parametrize("n", [(8, 7, 6), (6, 8, 7), (5, 6, 9)])
def test_folded_parseval_and_radial_sum(n, fold):
    g = O.Grid3D(n, BOX, fold)
    m = O.Multipliers.build(g)
    f = rnd((3,) + g.physical_shape, 5)
    q = rnd((3,) + g.physical_shape, 6)
    F, Q = O.to_spectral(f, fold=fold), O.to_spectral(q, fold=fold)
    assert abs(O.inner_spectral(F, Q, g) - O.inner_physical(f, q, g)) < 1e-13
    _, spec = O.radial_spectrum(F, m)
    assert abs(spec.sum() - O.inner_physical(f, f, g)) < 1e-12


@pytest.mark.parametrize("fold", [1, 2])
@pytest.mark.parametrize("n", [(8, 7, 6), (6, 8, 7), (5, 6, 9)])
def test_folded_reality_projection_is_the_round_trip(n, fold):
    """The projection onto the reality condition of the folded axis's
    self-conjugate planes is what a c2r / r2c round trip returns."""
    g = O.Grid3D(n, BOX, fold)
    rng = np.random.default_rng(7)
    h = rng.standard_normal(g.spectral_shape) + 1j * rng.standard_normal(g.spectral_shape)
    p = O.reality_projection(h, g)
    np.testing.assert_allclose(O.to_spectral(O.to_physical(h, g), fold=fold), p, atol=1e-12)
    assert O.reality_residual(p, g) < 1e-14
    assert O.reality_residual(h, g) > 0.1
