"""GPU parity for permuted real-FFT layouts (``RealFourierLayout((2,1,3))``,
``((3,2,1))`` and the other permutations, ``src/Fields.jl:219-239``): the
folded axis is y or z.  Transforms, operators, reductions, the reality
condition, the exponential kernel and the composed u x omega / product
against the oracle built for the same layout, through the C ABI
(``test/fields.jl:243-270`` is the reference's own test of the layouts)."""

import math
import sys

import numpy as np
import pytest

from oracle import gbg_oracle as O

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 1e-5}
DIMS = [(2, 1, 3), (3, 2, 1), (2, 3, 1), (3, 1, 2)]
BOX = (O.TWO_PI, 2 * O.TWO_PI, 3 * O.TWO_PI)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G(gbg):
    if gbg.device_count() < 1:
        pytest.skip("no GPU")
    return gbg


def rand_phys(g, order, seed):
    return np.random.default_rng(seed).standard_normal((3,) * order + g.physical_shape)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("n,L", [((8, 7, 6), BOX), ((16, 12, 10), (O.TWO_PI,) * 3),
                                 ((32, 32, 64), (O.TWO_PI,) * 3), ((17, 8, 6), BOX)])
def test_layout_transforms_and_operators(G, n, L, dims, dtype):
    say(f"[layout] n={n} dims={dims} dtype={np.dtype(dtype).name}")
    fold = dims[0] - 1
    g = O.Grid3D(n, L, fold)
    m = O.Multipliers.build(g)
    lay = G.RealFourierLayout(dims)
    grid = G.Grid3D(n, L)
    plans = G.FieldTransformPlans(grid, dtype, layout=lay)
    tol = TOL[dtype]
    F = lambda rep, order: G.Field(grid, rep, order, dtype, layout=lay)   # noqa: E731
    for order in (0, 1):
        f = rand_phys(g, order, 11 + order)
        ref = O.to_spectral(f, fold=fold)
        spec = G.to_spectral(F("physical", order).upload(f), plans)
        assert spec.shape == ref.shape
        got = spec.download()
        assert O.rel_l2(got, ref) < tol, "forward"
        back = G.to_physical(spec, plans)
        assert O.rel_l2(back.download(), f) < 10 * tol, "round trip"
        np.testing.assert_array_equal(spec.download(), got)
        rng = np.random.default_rng(5 + order)
        h = ref + 0.25j * (rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape))
        out = F("physical", order)
        G.to_physical_(out, F("spectral", order).upload(h), plans)      # 3-arg: consumes
        assert O.rel_l2(out.download(), O.to_physical(h, g)) < 10 * tol, "c2r"
    u = O.to_spectral(rand_phys(g, 1, 21), fold=fold)
    s = O.to_spectral(rand_phys(g, 0, 22), fold=fold)
    du, ds = F("spectral", 1).upload(u), F("spectral", 0).upload(s)
    assert O.rel_l2(G.gradient(ds, plans).download(), O.gradient(s, m)) < tol
    assert O.rel_l2(G.curl(du, plans).download(), O.curl(u, m)) < tol
    assert O.rel_l2(G.symmetric_gradient(du, plans).download(), O.symmetric_gradient(u, m)) < tol
    dv = F("spectral", 0)
    G.divergence_(dv, du, plans)
    assert O.rel_l2(dv.download(), O.divergence(u, m)) < tol
    lap = F("spectral", 1)
    G.laplacian_(lap, du, plans)
    assert O.rel_l2(lap.download(), O.laplacian(u, m)) < tol
    pr = F("spectral", 1)
    G.helmholtz_projection_(pr, du, plans)
    assert O.rel_l2(pr.download(), O.helmholtz_projection(u, m)) < tol
    band = G.spectral_band(plans, 1.5, 4.0)
    cb = F("spectral", 1)
    G.convolve_(cb, du, band, plans)
    assert O.rel_l2(cb.download(), O.convolve_band(u, m, 1.5, 4.0)) < tol
    ex = G.exponential_convolve(du, plans, coefficient=-0.01)
    assert O.rel_l2(ex.download(), O.exponential_convolve(u, m, -0.01)) < tol
    # Parseval weights on the folded axis, radial spectrum, reality condition
    rtol = 1e-12 if dtype == np.float64 else 1e-5
    assert abs(G.inner(du, du, plans) - O.inner_spectral(u, u, g)) <= rtol * O.inner_spectral(u, u, g)
    sh, sp = G.radial_spectrum(du, plans)
    _, sref = O.radial_spectrum(u, m)
    np.testing.assert_allclose(sp, sref, rtol=0, atol=rtol * np.abs(sref).max())
    rng = np.random.default_rng(9)
    h = u + 0.1 * (rng.standard_normal(u.shape) + 1j * rng.standard_normal(u.shape))
    dh = F("spectral", 1).upload(h)
    res = G.reality_residual(dh, plans)
    hr = dh.download().astype(np.complex128)          # the stored (rounded) entries
    assert abs(res - O.reality_residual(hr, g)) <= rtol * max(1.0, O.reality_residual(hr, g))
    G.reality_projection_(dh, plans)
    assert O.rel_l2(dh.download(), O.reality_projection(h, g)) < tol
    mf = F("spectral", 1).upload(u)
    G.mean_free_projection_(mf, plans)
    assert O.rel_l2(mf.download(), O.mean_free(u)) < tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims", [(2, 1, 3), (3, 2, 1)])
@pytest.mark.parametrize("n", [(16, 12, 10), (32, 32, 32)])
def test_layout_pipelines(G, n, dims, dtype):
    """u x omega (composed form, energy and enstrophy) and the dealiased
    product on a permuted layout."""
    say(f"[layout-pipelines] n={n} dims={dims} dtype={np.dtype(dtype).name}")
    fold = dims[0] - 1
    g = O.Grid3D(n, (O.TWO_PI,) * 3, fold)
    m = O.Multipliers.build(g)
    lay = G.RealFourierLayout(dims)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype, layout=lay)
    u = O.synthetic_field(g, 3, vector=True, m=m)
    kc = O.dealias_cutoff(min(n))
    du = G.Field(grid, "spectral", 1, dtype, layout=lay).upload(u)
    dn = G.Field(grid, "spectral", 1, dtype, layout=lay)
    E, Z = G.uxw_(dn, du, plans, kc)
    nref, Eref, Zref = O.uxw(u, g, kc, m)
    assert np.linalg.norm(nref) > 1e-3 * np.linalg.norm(u)
    assert O.rel_l2(dn.download(), nref) < 10 * TOL[dtype]
    assert abs(E - Eref) <= 1e-5 * Eref and abs(Z - Zref) <= 1e-5 * Zref
    # the integrating-factor combination on the same layout (composed form)
    dst = G.Field(grid, "spectral", 1, dtype, layout=lay).upload(u)
    G.uxw_combine_(dst, du, plans, kc, a=dst, delta=0.01, gamma=0.5, coefficient=-0.02)
    want = O.exponential_factor(m, -0.02) * (u + 0.01 * nref) + 0.5 * nref
    assert O.rel_l2(dst.download(), want) < 10 * TOL[dtype]
    a = O.synthetic_field(g, 4, m=m)
    da = G.Field(grid, "spectral", 0, dtype, layout=lay).upload(a)
    dp = G.Field(grid, "spectral", 0, dtype, layout=lay)
    G.dealiased_product_(dp, da, da, plans, kc)
    assert O.rel_l2(dp.download(), O.dealiased_product(a, a, g, kc, m)) < 10 * TOL[dtype]


def test_layout_refusals(G):
    say("[layout-refusals]")
    grid = G.Grid3D((8, 8, 8))
    with pytest.raises(G.ArgumentError):
        G.RealFourierLayout((1, 1, 3))
    L_ = G._lib.load()
    import ctypes as C
    h = C.c_void_p()
    st = L_.gbg_plan_create_layout((C.c_int * 3)(8, 8, 8), (C.c_double * 3)(1.0, 1.0, 1.0), 8,
                                   (C.c_int * 3)(2, 2, 3), C.byref(h))
    assert st == 1 and b"permutation" in L_.gbg_last_error()
    lay = G.RealFourierLayout((2, 1, 3))
    plans = G.FieldTransformPlans(grid, layout=lay)
    plain = G.FieldTransformPlans(grid)
    f = G.Field(grid, "spectral", 1, layout=lay)
    assert f.shape == (3, 8, 5, 8)
    with pytest.raises(G.ArgumentError):           # field layout != plans layout
        G.curl_(f, f, plain)
    with pytest.raises(G.ArgumentError):           # fields of different layouts
        G.curl_(G.Field(grid, "spectral", 1), f, plans)
    with pytest.raises(G.ArgumentError):           # (2, 3, 1) folds y too, but is another layout
        G.curl_(G.Field(grid, "spectral", 1, layout=G.RealFourierLayout((2, 3, 1))), f, plans)
    b = G.Batch(nstreams=1, autotune=False)
    with pytest.raises(G.UnsupportedError):
        b.add_vector(f)
    b.close()
