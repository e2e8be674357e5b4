"""The integrating-factor RK2 step of the consumer (``step!``,
``examples/HIST_CPU.jl:267-282``) with its combinations folded into the last
pass of u x omega (``gbg_uxw_combine``), against the oracle's step composed
from the same pieces the reference composes (``rhs!`` and
``exponential_convolve!``), on fused and composed grids, FP64 and FP32."""

import ctypes as C
import sys

import numpy as np
import pytest

from oracle import gbg_oracle as O

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 1e-5}
NU, DT = 0.02, 0.01


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G(gbg):
    if gbg.device_count() < 1:
        pytest.skip("no GPU")
    return gbg


def oracle_step(u, g, m, kc):
    """step! without forcing: k1 = N(u); u* = E(u + dt k1); k2 = N(u*);
    u_new = E u + dt/2 (E k1 + k2), E = exp(-nu dt |k|^2)."""
    E = O.exponential_factor(m, -NU * DT)
    k1 = O.uxw(u, g, kc, m)[0]
    us = E * (u + DT * k1)
    k2 = O.uxw(us, g, kc, m)[0]
    return k1, us, E * u + (DT / 2) * (E * k1 + k2)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [(32, 32, 32), (64, 64, 64), (128, 64, 32), (48, 48, 48), (16, 12, 10)])
def test_rk2_step_with_folded_combinations(G, n, dtype):
    say(f"[step] n={n} dtype={np.dtype(dtype).name}")
    g = O.Grid3D(n)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    tol = TOL[dtype]
    kc = O.dealias_cutoff(n[0])
    u = O.synthetic_field(g, 11, vector=True, m=m)
    k1_ref, us_ref, un_ref = oracle_step(u, g, m, kc)
    V = lambda: G.Field(grid, "spectral", 1, dtype)   # noqa: E731
    du, k1, us = V().upload(u), V(), V()
    # stage 1: k1 = N(u_n), u* = E (u_n + dt k1)
    E1, Z1 = G.uxw_combine_(us, du, plans, kc, n_hat=k1, a=du, alpha=1.0, delta=DT,
                            coefficient=-NU * DT)
    plain = V()
    E0, Z0 = G.uxw_(plain, du, plans, kc)
    assert O.rel_l2(k1.download(), plain.download()) < tol
    assert abs(E1 - E0) <= 1e-12 * E0 and abs(Z1 - Z0) <= 1e-12 * Z0
    assert O.rel_l2(k1.download(), k1_ref) < 10 * tol
    assert O.rel_l2(us.download(), us_ref) < 10 * tol
    # stage 2, in place on u_n: u_n <- E (u_n + dt/2 k1) + dt/2 N(u*)
    G.uxw_combine_(du, us, plans, kc, a=du, alpha=1.0, b=k1, beta=DT / 2, gamma=DT / 2,
                   coefficient=-NU * DT, reductions=False)
    assert O.rel_l2(du.download(), un_ref) < 10 * tol
    # the same step through the standalone calls (N^ stored, then the
    # exponential combination) gives the same field
    du2 = V().upload(u)
    G.uxw_(k1, du2, plans, kc)
    G.exponential_convolve_(us, 1.0, du2, DT, k1, plans, coefficient=-NU * DT)
    k2 = V()
    G.uxw_(k2, us, plans, kc)
    G.exponential_convolve_(du2, 1.0, du2, DT / 2, k1, plans, coefficient=-NU * DT)
    out = du2.download() + (DT / 2) * k2.download()
    assert O.rel_l2(du.download(), out) < 10 * tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("fused", [True, False])
def test_combination_terms(G, dtype, fused):
    """Each weight on its own against the oracle: beta b alone (N^ unused),
    gamma N^ alone (a zero factor exponent), and both inputs absent."""
    n = (64, 64, 64)
    say(f"[combination] fused={fused} dtype={np.dtype(dtype).name}")
    g = O.Grid3D(n)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    tol = TOL[dtype]
    kc = O.dealias_cutoff(n[0])
    u = O.synthetic_field(g, 5, vector=True, m=m)
    w = O.synthetic_field(g, 6, vector=True, m=m)
    N = O.uxw(u, g, kc, m)[0]
    E = O.exponential_factor(m, -0.3)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    dw = G.Field(grid, "spectral", 1, dtype).upload(w)
    dst = G.Field(grid, "spectral", 1, dtype)
    G.uxw_combine_(dst, du, plans, kc, b=dw, beta=2.0, coefficient=-0.3, fused=fused)
    assert O.rel_l2(dst.download(), E * (2.0 * w)) < tol
    G.uxw_combine_(dst, du, plans, kc, gamma=3.0, fused=fused)
    assert O.rel_l2(dst.download(), 3.0 * N) < 10 * tol
    G.uxw_combine_(dst, du, plans, kc, a=dw, alpha=-1.0, delta=0.5, gamma=0.25, coefficient=-0.3,
                   fused=fused)
    assert O.rel_l2(dst.download(), E * (-w + 0.5 * N) + 0.25 * N) < 10 * tol


def test_combination_refusals(G):
    say("[combination-refusals]")
    grid = G.Grid3D((32, 32, 32))
    plans = G.FieldTransformPlans(grid)
    u = G.Field(grid, "spectral", 1)
    d = G.Field(grid, "spectral", 1)
    with pytest.raises(G.AliasError):
        G.uxw_combine_(u, u, plans)
    with pytest.raises(G.AliasError):
        G.uxw_combine_(d, u, plans, n_hat=d)
    L = G._lib.load()
    big = G.Field(grid, "spectral", 2)
    cmb = G._lib.GbgUxwCombination(big.data + 64, big.data, 1.0, None, 0.0, 0.0, 0.0, 0.0)
    st = L.gbg_uxw_combine(plans.handle, u.data, None, 10.0, C.byref(cmb), None, None, 0, None)
    assert st == 4 and b"without being it" in L.gbg_last_error()
    cmb = G._lib.GbgUxwCombination(d.data, None, 1.0, None, 0.0, float("nan"), 0.0, 0.0)
    st = L.gbg_uxw_combine(plans.handle, u.data, None, 10.0, C.byref(cmb), None, None, 0, None)
    assert st == 1
    cmb = G._lib.GbgUxwCombination(big.data, None, 1.0, None, 0.0, 0.0, 0.0, 0.0)
    st = L.gbg_uxw_combine(plans.handle, u.data, big.data + 32, 10.0, C.byref(cmb), None, None, 0,
                           None)
    assert st == 4 and b"n_hat must not overlap dst" in L.gbg_last_error()
