"""The z-slab chain of u x omega's middle passes (``GBG_UXW_SLABS``): the
y-inverse, x and y-forward passes run a few z planes at a time over a ring of
small slots instead of two full-size buffers.  The same kernels compute every
value, so the results must equal the default fused schedule's bit for bit:
for several slab widths and slot counts, a last slab shorter than the others,
one slot (every slab on the caller's stream), FP64 and FP32 (whose zero-tile
kernel runs on a second stream beside the first pass), the
integrating-factor combination, and the 512^3 FP64 term.  One case is also
checked against the oracle, and the plan's device bytes show the smaller work
area."""

import os
import sys

import numpy as np
import pytest

from oracle import gbg_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G(gbg):
    if gbg.device_count() < 1:
        pytest.skip("no GPU")
    return gbg


def plans_with(G, grid, dtype, slab):
    """A plan made with GBG_UXW_SLAB=slab (planes per slab, slots) in force."""
    old = os.environ.pop("GBG_UXW_SLAB", None)
    try:
        if slab is not None:
            os.environ["GBG_UXW_SLAB"] = slab
        return G.FieldTransformPlans(grid, dtype)
    finally:
        os.environ.pop("GBG_UXW_SLAB", None)
        if old is not None:
            os.environ["GBG_UXW_SLAB"] = old


def same_bits(a, b):
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,slab", [
    ((64, 32, 128), "4,2"),     # 32 slabs on two streams
    ((64, 32, 128), "5,2"),     # last slab of 3 planes
    ((64, 32, 128), "7,3"),     # three slots, last slab of 2 planes
    ((64, 32, 128), "3,1"),     # one slot: every slab on the caller's stream
    ((32, 64, 96), "1,4"),      # one plane per slab, four slots
    ((128, 128, 64), None),     # the size rule: one slab holding every plane
])
def test_slabs_equal_default(G, n, slab, dtype):
    say(f"[slabs] n={n} slab={slab} dtype={np.dtype(dtype).name}")
    g = O.Grid3D(n)
    m = O.Multipliers.build(g)
    u = O.synthetic_field(g, 23, vector=True, m=m)
    grid = G.Grid3D(n)
    kc = O.dealias_cutoff(n[0])
    pd = G.FieldTransformPlans(grid, dtype)
    ps = plans_with(G, grid, dtype, slab)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    a, b = G.Field(grid, "spectral", 1, dtype), G.Field(grid, "spectral", 1, dtype)
    Ea, Za = G.uxw_(a, du, pd, kc)
    Eb, Zb = G.uxw_(b, du, ps, kc, slabs=True)
    ra, rb = a.download(), b.download()
    assert same_bits(ra, rb), f"{n} slab={slab}: slab chain differs from the default schedule"
    assert (Ea, Za) == (Eb, Zb)
    # a second call on the same plan reuses its streams and slots
    b.fill_zero()
    G.uxw_(b, du, ps, kc, reductions=False, slabs=True)
    assert same_bits(ra, b.download())
    # the input is read only
    assert same_bits(du.download(), u.astype(du.npdtype))


def test_slabs_against_oracle(G):
    n = (64, 32, 128)
    say(f"[slabs-oracle] n={n}")
    g = O.Grid3D(n)
    m = O.Multipliers.build(g)
    u = O.synthetic_field(g, 29, vector=True, m=m)
    kc = O.dealias_cutoff(n[0])
    want, E, Z = O.uxw(u, g, kc, m)
    grid = G.Grid3D(n)
    ps = plans_with(G, grid, np.float64, "6,2")
    du = G.Field(grid, "spectral", 1, np.float64).upload(u)
    dn = G.Field(grid, "spectral", 1, np.float64)
    Eg, Zg = G.uxw_(dn, du, ps, kc, slabs=True)
    got = dn.download()
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    assert abs(Eg - E) <= 1e-12 * abs(E) and abs(Zg - Z) <= 1e-12 * abs(Z)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_slabs_combination_equal_default(G, dtype):
    n = (64, 64, 64)
    say(f"[slabs-combine] n={n} dtype={np.dtype(dtype).name}")
    g = O.Grid3D(n)
    m = O.Multipliers.build(g)
    u = O.synthetic_field(g, 31, vector=True, m=m)
    grid = G.Grid3D(n)
    kc = O.dealias_cutoff(n[0])
    pd = G.FieldTransformPlans(grid, dtype)
    ps = plans_with(G, grid, dtype, "5,2")
    V = lambda: G.Field(grid, "spectral", 1, dtype)   # noqa: E731
    du = V().upload(u)
    outs = []
    for plans, slabs in ((pd, False), (ps, True)):
        dst, nh = V(), V()
        G.uxw_combine_(dst, du, plans, kc, n_hat=nh, a=du, alpha=1.0, delta=0.01,
                       coefficient=-2e-4, slabs=slabs)
        outs.append((dst.download(), nh.download()))
    assert same_bits(outs[0][0], outs[1][0]) and same_bits(outs[0][1], outs[1][1])


def test_slabs_512_fp64(G):
    """The 512^3 FP64 term: the size rule's 8 planes per slab on two slots,
    the same bits as the default schedule, and a work area smaller by the six
    full-y fields less the two slots."""
    from gbg_amd import workload
    n = (512, 512, 512)
    say(f"[slabs-512] n={n}")
    grid = G.Grid3D(n)
    kc = O.dealias_cutoff(512)
    pd = G.FieldTransformPlans(grid, np.float64)
    ps = G.FieldTransformPlans(grid, np.float64)
    du = G.Field(grid, "spectral", 1, np.float64).upload(workload.synthetic_field(n, 2001, vector=True))
    a, b = G.Field(grid, "spectral", 1, np.float64), G.Field(grid, "spectral", 1, np.float64)
    Ea, Za = G.uxw_(a, du, pd, kc)
    Eb, Zb = G.uxw_(b, du, ps, kc, slabs=True)
    assert (Ea, Za) == (Eb, Zb)
    assert same_bits(a.download(), b.download())
    nhp, ny, nz = 176, 512, 512            # kept kx 0..170, row pitch rounded up to 16
    saved = 6 * nz * ny * nhp * 16 - 2 * 9 * 8 * ny * nhp * 16
    assert pd.device_bytes() - ps.device_bytes() == saved
