"""gbg_uxw_noalloc, the u x omega entry point that allocates nothing (its work
area is the library's fixed device array), through the C ABI: its output
equals gbg_uxw's on the same input (and both meet the oracle's FP64 / FP32
bars), on two calls in a row where the second is smaller, so the second
runs over what the first left in the shared array.  Sizes beyond the array
are refused before any launch."""

import ctypes as C
import sys

import numpy as np
import pytest

from oracle import gbg_oracle as O

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 1e-5}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G(gbg):
    if gbg.device_count() < 1:
        pytest.skip("no GPU")
    return gbg


def noalloc(G, plans, dn, du, kc, flags=0):
    L = G._lib.load()
    E, Z = C.c_double(0.0), C.c_double(0.0)
    G.check(L.gbg_uxw_noalloc(plans.handle, du.data, dn.data, float(kc), C.byref(E), C.byref(Z),
                              flags, None))
    return E.value, Z.value


def one_case(G, n, dtype, seed, flags=0):
    g = O.Grid3D(n)
    m = O.Multipliers.build(g)
    u = O.synthetic_field(g, seed, vector=True, m=m)
    kc = O.dealias_cutoff(n[0])
    ref, E, Z = O.uxw(u, g, kc, m)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    dn = G.Field(grid, "spectral", 1, dtype)
    # the no-allocation call first, on a fresh plan (no cached tile list)
    e, z = noalloc(G, plans, dn, du, kc, flags)
    got = dn.download()
    # then the allocating entry point on the same plan and input
    da = G.Field(grid, "spectral", 1, dtype)
    ea, za = G.uxw_(da, du, plans, kc, fused=not flags)
    want = da.download()
    assert np.array_equal(got, want), f"{n}: noalloc differs from gbg_uxw"
    assert e == ea and z == za
    tol = TOL[dtype]
    assert O.rel_l2(got, ref) < tol
    assert abs(e - E) <= tol * abs(E) and abs(z - Z) <= tol * abs(Z)
    mask = O.band_mask(m, 0.0, kc)
    assert np.all(got[:, mask == 0] == 0)
    # a second no-allocation call on the same plan (now with its tile list cached)
    dn.fill_zero()
    e2, z2 = noalloc(G, plans, dn, du, kc, flags)
    assert np.array_equal(dn.download(), want) and e2 == ea and z2 == za


@pytest.mark.parametrize("dtype,sizes", [(np.float64, [(128, 128, 128), (64, 64, 64)]),
                                         (np.float32, [(128, 128, 128), (32, 32, 32)])])
def test_noalloc_matches_uxw_large_then_small(G, dtype, sizes):
    for i, n in enumerate(sizes):
        say(f"[noalloc] n={n} dtype={np.dtype(dtype).name} call {i}")
        one_case(G, n, dtype, 2001 + i)


def test_noalloc_composed_path(G):
    """A grid with a non-power-of-two axis runs the composed operators in the
    array; then a smaller one."""
    for i, n in enumerate([(48, 48, 48), (24, 16, 20)]):
        say(f"[noalloc-composed] n={n} call {i}")
        one_case(G, n, np.float64, 2011 + i, flags=G._lib.GBG_UXW_UNFUSED)


def test_noalloc_refuses_beyond_the_array(G):
    say("[noalloc-refusal] 256^3 FP64")
    n = (256, 256, 256)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, np.float64)
    du = G.Field(grid, "spectral", 1)
    dn = G.Field(grid, "spectral", 1)
    with pytest.raises(G.UnsupportedError, match="scratch"):
        noalloc(G, plans, dn, du, O.dealias_cutoff(256))
    with pytest.raises(G.AliasError):
        noalloc(G, plans, du, du, O.dealias_cutoff(256))
