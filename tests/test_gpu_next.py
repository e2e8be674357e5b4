"""The operators either side of the nonlinear term, through the C ABI against
the oracle: the exponential (integrating-factor) kernel, the radial spectrum
and the reality projection/residual.  FP64 rel-L2 <= 1e-12, FP32 <= 1e-5."""

import sys

import numpy as np
import pytest

from oracle import gbg_oracle as O

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 1e-5}
CDT = {np.float64: np.complex128, np.float32: np.complex64}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G(gbg):
    if gbg.device_count() < 1:
        pytest.skip("no GPU")
    return gbg


def rand_spec(shape, rng):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


GRIDS = [(16, 16, 16), (32, 24, 20), (64, 64, 64), (9, 8, 10)]
L_ANISO = (6.283185307179586, 9.0, 4.0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", GRIDS)
@pytest.mark.parametrize("order", [0, 1, 2])
def test_exponential_parity(G, n, order, dtype):
    say(f"[exp] n={n} order={order} dtype={np.dtype(dtype).name}")
    L = L_ANISO if n == (32, 24, 20) else None
    g = O.Grid3D(n) if L is None else O.Grid3D(n, L)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n) if L is None else G.Grid3D(n, L)
    plans = G.FieldTransformPlans(grid, dtype)
    rng = np.random.default_rng(11 + order)
    shape = (3,) * order + g.spectral_shape
    f = rand_spec(shape, rng).astype(CDT[dtype])
    d0 = rand_spec(shape, rng).astype(CDT[dtype])
    b = rand_spec(shape, rng).astype(CDT[dtype])
    fq, dq, bq = (x.astype(np.complex128) for x in (f, d0, b))
    c = -0.037
    df = G.Field(grid, "spectral", order, dtype).upload(f)
    dd = G.Field(grid, "spectral", order, dtype).upload(d0)
    db = G.Field(grid, "spectral", order, dtype).upload(b)
    tol = TOL[dtype]
    # plain, out of place
    out = G.exponential_convolve(df, plans, coefficient=c)
    assert O.rel_l2(out.download(), O.exponential_convolve(fq, m, c)) < tol
    # alpha/beta accumulate into dst
    G.exponential_convolve_(dd, df, plans, 0.5, -1.5, coefficient=c)
    assert O.rel_l2(dd.download(), O.exponential_convolve(fq, m, c, 0.5, -1.5, dq)) < tol
    # offset: w = |k|^2 stored as a complex scalar field, and a complex offset
    w = m.ksq().astype(CDT[dtype])
    dw = G.Field(grid, "spectral", 0, dtype).upload(w)
    out2 = G.exponential_convolve(df, plans, coefficient=c, offset=dw, offset_scale=0.011)
    assert O.rel_l2(out2.download(), O.exponential_convolve(fq, m, c, offset=w.astype(np.complex128),
                                                            offset_scale=0.011)) < tol
    wz = (rng.uniform(-1, 1, g.spectral_shape) + 1j * rng.uniform(-3, 3, g.spectral_shape)).astype(CDT[dtype])
    dwz = G.Field(grid, "spectral", 0, dtype).upload(wz)
    out3 = G.exponential_convolve(df, plans, coefficient=c, offset=dwz, offset_scale=0.7)
    assert O.rel_l2(out3.download(), O.exponential_convolve(fq, m, c, offset=wz.astype(np.complex128),
                                                            offset_scale=0.7)) < tol
    # two-operand form, and with b aliased to dst
    dc = df.similar()
    G.exponential_convolve_(dc, 0.3, df, 0.7, db, plans, coefficient=c)
    ref = O.exponential_combine(0.3, fq, 0.7, bq, m, c)
    assert O.rel_l2(dc.download(), ref) < tol
    G.exponential_convolve_(db, 0.3, df, 0.7, db, plans, coefficient=c)
    assert np.array_equal(db.download(), dc.download())
    # in place equals out of place bitwise
    inplace = df.similar().copy_from(df)
    G.exponential_convolve_(inplace, inplace, plans, coefficient=c)
    assert np.array_equal(inplace.download(), out.download())


def test_exponential_refusals(G):
    grid = G.Grid3D((16, 16, 16))
    plans = G.FieldTransformPlans(grid)
    u = G.Field(grid, "spectral", 1)
    lib = G.load()
    with pytest.raises(G.GbgError):   # partial overlap of destination and source
        G.check(lib.gbg_exponential_convolve(plans.handle, u.data + 16, u.data, 3, -0.1, None, 1.0,
                                             1.0, 0.0, None))
    with pytest.raises(G.GbgError):   # non-finite coefficient
        G.check(lib.gbg_exponential_convolve(plans.handle, u.data, u.data, 3, float("nan"), None,
                                             1.0, 1.0, 0.0, None))
    with pytest.raises(G.UnsupportedError):   # a physical offset has no method
        G.exponential_convolve(u, plans, coefficient=-0.1, offset=G.Field(grid, "physical", 0))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,L", [((16, 16, 16), None), ((9, 8, 10), None), ((32, 24, 20), L_ANISO),
                                 ((64, 64, 64), None), ((128, 32, 16), None)])
def test_radial_spectrum_parity(G, n, L, dtype):
    say(f"[radial] n={n} L={L} dtype={np.dtype(dtype).name}")
    g = O.Grid3D(n) if L is None else O.Grid3D(n, L)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n) if L is None else G.Grid3D(n, L)
    plans = G.FieldTransformPlans(grid, dtype)
    u = O.synthetic_field(g, 21, vector=True, m=m).astype(CDT[dtype])
    v = O.synthetic_field(g, 22, vector=True, m=m).astype(CDT[dtype])
    uq, vq = u.astype(np.complex128), v.astype(np.complex128)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    dv = G.Field(grid, "spectral", 1, dtype).upload(v)
    sh, spec = G.radial_spectrum(du, plans)
    rsh, rspec = O.radial_spectrum(uq, m)
    np.testing.assert_allclose(sh, rsh, rtol=1e-15)
    assert O.rel_l2(spec, rspec) < TOL[dtype]
    assert abs(spec.sum() - G.inner(du, du, plans)) < 1e-12 * abs(spec.sum())
    _, co = G.radial_spectrum(du, dv, plans)
    _, rco = O.radial_spectrum(uq, m, vq)
    assert np.max(np.abs(co - rco)) < TOL[dtype] * np.max(np.abs(rco))
    _, again = G.radial_spectrum(du, plans)
    assert np.array_equal(again, spec)          # fixed summation order


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [(8, 8, 8), (8, 7, 6), (7, 8, 8), (32, 24, 20)])
def test_reality_parity(G, n, dtype):
    say(f"[reality] n={n} dtype={np.dtype(dtype).name}")
    g = O.Grid3D(n)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    rng = np.random.default_rng(31)
    for order in (0, 1, 2):
        shape = (3,) * order + g.spectral_shape
        f = rand_spec(shape, rng).astype(CDT[dtype])
        df = G.Field(grid, "spectral", order, dtype).upload(f)
        r = G.reality_residual(df, plans)
        assert abs(r - O.reality_residual(f.astype(np.complex128), g)) <= 1e-6 * r if dtype == np.float32 \
            else abs(r - O.reality_residual(f, g)) <= 1e-15 * r
        # out of place: source untouched, destination equals the oracle's projection bitwise
        dp = G.reality_projection_(df.similar(), df, plans)
        assert np.array_equal(df.download(), f)
        want = O.reality_projection(f, g)
        assert np.array_equal(dp.download(), want)
        assert G.reality_residual(dp, plans) == 0.0
        # in place, idempotent
        G.reality_projection_(df, plans)
        assert np.array_equal(df.download(), want)
        G.reality_projection_(df, plans)
        assert np.array_equal(df.download(), want)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [(8, 8, 8), (9, 8, 10), (64, 48, 40)])
def test_inner_physical_parity(G, n, dtype):
    """Physical inner (src/Reductions.jl:838-853) against the oracle and, by
    Parseval, against the spectral inner of the transformed fields."""
    say(f"[inner-physical] n={n} dtype={np.dtype(dtype).name}")
    g = O.Grid3D(n)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    rng = np.random.default_rng(41)
    for order in (0, 1):
        shape = (3,) * order + g.physical_shape
        f = rng.standard_normal(shape).astype(dtype)
        h = rng.standard_normal(shape).astype(dtype)
        df = G.Field(grid, "physical", order, dtype).upload(f)
        dh = G.Field(grid, "physical", order, dtype).upload(h)
        got = G.inner(df, dh, plans)
        ref = O.inner_physical(f.astype(np.float64), h.astype(np.float64), g)
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))
        spec = G.inner(G.to_spectral(df, plans), G.to_spectral(dh, plans), plans)
        assert abs(got - spec) <= TOL[dtype] * max(1.0, abs(got))
    with pytest.raises(G.ArgumentError):
        G.inner(df, G.to_spectral(dh, plans), plans)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("key", ["512x512x512", "256x256x256", "128x128x256", "256x64x64", "48x48x48"])
def test_mask_digest_full_size(G, key, dtype):
    """The 2/3 mask the device applies (spectral_cutoff convolve of an all-ones
    field) at full size is bit-identical to the committed digest."""
    import hashlib
    import json
    import os
    say(f"[mask] {key} dtype={np.dtype(dtype).name}")
    with open(os.path.join(os.path.dirname(__file__), "golden", "mask_digests.json")) as fh:
        want = json.load(fh)[key]
    n = tuple(int(v) for v in key.split("x"))
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    ones = G.Field(grid, "spectral", 0, dtype).upload(np.ones(grid.spectral_shape(0), dtype=CDT[dtype]))
    masked = G.convolve_(ones.similar(), ones, G.spectral_cutoff(plans, want["kc"]), plans)
    got = masked.download()
    h = hashlib.sha256()
    kept = 0
    for kz in range(n[2]):
        plane = got[kz]
        bits = plane.real != 0
        assert np.all((plane.real == 1) | (plane.real == 0)) and np.all(plane.imag == 0)
        kept += int(bits.sum())
        h.update(np.packbits(bits.ravel()).tobytes())
    assert (kept, h.hexdigest()) == (want["kept"], want["sha256"])
