"""GPU parity: every entry point of libgbg against the FP64 oracle on the same
seeded inputs, through the C ABI.

Tolerances (relative L2 over stored entries): FP64 1e-12, FP32 1e-5 (the
bars of the spectral path); the 2/3 mask is compared bit-exactly (the set of
modes that are exactly zero)."""

import math
import sys

import numpy as np
import pytest

from oracle import gbg_oracle as O

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 1e-5}
GRIDS = [(8, 7, 6), (9, 9, 9), (7, 5, 3), (16, 12, 10), (32, 48, 40), (64, 64, 64), (30, 50, 36),
         (1, 4, 6)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G(gbg):
    if gbg.device_count() < 1:
        pytest.skip("no GPU")
    return gbg


def ogrid(n, L=(O.TWO_PI,) * 3):
    return O.Grid3D(tuple(n), tuple(L))


def rand_phys(g, order, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((3,) * order + g.physical_shape)


def rand_spec(g, order, seed):
    return O.to_spectral(rand_phys(g, order, seed))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", GRIDS)
def test_transforms(G, n, dtype):
    say(f"[transforms] n={n} dtype={np.dtype(dtype).name}")
    g = ogrid(n)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    for order in (0, 1):
        f = rand_phys(g, order, 11 + order)
        ref = O.to_spectral(f)
        phys = G.Field(grid, "physical", order, dtype).upload(f)
        spec = G.to_spectral(phys, plans)
        got = spec.download()
        assert O.rel_l2(got, ref) < TOL[dtype], "forward"
        # inverse, 4-arg form preserves the source
        back = G.to_physical(spec, plans)
        assert O.rel_l2(back.download(), f) < 10 * TOL[dtype], "round trip"
        np.testing.assert_array_equal(spec.download(), got)
        # inverse of a non-Hermitian half spectrum follows the c2r convention
        h = ref + 0.25 * np.conj(ref[..., ::-1, ::-1, :]) * 1j
        hs = G.Field(grid, "spectral", order, dtype).upload(h)
        out = G.Field(grid, "physical", order, dtype)
        G.to_physical_(out, hs, plans)          # 3-arg: consumes hs
        assert O.rel_l2(out.download(), O.to_physical(h, g)) < 10 * TOL[dtype], "c2r"


CHIRP_GRIDS = [(17, 19, 23), (34, 8, 6), (6, 97, 10), (5, 4, 127), (202, 6, 4), (1009, 4, 3)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", CHIRP_GRIDS)
def test_chirp_z_axes(G, n, dtype):
    """Axis lengths with a prime factor above 13 (chirp-z convolution form):
    forward, both inverse forms and the c2r convention against the oracle on
    every axis position, operators, the composed u x omega, and a tuned plan."""
    say(f"[chirp-z] n={n} dtype={np.dtype(dtype).name}")
    g = ogrid(n)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    tol = TOL[dtype]
    f = rand_phys(g, 1, 41)
    ref = O.to_spectral(f)
    spec = G.to_spectral(G.Field(grid, "physical", 1, dtype).upload(f), plans)
    got = spec.download()
    assert O.rel_l2(got, ref) < tol, "forward"
    assert O.rel_l2(G.to_physical(spec, plans).download(), f) < 10 * tol, "round trip"
    h = ref + 0.25 * np.conj(ref[..., ::-1, ::-1, :]) * 1j
    out = G.Field(grid, "physical", 1, dtype)
    G.to_physical_(out, G.Field(grid, "spectral", 1, dtype).upload(h), plans)
    assert O.rel_l2(out.download(), O.to_physical(h, g)) < 10 * tol, "c2r"
    assert O.rel_l2(G.curl(spec, plans).download(), O.curl(ref, m)) < tol
    u = O.synthetic_field(g, 7, vector=True, m=m)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    kc = O.dealias_cutoff(max(n))          # keeps modes on every axis: a non-vanishing N^
    dn = G.Field(grid, "spectral", 1, dtype)
    G.uxw_(dn, du, plans, kc, reductions=False)
    nref = O.uxw(u, g, kc, m)[0]
    assert np.linalg.norm(nref) > 1e-3 * np.linalg.norm(u)
    assert O.rel_l2(dn.download(), nref) < 10 * tol, "u x omega"
    if n[0] * n[1] * n[2] <= 20000:
        G.autotune(plans)
        spec2 = G.to_spectral(G.Field(grid, "physical", 1, dtype).upload(f), plans)
        assert O.rel_l2(spec2.download(), ref) < tol, "tuned forward"


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [(64, 64, 64), (32, 128, 256), (512, 32, 32), (1024, 16, 16),
                               (16, 1024, 8), (8, 16, 1024)])
def test_transforms_generic_vs_fast(G, n, dtype, monkeypatch):
    """Power-of-two lengths take the register-resident passes; the mixed-radix
    LDS passes (GBG_GENERIC_FFT=1 at plan creation) must agree with them and
    with the oracle.  Vector and tensor fields: the passes that transform
    three components per workgroup, in one and in three groups of three."""
    say(f"[generic-vs-fast] n={n} dtype={np.dtype(dtype).name}")
    g = ogrid(n)
    grid = G.Grid3D(n)
    fast = G.FieldTransformPlans(grid, dtype)
    monkeypatch.setenv("GBG_GENERIC_FFT", "1")
    generic = G.FieldTransformPlans(grid, dtype)
    monkeypatch.delenv("GBG_GENERIC_FFT")
    for order in (1, 2):
        f = rand_phys(g, order, 21 + order)
        ref = O.to_spectral(f)
        phys = G.Field(grid, "physical", order, dtype).upload(f)
        for plans in (fast, generic):
            spec = G.to_spectral(phys, plans)
            assert O.rel_l2(spec.download(), ref) < TOL[dtype]
            back = G.to_physical(spec, plans)
            assert O.rel_l2(back.download(), f) < 10 * TOL[dtype]
            out = G.Field(grid, "physical", order, dtype)
            G.to_physical_(out, spec, plans)   # 3-arg: consumes spec
            assert O.rel_l2(out.download(), f) < 10 * TOL[dtype]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,L", [((16, 12, 10), (O.TWO_PI,) * 3), ((8, 7, 6), (1.0, 2.5, 7.0)),
                                 ((8, 7, 5), (1.0, 2.5, 7.0)), ((64, 64, 64), (O.TWO_PI,) * 3)])
def test_operators(G, n, L, dtype):
    say(f"[operators] n={n} L={L} dtype={np.dtype(dtype).name}")
    g = ogrid(n, L)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n, L)
    plans = G.FieldTransformPlans(grid, dtype)
    tol = TOL[dtype]
    s = rand_spec(g, 0, 1)
    u = rand_spec(g, 1, 2)
    A = rand_spec(g, 2, 3)
    ds = G.Field(grid, "spectral", 0, dtype).upload(s)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    dA = G.Field(grid, "spectral", 2, dtype).upload(A)

    out = G.gradient(ds, plans).download()
    assert O.rel_l2(out, O.gradient(s, m)) < tol
    gt = G.gradient(du, plans).download()
    ref_gt = np.stack([np.stack([(1j * a) * u[i] for a in (m.a1(), m.a2(), m.a3())]) for i in range(3)])
    # tensor block (i, j) is stored at i + 3 j  ->  NumPy [j, i]
    assert O.rel_l2(gt, np.swapaxes(ref_gt, 0, 1)) < tol

    dv = G.Field(grid, "spectral", 0, dtype)
    G.divergence_(dv, du, plans)
    assert O.rel_l2(dv.download(), O.divergence(u, m)) < tol
    band = G.spectral_band(plans, 1.5, 4.0)
    pre = rand_spec(g, 0, 9)
    dv.upload(pre)
    G.divergence_(dv, du, plans, kernel=band, alpha=0.5, beta=2.0)
    ref = 0.5 * O.band_mask(m, 1.5, 4.0) * O.divergence(u, m) + 2.0 * pre
    assert O.rel_l2(dv.download(), ref) < tol
    dvec = G.Field(grid, "spectral", 1, dtype)
    G.divergence_(dvec, dA, plans)
    # tensor->vector: out_i = i sum_j a_j A_{ij}; A block (i,j) = NumPy [j, i]
    ref = np.stack([1j * ((m.a1() * A[0, i] + m.a2() * A[1, i]) + m.a3() * A[2, i]) for i in range(3)])
    assert O.rel_l2(dvec.download(), ref) < tol

    c = G.curl(du, plans).download()
    assert O.rel_l2(c, O.curl(u, m)) < tol
    alias = G.Field(grid, "spectral", 1, dtype).upload(u)
    G.curl_(alias, alias, plans)
    np.testing.assert_array_equal(alias.download(), c)          # aliased == non-aliased, bitwise
    ic = G.Field(grid, "spectral", 1, dtype)
    G.inverse_curl_(ic, du, plans)
    assert O.rel_l2(ic.download(), O.inverse_curl(u, m)) < tol

    lap = G.Field(grid, "spectral", 1, dtype)
    G.laplacian_(lap, du, plans)
    assert O.rel_l2(lap.download(), O.laplacian(u, m)) < tol
    G.inverse_laplacian_(lap, du, plans)
    assert O.rel_l2(lap.download(), O.inverse_laplacian(u, m)) < tol
    # scalar field (in FP32 with an even mode count: two modes per 16-byte
    # access), out of place and in place
    lap0 = G.Field(grid, "spectral", 0, dtype)
    G.laplacian_(lap0, ds, plans)
    assert O.rel_l2(lap0.download(), O.laplacian(s, m)) < tol
    G.inverse_laplacian_(lap0, ds, plans)
    assert O.rel_l2(lap0.download(), O.inverse_laplacian(s, m)) < tol
    s_in = G.Field(grid, "spectral", 0, dtype).upload(s)
    G.inverse_laplacian_(s_in, s_in, plans)
    np.testing.assert_array_equal(s_in.download(), lap0.download())

    hp = G.Field(grid, "spectral", 1, dtype)
    G.helmholtz_projection_(hp, du, plans)
    assert O.rel_l2(hp.download(), O.helmholtz_projection(u, m)) < tol
    pre3 = rand_spec(g, 1, 8)
    hp.upload(pre3)
    G.helmholtz_projection_(hp, du, plans, kernel=band, alpha=0.3, beta=1.0)
    ref = O.helmholtz_projection(u, m, O.band_mask(m, 1.5, 4.0), 0.3, 1.0, pre3)
    assert O.rel_l2(hp.download(), ref) < tol

    cv = G.Field(grid, "spectral", 1, dtype)
    kc = G.dealias_cutoff(n[0], L[0])
    G.convolve_(cv, du, G.spectral_cutoff(plans, kc), plans)
    got = cv.download()
    refm = O.band_mask(m, 0.0, kc)
    assert O.rel_l2(got, refm[None] * u) < tol
    assert np.array_equal(got[:, refm == 0] == 0, np.ones_like(got[:, refm == 0], bool))

    mf = G.Field(grid, "spectral", 1, dtype).upload(u)
    G.mean_free_projection_(mf, plans)
    assert O.rel_l2(mf.download(), O.mean_free(u)) < tol

    # physical products
    pu, pv = rand_phys(g, 1, 5), rand_phys(g, 1, 6)
    fu = G.Field(grid, "physical", 1, dtype).upload(pu)
    fv = G.Field(grid, "physical", 1, dtype).upload(pv)
    fx = G.Field(grid, "physical", 1, dtype)
    G.cross_(fx, fu, fv, plans)
    assert O.rel_l2(fx.download(), O.cross(pu.astype(dtype), pv.astype(dtype))) < tol
    a0, b0 = rand_phys(g, 0, 7), rand_phys(g, 0, 8)
    fa = G.Field(grid, "physical", 0, dtype).upload(a0)
    fb = G.Field(grid, "physical", 0, dtype).upload(b0)
    G.multiply_(fa, fa, fb, plans)
    assert O.rel_l2(fa.download(), (a0.astype(dtype) * b0.astype(dtype))) < tol

    # Parseval inner (FP64 accumulation)
    ip = G.inner(du, G.Field(grid, "spectral", 1, dtype).upload(pre3), plans)
    ref = O.inner_spectral(u.astype(plans.cdtype).astype(np.complex128),
                           pre3.astype(plans.cdtype).astype(np.complex128), g)
    assert abs(ip - ref) <= 1e-12 * abs(ref) + 1e-300


UXW_CASES = [((16, 16, 16), None), ((32, 32, 32), None), ((64, 64, 64), None), ((48, 32, 40), None),
             ((64, 64, 128), None), ((128, 128, 32), None), ((64, 64, 64), 5.0), ((64, 64, 64), 1000.0),
             ((32, 32, 64), 7.3), ((512, 32, 32), 170.5), ((256, 64, 32), 60.0),
             ((1024, 32, 32), 340.0), ((32, 1024, 32), 12.0), ((32, 32, 1024), 12.0),
             # Nx = 1024: 384 kept kx fill the FP32 x pass's staging area exactly; 501
             # exceed it (the unstaged form)
             ((1024, 32, 32), 383.5), ((1024, 32, 32), 500.0),
             # Ny = 512: the y passes of the 512^3 headline, k_uxw_y<T, 512, +-1, ...>
             ((64, 512, 64), None), ((32, 512, 64), 150.0)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,L", [((16, 12, 10), (O.TWO_PI,) * 3), ((8, 7, 5), (1.0, 2.5, 7.0)),
                                 ((9, 8, 6), (O.TWO_PI,) * 3), ((64, 64, 64), (O.TWO_PI,) * 3),
                                 ((1, 4, 6), (O.TWO_PI,) * 3)])
def test_symmetric_gradient(G, n, L, dtype):
    """symmetric_gradient! against the oracle: even and odd mode counts (FP32
    takes two modes per 16-byte access only for an even count), Nyquist
    axes, a non-2pi box; each off-diagonal value written to both blocks, and
    the diagonal equal to the gradient tensor's."""
    say(f"[symmetric-gradient] n={n} L={L} dtype={np.dtype(dtype).name}")
    g = ogrid(n, L)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n, L)
    plans = G.FieldTransformPlans(grid, dtype)
    u = rand_spec(g, 1, 31)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    S = G.symmetric_gradient(du, plans).download()
    ref = O.symmetric_gradient(u, m)
    assert O.rel_l2(S, ref) < TOL[dtype]
    for i in range(3):
        for j in range(3):
            np.testing.assert_array_equal(S[i, j], S[j, i])
    gt = G.gradient(du, plans).download()
    for i in range(3):
        assert O.rel_l2(S[i, i], gt[i, i]) < TOL[dtype]
    with pytest.raises(G.DimensionMismatch):
        G.symmetric_gradient_(G.Field(grid, "spectral", 1, dtype), du, plans)
    with pytest.raises(G.DimensionMismatch):
        G.symmetric_gradient_(G.Field(grid, "spectral", 2, dtype), G.Field(grid, "spectral", 0, dtype),
                              plans)
    L_ = G._lib.load()
    big = G.Field(grid, "spectral", 2, dtype)
    st = L_.gbg_symmetric_gradient(plans.handle, big.data, big.data + 16, None)
    assert st == 4 and b"cannot alias" in L_.gbg_last_error()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,kc", UXW_CASES)
@pytest.mark.parametrize("fused", [False, True])
def test_uxw_parity(G, n, kc, dtype, fused):
    say(f"[uxw] n={n} kc={kc} dtype={np.dtype(dtype).name} fused={fused}")
    g = ogrid(n)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    u = O.synthetic_field(g, 2001, vector=True, m=m)
    kc = O.dealias_cutoff(n[0]) if kc is None else kc
    ref, E, Z = O.uxw(u, g, kc, m)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    dn = G.Field(grid, "spectral", 1, dtype)
    e, z = G.uxw_(dn, du, plans, kc, fused=fused)
    got = dn.download()
    tol = TOL[dtype]
    assert O.rel_l2(got, ref) < tol
    assert abs(e - E) <= tol * abs(E) and abs(z - Z) <= tol * abs(Z)
    # the mask, bit-exactly: every mode outside it is exactly zero
    mask = O.band_mask(m, 0.0, kc)
    assert np.all(got[:, mask == 0] == 0)
    assert np.all(got[:, 0, 0, 0] == 0)
    # input untouched
    assert np.array_equal(du.download(), u.astype(plans.cdtype))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,L,kc", [((64, 32, 128), (1.0, 2.5, 7.0), 20.0),
                                    ((32, 64, 64), (3.0, 0.5, 2.0), 40.0),
                                    ((64, 64, 64), (O.TWO_PI,) * 3, 0.0),
                                    ((32, 32, 32), (O.TWO_PI,) * 3, 0.5)])
def test_uxw_boxes_and_tiny_cutoffs(G, n, L, kc, dtype):
    """Fused u x omega on boxes with unequal, non-2 pi sides (the mask then
    compares non-integer |k|^2 in FP64) and with cutoffs that keep only the
    mean mode or nothing beyond it: against the oracle, the mask bit-exact,
    the reductions within tolerance (exactly zero when nothing is kept)."""
    say(f"[uxw-boxes] n={n} L={L} kc={kc} dtype={np.dtype(dtype).name}")
    g = ogrid(n, L)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n, L)
    plans = G.FieldTransformPlans(grid, dtype)
    u = O.synthetic_field(g, 2005, vector=True, m=m)
    ref, E, Z = O.uxw(u, g, kc, m)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    dn = G.Field(grid, "spectral", 1, dtype)
    e, z = G.uxw_(dn, du, plans, kc, fused=True)
    got = dn.download()
    tol = TOL[dtype]
    mask = O.band_mask(m, 0.0, kc)
    assert np.all(got[:, mask == 0] == 0)
    if np.linalg.norm(ref) > 0:
        assert O.rel_l2(got, ref) < tol
    else:
        assert np.all(got == 0)
    if E > 0:
        assert abs(e - E) <= tol * abs(E) and abs(z - Z) <= tol * abs(Z)
    else:
        assert e == 0.0 and z == 0.0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [(256, 64, 64), (512, 32, 64), (1024, 32, 32)])
def test_x_pass_shapes(G, n, dtype, monkeypatch):
    """The x pass in both tile shapes (row-pair units per workgroup, forced
    with GBG_X_SHAPE at plan creation): bitwise the same N^ and product, and
    the oracle's; the autotuner keeps results unchanged."""
    say(f"[x-shapes] n={n} dtype={np.dtype(dtype).name}")
    g = ogrid(n)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n)
    u = O.synthetic_field(g, 2004, vector=True, m=m)
    a = O.synthetic_field(g, 1005, m=m)
    kc = O.dealias_cutoff(n[0])
    ref = O.uxw(u, g, kc, m)[0]
    pref = O.dealiased_product(a, a, g, kc, m)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    da = G.Field(grid, "spectral", 0, dtype).upload(a)
    outs = []
    for shape in ("0", "1", "tuned"):
        if shape == "tuned":
            plans = G.FieldTransformPlans(grid, dtype)
            G.autotune(plans)
        else:
            monkeypatch.setenv("GBG_X_SHAPE", shape)
            plans = G.FieldTransformPlans(grid, dtype)
            monkeypatch.delenv("GBG_X_SHAPE")
        dn = G.Field(grid, "spectral", 1, dtype)
        G.uxw_(dn, du, plans, kc)
        dp = G.Field(grid, "spectral", 0, dtype)
        G.dealiased_product_(dp, da, da, plans, kc)
        outs.append((dn.download(), dp.download()))
        assert O.rel_l2(outs[-1][0], ref) < TOL[dtype]
        assert O.rel_l2(outs[-1][1], pref) < TOL[dtype]
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [(64, 64, 64), (128, 64, 32)])
def test_zero_tile_fork_points(G, n, dtype, monkeypatch):
    """The zero tiles of the last pass beside each earlier pass on the second
    stream (or in stream order): bitwise the same N^, also with the
    integrating-factor combination writing over its input, and the oracle's."""
    say(f"[zero-fork] n={n} dtype={np.dtype(dtype).name}")
    g = ogrid(n)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    u = O.synthetic_field(g, 2006, vector=True, m=m)
    kc = O.dealias_cutoff(n[0])
    ref = O.uxw(u, g, kc, m)[0]
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    var = "GBG_ZERO_FORK_F64" if dtype == np.float64 else "GBG_ZERO_FORK_F32"
    outs = []
    for k in "01234":
        monkeypatch.setenv(var, k)
        dn = G.Field(grid, "spectral", 1, dtype)
        G.uxw_(dn, du, plans, kc)
        dst = G.Field(grid, "spectral", 1, dtype).upload(u)
        G.uxw_combine_(dst, du, plans, kc, a=dst, delta=0.01, coefficient=-0.002)
        outs.append((dn.download(), dst.download()))
        assert O.rel_l2(outs[-1][0], ref) < TOL[dtype]
    monkeypatch.delenv(var)
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1])


def _fp32_mask_case():
    n, L = (64, 64, 64), (1.3, 1.3, 1.3)
    g = ogrid(n, L)
    m = O.Multipliers.build(g)
    kc = 15 * (2 * math.pi / 1.3)
    rng = np.random.default_rng(7)
    u = rng.standard_normal((3,) + g.spectral_shape) + 1j * rng.standard_normal((3,) + g.spectral_shape)
    ref, E, Z = O.uxw(u, g, kc, m)
    return n, L, g, m, kc, u, ref, E, Z


@pytest.mark.parametrize("fused", [True, False])
def test_uxw_fp32_mask_semantics(G, fused):
    """FP32 u x omega on a 1.3 box with the cutoff on a lattice sphere (15
    lattice units): the output keeps exactly the FP64 predicate's modes, 81
    boundary modes more than the reference's Float32 predicate would (mode
    handling stays in FP64 in the FP32 form)."""
    say(f"[uxw-fp32-mask] 64^3 L=1.3 kc=15 units fused={fused}")
    n, L, g, m, kc, u, ref, E, Z = _fp32_mask_case()
    grid = G.Grid3D(n, L)
    plans = G.FieldTransformPlans(grid, np.float32)
    du = G.Field(grid, "spectral", 1, np.float32).upload(u)
    dn = G.Field(grid, "spectral", 1, np.float32)
    e, z = G.uxw_(dn, du, plans, kc, fused=fused)
    got = dn.download()
    assert O.rel_l2(got, ref) < 1e-5
    kept = np.any(got != 0, axis=0)
    kept[0, 0, 0] = True
    m64 = O.band_mask(m, 0.0, kc).astype(bool)
    m32 = O.band_mask_float32(g, 0.0, kc)
    assert np.array_equal(kept, m64)
    assert int((kept & ~m32).sum()) == 81


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [(32, 32, 32), (64, 64, 256), (64, 64, 512), (32, 32, 1024)])
def test_uxw_wide_offsets(G, n, dtype, monkeypatch):
    """The z-inverse pass of u x omega addresses a component with 32-bit element
    offsets when every component holds fewer than 2^32 modes, and with 64-bit
    ones otherwise.  GBG_WIDE_OFFSETS=1 at plan creation forces the 64-bit form
    at a size that fits: it must match the oracle and the 32-bit form."""
    say(f"[uxw-wide] n={n} dtype={np.dtype(dtype).name}")
    g = ogrid(n)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n)
    narrow = G.FieldTransformPlans(grid, dtype)
    monkeypatch.setenv("GBG_WIDE_OFFSETS", "1")
    wide = G.FieldTransformPlans(grid, dtype)
    monkeypatch.delenv("GBG_WIDE_OFFSETS")
    u = O.synthetic_field(g, 2002, vector=True, m=m)
    kc = O.dealias_cutoff(n[0])
    ref, E, Z = O.uxw(u, g, kc, m)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    tol = TOL[dtype]
    outs = []
    for plans in (narrow, wide):
        dn = G.Field(grid, "spectral", 1, dtype)
        e, z = G.uxw_(dn, du, plans, kc, fused=True)
        got = dn.download()
        assert O.rel_l2(got, ref) < tol
        assert abs(e - E) <= tol * abs(E) and abs(z - Z) <= tol * abs(Z)
        outs.append((got, e, z))
    assert O.rel_l2(outs[0][0], outs[1][0]) < tol
    assert abs(outs[0][1] - outs[1][1]) <= tol * abs(E) and abs(outs[0][2] - outs[1][2]) <= tol * abs(Z)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [(32, 32, 128), (16, 16, 512)])
def test_many_cutoffs(G, n, dtype):
    """A plan caches the output tile list of up to 8 cutoffs; further cutoffs
    run the last z pass over every tile (zeros written in-kernel), persistent
    in FP64 at these lengths.  Both forms, u x omega and the scalar product,
    against the oracle for ten cutoffs on one plan."""
    say(f"[many-cutoffs] n={n} dtype={np.dtype(dtype).name}")
    g = ogrid(n)
    m = O.Multipliers.build(g)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    u = O.synthetic_field(g, 2003, vector=True, m=m)
    a_hat = O.synthetic_field(g, 1003, m=m)
    b_hat = O.synthetic_field(g, 1004, m=m)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    ah = G.Field(grid, "spectral", 0, dtype).upload(a_hat)
    bh = G.Field(grid, "spectral", 0, dtype).upload(b_hat)
    tol = TOL[dtype]
    for kc in (3.5, 4.5, 5.5, 6.5, 7.5, 8.5, 9.5, 10.5, 11.5, 12.5):
        ref, E, Z = O.uxw(u, g, kc, m)
        dn = G.Field(grid, "spectral", 1, dtype)
        e, z = G.uxw_(dn, du, plans, kc, fused=True)
        got = dn.download()
        assert O.rel_l2(got, ref) < tol, kc
        assert abs(e - E) <= tol * abs(E) and abs(z - Z) <= tol * abs(Z)
        mask = O.band_mask(m, 0.0, kc)
        assert np.all(got[:, mask == 0] == 0)
        pref = O.dealiased_product(a_hat, b_hat, g, kc, m)
        ph = G.Field(grid, "spectral", 0, dtype)
        G.dealiased_product_(ph, ah, bh, plans, kc)
        pg = ph.download()
        assert O.rel_l2(pg, pref) < tol, kc
        assert np.all(pg[mask == 0] == 0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_config1_pipeline(G, dtype):
    """64^3 plain method: rfft of a and b, gradient of a, dealiased a*b, irfft."""
    say(f"[config1] dtype={np.dtype(dtype).name}")
    n = (64, 64, 64)
    g = ogrid(n)
    m = O.Multipliers.build(g)
    a_hat = O.synthetic_field(g, 1001, m=m)
    b_hat = O.synthetic_field(g, 1002, m=m)
    a = O.to_physical(a_hat, g)
    b = O.to_physical(b_hat, g)
    kc = O.dealias_cutoff(64)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    tol = TOL[dtype]
    fa = G.Field(grid, "physical", 0, dtype).upload(a)
    fb = G.Field(grid, "physical", 0, dtype).upload(b)
    ah = G.to_spectral(fa, plans)
    bh = G.to_spectral(fb, plans)
    assert O.rel_l2(ah.download(), O.to_spectral(a)) < tol
    assert O.rel_l2(bh.download(), O.to_spectral(b)) < tol
    gh = G.gradient(ah, plans)
    ref_gh = O.gradient(O.to_spectral(a), m)
    assert O.rel_l2(gh.download(), ref_gh) < tol
    gp = G.to_physical(gh, plans)
    assert O.rel_l2(gp.download(), O.to_physical(ref_gh, g)) < tol
    ph = G.Field(grid, "spectral", 0, dtype)
    G.dealiased_product_(ph, ah, bh, plans, kc)
    ref_ph = O.dealiased_product(O.to_spectral(a), O.to_spectral(b), g, kc, m)
    got = ph.download()
    assert O.rel_l2(got, ref_ph) < tol
    mask = O.band_mask(m, 0.0, kc)
    assert np.all(got[mask == 0] == 0)
    pp = G.to_physical(ph, plans)
    assert O.rel_l2(pp.download(), O.to_physical(ref_ph, g)) < tol


TWO_PI3 = (O.TWO_PI,) * 3
DP_CASES = [((64, 64, 64), None, TWO_PI3), ((32, 64, 128), None, TWO_PI3),
            ((128, 32, 64), None, TWO_PI3), ((64, 64, 64), 1e9, TWO_PI3),
            ((64, 64, 64), 3.5, TWO_PI3), ((64, 32, 128), 20.0, (1.0, 2.5, 7.0)),
            ((256, 256, 256), None, TWO_PI3), ((48, 40, 36), None, TWO_PI3)]


@pytest.mark.parametrize("n,kc,L", DP_CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dealiased_product_parity(G, n, kc, L, dtype, monkeypatch):
    """p^ = M F(F^-1(M a^) F^-1(M b^)) against the oracle: the fused five-pass
    form on power-of-two grids (cubic and not, an anisotropic box; the 2/3
    cutoff, no truncation, a cutoff of 3.5), the composed form on the others and on a
    plan restricted to the mixed-radix passes; both agree with each other; the
    removed modes are exactly zero; a^ = b^ (the square) is allowed."""
    say(f"[dealiased product] n={n} kc={kc} L={L} dtype={np.dtype(dtype).name}")
    g = ogrid(n, L)
    m = O.Multipliers.build(g)
    kc = O.dealias_cutoff(n[0]) if kc is None else kc
    a_hat = O.synthetic_field(g, 1001, m=m)
    b_hat = O.synthetic_field(g, 1002, m=m)
    grid = G.Grid3D(n, L)
    plans = G.FieldTransformPlans(grid, dtype)
    monkeypatch.setenv("GBG_GENERIC_FFT", "1")
    composed = G.FieldTransformPlans(grid, dtype)
    monkeypatch.delenv("GBG_GENERIC_FFT")
    tol = TOL[dtype]
    ah = G.Field(grid, "spectral", 0, dtype).upload(a_hat)
    bh = G.Field(grid, "spectral", 0, dtype).upload(b_hat)
    ref = O.dealiased_product(a_hat, b_hat, g, kc, m)
    mask = O.band_mask(m, 0.0, kc)
    outs = []
    for pl in (plans, composed):
        ph = G.Field(grid, "spectral", 0, dtype)
        G.dealiased_product_(ph, ah, bh, pl, kc)
        got = ph.download()
        assert O.rel_l2(got, ref) < tol
        assert np.all(got[mask == 0] == 0)
        outs.append(got)
    assert O.rel_l2(outs[0], outs[1]) < 2 * tol
    # the square, inputs aliased
    ph = G.Field(grid, "spectral", 0, dtype)
    G.dealiased_product_(ph, ah, ah, plans, kc)
    assert O.rel_l2(ph.download(), O.dealiased_product(a_hat, a_hat, g, kc, m)) < tol
    # inputs untouched
    assert np.array_equal(ah.download(), a_hat.astype(plans.cdtype))
    assert np.array_equal(bh.download(), b_hat.astype(plans.cdtype))


@pytest.mark.parametrize("fused", [False, True])
def test_uxw_golden_16(G, fused):
    """The committed 16^3 fixture: N^, E and Z of the seeded field."""
    import os
    say(f"[golden uxw 16] fused={fused}")
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "uxw_16.npz"))
    grid = G.Grid3D((16, 16, 16))
    plans = G.FieldTransformPlans(grid)
    du = G.Field(grid, "spectral", 1).upload(d["u_hat"])
    dn = G.Field(grid, "spectral", 1)
    e, z = G.uxw_(dn, du, plans, O.dealias_cutoff(16), fused=fused)
    assert O.rel_l2(dn.download(), d["n_hat"]) < 1e-12
    assert abs(e - float(d["energy"])) < 1e-12 * float(d["energy"])
    assert abs(z - float(d["enstrophy"])) < 1e-12 * float(d["enstrophy"])


def test_ops_golden_8x7x6(G):
    import os
    say("[golden ops 8x7x6]")
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ops_8x7x6.npz"))
    grid = G.Grid3D((8, 7, 6))
    plans = G.FieldTransformPlans(grid)
    phys = G.Field(grid, "physical", 1).upload(d["phys"])
    spec = G.to_spectral(phys, plans)
    assert O.rel_l2(spec.download(), d["spec"]) < 1e-12
    assert O.rel_l2(G.curl(spec, plans).download(), d["curl"]) < 1e-12
    pr = G.Field(grid, "spectral", 1)
    G.helmholtz_projection_(pr, spec, plans)
    assert O.rel_l2(pr.download(), d["proj"]) < 1e-12
    assert abs(G.inner(spec, spec, plans) - float(d["inner"])) < 1e-12 * float(d["inner"])


def test_refusals(G):
    say("[refusals]")
    grid = G.Grid3D((8, 8, 8))
    plans = G.FieldTransformPlans(grid)
    with pytest.raises(G.UnsupportedError):     # chirp-z length 4320 > one FP64 LDS pencil
        G.FieldTransformPlans(G.Grid3D((2053, 8, 8)))
    with pytest.raises(G.ArgumentError):
        G.FieldTransformPlans(grid, np.int32)
    s = G.Field(grid, "spectral", 1)
    with pytest.raises(G.AliasError):
        G.to_physical_(G.Field(grid, "physical", 1), s, plans, scratch=s)
    with pytest.raises(G.DimensionMismatch):
        G.inner(G.Field(grid, "spectral", 0), s, plans)
    with pytest.raises(G.ArgumentError):
        G.spectral_band(plans, 3.0, 1.0)
    other = G.FieldTransformPlans(G.Grid3D((8, 8, 4)))
    with pytest.raises(G.ArgumentError):
        G.curl_(s, s, other)
    with pytest.raises(G.UnsupportedError):
        G.curl_(G.Field(grid, "physical", 1), s, plans)
    # the fused pipelines refuse an output that overlaps an input
    with pytest.raises(G.AliasError):
        G.uxw_(s, s, plans)
    a = G.Field(grid, "spectral", 0)
    with pytest.raises(G.AliasError):
        G.dealiased_product_(a, a, G.Field(grid, "spectral", 0), plans)
    # to_physical! checks its scratch against the plans (src/Fields.jl:1601)
    with pytest.raises(G.ArgumentError):
        G.to_physical_(G.Field(grid, "physical", 1), s, plans,
                       scratch=G.Field(G.Grid3D((8, 8, 4)), "spectral", 1))
    with pytest.raises(G.ArgumentError):
        G.to_physical_(G.Field(grid, "physical", 1), s, plans,
                       scratch=G.Field(grid, "spectral", 1, np.float32))


def test_c_abi_overlap_refusals(G):
    """The C ABI refuses overlapping buffers on its own (no Python checks in
    between): a to_physical scratch over the physical destination, and batch
    items whose outputs overlap the input or each other."""
    import ctypes as C
    say("[refusals-c]")
    L = G._lib.load()
    grid = G.Grid3D((8, 8, 8))
    plans = G.FieldTransformPlans(grid)
    big = G.Field(grid, "spectral", 2)          # room for a vector field and more
    spec = G.Field(grid, "spectral", 1)
    st = L.gbg_to_physical(plans.handle, big.data, spec.data, 3, big.data + 64, None)
    assert st == 4 and b"scratch overlaps the physical destination" in L.gbg_last_error()
    b = G.Batch(nstreams=1, autotune=False)
    n = (C.c_int * 3)(8, 8, 8)
    Lb = (C.c_double * 3)(*grid.domain_size)
    cb = 5 * 8 * 8 * 16
    st = L.gbg_batch_add(b.handle, 0, n, Lb, 8, big.data, big.data + 16, big.data + 4 * cb, 2.0)
    assert st == 4                              # out0 overlaps the input
    st = L.gbg_batch_add(b.handle, 0, n, Lb, 8, big.data, big.data + cb, big.data + 16, 2.0)
    assert st == 4                              # out1 overlaps the input
    st = L.gbg_batch_add(b.handle, 0, n, Lb, 8, spec.data, big.data, big.data + cb, 2.0)
    assert st == 4                              # out1 inside out0
    b.close()
