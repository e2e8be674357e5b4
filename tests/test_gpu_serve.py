"""The serving loop: results alone equal results in a concurrent mix (bitwise),
both match the oracle, admission refuses jobs beyond the budget, and a tight
budget evicts idle plans without changing results."""

import sys
import time

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


GRIDS = [((32, 32, 32), np.float64), ((48, 40, 36), np.float32), ((64, 64, 64), np.float64),
         ((96, 32, 64), np.float32), ((64, 64, 64), np.float32)]


def make_requests(G):
    """(kind, out, in0, in1, oracle) for every job kind on every grid."""
    reqs = []
    for gi, (n, dt) in enumerate(GRIDS):
        g = O.Grid3D(n)
        m = O.Multipliers.build(g)
        grid = G.Grid3D(n)
        u = O.synthetic_field(g, 5001 + gi, vector=True, m=m)
        uq = u.astype(np.complex64 if dt == np.float32 else np.complex128).astype(np.complex128)
        phys = O.to_physical(uq, g)
        kc = O.dealias_cutoff(min(n))
        du = G.Field(grid, "spectral", 1, dt).upload(u)
        dp = G.Field(grid, "physical", 1, dt).upload(phys)
        da = G.Field(grid, "spectral", 0, dt).upload(u[0])
        db = G.Field(grid, "spectral", 0, dt).upload(u[1])
        S = G.Server

        def out(rep, order, grid=grid, dt=dt):
            return lambda: G.Field(grid, rep, order, dt)

        reqs.append((S.TO_SPECTRAL, out("spectral", 1), dp, None, kc,
                     O.to_spectral(phys.astype(dt).astype(np.float64))))
        reqs.append((S.TO_PHYSICAL, out("physical", 1), du, None, kc, phys))
        reqs.append((S.CURL, out("spectral", 1), du, None, kc, O.curl(uq, m)))
        reqs.append((S.INVERSE_LAPLACIAN, out("spectral", 1), du, None, kc, O.inverse_laplacian(uq, m)))
        reqs.append((S.DEALIASED_PRODUCT, out("spectral", 0), da, db, kc,
                     O.dealiased_product(uq[0], uq[1], g, kc, m)))
        reqs.append((S.UXW, out("spectral", 1), du, None, kc, O.uxw(uq, g, kc, m)[0]))
    return reqs


def test_serve_alone_equals_mix(G):
    say("[serve] alone vs mix")
    reqs = make_requests(G)
    alone = []
    srv = G.Server(hbm_budget_bytes=64 << 30, slots=1)
    for kind, mk, i0, i1, kc, ref in reqs:
        out = mk()
        t = srv.submit(kind, out, i0, i1, kc)
        srv.wait(t)
        got = out.download()
        tol = TOL[np.float64 if i0.dtype == np.float64 else np.float32]
        assert O.rel_l2(got, ref) < tol, f"kind {kind} grid {i0.grid}"
        alone.append(got)
    srv.close()
    # the mix: everything three times over, submitted at once to 4 slots
    srv = G.Server(hbm_budget_bytes=64 << 30, slots=4)
    subs = []
    for rep in range(3):
        order = np.random.default_rng(rep).permutation(len(reqs))
        for k in order:
            kind, mk, i0, i1, kc, _ = reqs[k]
            out = mk()
            subs.append((k, out, srv.submit(kind, out, i0, i1, kc)))
    for k, out, t in subs:
        lat = srv.wait(t)
        assert lat > 0
        assert np.array_equal(out.download(), alone[k]), f"request {k} differs in the mix"
    st = srv.stats()
    assert st["completed"] == len(subs) and st["p99_ms"] >= st["p50_ms"] > 0
    # non-blocking completion: poll each ticket until it reports done
    subs = []
    for k in range(len(reqs)):
        kind, mk, i0, i1, kc, _ = reqs[k]
        out = mk()
        subs.append((k, out, srv.submit(kind, out, i0, i1, kc)))
    for k, out, t in subs:
        lat = None
        for _ in range(200000):
            lat = srv.poll(t)
            if lat is not None:
                break
            time.sleep(0.0005)
        assert lat is not None and lat > 0, f"request {k} never completed"
        assert np.array_equal(out.download(), alone[k])
    with pytest.raises(G.ArgumentError):
        srv.poll(subs[0][2])          # the ticket is retired once reported done
    srv.close()


def test_serve_admission_and_eviction(G):
    say("[serve] admission and eviction")
    S = G.Server
    n = (64, 64, 64)
    grid = G.Grid3D(n)
    g = O.Grid3D(n)
    u = O.synthetic_field(g, 9, vector=True)
    du = G.Field(grid, "spectral", 1).upload(u)
    small = G.Server(hbm_budget_bytes=8 << 20, slots=2)
    with pytest.raises(G.GbgError):
        small.submit(S.UXW, G.Field(grid, "spectral", 1), du)     # needs more than 8 MiB
    small.close()
    # a budget for about two resident plans across four grids: plans are
    # evicted and recreated, results stay exact
    sizes = [(32, 32, 32), (32, 32, 64), (64, 32, 32), (32, 64, 32)]
    fields, refs = [], []
    for i, s in enumerate(sizes):
        gg = O.Grid3D(s)
        f = O.synthetic_field(gg, 100 + i, vector=True)
        fields.append(G.Field(G.Grid3D(s), "spectral", 1).upload(f))
        refs.append(O.uxw(f, gg, O.dealias_cutoff(32))[0])
    budget = 3 * (6 * 17 * 32 * 64 * 16 + 6 * 32 * 32 * 64 * 8) // 2
    srv = G.Server(hbm_budget_bytes=budget, slots=2)
    outs = []
    for rep in range(3):
        for i, f in enumerate(fields):
            o = G.Field(f.grid, "spectral", 1)
            outs.append((i, o, srv.submit(S.UXW, o, f, kc=O.dealias_cutoff(32))))
    for i, o, t in outs:
        srv.wait(t)
        assert O.rel_l2(o.download(), refs[i]) < 1e-12
        b, k = srv.resident()
        assert b <= budget
    srv.close()


def test_serve_submit_refusals(G):
    """Server.submit applies the refusals of the matching standalone call;
    the C ABI refuses an output that overlaps the input by itself."""
    import ctypes as C
    say("[serve] submit refusals")
    S = G.Server
    n = (16, 16, 16)
    grid = G.Grid3D(n)
    spec = G.Field(grid, "spectral", 1)
    phys = G.Field(grid, "physical", 1)
    srv = G.Server(hbm_budget_bytes=1 << 30, slots=1)
    with pytest.raises(G.UnsupportedError):
        srv.submit(S.TO_SPECTRAL, G.Field(grid, "spectral", 1), spec)        # needs a physical input
    with pytest.raises(G.UnsupportedError):
        srv.submit(S.TO_PHYSICAL, G.Field(grid, "spectral", 1), spec)       # needs a physical output
    with pytest.raises(G.ArgumentError):
        srv.submit(S.TO_SPECTRAL, G.Field(G.Grid3D((16, 16, 8)), "spectral", 1), phys)
    with pytest.raises(G.ArgumentError):
        srv.submit(S.CURL, G.Field(grid, "spectral", 1, np.float32), spec)
    with pytest.raises(G.DimensionMismatch):
        srv.submit(S.UXW, G.Field(grid, "spectral", 0), G.Field(grid, "spectral", 0))
    with pytest.raises(G.DimensionMismatch):
        srv.submit(S.TO_SPECTRAL, G.Field(grid, "spectral", 0), phys)
    with pytest.raises(G.AliasError):
        srv.submit(S.UXW, spec, spec)
    a = G.Field(grid, "spectral", 0)
    with pytest.raises(G.ArgumentError):
        srv.submit(S.DEALIASED_PRODUCT, G.Field(grid, "spectral", 0), a)    # second input missing
    with pytest.raises(G.AliasError):
        srv.submit(S.DEALIASED_PRODUCT, a, G.Field(grid, "spectral", 0), a)
    # the in-place curl is allowed, as curl! is alias-safe
    t = srv.submit(S.CURL, spec, spec)
    srv.wait(t)
    # C ABI: an output overlapping the input
    L = G._lib.load()
    j = G._lib.GbgJob()
    j.kind = S.UXW
    j.n = (C.c_int * 3)(*n)
    j.L = (C.c_double * 3)(*grid.domain_size)
    j.elem_bytes = 8
    j.ncomp = 3
    j.in0 = spec.data
    j.in1 = None
    j.out = spec.data + 16
    j.kc = 5.0
    tk = C.c_ulonglong(0)
    assert L.gbg_server_submit(srv.handle, C.byref(j), C.byref(tk)) == 4
    srv.close()


def test_serve_cached_plan_never_stalls(G):
    """A job whose grid already has a cached plan is admitted against the
    plan's real size: with a budget that holds the u x omega plan but not
    more, a curl then a u x omega job on that grid both finish; with a budget
    that cannot hold the plan beside the work area the u x omega job is
    refused at submission (it never waits for a completion that cannot come).
    The grid's long axes give the plan more than 64 KiB of tables."""
    say("[serve] cached plan admission")
    S = G.Server
    n = (1024, 1024, 32)
    grid = G.Grid3D(n)
    g = O.Grid3D(n)
    u = O.synthetic_field(g, 11, vector=True)
    kc = O.dealias_cutoff(32)
    du = G.Field(grid, "spectral", 1).upload(u)
    outs = [G.Field(grid, "spectral", 1) for _ in range(2)]

    def finish(srv, t, limit=60.0):
        t0 = time.time()
        while time.time() - t0 < limit:
            lat = srv.poll(t)
            if lat is not None:
                return lat
            time.sleep(0.001)
        raise AssertionError("job never completed")

    def resident_after(kind):
        srv = G.Server(hbm_budget_bytes=64 << 30, slots=1)
        finish(srv, srv.submit(kind, outs[0], du, kc=kc))
        b, _ = srv.resident()
        srv.close()
        return b

    r_curl, r_uxw = resident_after(S.CURL), resident_after(S.UXW)
    assert r_curl > 64 * 1024 and r_uxw > r_curl
    srv = G.Server(hbm_budget_bytes=r_uxw + (1 << 20), slots=1)
    finish(srv, srv.submit(S.CURL, outs[0], du, kc=kc))
    finish(srv, srv.submit(S.UXW, outs[1], du, kc=kc))
    assert O.rel_l2(outs[1].download(), O.uxw(u, g, kc)[0]) < 1e-12
    srv.close()
    tight = G.Server(hbm_budget_bytes=(r_uxw - r_curl) + 64 * 1024, slots=1)
    finish(tight, tight.submit(S.CURL, outs[0], du, kc=kc))
    with pytest.raises(G.GbgError):
        finish(tight, tight.submit(S.UXW, outs[1], du, kc=kc))
    tight.close()

