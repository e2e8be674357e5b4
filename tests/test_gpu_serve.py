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
