"""The batched, graph-captured form and the autotuner, against the oracle.

The graph is replayed several times with the program's own eager work and
synchronisations between replays; every replay is checked against the
eager result (bitwise) and against the oracle."""

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


ITEMS = [("s", (32, 32, 32)), ("v", (32, 32, 32)), ("s", (48, 48, 48)), ("v", (64, 64, 128)),
         ("v", (48, 32, 40)), ("s", (128, 64, 32)), ("v", (32, 32, 32))]


def oracle_item(kind, g, f, kc):
    m = O.Multipliers.build(g)
    if kind == "s":
        return O.gradient(f, m), O.dealiased_product(f, f, g, kc, m)
    grad = np.stack([np.stack([(1j * a) * f[i] for i in range(3)]) for a in (m.a1(), m.a2(), m.a3())])
    return grad, O.uxw(f, g, kc, m)[0]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_graph_replays(G, dtype):
    say(f"[batch] dtype={np.dtype(dtype).name} items={ITEMS}")
    batch = G.Batch(nstreams=3, autotune=True)
    refs, outs, fields = [], [], []
    for i, (kind, n) in enumerate(ITEMS):
        g = O.Grid3D(n)
        f = O.synthetic_field(g, 4001 + 1 + i, vector=(kind == "v"))
        kc = O.dealias_cutoff(min(n))
        grid = G.Grid3D(n)
        d = G.Field(grid, "spectral", 1 if kind == "v" else 0, dtype).upload(f)
        o0, o1 = (batch.add_vector if kind == "v" else batch.add_scalar)(d, kc)
        refs.append(oracle_item(kind, g, f.astype(np.dtype(d.npdtype)).astype(np.complex128), kc))
        outs.append((o0, o1))
        fields.append(d)
    batch.instantiate()
    eager = [(o0.download(), o1.download()) for o0, o1 in outs]
    for (e0, e1), (r0, r1) in zip(eager, refs):
        assert O.rel_l2(e0, r0) < TOL[dtype]
        assert O.rel_l2(e1, r1) < TOL[dtype]
    # unrelated eager work between replays
    gx = G.Grid3D((64, 64, 64))
    px = G.FieldTransformPlans(gx, dtype)
    ux = G.Field(gx, "spectral", 1, dtype).upload(O.synthetic_field(O.Grid3D((64, 64, 64)), 77, True))
    nx = G.Field(gx, "spectral", 1, dtype)
    for rep in range(3):
        for o0, o1 in outs:
            o0.fill_zero()
            o1.fill_zero()
        batch.launch()
        G.uxw_(nx, ux, px)                      # eager call, synchronises the default stream
        for k, ((o0, o1), (e0, e1)) in enumerate(zip(outs, eager)):
            g0, g1 = o0.download(), o1.download()
            assert np.array_equal(g0, e0), f"replay {rep} item {k} out0"
            assert np.array_equal(g1, e1), f"replay {rep} item {k} out1"
    batch.close()


def test_autotune_keeps_results(G):
    say("[autotune] 48x96x64 f64")
    n = (48, 96, 64)
    g = O.Grid3D(n)
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid)
    f = np.random.default_rng(5).standard_normal((3,) + g.physical_shape)
    ms = G.autotune(plans)
    assert ms > 0
    phys = G.Field(grid, "physical", 1).upload(f)
    spec = G.to_spectral(phys, plans)
    assert O.rel_l2(spec.download(), O.to_spectral(f)) < 1e-12
    assert O.rel_l2(G.to_physical(spec, plans).download(), f) < 1e-12
