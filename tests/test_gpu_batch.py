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


BATCH_SHAPES = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 128, 128),
                (128, 128, 256), (256, 64, 64), (256, 256, 256)]


@pytest.fixture(scope="module")
def batch96():
    """The batched workload: 96 fields, grids drawn from the eight listed
    shapes and scalar/vector 50/50 by master seed 4001, field i from seed
    4001 + 1 + i; the oracle's dealiased product (scalar) or u x omega (vector)
    of each, on the FP64 input."""
    rng = np.random.Generator(np.random.PCG64(4001))
    shapes = [BATCH_SHAPES[i] for i in rng.integers(0, len(BATCH_SHAPES), size=96)]
    kinds = ["v" if k else "s" for k in rng.integers(0, 2, size=96)]
    items = []
    for i, (n, kind) in enumerate(zip(shapes, kinds)):
        g = O.Grid3D(n)
        m = O.Multipliers.build(g)
        f = O.synthetic_field(g, 4001 + 1 + i, vector=(kind == "v"), m=m)
        kc = O.dealias_cutoff(min(n))
        ref = O.dealiased_product(f, f, g, kc, m) if kind == "s" else O.uxw(f, g, kc, m)[0]
        items.append((kind, n, f, kc, ref))
    say(f"[batch96] mix {[(s, shapes.count(s)) for s in BATCH_SHAPES]}, vectors "
        f"{kinds.count('v')}")
    return items


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_96_uneven_against_oracle(G, batch96, dtype):
    """Every item of the 96-field graph, after a replay, against the oracle."""
    say(f"[batch96] dtype={np.dtype(dtype).name}")
    batch = G.Batch(nstreams=4, autotune=True)
    outs = []
    for kind, n, f, kc, _ in batch96:
        d = G.Field(G.Grid3D(n), "spectral", 1 if kind == "v" else 0, dtype).upload(f)
        outs.append((batch.add_vector if kind == "v" else batch.add_scalar)(d, kc))
    batch.instantiate()
    for o0, o1 in outs:
        o0.fill_zero()
        o1.fill_zero()
    batch.launch()
    tol = TOL[dtype]
    for k, ((kind, n, f, kc, ref), (o0, o1)) in enumerate(zip(batch96, outs)):
        m = O.Multipliers.build(O.Grid3D(n))
        if kind == "s":
            grad = O.gradient(f, m)
        else:
            grad = np.stack([np.stack([(1j * a) * f[i] for i in range(3)])
                             for a in (m.a1(), m.a2(), m.a3())])
        e0 = O.rel_l2(o0.download(), grad)
        e1 = O.rel_l2(o1.download(), ref)
        assert e0 < tol and e1 < tol, f"item {k} {kind} {n}: {e0:.3e} {e1:.3e}"
    batch.close()
