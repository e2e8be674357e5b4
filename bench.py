#!/usr/bin/env python3
"""Benchmark: 512^3 dealiased u x omega fields/s on one MI355X.

One step = one dealiased rotational nonlinear term N^ = Q M P F(u x w) of a
512^3 Float64 solenoidal velocity (seeded k^-5/3 spectrum), computed by the
fused pass schedule of libgbg with the input already resident in HBM.
Prints ONE JSON line (rank 0).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 512]
                  [--dtype f64|f32] [--no-cpu-baseline] [--no-fp32]
"""

from __future__ import annotations

import argparse
import ctypes as C
import importlib.util
import json
import math
import os
import platform
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def load_package():
    name = "gbg_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg_dir = os.path.join(REPO, "gbgperiodicfields.jl_amd")
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(pkg_dir, "__init__.py"), submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


class HipEvents:
    """hipEvent timing on the stream the kernels are launched on."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]

    def create(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e.value

    def record(self, ev, stream):
        assert self.hip.hipEventRecord(ev, stream) == 0

    def elapsed(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return float(ms.value)


def kept_ranges(n, kc, L=None):
    """Index ranges the fused schedule keeps for cutoff kc, from FP64
    squared-wavenumber tables and the predicate (k1^2 + k2^2) + k3^2 <= kc^2
    the library applies: nhk kept kx (x folded), nyc kept ky, disk = kept
    (kx, ky) pencils, ball = kept modes."""
    L = L or (2.0 * math.pi,) * 3

    def sq(m, length, half):
        unit = (2.0 * math.pi) / length
        i = np.arange(m // 2 + 1 if half else m)
        f = i if half else np.where(i < (m + 1) // 2, i, i - m)
        k = unit * f.astype(np.float64)
        return k * k

    s1, s2, s3 = sq(n[0], L[0], True), sq(n[1], L[1], False), sq(n[2], L[2], False)
    hi2 = float(kc) * float(kc)
    nhk = max(1, int(np.count_nonzero(s1 <= hi2)))      # s1 grows with kx
    plane = s1[None, :] + s2[:, None]                 # (Ny, Nh)
    disk = int(np.count_nonzero(plane[:, :nhk] <= hi2))
    ball = sum(int(np.count_nonzero(plane + z <= hi2)) for z in s3)
    return {"nhk": nhk, "nyc": int(np.count_nonzero(s2 <= hi2)), "disk": disk, "ball": ball}


def uxw_stage_bytes(n, kc, elem):
    """Algorithmic bytes of each pass of the fused u x omega schedule: what the
    pass must read and write given the index ranges the 2/3 mask keeps
    (complex elements of 2 elem bytes).  With row = Nz Ny nhk (full-y rows of
    kept kx) and disk = Nz x (kept (kx, ky) pencils):
      K1  read u^ on the kept ball (3 comps), write u and the two kz terms
          of w on the disk (5; at Ny > 512 u and w, 6)
      K2  read those on the disk, write 6 full-y rows (u, w)
      K3  read 6 rows, write 3 rows (u x w transformed back)
      K4  read 3 rows, write 3 on the disk
      K5  read 3 on the disk, write the whole output (3 M, zeros included)."""
    r = kept_ranges(n, kc)
    sc = 2 * elem
    M = (n[0] // 2 + 1) * n[1] * n[2]
    row = n[2] * n[1] * r["nhk"]
    disk = n[2] * r["disk"]
    nz1 = 5 if n[1] <= 512 else 6    # fields from the z-inverse to the y-inverse
    return {"K1_z_inverse_curl": (3 * r["ball"] + nz1 * disk) * sc,
            "K2_y_inverse": (nz1 * disk + 6 * row) * sc,
            "K3_x_product": 9 * row * sc,
            "K4_y_forward": (3 * row + 3 * disk) * sc,
            "K5_z_forward_project": (3 * disk + 3 * M) * sc}


def product_bytes(n, kc, elem):
    """Algorithmic bytes of the fused dealiased scalar product, as above with
    two fields in and one out: P1 2 ball + 2 disk, P2 2 disk + 2 rows,
    P3 3 rows, P4 rows + disk, P5 disk + M."""
    r = kept_ranges(n, kc)
    M = (n[0] // 2 + 1) * n[1] * n[2]
    row = n[2] * n[1] * r["nhk"]
    disk = n[2] * r["disk"]
    return (2 * r["ball"] + 2 * disk + 2 * disk + 2 * row + 3 * row + row + disk + disk + M) * 2 * elem


def uxw_bytes_unpruned(n, elem):
    """42 component-sized transfers of M = (Nx/2+1) Ny Nz complex elements: the
    five passes as if nothing were masked (an equivalent-rate label only)."""
    M = (n[0] // 2 + 1) * n[1] * n[2]
    return 42 * M * 2 * elem


KERNEL_OF_STAGE = {"K1_z_inverse_curl": ("k_uxw_z_inverse<{t}, {n},",),
                   "K2_y_inverse": ("k_uxw_y<{t}, {n}, 1, 6",),
                   "K3_x_product": ("k_uxw_x<{t}, {n},",),
                   "K4_y_forward": ("k_uxw_y<{t}, {n}, -1, 3",),
                   "K5_z_forward_project": ("k_uxw_zero_tiles<{t}>", "k_uxw_z_forward<{t}, {n}, 3")}


def measured_traffic(stage, n, dtype):
    """HBM bytes per launch of a pass, from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (copy-kernel calibrated), or None if none match.  Each
    prefix of the pass names one kernel of the profiled run (its template
    arguments up to the ones that name the shape)."""
    path = os.path.join(REPO, "profiles", f"uxw_{n[0]}_hbm_traffic.json")
    if not os.path.exists(path) or len(set(n)) != 1:
        return None
    with open(path) as fh:
        data = json.load(fh)
    t = "double" if dtype == np.float64 else "float"
    kernels = data.get("kernels", {})
    total = 0
    for pre in KERNEL_OF_STAGE[stage]:
        pre = pre.format(t=t, n=n[0])
        hits = [k for k in kernels if k.startswith(pre) and
                (pre[-1] in ",>" or k[len(pre):len(pre) + 1] in (",", ">"))]
        if len(hits) != 1:
            return None
        total += int(kernels[hits[0]]["hbm_bytes"])
    return total


def run_gpu(G, n, dtype, steps, warmup, stream, ev, seed=2001):
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    t0 = time.time()
    from gbg_amd import workload  # noqa: E402
    u = workload.synthetic_field(n, seed, vector=True)
    gen_s = time.time() - t0
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    del u
    dn = G.Field(grid, "spectral", 1, dtype)
    kc = G.dealias_cutoff(n[0])
    for _ in range(warmup):
        G.uxw_(dn, du, plans, kc, reductions=False, stream=stream)
    stream.synchronize()
    evs = [ev.create() for _ in range(steps + 1)]
    t0 = time.perf_counter()
    ev.record(evs[0], stream.handle)
    for i in range(steps):
        G.uxw_(dn, du, plans, kc, reductions=False, stream=stream)
        ev.record(evs[i + 1], stream.handle)
    stream.synchronize()
    wall = time.perf_counter() - t0
    per = [ev.elapsed(evs[i], evs[i + 1]) for i in range(steps)]
    total_ms = ev.elapsed(evs[0], evs[-1])
    # per-pass durations: HIP events the library records between its kernels
    # on the same stream, averaged over 16 terms
    plans.set_profiling(True)
    runs = []
    for _ in range(16):
        G.uxw_(dn, du, plans, kc, reductions=False, stream=stream)
        runs.append(plans.stage_times())
    plans.set_profiling(False)
    stages = [float(np.mean(col)) for col in zip(*runs)]
    E, Z = G.uxw_(dn, du, plans, kc, reductions=True, stream=stream)
    return dict(plans=plans, du=du, dn=dn, per=per, total_ms=total_ms, wall=wall, stages=stages,
                E=E, Z=Z, gen_s=gen_s, kc=kc)


BATCH_SHAPES = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 128, 128),
                (128, 128, 256), (256, 64, 64), (256, 256, 256)]


def run_batch(G, stream, ev, steps, warmup, nitems=96, seed=4001, nstreams=4):
    """Batched form: nitems seeded fields of uneven grids, each gradient plus
    dealiased product (scalar) or gradient plus u x w (vector), one graph
    launch per batch, FP64 and FP32."""
    from gbg_amd import workload
    rng = np.random.Generator(np.random.PCG64(seed))
    shapes = [BATCH_SHAPES[i] for i in rng.integers(0, len(BATCH_SHAPES), size=nitems)]
    kinds = rng.integers(0, 2, size=nitems)
    specs = []
    t0 = time.time()
    for i, (n, k) in enumerate(zip(shapes, kinds)):
        specs.append(workload.synthetic_field(n, seed + 1 + i, vector=bool(k)))
    gen_s = time.time() - t0
    res = {"items": nitems, "nstreams": nstreams, "generation_s": round(gen_s, 1),
           "mix": {f"{n[0]}x{n[1]}x{n[2]}": int(sum(1 for s in shapes if s == n)) for n in BATCH_SHAPES},
           "vector_items": int(kinds.sum())}
    for dtype, tag in ((np.float64, "f64"), (np.float32, "f32")):
        batch = G.Batch(nstreams=nstreams, autotune=True)
        alg = 0
        for (n, k, f) in zip(shapes, kinds, specs):
            grid = G.Grid3D(n)
            d = G.Field(grid, "spectral", int(k), dtype).upload(f)
            (batch.add_vector if k else batch.add_scalar)(d, G.dealias_cutoff(min(n)))
            # gradient (4 M or 12 M component transfers) plus the pruned
            # u x omega or dealiased product passes
            es = np.dtype(dtype).itemsize
            M = (n[0] // 2 + 1) * n[1] * n[2]
            kcb = G.dealias_cutoff(min(n))
            alg += (12 * M * 2 * es + sum(uxw_stage_bytes(n, kcb, es).values()) if k
                    else 4 * M * 2 * es + product_bytes(n, kcb, es))
        t0 = time.time()
        batch.instantiate(stream)
        stream.synchronize()
        inst_s = time.time() - t0
        for _ in range(warmup):
            batch.launch(stream)
        stream.synchronize()
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, stream.handle)
        for _ in range(steps):
            batch.launch(stream)
        ev.record(e1, stream.handle)
        ms = ev.elapsed(e0, e1) / steps
        res[tag] = {"ms_per_batch": round(ms, 3), "fields_per_s": round(nitems / (ms * 1e-3), 1),
                    "bytes_alg": int(alg), "achieved_GBs": round(alg / (ms * 1e-3) / 1e9, 1),
                    "instantiate_s": round(inst_s, 2)}
        print(f"[batch] {tag}: {ms:.3f} ms per batch of {nitems}", file=sys.stderr, flush=True)
        batch.close()
    return res


def run_hbm(G, stream, ev, steps, nbytes=1 << 30):
    """Streaming rates the device sustains, for reading the pass rooflines:
    the runtime's device-to-device copy kernel (nbytes read + nbytes written)
    and its fill kernel (nbytes written), timed with events on the stream."""
    lib = G.load()
    src, dst = C.c_void_p(), C.c_void_p()
    G.check(lib.gbg_malloc(C.byref(src), nbytes))
    G.check(lib.gbg_malloc(C.byref(dst), nbytes))
    res = {"bytes": nbytes}
    try:
        for name, call, moved in (
                ("fill", lambda: lib.gbg_memset(dst.value, 1, nbytes, stream.handle), nbytes),
                ("copy", lambda: lib.gbg_memcpy_d2d(dst.value, src.value, nbytes, stream.handle),
                 2 * nbytes)):
            for _ in range(2):
                G.check(call())
            e0, e1 = ev.create(), ev.create()
            ev.record(e0, stream.handle)
            for _ in range(steps):
                G.check(call())
            ev.record(e1, stream.handle)
            ms = ev.elapsed(e0, e1) / steps
            res[name] = {"ms": round(ms, 4), "GBs": round(moved / (ms * 1e-3) / 1e9, 1)}
            print(f"[hbm] {name}: {res[name]}", file=sys.stderr, flush=True)
    finally:
        lib.gbg_free(src.value)
        lib.gbg_free(dst.value)
    return res


def run_ops(G, stream, ev, steps, warmup, n=(512, 512, 512), dtype=np.float64):
    """The operators either side of the nonlinear term at the headline size:
    the exponential kernel on a vector field (single- and two-operand forms),
    the radial spectrum and the reality projection, each with its algorithmic
    bytes (component = M complex elements) and fraction of the HBM peak."""
    from gbg_amd import workload
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    u = workload.synthetic_field(n, 2001, vector=True)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    dv = G.Field(grid, "spectral", 1, dtype).upload(u)
    dw = G.Field(grid, "spectral", 1, dtype)
    del u
    M = (n[0] // 2 + 1) * n[1] * n[2]
    comp = M * 2 * np.dtype(dtype).itemsize
    cases = [
        ("exponential_convolve (vector, in place)", 6 * comp,
         lambda: G.exponential_convolve_(dv, dv, plans, coefficient=-1e-6, stream=stream)),
        ("exponential_convolve two-operand (vector)", 9 * comp,
         lambda: G.exponential_convolve_(dw, 1.0, du, 0.5, dv, plans, coefficient=-1e-6,
                                         stream=stream)),
        ("radial_spectrum (vector)", 3 * comp, lambda: G.radial_spectrum(du, plans, stream=stream)),
        ("reality_projection (vector, in place)", None,
         lambda: G.reality_projection_(dv, plans, stream=stream)),
    ]
    res = {}
    for name, nbytes, call in cases:
        for _ in range(warmup):
            call()
        stream.synchronize()
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, stream.handle)
        for _ in range(steps):
            call()
        ev.record(e1, stream.handle)
        ms = ev.elapsed(e0, e1) / steps
        rec = {"ms": round(ms, 4)}
        if nbytes:
            gbs = nbytes / (ms * 1e-3) / 1e9
            rec.update({"bytes_alg": int(nbytes), "achieved_GBs": round(gbs, 1),
                        "frac": round(gbs / HBM_PEAK_GBS, 4)})
        res[name] = rec
        print(f"[ops] {name}: {rec}", file=sys.stderr, flush=True)
    return res


def run_step(G, stream, ev, steps, warmup, n=(512, 512, 512), dtype=np.float64, nu=0.02, dt=0.01):
    """One integrating-factor RK2 step of the consumer's step! (without forcing,
    examples/HIST_CPU.jl:267-282) on a 512^3 solenoidal field, two ways:
    folded -- two gbg_uxw_combine calls (stage 1 writes N_1 and u* = E(u + dt N_1),
    stage 2 writes u <- E(u + dt/2 N_1) + dt/2 N_2 in place); standalone -- two
    u x omega terms and three exponential-kernel passes (u* = E(u + dt N_1),
    u <- E(u + dt/2 N_1), u <- u + dt/2 N_2), the composition of reference calls."""
    from gbg_amd import workload
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    u = workload.synthetic_field(n, 2001, vector=True)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    del u
    k1, us, k2 = (G.Field(grid, "spectral", 1, dtype) for _ in range(3))
    kc = G.dealias_cutoff(n[0])
    c = -nu * dt

    def folded():
        G.uxw_combine_(us, du, plans, kc, n_hat=k1, a=du, delta=dt, coefficient=c,
                       reductions=False, stream=stream)
        G.uxw_combine_(du, us, plans, kc, a=du, b=k1, beta=dt / 2, gamma=dt / 2, coefficient=c,
                       reductions=False, stream=stream)

    def standalone():
        G.uxw_(k1, du, plans, kc, reductions=False, stream=stream)
        G.exponential_convolve_(us, 1.0, du, dt, k1, plans, coefficient=c, stream=stream)
        G.uxw_(k2, us, plans, kc, reductions=False, stream=stream)
        G.exponential_convolve_(du, 1.0, du, dt / 2, k1, plans, coefficient=c, stream=stream)
        G.exponential_convolve_(du, 1.0, du, dt / 2, k2, plans, coefficient=0.0, stream=stream)

    res = {}
    for name, call in (("folded", folded), ("standalone", standalone)):
        for _ in range(warmup):
            call()
        stream.synchronize()
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, stream.handle)
        for _ in range(steps):
            call()
        ev.record(e1, stream.handle)
        res[name] = {"ms_per_step": round(ev.elapsed(e0, e1) / steps, 4)}
        print(f"[step] {name}: {res[name]}", file=sys.stderr, flush=True)
    res["speedup"] = round(res["standalone"]["ms_per_step"] / res["folded"]["ms_per_step"], 4)
    return res


def run_rows(G, stream, ev, steps, warmup, n=(512, 512, 512), dtype=np.float64):
    """Each operator of the spectral path on its own at --size, timed by HIP
    events on the launch stream, with its algorithmic bytes (the fused-pass
    ideal: every pass reads and writes each participating component once;
    spectral component = M complex elements, physical component = P reals)
    and the fraction of the 8 TB/s HBM peak that implies."""
    from gbg_amd import workload
    grid = G.Grid3D(n)
    plans = G.FieldTransformPlans(grid, dtype)
    es = np.dtype(dtype).itemsize
    M = (n[0] // 2 + 1) * n[1] * n[2]
    P = n[0] * n[1] * n[2]
    sc, ph = M * 2 * es, P * es            # one spectral / one physical component
    u = workload.synthetic_field(n, 2001, vector=True)
    du = G.Field(grid, "spectral", 1, dtype).upload(u)
    dv = G.Field(grid, "spectral", 1, dtype).upload(u)
    dsrc = G.Field(grid, "spectral", 1, dtype).upload(u)     # consumed by the 3-arg inverse
    da = G.Field(grid, "spectral", 0, dtype).upload(u[0])
    db = G.Field(grid, "spectral", 0, dtype).upload(u[1])
    del u
    ds = G.Field(grid, "spectral", 0, dtype)
    dt = G.Field(grid, "spectral", 2, dtype)
    pu = G.to_physical(du, plans, stream=stream)
    pv = G.to_physical(dv, plans, stream=stream)
    pw = G.Field(grid, "physical", 1, dtype)
    pa = G.to_physical(da, plans, stream=stream)
    pb = G.to_physical(db, plans, stream=stream)
    pc = G.Field(grid, "physical", 0, dtype)
    band = G.spectral_cutoff(plans, G.dealias_cutoff(n[0]))
    stream.synchronize()
    cases = [
        ("to_spectral (vector)", 3 * (ph + 5 * sc), lambda: G.to_spectral_(dv, pu, plans, stream=stream)),
        ("to_physical 3-arg (vector)", 3 * (ph + 5 * sc),
         lambda: G.to_physical_(pw, dsrc, plans, stream=stream)),
        ("gradient scalar->vector", 4 * sc, lambda: G.gradient_(dv, da, plans, stream=stream)),
        ("gradient vector->tensor", 12 * sc, lambda: G.gradient_(dt, du, plans, stream=stream)),
        ("divergence vector->scalar", 4 * sc, lambda: G.divergence_(ds, du, plans, stream=stream)),
        ("curl", 6 * sc, lambda: G.curl_(dv, du, plans, stream=stream)),
        ("inverse_curl", 6 * sc, lambda: G.inverse_curl_(dv, du, plans, stream=stream)),
        ("laplacian (vector)", 6 * sc, lambda: G.laplacian_(dv, du, plans, stream=stream)),
        ("inverse_laplacian (scalar)", 2 * sc, lambda: G.inverse_laplacian_(ds, da, plans, stream=stream)),
        ("helmholtz_projection", 6 * sc, lambda: G.helmholtz_projection_(dv, du, plans, stream=stream)),
        ("convolve 2/3 cutoff (vector)", 6 * sc, lambda: G.convolve_(dv, du, band, plans, stream=stream)),
        ("cross (physical)", 9 * ph, lambda: G.cross_(pw, pu, pv, plans, stream=stream)),
        ("multiply (physical scalar)", 3 * ph, lambda: G.multiply_(pc, pa, pb, plans, stream=stream)),
        ("mean_free_projection (vector)", None, lambda: G.mean_free_projection_(dv, plans, stream=stream)),
        ("dealiased_product (scalar)", product_bytes(n, G.dealias_cutoff(n[0]), es),
         lambda: G.dealiased_product_(ds, da, db, plans, stream=stream)),
    ]
    res = {}
    for name, nbytes, call in cases:
        for _ in range(warmup):
            call()
        stream.synchronize()
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, stream.handle)
        for _ in range(steps):
            call()
        ev.record(e1, stream.handle)
        ms = ev.elapsed(e0, e1) / steps
        rec = {"ms": round(ms, 4)}
        if nbytes:
            gbs = nbytes / (ms * 1e-3) / 1e9
            rec.update({"bytes_alg": int(nbytes), "achieved_GBs": round(gbs, 1),
                        "frac": round(gbs / HBM_PEAK_GBS, 4)})
        res[name] = rec
        print(f"[rows] {name}: {rec}", file=sys.stderr, flush=True)
    # the reductions return a host scalar (a device-to-host copy and a stream
    # synchronisation per call): timed on the host clock
    for name, nbytes, call in [
            ("inner spectral (vector)", 6 * sc, lambda: G.inner(du, dv, plans, stream=stream)),
            ("inner physical (vector)", 6 * ph, lambda: G.inner(pu, pv, plans, stream=stream))]:
        for _ in range(warmup):
            call()
        t0 = time.perf_counter()
        for _ in range(steps):
            call()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        gbs = nbytes / (ms * 1e-3) / 1e9
        res[name] = {"ms": round(ms, 4), "bytes_alg": int(nbytes), "achieved_GBs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "clock": "host, includes the result copy"}
        print(f"[rows] {name}: {res[name]}", file=sys.stderr, flush=True)
    return res


def smooth_sizes(lo=32, hi=512):
    out = []
    for n in range(lo, hi + 1, 2):
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            out.append(n)
    return out


def run_serve(G, nreq=240, ngrids=12, seed=5001, slots=4, budget_gb=200.0):
    """Serving form: a Poisson stream of {rfft, irfft, curl, inverse Laplacian,
    dealiased product} requests on resident seeded fields of a pool of grids
    (axes drawn from the 2*3*5-smooth even sizes in [32, 512], about a third
    cubic), FP32 / FP64 50/50.  Capacity is measured with a closed burst; the
    Poisson stream then runs at 0.5, 0.8 and 1.0 of it."""
    import threading
    from gbg_amd import workload
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = smooth_sizes()
    grids = []
    while len(grids) < ngrids:
        if rng.random() < 1.0 / 3.0:
            n = (int(rng.choice(sizes)),) * 3
        else:
            n = tuple(int(v) for v in rng.choice(sizes, size=3))
        if n not in grids and (n[0] // 2 + 1) * n[1] * n[2] <= 40_000_000:
            grids.append(n)
    pool = {}
    t0 = time.time()
    for gi, n in enumerate(grids):
        u = workload.synthetic_field(n, seed + 1 + gi, vector=True)
        for dt in (np.float64, np.float32):
            grid = G.Grid3D(n)
            plans = G.FieldTransformPlans(grid, dt)
            du = G.Field(grid, "spectral", 1, dt).upload(u)
            da = G.Field(grid, "spectral", 0, dt).upload(u[0])
            db = G.Field(grid, "spectral", 0, dt).upload(u[1])
            dp = G.to_physical(du, plans)
            outs = {"s1": [G.Field(grid, "spectral", 1, dt) for _ in range(3)],
                    "s0": [G.Field(grid, "spectral", 0, dt) for _ in range(3)],
                    "p1": [G.Field(grid, "physical", 1, dt) for _ in range(3)]}
            pool[(gi, dt)] = (du, da, db, dp, outs)
            del plans
        print(f"[serve] grid {gi} {n} resident", file=sys.stderr, flush=True)
    gen_s = time.time() - t0
    S = G.Server
    kinds = [S.TO_SPECTRAL, S.TO_PHYSICAL, S.CURL, S.INVERSE_LAPLACIAN, S.DEALIASED_PRODUCT]

    def draw(r):
        gi = int(r.integers(0, len(grids)))
        dt = np.float64 if r.random() < 0.5 else np.float32
        return int(r.choice(kinds)), gi, dt

    def submit(srv, k, gi, dt, ctr):
        du, da, db, dp, outs = pool[(gi, dt)]
        c = ctr[(gi, dt)] = ctr.get((gi, dt), 0) + 1
        if k == S.TO_SPECTRAL:
            return srv.submit(k, outs["s1"][c % 3], dp)
        if k == S.TO_PHYSICAL:
            return srv.submit(k, outs["p1"][c % 3], du)
        if k == S.DEALIASED_PRODUCT:
            return srv.submit(k, outs["s0"][c % 3], da, db)
        return srv.submit(k, outs["s1"][c % 3], du)

    def run(rate, n, r):
        """rate None = closed burst (all submitted at once)."""
        srv = G.Server(int(budget_gb * (1 << 30)), slots=slots)
        # warm the per-slot plan caches (bursts of every request type on every
        # grid) so first-use plan creation stays out of the numbers; output
        # buffers are scratch rings, their contents are not inspected here
        ctr = {}
        for _ in range(2):
            tickets = [submit(srv, k, gi, dt, ctr) for gi in range(len(grids))
                       for dt in (np.float64, np.float32) for k in kinds for _ in range(slots)]
            for t in tickets:
                srv.wait(t)
        base = srv.stats()["completed"]
        reqs = [draw(r) for _ in range(n)]
        gaps = r.exponential(1.0 / rate, size=n) if rate else np.zeros(n)
        lats = []
        q = []
        done = threading.Event()

        def waiter():
            i = 0
            while i < n:
                if i < len(q):
                    lats.append(srv.wait(q[i]))
                    i += 1
                else:
                    time.sleep(0.0002)
            done.set()

        th = threading.Thread(target=waiter)
        th.start()
        t_start = time.perf_counter()
        t_next = t_start
        for (k, gi, dt), gap in zip(reqs, gaps):
            t_next += gap
            while True:
                now = time.perf_counter()
                if now >= t_next:
                    break
                time.sleep(min(t_next - now, 0.001))
            q.append(submit(srv, k, gi, dt, ctr))
        done.wait()
        th.join()
        wall = time.perf_counter() - t_start
        srv.close()
        lats = np.sort(np.array(lats))
        return {"requests": n, "p50_ms": round(float(np.percentile(lats, 50)), 3),
                "p99_ms": round(float(np.percentile(lats, 99)), 3),
                "fields_per_s": round(n / wall, 1), "offered_rate": None if rate is None else round(rate, 1),
                "warm_jobs": base}

    r = np.random.Generator(np.random.PCG64(seed + 100))
    cap = run(None, nreq, r)
    capacity = cap["fields_per_s"]
    res = {"grids": ["x".join(map(str, g)) for g in grids], "slots": slots, "budget_GB": budget_gb,
           "generation_s": round(gen_s, 1), "capacity_burst": cap, "poisson": {}}
    for frac in (0.5, 0.8, 1.0):
        res["poisson"][str(frac)] = run(frac * capacity, nreq, r)
        print(f"[serve] rate {frac} x capacity: {res['poisson'][str(frac)]}", file=sys.stderr, flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--cpu-workers", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-reps", type=int, default=3,
                    help="timed CPU-baseline terms per precision after one warm-up")
    ap.add_argument("--form", default="uxw",
                    choices=["uxw", "batch", "serve", "hbm", "ops", "rows", "step"],
                    help="uxw: the headline 512^3 term (default); batch: the 96-field batched "
                         "graph form; serve: the Poisson request stream; hbm: the streaming "
                         "copy/fill rates of the device; ops: exponential kernel, radial "
                         "spectrum and reality projection at --size; rows: every transform, "
                         "operator, product and reduction on its own at --size; step: one "
                         "integrating-factor RK2 step at --size, folded and standalone")
    ap.add_argument("--batch-streams", type=int, default=4,
                    help="streams the batched form spreads its items over (--form batch)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first copy BYTES device-to-device twice (a plain copy kernel of known "
                         "size, the scale for the HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the spectral path here runs on a single MI355X and has "
              "no multi-GPU form (a field would shard as z slabs with one all-to-all per "
              "transform direction). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    G = load_package()
    if G.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    n = (args.size,) * 3
    dtype = np.float64 if args.dtype == "f64" else np.float32
    elem = np.dtype(dtype).itemsize
    ev = HipEvents()
    stream = G.Stream()
    if args.form == "batch":
        res = run_batch(G, stream, ev, args.steps, args.warmup, nstreams=args.batch_streams)
        out = {"metric": "batched form: 96 uneven-grid fields (gradient + dealiased product) per "
                         "graph launch, fields/s", "value": res["f64"]["fields_per_s"],
               "unit": "fields/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
               "ms_per_step": res["f64"]["ms_per_batch"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64",
               "data": "synthetic (seeded k^-5/3 spectra, master seed 4001)",
               "config": {"workload": "96-field batch, grids drawn from " +
                          ", ".join("x".join(map(str, s)) for s in BATCH_SHAPES)},
               "batch": res}
        print(json.dumps(out), flush=True)
        return
    if args.form == "ops":
        res = run_ops(G, stream, ev, args.steps, args.warmup, n, dtype)
        first = res["exponential_convolve (vector, in place)"]
        out = {"metric": f"{n[0]}^3 exponential_convolve (vector) GB/s + radial spectrum, "
                         "reality projection", "value": first["achieved_GBs"], "unit": "GB/s",
               "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
               "ms_per_step": first["ms"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": args.dtype,
               "data": "synthetic (seeded k^-5/3 solenoidal spectrum, seed 2001)",
               "config": {"workload": f"{n[0]}x{n[1]}x{n[2]} vector field operators"},
               "roofline": {"bound": "hbm", "achieved": first["achieved_GBs"],
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": first["frac"],
                            "traffic": None},
               "ops": res}
        print(json.dumps(out), flush=True)
        return
    if args.form == "step":
        res = run_step(G, stream, ev, args.steps, args.warmup, n, dtype)
        out = {"metric": f"{n[0]}^3 integrating-factor RK2 step (two u x omega terms and the "
                         "exponential combinations), steps/s", "value": round(1e3 / res["folded"]["ms_per_step"], 3),
               "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
               "ms_per_step": res["folded"]["ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
               "data": "synthetic (seeded k^-5/3 solenoidal spectrum, seed 2001)",
               "config": {"workload": f"{n[0]}x{n[1]}x{n[2]} RK2 step, nu = 0.02, dt = 0.01"},
               "step": res}
        print(json.dumps(out), flush=True)
        return
    if args.form == "rows":
        res = run_rows(G, stream, ev, args.steps, args.warmup, n, dtype)
        first = res["to_spectral (vector)"]
        out = {"metric": f"{n[0]}^3 spectral-path operators on their own: GB/s by algorithmic "
                         "bytes (headline: to_spectral of a vector field)",
               "value": first["achieved_GBs"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": first["ms"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
               "data": "synthetic (seeded k^-5/3 solenoidal spectrum, seed 2001)",
               "config": {"workload": f"{n[0]}x{n[1]}x{n[2]} fields, one operator per line"},
               "roofline": {"bound": "hbm", "achieved": first["achieved_GBs"],
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": first["frac"],
                            "traffic": None},
               "rows": res}
        print(json.dumps(out), flush=True)
        return
    if args.form == "hbm":
        res = run_hbm(G, stream, ev, args.steps)
        out = {"metric": "HBM streaming rate: runtime copy kernel (1 GiB read + 1 GiB write)",
               "value": res["copy"]["GBs"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
               "warmup": 2, "ms_per_step": res["copy"]["ms"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "device buffers",
               "config": {"workload": "1 GiB device-to-device copy and 1 GiB fill"}, "hbm": res}
        print(json.dumps(out), flush=True)
        return
    if args.form == "serve":
        res = run_serve(G)
        p = res["poisson"]["0.8"]
        out = {"metric": "serving form: Poisson request stream, fields/s and p50/p99 latency",
               "value": p["fields_per_s"], "unit": "fields/s", "n_gpus": 1, "steps": p["requests"],
               "warmup": 0, "ms_per_step": round(1000.0 / p["fields_per_s"], 4),
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
               "data": "synthetic (seeded k^-5/3 spectra, seed 5001)",
               "config": {"workload": "requests {rfft, irfft, curl, inverse Laplacian, dealiased "
                                      "product} on 12 resident grids, FP32/FP64 50/50, offered at "
                                      "0.8 x measured capacity"},
               "latency_ms": {"p50": p["p50_ms"], "p99": p["p99_ms"]}, "serve": res}
        print(json.dumps(out), flush=True)
        return
    print(f"[bench] {n} {args.dtype} steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)
    if args.copy_calibration > 0:
        nb = args.copy_calibration
        src = C.c_void_p()
        dst = C.c_void_p()
        G.check(G.load().gbg_malloc(C.byref(src), nb))
        G.check(G.load().gbg_malloc(C.byref(dst), nb))
        G.check(G.load().gbg_memset(src.value, 1, nb, stream.handle))
        for _ in range(2):
            G.check(G.load().gbg_memcpy_d2d(dst.value, src.value, nb, stream.handle))
        stream.synchronize()
        G.check(G.load().gbg_free(src.value))
        G.check(G.load().gbg_free(dst.value))
        print(f"[bench] copy calibration: 2 x {nb} bytes device-to-device", file=sys.stderr, flush=True)

    r = run_gpu(G, n, dtype, args.steps, args.warmup, stream, ev)
    ms = r["total_ms"] / args.steps
    fields_s = 1000.0 / ms
    per = sorted(r["per"])
    p50 = per[len(per) // 2]
    p99 = per[min(len(per) - 1, int(math.ceil(0.99 * len(per))) - 1)]
    print(f"[bench] gpu {ms:.3f} ms/step, stages {r['stages']}", file=sys.stderr, flush=True)

    # dominant pass and its roofline, on the bytes the pruned pass schedule moves
    sb = uxw_stage_bytes(n, r["kc"], elem)
    names = list(sb.keys())
    stages = r["stages"]
    roof = None
    if len(stages) == len(names):
        i = int(np.argmax(stages))
        ach = sb[names[i]] / (stages[i] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": names[i], "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(names[i], n, dtype), "avg_ms": round(stages[i], 4),
                "bytes_alg": int(sb[names[i]])}
    B = sum(sb.values())
    pipeline_gbs = B / (ms * 1e-3) / 1e9
    M = (n[0] // 2 + 1) * n[1] * n[2]
    Bu = uxw_bytes_unpruned(n, elem)
    prec = "FP64" if dtype == np.float64 else "FP32"
    out = {
        "metric": f"{n[0]}^3 dealiased u x omega fields/s ({prec}) + % HBM roofline; FP32 vs FP64; "
                  "p99 latency",
        "value": round(fields_s, 3),
        "unit": "fields/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if dtype == np.float64 else "f32",
        "data": "synthetic (seeded k^-5/3 solenoidal spectrum, seed 2001)",
        "config": {"workload": f"dealiased u x omega, {n[0]}x{n[1]}x{n[2]} vector field, 2/3 rule, "
                               "Leray projection, energy/enstrophy", "grid": list(n),
                   "kc": r["kc"], "parallelism": "single GPU"},
        "latency_ms": {"p50": round(p50, 4), "p99": round(p99, 4)},
        "pipeline_roofline": {"bytes_alg": int(B), "achieved": round(pipeline_gbs, 1),
                              "frac": round(pipeline_gbs / HBM_PEAK_GBS, 4),
                              "stage_bytes_alg": {k: int(v) for k, v in sb.items()},
                              "compulsory_bytes": int(6 * M * 2 * elem),
                              "unpruned_equivalent": {
                                  "bytes": int(Bu), "GBs": round(Bu / (ms * 1e-3) / 1e9, 1),
                                  "note": "42 M s_c as if no mode were masked; a label, not a "
                                          "roofline"}},
        "stage_ms": [round(x, 4) for x in stages],
        "roofline": roof,
    }
    nh_gpu = r["dn"].download()
    E, Z = r["E"], r["Z"]
    del r

    n32 = None
    if dtype == np.float64 and not args.no_fp32:
        r32 = run_gpu(G, n, np.float32, args.steps, args.warmup, stream, ev)
        ms32 = r32["total_ms"] / args.steps
        n32 = r32["dn"].download()
        sb32 = uxw_stage_bytes(n, r32["kc"], 4)
        out["fp32"] = {"value": round(1000.0 / ms32, 3), "ms_per_step": round(ms32, 4),
                       "speedup_vs_fp64": round(ms / ms32, 3),
                       "stage_ms": [round(x, 4) for x in r32["stages"]],
                       "pipeline_frac": round(sum(sb32.values()) / (ms32 * 1e-3) / 1e9
                                              / HBM_PEAK_GBS, 4)}
        E32, Z32 = r32["E"], r32["Z"]
        del r32

    if args.no_cpu_baseline or n[0] > 512:
        if n32 is not None:
            # no host oracle at this size: FP32 against the FP64 GPU result,
            # which the parity tests hold to 1e-12 of the FP64 oracle
            num = den = 0.0
            for c in range(3):
                d = n32[c].astype(np.complex128) - nh_gpu[c]
                num += float(np.vdot(d, d).real)
                den += float(np.vdot(nh_gpu[c], nh_gpu[c]).real)
                del d
            out["rel_l2_fp32_vs_fp64_gpu"] = {"N_hat": math.sqrt(num / den),
                                             "energy": abs(E32 - E) / abs(E),
                                             "enstrophy": abs(Z32 - Z) / abs(Z)}
    else:
        os.environ["GBG_ORACLE_WORKERS"] = str(args.cpu_workers)
        sys.path.insert(0, REPO)
        from oracle import gbg_oracle as O   # the checker and CPU baseline only
        from gbg_amd import workload
        g = O.Grid3D(n)
        m = O.Multipliers.build(g)
        u = workload.synthetic_field(n, 2001, vector=True)
        kc = out["config"]["kc"]
        reps = args.cpu_reps
        t64 = []
        for _ in range(reps + 1):                  # the first run is the warm-up
            t0 = time.perf_counter()
            ref, Er, Zr = O.uxw(u, g, kc, m, workers=args.cpu_workers)
            t64.append(time.perf_counter() - t0)
        u32 = u.astype(np.complex64)
        del u
        t32 = []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            ref32 = cpu_uxw_single(u32, m, kc, args.cpu_workers)
            t32.append(time.perf_counter() - t0)
        del u32
        cpu_err32 = O.rel_l2(ref32.astype(np.complex128), ref)
        del ref32
        cpu_model = platform.processor() or ""
        try:
            with open("/proc/cpuinfo") as fh:
                for line in fh:
                    if line.startswith("model name"):
                        cpu_model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
        w64, w32 = sorted(t64[1:]), sorted(t32[1:])
        med64, med32 = float(np.median(w64)), float(np.median(w32))
        out["cpu_baseline"] = {
            "value": round(1.0 / med64, 4), "unit": "fields/s", "cores": args.cpu_workers,
            "kind": "port",
            "sample": f"{reps} timed full {n[0]}^3 u x omega terms after one warm-up, FP64 "
                      f"(the oracle: NumPy + scipy.fft pocketfft, workers={args.cpu_workers}); "
                      f"value from the median",
            "fp64_s": {"min": round(w64[0], 3), "median": round(med64, 3), "runs": reps},
            "fp32": {"value": round(1.0 / med32, 4), "min_s": round(w32[0], 3),
                     "median_s": round(med32, 3), "runs": reps,
                     "rel_l2_vs_fp64_oracle": cpu_err32},
            "fft_library": "pocketfft (scipy.fft " + _scipy_version() + ")",
            "fftw_on_host": _fftw_probe(),
            "cpu_model": cpu_model, "host_cpus": os.cpu_count()}
        out["rel_l2_vs_oracle"] = {"N_hat_fp64": O.rel_l2(nh_gpu, ref),
                                   "energy": abs(E - Er) / abs(Er), "enstrophy": abs(Z - Zr) / abs(Zr)}
        if n32 is not None:
            out["rel_l2_vs_oracle"].update({
                "N_hat_fp32": O.rel_l2(n32.astype(np.complex128), ref),
                "energy_fp32": abs(E32 - Er) / abs(Er), "enstrophy_fp32": abs(Z32 - Zr) / abs(Zr)})
        mask = O.band_mask(m, 0.0, kc)
        out["mask_exact"] = bool(np.all(nh_gpu[:, mask == 0] == 0))
    print(json.dumps(out), flush=True)


def _scipy_version():
    try:
        import scipy
        return scipy.__version__
    except Exception:
        return "?"


def _fftw_probe():
    """Whether the host's dynamic linker knows a libfftw3 (ldconfig -p)."""
    import subprocess
    try:
        out = subprocess.run(["ldconfig", "-p"], capture_output=True, text=True, timeout=10).stdout
    except Exception:
        return "unknown (ldconfig failed)"
    libs = sorted({ln.split()[0] for ln in out.splitlines() if "libfftw3" in ln})
    return libs or "none"


def cpu_uxw_single(u_hat, m, kc, workers):
    """The CPU baseline's u x omega term in single precision: the oracle's
    steps (examples/HIST_CPU.jl:229-247: M, curl, inverse transforms, cross
    product, forward transform, Leray projection, M, mean-free) on complex64 /
    float32 arrays (scipy.fft transforms in the input's precision), the
    derivative tables rounded to float32 and the FP64 mask predicate."""
    import scipy.fft as sf
    f32 = np.float32
    mask = ((m.ksq() <= float(kc) * float(kc))).astype(f32)
    a1, a2, a3 = (a.astype(f32) for a in (m.a1(), m.a2(), m.a3()))
    s = m.grid.physical_shape
    ut = mask[None] * u_hat
    wt = np.stack([1j * (a2 * ut[2] - a3 * ut[1]), 1j * (a3 * ut[0] - a1 * ut[2]),
                   1j * (a1 * ut[1] - a2 * ut[0])]).astype(np.complex64)
    u = sf.irfftn(ut, s=s, axes=(-3, -2, -1), workers=workers)
    w = sf.irfftn(wt, s=s, axes=(-3, -2, -1), workers=workers)
    del ut, wt
    nphys = np.stack([u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2],
                      u[0] * w[1] - u[1] * w[0]])
    del u, w
    nt = sf.rfftn(nphys, axes=(-3, -2, -1), workers=workers)
    del nphys
    sq = (a1 * a1 + a2 * a2) + a3 * a3
    lon = (a1 * nt[0] + a2 * nt[1] + a3 * nt[2]) / np.where(sq == 0, f32(1), sq)
    out = np.stack([mask * (nt[0] - a1 * lon), mask * (nt[1] - a2 * lon), mask * (nt[2] - a3 * lon)])
    out[:, 0, 0, 0] = 0
    return out

if __name__ == "__main__":
    main()
