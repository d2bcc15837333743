#!/usr/bin/env python3
"""
bench.py -- D3Q19 stream-collide throughput of the HIP solver on one MI355X.

Workload: the single-structure precision study lattice, D3Q19 at 256^3 with
populations stored as shifted FP32 (f - w) and FP64 flux sums, on a seeded
mask of overlapping 3:1 ellipsoids along a random axis with 15-degree jitter,
porosity drawn from U(0.3, 0.8).  A "step" is one stream-collide pass over
the whole lattice (every fluid cell updated once); the lattice stays resident
in HBM.  value = fluid-cell updates per second (MLUPS) over the timed region.

Timing: after the warmup steps, the K timed steps run twice.  First one kernel
at a time with a HIP event pair around each (the per-kernel averages the
roofline uses); then, after one untimed chunk, as replays of the captured
200-step hipGraph chunk, exactly as solves run: value and ms_per_step come
from this second pass (eager_ms_per_step reports the first).  The convergence
rules are those of a first phase that never converges (eps 0, a check every
2000 steps), so a chunk launches the finish kernel (flux sum and convergence
update) only around its checks, as a solve's chunks do between checks;
fp32_flux_every_step times the averaging phase, a finish after every step.

Printed on one JSON line (stdout): the metric, the roofline of the step
kernels (algorithmic bytes 2*19*4 = 152 B per fluid update over their average
duration from HIP events on the solver's stream), the FP64 rate on the same
lattice, and the reference CPU solver timed on this host on a bounded sample.
Progress goes to stderr.
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "flow-anizotropy_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_mask(n, seed, kind="ellipsoids"):
    from flowperm import masks
    if kind == "open":                       # diagnostic: every cell fluid
        log(f"mask {n}^3 open (all fluid)")
        return np.zeros((n, n, n), np.uint8), 1.0
    rng = np.random.default_rng(seed)
    por = 0.3 + 0.5 * rng.random()
    t0 = time.time()
    m = masks.ellipsoids((n, n, n), por, rmin=2.0, rmax=4.0, aspect=3.0, jitter_deg=15.0,
                         seed=seed)
    log(f"mask {n}^3 ellipsoids seed {seed}: target porosity {por:.3f}, "
        f"actual {1 - m.mean():.4f} ({time.time() - t0:.1f} s)")
    return m, por


TRAFFIC_PROFILE = "profiles/d3q19_f32_step_kernels_256cube_ellipsoids.md"


def measured_traffic(n_fluid):
    """HBM bytes per launch of each step kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this lattice (copy-calibrated counter
    runs of bench.py, not measured in this run), or None when the summary is
    for another lattice."""
    p = REPO / TRAFFIC_PROFILE
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("traffic-json:"):
            d = json.loads(line.split(":", 1)[1])
            if d["n_fluid"] == n_fluid:
                return d
    return None


def bench_steps(ctx, mask, prec, warmup, steps, tune=None, flux=0):
    from flowperm import _lib
    nz, ny, nx = mask.shape
    out = _lib.BenchOut()
    F = np.array([2.5e-7 * 40, 0.0, 0.0])
    t = None
    if tune is not None:
        t = _lib.Tuning(tune[0], tune[1], 200, tune[2], tune[3], tune[4],
                        *(tune[k] if len(tune) > k else 0 for k in (5, 6, 7, 8)))
    st = ctx.lib.fp_bench_steps(ctx.handle, _lib.u8ptr(mask), nx, ny, nz, 3, prec,
                                _lib.f64ptr(F), warmup, steps, C.byref(t) if t else None,
                                int(flux), C.byref(out))
    if st != 0:
        raise RuntimeError(f"fp_bench_steps failed: {st} {ctx.last_error()}")
    return out


# convergence flags of the batched and serving solves: the 64^3 parity
# configuration (--ff 40 --lag 1000 --eps 1e-6, reference defaults otherwise)
SOLVE_CFG = dict(ff=40, lag=1000, eps=1e-6, hold=2, min_steps=4000, max_steps=200000,
                 avg_steps=2000)


def k_matrix(row):
    return np.array([[row[f"K_{a}{b}"] for b in "xyz"] for a in "xyz"])


def k_error(r32, r64):
    """max over components of |K32 - K64| / max(|K64|, |trace K64|)"""
    K32, K64 = k_matrix(r32), k_matrix(r64)
    return float((np.abs(K32 - K64) / np.maximum(np.abs(K64), abs(np.trace(K64)))).max())


TUNED = {}


def apply_tuning(ctx, name, nbr_mode=None):
    """make the autotuned launch shapes of a precision the context's
    (nbr_mode overrides the neighbour-step form: 2 = per structure)"""
    from flowperm import _lib
    r = TUNED.get(name)
    if r:
        t = _lib.Tuning(r["local_cpt"], r["nbr_waves"], r["chunk_steps"], r["nontemporal"],
                        r["local_bpc"], r["nbr_bpc"],
                        r.get("nbr_mode", 2) if nbr_mode is None else nbr_mode, r.get("nbr_segs", 0),
                        r.get("local_wgs", 0), r.get("nbr_wgs", 0))
        ctx.lib.fp_set_tuning(ctx.handle, C.byref(t))


def plain_fp64(ctx, structs):
    """the plain path: one structure at a time, its three forcing runs one
    after another, FP64 (fp_solve3d)"""
    from flowperm import solver
    apply_tuning(ctx, "f64", nbr_mode=2)       # structures of every porosity
    rows, t0 = [], time.perf_counter()
    for s in structs:
        rows.append(solver.solve_permeability3d(s, ctx=ctx, precision="f64", concurrent=False,
                                                **SOLVE_CFG))
    return rows, time.perf_counter() - t0


def burst(ctx, structs, precision="f32", every=20.0):
    """every structure submitted at once to a request server: all forcing
    runs step together in one launch per step, each retiring on its own
    (a progress line every `every` seconds)"""
    from flowperm.server import Server
    apply_tuning(ctx, "f32" if precision == "f32" else "f64", nbr_mode=2)
    with Server(3, precision, ctx=ctx) as srv:
        t0 = time.perf_counter()
        ids = [srv.submit(s, **SOLVE_CFG) for s in structs]
        last = time.perf_counter()
        while srv.step():
            if time.perf_counter() - last > every:
                last = time.perf_counter()
                st = srv.stats()
                log(f"  ... {st['completed']}/{len(ids)} done, {st['live_instances']} runs live, "
                    f"{st['steps']} steps, {last - t0:.0f} s")
        res = {i: srv.poll(i) for i in ids}
        dt = time.perf_counter() - t0
        st = srv.stats()
    return [res[i][0] for i in ids], dt, st


def batched_full(ctx, n, m, seed, workers):
    """the whole batched set: n structures x 3 forcing runs resident and
    stepped together in FP32, against the plain FP64 path on the first m"""
    from flowperm import workloads
    t0 = time.time()
    structs = workloads.generate("batched", n, seed, workers)
    fluid = np.array([int((s == 0).sum()) for s in structs])
    log(f"batched-full: {n} structures ({fluid.sum() / 1e6:.0f} M fluid cells in all) "
        f"generated in {time.time() - t0:.0f} s")
    rows32, t32, st = burst(ctx, structs)
    log(f"batched FP32: {n} tensors in {t32:.1f} s = {n / t32:.3f} tensors/s, "
        f"peak device memory {st['bytes_peak'] / 2**30:.1f} GiB")
    rows64, t64 = plain_fp64(ctx, structs[:m])
    err = max(k_error(a, b) for a, b in zip(rows32[:m], rows64))
    t32m = sum(r["fluid_updates"] for r in rows32[:m]) / (st["fluid_updates"] / t32)
    return {"structures": n, "instances": 3 * n, "fluid_cells_total": int(fluid.sum()),
            "batched_fp32_tensors_per_s": n / t32, "batched_fp32_seconds": t32,
            "batched_fp32_mlups": st["fluid_updates"] / t32 / 1e6,
            "peak_device_bytes": st["bytes_peak"], "chunks": st["chunks"],
            "plain_fp64_structures": m, "plain_fp64_tensors_per_s": m / t64,
            "plain_fp64_mlups": sum(r["fluid_updates"] for r in rows64) / t64 / 1e6,
            "speedup_same_first_m": t64 / t32m,
            "speedup_note": "plain FP64 time on the first m against the batched FP32 time "
                            "those m structures' updates take at the batch's aggregate rate",
            "max_k_error_first_m": err}


def batched_sample(ctx, n, m, seed, workers, fp64_gbs=None):
    """The first n structures of the batched set (512 in all): all 3n forcing
    runs stepped together in FP32 by one request server, against the plain
    FP64 path (one structure at a time) on the first m of them.  The speedup
    compares the plain FP64 time of those m with the time their updates take
    at the batch's aggregate rate."""
    from flowperm import workloads
    t0 = time.time()
    structs = workloads.generate("batched", n, seed, workers)
    edges = [int(s.shape[0]) for s in structs]
    log(f"batched: {n} structures, edges {edges} ({time.time() - t0:.1f} s to generate)")
    rows32, t32, st = burst(ctx, structs)
    rate32 = st["fluid_updates"] / t32
    log(f"batched FP32: {n} tensors in {t32:.2f} s = {n / t32:.3f} tensors/s; "
        f"{rate32 / 1e6:.0f} MLUPS, {st['chunks']} chunks")
    rows64, t64 = plain_fp64(ctx, structs[:m])
    upd64 = sum(r["fluid_updates"] for r in rows64)
    t32m = sum(r["fluid_updates"] for r in rows32[:m]) / rate32
    log(f"plain FP64: {m} tensors in {t64:.2f} s = {m / t64:.3f} tensors/s; "
        f"{upd64 / t64 / 1e6:.0f} MLUPS; speedup on them {t64 / t32m:.2f}x")
    err = max(k_error(a, b) for a, b in zip(rows32[:m], rows64))
    fluid = [int((s == 0).sum()) for s in structs]
    steps32 = [max(r["steps_fx"], r["steps_fy"], r["steps_fz"]) for r in rows32]
    same_steps = sum(int(all(a[f"steps_{t}"] == b[f"steps_{t}"] for t in ("fx", "fy", "fz")))
                     for a, b in zip(rows32[:m], rows64))
    out = {
        "structures": n, "of_set": 512, "edges": edges, "mean_fluid_cells": float(np.mean(fluid)),
        "instances": 3 * n, "config": "ellipsoids 3:1, 15 deg jitter, porosity U(0.3, 0.8); "
                                      "--ff 40 --lag 1000 --eps 1e-6",
        "excluded": f"structures {n}..511 of the 512-structure set (bench.py --study batched-full "
                    f"solves all of them)",
        "batched_fp32_tensors_per_s": n / t32, "batched_fp32_seconds": t32,
        "batched_fp32_mlups": rate32 / 1e6,
        "plain_fp64_structures": m, "plain_fp64_tensors_per_s": m / t64,
        "plain_fp64_seconds": t64, "plain_fp64_mlups": upd64 / t64 / 1e6,
        "speedup": t64 / t32m,
        "speedup_note": "plain FP64 time of the first m structures over the time their updates "
                        "take at the batched FP32 aggregate rate",
        "max_k_error_vs_fp64": err, "structures_with_fp64_step_counts": same_steps,
        "max_steps": int(max(steps32)), "median_steps": float(np.median(steps32))}
    if fp64_gbs:
        # FP32 moves half the bytes per update: a bandwidth-bound FP32 batch is
        # 2 x (its bandwidth / FP64's) faster, so 3x needs 1.5x FP64's bandwidth
        out["bound_for_3x"] = {
            "fp64_step_kernel_GBs": round(fp64_gbs, 1),
            "fp32_GBs_needed": round(1.5 * fp64_gbs, 1),
            "fp32_frac_of_peak_needed": round(1.5 * fp64_gbs / HBM_PEAK_GBS, 3),
            "note": "speedup = 2 x BW(FP32) / BW(FP64) x batching gain on bandwidth-bound steps"}
    return out


def serving_sample(ctx, n, seed, workers, load=0.8):
    """n requests of the serving set: the burst capacity (all submitted at
    once), then Poisson arrivals at load x that capacity with latencies, and
    the plain FP64 sequential path on the same n requests"""
    from flowperm import workloads
    from flowperm.server import Server, poisson_arrivals, serve_schedule
    t0 = time.time()
    reqs = workloads.generate("serving", n, seed, workers)
    edges = [int(s.shape[0]) for s in reqs]
    log(f"serving: {n} requests, edges {edges} ({time.time() - t0:.1f} s to generate)")
    _, tb, _ = burst(ctx, reqs)
    capacity = n / tb
    rate = load * capacity
    arr = poisson_arrivals(rate, n, seed)
    apply_tuning(ctx, "f32", nbr_mode=2)
    with Server(3, "f32", ctx=ctx) as srv:
        out = serve_schedule(srv, reqs, arr, SOLVE_CFG)
        st = srv.stats()
    lat = np.array([o["latency"] for o in out])
    wait = np.array([o["t_admit"] - o["arrival"] for o in out])      # queueing before admission
    service = np.array([o["t_done"] - o["t_admit"] for o in out])    # admission to result
    span = max(o["t_done"] for o in out) - float(arr[0])
    rows64, t64 = plain_fp64(ctx, reqs)
    err = max(k_error(o["row"], r) for o, r in zip(out, rows64))
    res = {
        "requests": n, "edges": edges, "rate_per_s": rate, "load": load,
        "capacity_solves_per_s": capacity,
        "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
        "max_latency_s": float(lat.max()), "served_solves_per_s": n / span,
        "p50_wait_s": float(np.percentile(wait, 50)), "p99_wait_s": float(np.percentile(wait, 99)),
        "p99_service_s": float(np.percentile(service, 99)),
        "latency_note": "latency = wait (arrival to admission) + service (admission to result); "
                        "the tail is the service time of the largest low-porosity GRF requests",
        "peak_device_bytes": st["bytes_peak"],
        "plain_fp64_solves_per_s": n / t64, "capacity_vs_plain_fp64": capacity * t64 / n,
        "max_k_error_vs_fp64": err,
        "config": "edge log-uniform [32, 256], spheres / GRF (8, 8, 24) 50:50, porosity "
                  "U(0.2, 0.8); Poisson arrivals at load x the burst capacity of the same set"}
    log(f"serving: capacity {capacity:.3f}/s, rate {rate:.3f}/s, p50 {res['p50_latency_s']:.2f} s, "
        f"p99 {res['p99_latency_s']:.2f} s; plain FP64 {res['plain_fp64_solves_per_s']:.3f} solves/s")
    return res


def small_lattice_steps(ctx, seed):
    """per-step time of the reference's own lattice sizes solved one at a
    time (graph replays): D2Q9 256^2 and D3Q19 64^3 / 80^3, and whether the
    finish ran deferred into the step kernels, with the default launch shapes
    (2 cells per thread, 4-row slabs, each structure's cheaper neighbour-step
    form, as a context that was never tuned solves)"""
    from flowperm import _lib, masks
    res = {}
    shapes = _lib.Tuning(2, 4, 200, 1, 0, 0, 2)
    cases = [("d2q9_256", masks.ellipses((256, 256), 0.6, seed=seed + 3), 2,
              [("f32", 1), ("f32ref", 2), ("f64", 0)]),
             ("d3q19_64", masks.ellipsoids((64, 64, 64), 0.6, seed=seed + 4), 3, [("f32", 1), ("f64", 0)]),
             ("d3q19_80", masks.ellipsoids((80, 80, 80), 0.6, seed=seed + 4), 3, [("f32", 1), ("f64", 0)])]
    for name, m, lattice, precs in cases:
        if lattice == 3:
            nz, ny, nx = m.shape
        else:
            (ny, nx), nz = m.shape, 1
        m = np.ascontiguousarray(m, np.uint8)
        for pname, prec in precs:
            out = _lib.BenchOut()
            F = np.array([2.5e-7 * 40, 0.0, 0.0])
            st = ctx.lib.fp_bench_steps(ctx.handle, _lib.u8ptr(m), nx, ny, nz, lattice, prec,
                                        _lib.f64ptr(F), 200, 2000, C.byref(shapes), 0, C.byref(out))
            if st != 0:
                raise RuntimeError(f"fp_bench_steps ({name}) failed: {ctx.last_error()}")
            res[f"{name}_{pname}"] = {"us_per_step": round(out.total_ms / out.steps * 1e3, 3),
                                      "n_fluid": int(out.n_fluid), "deferred_finish": int(out.deferred)}
    log(f"small lattices (us per step): { {k: v['us_per_step'] for k, v in res.items()} }")
    return res


def k2_matrix(row):
    return np.array([[row["k_xx"], row["k_xy"]], [row["k_yx"], row["k_yy"]]])


def k2_error(a, b):
    A, B = k2_matrix(a), k2_matrix(b)
    return float((np.abs(A - B) / np.maximum(np.abs(B), abs(np.trace(B)))).max())


def bench2d(ctx, mask, prec, warmup, steps):
    from flowperm import _lib
    ny, nx = mask.shape
    out = _lib.BenchOut()
    F = np.array([2.5e-7, 0.0, 0.0])
    st = ctx.lib.fp_bench_steps(ctx.handle, _lib.u8ptr(mask), nx, ny, 1, 2, prec, _lib.f64ptr(F),
                                warmup, steps, None, 0, C.byref(out))
    if st != 0:
        raise RuntimeError(f"fp_bench_steps (2D) failed: {st} {ctx.last_error()}")
    return out


def solve_logged(ctx, mask, precision, every=20.0, **cfg):
    """one whole solve through a request server (the same result as the
    one-shot solve), with a progress line every `every` seconds"""
    from flowperm.server import Server
    lattice = 3 if mask.ndim == 3 else 2
    apply_tuning(ctx, "f64" if precision == "f64" else "f32")
    t0 = time.perf_counter()
    last = t0
    with Server(lattice, precision, ctx=ctx) as srv:
        rid = srv.submit(mask, **cfg)
        while srv.step():
            if time.perf_counter() - last > every:
                last = time.perf_counter()
                st = srv.stats()
                log(f"  ... {precision} {mask.shape}: {st['steps']} steps, "
                    f"{st['live_instances']} runs live, {last - t0:.0f} s")
        row, _ = srv.poll(rid)
    row["wall_s"] = time.perf_counter() - t0
    return row


def precision_study(ctx, seed, steps):
    """The single-structure precision study: full solves of the 256^3
    ellipsoid lattice (D3Q19) and of a 1024^2 ellipse lattice (D2Q9) in FP64
    and FP32-shifted (2D also in the reference's float arithmetic), with the
    stepping rates and the per-component K error against FP64."""
    from flowperm import masks, solver
    res = {}
    mask, _ = workload_mask(256, seed)
    apply_tuning(ctx, "f32")
    o32 = bench_steps(ctx, mask, 1, 20, steps)
    apply_tuning(ctx, "f64")
    o64 = bench_steps(ctx, mask, 0, 20, steps)
    runs = {}
    for prec in ("f64", "f32"):
        runs[prec] = solve_logged(ctx, mask, prec, **SOLVE_CFG)
        log(f"3D {prec}: steps {[runs[prec][f'steps_{t}'] for t in ('fx', 'fy', 'fz')]} "
            f"in {runs[prec]['wall_s']:.1f} s")
    K64, K32 = k_matrix(runs["f64"]), k_matrix(runs["f32"])
    rel = np.abs(K32 - K64) / np.maximum(np.abs(K64), abs(np.trace(K64)))
    res["d3q19_256"] = {
        "porosity": float(1 - mask.mean()), "n_fluid": int((mask == 0).sum()),
        "mlups": {"f32": o32.n_fluid * o32.steps / o32.total_ms / 1e3,
                  "f64": o64.n_fluid * o64.steps / o64.total_ms / 1e3},
        "step_kernel_GBs": {
            "f32": o32.n_fluid * 152 / (0.5 * (o32.local_ms + o32.nbr_ms) / 1e3) / 1e9,
            "f64": o64.n_fluid * 304 / (0.5 * (o64.local_ms + o64.nbr_ms) / 1e3) / 1e9},
        "K_f64": K64.tolist(), "K_f32": K32.tolist(),
        "rel_error_per_component": rel.tolist(), "max_rel_error": float(rel.max()),
        "steps": {p: [runs[p][f"steps_{t}"] for t in ("fx", "fy", "fz")] for p in runs},
        "solve_seconds": {p: runs[p]["wall_s"] for p in runs},
        "config": "--ff 40 --lag 1000 --eps 1e-6; error = |K32 - K64| / max(|K64|, |tr K64|)"}
    log(f"3D 256^3: max K error FP32 vs FP64 {rel.max():.2e}")

    rng = np.random.default_rng(seed)
    por = 0.3 + 0.5 * rng.random()
    m2 = masks.ellipses((1024, 1024), por, rmin=2.0, rmax=4.0, aspect=3.0, jitter_deg=15.0, seed=seed)
    rates, rows = {}, {}
    L0 = 4.0 * 1024 / 256
    for name, prec in (("f32ref", 2), ("f64", 0), ("f32", 1)):
        apply_tuning(ctx, "f64" if prec == 0 else "f32")
        o = bench2d(ctx, m2, prec, 20, steps)
        rates[name] = {"mlups": o.n_fluid * o.steps / o.total_ms / 1e3,
                       "step_kernel_GBs": o.n_fluid * o.bytes_per_update /
                                          (0.5 * (o.local_ms + o.nbr_ms) / 1e3) / 1e9}
        rows[name] = solve_logged(ctx, m2, name, L0=L0)
        log(f"2D {name}: {rates[name]['mlups']:.0f} MLUPS, steps "
            f"{rows[name]['steps_fx']}/{rows[name]['steps_fy']} in {rows[name]['wall_s']:.1f} s")
    res["d2q9_1024"] = {
        "porosity": float(1 - m2.mean()), "n_fluid": int((m2 == 0).sum()), "rates": rates,
        "K": {n: k2_matrix(r).tolist() for n, r in rows.items()},
        "max_rel_error_vs_f64": {n: k2_error(rows[n], rows["f64"]) for n in ("f32ref", "f32")},
        "steps": {n: [r["steps_fx"], r["steps_fy"]] for n, r in rows.items()},
        "solve_seconds": {n: r["wall_s"] for n, r in rows.items()},
        "config": "reference 2D defaults (eps 1e-3, lag 5000, hold 3, min 10000, avg 20000 "
                  "sampled every 50); L0 = 4 * 1024 / 256"}
    log(f"2D 1024^2: max K error vs FP64: {res['d2q9_1024']['max_rel_error_vs_f64']}")
    return res


def fixed_tuning(spec):
    """launch shapes given as
    cpt,waves,nontemporal,local_bpc,nbr_bpc[,nbr_mode[,nbr_segs[,local_wgs[,nbr_wgs]]]]"""
    v = [int(x) for x in spec.split(",")]
    c, w, n, lb, nb = v[:5]
    opt = [v[k] if len(v) > k else 0 for k in (5, 6, 7, 8)]
    return {"local_cpt": c, "nbr_waves": w, "nontemporal": n, "local_bpc": lb,
            "nbr_bpc": nb, "nbr_mode": opt[0], "nbr_segs": opt[1], "local_wgs": opt[2],
            "nbr_wgs": opt[3], "chunk_steps": 200, "fixed": True}


def large_lattice(ctx, seed, steps, solve, tune=""):
    """The 512^3 low-porosity Gaussian-random-field lattice: stepping rate in
    FP32 and, with solve, the full three-direction FP32 solve.  tune: fixed
    launch shapes instead of the autotuner (profiles see only those)."""
    from flowperm import solver, workloads
    t0 = time.time()
    m = workloads.large_grf(seed)
    log(f"512^3 GRF: porosity {1 - m.mean():.4f} ({time.time() - t0:.1f} s to generate)")
    from flowperm import autotune
    if tune:
        TUNED["f32"] = fixed_tuning(tune)
    else:
        TUNED["f32"] = autotune.tune(ctx, m, 3, 1, steps=10, use_cache=False)
        log(f"autotune f32 (512^3): {({k: TUNED['f32'][k] for k in autotune.FIELDS})}")
        log(f"  neighbour ms {TUNED['f32']['nbr_ms']}")
    apply_tuning(ctx, "f32")
    o = bench_steps(ctx, m, 1, 10, steps)
    res = {"porosity": float(1 - m.mean()), "n_fluid": int((m == 0).sum()),
           "mlups": o.n_fluid * o.steps / o.total_ms / 1e3,
           "ms_per_step": o.total_ms / o.steps,
           "step_kernel_GBs": o.n_fluid * 152 / (0.5 * (o.local_ms + o.nbr_ms) / 1e3) / 1e9,
           "local_ms": o.local_ms, "neighbour_ms": o.nbr_ms,
           "tuning": {k: TUNED["f32"].get(k, 0) for k in ("local_cpt", "nbr_waves", "nbr_bpc", "nbr_mode",
                                                          "nbr_segs", "local_wgs", "nbr_wgs")}}
    log(f"512^3: {res['mlups']:.0f} MLUPS, {res['step_kernel_GBs']:.0f} GB/s")
    if solve:
        row = solve_logged(ctx, m, "f32", **SOLVE_CFG)
        res["solve"] = {"seconds": row["wall_s"],
                        "steps": [row[f"steps_{t}"] for t in ("fx", "fy", "fz")],
                        "converged": [row[f"conv_{t}"] for t in ("fx", "fy", "fz")],
                        "K": k_matrix(row).tolist(), "k_mean": row["k_mean"], "fa": row["fa"],
                        "last_check": {t: {"step": row[f"check_{t}"]["step"],
                                           "drift_over_tol": row[f"check_{t}"]["ratio"][0],
                                           "hold": row[f"check_{t}"]["hold"]}
                                       for t in ("fx", "fy", "fz")},
                        "last_check_note": "drift/tol at the last first-phase check (the "
                                           "reference's progress line): how far from the eps "
                                           "test a run that reached max_steps stopped"}
        log(f"512^3 solve: {res['solve']}")
    return res


def split_errors(r32, r64, dirs, err, suffix):
    """max K error of the FP32 rows against FP64, apart for structures whose
    FP64 runs all converged and for the rest: a structure with no through
    path never converges and its K is a transient near zero, so an error
    relative to it says little about the arithmetic"""
    conv = [all(b[f"conv_{d}"] for d in dirs) for b in r64]
    e = [err(a, b) for a, b in zip(r32, r64)]
    ok = [x for x, c in zip(e, conv) if c]
    bad = [x for x, c in zip(e, conv) if not c]
    return {f"max_k_error_f32_vs_f64_converged{suffix}": max(ok) if ok else None,
            f"max_k_error_f32_vs_f64_unconverged{suffix}": max(bad) if bad else None,
            f"unconverged_f64{suffix}": len(bad)}


def small_batched(ctx, seed, n2=48, n3=24):
    """The reference's own dataset sizes: D2Q9 at 256^2 (its fixed lattice,
    2D defaults) and D3Q19 at 80^3 (--ff 40 --lag 1000 --eps 1e-6), each
    batched in one request server against the plain one-at-a-time path (the
    runs of a structure one after another)."""
    from flowperm import masks, solver
    res = {}
    rng = np.random.default_rng(seed)
    m2 = [masks.ellipses((256, 256), float(rng.uniform(0.4, 0.8)), seed=int(rng.integers(1 << 30)))
          for _ in range(n2)]
    m3 = [masks.ellipsoids((80, 80, 80), float(rng.uniform(0.4, 0.8)), seed=int(rng.integers(1 << 30)))
          for _ in range(n3)]
    from flowperm.server import Server

    def batched(ms, lattice, precision, cfg):
        apply_tuning(ctx, "f32", nbr_mode=2)
        with Server(lattice, precision, ctx=ctx) as srv:
            t0 = time.perf_counter()
            ids = [srv.submit(m, **cfg) for m in ms]
            last = t0
            while srv.step():
                if time.perf_counter() - last > 20:
                    last = time.perf_counter()
                    log(f"  ... batched {lattice}D {precision}: {srv.stats()['completed']}/{len(ms)} done")
            rows = [srv.poll(i)[0] for i in ids]
            return rows, time.perf_counter() - t0

    def plain(ms, lattice, precision, cfg, limit):
        apply_tuning(ctx, "f64" if precision == "f64" else "f32", nbr_mode=2)
        t0 = time.perf_counter()
        rows = []
        for m in ms[:limit]:
            f = solver.solve_permeability3d if lattice == 3 else solver.solve_permeability2d
            rows.append(f(m, ctx=ctx, precision=precision, concurrent=False, **cfg))
        return rows, time.perf_counter() - t0

    cfg2 = {}
    r32, t32 = batched(m2, 2, "f32", cfg2)
    rref, tref = batched(m2, 2, "f32ref", cfg2)
    p64, t64 = plain(m2, 2, "f64", cfg2, 8)
    pref, tpref = plain(m2, 2, "f32ref", cfg2, 8)
    res["d2q9_256"] = {
        "structures": n2, "config": "ellipses 3:1, porosity U(0.4, 0.8), reference 2D defaults",
        "batched_f32_tensors_per_s": n2 / t32, "batched_f32ref_tensors_per_s": n2 / tref,
        "plain_f64_tensors_per_s": 8 / t64, "plain_f32ref_tensors_per_s": 8 / tpref,
        "speedup_f32_vs_plain_f64": (n2 / t32) / (8 / t64),
        "speedup_f32ref_batched_vs_plain": (n2 / tref) / (8 / tpref),
        **split_errors(r32, p64, ("fx", "fy"), k2_error, ""),
        "f32ref_rows_equal_plain": all(a["k_xx"] == b["k_xx"] and a["k_yy"] == b["k_yy"]
                                       for a, b in zip(rref, pref)),
        "steps_differ_f32_vs_f64": sum((a["steps_fx"], a["steps_fy"]) != (b["steps_fx"], b["steps_fy"])
                                       for a, b in zip(r32, p64))}
    # the default eps 1e-3 lets the FP32 and FP64 runs stop at different
    # steps; at a tight eps the difference left is the arithmetic's
    tight = dict(eps=1e-6)
    t32r, _ = batched(m2[:8], 2, "f32", tight)
    t64r, _ = batched(m2[:8], 2, "f64", tight)
    res["d2q9_256"].update(split_errors(t32r, t64r, ("fx", "fy"), k2_error, "_eps1e-6"))
    log(f"2D 256^2: {res['d2q9_256']}")
    r32, t32 = batched(m3, 3, "f32", SOLVE_CFG)
    p64, t64 = plain(m3, 3, "f64", SOLVE_CFG, 6)
    res["d3q19_80"] = {
        "structures": n3, "config": "ellipsoids 3:1, porosity U(0.4, 0.8), --ff 40 --lag 1000 --eps 1e-6",
        "batched_f32_tensors_per_s": n3 / t32, "plain_f64_tensors_per_s": 6 / t64,
        "speedup_f32_vs_plain_f64": (n3 / t32) / (6 / t64),
        **split_errors(r32, p64, ("fx", "fy", "fz"), k_error, "")}
    log(f"3D 80^3: {res['d3q19_80']}")
    return res


def cpu_aggregate(ref, td, cores=None, extra_steps=100):
    """The reference's own layout on this host: one single-threaded solver
    process per core at once (scripts/solve_permeability3d.py:250), each on a
    64^3 sphere pack (porosity 0.6, the 64^3 parity lattice's recipe), runs of
    3x1 and 3x(1+extra_steps) steps; the difference of the slowest process's
    wall times over the summed lattice updates is the aggregate rate."""
    from flowperm import masks, structio
    cores = cores or min(16, os.cpu_count() or 1)
    m = masks.spheres((64, 64, 64), 0.6, seed=0)
    n_fluid = int(m.size - np.count_nonzero(m))
    raw = td / "agg.raw"
    structio.save_raw(raw, m)

    def wave(n):
        t0 = time.perf_counter()
        procs = [subprocess.Popen([str(ref), str(raw), str(td / f"agg{n}_{i}.csv"), "--quiet",
                                   "--max-steps", str(n), "--avg-steps", "0", "--lag", "1000000",
                                   "--ff", "40"], stdout=subprocess.DEVNULL,
                                  stderr=subprocess.PIPE) for i in range(cores)]
        for p in procs:
            p.wait(timeout=600)
            if p.returncode not in (0, 3):
                raise RuntimeError(f"reference failed: {p.stderr.read()[-200:]}")
        return time.perf_counter() - t0

    t1 = wave(1)
    tn = wave(1 + extra_steps)
    updates = cores * 3 * extra_steps * n_fluid
    agg = updates / (tn - t1) / 1e6
    return {"value": agg, "unit": "MLUPS", "cores": cores, "per_core": agg / cores,
            "sample": f"{cores} concurrent single-threaded lbm3d-perm processes on a 64^3 sphere "
                      f"pack ({n_fluid} fluid cells), 3x1 vs 3x{1 + extra_steps} steps each: "
                      f"{updates} fluid updates in {tn - t1:.2f} s"}


def cpu_baseline(mask, budget_s=30.0):
    """The reference solver (LMB3d/lbm3d-perm.cpp built by oracle/Makefile)
    on this host, one single-threaded process, on a bounded sample of the same
    workload: two runs of 1 and 3 steps per direction on the same 256^3 mask;
    their difference (6 lattice steps) isolates stepping from I/O and setup."""
    ref = REPO / "oracle" / "_ref" / "lbm3d-perm"
    from flowperm import structio
    n_fluid = int(mask.size - np.count_nonzero(mask))
    with tempfile.TemporaryDirectory() as td:
        raw = Path(td) / "m.raw"
        structio.save_raw(raw, mask)
        if ref.exists():
            times = {}
            for n in (1, 3):
                csv = Path(td) / f"o{n}.csv"
                t0 = time.perf_counter()
                p = subprocess.run([str(ref), str(raw), str(csv), "--quiet", "--max-steps", str(n),
                                    "--avg-steps", "0", "--lag", "1000000", "--ff", "40"],
                                   capture_output=True, text=True, timeout=600)
                times[n] = time.perf_counter() - t0
                if p.returncode not in (0, 3):
                    raise RuntimeError(f"reference failed: {p.stderr[-300:]}")
            dt = times[3] - times[1]
            updates = 3 * 2 * n_fluid
            out = {"value": updates / dt / 1e6, "unit": "MLUPS", "cores": 1, "kind": "reference",
                   "sample": f"lbm3d-perm built from the reference source (-O3 -march=broadwell "
                             f"-ffast-math) on the same 256^3 mask: runs of 3x1 and 3x3 steps "
                             f"(--max-steps 1 / 3 --avg-steps 0), difference = 6 lattice steps "
                             f"= {updates} fluid updates in {dt:.2f} s "
                             f"(wall {times[1]:.1f} s + {times[3]:.1f} s)"}
            try:
                out["all_cores"] = cpu_aggregate(ref, Path(td))
            except Exception as e:                      # reported, never fatal
                out["all_cores"] = {"error": str(e)[:200]}
            return out
        sys.path.insert(0, str(REPO))
        from oracle import binding
        t0 = time.perf_counter()
        binding.run_steps3d(mask, (1e-5, 0, 0), 1)
        t1 = time.perf_counter()
        binding.run_steps3d(mask, (1e-5, 0, 0), 3)
        t2 = time.perf_counter()
        dt = (t2 - t1) - (t1 - t0)
        return {"value": 2 * n_fluid / dt / 1e6, "unit": "MLUPS", "cores": 1, "kind": "port",
                "sample": "CPU oracle (oracle/lbm3d_oracle.c), 3 vs 1 steps on the same mask"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--mask", default="ellipsoids", choices=["ellipsoids", "open"],
                    help="workload mask (open = all fluid, a diagnostic)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 rate")
    ap.add_argument("--no-small", action="store_true", help="skip the small-lattice step times")
    ap.add_argument("--tune", default="", help="fixed launch shapes "
                    "'cpt,waves,nt,local_bpc,nbr_bpc[,nbr_mode[,nbr_segs[,local_wgs[,nbr_wgs]]]]' for both "
                    "precisions instead of "
                    "running the autotuner")
    ap.add_argument("--batched", type=int, default=48, help="structures of the batched set "
                    "solved batched in FP32 (0 = skip)")
    ap.add_argument("--batched-fp64", type=int, default=12, help="of those, how many the plain "
                    "FP64 path solves one at a time for the speedup and the K error")
    ap.add_argument("--serving", type=int, default=40, help="requests in the serving sample (0 = skip)")
    ap.add_argument("--workers", type=int, default=8, help="processes generating structures")
    ap.add_argument("--study", default="", choices=["", "precision", "large", "large-solve",
                                                     "batched-full", "small-batched", "small-steps"],
                    help="instead of the default line: the 256^3 / 1024^2 precision study, "
                         "the 512^3 GRF lattice (stepping rate; large-solve adds the full solve), "
                         "the whole batched set, the reference's dataset sizes batched, or the "
                         "per-step times of the small lattices (small-steps)")
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this solver has no multi-GPU path (a batch would "
              "shard by instance with no exchange between GPUs). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from flowperm import solver
    if a.study:
        ctx = solver.Context()
        from flowperm import autotune
        if a.study == "small-steps":             # per-step times of the small lattices only
            out = {"study": "small-steps", **small_lattice_steps(ctx, a.seed)}
            ctx.close()
            print(json.dumps(out), flush=True)
            return
        if not (a.tune and a.study.startswith("large")):
            mask, _ = workload_mask(256, a.seed)
            for prec, name in ((1, "f32"), (0, "f64")):
                TUNED[name] = autotune.tune(ctx, mask, 3, prec, steps=40, use_cache=False)
        steps = max(2, a.steps & ~1)
        if a.study == "precision":
            out = {"study": "precision", **precision_study(ctx, a.seed, steps)}
        elif a.study == "small-batched":
            out = {"study": "small-batched", **small_batched(ctx, a.seed)}
        elif a.study == "batched-full":
            out = {"study": "batched-full", **batched_full(ctx, 512, 24, a.seed, a.workers)}
        else:
            out = {"study": "large", **large_lattice(ctx, a.seed, steps, a.study == "large-solve",
                                                     a.tune)}
        ctx.close()
        print(json.dumps(out), flush=True)
        return
    mask, por_target = workload_mask(a.size, a.seed, a.mask)
    n_fluid = int(mask.size - np.count_nonzero(mask))
    ctx = solver.Context()
    # a step pair is one local and one neighbour step: even counts (an odd
    # request is rounded up, and the line reports what ran)
    steps = max(2, (a.steps + 1) & ~1)
    warm = max(0, (a.warmup + 1) & ~1)

    from flowperm import autotune
    tuned = {}
    for prec, name in ((1, "f32"), (0, "f64")):
        if prec == 0 and a.no_fp64:
            continue
        if a.tune:
            tuned[name] = fixed_tuning(a.tune)
            continue
        rec = autotune.tune(ctx, mask, 3, prec, steps=40, use_cache=False)
        tuned[name] = rec
        log(f"autotune {name}: best {({k: rec[k] for k in autotune.FIELDS})}")
        log(f"  local ms {rec['local_ms']}")
        log(f"  neighbour ms {rec['nbr_ms']}")
    TUNED.update(tuned)

    def shape(r):
        return (r["local_cpt"], r["nbr_waves"], r["nontemporal"], r["local_bpc"], r["nbr_bpc"],
                r.get("nbr_mode", 0), r.get("nbr_segs", 0), r.get("local_wgs", 0), r.get("nbr_wgs", 0))
    t32 = shape(tuned["f32"])
    log(f"FP32 D3Q19 {a.size}^3: warmup {warm}, timed {steps} steps")
    o32 = bench_steps(ctx, mask, 1, warm, steps, t32)
    ms_step = o32.total_ms / o32.steps
    mlups = o32.n_fluid * o32.steps / (o32.total_ms / 1e3) / 1e6
    kern_ms = 0.5 * (o32.local_ms + o32.nbr_ms)
    bytes_launch = o32.n_fluid * o32.bytes_per_update
    achieved = bytes_launch / (kern_ms / 1e3) / 1e9
    gbs = {"k_local": bytes_launch / (o32.local_ms / 1e3) / 1e9,
           "k_nbr": bytes_launch / (o32.nbr_ms / 1e3) / 1e9}
    dominant = "k_nbr" if o32.nbr_ms >= o32.local_ms else "k_local"
    log(f"FP32: {mlups:.1f} MLUPS, {ms_step:.4f} ms/step as graph replays "
        f"({o32.eager_total_ms / o32.steps:.4f} launched one by one); local {o32.local_ms:.4f} ms "
        f"({gbs['k_local'] / HBM_PEAK_GBS:.3f} of peak), neighbour {o32.nbr_ms:.4f} ms "
        f"({gbs['k_nbr'] / HBM_PEAK_GBS:.3f}), monitor {o32.monitor_ms:.4f} ms; "
        f"step kernels {achieved:.0f} GB/s = {achieved / HBM_PEAK_GBS:.3f} of peak")

    extra = {}
    of = bench_steps(ctx, mask, 1, warm, max(2, (steps // 4) & ~1), t32, flux=1)
    extra["fp32_flux_every_step"] = {
        "mlups": of.n_fluid * of.steps / (of.total_ms / 1e3) / 1e6,
        "ms_per_step": of.total_ms / of.steps, "local_ms": of.local_ms, "neighbour_ms": of.nbr_ms,
        "monitor_ms": of.monitor_ms,
        "note": "the averaging phase: exact flux sum and convergence update on every step"}
    log(f"FP32 with the flux on every step: {extra['fp32_flux_every_step']['mlups']:.1f} MLUPS; "
        f"local {of.local_ms:.4f} ms, neighbour {of.nbr_ms:.4f} ms, monitor {of.monitor_ms:.4f} ms")
    copy_ms = C.c_double(0)
    copy_bytes = 2 * 1024 ** 3
    if ctx.lib.fp_bench_copy(ctx.handle, copy_bytes, 20, C.byref(copy_ms)) == 0:
        extra["copy_GBs"] = round(2 * copy_bytes / (copy_ms.value / 1e3) / 1e9, 1)
        log(f"plain copy of {copy_bytes >> 20} MiB: {extra['copy_GBs']:.0f} GB/s (read + write)")
    if not a.no_fp64:
        s64 = max(2, (steps // 2) & ~1)
        t64 = shape(tuned["f64"])
        o64 = bench_steps(ctx, mask, 0, max(2, warm // 2), s64, t64)
        k64 = 0.5 * (o64.local_ms + o64.nbr_ms)
        extra["fp64"] = {
            "mlups": o64.n_fluid * o64.steps / (o64.total_ms / 1e3) / 1e6,
            "ms_per_step": o64.total_ms / o64.steps,
            "local_ms": o64.local_ms, "neighbour_ms": o64.nbr_ms,
            "step_kernel_GBs": o64.n_fluid * o64.bytes_per_update / (k64 / 1e3) / 1e9}
        log(f"FP64: {extra['fp64']['mlups']:.1f} MLUPS, {extra['fp64']['step_kernel_GBs']:.0f} GB/s")
    if not a.no_small:
        extra["small_lattices"] = small_lattice_steps(ctx, a.seed)
    if a.batched > 0:
        fp64_gbs = extra.get("fp64", {}).get("step_kernel_GBs")
        extra["batched"] = batched_sample(ctx, a.batched, min(a.batched, a.batched_fp64), a.seed,
                                          a.workers, fp64_gbs)
    if a.serving > 0:
        extra["serving"] = serving_sample(ctx, a.serving, a.seed, a.workers)
    ctx.close()

    traffic = measured_traffic(n_fluid) if a.mask == "ellipsoids" else None
    cpu = None
    if not a.no_cpu:
        log("CPU baseline: reference solver on this host ...")
        try:
            cpu = cpu_baseline(mask)
            log(f"CPU: {cpu['value']:.2f} MLUPS ({cpu['kind']})")
        except Exception as e:                       # reported, never fatal to the GPU number
            log(f"CPU baseline failed: {e}")

    line = {
        "metric": "D3Q19 MLUPS + %HBM peak at 256^3; batched K-tensor solves/s, p99 latency",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": int(o32.steps),
        "warmup": warm,
        "warmup_requested": a.warmup,
        "ms_per_step": round(ms_step, 5),
        # the same steps counted over every lattice node, solid ones included
        "all_node_mlups": round(mlups * mask.size / n_fluid, 2),
        "eager_ms_per_step": round(o32.eager_total_ms / o32.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"D3Q19 stream-collide, FP32 shifted populations (FP64 flux sums), "
                        f"{a.size}^3 periodic lattice of overlapping 3:1 ellipsoids "
                        f"(15 deg jitter, seed {a.seed}), porosity {1 - mask.mean():.4f}",
            "lattice": [a.size] * 3, "n_fluid": n_fluid, "n_cells": int(mask.size),
            "force": "2.5e-7 * 40 along x", "parallelism": "single GPU"},
        "roofline": {
            "bound": "hbm", "achieved": round(gbs[dominant], 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(gbs[dominant] / HBM_PEAK_GBS, 4),
            "traffic": (traffic[{"k_nbr": "nbr_bytes", "k_local": "local_bytes"}[dominant]]
                        if traffic else None),
            "kernel": f"{dominant}<D3Q19,F32> (the dominant step kernel: the longer of the two, "
                      f"one launch per step each): 152 B per fluid update x {n_fluid} fluid "
                      f"cells = {bytes_launch:.0f} B per launch over its mean event-timed duration",
            "traffic_source": (f"{TRAFFIC_PROFILE}: rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1 "
                               f"(factors from a copy of known size), committed counter runs of "
                               f"this lattice, not measured in this run") if traffic else None,
            "kernels": {
                "k_local": {"ms": round(o32.local_ms, 5), "achieved": round(gbs["k_local"], 1),
                            "frac": round(gbs["k_local"] / HBM_PEAK_GBS, 4),
                            "traffic": traffic["local_bytes"] if traffic else None},
                "k_nbr": {"ms": round(o32.nbr_ms, 5), "achieved": round(gbs["k_nbr"], 1),
                          "frac": round(gbs["k_nbr"] / HBM_PEAK_GBS, 4),
                          "traffic": traffic["nbr_bytes"] if traffic else None},
                "step_mean": {"ms": round(kern_ms, 5), "achieved": round(achieved, 1),
                              "frac": round(achieved / HBM_PEAK_GBS, 4)},
                "k_finish_ms": round(o32.monitor_ms, 5)},
            "copy_GBs": extra.get("copy_GBs")},
        "autotune": tuned,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
