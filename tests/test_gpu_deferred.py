"""
The deferred finish: on lattices with at most 1024 work items per step kernel
(up to 4096 while the flux is summed on few steps) each step kernel applies
the previous step's flux reduction and convergence update itself (every
workgroup for its own instance), and a chunk ends with a single finish launch,
instead of finish launches.  It must not change a single bit: every result
with the deferred finish on every chunk it applies to (finish plans off,
FLOWPERM_SPARSE_FINISH=0) equals the one with finish launches
(FLOWPERM_DEFER_FINISH=0), for both neighbour-step forms, a batch mixing them,
3D and 2D in every precision, with the velocity fields and the
convergence-check record.  fp_bench_steps reports which path ran.
"""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS3 = [f"K_{a}{b}" for a in "xyz" for b in "xyz"] + [f"steps_{t}" for t in ("fx", "fy", "fz")] + \
        [f"{p}q{a}_{t}" for p in ("", "se_") for a in "xyz" for t in ("fx", "fy", "fz")] + \
        [f"check_{t}" for t in ("fx", "fy", "fz")]
KEYS2 = ["k_xx", "k_xy", "k_yx", "k_yy", "steps_fx", "steps_fy", "fieldres_fx", "fieldres_fy",
         "check_fx", "check_fy"] + [f"{p}q{a}_{t}" for p in ("", "se_") for a in "xy" for t in ("fx", "fy")]


ENV = ("FLOWPERM_DEFER_FINISH", "FLOWPERM_SPARSE_FINISH")


def _with(env, fn):
    """fn() with exactly these finish settings"""
    old = {k: os.environ.pop(k, None) for k in ENV}
    try:
        os.environ.update(env)
        return fn()
    finally:
        for k, v in old.items():
            os.environ.pop(k, None)
            if v is not None:
                os.environ[k] = v


def _both(fn):
    """fn() with the deferred finish wherever it applies, then with finish
    launches"""
    return (_with({"FLOWPERM_SPARSE_FINISH": "0"}, fn),
            _with({"FLOWPERM_DEFER_FINISH": "0"}, fn))


def _ctx(mode):
    from flowperm import _lib, solver
    ctx = solver.Context()
    t = _lib.Tuning()
    ctx.lib.fp_get_tuning(ctx.handle, C.byref(t))
    t.nbr_mode = mode
    assert ctx.lib.fp_set_tuning(ctx.handle, C.byref(t)) == 0
    return ctx


def _bench(ctx, mask, lattice, prec, flux_every_step=0):
    from flowperm import _lib
    if lattice == 3:
        nz, ny, nx = mask.shape
    else:
        (ny, nx), nz = mask.shape, 1
    out = _lib.BenchOut()
    F = np.array([1e-5, 0.0, 0.0])
    m = np.ascontiguousarray(mask, np.uint8)
    st = ctx.lib.fp_bench_steps(ctx.handle, _lib.u8ptr(m), nx, ny, nz, lattice, prec, _lib.f64ptr(F),
                                2, 20, None, flux_every_step, C.byref(out))
    assert st == 0, ctx.last_error()
    return out


def test_the_deferred_path_runs_where_it_applies(gpu_ctx):
    """Without finish plans: up to 1024 partials per step kernel always; up
    to 4096 while the flux is summed on few steps (the bench's steps: a check
    every 2000); more never.  With them (the default), quiet steps take the
    plan, whose few finish launches cost less than the deferred update in
    every step kernel, and steps that sum the flux every time take the
    deferred finish where it applies (up to 1024 partials)."""
    from flowperm import masks
    small3 = masks.spheres((24, 20, 28), 0.6, 2, 4, seed=3)
    small2 = masks.ellipses((256, 256), 0.6, seed=3)
    medium3 = masks.spheres((80, 80, 80), 0.6, 3, 6, seed=3)     # ~600 local, 3200 neighbour items
    large3 = masks.spheres((160, 160, 160), 0.6, 3, 6, seed=3)   # ~4800 local items
    cases = ((small3, 3), (small2, 2), (medium3, 3), (large3, 3))
    print("deferred flags", flush=True)

    def flags(flux):
        return tuple(_bench(gpu_ctx, m, lat, 1, flux).deferred for m, lat in cases)
    a, b = _both(lambda: flags(0))
    assert a == (1, 1, 1, 0) and b == (0, 0, 0, 0)
    assert _with({}, lambda: flags(0)) == (0, 0, 0, 0)
    assert _with({}, lambda: flags(1)) == (1, 1, 0, 0)


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["row_segments", "compact", "per_structure"])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_deferred_finish_bitwise_3d(mode, precision):
    from flowperm import masks, solver
    ctx = _ctx(mode)
    try:
        cfg = dict(ff=40, eps=1e-6, lag=200, min_steps=400, avg_steps=200, max_steps=3000,
                   precision=precision)
        for shape, por, seed in (((24, 20, 28), 0.6, 5), ((12, 10, 70), 0.35, 6)):
            m = masks.spheres(shape, por, 2, 4, seed=seed) if por > 0.5 else \
                masks.grf(shape, porosity=por, corr=(3, 3, 6), seed=seed)
            print(f"3d mode {mode} {precision} {shape}", flush=True)
            a, b = _both(lambda: solver.solve_permeability3d(m, ctx=ctx, velocity=True, **cfg))
            for k in KEYS3:
                assert a[k] == b[k], k
            assert np.array_equal(a["velocity"], b["velocity"])
    finally:
        ctx.close()


def test_deferred_finish_bitwise_mixed_batch():
    """a dense and a sparse structure in one server batch: the two
    neighbour-step forms in one step, each instance's state handed on by its
    own form's kernel"""
    from flowperm import masks
    from flowperm.server import Server
    dense = masks.spheres((20, 20, 24), 0.8, 3, 5, seed=5)
    sparse = masks.grf((24, 28, 70), porosity=0.3, corr=(3, 3, 6), seed=6)
    cfg = dict(ff=40, eps=1e-6, lag=100, min_steps=300, avg_steps=100, max_steps=2000)

    ctx = _ctx(2)                      # each structure's cheaper form: both forms in one step

    def run():
        with Server(3, "f32", ctx=ctx) as srv:
            ids = [srv.submit(m, **cfg) for m in (dense, sparse)]
            while srv.step():
                pass
            return [srv.poll(i)[0] for i in ids]
    print("mixed batch", flush=True)
    try:
        a, b = _both(run)
    finally:
        ctx.close()
    for ra, rb in zip(a, b):
        for k in KEYS3:
            assert ra[k] == rb[k], k


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_deferred_finish_bitwise_medium_3d(gpu_ctx, precision):
    """an 80^3 lattice: more than 1024 neighbour-step partials, so the chunks
    of its averaging phase (flux on every step) run with a finish launch per
    step and the first phase's deferred; the whole solve equals the one with
    a finish launch after every step"""
    from flowperm import masks, solver
    m = masks.ellipsoids((80, 80, 80), 0.6, seed=4)
    cfg = dict(ff=40, eps=1e-6, lag=200, min_steps=400, avg_steps=300, max_steps=3000,
               precision=precision)
    print(f"80^3 {precision}", flush=True)
    a, b = _both(lambda: solver.solve_permeability3d(m, ctx=gpu_ctx, velocity=True, **cfg))
    for k in KEYS3:
        assert a[k] == b[k], k
    assert np.array_equal(a["velocity"], b["velocity"])


@pytest.mark.parametrize("precision", ["f64", "f32", "f32ref"])
def test_deferred_finish_bitwise_2d(gpu_ctx, precision):
    from flowperm import masks, solver
    m = masks.ellipses((48, 64), 0.6, rmin=1.5, rmax=3.0, seed=12, axis=(1, 2))
    cfg = dict(lag=500, min_steps=1000, max_steps=4000, avg_steps=600, avg_sample=50, validate=30.0)
    print(f"2d {precision}", flush=True)
    a, b = _both(lambda: solver.solve_permeability2d(m, ctx=gpu_ctx, precision=precision,
                                                     velocity=True, **cfg))
    for k in KEYS2 + ["val_q", "val_steps"]:
        assert a[k] == b[k], k
    assert np.array_equal(a["velocity"], b["velocity"])
