"""
The sparse finish schedule on the device: a batch that launches the finish
(exact flux sum and convergence update) after every step of its chunks gives
bit for bit the results of one that launches it only after the steps its
plan names (the default; FLOWPERM_SPARSE_FINISH=0 turns the plan off), for
3D and 2D solves in every precision, with lags that do not divide the chunk,
velocity fields, the convergence-check record, and a server batch whose
requests join at different steps.  Both runs take the finish launches of the
non-deferred path (FLOWPERM_DEFER_FINISH=0), where the plan applies.
"""
import os

import numpy as np
import pytest

from test_gpu_deferred import KEYS2, KEYS3

pytestmark = pytest.mark.gpu


def _both(fn):
    """fn() with a finish after every step, then with the finish plans"""
    old = {k: os.environ.pop(k, None) for k in ("FLOWPERM_DEFER_FINISH", "FLOWPERM_SPARSE_FINISH")}
    try:
        os.environ["FLOWPERM_DEFER_FINISH"] = "0"
        os.environ["FLOWPERM_SPARSE_FINISH"] = "0"
        a = fn()
        os.environ.pop("FLOWPERM_SPARSE_FINISH")
        b = fn()
    finally:
        for k, v in old.items():
            os.environ.pop(k, None)
            if v is not None:
                os.environ[k] = v
    return a, b


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_sparse_finish_bitwise_3d(gpu_ctx, precision):
    from flowperm import masks, solver
    m = masks.spheres((24, 20, 28), 0.6, 2, 4, seed=5)
    cfg = dict(ff=40, eps=1e-6, lag=37, min_steps=300, avg_steps=150, max_steps=3000,
               precision=precision)
    print(f"3d {precision}", flush=True)
    a, b = _both(lambda: solver.solve_permeability3d(m, ctx=gpu_ctx, velocity=True, **cfg))
    for k in KEYS3:
        assert a[k] == b[k], k
    assert np.array_equal(a["velocity"], b["velocity"])


def test_sparse_finish_bitwise_max_steps_3d(gpu_ctx):
    """a first phase that ends at max_steps (not a multiple of the lag) in
    the middle of a chunk"""
    from flowperm import masks, solver
    m = masks.ellipsoids((32, 32, 32), 0.6, seed=2)
    cfg = dict(ff=40, eps=0.0, lag=150, min_steps=0, avg_steps=90, max_steps=733, precision="f32")
    print("3d max_steps", flush=True)
    a, b = _both(lambda: solver.solve_permeability3d(m, ctx=gpu_ctx, velocity=True, **cfg))
    for k in KEYS3:
        assert a[k] == b[k], k
    assert a["steps_fx"] == 733 + 90
    assert np.array_equal(a["velocity"], b["velocity"])


def test_sparse_finish_bitwise_hold_zero(gpu_ctx):
    """--hold 0: the first check (step 37, long before min_steps) ends the
    first phase in the middle of a chunk"""
    from flowperm import masks, solver
    m = masks.spheres((24, 20, 28), 0.6, 2, 4, seed=9)
    cfg = dict(ff=40, eps=1e-6, lag=37, hold=0, min_steps=1000, avg_steps=150, max_steps=3000,
               precision="f32")
    print("3d hold 0", flush=True)
    a, b = _both(lambda: solver.solve_permeability3d(m, ctx=gpu_ctx, velocity=True, **cfg))
    for k in KEYS3:
        assert a[k] == b[k], k
    assert a["steps_fx"] == 37 + 150
    assert np.array_equal(a["velocity"], b["velocity"])


@pytest.mark.parametrize("precision", ["f64", "f32", "f32ref"])
def test_sparse_finish_bitwise_2d(gpu_ctx, precision):
    from flowperm import masks, solver
    m = masks.ellipses((48, 64), 0.6, rmin=1.5, rmax=3.0, seed=12, axis=(1, 2))
    cfg = dict(lag=130, min_steps=400, max_steps=3000, avg_steps=300, avg_sample=7, validate=30.0)
    print(f"2d {precision}", flush=True)
    a, b = _both(lambda: solver.solve_permeability2d(m, ctx=gpu_ctx, precision=precision,
                                                     velocity=True, **cfg))
    for k in KEYS2 + ["val_q", "val_steps"]:
        assert a[k] == b[k], k
    assert np.array_equal(a["velocity"], b["velocity"])


@pytest.mark.parametrize("lattice", [3, 2])
def test_default_modes_bitwise_small_lattices(gpu_ctx, lattice):
    """a small lattice with the default settings: its first-phase chunks
    take finish plans, its averaging chunks (3D: the flux on every step) the
    deferred finish, and the solve equals the one with a finish after every
    step"""
    from flowperm import masks, solver
    if lattice == 3:
        m = masks.spheres((24, 20, 28), 0.6, 2, 4, seed=7)
        cfg = dict(ff=40, eps=1e-6, lag=100, min_steps=300, avg_steps=250, max_steps=3000)

        def run():
            return solver.solve_permeability3d(m, ctx=gpu_ctx, velocity=True, precision="f32", **cfg)
        keys = KEYS3
    else:
        m = masks.ellipses((48, 64), 0.6, rmin=1.5, rmax=3.0, seed=13, axis=(1, 2))
        cfg = dict(lag=130, min_steps=400, max_steps=3000, avg_steps=300, avg_sample=7)

        def run():
            return solver.solve_permeability2d(m, ctx=gpu_ctx, precision="f32", velocity=True, **cfg)
        keys = KEYS2
    print(f"default modes {lattice}d", flush=True)
    a = run()
    old = {k: os.environ.pop(k, None) for k in ("FLOWPERM_DEFER_FINISH", "FLOWPERM_SPARSE_FINISH")}
    try:
        os.environ["FLOWPERM_DEFER_FINISH"] = "0"
        os.environ["FLOWPERM_SPARSE_FINISH"] = "0"
        b = run()
    finally:
        for k, v in old.items():
            os.environ.pop(k, None)
            if v is not None:
                os.environ[k] = v
    for k in keys:
        assert a[k] == b[k], k
    assert np.array_equal(a["velocity"], b["velocity"])


def test_sparse_finish_bitwise_staggered_server():
    """requests that join a running batch at different steps, with different
    lags: each chunk's plan is the union of the instances' events"""
    from flowperm import masks, solver
    from flowperm.server import Server
    ms = [masks.spheres((20, 20, 24), 0.7, 2, 4, seed=s) for s in (1, 2, 3)]
    cfgs = [dict(ff=40, eps=1e-6, lag=lag, min_steps=200, avg_steps=120, max_steps=2500)
            for lag in (50, 77, 300)]

    def run():
        ctx = solver.Context()
        try:
            with Server(3, "f32", ctx=ctx) as srv:
                ids = [srv.submit(ms[0], **cfgs[0])]
                for m, c in zip(ms[1:], cfgs[1:]):
                    for _ in range(2):
                        srv.step()
                    ids.append(srv.submit(m, **c))
                while srv.step():
                    pass
                return [srv.poll(i)[0] for i in ids]
        finally:
            ctx.close()
    print("staggered server", flush=True)
    a, b = _both(run)
    for ra, rb in zip(a, b):
        for k in KEYS3:
            assert ra[k] == rb[k], k


def test_random_configs_bitwise(gpu_ctx):
    """random convergence flags (lag, hold -1..3, min/max/avg steps, 2D
    sample interval and tolerances) on small lattices, 3D and 2D: the default
    settings (finish plans and the deferred finish, per chunk) against a
    finish after every step, bit for bit"""
    from flowperm import masks, solver
    rng = np.random.default_rng(2024)
    m3 = masks.spheres((20, 16, 24), 0.6, 2, 4, seed=11)
    m2 = masks.ellipses((40, 48), 0.6, rmin=1.5, rmax=3.0, seed=12, axis=(1, 2))
    dense = {"FLOWPERM_DEFER_FINISH": "0", "FLOWPERM_SPARSE_FINISH": "0"}
    for case in range(12):
        lag = int(rng.choice([7, 37, 100, 250, 1000]))
        common = dict(lag=lag, hold=int(rng.choice([-1, 0, 1, 2, 3])),
                      min_steps=int(rng.choice([0, 150, 600])),
                      max_steps=int(rng.choice([90, 801, 2500])),
                      avg_steps=int(rng.choice([0, 1, 77, 300])))
        if case % 2 == 0:
            cfg = dict(common, ff=40, eps=float(rng.choice([0.0, 1e-6, 1e-3])),
                       precision=str(rng.choice(["f32", "f64"])))

            def run():
                return solver.solve_permeability3d(m3, ctx=gpu_ctx, velocity=True, **cfg)
            keys = KEYS3
        else:
            cfg = dict(common, eps=float(rng.choice([0.0, 1e-3, 1e-2])),
                       ksigma=float(rng.choice([0.0, 3.0])), qfloor=float(rng.choice([0.0, 1e-2])),
                       avg_sample=int(rng.choice([1, 7, 50])),
                       precision=str(rng.choice(["f32", "f32ref", "f64"])))

            def run():
                return solver.solve_permeability2d(m2, ctx=gpu_ctx, velocity=True, **cfg)
            keys = KEYS2
        print(f"random case {case}: {cfg}", flush=True)
        a = run()
        old = {k: os.environ.pop(k, None) for k in dense}
        try:
            os.environ.update(dense)
            b = run()
        finally:
            for k, v in old.items():
                os.environ.pop(k, None)
                if v is not None:
                    os.environ[k] = v
        for k in keys:
            assert a[k] == b[k], (case, k, a[k], b[k])
        assert np.array_equal(a["velocity"], b["velocity"]), case
