"""
Neighbour items of S x-adjacent row segments (nbr_segs 1, 2, 4, 8): the same
results bit for bit whatever S and however many rows an item stacks, on rows
of 4 segments (3D) and 8 segments (2D), whose segment counts every S tried
divides, in FP32 and FP64, with the velocity fields.  The default (S = 4) is
also what every other GPU test runs; the reference fixtures with rows of 70
and 130 cells (S = 2 and 1) pin it to the oracle.  The same holds for the
step kernels held to fewer resident workgroups per CU (local_wgs, nbr_wgs).
"""
import ctypes as C

import numpy as np
import pytest

from test_gpu_deferred import KEYS2, KEYS3

pytestmark = pytest.mark.gpu


def _ctx(waves, segs):
    from flowperm import _lib, solver
    ctx = solver.Context()
    t = _lib.Tuning()
    ctx.lib.fp_get_tuning(ctx.handle, C.byref(t))
    t.nbr_mode, t.nbr_waves, t.nbr_segs = 0, waves, segs
    assert ctx.lib.fp_set_tuning(ctx.handle, C.byref(t)) == 0
    return ctx


SHAPES = [(4, 1), (4, 2), (4, 4), (8, 8), (8, 4), (2, 2)]


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_segments_per_item_bitwise_3d(precision):
    from flowperm import masks, solver
    m = masks.spheres((10, 12, 256), 0.6, 2, 4, seed=21)            # 4 segments per row
    cfg = dict(ff=40, eps=1e-6, lag=100, min_steps=200, avg_steps=100, max_steps=600,
               precision=precision)
    res = []
    for waves, segs in SHAPES:
        print(f"3d {precision} waves {waves} segs {segs}", flush=True)
        ctx = _ctx(waves, segs)
        try:
            res.append(solver.solve_permeability3d(m, ctx=ctx, velocity=True, **cfg))
        finally:
            ctx.close()
    for r in res[1:]:
        for k in KEYS3:
            assert r[k] == res[0][k], k
        assert np.array_equal(r["velocity"], res[0]["velocity"])


@pytest.mark.parametrize("precision", ["f32", "f32ref", "f64"])
def test_segments_per_item_bitwise_2d(precision):
    from flowperm import masks, solver
    m = masks.ellipses((40, 512), 0.6, rmin=1.5, rmax=3.0, seed=22, axis=(1, 2))   # 8 segments
    cfg = dict(lag=150, min_steps=300, max_steps=900, avg_steps=200, avg_sample=10)
    res = []
    for waves, segs in SHAPES:
        print(f"2d {precision} waves {waves} segs {segs}", flush=True)
        ctx = _ctx(waves, segs)
        try:
            res.append(solver.solve_permeability2d(m, ctx=ctx, precision=precision, velocity=True,
                                                   **cfg))
        finally:
            ctx.close()
    for r in res[1:]:
        for k in KEYS2:
            assert r[k] == res[0][k], k
        assert np.array_equal(r["velocity"], res[0]["velocity"])


def _ctx_caps(mode, cpt, local_wgs, nbr_wgs):
    from flowperm import _lib, solver
    ctx = solver.Context()
    t = _lib.Tuning()
    ctx.lib.fp_get_tuning(ctx.handle, C.byref(t))
    t.nbr_mode, t.local_cpt, t.local_wgs, t.nbr_wgs = mode, cpt, local_wgs, nbr_wgs
    assert ctx.lib.fp_set_tuning(ctx.handle, C.byref(t)) == 0
    back = _lib.Tuning()
    ctx.lib.fp_get_tuning(ctx.handle, C.byref(back))
    assert (back.local_wgs, back.nbr_wgs) == (local_wgs, nbr_wgs)
    return ctx


# (cells per thread, local_wgs, nbr_wgs): no cap, caps of 1-6 workgroups per CU,
# and 7, which no LDS size reaches on a 160 KiB CU (launched without a cap)
CAPS = [(2, 0, 0), (2, 2, 5), (1, 3, 6), (4, 2, 2), (1, 4, 3), (2, 1, 7)]


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("mode", [0, 1])
def test_workgroup_caps_bitwise_3d(precision, mode):
    """The local and neighbour kernels held to fewer resident workgroups per
    CU (local_wgs, nbr_wgs: LDS the kernels do not use) give the same solve bit
    for bit, in both neighbour forms; the first phase runs finish plans and the
    averaging phase the deferred finish, so both kernel variants are capped."""
    from flowperm import masks, solver
    m = masks.spheres((10, 12, 256), 0.6, 2, 4, seed=23)
    cfg = dict(ff=40, eps=1e-6, lag=100, min_steps=200, avg_steps=100, max_steps=600,
               precision=precision)
    res = []
    for cpt, lw, nw in CAPS:
        print(f"3d {precision} mode {mode} cpt {cpt} local_wgs {lw} nbr_wgs {nw}", flush=True)
        ctx = _ctx_caps(mode, cpt, lw, nw)
        try:
            res.append(solver.solve_permeability3d(m, ctx=ctx, velocity=True, **cfg))
        finally:
            ctx.close()
    for r in res[1:]:
        for k in KEYS3:
            assert r[k] == res[0][k], k
        assert np.array_equal(r["velocity"], res[0]["velocity"])


def test_workgroup_caps_bitwise_2d():
    from flowperm import masks, solver
    m = masks.ellipses((40, 512), 0.6, rmin=1.5, rmax=3.0, seed=24, axis=(1, 2))
    cfg = dict(lag=150, min_steps=300, max_steps=900, avg_steps=200, avg_sample=10)
    res = []
    for cpt, lw, nw in CAPS:
        print(f"2d cpt {cpt} local_wgs {lw} nbr_wgs {nw}", flush=True)
        ctx = _ctx_caps(0, cpt, lw, nw)
        try:
            res.append(solver.solve_permeability2d(m, ctx=ctx, precision="f32", velocity=True, **cfg))
        finally:
            ctx.close()
    for r in res[1:]:
        for k in KEYS2:
            assert r[k] == res[0][k], k
        assert np.array_equal(r["velocity"], res[0]["velocity"])
