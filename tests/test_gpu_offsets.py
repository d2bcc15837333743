"""
The neighbour kernels address populations with 32-bit byte offsets from the
instance's base pointer when its arrays span less than 4 GiB, and with 64-bit
addresses otherwise.  FLOWPERM_WIDE_OFFSETS=1 forces the 64-bit kernels on
small lattices; both must read and write the same slots, so every result is
bit-identical between them, in both neighbour-step forms, in 3D and in 2D.
"""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS3 = [f"K_{a}{b}" for a in "xyz" for b in "xyz"] + [f"steps_{t}" for t in ("fx", "fy", "fz")] + \
        [f"{p}q{a}_{t}" for p in ("", "se_") for a in "xyz" for t in ("fx", "fy", "fz")]
KEYS2 = ["k_xx", "k_xy", "k_yx", "k_yy", "steps_fx", "steps_fy"] + \
        [f"{p}q{a}_{t}" for p in ("", "se_") for a in "xy" for t in ("fx", "fy")]


def _ctx(mode):
    from flowperm import _lib, solver
    ctx = solver.Context()
    t = _lib.Tuning()
    ctx.lib.fp_get_tuning(ctx.handle, C.byref(t))
    t.nbr_mode = mode
    assert ctx.lib.fp_set_tuning(ctx.handle, C.byref(t)) == 0
    return ctx


@pytest.fixture(scope="module", params=[0, 1], ids=["row_segments", "compact"])
def form_ctx(request):
    ctx = _ctx(request.param)
    yield ctx
    ctx.close()


def _both(fn):
    """fn() with the 32-bit-offset kernels, then with the 64-bit ones."""
    old = os.environ.pop("FLOWPERM_WIDE_OFFSETS", None)
    try:
        a = fn()
        os.environ["FLOWPERM_WIDE_OFFSETS"] = "1"
        b = fn()
    finally:
        os.environ.pop("FLOWPERM_WIDE_OFFSETS", None)
        if old is not None:
            os.environ["FLOWPERM_WIDE_OFFSETS"] = old
    return a, b


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(12, 10, 65), (8, 6, 130)])
def test_wide_offsets_bitwise_3d(form_ctx, precision, shape):
    from flowperm import masks, solver
    m = masks.spheres(shape, 0.55, 2, 3, seed=sum(shape) + 1)
    cfg = dict(ff=40, eps=1e-6, lag=200, min_steps=400, avg_steps=200, max_steps=2000,
               precision=precision)
    print(f"shape {shape} {precision}", flush=True)
    a, b = _both(lambda: solver.solve_permeability3d(m, ctx=form_ctx, velocity=True, **cfg))
    for k in KEYS3:
        assert a[k] == b[k], k
    assert np.array_equal(a["velocity"], b["velocity"])


@pytest.mark.parametrize("precision", ["f64", "f32", "f32ref"])
def test_wide_offsets_bitwise_2d(form_ctx, precision):
    from flowperm import masks, solver
    m = masks.ellipses((40, 70), 0.6, rmin=1.5, rmax=3.0, seed=11, axis=(1, 2))
    cfg = dict(lag=500, min_steps=1000, max_steps=3000, avg_steps=600, avg_sample=50)
    print(f"2d {precision}", flush=True)
    a, b = _both(lambda: solver.solve_permeability2d(m, ctx=form_ctx, precision=precision,
                                                     velocity=True, **cfg))
    for k in KEYS2:
        assert a[k] == b[k], k
    assert np.array_equal(a["velocity"], b["velocity"])
