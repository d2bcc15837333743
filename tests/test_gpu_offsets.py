"""
The step kernels address populations in one of three ways: 32-bit byte offsets
from the instance's base pointer when its arrays span less than 4 GiB, 32-bit
element indices below 2^32 elements, and 64-bit element indices beyond
(226 M fluid cells in D3Q19).  FLOWPERM_WIDE_OFFSETS=1 and =2 force the two
wider forms on small lattices; all three must read and write the same slots, so
every result is bit-identical between them, in both neighbour-step forms, in
3D and in 2D.  The last test runs a lattice past 2^32 population elements
against its half, which the 32-bit element indices still cover.
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


def _modes(fn):
    """fn() with the byte-offset kernels, then with 32-bit and with 64-bit
    element indices."""
    old = os.environ.pop("FLOWPERM_WIDE_OFFSETS", None)
    try:
        out = [fn()]
        for v in ("1", "2"):
            os.environ["FLOWPERM_WIDE_OFFSETS"] = v
            out.append(fn())
    finally:
        os.environ.pop("FLOWPERM_WIDE_OFFSETS", None)
        if old is not None:
            os.environ["FLOWPERM_WIDE_OFFSETS"] = old
    return out


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(12, 10, 65), (8, 6, 130)])
def test_wide_offsets_bitwise_3d(form_ctx, precision, shape):
    from flowperm import masks, solver
    m = masks.spheres(shape, 0.55, 2, 3, seed=sum(shape) + 1)
    cfg = dict(ff=40, eps=1e-6, lag=200, min_steps=400, avg_steps=200, max_steps=2000,
               precision=precision)
    print(f"shape {shape} {precision}", flush=True)
    a, *rest = _modes(lambda: solver.solve_permeability3d(m, ctx=form_ctx, velocity=True, **cfg))
    for b in rest:
        for k in KEYS3:
            assert a[k] == b[k], k
        assert np.array_equal(a["velocity"], b["velocity"])


@pytest.mark.parametrize("precision", ["f64", "f32", "f32ref"])
def test_wide_offsets_bitwise_2d(form_ctx, precision):
    from flowperm import masks, solver
    m = masks.ellipses((40, 70), 0.6, rmin=1.5, rmax=3.0, seed=11, axis=(1, 2))
    cfg = dict(lag=500, min_steps=1000, max_steps=3000, avg_steps=600, avg_sample=50)
    print(f"2d {precision}", flush=True)
    a, *rest = _modes(lambda: solver.solve_permeability2d(m, ctx=form_ctx, precision=precision,
                                                          velocity=True, **cfg))
    for b in rest:
        for k in KEYS2:
            assert a[k] == b[k], k
        assert np.array_equal(a["velocity"], b["velocity"])


@pytest.mark.timeout(600)
def test_past_32bit_elements(form_ctx):
    """A 512 x 512 x 1024 periodic lattice of 236 M fluid cells (4.49 G FP32
    population elements, 64-bit indices) is its 512^3 lower half (118 M cells,
    32-bit element indices) stacked twice along z: every cell's update is the
    same, so after the same steps its velocity field is the half's twice over
    and its flux (an exact sum over twice the cells, divided by twice the cell
    count) is the half's to the bit."""
    from flowperm import solver
    rng = np.random.default_rng(2024)
    half = (rng.random((512, 512, 512), dtype=np.float32) >= 0.88).astype(np.uint8)
    n_half = int(half.size - half.sum())
    assert 113_100_000 < n_half < 2 ** 31                   # the doubled lattice passes 2^32 elements
    assert 19 * 2 * n_half > 2 ** 32
    cfg = dict(lag=10, min_steps=20, max_steps=20, avg_steps=0, precision="f32")
    F = (1e-5, 2e-6, -3e-6)
    print(f"half {half.shape} fluid {n_half}", flush=True)
    a = solver.solve_direction3d(half, F, ctx=form_ctx, velocity=True, **cfg)
    full = np.concatenate([half, half], axis=0)
    del half
    print(f"full {full.shape}", flush=True)
    b = solver.solve_direction3d(full, F, ctx=form_ctx, velocity=True, **cfg)
    del full
    print(f"steps {a['steps']} {b['steps']}", flush=True)
    assert a["steps"] == b["steps"] >= 20 and a["status"] == b["status"]
    assert np.array_equal(a["q"], b["q"]) and np.all(a["q"] != 0.0)
    va, vb = a["velocity"], b["velocity"]
    assert np.array_equal(vb[:512], va) and np.array_equal(vb[512:], va)
