"""
fp_solve_dir3d_noalloc, the allocation-free variant of fp_solve_dir3d (all
of its device memory in the library's fixed scratch array), against
fp_solve_dir3d bit for bit: two calls in a row, the second on a smaller
lattice (the array keeps the first call's contents), in FP64 and FP32; a
lattice whose buffers exceed the array and an invalid mask are rejected
before any device work.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import golden, golden_flags

pytestmark = pytest.mark.gpu


def _run(ctx, fn, mask, F, **cfg):
    from flowperm import _lib
    from flowperm.solver import config3d
    m = np.ascontiguousarray(mask, np.uint8)
    nz, ny, nx = m.shape
    c = config3d(**cfg)
    Fa = np.ascontiguousarray(F, np.float64)
    out = _lib.SolveOut()
    if fn == "noalloc":
        rc = ctx.lib.fp_solve_dir3d_noalloc(ctx.handle, _lib.u8ptr(m), nx, ny, nz, C.byref(c),
                                            _lib.f64ptr(Fa), C.byref(out))
    else:
        rc = ctx.lib.fp_solve_dir3d(ctx.handle, _lib.u8ptr(m), nx, ny, nz, C.byref(c),
                                    _lib.f64ptr(Fa), C.byref(out), None)
    return rc, out


def _same(a, b):
    return (list(a.q) == list(b.q) and list(a.se) == list(b.se) and a.steps == b.steps and
            a.converged == b.converged and a.status == b.status)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_noalloc_equals_solve_dir3d_on_two_calls_in_a_row(gpu_ctx, precision):
    big, _, meta = golden("c1_64_s0")                  # the 64^3 parity lattice
    small = golden("sph16_s0")[0]
    cfg = dict(golden_flags(meta), precision=precision)
    F = 2.5e-7 * cfg["ff"]
    for mask, Fv in ((big, (F, 0.0, 0.0)), (small, (0.0, 0.0, F))):
        print(f"noalloc {precision} {mask.shape} F {Fv}", flush=True)
        rc_a, a = _run(gpu_ctx, "noalloc", mask, Fv, **cfg)
        rc_b, b = _run(gpu_ctx, "regular", mask, Fv, **cfg)
        assert rc_a == rc_b and rc_a in (0, 3), (rc_a, rc_b, gpu_ctx.last_error())
        assert a.steps > 0
        assert _same(a, b), (list(a.q), list(b.q), a.steps, b.steps)


def test_noalloc_rejects_what_does_not_fit_and_bad_masks(gpu_ctx):
    open_cube = np.zeros((256, 256, 256), np.uint8)     # 2.6 GB of FP64 populations
    rc, out = _run(gpu_ctx, "noalloc", open_cube, (1e-5, 0, 0), precision="f64")
    assert rc == 16 and "too large" in gpu_ctx.last_error()
    bad = np.zeros((8, 8, 8), np.uint8)
    bad[1, 2, 3] = 2
    rc, _ = _run(gpu_ctx, "noalloc", bad, (1e-5, 0, 0))
    assert rc == 11
    # and the entry point still works after the rejections
    m = golden("sph16_s0")[0]
    cfg = dict(ff=40, eps=1e-6, lag=200, min_steps=400, avg_steps=200, max_steps=4000)
    rc_a, a = _run(gpu_ctx, "noalloc", m, (1e-5, 0, 0), **cfg)
    rc_b, b = _run(gpu_ctx, "regular", m, (1e-5, 0, 0), **cfg)
    assert rc_a == rc_b and _same(a, b)
