"""
Batched solving: every job's result is bit-identical whether it runs alone
or among others of different sizes (FP64 and FP32): the flux sums are exact
and each job's lattice evolves independently.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(ctx, prec, items):
    from flowperm import _lib
    from flowperm.solver import config3d
    lib = ctx.lib
    st = C.c_int(0)
    b = lib.fp_batch_create(ctx.handle, 3, prec, C.byref(st))
    assert b, st.value
    ids = []
    for mask, F, cfg in items:
        m = np.ascontiguousarray(mask, np.uint8)
        nz, ny, nx = m.shape
        c = config3d(**cfg)
        Fa = np.asarray(F, np.float64)
        ids.append(lib.fp_batch_add3d(b, _lib.u8ptr(m), nx, ny, nz, C.byref(c), _lib.f64ptr(Fa)))
        assert ids[-1] >= 0
    assert lib.fp_batch_run(b) == 0
    outs = []
    for i in ids:
        o = _lib.SolveOut()
        assert lib.fp_batch_result(b, i, C.byref(o)) == 0
        outs.append((np.array(o.q[:]), np.array(o.se[:]), o.steps, o.converged))
    lib.fp_batch_destroy(b)
    return outs


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_equals_alone(gpu_ctx, prec):
    from conftest import golden
    from flowperm import masks
    F = 2.5e-7 * 40
    cfg = dict(eps=1e-6, lag=200, min_steps=400, max_steps=20000, avg_steps=200)
    items = []
    for name in ("sph16_s0", "sph20_fast_s2", "ell_x24y20z16_fast_s5"):
        m, _, _ = golden(name)
        for d in range(3):
            Fv = [0.0, 0.0, 0.0]
            Fv[d] = F
            items.append((m, Fv, cfg))
    items.append((masks.slit(12, 5), [F, 0, 0], dict(cfg, max_steps=3000)))
    mixed = _batch(gpu_ctx, prec, items)
    for k, it in enumerate(items):
        alone = _batch(gpu_ctx, prec, [it])[0]
        print(f"prec {prec} item {k} steps {alone[2]}", flush=True)
        assert np.array_equal(mixed[k][0], alone[0]) and np.array_equal(mixed[k][1], alone[1])
        assert mixed[k][2] == alone[2] and mixed[k][3] == alone[3]


def test_batch_fp64_equals_single_solver(gpu_ctx):
    from conftest import golden, golden_flags
    from flowperm import solver
    m, _, meta = golden("sph20_fast_s2")
    cfg = golden_flags(meta)
    F = 2.5e-7 * cfg["ff"]
    out = _batch(gpu_ctx, 0, [(m, [0, F, 0], cfg)])[0]
    s = solver.solve_direction3d(m, [0, F, 0], ctx=gpu_ctx, **cfg)
    assert np.array_equal(out[0], s["q"]) and out[2] == s["steps"]
