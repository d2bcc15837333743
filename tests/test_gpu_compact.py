"""
The compact neighbour step (fp_tuning.nbr_mode = 1: 64 fluid cells per
wavefront, per-cell neighbour table) against the row-segment step and the
oracle.  Both forms read and write exactly the same population slots with the
same arithmetic, so every result is bit-identical between them.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import golden, golden_flags

pytestmark = pytest.mark.gpu

KEYS = [f"K_{a}{b}" for a in "xyz" for b in "xyz"] + [f"steps_{t}" for t in ("fx", "fy", "fz")] + \
       [f"{p}q{a}_{t}" for p in ("", "se_") for a in "xyz" for t in ("fx", "fy", "fz")]


@pytest.fixture(scope="module")
def compact_ctx():
    from flowperm import _lib, solver
    ctx = solver.Context()
    t = _lib.Tuning()
    ctx.lib.fp_get_tuning(ctx.handle, C.byref(t))
    t.nbr_mode = 1
    t.nbr_bpc = 8
    assert ctx.lib.fp_set_tuning(ctx.handle, C.byref(t)) == 0
    yield ctx
    ctx.close()


@pytest.mark.parametrize("name", ["sph16_s0", "ell_x24y20z16_fast_s5", "sph_x28y24z20_s1"])
def test_compact_fp64_bit_exact_against_oracle(compact_ctx, oracle, name):
    from flowperm import solver
    mask, _, meta = golden(name)
    cfg = golden_flags(meta)
    print(f"case {name} {mask.shape}", flush=True)
    rc, T = oracle.solve_tensor3d(mask, **cfg)
    r = solver.solve_permeability3d(mask, ctx=compact_ctx, **cfg)
    assert r["exit_code"] == rc
    K = np.array(T.K[:]).reshape(3, 3)
    for d, t in enumerate(("fx", "fy", "fz")):
        assert r[f"steps_{t}"] == T.run[d].steps
        for k, a in enumerate("xyz"):
            assert r[f"q{a}_{t}"] == T.run[d].q[k]
    for i, a in enumerate("xyz"):
        for j, b in enumerate("xyz"):
            assert r[f"K_{a}{b}"] == K[i, j]


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(12, 10, 65), (9, 14, 30), (8, 6, 130), (5, 7, 64)])
def test_compact_equals_row_segment_bitwise(gpu_ctx, compact_ctx, precision, shape):
    from flowperm import masks, solver
    m = masks.spheres(shape, 0.55, 2, 3, seed=sum(shape))
    cfg = dict(ff=40, eps=1e-6, lag=200, min_steps=400, avg_steps=200, max_steps=3000,
               precision=precision)
    print(f"shape {shape} {precision}", flush=True)
    a = solver.solve_permeability3d(m, ctx=gpu_ctx, **cfg)
    b = solver.solve_permeability3d(m, ctx=compact_ctx, **cfg)
    for k in KEYS:
        assert a[k] == b[k], k


def test_compact_velocity_export_equal(gpu_ctx, compact_ctx):
    from flowperm import solver
    mask, _, meta = golden("sph12_velocity_s4")
    cfg = golden_flags(meta)
    a = solver.solve_permeability3d(mask, ctx=gpu_ctx, velocity=True, **cfg)
    b = solver.solve_permeability3d(mask, ctx=compact_ctx, velocity=True, **cfg)
    assert np.array_equal(a["velocity"], b["velocity"])


@pytest.mark.parametrize("precision", ["f32ref", "f64"])
def test_compact_2d_equals_row_segment(gpu_ctx, compact_ctx, precision):
    from flowperm import masks, solver
    m = masks.ellipses((70, 90), 0.6, seed=4)
    cfg = dict(lag=500, min_steps=1000, max_steps=6000, avg_steps=1000, precision=precision)
    a = solver.solve_permeability2d(m, ctx=gpu_ctx, **cfg)
    b = solver.solve_permeability2d(m, ctx=compact_ctx, **cfg)
    for k in ("k_xx", "k_xy", "k_yx", "k_yy", "steps_fx", "steps_fy", "fieldres_fx", "fieldres_fy"):
        assert a[k] == b[k], k


def test_compact_batch_mix_equals_alone(compact_ctx):
    from flowperm import masks
    from flowperm.server import Server
    cfg = dict(ff=40, eps=1e-6, lag=200, min_steps=400, avg_steps=200, max_steps=4000)
    ms = [masks.spheres((16, 18, 20), 0.6, 3, 5, seed=1), masks.ellipsoids((24, 16, 70), 0.5, seed=2)]
    alone = []
    for m in ms:
        with Server(3, "f32", ctx=compact_ctx) as s:
            rid = s.submit(m, **cfg)
            alone.append(s.run()[rid][0])
    with Server(3, "f32", ctx=compact_ctx) as s:
        ids = [s.submit(m, **cfg) for m in ms]
        res = s.run(ids)
    for i, rid in enumerate(ids):
        for k in KEYS:
            assert res[rid][0][k] == alone[i][k], k


def test_auto_form_mixed_batch_equals_row_segment(gpu_ctx):
    """nbr_mode 2 (the default): a dense and a sparse structure in one batch
    take different forms (both kernels run each even step) and every result
    equals the row-segment form's bit for bit"""
    import ctypes as C
    from flowperm import _lib, masks, solver
    from flowperm.server import Server
    dense = masks.spheres((20, 20, 24), 0.8, 3, 5, seed=5)
    sparse = masks.grf((24, 28, 70), porosity=0.3, corr=(3, 3, 6), seed=6)
    print("porosities", 1 - dense.mean(), 1 - sparse.mean(), flush=True)
    # on either side of COMPACT_BELOW_POROSITY (0.68, csrc/lbm_common.h)
    assert 1 - sparse.mean() < 0.68 < 1 - dense.mean()
    cfg = dict(ff=40, eps=1e-6, lag=200, min_steps=400, avg_steps=200, max_steps=2000)
    seg = solver.Context()
    t = _lib.Tuning()
    seg.lib.fp_get_tuning(seg.handle, C.byref(t))
    t.nbr_mode = 0
    assert seg.lib.fp_set_tuning(seg.handle, C.byref(t)) == 0
    try:
        ref = [solver.solve_permeability3d(m, ctx=seg, precision="f32", **cfg) for m in (dense, sparse)]
    finally:
        seg.close()
    with Server(3, "f32", ctx=gpu_ctx) as s:
        ids = [s.submit(m, **cfg) for m in (dense, sparse)]
        res = s.run(ids)
    for i, rid in enumerate(ids):
        for k in KEYS:
            assert res[rid][0][k] == ref[i][k], (i, k)
