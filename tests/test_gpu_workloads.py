"""
The measured workloads at their real shapes, at fixed step counts (a full
solve takes the CPU oracle hours; these check what holds at any size):

* the batched set (3:1 ellipsoids, edges {48 ... 160}, porosity 0.3-0.8): six
  structures x 3 forcing runs in one FP32 batch, mixed neighbour-step forms;
  every row equals the same structure solved alone bit for bit, is within 1e-4
  of FP64 at the same steps, and FP64 steps of its largest lattice equal the
  oracle bit for bit;
* the 512^3 Gaussian random field (porosity 0.15-0.25, its percolating
  cluster, ~10 GB of FP32 populations dense): both neighbour-step forms equal
  bit for bit (y-band item dealing), FP64 steps equal to the oracle, FP32
  within 1e-4 of FP64;
* the serving set (edges 32-256, spheres and GRF): requests submitted while
  others step, each row equal to the request alone, FP32 within 1e-4 of FP64.

These replace the reference's one-process-per-structure pool
(scripts/solve_permeability3d.py:250-259) at the sizes it would run.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# 400 steps of the first phase (min_steps is never reached) + 100 averaged
FIXED = dict(ff=40, lag=100, min_steps=10 ** 6, max_steps=400, avg_steps=100)
STEPS6 = dict(ff=40, lag=2, min_steps=10 ** 6, max_steps=6, avg_steps=0)
F6 = (1e-5, 0.0, 0.0)
KEYS = [f"K_{a}{b}" for a in "xyz" for b in "xyz"] + \
       [f"{p}q{a}_{t}" for p in ("", "se_") for a in "xyz" for t in ("fx", "fy", "fz")] + \
       [f"steps_{t}" for t in ("fx", "fy", "fz")] + [f"conv_{t}" for t in ("fx", "fy", "fz")] + \
       ["k1", "k2", "k3", "fa", "recip_resid", "porosity"]


def kerr(a, b):
    Ka = np.array([[a[f"K_{r}{c}"] for c in "xyz"] for r in "xyz"])
    Kb = np.array([[b[f"K_{r}{c}"] for c in "xyz"] for r in "xyz"])
    return float((np.abs(Ka - Kb) / np.maximum(np.abs(Kb), abs(np.trace(Kb)))).max())


def serve(gpu_ctx, precision, masks, stagger=0):
    """rows of masks solved together by one server; the first `stagger`
    requests step one chunk before the rest are submitted"""
    from flowperm.server import Server
    with Server(3, precision, ctx=gpu_ctx, chunk_steps=100) as s:
        ids = [s.submit(m, **FIXED) for m in masks[:stagger or len(masks)]]
        if stagger:
            s.step()
            ids += [s.submit(m, **FIXED) for m in masks[stagger:]]
        res = s.run(ids)
    return [res[i][0] for i in ids]


def ctx_with_form(mode):
    from flowperm import _lib, solver
    ctx = solver.Context()
    t = _lib.Tuning()
    ctx.lib.fp_get_tuning(ctx.handle, C.byref(t))
    t.nbr_mode = mode
    assert ctx.lib.fp_set_tuning(ctx.handle, C.byref(t)) == 0
    return ctx


# ---------------------------------------------------------------------------
# batched set
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def batched_set():
    from flowperm import workloads
    return [workloads.batched_structure(i) for i in range(6)]


@pytest.mark.timeout(600)
def test_batched_set_in_one_batch_equals_alone_and_fp64(gpu_ctx, batched_set):
    desc = [(m.shape[0], round(1 - float(m.mean()), 3)) for m in batched_set]
    print(f"batched set (edge, porosity): {desc}", flush=True)
    mixed = serve(gpu_ctx, "f32", batched_set)
    ref64 = serve(gpu_ctx, "f64", batched_set)
    worst = 0.0
    for i, m in enumerate(batched_set):
        alone = serve(gpu_ctx, "f32", [m])[0]
        bad = [k for k in KEYS if mixed[i][k] != alone[k]]
        assert not bad, (i, bad)
        assert mixed[i]["steps_fx"] == 500
        worst = max(worst, kerr(mixed[i], ref64[i]))
    print(f"batched set FP32 vs FP64 at 500 steps: max K error {worst:.2e}", flush=True)
    assert worst <= 1e-4


@pytest.mark.timeout(600)
def test_batched_set_largest_lattice_fp64_steps_bit_exact(gpu_ctx, oracle, batched_set):
    from flowperm import solver
    m = max(batched_set, key=lambda a: (a.size, -float(a.mean())))
    print(f"{m.shape} FP64 6 steps against the oracle", flush=True)
    g = solver.solve_direction3d(m, F6, ctx=gpu_ctx, velocity=True, precision="f64", **STEPS6)
    o = oracle.solve_dir3d(m, F6, velocity=True, **STEPS6)
    assert g["steps"] == o["steps"] == 6
    assert np.array_equal(g["q"], o["q"])
    assert np.array_equal(g["velocity"], o["velocity"])


# ---------------------------------------------------------------------------
# 512^3 Gaussian random field
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def large():
    from flowperm import workloads
    m = workloads.large_grf(0)
    assert 0.15 <= 1 - float(m.mean()) <= 0.25
    return m


@pytest.mark.timeout(900)
def test_large_grf_forms_bitwise_equal_and_fp64_steps_equal_the_oracle(oracle, large):
    from flowperm import solver
    print(f"512^3 GRF porosity {1 - float(large.mean()):.4f}", flush=True)
    out = {}
    for mode in (0, 1):
        ctx = ctx_with_form(mode)
        try:
            for prec in ("f32", "f64"):
                print(f"512^3 {prec} 6 steps, form {mode}", flush=True)
                out[mode, prec] = solver.solve_direction3d(large, F6, ctx=ctx, velocity=True,
                                                           precision=prec, **STEPS6)
        finally:
            ctx.close()
    for prec in ("f32", "f64"):
        a, b = out[0, prec], out[1, prec]
        assert a["steps"] == b["steps"] == 6
        assert np.array_equal(a["q"], b["q"]) and np.array_equal(a["velocity"], b["velocity"]), prec
    del out[0, "f32"], out[1, "f32"], out[1, "f64"]
    print("512^3 FP64 6 steps on the CPU oracle", flush=True)
    o = oracle.solve_dir3d(large, F6, velocity=True, **STEPS6)
    g = out[0, "f64"]
    assert np.array_equal(g["q"], o["q"])
    assert np.array_equal(g["velocity"], o["velocity"])


@pytest.mark.timeout(600)
def test_large_grf_fp32_within_1e4_of_fp64(gpu_ctx, large):
    from flowperm import solver
    r32 = solver.solve_permeability3d(large, ctx=gpu_ctx, precision="f32", **FIXED)
    r64 = solver.solve_permeability3d(large, ctx=gpu_ctx, precision="f64", **FIXED)
    err = kerr(r32, r64)
    print(f"512^3 GRF FP32 vs FP64 at 500 steps: max K error {err:.2e}", flush=True)
    assert r32["steps_fx"] == 500 and err <= 1e-4


# ---------------------------------------------------------------------------
# serving set
# ---------------------------------------------------------------------------

@pytest.mark.timeout(600)
def test_serving_set_staggered_equals_alone_and_fp64(gpu_ctx):
    from flowperm import workloads
    ms = [workloads.serving_structure(i) for i in range(8)]
    print(f"serving set (edge, porosity): "
          f"{[(m.shape[0], round(1 - float(m.mean()), 3)) for m in ms]}", flush=True)
    mixed = serve(gpu_ctx, "f32", ms, stagger=3)
    ref64 = serve(gpu_ctx, "f64", ms, stagger=3)
    worst = 0.0
    for i, m in enumerate(ms):
        alone = serve(gpu_ctx, "f32", [m])[0]
        bad = [k for k in KEYS if mixed[i][k] != alone[k]]
        assert not bad, (i, bad)
        worst = max(worst, kerr(mixed[i], ref64[i]))
    print(f"serving set FP32 vs FP64 at 500 steps: max K error {worst:.2e}", flush=True)
    assert worst <= 1e-4
