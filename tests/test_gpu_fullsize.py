"""
The hot path at the workload sizes: the 256^3 D3Q19 lattice of the throughput
benchmark (3:1 ellipsoids, seed 0, porosity 0.618) and a 1024^2 D2Q9 lattice.
A full solve there takes the CPU oracle hours, so these check what holds at any
size:

* a few FP64 (3D) and reference-float / FP64 (2D) steps bit for bit against the
  oracle, flux and exported velocity field;
* FP32 rows of a structure solved alone equal to the same structure solved inside
  a batch with another one (the flux sums are exact, so order-independent), at a
  fixed step count, and within 1e-4 of FP64 at the same steps;
* the transpose covariance K' = P K P^T (scripts/rotation_covariance3d.py) at a
  fixed step count.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# a fixed number of steps: the hold test needs min_steps, which is never reached
FIXED3 = dict(ff=40, lag=100, min_steps=10 ** 6, max_steps=400, avg_steps=100)


@pytest.fixture(scope="module")
def lattice256():
    from flowperm import masks
    rng = np.random.default_rng(0)
    por = 0.3 + 0.5 * rng.random()
    return masks.ellipsoids((256,) * 3, por, rmin=2.0, rmax=4.0, aspect=3.0, jitter_deg=15.0,
                            seed=0)


STEPS6 = dict(ff=40, lag=2, min_steps=10 ** 6, max_steps=6, avg_steps=0)
F6 = (1e-5, 0.0, 0.0)


@pytest.fixture(scope="module")
def oracle256(oracle, lattice256):
    return oracle.solve_dir3d(lattice256, F6, velocity=True, **STEPS6)


# both neighbour-step forms: at this size (>= 400 MB of populations) their work
# items are dealt to the XCDs in y bands
@pytest.mark.parametrize("mode", [0, 1], ids=["row_segments", "compact"])
def test_256cube_fp64_steps_bit_exact(oracle256, lattice256, mode):
    import ctypes as C
    from flowperm import _lib, solver
    ctx = solver.Context()
    try:
        t = _lib.Tuning()
        ctx.lib.fp_get_tuning(ctx.handle, C.byref(t))
        t.nbr_mode = mode
        assert ctx.lib.fp_set_tuning(ctx.handle, C.byref(t)) == 0
        print(f"256^3 FP64 6 steps, form {mode}, porosity {1 - lattice256.mean():.4f}", flush=True)
        g = solver.solve_direction3d(lattice256, F6, ctx=ctx, velocity=True, precision="f64", **STEPS6)
    finally:
        ctx.close()
    o = oracle256
    assert g["steps"] == o["steps"] == 6
    assert np.array_equal(g["q"], o["q"])
    assert np.array_equal(g["velocity"], o["velocity"])
    assert np.all(g["velocity"][lattice256 == 1] == 0)


def test_4gib_fp64_wide_addressing_steps_bit_exact(gpu_ctx, oracle):
    """A lattice whose FP64 populations span more than 4 GiB: the neighbour
    kernels switch to 64-bit addresses by themselves (28.9 M fluid cells,
    3% isolated solid voxels, row-segment form, y bands)."""
    from flowperm import solver
    rng = np.random.default_rng(11)
    m = (rng.random((300, 310, 320)) < 0.03).astype(np.uint8)
    n_fluid = int((m == 0).sum())
    stride = (n_fluid + 1023) // 1024 * 1024
    assert 19 * 8 * stride > 2 ** 32
    cfg = dict(ff=40, lag=2, min_steps=10 ** 6, max_steps=4, avg_steps=0)
    print(f"300x310x320 FP64 4 steps, {n_fluid} fluid cells, {19 * 8 * stride / 2 ** 30:.2f} GiB", flush=True)
    g = solver.solve_direction3d(m, F6, ctx=gpu_ctx, precision="f64", **cfg)
    o = oracle.solve_dir3d(m, F6, **cfg)
    assert g["steps"] == o["steps"] == 4
    assert np.array_equal(g["q"], o["q"])


def test_256cube_fp32_alone_equals_batched_and_fp64(gpu_ctx, lattice256):
    from flowperm import masks, solver
    from flowperm.server import Server
    alone = solver.solve_permeability3d(lattice256, ctx=gpu_ctx, precision="f32", **FIXED3)
    other = masks.spheres((64, 64, 64), 0.6, seed=1)
    with Server(3, "f32", ctx=gpu_ctx) as srv:
        a = srv.submit(other, **FIXED3)
        b = srv.submit(lattice256, **FIXED3)
        while srv.step():
            pass
        srv.poll(a)
        batched = srv.poll(b)[0]
    keys = [k for k in alone if k.startswith(("q", "se_", "K_", "k", "steps_", "conv_"))]
    assert keys and all(alone[k] == batched[k] for k in keys), \
        [k for k in keys if alone[k] != batched[k]]
    assert alone["steps_fx"] == 500
    ref = solver.solve_permeability3d(lattice256, ctx=gpu_ctx, precision="f64", **FIXED3)
    K32 = np.array([[alone[f"K_{r}{c}"] for c in "xyz"] for r in "xyz"])
    K64 = np.array([[ref[f"K_{r}{c}"] for c in "xyz"] for r in "xyz"])
    err = np.abs(K32 - K64) / np.maximum(np.abs(K64), abs(np.trace(K64)))
    print(f"256^3 FP32 vs FP64 at 500 steps: max error {err.max():.2e}", flush=True)
    assert err.max() <= 1e-4


def test_256cube_transpose_covariance(gpu_ctx, lattice256):
    from flowperm import covariance as cv
    res = cv.solve_orientations(lattice256, ["yxz"], precision="f32", ctx=gpu_ctx, **FIXED3)
    ok, lines = cv.compare(res, ["yxz"], tol=1e-6, min_scale=1e-3)
    print(lines, flush=True)
    assert ok, lines
    assert lines[0]["steps"] == lines[0]["steps_expected"]


@pytest.mark.parametrize("prec,double", [("f32ref", False), ("f64", True)])
def test_1024sq_d2q9_bit_exact(gpu_ctx, oracle, prec, double):
    from flowperm import masks, solver
    m = masks.ellipses((1024, 1024), 0.6, seed=3)
    cfg = dict(ff=1.0, lag=50, min_steps=10 ** 6, max_steps=100, avg_steps=100, avg_sample=25,
               L0=16.0)
    print(f"1024^2 {prec}, porosity {1 - m.mean():.4f}", flush=True)
    rc, T = oracle.solve_tensor2d(m, double=double, **cfg)
    r = solver.solve_permeability2d(m, ctx=gpu_ctx, precision=prec, velocity=True, **cfg)
    assert r["exit_code"] == rc
    for d, t in enumerate(("fx", "fy")):
        assert r[f"steps_{t}"] == T.run[d].steps
        assert r[f"qx_{t}"] == T.run[d].qx and r[f"qy_{t}"] == T.run[d].qy
        assert r[f"se_qx_{t}"] == T.run[d].se_qx and r[f"se_qy_{t}"] == T.run[d].se_qy
        assert r[f"fieldres_{t}"] == T.run[d].fieldres
    o = oracle.solve_dir2d(m, (2.5e-7, 0.0), velocity=True, double=double, **cfg)
    assert np.array_equal(r["velocity"][0], o["velocity"])
