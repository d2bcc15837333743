"""
D2Q9 on the GPU against the 2D oracle (the reference's own scheme restated:
LMB2d/lbm.cpp + main.cpp).  The float variant is the reference's arithmetic
and is compared bit for bit with the float oracle; the double variant with the
double oracle; the shifted FP32 variant against FP64 within tolerance.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG = dict(lag=500, min_steps=1000, max_steps=4000, avg_steps=600, avg_sample=50)


def mask24(seed=4, por=0.7):
    from flowperm import masks
    return masks.ellipses((24, 28), por, rmin=1.5, rmax=2.5, seed=seed, axis=(1, 2))


@pytest.mark.parametrize("prec,double", [("f32ref", False), ("f64", True)])
def test_2d_bit_exact_against_oracle(gpu_ctx, oracle, prec, double):
    from flowperm import solver
    m = mask24()
    for ff in (1.0, 40.0):
        print(f"2d {prec} ff={ff} shape {m.shape}", flush=True)
        rc, T = oracle.solve_tensor2d(m, double=double, ff=ff, L0=4.0, **CFG)
        r = solver.solve_permeability2d(m, ctx=gpu_ctx, precision=prec, ff=ff, L0=4.0,
                                        velocity=True, **CFG)
        assert r["exit_code"] == rc
        for d, t in enumerate(("fx", "fy")):
            assert r[f"steps_{t}"] == T.run[d].steps
            assert r[f"conv_{t}"] == T.run[d].converged
            assert r[f"qx_{t}"] == T.run[d].qx and r[f"qy_{t}"] == T.run[d].qy
            assert r[f"se_qx_{t}"] == T.run[d].se_qx and r[f"se_qy_{t}"] == T.run[d].se_qy
            assert r[f"fieldres_{t}"] == T.run[d].fieldres
        for key, okey in (("k_xx", "k_xx"), ("k_xy", "k_xy"), ("k_yx", "k_yx"), ("k_yy", "k_yy"),
                          ("theta_deg", "theta"), ("anisotropy", "aniso"), ("recip_resid", "recip")):
            assert r[key] == getattr(T, okey), key
        F = 2.5e-7 * ff
        o = oracle.solve_dir2d(m, (F, 0.0), velocity=True, double=double, L0=4.0, ff=ff, **CFG)
        assert np.array_equal(r["velocity"][0], o["velocity"])


def test_2d_fp64_slit_closed_forms(gpu_ctx):
    from flowperm import solver
    L, h = 32, 10
    m = np.zeros((L, L), np.uint8)
    m[:L - h, :] = 1
    r = solver.solve_permeability2d(m, ctx=gpu_ctx, precision="f64", eps=1e-9, lag=2000,
                                    min_steps=4000, max_steps=60000, avg_steps=2000, L0=4.0)
    mu = float(np.float32(0.165))
    assert abs(r["k_xx"] - mu * h ** 3 / 32.0) / (mu * h ** 3 / 32.0) < 1e-8
    assert abs(r["k_yy"] + mu * h / 32.0) / (mu * h / 32.0) < 1e-8


def test_2d_fp32_shifted_close_to_fp64(gpu_ctx):
    from flowperm import solver
    m = mask24(seed=6)
    cfg = dict(ff=40.0, eps=1e-6, lag=1000, min_steps=2000, max_steps=30000, avg_steps=1000,
               L0=4.0)
    a = solver.solve_permeability2d(m, ctx=gpu_ctx, precision="f64", **cfg)
    b = solver.solve_permeability2d(m, ctx=gpu_ctx, precision="f32", **cfg)
    tr = abs(a["k_xx"] + a["k_yy"])
    for k in ("k_xx", "k_xy", "k_yx", "k_yy"):
        err = abs(b[k] - a[k]) / max(abs(a[k]), tr)
        print(k, err, flush=True)
        assert err < 1e-4, k
