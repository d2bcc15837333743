"""
The 2D convergence rules on the GPU, on lattices whose outcome is fixed by the
rule text (LMB2d/main.cpp:99-235) or by a closed form, each also bit for bit
against the CPU oracle:

* a slit at the reference defaults stops at the earliest step the rules allow,
  10000 + 2 * 5000 + 20000 = 40000, with fieldres 0 (the field is a fixed point);
* a direction without a through path converges there too, to
  q = -F n_fluid / (2 LX L0^2) (the fluid at rest reads u = -F/2 under the 2D
  forcing convention);
* max_steps below that minimum ends with conv 0 and exit 3 at max_steps + avg_steps;
* max_steps 0: the snapshot is the rest field, so fieldres is 1 exactly, or 10
  when no step runs at all.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


# the flux normalisation length of the reference's 256^2 lattice (lbm.h:24-27);
# the Python surface would scale it with the lattice, the oracle takes it as given
L0 = 4.0


def slit2d(L=32, h=6):
    m = np.zeros((L, L), np.uint8)
    m[:L - h, :] = 1
    return m


def _check_against_oracle(r, T):
    for d, t in enumerate(("fx", "fy")):
        assert r[f"steps_{t}"] == T.run[d].steps and r[f"conv_{t}"] == T.run[d].converged
        assert r[f"qx_{t}"] == T.run[d].qx and r[f"qy_{t}"] == T.run[d].qy
        assert r[f"se_qx_{t}"] == T.run[d].se_qx and r[f"se_qy_{t}"] == T.run[d].se_qy
        assert r[f"fieldres_{t}"] == T.run[d].fieldres


@pytest.mark.parametrize("prec,double", [("f32ref", False), ("f64", True)])
def test_slit_stops_at_the_minimum_at_reference_defaults(gpu_ctx, oracle, prec, double):
    from flowperm import solver
    m = slit2d()
    print(f"2D slit at the reference defaults, {prec}", flush=True)
    r = solver.solve_permeability2d(m, ctx=gpu_ctx, precision=prec, L0=L0)
    rc, T = oracle.solve_tensor2d(m, double=double, L0=L0)
    assert r["exit_code"] == rc == 0
    for t in ("fx", "fy"):
        assert r[f"steps_{t}"] == 40000 and r[f"conv_{t}"] == 1 and r[f"fieldres_{t}"] == 0.0
    _check_against_oracle(r, T)


def test_direction_without_a_through_path_converges_to_the_closed_form(gpu_ctx, oracle):
    from flowperm import solver
    m = np.ones((32, 32), np.uint8)
    m[:, 10:16] = 0
    r = solver.solve_permeability2d(m, ctx=gpu_ctx, precision="f64", L0=L0)
    rc, T = oracle.solve_tensor2d(m, double=True, L0=L0)
    want = -2.5e-7 * int((m == 0).sum()) / (2 * 32 * L0 ** 2)
    assert r["steps_fx"] == 40000 and r["conv_fx"] == 1
    assert abs(r["qx_fx"] - want) <= 1e-9 * abs(want)
    _check_against_oracle(r, T)


@pytest.mark.parametrize("prec,double", [("f32ref", False), ("f64", True)])
def test_max_steps_below_the_minimum_is_exit_3(gpu_ctx, oracle, prec, double):
    from flowperm import solver
    m = slit2d()
    r = solver.solve_permeability2d(m, ctx=gpu_ctx, precision=prec, max_steps=15000, L0=L0)
    rc, T = oracle.solve_tensor2d(m, double=double, max_steps=15000, L0=L0)
    assert r["exit_code"] == rc == 3
    for t in ("fx", "fy"):
        assert r[f"steps_{t}"] == 35000 and r[f"conv_{t}"] == 0
    _check_against_oracle(r, T)


@pytest.mark.parametrize("prec", ["f32ref", "f64", "f32"])
@pytest.mark.parametrize("avg,want", [(100, 1.0), (0, 10.0)])
def test_no_first_phase_snapshots_the_rest_field(gpu_ctx, oracle, prec, avg, want):
    """fieldres from the rest-state snapshot (resetdistributions zeroes U, V,
    LMB2d/lbm.cpp:200-211): the snapshot slot must hold the rest field when no
    first-phase step writes it."""
    from flowperm import solver
    m = slit2d()
    cfg = dict(max_steps=0, avg_steps=avg, avg_sample=25, L0=L0)
    r = solver.solve_permeability2d(m, ctx=gpu_ctx, precision=prec, **cfg)
    for t in ("fx", "fy"):
        assert r[f"fieldres_{t}"] == want and r[f"steps_{t}"] == avg and r[f"conv_{t}"] == 0
    if prec != "f32":
        rc, T = oracle.solve_tensor2d(m, double=prec == "f64", **cfg)
        _check_against_oracle(r, T)
