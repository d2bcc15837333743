"""
The convergence rules of solve() pinned by hand-derived stop steps.

The 2D reference cannot be built here (LMB2d needs Magick++), so its stop rule
(LMB2d/main.cpp:99-235: flux sampled every avg_sample steps, block means and
their standard errors compared every lag steps, tol = max(eps * max(|m|,
qfloor * |m|_2), ksigma * sqrt(se^2 + se_prev^2)), hold consecutive passes
from min_steps on, then avg_steps of averaging) is checked here against stop
steps derived by hand from that text, on planted flux sequences chosen so that
a plausible misreading of one clause lands on a different step:

* the first check only records the previous block (no hold yet);
* ``bn < 2`` skips a check *without* resetting the block;
* the k-sigma term, the qfloor term and ``s >= min_steps`` each decide a case;
* a failed check resets hold to 0;
* a non-finite sample fails the run at that step.

Both the CPU oracle (oracle/, float and double builds) and the product's own
state machine (the code the finish kernel runs on the device, replayed on the
host through fp_replay_convergence2d / 3d) must reproduce every derived value.
The 3D rules (LMB3d/lbm3d-perm.cpp:320-398) get the same treatment.

Then the 2D rules on real lattices (CPU oracle; the same cases run on the GPU
in test_gpu_2d_rules.py): the minimum possible step count at the reference
defaults, a direction without a through path (which converges, at that
minimum, to the closed-form flux of the 2D forcing convention), the max_steps
path (exit 3), and field_residual against closed forms.
"""
import ctypes as C

import numpy as np
import pytest

# a short 2D schedule: lag 100, samples every 10 steps, hold 3 from step 200
CFG2 = dict(eps=1e-3, lag=100, hold=3, min_steps=200, max_steps=2000, avg_steps=100,
            avg_sample=10, qfloor=1e-2, ksigma=3.0)
A = 2.0 ** -10                  # a flux magnitude whose sums are exact in binary


def _lib():
    from flowperm import _lib
    return _lib.load()


def replay2d_product(seq, **cfg):
    from flowperm import _lib
    from flowperm.solver import config2d
    q = np.ascontiguousarray(seq, np.float64).reshape(-1, 2)
    c = config2d(**cfg)
    out = _lib.SolveOut()
    rc = _lib.load().fp_replay_convergence2d(C.byref(c), _lib.f64ptr(q), q.shape[0], C.byref(out))
    return rc, {"q": np.array(out.q[:2]), "se": np.array(out.se[:2]), "steps": out.steps,
                "converged": out.converged, "status": out.status, "check": check_of(out)}


def replay3d_product(seq, **cfg):
    from flowperm import _lib
    from flowperm.solver import config3d
    q = np.ascontiguousarray(seq, np.float64).reshape(-1, 3)
    c = config3d(**cfg)
    out = _lib.SolveOut()
    rc = _lib.load().fp_replay_convergence3d(C.byref(c), _lib.f64ptr(q), q.shape[0], C.byref(out))
    return rc, {"q": np.array(out.q[:]), "se": np.array(out.se[:]), "steps": out.steps,
                "converged": out.converged, "status": out.status, "check": check_of(out)}


def check_of(out):
    return {"step": out.check_step, "checks": out.checks, "q": np.array(out.check_q[:]),
            "ratio": np.array(out.check_ratio[:]), "hold": out.check_hold}


def all_replays2d(oracle, seq, cfg):
    """(product status, product result, [oracle float, oracle double])"""
    rc, p = replay2d_product(seq, **cfg)
    o = [oracle.replay2d(seq, double=d, **cfg) for d in (False, True)]
    return rc, p, o


def seq2(n, fx, fy=lambda s: 0.0):
    """flux after step s = (fx(s), fy(s)), s = 1..n"""
    return np.array([[fx(s), fy(s)] for s in range(1, n + 1)], np.float64)


def alternating(s, d):
    """+d on odd samples, -d on even ones (samples every 10 steps)"""
    return d if (s // 10) % 2 else -d


# name, config changes, flux sequence, (steps, converged) derived from main.cpp:99-235
CASES2 = [
    # constant flux: checks at 100 (records), 200, 300, 400 (hold 1, 2, 3) -> 400 + 100
    ("constant", {}, seq2(3000, lambda s: A), (500, 1)),
    # min_steps 450: the checks at 200-400 reset hold; 500, 600, 700 count -> 700 + 100
    ("min_steps", {"min_steps": 450}, seq2(3000, lambda s: A), (800, 1)),
    # lag 100 = avg_sample: one sample per lag, so a check has bn = 1 and is skipped
    # without resetting the block; every second lag sees bn = 2: records at 200,
    # holds at 400, 600, 800 -> 800 + 100
    ("bn_lt_2", {"avg_sample": 100}, seq2(3000, lambda s: A), (900, 1)),
    # one block (301-400) offset by 10 eps |m|: the checks at 400 and 500 fail and
    # reset hold (2 -> 0); 600, 700, 800 count -> 800 + 100
    ("hold_reset", {"ksigma": 0.0},
     seq2(3000, lambda s: A * (1 + 1e-2) if 300 < s <= 400 else A), (900, 1)),
    # block means drift by 10 eps |m| per block while the samples alternate by
    # +-0.1 |m|: se = 0.1 |m| / 3, so only ksigma * sqrt(se^2 + se_prev^2)
    # (0.14 |m|) covers the drift -> stops like "constant"
    ("ksigma_covers_drift", {},
     seq2(3000, lambda s: A * (1 + 1e-2 * ((s - 1) // 100)) + A * alternating(s, 0.1)), (500, 1)),
    # the same with ksigma 0: the drift fails every check -> max_steps + avg_steps
    ("ksigma_off", {"ksigma": 0.0},
     seq2(3000, lambda s: A * (1 + 1e-2 * ((s - 1) // 100)) + A * alternating(s, 0.1)), (2100, 0)),
    # q_y drifts by 1e-6 |q_x| per block: eps * |m_y| is far below it, but
    # eps * qfloor * |m| (1e-5 |q_x|) covers it -> stops like "constant"
    ("qfloor_covers_small_component", {},
     seq2(3000, lambda s: A, lambda s: A * 1e-6 * ((s - 1) // 100 + 1)), (500, 1)),
    ("qfloor_off", {"qfloor": 0.0},
     seq2(3000, lambda s: A, lambda s: A * 1e-6 * ((s - 1) // 100 + 1)), (2100, 0)),
    # a component that is exactly zero converges only through the qfloor term
    ("zero_component_needs_qfloor", {"qfloor": 0.0}, seq2(3000, lambda s: A), (2100, 0)),
]


@pytest.mark.parametrize("name,change,seq,want", CASES2, ids=[c[0] for c in CASES2])
def test_2d_rules_stop_where_derived(oracle, name, change, seq, want):
    cfg = dict(CFG2, **change)
    rc, p, orc = all_replays2d(oracle, seq, cfg)
    steps, conv = want
    for r in [p] + orc:
        assert (r["steps"], r["converged"], r["status"]) == (steps, conv, 0), (name, r)
    assert rc == (0 if conv else 3)
    for o in orc:                               # same statistics, bit for bit
        assert np.array_equal(o["q"], p["q"]) and np.array_equal(o["se"], p["se"])


def test_2d_rules_average_and_standard_error_closed_form(oracle):
    """Phase 2 samples alternating A(1 +- 1/4): mean A exactly, and
    se = sqrt(var / n) with var = n (A/4)^2 / (n - 1), n = avg_steps / avg_sample."""
    seq = seq2(3000, lambda s: A if s <= 400 else A * (1 + alternating(s, 0.25)),
               lambda s: -A)
    rc, p, orc = all_replays2d(oracle, seq, CFG2)
    n = CFG2["avg_steps"] // CFG2["avg_sample"]
    se = np.sqrt(n * (A / 4) ** 2 / (n - 1) / n)
    for r in [p] + orc:
        assert r["steps"] == 500 and r["converged"] == 1
        assert r["q"][0] == A and r["q"][1] == -A
        assert abs(r["se"][0] - se) <= 1e-15 * se and r["se"][1] == 0.0


def test_2d_rules_short_averaging_reads_the_final_flux(oracle):
    """avg_steps < avg_sample: no phase-2 sample, q is volumeavg after the
    last step (main.cpp:203-207) -- here the planted value of step 505."""
    cfg = dict(CFG2, avg_steps=5)
    seq = seq2(3000, lambda s: A if s <= 400 else A * (1 + s * 2.0 ** -20))
    rc, p, orc = all_replays2d(oracle, seq, cfg)
    for r in [p] + orc:
        assert r["steps"] == 405
        assert r["q"][0] == A * (1 + 405 * 2.0 ** -20) and r["se"][0] == 0.0


def test_2d_rules_nonfinite_sample_fails_at_its_step(oracle):
    seq = seq2(3000, lambda s: float("nan") if s == 250 else A)
    rc, p, orc = all_replays2d(oracle, seq, CFG2)
    assert rc == 13 and p["status"] == 13        # FP_ERR_NONFINITE: exit code 2
    for r in [p] + orc:
        assert r["steps"] == 250
    for o in orc:
        assert o["status"] == 2


def test_2d_rules_without_first_phase():
    """max_steps 0: phase 1 runs no step, converged = (0 >= hold) = 0, then avg."""
    rc, p = replay2d_product(seq2(300, lambda s: A), **dict(CFG2, max_steps=0))
    assert (p["steps"], p["converged"], rc) == (100, 0, 3)


CFG3 = dict(eps=1e-6, lag=100, hold=2, min_steps=200, max_steps=2000, avg_steps=100)


def seq3(n, f):
    return np.array([f(s) for s in range(1, n + 1)], np.float64)


# LMB3d/lbm3d-perm.cpp:320-398: drift = max_c |q_c - qprev_c| < eps |q|, checks
# every lag from min_steps on, hold consecutive passes, then avg_steps with a flux
# every step.
CASES3 = [
    # 100 records, 200 and 300 count -> 300 + 100 (the reference's 8000 at lag
    # 2000, min 4000, hold 2, avg 2000 is the same rule)
    ("constant", {}, lambda s: [A, -A / 2, 0.0], (400, 1)),
    ("min_steps", {"min_steps": 350}, lambda s: [A, -A / 2, 0.0], (600, 1)),
    ("hold_reset", {}, lambda s: [A * (1 + 1e-5) if 200 < s <= 300 else A, 0.0, 0.0], (700, 1)),
    # drift 2e-6 |q| per check: above eps 1e-6 -> never
    ("drifting", {}, lambda s: [A * (1 + 2e-6 * (s // 100)), 0.0, 0.0], (2100, 0)),
    # the norm covers all three components: drift 0.5e-6 A in z against |q| ~ A
    ("norm_of_all_components", {}, lambda s: [A, 0.0, A * 0.5e-6 * (s // 100)], (400, 1)),
]


@pytest.mark.parametrize("name,change,f,want", CASES3, ids=[c[0] for c in CASES3])
def test_3d_rules_stop_where_derived(oracle, name, change, f, want):
    cfg = dict(CFG3, **change)
    seq = seq3(3000, f)
    rc, p = replay3d_product(seq, **cfg)
    o = oracle.replay3d(seq, **cfg)
    for r in (p, o):
        assert (r["steps"], r["converged"], r["status"]) == (want[0], want[1], 0), (name, r)
    assert np.array_equal(p["q"], o["q"]) and np.array_equal(p["se"], o["se"])


def test_3d_rules_mean_of_the_averaging_phase(oracle):
    seq = seq3(3000, lambda s: [A if s <= 300 else A * (1 + alternating(s * 10, 0.25)), 0.0, 0.0])
    rc, p = replay3d_product(seq, **CFG3)
    o = oracle.replay3d(seq, **CFG3)
    n = CFG3["avg_steps"]
    se = np.sqrt(n * (A / 4) ** 2 / (n - 1) / n)
    for r in (p, o):
        assert r["steps"] == 400 and r["q"][0] == A
        assert abs(r["se"][0] - se) <= 1e-15 * se


def test_last_check_values_3d():
    """The values of the reference's progress line at the last check
    (LMB3d/lbm3d-perm.cpp:353-357: drift/tol = max_c |q_c - qprev_c| /
    (eps |q|)), derived by hand.  A flux that grows by 2e-6 A per check never
    converges: the last check is at max_steps 2000, the 20th, with q = A (1 +
    4e-5), drift 2e-6 A against eps |q| = 1e-6 A (1 + 4e-5), so drift/tol =
    2 / 1.00004, hold 0.  A constant flux stops at its third check (300) with
    drift 0 and hold 2."""
    rc, p = replay3d_product(seq3(3000, lambda s: [A * (1 + 2e-6 * (s // 100)), 0.0, 0.0]), **CFG3)
    c = p["check"]
    assert (c["step"], c["checks"], c["hold"]) == (2000, 20, 0)
    assert c["q"][0] == A * (1 + 2e-6 * 20) and c["q"][1] == c["q"][2] == 0.0
    assert abs(c["ratio"][0] - 2 / 1.00004) <= 1e-9 and c["ratio"][1] == 0.0
    rc, p = replay3d_product(seq3(3000, lambda s: [A, -A / 2, 0.0]), **CFG3)
    c = p["check"]
    assert (c["step"], c["checks"], c["hold"]) == (300, 3, 2)
    assert list(c["q"]) == [A, -A / 2, 0.0] and list(c["ratio"]) == [0.0, 0.0]


def test_last_check_values_2d():
    """LMB2d/main.cpp:167-172 at the last check, by hand.  Block means that
    grow by 1e-2 A per 100-step block (samples alternating +-0.1 A about them)
    with ksigma 0 never converge: the last check is at max_steps 2000, the
    20th, block mean qx = 1.19 A against 1.18 A before, tol = eps |qx| =
    1.19e-3 A (qfloor 1e-2 |m| is smaller), dx/tol = 0.01 / 1.19e-3; qy = 0
    gives dy = 0 and dy/tol 0.  Before any previous block (the first check)
    both ratios are 0."""
    seq = seq2(3000, lambda s: A * (1 + 1e-2 * ((s - 1) // 100)) + A * alternating(s, 0.1))
    rc, p = replay2d_product(seq, **dict(CFG2, ksigma=0.0))
    c = p["check"]
    assert (c["step"], c["checks"], c["hold"]) == (2000, 20, 0)
    assert abs(c["q"][0] - 1.19 * A) <= 1e-14 * A and c["q"][1] == 0.0
    assert abs(c["ratio"][0] - 0.01 / 1.19e-3) <= 1e-9 * 8.4 and c["ratio"][1] == 0.0
    rc, p = replay2d_product(seq2(250, lambda s: A), **dict(CFG2, max_steps=150, avg_steps=0))
    c = p["check"]
    assert (c["step"], c["checks"], c["hold"]) == (100, 1, 0) and list(c["ratio"]) == [0.0, 0.0]


def test_replay_rejects_a_short_sequence():
    rc, _ = replay2d_product(seq2(150, lambda s: A), **CFG2)
    assert rc == 16                              # FP_ERR_ARG


# ---------------------------------------------------------------------------
# the 2D rules on real lattices (CPU oracle)
# ---------------------------------------------------------------------------

def min_steps_at(cfg):
    """The earliest stop the rules allow: the first check that can count is
    the first lag multiple >= min_steps that has a previous block; hold
    consecutive passes; then the averaging phase."""
    first = max(cfg["min_steps"], 2 * cfg["lag"])
    first = -(-first // cfg["lag"]) * cfg["lag"]
    return first + (cfg["hold"] - 1) * cfg["lag"] + cfg["avg_steps"]


DEFAULTS2 = dict(eps=1e-3, lag=5000, hold=3, min_steps=10000, max_steps=500000, avg_steps=20000,
                 avg_sample=50, qfloor=1e-2, ksigma=3.0)


def slit2d(L=32, h=6):
    m = np.zeros((L, L), np.uint8)
    m[:L - h, :] = 1
    return m


@pytest.mark.parametrize("double", [False, True])
def test_oracle2d_slit_stops_at_the_minimum_at_reference_defaults(oracle, double):
    """A slit settles within a few hundred steps, so at the defaults it stops
    at the earliest step the rules allow: 10000 + 2 * 5000 + 20000 = 40000."""
    assert min_steps_at(DEFAULTS2) == 40000
    rc, T = oracle.solve_tensor2d(slit2d(), double=double)
    assert rc == 0
    for d in range(2):
        assert T.run[d].steps == 40000 and T.run[d].converged == 1
        assert T.run[d].fieldres == 0.0            # the steady field is a fixed point


def test_oracle2d_direction_without_a_through_path(oracle):
    """Fluid columns along y, forced along x: no through path.  Under the 2D
    convention (shifted equilibrium velocity, raw first moment reported) the
    fluid comes to rest and every fluid node reads u = -F / (2 rho), rho = 1,
    so q_x = -F n_fluid / (2 LX L0^2); the run converges, at the minimum."""
    m = np.ones((32, 32), np.uint8)
    m[:, 10:16] = 0
    rc, T = oracle.solve_tensor2d(m, double=True)
    F = 2.5e-7
    want = -F * int((m == 0).sum()) / (2 * 32 * 4.0 ** 2)
    assert T.run[0].steps == 40000 and T.run[0].converged == 1
    assert abs(T.run[0].qx - want) <= 1e-9 * abs(want)
    assert abs(T.run[0].qy) <= 1e-10 * abs(want)


@pytest.mark.parametrize("double", [False, True])
def test_oracle2d_max_steps_below_the_minimum_is_exit_3(oracle, double):
    cfg = dict(max_steps=15000)
    rc, T = oracle.solve_tensor2d(slit2d(), double=double, **cfg)
    assert rc == 3
    for d in range(2):
        assert T.run[d].steps == 15000 + 20000 and T.run[d].converged == 0


@pytest.mark.parametrize("avg,want", [(100, 1.0), (0, 10.0)])
def test_oracle2d_no_first_phase_snapshots_the_rest_field(oracle, avg, want):
    """max_steps 0: the snapshot is the rest field U = V = 0
    (resetdistributions, LMB2d/lbm.cpp:200-211), so fieldres is
    sqrt(sum u^2 / sum u^2) = 1 exactly, or 10 when no step runs at all."""
    r = oracle.solve_dir2d(slit2d(), (2.5e-7, 0.0), max_steps=0, avg_steps=avg, avg_sample=25)
    assert r["fieldres"] == want
    assert r["steps"] == avg and r["converged"] == 0


def test_field_residual_closed_forms(oracle):
    """fp_field_residual2d and the oracle's field_residual: cur = 2 snap gives
    sqrt(1/4) = 0.5 exactly; cur = snap gives 0; a zero field gives 10; a
    random pair equals the double-precision formula to rounding."""
    from flowperm import _lib
    lib = _lib.load()
    rng = np.random.default_rng(2)
    snap = rng.normal(0, 1e-4, (9, 7, 2)).astype(np.float32)

    def product(cur, sn):
        c = np.ascontiguousarray(cur, np.float32)
        s = np.ascontiguousarray(sn, np.float32)
        return lib.fp_field_residual2d(7, 9, _lib.f32ptr(c), _lib.f32ptr(s))

    for fn in (product, oracle.field_residual2d,
               lambda c, s: oracle.field_residual2d(c, s, double=True)):
        assert fn(2 * snap, snap) == 0.5
        assert fn(snap, snap) == 0.0
        assert fn(np.zeros_like(snap), snap) == 10.0
        cur = rng.normal(0, 1e-4, snap.shape).astype(np.float32)
        want = np.sqrt(((cur.astype(np.float64) - snap) ** 2).sum() / (cur.astype(np.float64) ** 2).sum())
        assert abs(fn(cur, snap) - want) <= 2 ** -23 * want


def test_oracle2d_tensor_test_is_linear_in_the_force(oracle):
    """--validate DEG (LMB2d/main.cpp:360-380): the solve at angle DEG against
    K F.  Its residual is the lattice's O(Re) non-linearity: small, and
    proportional to F (the 3D check is test_oracle_tensor_test_and_reciprocity)."""
    from flowperm import masks
    m = masks.ellipses((24, 24), 0.7, rmin=1.5, rmax=2.5, seed=4, axis=(1, 2))
    base = dict(double=True, eps=1e-10, ksigma=0.0, lag=500, min_steps=1000, max_steps=60000,
                avg_steps=50, avg_sample=1)
    errs = {}
    for ff in (1.0, 4.0):
        rc, T = oracle.solve_tensor2d(m, ff=ff, validate=30.0, **base)
        assert T.validated == 1
        th = np.deg2rad(30.0)
        F = T.force
        K = np.array([[T.k_xx, T.k_off], [T.k_off, T.k_yy]])
        pred = K @ np.array([F * np.cos(th), F * np.sin(th)]) / T.mu
        q = np.array([T.val.qx, T.val.qy])
        assert np.allclose(pred, T.val_pred[:], rtol=1e-14, atol=0)
        errs[ff] = np.linalg.norm(q - pred) / np.linalg.norm(q)
        assert abs(errs[ff] - T.val_err) <= 1e-12
    assert errs[1.0] < 1e-6
    assert 3.9 < errs[4.0] / errs[1.0] < 4.1
