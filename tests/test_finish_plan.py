"""
The sparse finish schedule, checked on the host (fp_verify_finish_plan, no
GPU).  A batch launches the per-step finish (the exact flux sum and the
convergence update) only after steps on which some instance's rules can act;
the next finish counts the quiet steps in between.  The check runs the
device's convergence rules for a batch of instances twice -- with a finish
after every step, as a solve without the schedule runs, and with the
schedules of its chunks -- and compares what every step kernel reads (phase,
flux sum, velocity write) and every state after a finish, bit for bit.

The flux sequences converge at different speeds, some never (eps = 0), some
turn non-finite, and the instances join at different batch steps, so that the
schedules meet convergence checks, phase ends at a check and at max_steps,
averaging phases (3D: a flux sum every step; 2D: every avg_sample steps),
failures and retired instances.  A planted wrong schedule (one finish per
chunk) must be caught.
"""
import ctypes as C

import numpy as np
import pytest


def flux_sequences(rng, n, nsteps, nd):
    t = np.arange(1, nsteps + 1)[None, :, None]
    qinf = rng.uniform(0.5, 2.0, size=(n, 1, 3)) * rng.choice([-1.0, 1.0], size=(n, 1, 3))
    tau = rng.uniform(20, 400, size=(n, 1, 1))
    amp = rng.uniform(0.2, 2.0, size=(n, 1, 1))
    noise = rng.choice([0.0, 1e-4, 1e-2], size=(n, 1, 1))
    q = qinf * (1 + amp * np.exp(-t / tau)) + noise * rng.standard_normal((n, nsteps, 3))
    if nd == 2:
        q[:, :, 2] = 0.0
    for i in range(0, n, 7):                       # a non-finite flux somewhere
        q[i, rng.integers(nsteps // 4, nsteps), rng.integers(0, nd)] = np.nan
    return np.ascontiguousarray(q)


def configs3d(rng, n):
    from flowperm import _lib
    out = (_lib.Config3D * n)()
    for c in out:
        c.eps = float(rng.choice([0.0, 1e-4, 1e-3, 1e-2, 0.05]))
        c.lag = int(rng.choice([1, 2, 3, 7, 16, 50, 199]))
        c.hold = int(rng.choice([-1, 0, 1, 1, 2, 3]))
        c.min_steps = int(rng.choice([0, 20, 300]))
        c.max_steps = int(rng.choice([150, 701, 5000]))
        c.avg_steps = int(rng.choice([0, 1, 5, 40]))
    return out


def configs2d(rng, n):
    from flowperm import _lib
    out = (_lib.Config2D * n)()
    for c in out:
        c.eps = float(rng.choice([0.0, 1e-3, 1e-2, 0.05]))
        c.ksigma = float(rng.choice([0.0, 1.0, 3.0]))
        c.qfloor = float(rng.choice([0.0, 0.1]))
        c.L0 = 1.0
        c.lag = int(rng.choice([1, 4, 9, 30, 120]))
        c.avg_sample = int(rng.choice([1, 3, 10]))
        c.hold = int(rng.choice([-1, 0, 1, 1, 2, 3]))
        c.min_steps = int(rng.choice([0, 50, 400]))
        c.max_steps = int(rng.choice([160, 903, 5000]))
        c.avg_steps = int(rng.choice([0, 7, 40, 150]))
    return out


def run(cfgs, lattice, q, start, chunk, replays, plant=0):
    from flowperm import _lib
    lib = _lib.load()
    n, nsteps = q.shape[0], q.shape[1]
    st = (C.c_int64 * n)(*[int(s) for s in start])
    prob, fin = C.c_int64(-1), C.c_int64(-1)
    c3 = cfgs if lattice == 3 else None
    c2 = cfgs if lattice == 2 else None
    rc = lib.fp_verify_finish_plan(c3, c2, n, _lib.f64ptr(q), nsteps, st, chunk, replays, plant,
                                   C.byref(prob), C.byref(fin))
    assert rc == 0
    return prob.value, fin.value


@pytest.mark.parametrize("lattice", [3, 2])
@pytest.mark.parametrize("chunk,replays", [(200, 1), (200, 3), (40, 5), (6, 4), (2, 1)])
def test_schedule_matches_a_finish_after_every_step(lattice, chunk, replays):
    rng = np.random.default_rng(100 * lattice + chunk + replays)
    n, nsteps = 24, 2400
    q = flux_sequences(rng, n, nsteps, lattice)
    cfgs = configs3d(rng, n) if lattice == 3 else configs2d(rng, n)
    rnd = chunk * replays
    start = rng.integers(0, nsteps // rnd // 2, size=n) * rnd
    prob, fin = run(cfgs, lattice, q, start, chunk, replays)
    assert prob == 0
    assert fin <= nsteps // rnd * rnd


@pytest.mark.parametrize("lattice", [3, 2])
def test_planted_schedule_is_caught(lattice):
    rng = np.random.default_rng(7 + lattice)
    n, nsteps = 12, 1200
    q = flux_sequences(rng, n, nsteps, lattice)
    cfgs = configs3d(rng, n) if lattice == 3 else configs2d(rng, n)
    prob, _ = run(cfgs, lattice, q, np.zeros(n), 200, 1, plant=1)
    assert prob > 0


def test_quiet_chunks_take_one_or_two_finishes():
    """pure stepping at the default 3D lag (2000 steps): a 200-step chunk
    needs its closing finish only, plus the one before a check step"""
    from flowperm import _lib
    rng = np.random.default_rng(3)
    q = flux_sequences(rng, 1, 4800, 3)
    q[:] = 1.0
    cfg = (_lib.Config3D * 1)()
    _lib.load().fp_config3d_default(cfg)
    cfg[0].eps = 0.0
    cfg[0].max_steps = 10 ** 12
    assert cfg[0].lag == 2000
    prob, fin = run(cfg, 3, q, [0], 200, 6)
    assert prob == 0
    # rounds of 6 x 200 steps: checks at steps 2000 and 4000 fall in the
    # second and fourth rounds, where the plan gains the finish before them
    assert fin == 6 + 12 + 6 + 12
    # a tolerance that can be met: from the first check at which hold can
    # reach its bound a chunk takes a finish after every step.  lag 1000,
    # hold 1, a constant flux, chunks of 300: the check at step 1000 (j = 99
    # of the fourth chunk) has no previous flux, so hold stays 0; the one at
    # 2000 (j = 199 of the seventh) ends the run (avg_steps 0).  Later chunks
    # hold no live instance and keep only their closing finish.
    cfg[0].eps = 1e-3
    cfg[0].lag = 1000
    cfg[0].hold = 1
    cfg[0].min_steps = 0
    cfg[0].avg_steps = 0
    prob, fin = run(cfg, 3, q, [0], 300, 1)
    assert prob == 0
    assert fin == 3 + (300 - 98) + 2 + (300 - 198) + 9
    # hold grows only at checks at or after min_steps: with min_steps 3500
    # the checks at 1000, 2000 and 3000 take the finishes before and after
    # them only, and the chunk of the check at 4000 (j = 99 of the 14th),
    # which ends the run, is dense from there
    cfg[0].min_steps = 3500
    prob, fin = run(cfg, 3, q, [0], 300, 1)
    assert prob == 0
    assert fin == 10 + 3 + 3 + 2 + (300 - 98) + 2


def test_2d_averaging_phase_takes_finishes_around_its_samples():
    """a 2D averaging phase (LMB2d/main.cpp:194-206) sums the flux every
    avg_sample steps: its chunks need finishes at its samples only.  max_steps
    200 ends the first phase at the end of the first chunk; 1000 averaging
    steps with a sample every 50 follow.  Chunks of 200: the first phase's
    chunk has its 4 block samples (and the end at 200), each averaging chunk
    its 4 samples (the last at the phase's end), each a finish before and
    after: 8 finishes; the two chunks after the end keep their closing one."""
    from flowperm import _lib
    cfg = (_lib.Config2D * 1)()
    _lib.load().fp_config2d_default(cfg)
    c = cfg[0]
    c.eps, c.ksigma, c.L0 = 0.0, 0.0, 1.0
    c.lag, c.min_steps, c.max_steps = 1000, 0, 200
    c.avg_steps, c.avg_sample = 1000, 50
    q = np.ones((1, 1600, 3))
    q[:, :, 2] = 0.0
    prob, fin = run(cfg, 2, q, [0], 200, 1)
    assert prob == 0
    assert fin == 6 * 8 + 2


def test_hold_zero_ends_the_first_phase_at_any_check():
    """with --hold 0 the first check ends the first phase whatever min_steps
    says (hold >= 0 holds after the reset, LMB3d/lbm3d-perm.cpp:345-352): lag
    50, min_steps 1000.  The first chunk is dense from the step before the
    check at 50 (152 finishes); the 100 averaging steps end at 150; the two
    chunks after keep their closing finish."""
    from flowperm import _lib
    cfg = (_lib.Config3D * 1)()
    _lib.load().fp_config3d_default(cfg)
    c = cfg[0]
    c.eps, c.hold, c.lag, c.min_steps, c.max_steps, c.avg_steps = 1e-3, 0, 50, 1000, 10 ** 6, 100
    q = np.ones((1, 600, 3))
    prob, fin = run(cfg, 3, q, [0], 200, 1)
    assert prob == 0
    assert fin == (200 - 48) + 1 + 1


def test_random_batches():
    """120 random batches: both lattices, chunks of 2-202 steps replayed up
    to 7 times, 1-29 instances joining at random chunk boundaries"""
    bad = []
    for seed in range(120):
        rng = np.random.default_rng(1000 + seed)
        lattice = 3 if seed % 2 else 2
        chunk = int(rng.choice([2, 4, 6, 10, 40, 200, 202]))
        replays = int(rng.choice([1, 2, 3, 7]))
        n, nsteps = int(rng.integers(1, 30)), 3000
        q = flux_sequences(rng, n, nsteps, lattice)
        cfgs = configs3d(rng, n) if lattice == 3 else configs2d(rng, n)
        rnd = chunk * replays
        start = rng.integers(0, max(1, nsteps // rnd // 2), size=n) * rnd
        prob, _ = run(cfgs, lattice, q, start, chunk, replays)
        if prob:
            bad.append((seed, prob))
    assert not bad, bad
