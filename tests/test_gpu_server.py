"""
The request server (fp_server_*): continuous batching with admission by
device-memory footprint.

A request's result must not depend on what else is resident: every forcing
run is stepped by the same kernels on its own arrays and its flux sums are
exact integers, so a request answered in a mix equals the same request solved
alone bit for bit, in FP64 and in FP32, and equals fp_solve3d / fp_solve2d
(which run the directions one after another).
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

CFG = dict(ff=40, eps=1e-6, lag=200, min_steps=400, avg_steps=200, max_steps=6000)
KEYS3 = [f"K_{a}{b}" for a in "xyz" for b in "xyz"] + \
        [f"{p}q{a}_{t}" for p in ("", "se_") for a in "xyz" for t in ("fx", "fy", "fz")] + \
        [f"steps_{t}" for t in ("fx", "fy", "fz")] + [f"conv_{t}" for t in ("fx", "fy", "fz")] + \
        ["k1", "k2", "k3", "fa", "recip_resid", "porosity"]


def masks3d():
    from flowperm import masks
    out = [golden("sph16_s0")[0],
           masks.spheres((20, 24, 28), 0.6, 3, 5, seed=3),
           masks.ellipsoids((24, 16, 20), 0.55, seed=4),
           masks.slit(12, 6)]
    return out


def alone(gpu_ctx, mask, precision, **extra):
    from flowperm.server import Server
    with Server(3, precision, ctx=gpu_ctx) as s:
        rid = s.submit(mask, **CFG, **extra)
        return s.run()[rid]


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_alone_equals_mix_bitwise(gpu_ctx, precision):
    from flowperm.server import Server
    ms = masks3d()
    ref = [alone(gpu_ctx, m, precision)[0] for m in ms]
    with Server(3, precision, ctx=gpu_ctx, chunk_steps=100) as s:
        # staggered: two requests, one chunk, then the rest join the running batch
        ids = [s.submit(ms[0], **CFG), s.submit(ms[1], **CFG)]
        s.step()
        ids += [s.submit(m, **CFG) for m in ms[2:]]
        res = s.run(ids)
        st = s.stats()
    print(f"{precision}: stats {st}", flush=True)
    assert st["completed"] == len(ms) and st["queued"] == 0 and st["bytes_in_use"] == 0
    for i, rid in enumerate(ids):
        row, info = res[rid]
        assert info["t_admit"] >= info["t_submit"] and info["t_done"] >= info["t_admit"]
        for k in KEYS3:
            assert row[k] == ref[i][k], (i, k)


def test_server_fp64_equals_plain_solve(gpu_ctx):
    """the server's FP64 row is the one-shot fp_solve3d row (bitwise), also
    with the --validate run stepping beside the three forcing runs"""
    from flowperm import solver
    m = masks3d()[1]
    row, _ = alone(gpu_ctx, m, "f64", validate=True)
    ref = solver.solve_permeability3d(m, ctx=gpu_ctx, validate=True, **CFG)
    for k in KEYS3 + ["val_err", "exit_code"]:
        assert row[k] == ref[k], k
    assert row["val_q"] == ref["val_q"]


def test_budget_admits_in_arrival_order(gpu_ctx):
    """a budget that holds one request at a time: the second is admitted only
    after the first has finished, and both results are unchanged"""
    from flowperm.server import Server
    ms = masks3d()[:2]
    ref = [alone(gpu_ctx, m, "f32")[0] for m in ms]
    with Server(3, "f32", ctx=gpu_ctx) as probe:
        rid = probe.submit(ms[0], **CFG)
        probe.step()
        big = max(probe.stats()["bytes_in_use"], 1)
        probe.run()
    with Server(3, "f32", ctx=gpu_ctx) as probe:
        rid = probe.submit(ms[1], **CFG)
        probe.step()
        big = max(big, probe.stats()["bytes_in_use"])
        probe.run()
    with Server(3, "f32", budget_bytes=big + 1024, ctx=gpu_ctx) as s:
        ids = [s.submit(m, **CFG) for m in ms]
        s.step()
        st = s.stats()
        assert st["resident"] == 1 and st["queued"] == 1
        res = s.run(ids)
        assert s.stats()["bytes_peak"] <= big + 1024
    (r0, i0), (r1, i1) = res[ids[0]], res[ids[1]]
    assert i1["t_admit"] >= i0["t_done"]
    for k in KEYS3:
        assert r0[k] == ref[0][k] and r1[k] == ref[1][k], k


def test_rejections(gpu_ctx):
    from flowperm.server import Server
    from flowperm._lib import FlowpermError
    with Server(3, "f32", budget_bytes=4096, ctx=gpu_ctx) as s:
        with pytest.raises(FlowpermError) as e:
            s.submit(masks3d()[0], **CFG)
        assert e.value.status == 14                    # FP_ERR_OOM: exceeds the budget alone
    with Server(3, "f32", ctx=gpu_ctx) as s:
        bad = np.zeros((4, 4, 4), np.uint8)
        bad[0, 0, 0] = 3
        with pytest.raises(FlowpermError) as e:
            s.submit(bad, **CFG)
        assert e.value.status == 11
        with pytest.raises(FlowpermError) as e:
            s.submit(np.ones((4, 4, 4), np.uint8), **CFG)
        assert e.value.status == 12
        assert s.step() == 0


def test_server_2d_equals_plain_solve(gpu_ctx):
    from flowperm import masks, solver
    from flowperm.server import Server
    cfg = dict(ff=1.0, eps=1e-3, lag=500, min_steps=1000, max_steps=8000, avg_steps=1000,
               avg_sample=50, validate=30.0)
    ms = [masks.ellipses((64, 64), 0.6, seed=1), masks.ellipses((48, 80), 0.5, seed=2)]
    with Server(2, "f32ref", ctx=gpu_ctx) as s:
        ids = [s.submit(m, **cfg) for m in ms]
        res = s.run(ids)
    for m, rid in zip(ms, ids):
        row, _ = res[rid]
        ref = solver.solve_permeability2d(m, ctx=gpu_ctx, **cfg)
        for k in ("k_xx", "k_yx", "k_xy", "k_yy", "steps_fx", "steps_fy", "fieldres_fx",
                  "fieldres_fy", "theta_deg", "val_err", "porosity"):
            assert row[k] == ref[k], k


def test_poisson_schedule_latencies(gpu_ctx):
    """Requests arriving on a Poisson schedule while others step: timestamps
    in order, and every row equal bit for bit to the same request alone."""
    from flowperm import masks
    from flowperm.server import Server, poisson_arrivals, serve_schedule
    ms = [masks.spheres((n, n, n), 0.6, 3, 5, seed=i) for i, n in enumerate((16, 20, 24, 16, 28))]
    arr = poisson_arrivals(50.0, len(ms), seed=1)
    with Server(3, "f32", ctx=gpu_ctx, chunk_steps=100) as s:
        out = serve_schedule(s, ms, arr, CFG)
    lat = [o["latency"] for o in out]
    print("latencies", [round(x, 4) for x in lat], flush=True)
    for m, o in zip(ms, out):
        assert o["t_submit"] >= o["arrival"] - 1e-6
        assert o["t_done"] >= o["t_admit"] >= o["t_submit"]
        assert o["latency"] > 0
        assert o["row"]["conv_fx"] == 1
        ref = alone(gpu_ctx, m, "f32")[0]
        for k in KEYS3:
            assert o["row"][k] == ref[k], k


def _footprint(gpu_ctx, m):
    from flowperm.server import Server
    with Server(3, "f32", ctx=gpu_ctx) as probe:
        probe.submit(m, **CFG)
        probe.step()
        b = probe.stats()["bytes_in_use"]
        probe.run()
    return b


def test_small_requests_pass_a_waiting_large_one_up_to_the_bypass_limit(gpu_ctx):
    """A request that does not fit yet is passed by later ones that do (they
    need no memory it is waiting for), at most max_bypass times; then arrival
    order holds again.  Every result is the one the request gives alone."""
    from flowperm import masks
    from flowperm.server import Server
    big = [masks.spheres((40, 40, 40), 0.6, 3, 5, seed=s) for s in (11, 12)]
    small = [masks.spheres((16, 16, 16), 0.6, 3, 5, seed=s) for s in (13, 14)]
    nb, ns = _footprint(gpu_ctx, big[0]), max(_footprint(gpu_ctx, m) for m in small)
    assert nb > 3 * ns
    budget = nb + 2 * ns + 4096                   # one large + two small, not two large
    order = [big[0], big[1], small[0], small[1]]
    ref = [alone(gpu_ctx, m, "f32")[0] for m in order]
    for bypass, want_resident in ((1, 2), (4, 3)):
        with Server(3, "f32", budget_bytes=budget, ctx=gpu_ctx, max_bypass=bypass) as s:
            ids = [s.submit(m, **CFG) for m in order]
            s.step()
            st = s.stats()
            print(f"max_bypass {bypass}: resident {st['resident']} queued {st['queued']}", flush=True)
            # bypass 1: the first small request passes the waiting large one, the
            # second may not; bypass 4: both pass it
            assert st["resident"] == want_resident
            res = s.run(ids)
        info = [res[i][1] for i in ids]
        assert info[2]["t_admit"] < info[1]["t_admit"]
        if bypass == 1:
            assert info[3]["t_admit"] >= info[1]["t_admit"]
        for i, rid in enumerate(ids):
            for k in KEYS3:
                assert res[rid][0][k] == ref[i][k], (bypass, i, k)


def test_strict_arrival_order_when_asked(gpu_ctx):
    from flowperm import masks
    from flowperm.server import Server
    big = masks.spheres((40, 40, 40), 0.6, 3, 5, seed=11)
    small = masks.spheres((16, 16, 16), 0.6, 3, 5, seed=13)
    nb = _footprint(gpu_ctx, big)
    with Server(3, "f32", budget_bytes=nb + 4096 + nb // 4, ctx=gpu_ctx, max_bypass=-1) as s:
        ids = [s.submit(big, **CFG), s.submit(big, **CFG), s.submit(small, **CFG)]
        s.step()
        assert s.stats()["resident"] == 1 and s.stats()["queued"] == 2
        res = s.run(ids)
    assert res[ids[2]][1]["t_admit"] >= res[ids[1]][1]["t_admit"]


def test_failed_chunk_ends_resident_requests_and_the_server_goes_on(gpu_ctx):
    """A chunk that fails (a table rebuild forced to fail, as a device error
    would, fp_test_fail_rebuilds) ends every resident request with
    FP_ERR_DEVICE and returns its memory.  Nothing of those requests stays in
    the batch: stepping again is idle, a failed request polled late keeps its
    status, and a new request solves exactly as alone."""
    from flowperm import _lib
    from flowperm._lib import FlowpermError
    from flowperm.server import Server
    ms = masks3d()[:2]
    ref = alone(gpu_ctx, ms[0], "f32")[0]
    lib = gpu_ctx.lib
    with Server(3, "f32", ctx=gpu_ctx) as s:
        ids = [s.submit(m, **CFG) for m in ms]
        lib.fp_test_fail_rebuilds(1)
        try:
            with pytest.raises(FlowpermError) as e:
                s.step()
        finally:
            lib.fp_test_fail_rebuilds(0)
        assert e.value.status == _lib.FP_ERR_DEVICE
        st = s.stats()
        assert st["bytes_in_use"] == 0 and st["resident"] == 0 and st["queued"] == 0
        assert s.step() == 0
        with pytest.raises(FlowpermError) as e0:
            s.poll(ids[0])
        assert e0.value.status == _lib.FP_ERR_DEVICE
        rid = s.submit(ms[0], **CFG)
        s.step()
        res = s.run([rid])
        with pytest.raises(FlowpermError) as e1:
            s.poll(ids[1])
        assert e1.value.status == _lib.FP_ERR_DEVICE
        assert s.stats()["bytes_in_use"] == 0
    for k in KEYS3:
        assert res[rid][0][k] == ref[k], k
