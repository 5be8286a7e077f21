"""Serving engine: continuous batching of device-resident session QPs.

The engine fixes the tile shape, so a QP's partial sums are reduced in the
same order whatever else is in the batch: every result must be bit-identical
alone and in any mix, and identical to the direct solve with that shape."""
import numpy as np
import pytest

import mprgp_amd as M
from mprgp_amd import serving

pytestmark = pytest.mark.gpu


def sessions(ctx, seed=11, count=6, dof_range=(1.5e3, 9e3)):
    return serving.make_sessions(ctx, count=count, seed=seed, dof_range=dof_range, reassemble=2)


def prec(s):
    return M.MIXED if s.sid % 2 == 0 else M.FP64


def test_engine_equals_direct_solve(ctx):
    a, b = sessions(ctx, seed=12, count=2), sessions(ctx, seed=12, count=2)
    eng = M.Engine(ctx, max_active=4, tiles=4)
    for sa, sb in zip(a, b):
        sa.prepare()
        sb.prepare()
        rc, st = sa.qp.solve(serving.solve_params(sa, 1e-6, prec(sa), tiles=4))
        t = eng.submit(sb.qp, serving.solve_params(sb, 1e-6, prec(sb)))
        rc2, st2 = eng.wait(t)
        assert rc == rc2 == 0
        assert st["iterations"] == st2["iterations"]
        np.testing.assert_array_equal(sa.qp.get_x(), sb.qp.get_x())
    eng.close()


# "small": every launch has few tiles (the tail variant of the step kernel);
# "streamed": sessions of 30k-60k DoF in one-slice tiles, so the mix launches
# the streamed variant (threshold 0 bytes) and a session alone the tail one.
@pytest.mark.parametrize("case", ["small", "streamed"])
def test_sessions_alone_equal_mix(ctx, case, monkeypatch):
    steps = 5
    if case == "small":
        alone, mix = sessions(ctx), sessions(ctx)
        eng = M.Engine(ctx, max_active=8, tiles=4, graph_steps=8)
    else:
        monkeypatch.setenv("MPRGP_STREAM_BYTES", "0")
        alone, mix = sessions(ctx, seed=13, dof_range=(3e4, 6e4)), sessions(ctx, seed=13, dof_range=(3e4, 6e4))
        eng = M.Engine(ctx, max_active=8, tiles=1, graph_steps=8)
    for s in alone:  # one session at a time
        for _ in range(steps):
            s.prepare()
            rc, st = eng.wait(eng.submit(s.qp, serving.solve_params(s, 1e-6, prec(s))))
            assert rc == 0, st
            s.complete(keep_history=True, rc=rc, stats=st)
    # all sessions at once; a finished session immediately issues its next
    # step, so the sessions drift apart and every step meets a different mix
    rng = np.random.default_rng(3)
    order = rng.permutation(len(mix))
    live = {}
    for i in order:
        s = mix[i]
        s.prepare()
        live[eng.submit(s.qp, serving.solve_params(s, 1e-6, prec(s)))] = s
    peak = 0
    while live:
        peak = max(peak, eng.info()["active"])
        t, rc, st = eng.wait_any(-1)
        assert rc == 0, st
        s = live.pop(t)
        s.complete(keep_history=True, rc=rc, stats=st)
        if s.step < steps:
            s.prepare()
            live[eng.submit(s.qp, serving.solve_params(s, 1e-6, prec(s)))] = s
    info = eng.info()
    print(f"engine: {info}")
    assert info["peak_active"] > 1
    for sa, sm in zip(alone, mix):
        assert len(sa.history) == len(sm.history) == steps
        for k, (xa, xm) in enumerate(zip(sa.history, sm.history)):
            np.testing.assert_array_equal(xa, xm, err_msg=f"session {sa.sid} step {k}")
    eng.close()


def test_budget_caps_admission(ctx):
    ss = sessions(ctx, seed=13, count=5)
    for s in ss:
        s.prepare()
    need = [ss[0].qp, ss[1].qp]
    import mprgp_amd._lib as L
    # budget: the two largest working sets side by side, never three
    sizes = sorted(((s.fem.matrix.info()["bytes"]) for s in ss), reverse=True)
    eng = M.Engine(ctx, max_active=8, tiles=4, hbm_budget=int(3.2 * sizes[0]))
    ts = [eng.submit(s.qp, serving.solve_params(s, 1e-6, M.FP64)) for s in ss]
    for t in ts:
        rc, st = eng.wait(t)
        assert rc == 0
    info = eng.info()
    print(f"budget {3.2 * sizes[0]:.0f} B: {info}")
    assert info["peak_bytes"] <= 3.2 * sizes[0]
    assert info["completed"] == len(ss)
    eng.close()


def test_submit_errors(ctx):
    ss = sessions(ctx, seed=14, count=1)
    s = ss[0]
    s.prepare()
    eng = M.Engine(ctx, max_active=2, tiles=4, hbm_budget=1024)
    with pytest.raises(M.MprgpError) as e:
        eng.submit(s.qp, serving.solve_params(s, 1e-6, M.FP64))
    assert e.value.code == -7  # EBUDGET
    eng.close()
    eng = M.Engine(ctx, max_active=2, tiles=4)
    t = eng.submit(s.qp, serving.solve_params(s, 1e-6, M.FP64))
    with pytest.raises(M.MprgpError) as e:
        eng.submit(s.qp, serving.solve_params(s, 1e-6, M.FP64))
    assert e.value.code == -8  # busy
    with pytest.raises(M.MprgpError) as e:
        s.qp.set_x(s.scene.x_n.ravel())
    assert e.value.code == -8
    rc, st = eng.wait(t)
    assert rc == 0
    s.qp.set_x(s.scene.x_n.ravel())  # free again once collected
    eng.close()


def test_open_loop_small(ctx):
    ss = sessions(ctx, seed=15, count=8)
    eng = M.Engine(ctx, max_active=16, tiles=4)
    out = serving.run_open_loop(eng, ss, rate=300.0, duration=1.0, precision=M.MIXED, workers=2)
    print(out)
    assert out["converged"] == out["requests"]
    assert out["p99_ms"] >= out["p50_ms"] > 0
    eng.close()


def test_device_sessions_equal_host_sessions(ctx):
    """Sessions whose state, rows and velocity updates stay on the device
    (begin_step/end_step) step bit-identically to sessions that keep the
    state on the host, assemble from host arrays and upload the host
    collider's rows; re-assembly steps (every 2nd) read the device state back
    for the rotation."""
    dev = serving.make_sessions(ctx, count=3, seed=15, dof_range=(1.5e3, 9e3), reassemble=2)
    host = serving.make_sessions(ctx, count=3, seed=15, dof_range=(1.5e3, 9e3), reassemble=2,
                                 device_state=False)
    for sd, sh in zip(dev, host):
        for step in range(5):
            for s in (sd, sh):
                s.prepare()
                rc, st = s.qp.solve(serving.solve_params(s, 1e-6, prec(s), tiles=4))
                assert rc == 0, (s.sid, step, st)
                if s is sd:
                    nd = s.qp.get_rows()
                else:
                    nh = s.qp.get_rows()
                s.complete(keep_history=True, rc=rc, stats=st)
            for a, b in zip(nd, nh):
                np.testing.assert_array_equal(a, b)
            np.testing.assert_array_equal(sd.history[-1], sh.history[-1], err_msg=f"{sd.sid} {step}")
        xd, vd = sd.fem.get_state()
        np.testing.assert_array_equal(vd.reshape(-1, 3), sh.scene.v_n)


def test_config5_full_scale(ctx):
    """The serving workload at its stated scale: 64 sessions (beams and
    blobs, DoF log-uniform in [3k, 150k], each on its own context),
    corotated re-assembly every 10 steps, a 200 GB HBM budget, Poisson
    arrivals at 2000 requests/s for 1 s, mixed form, the simulator's
    iteration cap (MprgpFemSolver.cpp:189).  Every request converges, p50/p99
    are reported with an error bound per solve, and the 4 largest sessions,
    replayed alone step by step on the same engine, reproduce their positions
    in the mix bit for bit."""
    ss = serving.make_sessions(ctx, count=64, seed=5, dof_range=(3e3, 1.5e5), reassemble=10,
                               own_contexts=True)
    dof = [3 * s.n_nodes for s in ss]
    big = set(int(i) for i in np.argsort(dof)[-4:])
    eng = M.Engine(ctx, max_active=256, tiles=4, hbm_budget=int(200e9))
    out = serving.run_open_loop(eng, ss, rate=2000.0, duration=1.0, precision=M.MIXED, seed=16,
                                keep_history=big)
    print({k: v for k, v in out.items() if k != "engine"}, f"DoF {min(dof)}..{max(dof)} total {sum(dof)}")
    assert min(dof) >= 3000 and max(dof) <= 150000 * 1.05 and max(dof) >= 60000
    assert out["requests"] >= 1500
    assert out["converged"] == out["requests"]
    assert out["p99_ms"] >= out["p50_ms"] > 0
    assert out["error_bound_m"]["reported"] == out["requests"]
    assert out["engine"]["peak_bytes"] <= 200e9
    solo = serving.make_sessions(ctx, count=64, seed=5, dof_range=(3e3, 1.5e5), reassemble=10,
                                 own_contexts=True, only=big)
    for s in solo:
        mixed_run = ss[s.sid]
        assert 3 * s.n_nodes == dof[s.sid] and len(mixed_run.history) >= 5
        for k in range(len(mixed_run.history)):
            s.prepare()
            rc, st = eng.wait(eng.submit(s.qp, serving.solve_params(s, 1e-6, M.MIXED)))
            assert rc == 0, st
            s.complete(keep_history=True, rc=rc, stats=st)
            np.testing.assert_array_equal(s.history[k], mixed_run.history[k],
                                          err_msg=f"session {s.sid} ({dof[s.sid]} DoF) step {k}")
    eng.close()


def test_fp64_error_sample(ctx):
    """Served mixed solves against FP64 solves of the same QPs to 1e-10: the
    per-solve error bound holds and the error is small."""
    ss = sessions(ctx, seed=17, count=6, dof_range=(3e3, 3e4))
    eng = M.Engine(ctx, max_active=8, tiles=4)
    for _ in range(3):  # a few steps in
        for s in ss:
            s.prepare()
            rc, st = eng.wait(eng.submit(s.qp, serving.solve_params(s, 1e-6, M.MIXED)))
            s.complete(rc=rc, stats=st)
    out = serving.fp64_error_sample(eng, ss, n=6)
    print(out)
    assert out["bound_holds"] and out["error_bound_max_m"] > 0
    assert out["rel_err_max"] <= 1e-4
    eng.close()
