"""Request server (continuous batching) on the GPU.

A request's result must not depend on what shares its launches: the same
request is solved alone, in a mix of requests of other sizes, axis counts
and iteration counts that arrive while it is active, and by a standalone
context.  Batching only puts several agents into one launch; each agent's
blocks, sums and reductions are its own, so the three results are compared
bit for bit (a stricter bar than the FP32 tolerance, 1e-5 relative L2,
which the oracle comparison uses).  Admission follows arrival order (or
smallest work first, with a bound on the wait) within the HBM budget and
the slot count.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SOLVE_TOL = 1e-5


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def req(mtr, X, Y, Z, phis, seed, M=3, S=2):
    return mtr.synth.make_problem("srv", X, Y, Z, 6, phis, seed=seed, amin=2, amax=6,
                                  theta=np.linspace(-60, 60, 9), M=M, S=S, lam=0.5, rho=0.5)


def drain(srv, n):
    steps = 0
    while srv.status()["done"] < n:
        srv.step()
        steps += 1
        assert steps < 1000
    return steps


def test_result_is_the_same_alone_in_a_mix_and_standalone(mtr, orc):
    pbs = {
        "a": req(mtr, 40, 36, 12, [0.0, 90.0], 41),
        "b": req(mtr, 48, 40, 16, [0.0, 60.0, 120.0], 42),  # the probe
        "c": req(mtr, 32, 32, 10, [0.0], 43, M=2),
        "d": req(mtr, 36, 44, 14, [0.0, 45.0, 90.0, 135.0], 44, M=4, S=3),
    }
    say("serve alone")
    with mtr.Server(0) as srv:
        rb = srv.submit_problem(pbs["b"])
        drain(srv, 1)
        alone, _ = srv.result(rb)
    say("serve mix")
    with mtr.Server(0) as srv:
        ids = {"a": srv.submit_problem(pbs["a"])}
        srv.step()
        ids["b"] = srv.submit_problem(pbs["b"])
        ids["c"] = srv.submit_problem(pbs["c"])
        srv.step()
        assert srv.status()["active"] >= 2  # b runs alongside a
        ids["d"] = srv.submit_problem(pbs["d"])
        drain(srv, 4)
        mix = {k: srv.result(r)[0] for k, r in ids.items()}
    assert np.array_equal(alone, mix["b"])
    say("standalone context")
    pb = pbs["b"]
    with mtr.Context(0) as c:
        c.set_volume(pb.vol)
        for a in range(pb.K):
            c.add_agent(pb.prjs[a], pb.xcs[a], pb.ycs[a], pb.ps[a])
        solo = c.mace_solve(pb.M, pb.S, pb.lam, pb.rho)
    assert np.array_equal(solo, alone)
    for k, pb in pbs.items():
        ref = orc.mace(pb.prjs, pb.vol, pb.xcs, pb.ycs, pb.ps, pb.M, pb.S, pb.lam, pb.rho)
        err = orc.rel_l2(mix[k], ref)
        say(f"request {k}: rel_l2 vs oracle {err:.2e}")
        assert err < SOLVE_TOL


def test_admission_follows_hbm_budget_in_arrival_order(mtr):
    pbs = [req(mtr, 40, 40, 12, [0.0, 90.0], 50 + q, M=2) for q in range(3)]
    with mtr.Server(0) as probe:
        need = probe.footprint(pbs[0].vol, pbs[0].prjs)
    budget = int(need * 1.5)  # room for one request at a time
    say("budget", need, budget)
    with mtr.Server(0, hbm_budget=budget, max_active=8) as srv:
        ids = [srv.submit_problem(p) for p in pbs]
        assert srv.status()["pending"] == 3
        srv.step()
        st = srv.status()
        assert st["active"] == 1 and st["pending"] == 2 and 0 < st["hbm_in_use"] <= budget
        drain(srv, 3)
        t = [srv.times(r) for r in ids]
        assert srv.status()["hbm_in_use"] == 0
    # admitted in arrival order, each after its predecessor retired
    for q in range(1, 3):
        assert t[q][1] >= t[q - 1][2]


def test_smallest_first_admission_order_and_results(mtr):
    """With one slot, smallest-first admission runs the queued requests in
    order of work (voxel-view updates of the whole solve) instead of
    arrival, a request that has waited past the bound goes first, and every
    result equals the one from arrival-order admission bit for bit."""
    pbs = [req(mtr, 40, 40, 12, [0.0, 90.0], 70, M=2),  # largest
           req(mtr, 24, 24, 8, [0.0], 71, M=2),  # smallest
           req(mtr, 32, 32, 10, [0.0, 90.0], 72, M=2)]
    res = {}
    for policy, wait in ((mtr.ADMIT_ARRIVAL, 0.0), (mtr.ADMIT_SMALLEST, 1e6), (mtr.ADMIT_SMALLEST, 0.0)):
        say("admission policy", policy, "max wait", wait)
        with mtr.Server(0, max_active=1) as srv:
            srv.set_admission(policy, wait)
            ids = [srv.submit_problem(p) for p in pbs]
            drain(srv, 3)
            t = [srv.times(r)[1] for r in ids]
            res[policy, wait] = [srv.result(r)[0] for r in ids]
        order = sorted(range(3), key=lambda q: t[q])
        if policy == mtr.ADMIT_SMALLEST and wait > 0:
            assert order == [1, 2, 0]
        else:  # arrival order, or every request past a zero wait bound
            assert order == [0, 1, 2]
    for key, got in res.items():
        for q in range(3):
            assert np.array_equal(got[q], res[mtr.ADMIT_ARRIVAL, 0.0][q]), (key, q)
    with mtr.Server(0) as srv:
        for bad in ((2, 0.0), (mtr.ADMIT_SMALLEST, -1.0), (mtr.ADMIT_SMALLEST, float("nan"))):
            with pytest.raises(mtr.MtrError):
                srv.set_admission(*bad)


def test_slot_limit_and_oversized_request(mtr):
    pbs = [req(mtr, 24, 24, 8, [0.0], 60 + q, M=3) for q in range(4)]
    with mtr.Server(0, max_active=2) as srv:
        for p in pbs:
            srv.submit_problem(p)
        srv.step()
        assert srv.status()["active"] == 2 and srv.status()["pending"] == 2
        drain(srv, 4)
    with mtr.Server(0) as probe:
        need = probe.footprint(pbs[0].vol, pbs[0].prjs)
    with mtr.Server(0, hbm_budget=need // 2) as srv:
        with pytest.raises(mtr.MtrError) as ei:
            srv.submit_problem(pbs[0])
        assert ei.value.code == -3


def test_zero_step_request_returns_the_initial_model(mtr):
    pb = req(mtr, 30, 26, 9, [0.0, 45.0], 70, M=0)
    with mtr.Server(0) as srv:
        r = srv.submit_problem(pb)
        drain(srv, 1)
        got, _ = srv.result(r)
    with mtr.Context(0) as c:
        c.set_volume(pb.vol)
        for a in range(pb.K):
            c.add_agent(pb.prjs[a], pb.xcs[a], pb.ycs[a], pb.ps[a])
        ref = c.mace_solve(0, pb.S, pb.lam, pb.rho)
    assert np.array_equal(got, ref)


def test_background_preparation_takes_memory_at_submission(mtr, monkeypatch):
    """With background preparation (the default) a submitted request is
    queued for the worker at once, so its footprint counts as in use before
    any tick; with MTR_SERVER_PREFETCH=0 nothing is taken until the tick
    that admits it.  Both give the same result, and a mix of requests
    prepared while others run equals the same mix prepared inside ticks."""
    pbs = [req(mtr, 40, 40, 12, [0.0, 90.0], 80, M=2), req(mtr, 32, 28, 10, [45.0], 81, M=3),
           req(mtr, 36, 36, 12, [0.0, 60.0, 120.0], 82, M=2)]
    res = {}
    for pf in ("1", "0"):
        monkeypatch.setenv("MTR_SERVER_PREFETCH", pf)
        say("server preparation", pf)
        with mtr.Server(0, max_active=2) as srv:
            need = srv.footprint(pbs[0].vol, pbs[0].prjs)
            ids = [srv.submit_problem(pbs[0])]
            used = srv.status()["hbm_in_use"]
            assert used == (need if pf == "1" else 0)
            srv.step()
            ids += [srv.submit_problem(p) for p in pbs[1:]]
            drain(srv, 3)
            res[pf] = [srv.result(r)[0] for r in ids]  # waits for each hand-over
            assert srv.status()["hbm_in_use"] == 0
    for a, b in zip(res["1"], res["0"]):
        assert np.array_equal(a, b)


def test_a_step_returns_while_it_runs_and_the_next_retires(mtr):
    # a tick returns with its launches in flight; the next step waits for
    # them and hands the requests that finished to the retirer, so a request
    # of M steps is done after M + 1 step calls, and its result is ready
    # only then (the result call waits for the retirer's copy)
    pb = req(mtr, 32, 28, 10, [0.0, 90.0], 71, M=2, S=2)
    with mtr.Server(0) as srv:
        rid = srv.submit_problem(pb)
        assert srv.step() == 0 and srv.status()["active"] == 1
        assert srv.step() == 0 and srv.status()["done"] == 0
        with pytest.raises(mtr.MtrError):
            srv.result(rid)  # the last step is still in flight
        assert srv.step() == 1
        st = srv.status()
        assert st["done"] == 1 and st["active"] == 0
        out, _ = srv.result(rid)
        assert srv.status()["hbm_in_use"] == 0  # retired: its device memory is back
        ts, ta, td = srv.times(rid)
        assert ts <= ta <= td
    with mtr.Context(0) as c:
        c.set_volume(pb.vol)
        for a in range(pb.K):
            c.add_agent(pb.prjs[a], pb.xcs[a], pb.ycs[a], pb.ps[a])
        ref = c.mace_solve(pb.M, pb.S, pb.lam, pb.rho)
    assert np.array_equal(out, ref)
