"""Convergence control and checkpoint/resume of the fast MACE solve.

The reference always runs its -n MACE steps in one process
(MultiAxis.cpp:40-71) and keeps no state between runs; `mtr_mace_continue`
adds the two controls a long solve needs, on the same step kernels:

* continuing from the state (w_a, w-bar) is bit-identical to a solve that
  never stopped, in the same context or in a new one restored from a
  checkpoint (get_volume/set_volume_data, or a checkpoint file), in every
  storage precision;
* with rel_tol > 0 the loop stops after the first step whose FP64 consensus
  difference sum_a ||w_a - w-bar||^2 changed by at most rel_tol of the
  previous step's value, and the result is that of a solve of that many steps;
* a checkpoint file of other inputs, or a damaged one, is refused before
  any state is replaced.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_STATE = -1, -6


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def problem(mtr, seed=31):
    # a row agent, a column agent and a separable agent
    return mtr.synth.make_problem("resume", 34, 30, 14, 6, [0.0, 45.0, 90.0], seed=seed, amin=2, amax=6,
                                  theta=np.linspace(-60, 60, 9))


def load(mtr, pb, precision=None):
    c = mtr.Context(0) if precision is None else mtr.Context(0, precision)
    c.set_volume(pb.vol)
    for a in range(pb.K):
        c.add_agent(pb.prjs[a], pb.xcs[a], pb.ycs[a], pb.ps[a])
    return c


PRECS = [("fp32", 0), ("bf16", 1), ("fp16", 2)]


@pytest.mark.parametrize("name,prec", PRECS, ids=[p[0] for p in PRECS])
def test_continue_in_place_equals_one_solve(mtr, name, prec):
    pb = problem(mtr)
    say("continue in place", name)
    with load(mtr, pb, prec) as c:
        whole = c.mace_solve(4, 3, 0.5, 0.5)
        d_whole = c.consensus_diff2()
    with load(mtr, pb, prec) as c:
        c.mace_init()
        _, n1 = c.mace_continue(1, 3, 0.5, 0.5)
        _, n2 = c.mace_continue(3, 3, 0.5, 0.5)
        parts = c.get_volume(mtr.VOL_WBAR)
        assert (n1, n2) == (1, 3)
        assert c.consensus_diff2() == d_whole
    assert np.array_equal(whole, parts)


@pytest.mark.parametrize("name,prec", PRECS, ids=[p[0] for p in PRECS])
def test_consensus_difference_matches_an_independent_fp64_sum(mtr, name, prec):
    # the early-stop quantity sum_a ||w_a - w-bar||^2 (the consensus of
    # MultiAxis.cpp:61-66) against a sum formed in FP64 numpy from the
    # stored state after every step, for every agent kind and storage type
    pb = problem(mtr, seed=36)
    say("consensus difference vs FP64 numpy", name)
    with load(mtr, pb, prec) as c:
        c.mace_init()
        for step in range(4):
            c.mace_continue(1, 2, 0.5, 0.5, out=False)
            wbar = c.get_volume(mtr.VOL_WBAR).astype(np.float64)
            ref = 0.0
            for a in range(pb.K):
                d = c.get_volume(mtr.VOL_W, a).astype(np.float64) - wbar
                ref += float(np.dot(d.ravel(), d.ravel()))
            got = c.consensus_diff2()
            say(f"  step {step + 1}: device {got:.17e}, numpy {ref:.17e}")
            assert ref > 0
            assert abs(got - ref) <= 1e-12 * ref


@pytest.mark.parametrize("name,prec", PRECS, ids=[p[0] for p in PRECS])
def test_resume_from_checkpoint_in_a_new_context(mtr, name, prec):
    pb = problem(mtr, seed=32)
    say("resume from checkpoint", name)
    with load(mtr, pb, prec) as c:
        whole = c.mace_solve(3, 2, 0.6, 0.4)
    with load(mtr, pb, prec) as c:
        c.mace_solve(1, 2, 0.6, 0.4, out=False)
        ckpt = [c.get_volume(mtr.VOL_W, a) for a in range(pb.K)] + [c.get_volume(mtr.VOL_WBAR)]
    with load(mtr, pb, prec) as c:
        for a in range(pb.K):
            c.set_volume_data(mtr.VOL_W, a, ckpt[a])
        c.set_volume_data(mtr.VOL_WBAR, 0, ckpt[-1])
        resumed, n = c.mace_continue(2, 2, 0.6, 0.4)
    assert n == 2
    assert np.array_equal(whole, resumed)


def test_zero_steps_leave_the_state(mtr):
    pb = problem(mtr, seed=33)
    with load(mtr, pb) as c:
        w1 = c.mace_solve(1, 2, 0.5, 0.5)
        out, n = c.mace_continue(0, 2, 0.5, 0.5)
        assert n == 0 and np.array_equal(out, w1)


def test_early_stop_follows_the_consensus_difference(mtr):
    pb = problem(mtr, seed=34)
    M, S, lam, rho = 12, 2, 0.5, 0.5
    say("early stop trajectory")
    # the trajectory of the consensus difference, one step per call
    d = []
    with load(mtr, pb) as c:
        c.mace_init()
        for _ in range(M):
            c.mace_continue(1, S, lam, rho, out=False)
            d.append(c.consensus_diff2())
    assert all(np.isfinite(d)) and d[-1] < d[0]
    changes = [abs(d[m] - d[m - 1]) / d[m - 1] for m in range(1, M)]
    # a tolerance that stops strictly inside the run
    tol = float(np.median(changes))
    expect = next(m + 1 for m in range(1, M) if abs(d[m] - d[m - 1]) <= tol * d[m - 1])
    assert 2 <= expect < M
    say("early stop at", expect, "of", M, "tol", tol)
    with load(mtr, pb) as c:
        c.mace_init()
        stopped, n = c.mace_continue(M, S, lam, rho, rel_tol=tol)
        assert c.consensus_diff2() == d[n - 1]
    assert n == expect
    with load(mtr, pb) as c:
        assert np.array_equal(stopped, c.mace_solve(n, S, lam, rho))
    # rel_tol 0 never stops early
    with load(mtr, pb) as c:
        c.mace_init()
        assert c.mace_continue(M, S, lam, rho, rel_tol=0.0, out=False)[1] == M


def test_continue_errors(mtr):
    pb = problem(mtr, seed=35)
    with mtr.Context(0) as c:
        c.set_volume(pb.vol)
        for call in (lambda: c.mace_continue(1, 1, 0.5, 0.5),):
            with pytest.raises(Exception) as ei:
                call()
            assert ei.value.code == ERR_STATE
        c.add_agent(pb.prjs[0], pb.xcs[0], pb.ycs[0], pb.ps[0])
        for bad in (dict(M=-1), dict(S=-1), dict(rel_tol=-1e-3), dict(rel_tol=float("nan"))):
            kw = dict(M=1, S=1, lam=0.5, rho=0.5, rel_tol=0.0)
            kw.update(bad)
            with pytest.raises(Exception) as ei:
                c.mace_continue(**kw)
            assert ei.value.code == ERR_ARG, bad
        # without init the state starts at zero: a solve from the zero volume
        out, n = c.mace_continue(1, 1, 0.5, 0.5)
        assert n == 1 and np.isfinite(out).all()


def test_checkpoint_file_round_trip(mtr, tmp_path):
    pb = problem(mtr, seed=36)
    path = str(tmp_path / "mace.ckpt")
    say("checkpoint file")
    with load(mtr, pb) as c:
        whole = c.mace_solve(3, 2, 0.5, 0.5)
    with load(mtr, pb) as c:
        c.mace_solve(2, 2, 0.5, 0.5, out=False)
        c.checkpoint_save(path, 2)
    n = pb.vol.X * pb.vol.Y * pb.vol.Z
    raw = np.fromfile(path, np.uint8)
    assert raw.size == 64 + 4 * n * (pb.K + 1) and raw[:8].tobytes() == b"MTRCKPT1"
    with load(mtr, pb) as c:
        assert c.checkpoint_load(path) == 2
        resumed, steps = c.mace_continue(1, 2, 0.5, 0.5)
    assert steps == 1 and np.array_equal(whole, resumed)


def test_bad_checkpoints_leave_the_state(mtr, tmp_path):
    pb = problem(mtr, seed=37)
    good = tmp_path / "good.ckpt"
    with load(mtr, pb) as c:
        c.mace_solve(1, 1, 0.5, 0.5, out=False)
        c.checkpoint_save(str(good), 1)
    raw = good.read_bytes()
    cases = {
        "magic": (b"XXXXXXXX" + raw[8:], -2),
        "short": (raw[:-4], -2),
        "long": (raw + b"\0\0\0\0", -2),
        "nan": (raw[:100] + np.float32(np.nan).tobytes() + raw[104:], -5),
    }
    say("bad checkpoints")
    with load(mtr, pb) as c:
        c.mace_init()
        before = [c.get_volume(mtr.VOL_W, a) for a in range(pb.K)] + [c.get_volume(mtr.VOL_WBAR)]
        for name, (data, code) in cases.items():
            p = tmp_path / f"{name}.ckpt"
            p.write_bytes(data)
            with pytest.raises(Exception) as ei:
                c.checkpoint_load(str(p))
            assert ei.value.code == code, name
        after = [c.get_volume(mtr.VOL_W, a) for a in range(pb.K)] + [c.get_volume(mtr.VOL_WBAR)]
        assert all(np.array_equal(x, y) for x, y in zip(before, after))
        with pytest.raises(Exception) as ei:
            c.checkpoint_load(str(tmp_path / "missing.ckpt"))
        assert ei.value.code == ERR_ARG
    # other inputs: one agent fewer, or other coefficients
    with mtr.Context(0) as c:
        c.set_volume(pb.vol)
        c.add_agent(pb.prjs[0], pb.xcs[0], pb.ycs[0], pb.ps[0])
        with pytest.raises(Exception) as ei:
            c.checkpoint_load(str(good))
        assert ei.value.code == -2
    with mtr.Context(0) as c:
        c.set_volume(pb.vol)
        for a in range(pb.K):
            xc = pb.xcs[a].copy()
            xc[0] += 0.25 if a == 1 else 0.0
            c.add_agent(pb.prjs[a], xc, pb.ycs[a], pb.ps[a])
        with pytest.raises(Exception) as ei:
            c.checkpoint_load(str(good))
        assert ei.value.code == -2
    # another storage precision would not continue bit-exactly
    with load(mtr, pb, 1) as c:
        with pytest.raises(Exception) as ei:
            c.checkpoint_load(str(good))
        assert ei.value.code == -2
    # a save replaces the file whole, through a temporary file beside it
    with load(mtr, pb) as c:
        c.mace_solve(2, 1, 0.5, 0.5, out=False)
        c.checkpoint_save(str(good), 2)
        assert not (tmp_path / "good.ckpt.tmp").exists()
        assert c.checkpoint_load(str(good)) == 2
    assert good.stat().st_size == len(raw)
