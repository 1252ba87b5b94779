"""Error returns of the C ABI on a live GPU context.

The reference exits the process on any CUDA error (`checkCudaErrors`,
FunctionOnGPU.cuh:24-30) and ignores bad options (MultiAxis.cpp:117-120);
here every failure is a negative status (raised by the binding as MtrError
carrying that code), nothing is launched for a rejected call, and the
context stays usable afterwards.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ERR_SHAPE, ERR_NONFINITE, ERR_STATE, ERR_RANGE = -2, -5, -6, -7


def small(mtr, seed=9):
    return mtr.synth.make_problem("e", 20, 18, 10, 4, [0.0], seed=seed, amin=2, amax=5,
                                  theta=np.linspace(-60, 60, 7))


def code_of(call):
    with pytest.raises(Exception) as ei:
        call()
    assert isinstance(ei.value, Exception) and hasattr(ei.value, "code"), ei.value
    return ei.value.code


def test_call_order_and_index_errors(mtr, orc):
    pb = small(mtr)
    prj, xc, yc, p = pb.prjs[0], pb.xcs[0], pb.ycs[0], pb.ps[0]
    with mtr.Context(0) as c:
        # an agent needs the volume box; a solve needs agents
        assert code_of(lambda: c.add_agent(prj, xc, yc, p)) == ERR_STATE
        c.set_volume(pb.vol)
        assert code_of(lambda: c.set_volume(pb.vol)) == ERR_STATE
        assert code_of(lambda: c.mace_solve(1, 1, 0.5, 0.5)) == ERR_STATE
        a = c.add_agent(prj, xc, yc, p)
        # norms exist only after an iteration has run
        assert code_of(lambda: c.residual_norm2(a)) == ERR_STATE
        assert code_of(lambda: c.consensus_diff2()) == ERR_STATE
        assert code_of(lambda: c.bpt(a + 1)) == ERR_RANGE
        assert code_of(lambda: c.sirt(-1, 1, 0.5)) == ERR_RANGE
        # the context still works, and agrees with the oracle
        xb = c.bpt(a)
        assert orc.rel_l2(xb, orc.bpt(prj, pb.vol, xc, yc, p)) < 2e-6


def test_bad_shapes_and_nonfinite_inputs(mtr):
    pb = small(mtr)
    prj, xc, yc, p = pb.prjs[0], pb.xcs[0], pb.ycs[0], pb.ps[0]
    with mtr.Context(0) as c:
        assert code_of(lambda: c.set_volume(mtr.volume_size(0, 4, 4))) == ERR_SHAPE
        c.set_volume(pb.vol)
        assert code_of(lambda: c.agent_footprint(mtr.projection_size(prj.X, 0, prj.AngN))) == ERR_SHAPE
        for which, idx, val in (("xc", 4 * 3 + 1, np.nan), ("xc", 0, np.inf), ("yc", 4 * 5 + 2, -np.inf)):
            bx, by = xc.copy(), yc.copy()
            (bx if which == "xc" else by)[idx] = val
            assert code_of(lambda: c.add_agent(prj, bx, by, p)) == ERR_NONFINITE
        bad = np.array(p, np.float32).reshape(-1)
        bad[prj.X * prj.Y * 3 + 7] = np.nan
        assert code_of(lambda: c.add_agent(prj, xc, yc, bad)) == ERR_NONFINITE
        # rejected agents leave nothing behind; a good one is agent 0
        assert c.add_agent(prj, xc, yc, p) == 0
        assert np.isfinite(c.bpt(0)).all()


def test_server_rejects_nonfinite_requests_at_submission(mtr):
    pb = small(mtr)
    bad = np.array(pb.ps[0], np.float32).reshape(-1)
    bad[5] = np.inf
    badx = pb.xcs[0].copy()
    badx[2] = np.nan
    with mtr.Server(0) as srv:
        assert code_of(lambda: srv.submit(pb.vol, pb.prjs, pb.xcs, pb.ycs, [bad], 1, 1, 0.5, 0.5)) == ERR_NONFINITE
        assert code_of(lambda: srv.submit(pb.vol, pb.prjs, [badx], pb.ycs, pb.ps, 1, 1, 0.5, 0.5)) == ERR_NONFINITE
        assert srv.status()["pending"] == 0
        rid = srv.submit(pb.vol, pb.prjs, pb.xcs, pb.ycs, pb.ps, 1, 1, 0.5, 0.5)
        for _ in range(100):
            if srv.status()["done"] >= 1:
                break
            srv.step()
        assert srv.status()["done"] == 1
        out, _ = srv.result(rid)
        assert np.isfinite(out).all()
