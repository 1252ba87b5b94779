"""Split-precision mode on the GPU: projections, residuals and volumes stored
in BF16 or FP16, ray and back-projection sums accumulated in FP32, residual
norms and consensus differences reduced in FP64.

Each result is compared with the FP64 oracle on the same seeded inputs and
its relative L2 error is printed.  Stated bounds (relative L2 against the
FP64 oracle, 3-agent MACE, 3 steps of 4 SIRT iterations):
  BF16 storage  <= 2e-2   (8-bit significand: unit roundoff 2^-9 = 2e-3 per store)
  FP16 storage  <= 3e-3   (11-bit significand: unit roundoff 2^-12 = 2.4e-4 per store)
  FP32 storage  <= 1e-5
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOUND = {"fp32": 1e-5, "bf16": 2e-2, "fp16": 3e-3}


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def problem(mtr, seed=21):
    return mtr.synth.make_problem("prec", 30, 26, 14, 8, [0.0, 90.0, 45.0], seed=seed, amin=2, amax=6,
                                  theta=np.linspace(-60, 60, 11))


def solve(mtr, pb, prec, M=3, S=4):
    with mtr.Context(0, prec) as c:
        c.set_volume(pb.vol)
        for a in range(pb.K):
            c.add_agent(pb.prjs[a], pb.xcs[a], pb.ycs[a], pb.ps[a])
        out = c.mace_solve(M, S, 0.5, 0.5)
        work = c.get_volume(mtr.VOL_WORK, 0)
        return out, work, c.consensus_diff2(), [c.residual_norm2(a) for a in range(pb.K)]


def precs(mtr):
    return {"fp32": mtr.PREC_FP32, "bf16": mtr.PREC_BF16, "fp16": mtr.PREC_FP16}


@pytest.fixture(scope="module")
def reference(mtr, orc):
    pb = problem(mtr)
    ref64 = orc.mace_f64(pb.prjs, pb.vol, pb.xcs, pb.ycs, pb.ps, 3, 4, 0.5, 0.5)
    return pb, ref64


@pytest.mark.parametrize("name", ["fp32", "bf16", "fp16"])
def test_mace_error_against_fp64_oracle(mtr, orc, reference, name):
    pb, ref64 = reference
    out, _, d2, r2 = solve(mtr, pb, precs(mtr)[name])
    err = orc.rel_l2(out, ref64)
    say(f"{name}: MACE rel_l2 vs FP64 oracle = {err:.3e} (bound {BOUND[name]:.0e}); "
        f"consensus diff2 {d2:.6e}, residual norms {[f'{v:.6e}' for v in r2]}")
    assert np.isfinite(out).all()
    assert err < BOUND[name]


def test_storage_is_really_reduced(mtr, reference):
    # every stored value of the working volume is representable in the
    # storage type (a BF16 value has 16 zero low bits as FP32; an FP16 value
    # survives a round trip through float16)
    pb, _ = reference
    _, w_bf, _, _ = solve(mtr, pb, mtr.PREC_BF16, M=1, S=1)
    assert np.all((w_bf.view(np.uint32) & 0xFFFF) == 0)
    _, w_h, _, _ = solve(mtr, pb, mtr.PREC_FP16, M=1, S=1)
    assert np.array_equal(w_h.astype(np.float16).astype(np.float32), w_h)
    _, w_f, _, _ = solve(mtr, pb, mtr.PREC_FP32, M=1, S=1)
    assert np.any((w_f.view(np.uint32) & 0xFFFF) != 0)


@pytest.mark.parametrize("name", ["bf16", "fp16"])
def test_fp64_reductions_track_fp32(mtr, reference, name):
    # the convergence quantities are reduced in FP64 from the stored values;
    # in reduced storage they stay close to the FP32 run's
    pb, _ = reference
    _, _, d32, r32 = solve(mtr, pb, mtr.PREC_FP32)
    _, _, dlo, rlo = solve(mtr, pb, precs(mtr)[name])
    tol = 0.02 if name == "bf16" else 0.005
    say(f"{name}: diff2 {dlo:.6e} vs fp32 {d32:.6e}; r2 {rlo} vs {r32}")
    assert dlo == pytest.approx(d32, rel=tol)
    for a, b in zip(rlo, r32):
        assert a == pytest.approx(b, rel=tol)


@pytest.mark.parametrize("name", ["bf16", "fp16"])
def test_split_precision_is_deterministic(mtr, reference, name):
    pb, _ = reference
    a = solve(mtr, pb, precs(mtr)[name], M=2, S=2)
    b = solve(mtr, pb, precs(mtr)[name], M=2, S=2)
    assert np.array_equal(a[0], b[0]) and a[2] == b[2] and a[3] == b[3]


@pytest.mark.parametrize("name", ["fp32", "bf16", "fp16"])
def test_deep_column_agent_against_fp64_oracle(mtr, orc, name):
    # oblique axis with 300-plane columns and tilts in 5-degree steps: the
    # column-plan forward projection's multi-plane lanes and grouped reads
    # in every storage type (BPT, then 2 SIRT iterations)
    pb = mtr.synth.make_problem("deep", 24, 20, 300, 6, [45.0], seed=23, amin=2, amax=6,
                                theta=np.linspace(-60, 60, 25), nu=330)
    prj, vol, xc, yc, p = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], pb.ps[0]
    x0 = orc.bpt_f64(prj, vol, xc, yc, p)
    ref = orc.sirt_f64(prj, vol, xc, yc, p, x0, 2, 0.5)
    with mtr.Context(0, precs(mtr)[name]) as c:
        c.set_volume(vol)
        c.add_agent(prj, xc, yc, p)
        assert c.agent_kernels(0) == (mtr.FPK_COLUMN, mtr.BPK_COLUMN)
        c.bpt(0, out=False)
        got = c.sirt(0, 2, 0.5)
    err = orc.rel_l2(got, ref)
    say(f"{name}: deep column SIRT rel_l2 vs FP64 oracle = {err:.3e} (bound {BOUND[name]:.0e})")
    assert err < BOUND[name]
