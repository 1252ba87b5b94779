"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Floating-point tolerances (relative L2 = ||gpu - oracle|| / ||oracle||):
  single operators (FP, BP, BPT)  <= 2e-6   (FP32 summation order only)
  SIRT / MACE reconstructions     <= 1e-5   (the FP32-mode bar)
Element-wise MACE kernels and the residual are bit-exact.
"""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OP_TOL = 2e-6
SOLVE_TOL = 1e-5


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def problem(mtr, phis, kind=None, X=24, Y=20, Z=12, seed=5, views=9, nu=None, nv=None, origin=(0, 0, 0)):
    kind = mtr.synth.GEOM_SLAB if kind is None else kind
    pb = mtr.synth.make_problem("t", X, Y, Z, 8, list(phis), kind=kind, seed=seed, amin=2, amax=6,
                                theta=np.linspace(-60, 60, views), nu=nu, nv=nv)
    if origin != (0, 0, 0):
        # same data, shifted box: the geometry constant absorbs the shift
        xs, ys, zs = origin
        pb.vol = mtr.volume_size(X, Y, Z, xs, ys, zs)
        for a in range(pb.K):
            for v in range(pb.prjs[a].AngN):
                pb.xcs[a][4 * v] -= pb.xcs[a][4 * v + 1] * xs + pb.xcs[a][4 * v + 2] * ys + pb.xcs[a][4 * v + 3] * zs
                pb.ycs[a][4 * v] -= pb.ycs[a][4 * v + 1] * xs + pb.ycs[a][4 * v + 2] * ys + pb.ycs[a][4 * v + 3] * zs
    return pb


GEOMS = [
    ("axisZ", dict(phis=[0.0], kind=1)),
    ("slab0", dict(phis=[0.0])),
    ("slab90", dict(phis=[90.0])),
    ("slab45", dict(phis=[45.0])),
    ("slab120", dict(phis=[120.0])),
    ("ragged", dict(phis=[30.0], X=37, Y=23, Z=11, nu=29, nv=41)),
    ("origin", dict(phis=[90.0], origin=(5, -3, 2))),
    ("small_det", dict(phis=[0.0], nu=15, nv=9)),
]


def random_general(mtr, rng, A, nu, nv, vol):
    from test_oracle import random_affine  # noqa: WPS433

    return random_affine(rng, A, nu, nv, vol)


@pytest.mark.parametrize("name,kw", GEOMS, ids=[g[0] for g in GEOMS])
def test_fp_matches_oracle(mtr, orc, ctx, name, kw):
    pb = problem(mtr, **kw)
    rng = np.random.default_rng(1)
    x = rng.uniform(0.0, 2.0, (pb.vol.Z, pb.vol.Y, pb.vol.X)).astype(np.float32)
    say("fp", name)
    s, c = ctx.op_fp(pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], x)
    s0, c0 = orc.fp(pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], x)
    assert orc.rel_l2(c, c0) < OP_TOL
    assert orc.rel_l2(s, s0) < OP_TOL
    assert np.array_equal(c == 0, c0 == 0)


def test_fp_general_affine(mtr, orc, ctx):
    rng = np.random.default_rng(3)
    vol = mtr.volume_size(21, 17, 13, 2, -4, 1)
    prj = mtr.projection_size(23, 19, 6)
    xc, yc = random_general(mtr, rng, prj.AngN, prj.X, prj.Y, vol)
    x = rng.uniform(-1, 1, (vol.Z, vol.Y, vol.X)).astype(np.float32)
    say("fp general")
    s, c = ctx.op_fp(prj, vol, xc, yc, x)
    s0, c0 = orc.fp(prj, vol, xc, yc, x)
    assert orc.rel_l2(c, c0) < OP_TOL
    assert orc.rel_l2(s, s0) < OP_TOL


@pytest.mark.parametrize("name,kw", GEOMS, ids=[g[0] for g in GEOMS])
def test_bp_and_bpt_match_oracle(mtr, orc, ctx, name, kw):
    pb = problem(mtr, **kw)
    prj, vol, xc, yc = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0]
    rng = np.random.default_rng(2)
    r = rng.normal(size=(prj.AngN, prj.Y, prj.X)).astype(np.float32)
    x0 = rng.uniform(0, 1, (vol.Z, vol.Y, vol.X)).astype(np.float32)
    say("bp", name)
    xb = ctx.op_bp(prj, vol, xc, yc, r, x0, 0.5)
    assert orc.rel_l2(xb, orc.bp(prj, vol, xc, yc, r, x0, 0.5)) < OP_TOL
    say("bpt", name)
    xt = ctx.op_bpt(prj, vol, xc, yc, pb.ps[0])
    assert orc.rel_l2(xt, orc.bpt(prj, vol, xc, yc, pb.ps[0])) < OP_TOL


def test_bp_general_affine(mtr, orc, ctx):
    rng = np.random.default_rng(4)
    vol = mtr.volume_size(70, 9, 19, -7, 3, 0)
    prj = mtr.projection_size(31, 27, 5)
    xc, yc = random_general(mtr, rng, prj.AngN, prj.X, prj.Y, vol)
    r = rng.normal(size=(prj.AngN, prj.Y, prj.X)).astype(np.float32)
    x0 = rng.normal(size=(vol.Z, vol.Y, vol.X)).astype(np.float32)
    say("bp general")
    xb = ctx.op_bp(prj, vol, xc, yc, r, x0, 1.0)
    assert orc.rel_l2(xb, orc.bp(prj, vol, xc, yc, r, x0, 1.0)) < OP_TOL


def test_residual_is_bit_exact(orc, ctx):
    rng = np.random.default_rng(5)
    p = rng.normal(size=4099).astype(np.float32)
    s = rng.normal(size=4099).astype(np.float32)
    c = rng.uniform(0, 3, 4099).astype(np.float32)
    c[::7] = 0
    say("residual")
    assert np.array_equal(ctx.op_residual(p, s, c), orc.residual(p, s, c))


def test_line_golden_on_gpu(mtr, ctx):
    g = json.load(open(os.path.join(GOLD, "line.json")))
    vol = mtr.volume_size(*g["vol"])
    prj = mtr.projection_size(*g["prj"])
    x = np.array(g["x"], np.float32).reshape(1, 1, 2)
    say("line golden")
    s, c = ctx.op_fp(prj, vol, g["xc"], g["yc"], x)
    assert s.ravel().tolist() == g["s"] and c.ravel().tolist() == g["c"]
    r = ctx.op_residual(np.array(g["p"], np.float32), s.ravel(), c.ravel())
    assert r.tolist() == g["r"]
    xb = ctx.op_bp(prj, vol, g["xc"], g["yc"], r.reshape(s.shape), x, g["lam"])
    assert xb.ravel().tolist() == g["x_after_sirt1"]
    xt = ctx.op_bpt(prj, vol, g["xc"], g["yc"], np.array(g["p"], np.float32).reshape(s.shape))
    assert xt.ravel().tolist() == g["bpt"]


# Resident-agent kernels (selected from the geometry) against the oracle:
# streaming FP / separable BP for row-separable agents, the column plan for
# oblique in-plane axes, the general kernels otherwise.
AGENT_GEOMS = GEOMS + [
    ("col60_deep", dict(phis=[60.0], X=40, Y=36, Z=75)),           # two 64-plane chunks per column
    ("col135_small_det", dict(phis=[135.0], X=30, Y=30, Z=20, nu=17, nv=13)),  # footprint off the detector
    ("col45_origin", dict(phis=[45.0], origin=(3, -2, 5))),
    # boxes whose Z range straddles a multiple of the 16-plane back-projection
    # runs (one run more than ceil(Z / 16)), wide enough for 24+ tiles
    ("col45_straddle", dict(phis=[45.0], X=128, Y=96, Z=8, origin=(0, 0, 12))),
    ("col135_negz", dict(phis=[135.0], X=96, Y=128, Z=10, origin=(-4, 3, -5))),
    ("col75_wide", dict(phis=[75.0], X=150, Y=21, Z=9)),            # several back-projection tiles
    # deep columns, tilts in 5-degree steps: forward-projection chunks read
    # in groups of 4 (Z >= 256) and 2 (Z >= 128), lanes owning 4 planes
    # (|sin theta| < 1/4) and 2 planes (< 1/2, and > 1/2 without a scan) in
    # groups of 1 and 2 chunks, whole chunks after the last group, and the
    # partial last chunk
    ("col30_z300", dict(phis=[30.0], X=40, Y=36, Z=300, views=25, nu=360)),
    ("col45_z256", dict(phis=[45.0], X=24, Y=28, Z=256, views=25, nu=300)),
    ("col120_z160", dict(phis=[120.0], X=36, Y=40, Z=160, views=25, nu=200)),
    ("col60_z520", dict(phis=[60.0], X=16, Y=20, Z=520, views=25, nu=560)),
    ("col135_z400", dict(phis=[135.0], X=20, Y=16, Z=400, views=25, nu=440)),
]


def expected_kernels(mtr, name):
    if name in ("axisZ", "slab0"):
        return mtr.FPK_STREAM, mtr.BPK_SEPARABLE
    if name in ("slab90", "origin"):
        # planes of the box fall off the detector's rows: the separable
        # back-projection needs every plane on a row
        return mtr.FPK_STREAM, mtr.BPK_GENERAL
    # oblique axes, and the 15 x 9 detector whose separable rays would need
    # more than four candidates per line
    return mtr.FPK_COLUMN, mtr.BPK_COLUMN


@pytest.mark.parametrize("name,kw", AGENT_GEOMS, ids=[g[0] for g in AGENT_GEOMS])
def test_agent_kernels_match_oracle(mtr, orc, name, kw):
    pb = problem(mtr, **kw)
    prj, vol, xc, yc, p = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], pb.ps[0]
    say("agent kernels", name)
    with load_ctx(mtr, pb) as c:
        say("kernels", c.agent_kernels(0))
        assert c.agent_kernels(0) == expected_kernels(mtr, name)
        xb = c.bpt(0)
        x_ref = orc.bpt(prj, vol, xc, yc, p)
        assert orc.rel_l2(xb, x_ref) < OP_TOL
        x1 = c.sirt(0, 1, 0.7)
        assert orc.rel_l2(x1, orc.sirt(prj, vol, xc, yc, p, x_ref, 1, 0.7)) < OP_TOL
        # FP64 norm of the residual formed by the forward projection of the
        # volume that iteration started from
        s0, c0 = orc.fp(prj, vol, xc, yc, xb)
        r = orc.residual(p, s0, c0).astype(np.float64)
        assert c.residual_norm2(0) == pytest.approx(float((r * r).sum()), rel=1e-5)


CS_GEOMS = [
    ("col45", dict(phis=[45.0], X=70, Y=66, Z=40, views=13)),
    ("col135_deep", dict(phis=[135.0], X=40, Y=36, Z=300, views=13, nu=340)),
    ("col60_small_det", dict(phis=[60.0], X=30, Y=30, Z=20, nu=17, nv=13)),
    ("col45_straddle", dict(phis=[45.0], X=128, Y=96, Z=8, origin=(0, 0, 12))),
    # rows of the box's corners off the detector (columns outside every strip)
    ("col45_corners", dict(phis=[45.0], X=30, Y=26, Z=14, views=9)),
]


@pytest.mark.parametrize("name,kw", CS_GEOMS, ids=[g[0] for g in CS_GEOMS])
@pytest.mark.parametrize("prec", [0, 1], ids=["fp32", "bf16"])
def test_strip_tiled_back_projection_equals_rectangular_tiles(mtr, name, kw, prec, monkeypatch):
    # the column plan's back-projection on strip tiles (64 columns of one
    # strip per block) and on 64 x 8 column tiles: same per-voxel
    # arithmetic and order, so BPT, normaliser and SIRT volumes agree bit
    # for bit
    pb = problem(mtr, **kw)
    outs = []
    for strip in ("1", "0"):
        monkeypatch.setenv("MTR_BP_COL_STRIP", strip)
        say("strip tiles", name, prec, strip)
        with load_ctx(mtr, pb, prec) as c:
            assert c.agent_kernels(0) == (mtr.FPK_COLUMN, mtr.BPK_COLUMN)
            outs.append((c.bpt(0), c.sirt(0, 2, 0.6)))
    assert np.isfinite(outs[0][0]).all() and np.isfinite(outs[0][1]).all()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def load_ctx(mtr, pb, precision=0):
    c = mtr.Context(0, precision)
    c.set_volume(pb.vol)
    for a in range(pb.K):
        c.add_agent(pb.prjs[a], pb.xcs[a], pb.ycs[a], pb.ps[a])
    return c


@pytest.mark.parametrize("phi,kind", [(0.0, 1), (90.0, 0), (45.0, 0)])
def test_sirt_matches_oracle(mtr, orc, phi, kind):
    pb = problem(mtr, [phi], kind=kind, X=32, Y=28, Z=16, views=13)
    prj, vol, xc, yc, p = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], pb.ps[0]
    say("sirt", phi, kind)
    with load_ctx(mtr, pb) as c:
        x0 = c.bpt(0)
        assert orc.rel_l2(x0, orc.bpt(prj, vol, xc, yc, p)) < OP_TOL
        out = c.sirt(0, 10, 1.0)
        x9 = orc.sirt(prj, vol, xc, yc, p, orc.bpt(prj, vol, xc, yc, p), 9, 1.0)
        ref = orc.sirt(prj, vol, xc, yc, p, x9, 1, 1.0)
        assert orc.rel_l2(out, ref) < SOLVE_TOL
        # FP64 residual norm of the last iteration against the oracle's
        s, cc = orc.fp(prj, vol, xc, yc, x9)
        r = orc.residual(p, s, cc).astype(np.float64)
        assert c.residual_norm2(0) == pytest.approx(float((r * r).sum()), rel=1e-4)


def test_mace_golden_on_gpu(mtr):
    g = json.load(open(os.path.join(GOLD, "mace.json")))
    vol = mtr.volume_size(1, 1, 1)
    prj = mtr.projection_size(1, 1, 1)
    xc, yc = [1.0, 1.0, 0.0, 0.0], [1.0, 0.0, 1.0, 0.0]
    for form in (mtr.FORM_FAST, mtr.FORM_PLAIN):
        for case in g["cases"]:
            for M, expect in enumerate(case["wbar"]):
                say("mace golden", form, case["lam"], M)
                with mtr.Context(0) as c:
                    c.set_volume(vol)
                    for q in g["p"]:
                        c.add_agent(prj, xc, yc, np.array([[[q]]], np.float32))
                    out = c.mace_solve(M, g["S"], case["lam"], g["rho"], form=form)
                    assert float(out.ravel()[0]) == expect
                    assert float(c.get_volume(mtr.VOL_W, 0).ravel()[0]) == case["w1"][M]
                    assert float(c.get_volume(mtr.VOL_W, 1).ravel()[0]) == case["w2"][M]


@pytest.mark.parametrize("form", [0, 1], ids=["fast", "plain"])
def test_mace_matches_oracle(mtr, orc, form):
    pb = problem(mtr, [0.0, 90.0, 45.0], X=30, Y=26, Z=14, views=11, seed=8)
    say("mace", form)
    with load_ctx(mtr, pb) as c:
        out = c.mace_solve(3, 4, 0.5, 0.5, form=form)
    ref = orc.mace(pb.prjs, pb.vol, pb.xcs, pb.ycs, pb.ps, 3, 4, 0.5, 0.5)
    ref64 = orc.mace_f64(pb.prjs, pb.vol, pb.xcs, pb.ycs, pb.ps, 3, 4, 0.5, 0.5)
    assert orc.rel_l2(out, ref) < SOLVE_TOL
    assert orc.rel_l2(out, ref64) < SOLVE_TOL


def test_mace_steps_through_api_match_solve(mtr, orc):
    # the reference-shaped single calls (BPT, z, SIRT, Mann, mean) compose
    # to the same result as the fused solve
    pb = problem(mtr, [0.0, 90.0], X=20, Y=20, Z=10, views=7, seed=9)
    say("mace api")
    with load_ctx(mtr, pb) as c:
        fused = c.mace_solve(2, 3, 0.5, 0.5)
    with load_ctx(mtr, pb) as c:
        c.mace_init()
        for _ in range(2):
            for a in range(pb.K):
                c.form_z(a)
                c.sirt(a, 3, 0.5, out=False)
                c.mann(a, 0.5)
            c.mean()
        stepped = c.get_volume(mtr.VOL_WBAR)
    assert np.array_equal(fused, stepped)


def test_kind_sorted_batches_match_per_agent_calls(mtr):
    # agents of alternating kernel kinds (separable, column, separable,
    # column): the fused solve launches each kind once per iteration through
    # the kind-sorted tables, the single calls one agent at a time
    pb = problem(mtr, [0.0, 45.0, 90.0, 135.0], X=28, Y=24, Z=10, views=7, seed=19)
    say("kind-sorted batches")
    with load_ctx(mtr, pb) as c:
        kinds = [c.agent_kernels(a)[0] for a in range(pb.K)]
        assert kinds[0] == kinds[2] != kinds[1] == kinds[3]
        fused = c.mace_solve(2, 2, 0.5, 0.5)
    with load_ctx(mtr, pb) as c:
        c.mace_init()
        for _ in range(2):
            for a in range(pb.K):
                c.form_z(a)
                c.sirt(a, 2, 0.5, out=False)
                c.mann(a, 0.5)
            c.mean()
        stepped = c.get_volume(mtr.VOL_WBAR)
    assert np.array_equal(fused, stepped)


def test_kind_streams_give_identical_results(mtr, monkeypatch):
    # the column-plan agents iterate on the context's stream and the others
    # on a second one (MTR_KIND_STREAMS=1, the default) or all on one: each
    # agent's iterations touch only its own buffers, so the solve, its
    # consensus difference and every agent's state are the same bit for bit,
    # replayed from a graph or issued from the host (timing keeps one stream)
    pb = problem(mtr, [0.0, 45.0, 90.0, 135.0], X=40, Y=36, Z=16, views=9, seed=23)
    out = {}
    for ks in ("1", "0"):
        monkeypatch.setenv("MTR_KIND_STREAMS", ks)
        say("kind streams", ks)
        with load_ctx(mtr, pb) as c:
            g = c.mace_solve(3, 3, 0.5, 0.5)
            d2 = c.consensus_diff2()
            w = [c.get_volume(mtr.VOL_W, a) for a in range(pb.K)]
            c.use_graphs(False)
            e = c.mace_solve(2, 3, 0.5, 0.5)
            c.timing(True)
            t = c.mace_solve(2, 3, 0.5, 0.5)
            c.timing(False)
        out[ks] = (g, d2, w, e, t)
    g1, d1, w1, e1, t1 = out["1"]
    g0, d0, w0, e0, t0 = out["0"]
    assert np.array_equal(g1, g0) and d1 == d0
    assert all(np.array_equal(x, y) for x, y in zip(w1, w0))
    assert np.array_equal(e1, e0) and np.array_equal(t1, t0) and np.array_equal(e1, t1)


def test_elementwise_bit_exact(mtr, orc):
    pb = problem(mtr, [0.0, 90.0], X=16, Y=12, Z=6, views=5)
    rng = np.random.default_rng(12)
    shape = (pb.vol.Z, pb.vol.Y, pb.vol.X)
    w = [rng.normal(size=shape).astype(np.float32) for _ in range(2)]
    wb = rng.normal(size=shape).astype(np.float32)
    v = rng.normal(size=shape).astype(np.float32)
    say("elementwise")
    with load_ctx(mtr, pb) as c:
        for a in range(2):
            c.set_volume_data(mtr.VOL_W, a, w[a])
        c.set_volume_data(mtr.VOL_WBAR, 0, wb)
        c.form_z(0)
        z = c.get_volume(mtr.VOL_WORK, 0)
        assert np.array_equal(z, orc.form_z(w[0], wb))
        c.set_volume_data(mtr.VOL_WORK, 0, v)
        c.mann(0, 0.3)
        assert np.array_equal(c.get_volume(mtr.VOL_W, 0), orc.mann(v, w[0], orc.form_z(w[0], wb), 0.3))
        c.mean()
        assert np.array_equal(c.get_volume(mtr.VOL_WBAR), orc.mean([c.get_volume(mtr.VOL_W, 0), w[1]]))


def test_solve_is_deterministic(mtr):
    pb = problem(mtr, [0.0, 45.0], X=40, Y=36, Z=12, views=9, seed=13)
    say("determinism")
    outs = []
    for _ in range(2):
        with load_ctx(mtr, pb) as c:
            outs.append(c.mace_solve(2, 3, 0.5, 0.5))
            outs.append(c.consensus_diff2())
    assert np.array_equal(outs[0], outs[2]) and outs[1] == outs[3]


def test_c1_full_size_matches_oracle(mtr, orc):
    pb = mtr.synth.config("c1")
    say("c1 full size")
    with load_ctx(mtr, pb) as c:
        out = c.mace_solve(pb.M, pb.S, pb.lam, pb.rho)
    ref = orc.mace(pb.prjs, pb.vol, pb.xcs, pb.ycs, pb.ps, pb.M, pb.S, pb.lam, pb.rho)
    err = orc.rel_l2(out, ref)
    say("c1 rel_l2 vs oracle", err)
    assert err < SOLVE_TOL


def test_c2_full_size_one_step_matches_oracle(mtr, orc):
    pb = mtr.synth.config("c2")
    say("c2 full size, M=1 S=1")
    with load_ctx(mtr, pb) as c:
        out = c.mace_solve(1, 1, pb.lam, pb.rho)
    ref = orc.mace(pb.prjs, pb.vol, pb.xcs, pb.ycs, pb.ps, 1, 1, pb.lam, pb.rho)
    err = orc.rel_l2(out, ref)
    say("c2 rel_l2 vs oracle", err)
    assert err < SOLVE_TOL


def test_c2_full_solve_properties(mtr):
    pb = mtr.synth.config("c2")
    say("c2 full solve")
    with load_ctx(mtr, pb) as c:
        out = c.mace_solve(pb.M, pb.S, pb.lam, pb.rho)
        d2 = c.consensus_diff2()
        r2 = [c.residual_norm2(a) for a in range(pb.K)]
    assert np.isfinite(out).all() and np.isfinite(d2) and all(np.isfinite(r2))
    assert float(np.abs(out).max()) > 0


@pytest.mark.parametrize("phi,kind", [(0.0, 0), (90.0, 0), (0.0, 1), (45.0, 0)], ids=["W=Y", "W=X", "W=Z", "general"])
def test_working_volume_layout_round_trip(mtr, phi, kind):
    # agents keep their working volume with the plane axis contiguous; the
    # API always exchanges the reference's [Z][Y][X] layout
    pb = problem(mtr, [phi], kind=kind, X=37, Y=29, Z=21, views=5, seed=31)
    rng = np.random.default_rng(7)
    v = rng.normal(size=(pb.vol.Z, pb.vol.Y, pb.vol.X)).astype(np.float32)
    say("layout round trip", phi, kind)
    with load_ctx(mtr, pb) as c:
        c.set_volume_data(mtr.VOL_WORK, 0, v)
        assert np.array_equal(c.get_volume(mtr.VOL_WORK, 0), v)
        out = c.sirt(0, 0, 0.5, x0=v)
        assert np.array_equal(out, v)


def test_graph_replays_match_host_issued_launches(mtr):
    # A fast solve captures one MACE step into a HIP graph and replays it M
    # times; every replay count is checked against host-issued launches,
    # with the context's own eager work (BPT, initial model, host-issued
    # solves) and synchronisations between the replays.
    pb = problem(mtr, [0.0, 90.0, 45.0], X=30, Y=26, Z=14, views=9, seed=61)
    say("graph replays")
    with load_ctx(mtr, pb) as eager, load_ctx(mtr, pb) as graph:
        eager.use_graphs(False)
        for M in (1, 2, 3):
            ref = eager.mace_solve(M, 3, 0.5, 0.5)
            got = graph.mace_solve(M, 3, 0.5, 0.5)
            assert np.array_equal(got, ref), M
            assert graph.consensus_diff2() == eager.consensus_diff2()
            for a in range(pb.K):
                assert np.array_equal(graph.get_volume(mtr.VOL_W, a), eager.get_volume(mtr.VOL_W, a))
        # host-issued solve on the graph context, then replays again
        graph.use_graphs(False)
        mid = graph.mace_solve(2, 3, 0.5, 0.5)
        graph.use_graphs(True)
        again = graph.mace_solve(2, 3, 0.5, 0.5)
        assert np.array_equal(mid, again)
        # new solve parameters capture a new graph
        assert np.array_equal(graph.mace_solve(2, 2, 0.4, 0.6), eager.mace_solve(2, 2, 0.4, 0.6))


def test_autotuned_tiles_do_not_change_results(mtr):
    # the tiles only regroup work: every voxel still sums its views in order
    pb = problem(mtr, [0.0, 90.0, 60.0, 135.0], X=70, Y=66, Z=40, views=9, seed=62)
    say("autotune")
    with load_ctx(mtr, pb) as tuned, load_ctx(mtr, pb) as fixed:
        fixed.autotune(False)
        a = tuned.mace_solve(2, 2, 0.5, 0.5)
        b = fixed.mace_solve(2, 2, 0.5, 0.5)
        say("tiles", [tuned.agent_tiles(q) for q in range(pb.K)], [fixed.agent_tiles(q) for q in range(pb.K)])
        assert np.array_equal(a, b)
        for q in range(pb.K):
            run, lanes = tuned.agent_tiles(q)
            assert run in (0, 8, 16, 32) and lanes in (0, 1, 2)  # 0: the general kernel


STREAM_GEOMS = [
    ("slab0", dict(phis=[0.0])),
    ("slab90", dict(phis=[90.0])),
    ("ragged", dict(phis=[0.0], X=37, Y=23, Z=11, nu=29, nv=41)),
    ("origin", dict(phis=[90.0], origin=(5, -3, 2))),
    ("axisZ", dict(phis=[0.0], kind=1, nu=27)),
    ("wide", dict(phis=[90.0], X=70, Y=66, Z=20, nu=83, nv=90, views=21)),
]


@pytest.mark.parametrize("name,kw", STREAM_GEOMS, ids=[g[0] for g in STREAM_GEOMS])
def test_stream_fp_columns_per_wave_give_identical_sums(mtr, orc, name, kw):
    # the streaming FP loads each voxel row once for G adjacent detector
    # columns; every column keeps its own tap order, so G = 1, 2, 4 agree
    # bit for bit (through SIRT, which back-projects the residual)
    pb = problem(mtr, **kw)
    say("stream G", name)
    outs = []
    old = os.environ.get("MTR_FP_COLS")
    try:
        for g in ("1", "2", "4"):
            os.environ["MTR_FP_COLS"] = g
            with load_ctx(mtr, pb) as c:
                assert c.agent_kernels(0)[0] == mtr.FPK_STREAM
                c.bpt(0, out=False)
                x = c.sirt(0, 2, 0.5)
                outs.append((x, c.residual_norm2(0)))
    finally:
        if old is None:
            os.environ.pop("MTR_FP_COLS", None)
        else:
            os.environ["MTR_FP_COLS"] = old
    for x, r2 in outs[1:]:
        # the volumes bit for bit; the FP64 residual norm sums per-block
        # partials, whose grouping follows the block count of each G
        assert np.array_equal(x, outs[0][0])
        assert r2 == pytest.approx(outs[0][1], rel=1e-12)
    prj, vol, xc, yc, p = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], pb.ps[0]
    ref = orc.sirt(prj, vol, xc, yc, p, orc.bpt(prj, vol, xc, yc, p), 2, 0.5)
    assert orc.rel_l2(outs[0][0], ref) < SOLVE_TOL


SEG_GEOMS = [
    ("slab0", dict(phis=[0.0], Z=21)),
    ("slab90", dict(phis=[90.0], Z=19)),
    ("wide", dict(phis=[0.0], X=70, Y=66, Z=23, nu=83, nv=90, views=21)),
    ("ragged", dict(phis=[0.0], X=37, Y=23, Z=17, nu=29, nv=41)),
    ("origin", dict(phis=[90.0], Z=18, origin=(5, -3, 2))),
    ("axisZ", dict(phis=[0.0], kind=1, nu=27)),
]


@pytest.mark.parametrize("name,kw", SEG_GEOMS, ids=[g[0] for g in SEG_GEOMS])
def test_stream_fp_in_z_segments(mtr, orc, name, kw, monkeypatch):
    # the streaming FP of a slab in Z segments (4 planes here, 128 at full
    # size): each segment sums its taps in (line, point) order into partial
    # ray sums that a second kernel adds in segment order and turns into
    # the residual.  Against whole rays (FP32 reassociation only) and the
    # oracle; an agent whose plane axis is Z is never segmented.
    pb = problem(mtr, **kw)
    # nonzero data on every pixel, also on rows of planes outside the box
    # (no voxel: r = p there)
    pb.ps[0] = (pb.ps[0] + np.random.default_rng(9).uniform(0.0, 0.05, pb.ps[0].shape)).astype(np.float32)
    prj, vol, xc, yc, p = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], pb.ps[0]
    outs = {}
    for seg, order in (("0", "1"), ("4", "1"), ("4", "0")):
        # block order: by ray crossing (1) or view fastest (0)
        monkeypatch.setenv("MTR_FP_STREAM_SEG", seg)
        monkeypatch.setenv("MTR_FP_STREAM_ORDER", order)
        say("stream segments", name, seg, order)
        with load_ctx(mtr, pb) as c:
            assert c.agent_kernels(0)[0] == mtr.FPK_STREAM
            c.timing(True)
            x0 = c.bpt(0)
            x = c.sirt(0, 1, 0.5)
            outs[seg if order == "1" else "4v"] = (x0, x, c.residual_norm2(0),
                                                    c.timing_get(mtr.K_FP_STREAM_RES)[0])
    # the block order changes which block sums a ray, not its sum
    assert np.array_equal(outs["4"][1], outs["4v"][1])
    assert outs["4"][2] == pytest.approx(outs["4v"][2], rel=1e-12)
    assert outs["0"][3] == 0
    assert (outs["4"][3] > 0) == (kw.get("kind") != 1)
    assert np.array_equal(outs["0"][0], outs["4"][0])  # BPT: no FP
    # residual ||p - s/c||^2 of the BPT volume against FP64 numpy on the
    # oracle's sums
    s0, c0 = orc.fp(prj, vol, xc, yc, outs["4"][0])
    r = orc.residual(p, s0, c0).astype(np.float64)
    ref_r2 = float(np.dot(r.ravel(), r.ravel()))
    assert outs["4"][2] == pytest.approx(ref_r2, rel=1e-5)
    assert outs["4"][2] == pytest.approx(outs["0"][2], rel=1e-6)
    assert orc.rel_l2(outs["4"][1], outs["0"][1]) < OP_TOL
    ref = orc.sirt(prj, vol, xc, yc, p, orc.bpt(prj, vol, xc, yc, p), 1, 0.5)
    assert orc.rel_l2(outs["4"][1], ref) < SOLVE_TOL


def test_column_fp_row_window_on_a_wide_detector(mtr, orc):
    # a 1000-column detector: whole rows would leave fewer than six blocks
    # per CU, so each wave keeps a window of its rows that the strip walk
    # writes out as it moves; against the oracle, and against whole rows
    pb = problem(mtr, [45.0], X=600, Y=600, Z=48, views=7, seed=27, nu=1000, nv=1000)
    prj, vol, xc, yc, p = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], pb.ps[0]
    say("column FP window")
    outs = []
    old = os.environ.get("MTR_FP_COL_WINDOW")
    try:
        for w in ("1", "0"):
            os.environ["MTR_FP_COL_WINDOW"] = w
            with load_ctx(mtr, pb) as c:
                assert c.agent_kernels(0)[0] == mtr.FPK_COLUMN
                win = c.agent_fp_window(0)
                assert (win > 0) == (w == "1") and win < prj.X + 2
                c.bpt(0, out=False)
                outs.append(c.sirt(0, 2, 0.5))
    finally:
        if old is None:
            os.environ.pop("MTR_FP_COL_WINDOW", None)
        else:
            os.environ["MTR_FP_COL_WINDOW"] = old
    ref = orc.sirt(prj, vol, xc, yc, p, orc.bpt(prj, vol, xc, yc, p), 2, 0.5)
    assert orc.rel_l2(outs[0], ref) < SOLVE_TOL
    assert orc.rel_l2(outs[1], ref) < SOLVE_TOL
    assert orc.rel_l2(outs[0], outs[1]) < 1e-6


# deep columns on a 1000-column detector: the row window together with
# lanes owning 4 and 2 planes, grouped chunk reads and the partial last
# chunk, for in-plane axes whose u rises and falls along the memory order
# (the strip walk runs backwards for the latter)
WINDOW_DEEP = [
    ("col30_z300", dict(phis=[30.0], X=40, Y=36, Z=300, views=13)),
    ("col135_z400", dict(phis=[135.0], X=20, Y=16, Z=400, views=13)),
    ("col60_z520", dict(phis=[60.0], X=16, Y=20, Z=520, views=13)),
]


@pytest.mark.parametrize("name,kw", WINDOW_DEEP, ids=[g[0] for g in WINDOW_DEEP])
def test_column_fp_row_window_on_deep_columns(mtr, orc, name, kw):
    pb = problem(mtr, seed=29, nu=1000, **kw)
    prj, vol, xc, yc, p = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], pb.ps[0]
    say("column FP window, deep", name)
    outs = []
    old = os.environ.get("MTR_FP_COL_WINDOW")
    try:
        for w in ("1", "0"):
            os.environ["MTR_FP_COL_WINDOW"] = w
            with load_ctx(mtr, pb) as c:
                assert c.agent_kernels(0)[0] == mtr.FPK_COLUMN
                win = c.agent_fp_window(0)
                assert (win > 0) == (w == "1") and win < prj.X + 2
                c.bpt(0, out=False)
                outs.append(c.sirt(0, 2, 0.5))
    finally:
        if old is None:
            os.environ.pop("MTR_FP_COL_WINDOW", None)
        else:
            os.environ["MTR_FP_COL_WINDOW"] = old
    ref = orc.sirt(prj, vol, xc, yc, p, orc.bpt(prj, vol, xc, yc, p), 2, 0.5)
    assert orc.rel_l2(outs[0], ref) < SOLVE_TOL
    assert orc.rel_l2(outs[1], ref) < SOLVE_TOL
    assert orc.rel_l2(outs[0], outs[1]) < 1e-6


TINY = [
    ("tiny_slab0", dict(phis=[0.0], X=5, Y=3, Z=2, nu=4, nv=3, views=3)),
    ("tiny_slab90", dict(phis=[90.0], X=3, Y=5, Z=2, nu=4, nv=3, views=3)),
    ("tiny_axisZ", dict(phis=[0.0], kind=1, X=3, Y=4, Z=2, nu=3, nv=3, views=3)),
    ("tiny_col45", dict(phis=[45.0], X=4, Y=5, Z=3, nu=6, nv=6, views=3)),
    ("one_voxel", dict(phis=[45.0], X=1, Y=1, Z=1, nu=2, nv=2, views=2)),
    ("one_plane", dict(phis=[0.0], X=7, Y=6, Z=1, nu=8, nv=7, views=5)),
]


@pytest.mark.parametrize("name,kw", TINY, ids=[g[0] for g in TINY])
def test_tiny_problems_match_oracle(mtr, orc, name, kw):
    # a handful of voxels and detector pixels: every tile, strip and chunk is
    # partial, most views' footprints are cut by the detector's edges
    pb = problem(mtr, **kw)
    prj, vol, xc, yc, p = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], pb.ps[0]
    say("tiny", name)
    with load_ctx(mtr, pb) as c:
        say("kernels", c.agent_kernels(0))
        xb = c.bpt(0)
        x_ref = orc.bpt(prj, vol, xc, yc, p)
        assert orc.rel_l2(xb, x_ref) < OP_TOL
        x2 = c.sirt(0, 2, 0.7)
        assert orc.rel_l2(x2, orc.sirt(prj, vol, xc, yc, p, x_ref, 2, 0.7)) < SOLVE_TOL
    pb2 = problem(mtr, **dict(kw, phis=[kw["phis"][0], kw["phis"][0] + 90.0]))
    with load_ctx(mtr, pb2) as c:
        out = c.mace_solve(2, 2, 0.5, 0.5)
    ref = orc.mace(pb2.prjs, pb2.vol, pb2.xcs, pb2.ycs, pb2.ps, 2, 2, 0.5, 0.5)
    assert orc.rel_l2(out, ref) < SOLVE_TOL
