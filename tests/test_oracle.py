"""Pins the CPU oracle (oracle/mt_oracle.c) without the reference.

The reference ships no tests, fixtures or data and cannot be built here (it
needs CUDA and MPI), so each oracle function is checked against what the
mathematics fixes: hand-derived golden vectors (tests/golden/), closed
forms, an independent dense tent-function matrix built in long double,
adjointness, constant fields, fixed points, affinity and the MACE
reductions.  Each check is chosen so that a dropped term, a wrong sign or
index, or a transposed operand fails it.
"""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    with open(os.path.join(GOLD, name + ".json")) as f:
        return json.load(f)


def box(mtr, X, Y, Z, xs=0, ys=0, zs=0):
    return mtr.volume_size(X, Y, Z, xs, ys, zs)


def random_affine(rng, A, nu, nv, vol, jitter=0.3):
    """General (non-separable) affine views roughly centred on the detector."""
    xc = np.empty(4 * A)
    yc = np.empty(4 * A)
    cx = (vol.Xstart + vol.Xend - 1) / 2
    cy = (vol.Ystart + vol.Yend - 1) / 2
    cz = (vol.Zstart + vol.Zend - 1) / 2
    for a in range(A):
        gu = rng.normal(size=3) * jitter + np.array([0.8, 0.3, 0.5])
        gv = rng.normal(size=3) * jitter + np.array([-0.3, 0.9, 0.1])
        xc[4 * a + 1: 4 * a + 4] = gu
        yc[4 * a + 1: 4 * a + 4] = gv
        xc[4 * a] = (nu - 1) / 2 + 1 - gu @ [cx, cy, cz] + rng.uniform(-0.5, 0.5)
        yc[4 * a] = (nv - 1) / 2 + 1 - gv @ [cx, cy, cz] + rng.uniform(-0.5, 0.5)
    return xc, yc


def dense_matrix(prj, vol, xc, yc):
    """A[(a,k,i), j] = T(u_j - i) T(v_j - k), T(t) = max(0, 1 - |t|), in long double."""
    ld = np.longdouble
    Z, Y, X = np.meshgrid(np.arange(vol.Zstart, vol.Zend), np.arange(vol.Ystart, vol.Yend),
                          np.arange(vol.Xstart, vol.Xend), indexing="ij")
    X, Y, Z = (q.ravel().astype(ld) for q in (X, Y, Z))
    rows = []
    ii = np.arange(prj.X, dtype=ld)
    kk = np.arange(prj.Y, dtype=ld)
    for a in range(prj.AngN):
        c = np.asarray(xc[4 * a: 4 * a + 4], dtype=ld)
        d = np.asarray(yc[4 * a: 4 * a + 4], dtype=ld)
        u = c[0] + c[1] * X + c[2] * Y + c[3] * Z - 1
        v = d[0] + d[1] * X + d[2] * Y + d[3] * Z - 1
        tu = np.maximum(0, 1 - np.abs(u[None, :] - ii[:, None]))  # [nu, N]
        tv = np.maximum(0, 1 - np.abs(v[None, :] - kk[:, None]))  # [nv, N]
        rows.append((tv[:, None, :] * tu[None, :, :]).reshape(prj.Y * prj.X, -1))
    return np.concatenate(rows, axis=0)


# ---------------------------------------------------------------- geometry
def test_weight_golden(orc):
    g = gold("geometry")
    for case in g["cases"]:
        i0, k0, du, dv = orc.weight(*case["XYZ"], 0, g["xc"], g["yc"])
        assert (i0, k0) == (case["x_min"], case["y_min"])
        assert du == case["du"] and dv == case["dv"]


def test_weight_integer_aligned_closed_form(orc):
    # u = X + 3, v = Z - 2: integral everywhere -> du = dv = 0 exactly.
    xc = [4.0, 1.0, 0.0, 0.0]
    yc = [-1.0, 0.0, 0.0, 1.0]
    for X, Y, Z in [(0, 0, 0), (5, -7, 3), (-9, 2, -4)]:
        i0, k0, du, dv = orc.weight(X, Y, Z, 0, xc, yc)
        assert (i0, k0, du, dv) == (X + 3, Z - 2, 0.0, 0.0)


def test_weight_brute_force_long_double(orc):
    rng = np.random.default_rng(7)
    for _ in range(300):
        xc = rng.uniform(-3, 3, 4)
        yc = rng.uniform(-3, 3, 4)
        X, Y, Z = (int(q) for q in rng.integers(-40, 40, 3))
        i0, k0, du, dv = orc.weight(X, Y, Z, 0, xc, yc)
        ld = np.longdouble
        u = ld(xc[0]) + ld(xc[1]) * X + ld(xc[2]) * Y + ld(xc[3]) * Z - 1
        v = ld(yc[0]) + ld(yc[1]) * X + ld(yc[2]) * Y + ld(yc[3]) * Z - 1
        if abs(u - np.round(u)) > 1e-9:
            assert i0 == int(np.floor(u))
            assert abs(du - float(u - np.floor(u))) < 1e-12
        if abs(v - np.round(v)) > 1e-9:
            assert k0 == int(np.floor(v))
            assert abs(dv - float(v - np.floor(v))) < 1e-12


# ------------------------------------------------------- operators (tiny)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fp_bp_match_dense_tent_matrix(mtr, orc, seed):
    rng = np.random.default_rng(seed)
    vol = box(mtr, 5, 4, 3, xs=seed, ys=-1, zs=2)
    prj = mtr.projection_size(7, 5, 3)
    xc, yc = random_affine(rng, prj.AngN, prj.X, prj.Y, vol)
    Am = dense_matrix(prj, vol, xc, yc)
    x = rng.uniform(0.1, 2.0, (vol.Z, vol.Y, vol.X))
    r = rng.normal(size=(prj.AngN, prj.Y, prj.X))
    s_ref = (Am @ x.ravel().astype(np.longdouble)).astype(np.float64)
    c_ref = Am.sum(axis=1).astype(np.float64)
    # f64 mode to rounding; ref mode (FP32 contributions and sums) to 1e-6.
    s64, c64 = orc.fp_f64(prj, vol, xc, yc, x)
    np.testing.assert_allclose(s64.ravel(), s_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(c64.ravel(), c_ref, rtol=1e-12, atol=1e-12)
    s32, c32 = orc.fp(prj, vol, xc, yc, x.astype(np.float32))
    np.testing.assert_allclose(s32.ravel(), s_ref, rtol=2e-6, atol=2e-6)
    np.testing.assert_allclose(c32.ravel(), c_ref, rtol=2e-6, atol=2e-6)
    # un-normalised back-projection = A^T r, normaliser = A^T 1
    S, C = orc.bp_sums_f64(prj, vol, xc, yc, r)
    np.testing.assert_allclose(S.ravel(), (Am.T @ r.ravel().astype(np.longdouble)).astype(np.float64), rtol=1e-12,
                               atol=1e-12)
    np.testing.assert_allclose(C.ravel(), Am.sum(axis=0).astype(np.float64), rtol=1e-12, atol=1e-12)
    # normalised BP update and BPT
    Cd = Am.sum(axis=0).astype(np.float64)
    Sd = (Am.T @ r.ravel().astype(np.longdouble)).astype(np.float64)
    upd = np.where(Cd != 0, Sd / np.where(Cd != 0, Cd, 1), 0.0)
    xb = orc.bp(prj, vol, xc, yc, r.astype(np.float32), x.astype(np.float32), 0.5)
    np.testing.assert_allclose(xb.ravel(), x.astype(np.float32).ravel() + 0.5 * upd, rtol=1e-6, atol=1e-6)
    pb = np.abs(r).astype(np.float32)
    Sp = (Am.T @ pb.ravel().astype(np.longdouble)).astype(np.float64)
    xt = orc.bpt(prj, vol, xc, yc, pb)
    np.testing.assert_allclose(xt.ravel(), np.where(Cd != 0, Sp / np.where(Cd != 0, Cd, 1), 0.0), rtol=1e-6,
                               atol=1e-6)


def test_adjointness_f64(mtr, orc):
    rng = np.random.default_rng(11)
    vol = box(mtr, 9, 8, 6, xs=-3, ys=2, zs=1)
    prj = mtr.projection_size(10, 9, 4)
    xc, yc = random_affine(rng, prj.AngN, prj.X, prj.Y, vol)
    x = rng.normal(size=(vol.Z, vol.Y, vol.X))
    y = rng.normal(size=(prj.AngN, prj.Y, prj.X))
    s, _ = orc.fp_f64(prj, vol, xc, yc, x)
    S, _ = orc.bp_sums_f64(prj, vol, xc, yc, y)
    lhs, rhs = float((s * y).sum()), float((x * S).sum())
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_constant_fields(mtr, orc):
    rng = np.random.default_rng(3)
    vol = box(mtr, 12, 10, 7)
    prj = mtr.projection_size(11, 9, 5)
    xc, yc = random_affine(rng, prj.AngN, prj.X, prj.Y, vol)
    s, c = orc.fp(prj, vol, xc, yc, np.ones((vol.Z, vol.Y, vol.X), np.float32))
    assert np.array_equal(s, c)  # fl32(w * 1) == fl32(w) term by term
    assert (c > 0).any()
    # r == 1 -> S == C -> x += lambda exactly where C != 0
    x0 = rng.uniform(-1, 1, (vol.Z, vol.Y, vol.X)).astype(np.float32)
    x1 = orc.bp(prj, vol, xc, yc, np.ones((prj.AngN, prj.Y, prj.X), np.float32), x0, 0.25)
    _, C = orc.bp_sums_f64(prj, vol, xc, yc, np.ones((prj.AngN, prj.Y, prj.X)))
    expect = np.where(C != 0, (x0.astype(np.float64) + 0.25).astype(np.float32), x0)
    assert np.array_equal(x1, expect)


def test_integer_shift_is_mean_along_ray(mtr, orc):
    # theta = 0 slab geometry with detector = XY: u = X, v = Y, every voxel
    # lands on one pixel with weight 1; s = sum over Z, c = Z, s/c = mean.
    vol = box(mtr, 8, 6, 5)
    prj = mtr.projection_size(8, 6, 1)
    xc, yc = mtr.synth.geometry(vol, prj, mtr.synth.GEOM_SLAB, 0.0, [0.0])
    rng = np.random.default_rng(5)
    x = rng.integers(-8, 9, (vol.Z, vol.Y, vol.X)).astype(np.float32)
    s, c = orc.fp(prj, vol, xc, yc, x)
    assert np.array_equal(c[0], np.full((6, 8), 5.0, np.float32))
    assert np.array_equal(s[0], x.sum(axis=0))
    r = orc.residual(np.zeros_like(s), s, c)
    np.testing.assert_allclose(r[0], -x.mean(axis=0), rtol=1e-6, atol=1e-6)


def test_residual_known_answer(orc):
    p = np.array([3.0, 4.0, 5.0, 7.0], np.float32)
    s = np.array([1.0, 4.0, 3.0, 2.0], np.float32)
    c = np.array([0.5, 1.0, 0.5, 0.0], np.float32)  # c == 0 keeps s as is
    r = orc.residual(p, s, c)
    assert r.tolist() == [1.0, 0.0, -1.0, 5.0]


def test_line_golden(mtr, orc):
    g = gold("line")
    vol = box(mtr, *g["vol"])
    prj = mtr.projection_size(*g["prj"])
    x = np.array(g["x"], np.float32).reshape(1, 1, 2)
    s, c = orc.fp(prj, vol, g["xc"], g["yc"], x)
    assert s.ravel().tolist() == g["s"] and c.ravel().tolist() == g["c"]
    r = orc.residual(np.array(g["p"], np.float32).reshape(s.shape), s, c)
    assert r.ravel().tolist() == g["r"]
    xb = orc.bp(prj, vol, g["xc"], g["yc"], r, x, g["lam"])
    assert xb.ravel().tolist() == g["x_after_sirt1"]
    xs = orc.sirt(prj, vol, g["xc"], g["yc"], np.array(g["p"], np.float32).reshape(s.shape), x, 1, g["lam"])
    assert xs.ravel().tolist() == g["x_after_sirt1"]
    xt = orc.bpt(prj, vol, g["xc"], g["yc"], np.array(g["p"], np.float32).reshape(s.shape))
    assert xt.ravel().tolist() == g["bpt"]


# --------------------------------------------------------------- SIRT
def small_problem(mtr, phis=(0.0,), kind=None, X=16, Y=14, Z=10, seed=4, views=7):
    kind = mtr.synth.GEOM_SLAB if kind is None else kind
    return mtr.synth.make_problem("t", X, Y, Z, 6, list(phis), kind=kind, seed=seed, amin=2, amax=5,
                                  theta=np.linspace(-60, 60, views))


def test_sirt_fixed_point_f64(mtr, orc):
    pb = small_problem(mtr, phis=(30.0,))
    prj, vol, xc, yc = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0]
    rng = np.random.default_rng(9)
    xs = rng.uniform(0, 1, (vol.Z, vol.Y, vol.X))
    s, c = orc.fp_f64(prj, vol, xc, yc, xs)
    p = np.where(c != 0, s / np.where(c != 0, c, 1), 0.0)
    # p must be representable: only exact FP32 values keep r == 0 in f64.
    xs32 = xs.astype(np.float32).astype(np.float64)
    s, c = orc.fp_f64(prj, vol, xc, yc, xs32)
    p = np.where(c != 0, s / np.where(c != 0, c, 1), 0.0)
    p32 = p.astype(np.float32)
    out = orc.sirt_f64(prj, vol, xc, yc, p32, xs32, 3, 1.0)
    # residual is p32 - p (rounding only) -> the update is tiny, not zero
    assert np.max(np.abs(out - xs32)) < 1e-6
    # with the exact p the residual vanishes identically in f64 FP: use a
    # geometry where s/c is exactly representable (integer shift).
    vol2 = box(mtr, 6, 5, 4)
    prj2 = mtr.projection_size(6, 5, 1)
    xc2, yc2 = mtr.synth.geometry(vol2, prj2, mtr.synth.GEOM_SLAB, 0.0, [0.0])
    x2 = np.tile(rng.integers(-4, 5, (1, 5, 6)).astype(np.float64), (4, 1, 1))  # constant along rays
    s2, c2 = orc.fp_f64(prj2, vol2, xc2, yc2, x2)
    p2 = (s2 / c2).astype(np.float32)
    out2 = orc.sirt_f64(prj2, vol2, xc2, yc2, p2, x2, 4, 0.7)
    assert np.array_equal(out2, x2)


def test_sirt_affine_in_x0_f64(mtr, orc):
    pb = small_problem(mtr, phis=(45.0,))
    prj, vol, xc, yc, p = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], pb.ps[0]
    rng = np.random.default_rng(2)
    x, y = rng.normal(size=(2, vol.Z, vol.Y, vol.X))
    al = 0.3
    fx = orc.sirt_f64(prj, vol, xc, yc, p, x, 4, 0.6)
    fy = orc.sirt_f64(prj, vol, xc, yc, p, y, 4, 0.6)
    fxy = orc.sirt_f64(prj, vol, xc, yc, p, al * x + (1 - al) * y, 4, 0.6)
    np.testing.assert_allclose(fxy, al * fx + (1 - al) * fy, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("lam,n", [(1.0, 1), (0.5, 3), (0.25, 5)])
def test_sirt_one_to_one_geometry_closed_form(mtr, orc, lam, n):
    # u = X, v = Y, one view, Z = 1: s = x, c = 1, S = r, C = 1, so
    # SIRT^n(x0) = p + (1 - lam)^n (x0 - p) for dyadic inputs exactly.
    vol = box(mtr, 4, 3, 1)
    prj = mtr.projection_size(4, 3, 1)
    xc, yc = [1.0, 1.0, 0.0, 0.0], [1.0, 0.0, 1.0, 0.0]
    p = np.arange(12, dtype=np.float32).reshape(1, 3, 4)
    x0 = np.full((1, 3, 4), 8.0, np.float32)
    out = orc.sirt(prj, vol, xc, yc, p, x0, n, lam)
    np.testing.assert_array_equal(out.reshape(3, 4), (p[0] + (1 - lam) ** n * (8.0 - p[0])).astype(np.float32))


# --------------------------------------------------------------- MACE
def test_mace_golden_one_voxel(mtr, orc):
    g = gold("mace")
    vol = box(mtr, 1, 1, 1)
    prj = mtr.projection_size(1, 1, 1)
    xc, yc = [1.0, 1.0, 0.0, 0.0], [1.0, 0.0, 1.0, 0.0]
    ps = [np.array([[[q]]], np.float32) for q in g["p"]]
    for case in g["cases"]:
        for M, expect in enumerate(case["wbar"]):
            out = orc.mace([prj, prj], vol, [xc, xc], [yc, yc], ps, M, g["S"], case["lam"], g["rho"])
            assert float(out.ravel()[0]) == expect


def test_mace_single_agent_reduces_to_sirt(mtr, orc):
    pb = small_problem(mtr, phis=(0.0,))
    prj, vol, xc, yc, p = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], pb.ps[0]
    M, S = 3, 2
    w = orc.mace_f64([prj], vol, [xc], [yc], [p], M, S, 0.5, 0.5)
    x0 = orc.bpt_f64(prj, vol, xc, yc, p)
    ref = orc.sirt_f64(prj, vol, xc, yc, p, x0, M * S, 0.5)
    np.testing.assert_allclose(w, ref, rtol=1e-12, atol=1e-12)
    w32 = orc.mace([prj], vol, [xc], [yc], [p], M, S, 0.5, 0.5)
    assert orc.rel_l2(w32, ref) < 1e-6


def test_mace_rho_zero_keeps_initial_model(mtr, orc):
    pb = small_problem(mtr, phis=(0.0, 90.0))
    B = sum(orc.bpt(pb.prjs[a], pb.vol, pb.xcs[a], pb.ycs[a], pb.ps[a]) for a in range(2))
    out = orc.mace(pb.prjs, pb.vol, pb.xcs, pb.ycs, pb.ps, 3, 2, 0.5, 0.0)
    np.testing.assert_array_equal(out, B.astype(np.float32))


def test_mace_identical_agents(mtr, orc):
    # K identical agents stay identical: w = wbar, z = wbar and with rho = 0.5
    # w' = v, so wbar_M = SIRT^(S*M)(K * BPT).
    pb = small_problem(mtr, phis=(60.0,))
    prj, vol, xc, yc, p = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], pb.ps[0]
    w = orc.mace_f64([prj, prj], vol, [xc, xc], [yc, yc], [p, p], 2, 3, 0.5, 0.5)
    x0 = 2 * orc.bpt_f64(prj, vol, xc, yc, p)
    ref = orc.sirt_f64(prj, vol, xc, yc, p, x0, 6, 0.5)
    np.testing.assert_allclose(w, ref, rtol=1e-12, atol=1e-12)


def test_elementwise_known_answers(orc):
    w = np.array([1.0, 2.0, -3.0], np.float32)
    wb = np.array([0.5, 4.0, 1.0], np.float32)
    assert orc.form_z(w, wb).tolist() == [0.0, 6.0, 5.0]
    v = np.array([2.0, 1.0, 0.0], np.float32)
    z = np.array([1.0, 1.0, 1.0], np.float32)
    # 0.5*(2v - z) + 0.5*w
    assert orc.mann(v, w, z, 0.5).tolist() == [2.0, 1.5, -2.0]
    assert orc.mean([w, wb]).tolist() == [0.75, 3.0, -1.0]


def test_divide_slice_golden_and_invariants(orc, mtr):
    g = gold("slices")
    assert [list(orc.divide_slice(i, g["N"], g["Z"], g["S"])) for i in range(g["N"])] == g["parts"]
    for Z in range(0, 40):
        for N in range(1, 9):
            parts = [orc.divide_slice(i, N, Z, 3) for i in range(N)]
            assert parts[0][0] == 3 and parts[-1][1] == 3 + Z
            assert all(parts[i][1] == parts[i + 1][0] for i in range(N - 1))
            h = [e - s for s, e in parts]
            assert max(h) - min(h) <= 1 and h == sorted(h, reverse=True)
            assert parts == [mtr.divide_slice(i, N, Z, 3) for i in range(N)]


def test_ref_mode_tracks_f64_mode(mtr, orc):
    pb = small_problem(mtr, phis=(0.0, 45.0), X=20, Y=18, Z=8)
    w32 = orc.mace(pb.prjs, pb.vol, pb.xcs, pb.ycs, pb.ps, 2, 3, 0.5, 0.5)
    w64 = orc.mace_f64(pb.prjs, pb.vol, pb.xcs, pb.ycs, pb.ps, 2, 3, 0.5, 0.5)
    assert orc.rel_l2(w32, w64) < 1e-6


def test_slab_fp_sums_to_full(mtr, orc):
    # the reference splits Z into slabs and sums the partial projections
    # (FunctionOnGPU.cu:225-226): slab partials must add up to the full FP.
    pb = small_problem(mtr, phis=(30.0,))
    prj, vol, xc, yc = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0]
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 1, (vol.Z, vol.Y, vol.X)).astype(np.float32)
    s, c = orc.fp(prj, vol, xc, yc, x)
    acc_s = np.zeros_like(s, dtype=np.float64)
    acc_c = np.zeros_like(c, dtype=np.float64)
    for i in range(3):
        zs, ze = orc.divide_slice(i, 3, vol.Z, vol.Zstart)
        ss, cc = orc.fp_slab(prj, vol, zs, ze, xc, yc, x[zs - vol.Zstart: ze - vol.Zstart])
        acc_s += ss
        acc_c += cc
    np.testing.assert_allclose(acc_s, s, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(acc_c, c, rtol=1e-5, atol=1e-5)


def test_oracle_runs_clean_under_asan_and_ubsan():
    """Every oracle function on tiny ragged problems (offset boxes,
    detectors narrower than the box, tilted and general affine views)
    built with AddressSanitizer and UndefinedBehaviorSanitizer: no
    out-of-bounds access, overflow or other undefined behaviour (the check
    aborts on the first report)."""
    import subprocess

    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    b = subprocess.run(["make", "-s", "-C", here, "sanitize"], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stdout + b.stderr
    r = subprocess.run([os.path.join(here, "_san", "sanitize_check")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr
    assert "12 of 12 cases clean" in r.stdout
