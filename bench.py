#!/usr/bin/env python3
"""Benchmark: MACE reconstructions per second on one MI355X.

A step is one whole reconstruction of the dual-axis MACE workload
(512x512x128 volume, two tilt series at in-plane axes 0 and 90 degrees,
61 views each over +-60 degrees on a 512x512 detector, 1% noise, 10 MACE
steps of 5 SIRT iterations, lambda 0.5, rho 0.5): BPT of both series, the
summed initial model, then 10 x (2 x 5 SIRT iterations + fused consensus).
Inputs are synthetic (seeded ellipsoid phantom, analytic ray sums) and are
resident in HBM before the timed region starts.

Prints one JSON line.  Per-kernel times come from HIP events recorded on
the library's own stream around every launch inside the timed region.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
HBM_PEAK_GBS = 8000.0
# FP32 vector lanes: 256 CU x 64 lanes x 2.4 GHz (non-packed issue).
LANE_OPS_PEAK = 256 * 64 * 2.4e9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(pb, threads: int):
    """Times the CPU oracle on a bounded sample and scales to recon/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # the CPU restatement (timed baseline only)

    used = orc.set_threads(threads) if threads > 0 else orc.set_threads(0)
    prj, vol, xc, yc, p = pb.prjs[0], pb.vol, pb.xcs[0], pb.ycs[0], pb.ps[0]
    t0 = time.perf_counter()
    x = orc.bpt(prj, vol, xc, yc, p)
    t_bpt = time.perf_counter() - t0
    t0 = time.perf_counter()
    xs = orc.sirt(prj, vol, xc, yc, p, x, 1, pb.lam)
    t_sirt = time.perf_counter() - t0
    w = x.copy()
    t0 = time.perf_counter()
    z = orc.form_z(w, x)
    w = orc.mann(xs, w, z, pb.rho)
    wb = orc.mean([w, x])
    t_cons = time.perf_counter() - t0
    K, M, S = pb.K, pb.M, pb.S
    # consensus per step: K x (z + Mann) + mean; the sample above did one of each
    t_recon = K * t_bpt + M * K * S * t_sirt + M * K * t_cons
    del wb
    return {
        "value": 1.0 / t_recon,
        "unit": "reconstructions/s",
        "cores": used,
        "kind": "port",
        "sample": (f"oracle BPT + 1 SIRT iteration + 1 consensus of agent 0 at full size "
                   f"({t_bpt:.2f}+{t_sirt:.2f}+{t_cons:.2f} s), scaled to {K} BPT + {M*K*S} SIRT iterations "
                   f"+ {M} consensus steps"),
        "seconds_sample": t_bpt + t_sirt + t_cons,
    }, (x, xs)


def load_traffic():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4"])
    ap.add_argument("--plain-steps", type=int, default=1, help="plain per-agent loop reconstructions to time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp16"])
    ap.add_argument("--calibrate-copy-mb", type=int, default=256,
                    help="size of the streaming copy run once for traffic-counter calibration (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); multi-GPU agent sharding is a "
              "design, not a build, here.", file=sys.stderr)
        return 2

    mtr = importlib.import_module("multi-tilt-reconstuction_amd")
    t0 = time.perf_counter()
    pb = mtr.synth.config(args.config, threads=16)
    log(f"generated {args.config}: {pb.vol.X}x{pb.vol.Y}x{pb.vol.Z}, K={pb.K} in {time.perf_counter()-t0:.1f}s")
    prec = {"fp32": mtr.PREC_FP32, "bf16": mtr.PREC_BF16, "fp16": mtr.PREC_FP16}[args.precision]
    N = pb.N
    Ps = pb.P
    A = pb.prjs[0].AngN

    ctx = mtr.Context(0, prec)
    copy_cal = None
    if args.calibrate_copy_mb > 0:
        nbytes = args.calibrate_copy_mb << 20
        ms = ctx.calibrate_copy(nbytes)
        copy_cal = {"bytes": nbytes, "ms": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}
    ctx.set_volume(pb.vol)
    for a in range(pb.K):
        ctx.add_agent(pb.prjs[a], pb.xcs[a], pb.ycs[a], pb.ps[a])
    ctx.sync()

    for i in range(args.warmup):
        ctx.mace_solve(pb.M, pb.S, pb.lam, pb.rho, out=False)
        ctx.sync()
        log(f"warmup {i} done")

    ctx.timing_reset()
    ctx.timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.mace_solve(pb.M, pb.S, pb.lam, pb.rho, out=False)
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.timing(False)
    log(f"timed {args.steps} reconstructions in {wall:.3f}s")

    ktimes = {name: ctx.timing_get(k) for k, name in enumerate(mtr.KERNEL_NAMES)}
    K, M, S = pb.K, pb.M, pb.S
    steps = args.steps
    # Algorithmic HBM bytes per reconstruction (FP32 storage; 2-byte storage halves it).
    es = 4 if prec == mtr.PREC_FP32 else 2
    fp_bytes = sum(M * S * es * (N + 2 * P) for P in Ps)          # read x, read p, write r
    bp_bytes = sum(M * S * es * (2 * N + P) for P in Ps)          # read r, read+write x
    bpt_bytes = sum(es * (N + P) for P in Ps)
    cons_bytes = M * es * (4 * K + 2) * N                         # v, w, wbar in; w, z, wbar out
    vvu = {"fp": M * S * K * N * A, "bp": M * S * K * N * A, "bpt": K * N * A}
    per = {}
    for name, nbytes in (("fp", fp_bytes), ("bp", bp_bytes), ("bpt", bpt_bytes), ("consensus", cons_bytes)):
        n, ms = ktimes[name]
        if n == 0 or ms <= 0:
            continue
        per[name] = {
            "launches": n,
            "avg_ms": ms / n,
            "total_ms": ms,
            "bytes_per_launch": nbytes * steps / n,
            "GBps": nbytes * steps / (ms * 1e-3) / 1e9,
        }
        if name in vvu:
            per[name]["vvu_per_s"] = vvu[name] * steps / (ms * 1e-3)
    dom = max(("fp", "bp"), key=lambda q: per.get(q, {}).get("total_ms", 0.0))
    d = per[dom]
    traffic = load_traffic()
    tr = None
    if traffic and args.config in traffic and dom in traffic[args.config]:
        tr = traffic[args.config][dom]
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": d["GBps"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": d["GBps"] / HBM_PEAK_GBS,
        "traffic": tr,
        "bytes_per_launch": d["bytes_per_launch"],
        "avg_launch_ms": d["avg_ms"],
        "vvu_per_s": d.get("vvu_per_s"),
        "note": "projector is VALU/LDS-bound at 61 views; HBM fraction is on algorithmic bytes",
    }
    if "consensus" in per:
        roofline["consensus_GBps"] = per["consensus"]["GBps"]
        roofline["consensus_frac"] = per["consensus"]["GBps"] / HBM_PEAK_GBS

    plain = None
    if args.plain_steps > 0 and prec == mtr.PREC_FP32:
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            ctx.mace_solve(M, S, pb.lam, pb.rho, form=mtr.FORM_PLAIN, out=False)
        ctx.sync()
        tp = (time.perf_counter() - t0) / args.plain_steps
        plain = {"recon_per_s": 1.0 / tp, "ms_per_recon": tp * 1e3}
        log(f"plain loop: {tp*1e3:.1f} ms per reconstruction")

    cpu = None
    parity = None
    if not args.no_cpu:
        log("timing the CPU oracle sample ...")
        cpu, (x_bpt, x_s1) = cpu_baseline(pb, args.cpu_threads)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        g_bpt = ctx.bpt(0)
        g_s1 = ctx.sirt(0, 1, pb.lam)
        parity = {"bpt_rel_l2": orc.rel_l2(g_bpt, x_bpt), "sirt1_rel_l2": orc.rel_l2(g_s1, x_s1)}
        log(f"parity sample: {parity}")
    ctx.close()

    value = steps / wall
    out = {
        "metric": "MACE reconstructions/sec (dual-axis 512x512x128, 2x61 views, 10 MACE x 5 SIRT)",
        "value": value,
        "unit": "reconstructions/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == mtr.PREC_FP32 else ("bf16" if prec == mtr.PREC_BF16 else "f16"),
        "data": "synthetic (seeded ellipsoid phantom, analytic ray sums, 1% noise)",
        "config": {
            "workload": f"{args.config}: dual-axis MACE {pb.vol.X}x{pb.vol.Y}x{pb.vol.Z}, K={K}, {A} views/axis, "
                        f"{pb.prjs[0].X}x{pb.prjs[0].Y} detector, M={M}, S={S}, lambda={pb.lam}, rho={pb.rho}",
            "parallelism": "single GPU, all agents resident",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": per,
        "plain_loop": plain,
        "speedup_vs_plain": (value / plain["recon_per_s"]) if plain else None,
        "parity_sample": parity,
        "copy_calibration": copy_cal,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
