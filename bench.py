"""Benchmark: bead-steps/s of the full per-timestep hydrodynamic solve on one MI355X.

Workload (default): 20,000 beads = 200 probes (radius 5a) + 990 chains x 20 beads (radius a),
periodic box 320a, harmonic bonds, full monomer-monomer HI, kT = 1 (reference units: beta = 0.2,
lengths x 0.2).  A step is one iteration of PolyStokes::run: predictor solve, Brownian slip by the
Lanczos square root, the two RFD probe solves, the corrector solve and the integrator.  Inputs are
synthetic (seeded), resident in HBM before timing.

The headline value is the mixed-precision form (FP32 tiles, FP64 MINRES/Lanczos recurrences and
FP64 iterative refinement); the plain FP64 form is timed beside it, and both are checked against an
FP64 run at tight tolerances (relative velocity error).  The roofline entry is for the dominant
kernel (the monomer-monomer tile kernel), timed live with HIP events on the library's stream.  The
CPU baseline is the FP64 oracle on the host cores, timed on a bounded sample (a few matrix-free
applies) and scaled by the applies per step the FP64 form performs.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_ctx(Context, pd, x0, **over):
    d = dict(pd)
    d.update(over)
    c = Context(d)
    c.initial_configuration(x0)
    return c


def timed_run(ctx, steps, warmup):
    if warmup:
        ctx.run(warmup)
    c0 = ctx.counters()
    ctx.state()   # synchronise
    ctx.profile_begin()
    t0 = time.perf_counter()
    ctx.run(steps)
    ctx.state()
    dt = time.perf_counter() - t0
    prof = ctx.profile_end()
    c1 = ctx.counters()
    per = {k: (c1[k] - c0[k]) / steps for k in ("applies", "minres_iterations", "lanczos_iterations",
                                                 "saddle_solves", "mob_calls")}
    return dt, prof, per


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-applies", type=int, default=3)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--solver", default="minres", choices=["minres", "block"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the solve is not sharded across GPUs",
              file=sys.stderr)
        sys.exit(2)

    from polystokes_brownian_amd import configs
    from polystokes_brownian_amd.polystokes import Context

    w = configs.workload(args.workload)
    pd = configs.params_dict(w)
    pd["solver"] = 1 if args.solver == "block" else 0
    t0 = time.perf_counter()
    x0 = configs.positions(w)
    log(f"workload {args.workload}: Np={w['Np']} (Nc={w['Nc']}, Nm={w['Nm']}), configuration {time.perf_counter() - t0:.1f}s")
    Np, Nm = w["Np"], w["Nm"]

    # ---- accuracy reference: one FP64 step at tight tolerances
    ref = make_ctx(Context, pd, x0, ksp_rtol=1e-10, sqrt_rtol=pd["sqrt_rtol"], solver=0)
    name, ncu = ref.device_info()
    log(f"device {name} ({ncu} CUs)")
    t0 = time.perf_counter()
    ref.run(1)
    up_ref = ref.vectors()["up"]
    log(f"reference FP64 step (rtol 1e-10): {time.perf_counter() - t0:.2f}s")
    ref.close()

    results = {}
    forms = [("mixed", 1)] + ([] if args.skip_fp64 else [("fp64", 0)])
    for form, prec in forms:
        ctx = make_ctx(Context, pd, x0, precision=prec, ir_rtol=1e-7 if prec else pd["ksp_rtol"])
        ctx.run(1)
        err = float(np.linalg.norm(ctx.vectors()["up"] - up_ref) / np.linalg.norm(up_ref))
        ctx.close()
        ctx = make_ctx(Context, pd, x0, precision=prec, ir_rtol=1e-7 if prec else pd["ksp_rtol"])
        dt, prof, per = timed_run(ctx, args.steps, args.warmup)
        ctx.close()
        bs = Np * args.steps / dt
        aa_avg = prof["aa_ms"] / max(prof["aa_launches"], 1) * 1e-3
        flops = 18.0 * Nm * Nm   # 2 * nnz of the monomer-monomer block (3Nm)^2
        results[form] = dict(value=bs, ms_per_step=1e3 * dt / args.steps, vel_rel_err=err, per_step=per,
                             aa_ms_avg=aa_avg * 1e3, aa_launches=prof["aa_launches"],
                             apply_ms_avg=prof["apply_ms"] / max(prof["apply_launches"], 1),
                             achieved=flops / aa_avg if aa_avg > 0 else 0.0)
        log(f"{form}: {bs:.4g} bead-steps/s, {1e3 * dt / args.steps:.1f} ms/step, vel rel err {err:.2e}, "
            f"per step {per}, k_aa avg {aa_avg * 1e3:.3f} ms, apply avg {results[form]['apply_ms_avg']:.3f} ms")

    cpu = None
    if not args.no_cpu:
        from oracle.pyoracle import Oracle, get_threads
        o = Oracle(pd, dense=False)
        o.initial_configuration(x0)
        o.check_dist()
        o.mob()
        f = np.random.default_rng(0).normal(size=o.n)
        o.mob_mult(f)
        t0 = time.perf_counter()
        for _ in range(args.cpu_applies):
            o.mob_mult(f)
        t_apply = (time.perf_counter() - t0) / args.cpu_applies
        applies = results.get("fp64", results["mixed"])["per_step"]["applies"]
        t_step = t_apply * applies
        cpu = {"value": Np / t_step, "unit": "bead-steps/s", "cores": get_threads(), "kind": "port",
               "sample": f"{args.cpu_applies} FP64 matrix-free mobility applies of the oracle at N={Np} "
                         f"({t_apply:.3f} s each), scaled by the {applies:.0f} applies per step of the FP64 form"}
        log(f"cpu baseline {cpu}")

    head = results["mixed"]
    fp64 = results.get("fp64")
    roof = {"bound": "mfma", "achieved": head["achieved"] / 1e12, "peak": PEAK["fp32"] / 1e12, "unit": "TFLOP/s",
            "frac": head["achieved"] / PEAK["fp32"], "traffic": None,
            "kernel": "k_aa<float> (monomer-monomer tiles, VALU-bound)",
            "flops_per_launch": 18.0 * Nm * Nm, "avg_launch_ms": head["aa_ms_avg"]}
    out = {
        "metric": "bead-steps/s at N=20k: mixed-prec vs FP64 GPU at 1e-6 rel err; % roofline",
        "value": head["value"], "unit": "bead-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {Np} beads ({w['Nc']} probes + {w['Npoly']} chains x 20), "
                               f"box {w['L']:.1f}, full mm-HI, saddle solver {args.solver} (rtol {pd['ksp_rtol']}), "
                               f"Lanczos m<={pd['sqrt_kmax']}",
                   "global_batch": Np, "seq_len": 1, "parallelism": "single-gpu"},
        "mixed_vel_rel_err": head["vel_rel_err"],
        "fp64_value": fp64["value"] if fp64 else None,
        "fp64_vel_rel_err": fp64["vel_rel_err"] if fp64 else None,
        "mixed_over_fp64": head["value"] / fp64["value"] if fp64 else None,
        "per_step": head["per_step"],
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if fp64:
        out["roofline_fp64"] = {"achieved": fp64["achieved"] / 1e12, "peak": PEAK["fp64"] / 1e12,
                                "frac": fp64["achieved"] / PEAK["fp64"], "avg_launch_ms": fp64["aa_ms_avg"]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
