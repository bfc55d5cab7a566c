"""Benchmark: bead-steps/s of the full per-timestep hydrodynamic solve on one MI355X.

Workload (default c3): 20,000 beads = 200 probes (radius 5a) + 990 chains x 20 beads (radius a),
periodic box 320a, harmonic bonds, full monomer-monomer HI, kT = 1, in the reference's units
(beta = 0.2, lengths x 0.2, time x beta^3).  A step is one iteration of PolyStokes::run: predictor
solve, Brownian slip by the Lanczos square root (m <= 32, rtol 1e-6), the two RFD probe solves,
the corrector solve and the integrator.  Inputs are synthetic (seeded), resident in HBM before the
timed region; the timed region is exactly `steps` full steps after `warmup` untimed ones.

Forms timed (both with exact block-eliminated saddle solves):
  mixed/block  (headline)  every mobility apply on FP32 tiles (fixed-point separations), FP64
                           Lanczos recurrences and tridiagonal square root, FP64 stresslet solves,
                           FP64 colloid pairs, no refinement sweep;
  fp64/block               the same algorithm in FP64 (the FP64 GPU form the ratio is against);
  fp64/minres (--minres)   the reference algorithm (Jacobi MINRES on the saddle system), one step.
Each form is timed launched kernel by kernel and as replayed step graphs; the faster is its
figure.  Accuracy: one step of each form from the same state and noise as one step of the FP64
oracle (velocity, displacement, bond-length distribution) and the Lanczos square root on the same xi,
relative to the oracle.  Roofline: the fused apply kernel (monomer tile pairs + monomer-colloid
blocks) of the headline form, its launches timed by their own dispatch events on the library's
stream in a separate profiled run.  CPU baseline: the oracle's step of the same algorithm, timed on
the host threads it uses, and the one-core rate extrapolated from one timed apply.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3] [--no-cpu] [--minres]
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
# FP32 fused multiply-adds at one wave64 instruction per CU per cycle (256 CUs x 64 lanes x 2 flop x
# 2.4 GHz): half the published FP32 figure, which counts two VALU instructions issued per CU per cycle
# (gfx950 issues two independent FP32 instructions in one cycle; packed FP32 is not used here).
PEAK_FP32_VALU = 78.6e12
# HBM bytes per launch of the fused apply kernel, from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE
# of `bench.py --profile-form <form> --steps 2 --warmup 0 --no-cpu --calibrate-copy 256`, each scaled by
# what the counter reads for the 256 MiB calibration copy; keyed by (workload, tiles).  None until
# measured for that kernel.  profiles/k_apply_sym_hbm_traffic_n20000.md: FP32, 9,107.5 KiB fetched and
# 19,761.0 KiB written per launch (FETCH_SIZE 0.50x, WRITE_SIZE 1.00x of the calibration copy).
TRAFFIC_APPLY = {("c3", "fp32"): (9107.5 + 19761.0) * 1024}
# VALU issue of the fused apply kernel from rocprofv3 --pmc SQ_INSTS_VALU SQ_ACTIVE_INST_VALU2 ... of
# `bench.py --profile-form mixed/block --steps 1 --warmup 0 --no-cpu` (profiles/k_apply_sym_valu_issue_n20000.md):
# instructions per launch, and the share issued two per SIMD quad-cycle (gfx950 dual issue).
VALU_ISSUE = {("c3", "fp32"): {"instructions_per_launch": 137.19e6, "dual_issued_fraction": 0.652,
                               "instructions_per_issue_quad_cycle": 1.484}}


def measured_peaks():
    """VALU issue rates measured on this GPU by the probe library (polystokes_brownian_amd/libpsprobe.so,
    csrc/peak_probe.hip): wave64 instructions per CU per shader cycle of independent chains of one
    instruction form, and the shader clock read in the kernel."""
    import ctypes as C
    path = os.path.join(ROOT, "polystokes_brownian_amd", "libpsprobe.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.psp_valu_rate.argtypes = [C.c_int] + [C.POINTER(C.c_double)] * 3
    names = {0: "v_fmac_f32", 1: "v_fma_f32", 2: "v_fmac_f32+v_mul_f32", 3: "v_fma_f64",
             4: "v_mov_b32_dpp wave_rol:1", 5: "v_rsq_f32"}
    out = {}
    for form, name in names.items():
        ms, ghz, ipc = C.c_double(), C.c_double(), C.c_double()
        rc = L.psp_valu_rate(form, C.byref(ms), C.byref(ghz), C.byref(ipc))
        if rc != 0:
            return {"error": f"psp_valu_rate({form}) returned {rc}"}
        out[name] = {"instr_per_cu_cycle": round(ipc.value, 3), "shader_ghz": round(ghz.value, 3)}
    log(f"VALU issue probe: {out}")
    return out


def ALGO_BYTES(Nm, Nc, elt):
    """Algorithmic HBM bytes of one fused apply launch: the monomer positions (4 elements each) and
    the input vector's monomer and colloid entries read once (FP64), the monomer partial sums
    written once per source tile (elt bytes)."""
    return 16 * Nm + 8 * (3 * Nm + 11 * Nc) + elt * 3 * Nm


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_ctx(Context, pd, x0, **over):
    d = dict(pd)
    d.update(over)
    c = Context(d)
    c.initial_configuration(x0)
    return c


def timed_run(ctx, steps, warmup, label, profile=True):
    """W warmup steps, then exactly `steps` steps in one run() call (as PolyStokes.run drives them:
    the noise of step k+1 is drawn on the host while the GPU works on step k).  With `profile`, the
    library times the kernels it reports (ps_profile) -- which also makes ps_run launch every kernel
    itself rather than replay a step graph."""
    if warmup > 0:
        t0 = time.perf_counter()
        ctx.run(warmup)
        log(f"{label}: {warmup} warmup step(s) {time.perf_counter() - t0:.2f}s")
    c0 = ctx.counters()
    ctx.state()   # synchronise
    if profile:
        ctx.profile_begin()
    log(f"{label}: timing {steps} step(s)")
    t0 = time.perf_counter()
    ctx.run(steps)   # ps_run synchronises the library's stream at its end
    ctx.state()
    dt = time.perf_counter() - t0
    log(f"{label}: {steps} timed step(s) {dt:.3f}s")
    prof = ctx.profile_end() if profile else None
    c1 = ctx.counters()
    per = {k: (c1[k] - c0[k]) / steps for k in ("applies", "minres_iterations", "lanczos_iterations",
                                                 "saddle_solves", "mob_calls", "stresslet_iterations")}
    return dt, prof, per


def fixed_state_timing(Context, pd, x0, over, steps, label):
    """Steps that each start from the same state x0 (restored outside the timed region, the noise
    stream running on), one ps_run(1) per step: the mixed and FP64 forms then do the same work per
    step (the same stresslet iterations), so their ratio is not blurred by two trajectories that part
    ways after a few steps."""
    ctx = make_ctx(Context, pd, x0, **over)
    ctx.run(1)   # warmup
    dt, its = 0.0, 0.0
    for _ in range(steps):
        ctx.initial_configuration(x0)
        ctx.state()
        c0 = ctx.counters()
        t0 = time.perf_counter()
        ctx.run(1)
        ctx.state()
        dt += time.perf_counter() - t0
        its += ctx.counters()["stresslet_iterations"] - c0["stresslet_iterations"]
    ctx.close()
    log(f"{label} from x0: {1e3 * dt / steps:.2f} ms/step, {its / steps:.1f} stresslet iterations per step")
    return {"ms_per_step": 1e3 * dt / steps, "stresslet_iterations": its / steps, "steps": steps}


def lub_effect(Context, pd, x0, over):
    """The lubricated form against the same form without lubrication, one step each from x0 with the
    same noise: colloid pairs inside the cutoff at x0 and the relative change of the step's velocities
    (all beads, and the colloids' U, W)."""
    vs = {}
    for lub in (0, 1):
        c = make_ctx(Context, pd, x0, **dict(over, lub=lub))
        c.run(1)
        vs[lub] = c.vectors()
        c.close()
    c = make_ctx(Context, pd, x0, **over)
    c.check_dist()
    c.mob()
    npairs = c.lub_build(pd.get("lub_rcut", 4.0))
    c.close()
    nm3 = 3 * pd["Nm"]
    a, b = vs[1]["v_det_n"], vs[0]["v_det_n"]
    return {"colloid_pairs_within_cutoff": int(npairs),
            "first_solve_colloid_vel_rel_change": rel(a[nm3:], b[nm3:]),
            "first_solve_vel_rel_change": rel(a, b),
            "step_vel_rel_change": rel(vs[1]["up"], vs[0]["up"]),
            "note": "first solve: the deterministic predictor solve at x0; the monomer rows respond to the sum "
                    "of all colloid forces through the carried translation tile of mob()"}


def mixed_vs_gpu_fp64(Context, pd, x0, forms):
    """Without the oracle (sizes it cannot finish in minutes): one step of each form from x0 with the
    same noise, the mixed form's velocity and displacement against the GPU FP64 form's."""
    out = {}
    for form in ("mixed/block", "fp64/block"):
        c = make_ctx(Context, pd, x0, **forms[form])
        c.run(1)
        v = c.vectors()
        out[form] = (v["up"], c.state()[0], v["v_det_n"])
        c.close()
    (um, xm, vm), (ud, xd, vd) = out["mixed/block"], out["fp64/block"]
    return {"vel_rel_err": rel(um, ud), "disp_rel_err": rel(xm - x0, xd - x0),
            "first_solve_vel_rel_err": rel(vm, vd),
            "reference": "the GPU FP64 block form (the oracle is not run at this size)"}


def one_step_up(Context, pd, x0, **over):
    ctx = make_ctx(Context, pd, x0, **over)
    ctx.run(1)
    up = ctx.vectors()["up"]
    ctx.close()
    return up


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def bond_lengths(x, Nm, Nmono, L):
    """Lengths of the chain bonds (consecutive monomers of each chain), minimum image in a box."""
    xm = np.asarray(x[: 3 * Nm]).reshape(-1, Nmono, 3)
    d = xm[:, 1:] - xm[:, :-1]
    if L:
        d -= L * np.rint(d / L)
    return np.linalg.norm(d, axis=2).ravel()


PEAK_MFMA_16BIT = 2.5e15   # dense BF16 / FP16 MFMA, MI355X
HBM_PEAK = 8.0e12


def dense_bench(args, configs, Context):
    """Stored-dense M (16-bit, MFMA operand order) applied to 16 vectors at once: HBM GB/s of the
    apply kernel (it reads the stored M once per call), MFMA utilisation, and the error of M X
    against the FP64 on-the-fly apply."""
    wl = args.workload if args.workload != "c3" else "c5dense"
    w = configs.workload(wl)
    pd = configs.params_dict(w)
    x0 = configs.positions(w)
    ctx = make_ctx(Context, pd, x0)
    ctx.check_dist()
    ctx.mob()
    n = ctx.n
    t0 = time.perf_counter()
    ctx.dense_build(args.dense)
    ctx.state()
    t_build = time.perf_counter() - t0
    log(f"dense {args.dense}: n = {n}, build {t_build * 1e3:.1f} ms (incl. allocation)")
    X = np.random.default_rng(7).normal(size=(16, n))
    Y = ctx.dense_mult(X)   # warmup
    reps = max(args.steps, 5)
    ctx.profile_begin()
    for _ in range(reps):
        Y = ctx.dense_mult(X)
    prof = ctx.profile_end()
    t = prof["dense_ms"] / max(prof["dense_launches"], 1) * 1e-3
    ref = np.stack([ctx.mob_mult(X[j]) for j in range(16)])
    err = rel(Y, ref)
    col_err = max(rel(Y[j], ref[j]) for j in range(16))
    npad = (n + 31) // 32 * 32
    bytes_m = npad * npad * 2
    flops = 2.0 * n * n * 16
    out = {"metric": "stored-dense 16-bit MFMA multi-vector apply (16 vectors)", "value": bytes_m / t / 1e9,
           "unit": "GB/s", "higher_is_better": True, "n_gpus": 1, "steps": reps, "warmup": 1,
           "ms_per_step": t * 1e3, "dtype": args.dense + " storage, fp32 accumulation", "data": "synthetic",
           "config": {"workload": f"{wl}: {w['Np']} beads, n = {n}, 16 vectors"},
           "roofline": {"bound": "hbm", "achieved": bytes_m / t / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": bytes_m / t / HBM_PEAK, "traffic": bytes_m},
           "mfma_tflops": flops / t / 1e12, "mfma_utilisation": flops / t / PEAK_MFMA_16BIT,
           "rel_err_vs_fp64_apply": err, "worst_vector_rel_err": col_err, "build_ms": t_build * 1e3}
    ctx.close()
    print(json.dumps(out), flush=True)


# The reference's only published throughput: mm_HI=False, Nm = 10,000 monomers in 5000 dumbbells + 1
# trapped colloid, beta 0.001, dt 1e-5, box 10^3, 300 steps, serial CPU.  Best figure: MINRES with a
# single-threaded OpenBLAS, 23.6 s / 300 steps (docs/profiling_baseline_serial.md:109-112, :128).
ARROW_PUBLISHED = 10001 * 300 / 23.6


def arrow_bench(args, configs, Context, steps):
    """The mm_HI=False (arrowhead) regime on the published case (workload arrow10k).  Forms: the
    reference algorithm (Jacobi MINRES rtol 1e-6, diagonal frozen at the first solve) and exact block
    elimination; both FP64.  The timed steps run as ps_run drives them (the block form replays each
    step as a HIP graph); a further profiled run times the dominant kernels by their own dispatch
    events.  Roofline (HBM-bound): for the block form one block-eliminated solve, k_ar_bs1 + k_ar_bs2,
    per solve 8 B x 22 x 3Nm (each pass reads the 8 stored rows of M^cm; pass 1 reads b_l and writes
    f_m, pass 2 reads f_m and b_u and writes u_m and its copy); for MINRES the apply's DOF pass,
    8 B x 10 x 3Nm.  CPU baseline: the oracle on one host core over a bounded number of steps."""
    w = configs.workload("arrow10k")
    pd = configs.params_dict(w)
    x0 = configs.positions(w)
    Np, nm3 = w["Np"], 3 * w["Nm"]
    forms = {"fp64/block": dict(solver=1), "fp64/minres": dict(solver=0)}
    if args.profile_form:
        ctx = make_ctx(Context, pd, x0, **forms[args.profile_form])
        dt, _, per = timed_run(ctx, steps, args.warmup, "arrow " + args.profile_form, profile=False)
        ctx.close()
        log(f"arrow {args.profile_form}: {Np * steps / dt:.4g} bead-steps/s, {1e3 * dt / steps:.3f} ms/step, "
            f"per step {per}")
        return None
    res = {}
    for form, over in forms.items():
        ctx = make_ctx(Context, pd, x0, **over)
        dt, _, per = timed_run(ctx, steps, args.warmup, "arrow " + form, profile=False)
        _, prof, _ = timed_run(ctx, min(steps, 50), 0, "arrow " + form + " kernel timing", profile=True)
        ctx.close()
        if form == "fp64/block":
            kms, kn = prof["arsolve_ms"], prof["arsolve_launches"]
        else:
            kms, kn = prof["arrow_ms"], prof["arrow_launches"]
        avg = kms / max(kn, 1) * 1e-3
        res[form] = dict(value=Np * steps / dt, ms_per_step=1e3 * dt / steps, per_step=per,
                         kernel_ms_avg=avg * 1e3, kernel_launches=kn)
        log(f"arrow {form}: {res[form]['value']:.4g} bead-steps/s, {res[form]['ms_per_step']:.3f} ms/step, "
            f"per step {per}, dominant kernel avg {avg * 1e6:.2f} us over {kn} launches")
    up_b = one_step_up(Context, pd, x0, **forms["fp64/block"])
    up_m = one_step_up(Context, pd, x0, **forms["fp64/minres"])
    head = res["fp64/block"]
    bytes_launch = 8.0 * 22 * nm3   # one block-eliminated solve: both passes
    achieved = bytes_launch / (head["kernel_ms_avg"] * 1e-3) if head["kernel_ms_avg"] > 0 else 0.0
    cpu = None
    if not args.no_cpu:
        from oracle.pyoracle import Oracle, get_threads, set_threads
        nthr = get_threads()
        set_threads(1)
        o = Oracle(pd)
        o.initial_configuration(x0)
        ncpu = args.arrow_cpu_steps
        t0 = time.perf_counter()
        o.run(ncpu)
        t_cpu = (time.perf_counter() - t0) / ncpu
        set_threads(nthr)
        cpu = {"value": Np / t_cpu, "unit": "bead-steps/s", "cores": 1, "kind": "port",
               "sample": f"{ncpu} steps of the oracle's arrowhead regime (the reference algorithm, Jacobi MINRES "
                         f"rtol 1e-6) on the same configuration, one host core, {t_cpu * 1e3:.1f} ms/step"}
        log(f"arrow cpu baseline {cpu}")
    return {
        "metric": "bead-steps/s, mm_HI=False arrowhead, the reference's published case",
        "value": head["value"], "unit": "bead-steps/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": head["value"] / ARROW_PUBLISHED,
        "baseline_published": {"value": ARROW_PUBLISHED, "unit": "bead-steps/s",
                               "source": "docs/profiling_baseline_serial.md:109-112 (MINRES + OpenBLAS, 1 core)"},
        "dtype": "fp64", "data": "synthetic",
        "config": {"workload": f"arrow10k: {w['Nm']} monomers in {w['Npoly']} dumbbells + 1 trapped colloid, "
                               f"beta {w['beta']}, dt {w['dt']}, box {w['L']}, kbond {w['kbond']}, kT {w['kT']}, "
                               f"mm_HI off; exact FP64 block-eliminated saddle solves",
                   "global_batch": Np, "seq_len": 1, "parallelism": "single-gpu"},
        "reference_algorithm": dict(res["fp64/minres"], vs_baseline=res["fp64/minres"]["value"] / ARROW_PUBLISHED),
        "block_vs_minres_velocity_rel_diff": rel(up_b, up_m),
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": "k_ar_bs1 + k_ar_bs2 (one block-eliminated saddle solve, first dispatch to last)",
                     "bytes_per_launch": bytes_launch, "avg_launch_ms": head["kernel_ms_avg"]},
        "cpu_baseline": cpu,
    }


def batch_bench(args, configs):
    """Batched replicas (256 unbounded systems, N_r log-uniform in [100, 3000], 10% probes, 20-bead
    chains, C3 parameters, seeds 1000 + r): one persistent kernel per step, replayed as a captured HIP
    graph (noise copy, team reset, step kernel), replicas far above the
    mean work on teams of workgroups, the packing re-tuned from measured per-replica times.  Forms:
    mixed (FP32 monomer tiles in the 32-iteration Lanczos square root, FP64 solves and recurrences;
    headline) and FP64.  value = aggregate bead-steps/s = sum_r N_r x steps / (sum of the timed
    steps' wall times).  Every timed step starts from the initial configurations (restored before
    the step, outside the timed region): with the reference's carried monomer-colloid tile these
    systems leave the well-posed regime within two steps (overlapping probes, stresslet solves that
    stop at their cap), so later steps would time that breakdown rather than the step.  The first
    (untimed) step of each form is the accuracy step and tunes the schedule.  Accuracy: per-replica
    relative velocity error of mixed against FP64 from the same state and noise, the square root
    with FP32 against FP64 tiles on the same xi, and the bond-length distribution after the first
    step.  Roofline
    (VALU-bound tiles): algorithmic work per step = sum_r 37 x 2 n_r^2 flop (5 full applies of the
    saddle solves + 32 Lanczos applies; the stresslet-row applies are not counted), over the step
    time.  CPU baseline: the oracle (one host core) stepping a sample of replicas, scaled to the
    batch by sum N_r^2."""
    from polystokes_brownian_amd.polystokes import Batch
    t0 = time.perf_counter()
    params, reps = configs.replica_workload(R=args.batch_replicas)
    N = np.array([r["Nc"] + r["Npoly"] * r["Nmono"] for r in reps], dtype=np.int64)
    log(f"batch: {len(reps)} replicas, {N.sum()} beads (max {N.max()}), configurations {time.perf_counter() - t0:.1f}s")
    forms = {"mixed": dict(precision=1), "fp64": dict(precision=0)}
    steps = args.batch_steps
    res = {}
    ups = {}
    sqrt_errs, bonds = None, {}
    for form, over in forms.items():
        b = Batch(dict(params, **over), reps)
        if sqrt_errs is None:
            # the square root on the same xi at the initial positions: FP32 against FP64 tiles
            b.op(4)
            rng = np.random.default_rng(7)
            xis = [rng.normal(size=s_["n"]) for s_ in b.sizes]
            w64, _ = b.op(3, xis)
            w32, _ = b.op(6, xis)
            sqrt_errs = np.array([rel(a, c) for a, c in zip(w32, w64)])
        t1 = time.perf_counter()
        b.run(1)   # accuracy step (also the warmup and the first measured schedule)
        log(f"batch {form}: first step {time.perf_counter() - t1:.3f}s")
        ups[form] = [b.vectors(r)["up"] for r in range(b.R)]
        bonds[form] = [bond_lengths(b.state(r)[0], rep["Npoly"] * rep["Nmono"], rep["Nmono"], None)
                       for r, rep in enumerate(reps)]
        dt, its, unconv = 0.0, 0, 0
        for k in range(steps):
            for r, rep in enumerate(reps):
                b.set_state(r, rep["x0"])
            t1 = time.perf_counter()
            b.run(1)
            dt += time.perf_counter() - t1
            st = b.stats()
            its = max(its, int(st[:, 0].max()))
            unconv += int(st[:, 3].sum())
        us, team = b.times()
        ginfo = b.graph_info()
        b.close()
        res[form] = dict(value=float(N.sum()) * steps / dt, ms_per_step=1e3 * dt / steps,
                         lanczos_iterations_max=int(st[:, 1].max()), minres_iterations_max=its,
                         unconverged_solves=unconv, teams=int((team > 1).sum()),
                         team_workgroups=int(team[team > 1].sum()), team_max=int(team.max()),
                         replica_ms_max=float(us.max() / 1e3), replica_ms_mean=float(us.mean() / 1e3),
                         step_graphs=ginfo)
        log(f"batch {form}: {res[form]['value']:.4g} bead-steps/s, {res[form]['ms_per_step']:.2f} ms/step, "
            f"{res[form]['teams']} teams, replica ms max {res[form]['replica_ms_max']:.1f}")
    errs = np.array([rel(a, c) for a, c in zip(ups["mixed"], ups["fp64"])])
    # bond lengths after the first step, every chain bond of every replica
    ba = np.concatenate([v for v in bonds["mixed"] if v.size])
    bb = np.concatenate([v for v in bonds["fp64"] if v.size])
    r0 = float(params["r0"])
    bond_drift = {"steps": 1, "bonds": int(bb.size), "fp64_mean": float(bb.mean()), "fp64_std": float(bb.std()),
                  "mean_rel": float(abs(ba.mean() - bb.mean()) / bb.mean()),
                  "std_rel": float(abs(ba.std() - bb.std()) / bb.std()),
                  "max_bond_change_over_r0": float(np.abs(ba - bb).max() / r0)}
    log(f"batch mixed vs fp64: sqrt {sqrt_errs.max():.3e} (max), bonds {bond_drift}")
    n = np.array([3 * (r["Npoly"] * r["Nmono"]) + 11 * r["Nc"] for r in reps], dtype=np.float64)
    flops = float(np.sum(37.0 * 2.0 * n * n))
    head = res["mixed"]
    achieved = flops / (head["ms_per_step"] * 1e-3)
    cpu = None
    if not args.no_cpu:
        from oracle.pyoracle import Oracle, get_threads, set_threads
        nthr = get_threads()
        set_threads(1)
        order = np.argsort(N)
        sample = [int(order[k]) for k in np.linspace(0, len(order) - 1, 6).astype(int)[:5]]
        t_cpu = 0.0
        for r in sample:
            rep = reps[r]
            pd = dict(params, Nc=rep["Nc"], Npoly=rep["Npoly"], Nmono_per_chain=rep["Nmono"],
                      Nm=rep["Npoly"] * rep["Nmono"], Np=int(N[r]), seed=rep["seed"], solver=1)
            o = Oracle(pd)
            o.initial_configuration(rep["x0"])
            t1 = time.perf_counter()
            o.run(1)
            t_cpu += time.perf_counter() - t1
        set_threads(nthr)
        scale = float(np.sum(N.astype(float) ** 2)) / float(np.sum(N[sample].astype(float) ** 2))
        t_step = t_cpu * scale
        cpu = {"value": float(N.sum()) / t_step, "unit": "bead-steps/s", "cores": 1, "kind": "port",
               "sample": f"one FP64 step of the oracle (block solves, the same Lanczos) for {len(sample)} replicas "
                         f"(N = {', '.join(str(int(N[r])) for r in sample)}) on one host core, {t_cpu:.2f} s, scaled to "
                         f"the 256 replicas by sum N_r^2 ({t_step:.1f} s per batch step)"}
        log(f"batch cpu baseline {cpu}")
    return {
        "metric": "aggregate bead-steps/s, 256 batched replicas (FP32 square-root tiles, FP64 recurrences)",
        "value": head["value"], "unit": "bead-steps/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64", "data": "synthetic",
        "config": {"workload": f"batched replicas: {len(reps)} unbounded systems, {int(N.sum())} beads, N_r log-uniform "
                               f"in [100, 3000] ({int(N.min())}..{int(N.max())}), 10% probes, 20-bead chains, C3 "
                               f"parameters, seeds 1000 + r, Lanczos m = 32, block-eliminated solves",
                   "global_batch": int(N.sum()), "seq_len": 1, "parallelism": "single-gpu"},
        "fp64_value": res["fp64"]["value"], "fp64_ms_per_step": res["fp64"]["ms_per_step"],
        "mixed_over_fp64": head["value"] / res["fp64"]["value"],
        "schedule": {form: {k: v for k, v in r.items() if k not in ("value", "ms_per_step")} for form, r in res.items()},
        "timed_steps_start_from": "the initial configurations (restored outside the timed region)",
        "per_replica_vel_rel_err_vs_fp64": {"max": float(errs.max()), "median": float(np.median(errs))},
        "per_replica_sqrt_rel_err_vs_fp64": {"max": float(sqrt_errs.max()), "median": float(np.median(sqrt_errs))},
        "bond_length_drift_vs_fp64": bond_drift,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK["fp32"] / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK["fp32"], "traffic": None,
                     "kernel": "k_batch_step (the whole step of every replica; VALU-bound tile generation)",
                     "flops_per_launch": flops, "avg_launch_ms": head["ms_per_step"],
                     "frac_of_single_issue_fp32_fma_rate": achieved / PEAK_FP32_VALU},
        "cpu_baseline": cpu,
    }


def oracle_reference(pd, x0, w, xi):
    """The FP64 oracle (the checker) on the same inputs as the GPU forms: one step of the block
    algorithm from x0 with the same noise stream, and the Lanczos square root M^{1/2} xi at x0.  Also
    the CPU baseline: the step is timed on every host thread the oracle uses."""
    from oracle.pyoracle import Oracle, get_threads
    o = Oracle(dict(pd, solver=1), dense=False)
    o.initial_configuration(x0)
    t0 = time.perf_counter()
    o.run(1)
    t_step = time.perf_counter() - t0
    vec = o.vectors()
    x1 = o.state()[0]
    ctr = o.counters()
    log(f"oracle: one FP64 step of the block algorithm in {t_step:.2f}s on {get_threads()} threads")
    o2 = Oracle(dict(pd, solver=1), dense=False)
    o2.initial_configuration(x0)
    o2.check_dist()
    o2.mob()
    t0 = time.perf_counter()
    W, info = o2.sqrt_apply(xi)
    log(f"oracle: square root ({info.iterations} Lanczos vectors) in {time.perf_counter() - t0:.2f}s")
    return dict(up=vec["up"], x1=x1, W=W, lanczos=info.iterations, t_step=t_step, threads=get_threads(),
                applies=ctr["applies"], bonds=bond_lengths(x1, w["Nm"], w["Nmono"], w.get("L")))


def errors_vs_oracle(Context, pd, x0, w, over, ref, xi):
    """One step of a GPU form from x0 against the oracle's step (same state and noise): the velocity
    and displacement, the Lanczos square root on the same xi, and the bond-length distribution."""
    c = make_ctx(Context, pd, x0, **over)
    c.check_dist()
    c.mob()
    W, info = c.sqrt_apply(xi)
    c.close()
    c = make_ctx(Context, pd, x0, **over)
    c.run(1)
    up = c.vectors()["up"]
    x1 = c.state()[0]
    c.close()
    b = bond_lengths(x1, w["Nm"], w["Nmono"], w.get("L"))
    bo = ref["bonds"]
    return {
        "vel_rel_err": rel(up, ref["up"]),
        "disp_rel_err": rel(x1 - x0, ref["x1"] - x0),
        "sqrt_rel_err": rel(W, ref["W"]),
        "sqrt_lanczos_vectors": [info.iterations, ref["lanczos"]],
        "bond_length_drift": {"steps": 1, "bonds": int(bo.size), "oracle_mean": float(bo.mean()),
                              "oracle_std": float(bo.std()), "oracle_max": float(bo.max()),
                              "mean_rel": float(abs(b.mean() - bo.mean()) / bo.mean()),
                              "std_rel": float(abs(b.std() - bo.std()) / bo.std()),
                              "max_bond_change_over_r0": float(np.abs(b - bo).max() / w["r0"])},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle: no CPU baseline, no errors vs the oracle")
    ap.add_argument("--minres", action="store_true",
                    help="also time the reference algorithm (Jacobi MINRES on the saddle system) for one step")
    ap.add_argument("--lub", action="store_true",
                    help="also time the mixed block form with lubrication in the saddle (lub = 1), and its effect "
                         "on the step's velocities")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device (a plain copy of known size, to scale HBM counters)")
    ap.add_argument("--dense", default=None, metavar="STORAGE", choices=["bf16", "f16"],
                    help="instead: the stored-dense MFMA multi-vector apply (16 vectors) at --workload "
                         "(default c5dense, N = 8,000) with STORAGE; prints its own JSON line")
    ap.add_argument("--profile-form", default=None,
                    help="only run warmup + timed steps of this form (for a profiler), no JSON line")
    ap.add_argument("--profile-events", action="store_true",
                    help="with --profile-form: time the apply kernel's launches with the library's events over "
                         "the timed steps (the measurement the JSON line's roofline uses) and log the average")
    ap.add_argument("--arrow", action="store_true",
                    help="instead: the mm_HI=False arrowhead regime on the reference's published case; prints "
                         "its own JSON line")
    ap.add_argument("--arrow-steps", type=int, default=300, help="timed steps of the arrowhead case (300 as published)")
    ap.add_argument("--arrow-cpu-steps", type=int, default=20, help="oracle steps of the arrowhead CPU baseline")
    ap.add_argument("--no-arrow", action="store_true", help="omit the arrowhead case from the default line")
    ap.add_argument("--batch", action="store_true",
                    help="instead: the batched-replica workload (256 replicas); prints its own JSON line")
    ap.add_argument("--batch-replicas", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--peaks", action="store_true",
                    help="instead: only the VALU issue-rate probe (instructions per CU per cycle, shader clock)")
    ap.add_argument("--no-batch", action="store_true", help="omit the batched replicas from the default line")
    ap.add_argument("--no-c5", action="store_true", help="omit the 100,000-bead workload from the default line")
    ap.add_argument("--c5-steps", type=int, default=2, help="timed steps of the 100,000-bead workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the solve is not sharded across GPUs",
              file=sys.stderr)
        sys.exit(2)

    from polystokes_brownian_amd import configs
    from polystokes_brownian_amd.polystokes import Context

    if args.calibrate_copy > 0:
        import torch
        nel = args.calibrate_copy * (1 << 20) // 8
        src = torch.randn(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {nel * 8} bytes read and written per dispatch, 2 dispatches")
        del src, dst

    if args.peaks:
        print(json.dumps({"valu_issue_probe": measured_peaks()}), flush=True)
        return
    if args.dense:
        return dense_bench(args, configs, Context)
    if args.batch:
        print(json.dumps(batch_bench(args, configs)), flush=True)
        return
    if args.arrow:
        line = arrow_bench(args, configs, Context, args.arrow_steps)
        if line is not None:
            print(json.dumps(line), flush=True)
        return

    out = workload_line(args, args.workload, args.steps, args.warmup, with_cpu=not args.no_cpu, graphs=(0, 1),
                        minres=args.minres, lub=args.lub, profile_form=args.profile_form,
                        profile_events=args.profile_events)
    if out is None:
        return
    if args.workload == "c3" and not args.no_c5:
        # the large matrix-free workload (100,000 beads): GPU forms only, errors against the GPU FP64 form
        out["large_matrix_free_c5"] = workload_line(args, "c5", args.c5_steps, 1, with_cpu=False, graphs=(0,))
    if args.workload == "c3" and not args.no_batch:
        out["batched_replicas_c4"] = batch_bench(args, configs)
    if not args.no_arrow:
        out["arrowhead_published_case"] = arrow_bench(args, configs, Context, args.arrow_steps)
    out["valu_issue_probe"] = measured_peaks()
    print(json.dumps(out), flush=True)


def workload_line(args, wname, steps, warmup, with_cpu, graphs, minres=False, lub=False, profile_form=None,
                  profile_events=False):
    """The timed forms of one single-system workload and their JSON object (None with profile_form)."""
    from polystokes_brownian_amd import configs
    from polystokes_brownian_amd.polystokes import Context
    w = configs.workload(wname)
    pd = configs.params_dict(w)
    t0 = time.perf_counter()
    x0 = configs.positions(w)
    Np, Nm, Nc = w["Np"], w["Nm"], w["Nc"]
    log(f"workload {wname}: Np={Np} (Nc={Nc}, Nm={Nm}), configuration {time.perf_counter() - t0:.1f}s")

    # mixed: every mobility apply on FP32 tiles (fixed-point separations), FP64 recurrences, stresslet
    # solves and colloid pairs, no refinement sweep (the block elimination is exact up to the apply)
    forms = {
        "mixed/block": dict(precision=1, solver=1, ir_maxit=0),
        "fp64/block": dict(precision=0, solver=1),
    }
    if minres:
        forms["fp64/minres"] = dict(precision=0, solver=0)
    if lub:
        # lubrication in the saddle (opt-in physics, not the reference's system): eager only, its
        # colloid saddle solves poll on the host
        forms["mixed/block+lub"] = dict(precision=1, solver=1, ir_maxit=0, lub=1)
    if profile_form:
        ctx = make_ctx(Context, pd, x0, graph=0, **forms[profile_form])
        dt, prof, per = timed_run(ctx, steps, warmup, profile_form, profile=profile_events)
        log(f"{profile_form}: {Np * steps / dt:.4g} bead-steps/s, {1e3 * dt / steps:.2f} ms/step, "
            f"per step {per}")
        if prof:
            log(f"{profile_form}: apply kernel {prof['aa_launches']} launches, average "
                f"{prof['aa_ms'] / max(prof['aa_launches'], 1) * 1e3:.2f} us (library events)")
        ctx.close()
        return None

    # timed steps: each form eagerly (graph 0) and as replayed step graphs (graph 1); the faster is
    # the form's figure, both are reported
    results = {}
    for form, over in forms.items():
        if form == "fp64/minres":
            ctx = make_ctx(Context, pd, x0, **over)
            dt, _, per = timed_run(ctx, 1, 0, form, profile=False)
            ctx.close()
            results[form] = dict(value=Np / dt, ms_per_step=1e3 * dt, per_step=per)
            continue
        if form.endswith("+lub"):
            ctx = make_ctx(Context, pd, x0, graph=0, **over)
            dt, _, per = timed_run(ctx, steps, warmup, form, profile=False)
            ctx.close()
            results[form] = dict(value=Np * steps / dt, ms_per_step=1e3 * dt / steps, per_step=per,
                                 graph=0)
            results[form]["effect_vs_no_lubrication"] = lub_effect(Context, pd, x0, over)
            log(f"{form}: {results[form]['value']:.4g} bead-steps/s, {results[form]['ms_per_step']:.2f} ms/step, "
                f"per step {per}, {results[form]['effect_vs_no_lubrication']}")
            continue
        modes = {}
        for graph in graphs:
            ctx = make_ctx(Context, pd, x0, graph=graph, **over)
            dt, _, per = timed_run(ctx, steps, warmup, f"{form} graph={graph}", profile=False)
            ctx.close()
            modes[graph] = dict(value=Np * steps / dt, ms_per_step=1e3 * dt / steps, per_step=per)
        best = max(modes, key=lambda g: modes[g]["value"])
        results[form] = dict(modes[best], graph=best)
        for g in modes:
            results[form]["graph_ms_per_step" if g else "eager_ms_per_step"] = modes[g]["ms_per_step"]
        # kernel timing in a separate run (the library records events around the launches, which
        # turns graph replay off): the fused monomer apply kernel, and each whole apply
        ctx = make_ctx(Context, pd, x0, graph=0, **over)
        _, prof, _ = timed_run(ctx, steps, warmup, f"{form} kernel timing", profile=True)
        results[form]["monomer_kernel"] = ctx.kernel_choice()
        ctx.close()
        results[form]["apply_kernel_ms_avg"] = prof["aa_ms"] / max(prof["aa_launches"], 1)
        results[form]["apply_kernel_launches"] = prof["aa_launches"]
        results[form]["apply_ms_avg"] = prof["apply_ms"] / max(prof["apply_launches"], 1)
        timing = ", ".join(f"graph={g} {m['ms_per_step']:.2f}" for g, m in modes.items())
        log(f"{form}: {results[form]['value']:.4g} bead-steps/s, {results[form]['ms_per_step']:.2f} ms/step "
            f"({timing}), per step {per}, apply kernel {results[form]['apply_kernel_ms_avg'] * 1e3:.1f} us")

    # both forms on steps that each start from x0 (the same work per step)
    fixed = None
    if "mixed/block" in forms and "fp64/block" in forms:
        fm = fixed_state_timing(Context, pd, x0, forms["mixed/block"], steps, "mixed/block")
        ff = fixed_state_timing(Context, pd, x0, forms["fp64/block"], steps, "fp64/block")
        fixed = {"mixed/block": fm, "fp64/block": ff, "mixed_over_fp64": ff["ms_per_step"] / fm["ms_per_step"],
                 "note": "each step from the initial configuration (restored outside the timed region), one "
                         "ps_run(1) per step, launched kernel by kernel"}
    # accuracy against the FP64 oracle, and the CPU baseline (the oracle's step on the host cores)
    errs, cpu, cpu1 = {}, None, None
    if with_cpu:
        xi = np.random.default_rng(7).normal(size=3 * Nm + 11 * Nc)
        ref = oracle_reference(pd, x0, w, xi)
        for form, over in forms.items():
            if form == "fp64/minres" or form.endswith("+lub"):
                continue
            errs[form] = errors_vs_oracle(Context, pd, x0, w, over, ref, xi)
            log(f"{form} vs oracle: {errs[form]}")
        cpu = {"value": Np / ref["t_step"], "unit": "bead-steps/s", "cores": ref["threads"], "kind": "port",
               "sample": f"one full FP64 step of the oracle (the block algorithm: {ref['applies']} matrix-free "
                         f"applies, the stresslet solves, the Lanczos square root of {ref['lanczos']} vectors, "
                         f"the Verlet lists) at N={Np} on {ref['threads']} host threads: {ref['t_step']:.2f} s"}
        from oracle.pyoracle import Oracle, get_threads, set_threads
        nthr = get_threads()
        set_threads(1)
        o = Oracle(pd, dense=False)
        o.initial_configuration(x0)
        o.check_dist()
        o.mob()
        f = np.random.default_rng(0).normal(size=o.n)
        t0 = time.perf_counter()
        o.mob_mult(f)
        t_apply = time.perf_counter() - t0
        set_threads(nthr)
        cpu1 = {"value": Np / (t_apply * ref["applies"]), "unit": "bead-steps/s", "cores": 1, "kind": "port",
                "extrapolated": True,
                "sample": f"one FP64 matrix-free apply of the oracle at N={Np} on one host core ({t_apply:.2f} s), "
                          f"times the {ref['applies']} applies of one block-algorithm step (the step's other work "
                          f"is under 2% of it on the host)"}
        log(f"cpu baseline {cpu}; one core {cpu1}")
    errs_gpu = None
    if not with_cpu:
        errs_gpu = mixed_vs_gpu_fp64(Context, pd, x0, forms)
        log(f"mixed/block vs GPU FP64: {errs_gpu}")

    head = results["mixed/block"]
    f64 = results["fp64/block"]
    # the fused apply kernel's algorithmic work: the monomer-monomer block and both monomer-colloid
    # blocks of M applied to one vector, 2 flop per entry
    # (the ordered kernel, where the measured choice picks it, is timed without the monomer-colloid
    # kernels that follow it, so its work is the monomer-monomer block alone)
    n_mm, n_mc = 3 * Nm, 2 * 3 * Nm * 11 * Nc
    sym32 = head["monomer_kernel"]["fp32"] == "symmetric"
    sym64 = f64["monomer_kernel"]["fp64"] == "symmetric"
    flops = 2.0 * (n_mm * n_mm + (n_mc if sym32 else 0))
    flops64 = 2.0 * (n_mm * n_mm + (n_mc if sym64 else 0))
    ach = flops / (head["apply_kernel_ms_avg"] * 1e-3)
    ach64 = flops64 / (f64["apply_kernel_ms_avg"] * 1e-3)
    traffic = TRAFFIC_APPLY.get((wname, "fp32")) if sym32 else None
    kname = ("k_apply_sym<float, true> (monomer-monomer tile pairs on FP32 tiles with fixed-point separations, "
             "the monomer-colloid blocks in the same launch); VALU-bound tile generation" if sym32 else
             "k_aa<float> (ordered monomer-monomer tiles on FP32); VALU-bound tile generation")
    roof = {"bound": "valu", "achieved": ach / 1e12, "peak": PEAK["fp32"] / 1e12, "unit": "TFLOP/s",
            "frac": ach / PEAK["fp32"], "traffic": traffic, "kernel": kname,
            "kernel_choice": head["monomer_kernel"],
            "flops_per_launch": flops, "avg_launch_ms": head["apply_kernel_ms_avg"],
            "frac_of_single_issue_fp32_fma_rate": ach / PEAK_FP32_VALU,
            "valu_issue": VALU_ISSUE.get((wname, "fp32")) if sym32 else None,
            "traffic_algorithmic_bytes": ALGO_BYTES(Nm, Nc, 4)}
    out = {
        "metric": "bead-steps/s at N=20k: mixed-prec vs FP64 GPU at 1e-6 rel err; % roofline",
        "value": head["value"], "unit": "bead-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64", "data": "synthetic",
        "config": {"workload": f"{wname}: {Np} beads ({Nc} probes + {w['Npoly']} chains x 20), "
                               f"box {w['L']:.1f}, full mm-HI, mixed precision (every mobility apply on FP32 "
                               f"tiles, FP64 recurrences and stresslet solves), exact block-eliminated saddle "
                               f"solves, Lanczos m<={pd['sqrt_kmax']} rtol {pd['sqrt_rtol']}",
                   "global_batch": Np, "seq_len": 1, "parallelism": "single-gpu"},
        "errors_vs_fp64_oracle": errs if errs else None,
        "errors_mixed_vs_gpu_fp64": errs_gpu,
        "fp64_value": f64["value"],
        "mixed_over_fp64": head["value"] / f64["value"],
        "from_fixed_state": fixed,
        "forms": results,
        "roofline": roof,
        "roofline_fp64": {"bound": "valu", "achieved": ach64 / 1e12, "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s",
                          "frac": ach64 / PEAK["fp64"], "avg_launch_ms": f64["apply_kernel_ms_avg"],
                          "traffic": TRAFFIC_APPLY.get((wname, "fp64"))},
        "cpu_baseline": cpu,
        "cpu_baseline_one_core": cpu1,
    }
    if "fp64/minres" in results:
        out["reference_algorithm_fp64_minres"] = dict(
            results["fp64/minres"],
            note="Jacobi MINRES on the saddle system (rtol 1e-6, cap 10000 iterations per solve) does not "
                 "converge at this size; one step")
    return out




if __name__ == "__main__":
    main()
