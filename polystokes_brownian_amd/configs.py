"""Synthetic workloads in the reference's units.

Unit mapping (one decision, used by every workload): the reference's length unit is the probe
(colloid) radius and its mobility unit sets 6*pi*eta = 1 (src/mob.cpp:65-86, :572-583).  The
workloads are specified with bead radius a = 1 and probe radius 5a, so
    beta = 0.2,  lengths x 0.2,  spring constants x 25 (energy / length^2),
    time x beta^3 = 0.008 (tau is the bead's Stokes time 6*pi*eta*a^3/kT),
so dt = 1e-3 tau -> 8e-6, kbond = 100 -> 2500, r0 = 2.1a -> 0.42, ktrap = 500 -> 12500,
Lmax = 3a -> 0.6, box 60a -> 12, 150a -> 30, 320a -> 64, 550a -> 110.

Initial configurations follow the reference's generator conventions (seeded uniform probe centres,
isotropic random-walk chains with minimum-image overlap rejection; test_sim_dumbbells_thermal.py
:93-174): bead-bead >= 2*beta, bead-probe >= 1+beta, probe-probe >= 2.
"""
import math

import numpy as np

BETA = 0.2
LSCALE = BETA
KSCALE = 1.0 / BETA ** 2
TSCALE = BETA ** 3


def _min_image(d, L):
    return d - L * np.rint(d / L)


def _wrap(x, L):
    return x - L * np.rint(x / L)


def make_configuration(Nc, Npoly, Nmono, L, r0, seed, probe_positions=None, beta=BETA, max_tries=2000,
                       unwrap=False):
    """Positions (3*Np,) for Nm = Npoly*Nmono monomers followed by Nc probes, wrapped in [-L/2, L/2).
    With `unwrap`, chains are returned as the continuous walks they were grown as (for unbounded
    systems): the overlap tests used minimum images, and unbounded distances are never shorter."""
    rng = np.random.default_rng(seed)
    Nm = Npoly * Nmono
    x = np.zeros((Nm + Nc, 3))
    # probes
    if probe_positions is not None:
        probes = np.asarray(probe_positions, dtype=float).reshape(Nc, 3)
    else:
        probes = np.zeros((Nc, 3))
        k = 0
        tries = 0
        while k < Nc:
            c = rng.uniform(-L / 2, L / 2, 3)
            if k == 0 or np.all(np.linalg.norm(_min_image(probes[:k] - c, L), axis=1) >= 2.0 + 0.05):
                probes[k] = c
                k += 1
            tries += 1
            if tries > 1000 * max(Nc, 1):
                raise RuntimeError("could not place probes")
    x[Nm:] = probes
    # cell grid for fast overlap tests of monomers
    cell = 1.0 + beta
    ncell = max(1, int(L // cell))
    grid = {}

    def key(p):
        return tuple(((np.floor((_wrap(p, L) + L / 2) / (L / ncell))).astype(int)) % ncell)

    def ok(p, skip=-1):
        dprobe = np.linalg.norm(_min_image(probes - p, L), axis=1) if Nc else np.zeros(0)
        if np.any(dprobe < 1.0 + beta):
            return False
        kx, ky, kz = key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get(((kx + dx) % ncell, (ky + dy) % ncell, (kz + dz) % ncell), ()):
                        if j == skip:
                            continue
                        if np.linalg.norm(_min_image(x[j] - p, L)) < 2 * beta:
                            return False
        return True

    def add(i, p):
        x[i] = p
        grid.setdefault(key(p), []).append(i)

    def remove(i):
        grid[key(x[i])].remove(i)

    for c in range(Npoly):
        for attempt in range(max_tries):
            start = rng.uniform(-L / 2, L / 2, 3)
            if not ok(start):
                continue
            placed = [c * Nmono]
            add(c * Nmono, start)
            good = True
            for m in range(1, Nmono):
                i = c * Nmono + m
                prev = x[i - 1]
                for _ in range(200):
                    v = rng.normal(size=3)
                    v /= np.linalg.norm(v)
                    p = prev + r0 * v
                    if ok(p, skip=i - 1):
                        add(i, p)
                        placed.append(i)
                        break
                else:
                    good = False
                    break
            if good:
                break
            for i in placed:
                remove(i)
        else:
            raise RuntimeError("could not place chain %d" % c)
    if unwrap:
        return x.reshape(-1).copy()
    return _wrap(x, L).reshape(-1)


def workload(name):
    """Parameter dict + generator arguments for a named workload."""
    base = dict(
        mm_HI=1, fene=0, record_forces=1, tether=0, mono_ev=0, seed=12345, restart=0.0,
        kT=1.0, epsilon=1.0, beta=BETA, kbond=100.0 * KSCALE, r0=2.1 * LSCALE, Lmax=3.0 * LSCALE,
        tau=1000.0, ktrap=500.0 * KSCALE, tstart=0.0, trun=1e9, weaken_trap=-1.0,
        dt=1e-3 * TSCALE, samplerate=1 << 30,
    )
    if name == "c1":
        d = dict(base, Nc=2, Npoly=11, Nmono=20, L=60 * LSCALE, chain_HI=1, ksp_rtol=1e-10,
                 sqrt_rtol=1e-8, sqrt_kmax=64, steps=100, probes=[[-3.0, 0, 0], [3.0, 0, 0]], ic_seed=42)
    elif name == "c2":
        d = dict(base, Nc=20, Npoly=99, Nmono=20, L=150 * LSCALE, chain_HI=0, fene=1, ksp_rtol=1e-10,
                 sqrt_rtol=1e-8, sqrt_kmax=64, steps=200, probes=None, ic_seed=42)
    elif name == "c3":
        d = dict(base, Nc=200, Npoly=990, Nmono=20, L=320 * LSCALE, chain_HI=0, ksp_rtol=1e-6,
                 sqrt_rtol=1e-6, sqrt_kmax=32, steps=2, probes=None, ic_seed=42)
    elif name == "c5":
        d = dict(base, Nc=1000, Npoly=4950, Nmono=20, L=550 * LSCALE, chain_HI=0, ksp_rtol=1e-6,
                 sqrt_rtol=0.0, sqrt_kmax=40, steps=1, probes=None, ic_seed=42)
    elif name == "c5dense":
        d = dict(base, Nc=80, Npoly=396, Nmono=20, L=None, chain_HI=0, ksp_rtol=1e-6,
                 sqrt_rtol=1e-6, sqrt_kmax=32, steps=1, probes=None, ic_seed=7)
        # same density as the 100k case
        d["L"] = 550 * LSCALE * (8000 / 100000) ** (1 / 3)
    elif name in ("arrow10k", "arrow_thermal"):
        # monomer-monomer HI off (the arrowhead operator), one trapped colloid at the origin and
        # free dumbbells, in the reference's own units.  arrow10k is the case of the reference's
        # published serial profile (docs/profiling_baseline_serial.md:6-16, scratch_profile.py):
        # 5000 dumbbells, beta 0.001, dt 1e-5, box 10, kbond 1, 300 steps.  arrow_thermal follows
        # test_sim_dumbbells_thermal.py:455-520 (beta 0.1, dt 1e-3) with fewer dumbbells.
        beta = 0.001 if name == "arrow10k" else 0.1
        nd = 5000 if name == "arrow10k" else 500
        r0 = 2.0 ** (1.0 / 6.0) * (2.0 * beta) + (0.03 if name == "arrow10k" else 0.0)
        d = dict(base, mm_HI=0, chain_HI=0, record_forces=0, Nc=1, Npoly=nd, Nmono=2, L=10.0, beta=beta,
                 dt=1e-5 if name == "arrow10k" else 1e-3, kbond=1.0, r0=r0, Lmax=1.5 * r0, tau=1000.0,
                 epsilon=5.0 if name == "arrow10k" else 1.0, ktrap=1.0, ksp_rtol=1e-6, sqrt_rtol=1e-6,
                 sqrt_kmax=32, steps=300, probes=None, ic_seed=2024, dumbbells=True,
                 min_sep=2.0 ** (1.0 / 6.0) * (2.0 * beta))
    else:
        raise KeyError(name)
    d["Nm"] = d["Npoly"] * d["Nmono"]
    d["Np"] = d["Nm"] + d["Nc"]
    return d


def small(Nc, Npoly, Nmono, L, seed=1, chain_HI=0, **over):
    """A small workload for parity tests."""
    d = workload("c1")
    d.update(dict(Nc=Nc, Npoly=Npoly, Nmono=Nmono, L=L, chain_HI=chain_HI, probes=None, ic_seed=seed))
    d.update(over)
    d["Nm"] = d["Npoly"] * d["Nmono"]
    d["Np"] = d["Nm"] + d["Nc"]
    return d


def params_dict(w):
    """Split a workload into the ps_params fields."""
    keys = ["mm_HI", "chain_HI", "fene", "record_forces", "tether", "mono_ev", "seed", "restart", "kT",
            "epsilon", "beta", "kbond", "r0", "Lmax", "tau", "ktrap", "tstart", "trun", "weaken_trap", "dt",
            "samplerate", "Nc", "Nm", "Np", "Npoly", "ksp_rtol", "sqrt_rtol", "sqrt_kmax"]
    d = {k: w[k] for k in keys if k in w}
    d["Nmono_per_chain"] = w["Nmono"]
    d["box"] = [w["L"]] * 3 if w.get("L") else None
    d["tmax"] = w["dt"] * w.get("steps", 1)
    return d


def dumbbell_configuration(n_dumbbell, r0, kbond, kT, L, seed, clearance=1.5, min_sep=0.0, max_tries=1000,
                           min_bond=0.0):
    """Positions (3*(2 n_dumbbell + 1),) of free dumbbells around one colloid at the origin, drawn as
    the reference's thermal-dumbbell generator draws them (test_sim_dumbbells_thermal.py:93-174):
    centre uniform in the box, bond vector r0 u + N(0, kT/kbond) per component with u isotropic,
    beads wrapped into the box, rejected when closer than `clearance` to the colloid or closer than
    `min_sep` (minimum image) to a bead of an earlier dumbbell.  `min_bond` (not in the reference's
    generator) also rejects bond vectors shorter than it, e.g. inside the monomer WCA core."""
    rng = np.random.default_rng(seed)
    Lv = np.full(3, float(L))
    sigma = math.sqrt(kT / kbond)
    beads = np.zeros((2 * n_dumbbell, 3))
    ncell = max(1, int(L // min_sep)) if min_sep > 0 else 1
    grid = {}

    def key(p):
        return tuple((np.floor((p + Lv / 2) / (Lv / ncell)).astype(int)) % ncell)

    def clear(p):
        if min_sep <= 0:
            return True
        kx, ky, kz = key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get(((kx + dx) % ncell, (ky + dy) % ncell, (kz + dz) % ncell), ()):
                        if np.linalg.norm(_min_image(q - p, Lv)) < min_sep:
                            return False
        return True

    for dd in range(n_dumbbell):
        for _ in range(max_tries):
            c = rng.uniform(-Lv / 2, Lv / 2)
            u = rng.normal(size=3)
            v = r0 * u / np.linalg.norm(u) + rng.normal(scale=sigma, size=3)
            if np.linalg.norm(v) < min_bond:
                continue
            b0, b1 = _wrap(c - 0.5 * v, Lv), _wrap(c + 0.5 * v, Lv)
            if np.linalg.norm(b0) < clearance or np.linalg.norm(b1) < clearance:
                continue
            if not (clear(b0) and clear(b1)):
                continue
            beads[2 * dd], beads[2 * dd + 1] = b0, b1
            for b in (b0, b1):
                grid.setdefault(key(b), []).append(b)
            break
        else:
            raise RuntimeError("could not place dumbbell %d" % dd)
    return np.concatenate([beads.reshape(-1), np.zeros(3)])


def positions(w):
    if w.get("dumbbells"):
        return dumbbell_configuration(w["Npoly"], w["r0"], w["kbond"], w["kT"], w["L"], w["ic_seed"],
                                      min_sep=w.get("min_sep", 0.0))
    return make_configuration(w["Nc"], w["Npoly"], w["Nmono"], w["L"], w["r0"], w["ic_seed"],
                              probe_positions=w.get("probes"))


def replica_sizes(R=256, seed=1000, lo=100, hi=3000):
    """Bead counts of the batched-replica workload: N_r = round(exp(U(ln lo, ln hi))), 10% probes,
    the rest 20-bead chains."""
    out = []
    for r in range(R):
        rng = np.random.default_rng(seed + r)
        N = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        Nc = max(1, int(round(0.1 * N)))
        Npoly = max(1, (N - Nc) // 20)
        out.append((Nc, Npoly, 20))
    return out


def replica_workload(R=256, seed=1000, lo=100, hi=3000, sizes=None):
    """The batched-replica workload: R unbounded systems with the C3 parameters (harmonic bonds, full
    monomer hydrodynamics, no chain_HI), sizes from replica_sizes, each at C3's bead density (box
    side 320a (N_r / 20 000)^(1/3) for the placement), positions and noise seeded seed + r.  Lanczos
    runs 32 iterations.  Returns (params dict, list of replica dicts)."""
    w = workload("c3")
    w.update(L=None, sqrt_kmax=32, sqrt_rtol=0.0)
    params = params_dict(w)
    reps = []
    for r, (Nc, Npoly, Nmono) in enumerate(sizes if sizes is not None else replica_sizes(R, seed, lo, hi)):
        N = Nc + Npoly * Nmono
        L = 320 * LSCALE * (N / 20000.0) ** (1.0 / 3.0)
        x0 = make_configuration(Nc, Npoly, Nmono, L, w["r0"], seed + r, unwrap=True)
        reps.append(dict(Nc=Nc, Npoly=Npoly, Nmono=Nmono, seed=seed + r, x0=x0))
    return params, reps

