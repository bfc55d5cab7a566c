"""Batched replicas on the CPU: the replica workload generator and the host-side validation of
ps_batch_create (no device call is made before it)."""
import math

import numpy as np
import pytest

from polystokes_brownian_amd import configs


def test_replica_sizes_follow_the_workload_definition():
    sizes = configs.replica_sizes(R=256, seed=1000)
    assert len(sizes) == 256
    N = np.array([nc + npoly * nm for nc, npoly, nm in sizes])
    for (nc, npoly, nm) in sizes:
        assert nm == 20 and nc >= 1 and npoly >= 1
    # N_r = round(exp(U(ln 100, ln 3000))) before the chains are cut to whole 20-bead chains
    assert N.min() >= 100 - 20 and N.max() <= 3000
    assert abs(np.median(np.log(N)) - 0.5 * (math.log(100) + math.log(3000))) < 0.3
    # 10% probes
    frac = np.array([nc for nc, _, _ in sizes]) / N
    assert np.all(np.abs(frac - 0.1) < 0.03)


def test_replica_configurations_are_continuous_chains_without_overlaps():
    params, reps = configs.replica_workload(sizes=[(3, 4, 20), (2, 6, 20)], seed=2000)
    assert params["box"] is None and params["mm_HI"] == 1 and params["sqrt_rtol"] == 0.0
    for rep in reps:
        x = rep["x0"].reshape(-1, 3)
        Nm = rep["Npoly"] * rep["Nmono"]
        for c in range(rep["Npoly"]):
            bonds = np.linalg.norm(np.diff(x[c * 20:(c + 1) * 20], axis=0), axis=1)
            np.testing.assert_allclose(bonds, params["r0"], rtol=1e-12)   # unwrapped: every bond r0
        d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2) + np.eye(len(x)) * 1e9
        beta = params["beta"]
        assert d[:Nm, :Nm].min() >= 2 * beta - 1e-12
        if rep["Nc"]:
            assert d[:Nm, Nm:].min() >= 1 + beta - 1e-12
            if rep["Nc"] > 1:
                assert d[Nm:, Nm:].min() >= 2.0


def test_batch_create_rejects_a_box_and_the_arrowhead_regime():
    from polystokes_brownian_amd.polystokes import Batch, PsError
    params, reps = configs.replica_workload(sizes=[(3, 4, 20), (1, 2, 20)], seed=2000)
    with pytest.raises(PsError, match="unbounded"):
        Batch(dict(params, box=[10.0] * 3), reps)
    with pytest.raises(PsError, match="mm_HI"):
        Batch(dict(params, mm_HI=0), reps)
    with pytest.raises(PsError, match="sqrt_kmax"):
        Batch(dict(params, sqrt_kmax=64), reps)


def _team_time(t1, g):
    # the planner's team model: t1 (0.08 + 0.92/g) 1.05 on g > 1 workgroups
    return t1 if g <= 1 else t1 * (0.08 + 0.92 / g) * 1.05


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_schedule_planner_invariants(seed):
    """ps_batch_plan (the step schedule's planner, host only): every replica placed once, teams of at
    most cap that fit in the workgroups, every workgroup's load within the makespan target, which is
    at or above both lower bounds, and each team the smallest that meets it."""
    from polystokes_brownian_amd.polystokes import batch_plan
    rng = np.random.default_rng(seed)
    R, W = 256, 256
    N = np.exp(rng.uniform(np.log(100), np.log(3000), R))
    cost = 1e-3 * (N * N + 30 * (0.1 * N) ** 2) + 0.5 * N + 400
    cap = np.minimum(16, np.maximum(1, (N // 64 + N // 160) // 2)).astype(int)
    team, wg, load, T = batch_plan(cost, cap, W)
    assert np.all((team >= 1) & (team <= cap))
    assert np.all((wg >= 0) & (wg < W))
    assert team[team > 1].sum() <= W
    assert load.max() <= T * (1 + 1e-12)
    assert T >= cost.sum() / W * (1 - 1e-12)
    assert T >= max(_team_time(c, k) for c, k in zip(cost, cap)) * (1 - 1e-12)
    for c, g in zip(cost, team):
        if g == 1:
            assert c <= T * (1 + 1e-12)
        else:
            assert _team_time(c, g) <= T * (1 + 1e-12)
            assert g == 2 or _team_time(c, g - 1) > T
    # the loads account for every replica's work exactly once
    work = sum(_team_time(c, g) * g if g > 1 else c for c, g in zip(cost, team))
    assert abs(load.sum() - work) <= 1e-9 * work


def test_schedule_planner_small_cases():
    from polystokes_brownian_amd.polystokes import batch_plan
    # equal replicas, more workgroups than replicas: one each, makespan = one replica
    team, wg, load, T = batch_plan(np.ones(10), np.ones(10, int), 16)
    assert np.all(team == 1) and len(set(wg)) == 10 and abs(T - 1.0) < 1e-6
    # more replicas than workgroups: packed two per workgroup
    team, wg, load, T = batch_plan(np.ones(8), np.ones(8, int), 4)
    assert np.all(team == 1) and abs(T - 2.0) < 1e-6 and np.allclose(load, 2.0)
    # one large replica gets a team; the small ones fill the rest
    team, wg, load, T = batch_plan(np.array([100.0] + [1.0] * 20), np.array([16] + [1] * 20), 16)
    assert team[0] > 1 and np.all(team[1:] == 1)
    assert T < 100.0 and load.max() <= T * (1 + 1e-12)
