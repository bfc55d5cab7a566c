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
