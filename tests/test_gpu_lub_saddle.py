"""Lubrication in the saddle (parameter lub = 1): GPU saddle operator and saddle solves against a dense
numpy solve of [[M, B], [B^T, -R]], with M the oracle's dense mobility (oracle/psoracle.cpp, the
reference's mob() assembly) and R the force/torque-velocity block of the oracle's dense pair
lubrication (oracle/lub.py) at the same positions.  B = [-I_nu; 0] (src/arrays.cpp:349-408); the
reference leaves the lower-right block as a TODO (src/saddle.cpp:53).

Tolerances: the operator to 1e-12 (FP64 both sides); the block solve to 1e-9 (its colloid MINRES stops
at 1e-12 of the preconditioned residual); the reference's MINRES at ksp_rtol 1e-12 to 1e-8.
"""
import numpy as np
import pytest

from polystokes_brownian_amd import configs
from polystokes_brownian_amd.polystokes import Context, PsError

from test_gpu_lubrication import CASES, clustered_probes

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def setup(name, **over):
    Nc, Npoly, Nmono, L, seed = CASES[name]
    probes = clustered_probes(Nc, L, seed)
    Lgen = L if L is not None else 40.0
    w = configs.small(Nc, Npoly, Nmono, Lgen, seed=seed, probes=probes)
    if L is None:
        w["L"] = None
    x0 = configs.make_configuration(Nc, Npoly, Nmono, Lgen, w["r0"], seed, probe_positions=probes)
    pd = dict(configs.params_dict(w), **over)
    return pd, x0, Npoly * Nmono, Nc, (None if L is None else [L] * 3)


def dense_saddle(pd, x0, Nm, Nc, box, rcut=4.0):
    """[[M, B], [B^T, -R]] from the oracle's dense M and dense lubrication."""
    from oracle import lub as olub
    from oracle.pyoracle import Oracle
    o = Oracle(dict(pd, lub=0))
    o.initial_configuration(x0)
    o.check_dist()
    o.mob()
    M = o.mob_dense()
    n, nu = 3 * Nm + 11 * Nc, 3 * Nm + 6 * Nc
    R = olub.resistance(x0.reshape(-1, 3)[Nm:], rcut, box=box)[: 6 * Nc, : 6 * Nc]
    A = np.zeros((n + nu, n + nu))
    A[:n, :n] = M
    A[:nu, n:] = -np.eye(nu)
    A[n:, :nu] = -np.eye(nu)
    A[n + 3 * Nm:, n + 3 * Nm:] = -R
    return A, R


def gpu_ctx(pd, x0):
    g = Context(pd)
    g.initial_configuration(x0)
    g.check_dist()
    g.mob()
    return g


@pytest.mark.parametrize("name", list(CASES))
def test_lubricated_saddle_operator(name):
    pd, x0, Nm, Nc, box = setup(name, lub=1, solver=1)
    A, R = dense_saddle(pd, x0, Nm, Nc, box)
    assert np.abs(R).max() > 0.1   # near-contact pairs present
    g = gpu_ctx(pd, x0)
    rng = np.random.default_rng(5)
    for _ in range(3):
        x = rng.normal(size=A.shape[0])
        y = g.saddle_mult(x)
        print("case", name, "operator rel", rel(y, A @ x), flush=True)
        assert rel(y, A @ x) < 1e-12
    g.close()
    # lub = 0 keeps the reference's operator
    g0 = gpu_ctx(dict(pd, lub=0), x0)
    x = rng.normal(size=A.shape[0])
    A0 = A.copy()
    A0[-6 * Nc:, -6 * Nc:] = 0.0
    assert rel(g0.saddle_mult(x), A0 @ x) < 1e-12
    g0.close()


@pytest.mark.parametrize("name,solver,tol", [("box16", 1, 1e-9), ("box_tight", 1, 1e-9), ("nobox", 1, 1e-9),
                                             ("box16", 0, 1e-8)])
def test_lubricated_saddle_solve_matches_dense(name, solver, tol):
    pd, x0, Nm, Nc, box = setup(name, lub=1, solver=solver, ksp_rtol=1e-12)
    A, _ = dense_saddle(pd, x0, Nm, Nc, box)
    g = gpu_ctx(pd, x0)
    rng = np.random.default_rng(9)
    for k in range(2):
        b = rng.normal(size=A.shape[0])
        want = np.linalg.solve(A, b)
        x, info, rc = g.saddle_solve(b)
        print("case", name, "solver", solver, "rhs", k, "rel", rel(x, want), "its", info.iterations, flush=True)
        assert rc == 0
        assert rel(x, want) < tol
    if solver == 1:
        c = g.counters()
        assert c["stresslet_iterations"] > 0 and c["stresslet_capped"] == 0
    g.close()


def test_lubrication_changes_the_solution_and_off_matches_reference():
    """The lubricated solve differs from the reference's (lub = 0) where pairs are close, and lub = 0
    is the unlubricated dense solve."""
    pd, x0, Nm, Nc, box = setup("box16", solver=1)
    A, _ = dense_saddle(pd, x0, Nm, Nc, box)
    A0 = A.copy()
    A0[-6 * Nc:, -6 * Nc:] = 0.0
    b = np.random.default_rng(4).normal(size=A.shape[0])
    g0 = gpu_ctx(dict(pd, lub=0), x0)
    x0s, _, _ = g0.saddle_solve(b)
    g0.close()
    g1 = gpu_ctx(dict(pd, lub=1), x0)
    x1s, _, _ = g1.saddle_solve(b)
    g1.close()
    assert rel(x0s, np.linalg.solve(A0, b)) < 1e-9
    assert rel(x1s, x0s) > 1e-3


def test_lubricated_steps_run():
    """Steps of a lubricated run (block solver, pairs rebuilt at every mob()): finite, and not graphed."""
    pd, x0, Nm, Nc, box = setup("box16", lub=1, solver=1)
    g = Context(pd)
    g.initial_configuration(x0)
    g.run(3)
    x = g.state()[0]
    assert np.all(np.isfinite(x))
    c = g.counters()
    assert c["mob_calls"] >= 12 and c["stresslet_iterations"] > 0
    g.close()


def test_lubrication_parameter_errors():
    pd, x0, Nm, Nc, box = setup("box16")
    with pytest.raises(PsError):
        Context(dict(pd, lub=2))
    with pytest.raises(PsError):
        Context(dict(pd, lub=1, lub_rcut=0.0))
    g = Context(dict(pd, lub=1, lub_rcut=4.5))   # beyond the last node: refused at the first mob()
    g.initial_configuration(x0)
    g.check_dist()
    with pytest.raises(PsError):
        g.mob()
    g.close()


def test_cell_list_pairs_match_all_pairs():
    """The cell-list pair build against the oracle's all-pairs scan on many colloids (grids of many
    cells, periodic and open)."""
    from oracle import lub as olub
    from test_gpu_lubrication import build
    for Nc, L, seed in ((300, 30.0, 1), (200, None, 2), (120, 11.0, 3)):
        CASES[f"_many{seed}"] = (Nc, 1, 4, L, seed)
        g, xc, bx, npairs = build(f"_many{seed}")
        R = olub.resistance(xc, 4.0, box=bx)
        want = sum(1 for i in range(Nc) for j in range(i + 1, Nc) if np.any(R[3 * i: 3 * i + 3, 3 * j: 3 * j + 3]))
        uw = np.random.default_rng(seed).normal(size=6 * Nc)
        print("cells", Nc, L, "pairs", npairs, want, flush=True)
        assert npairs == want
        assert rel(g.lub_apply(uw), R[:, : 6 * Nc] @ uw) < 1e-12
        g.close()
        del CASES[f"_many{seed}"]
