"""The oracle's mm_HI=False (arrowhead) regime against independent restatements (CPU).

Pins, each independent of the oracle's own code path:
  - M^cm: its rotation and stresslet rows equal the corresponding blocks of the mm_HI grand mobility
    (the same mobility_AB, pinned by test_oracle.py), its translation rows are zero, and M^cc starts
    as the colloid self block;
  - the Schur floor (src/slip_vel.cpp:81-203) against numpy's eigh;
  - the noise (src/slip_vel.cpp:205-349): its linear map L satisfies L L^T = the arrowhead mobility
    with the floored M^cc;
  - the monomer cell list (src/cell_list.cpp) against the binning formula, and its WCA forces
    against a brute-force sum over all monomer pairs;
  - the saddle solves against a dense direct solve.
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from polystokes_brownian_amd import configs


def arrow_case(nd=30, beta=0.1, L=6.0, seed=5, mono_ev=0, clearance=1.5, **over):
    w = configs.workload("arrow_thermal")
    w.update(dict(Npoly=nd, Nm=2 * nd, Np=2 * nd + 1, beta=beta, L=L, ic_seed=seed, mono_ev=mono_ev))
    w.update(over)
    r0 = w["r0"]
    x0 = configs.dumbbell_configuration(nd, r0, w["kbond"], w["kT"], L if L else 10.0, seed, clearance=clearance)
    if L is None:
        w["L"] = None
    return w, x0


def make(w, x0, **over):
    pd = configs.params_dict(w)
    pd.update(over)
    o = Oracle(pd)
    o.initial_configuration(x0)
    return o, pd


def test_arrow_rejects_several_colloids():
    w = configs.small(2, 3, 4, 8.0, seed=1)
    pd = dict(configs.params_dict(w), mm_HI=0)
    with pytest.raises(Exception):
        Oracle(pd)


def test_arrow_vlist_has_only_monomer_colloid_pairs():
    w, x0 = arrow_case(nd=40, clearance=1.0)
    o, _ = make(w, x0)
    o.check_dist()
    rows, parts, _ = o.vlist()
    Nm = 2 * 40
    assert np.all(parts >= Nm)
    x = x0.reshape(-1, 3)
    d = x[:Nm] - x[Nm]
    d -= w["L"] * np.rint(d / w["L"])
    r = np.sqrt((d * d).sum(1))
    want = np.nonzero(r < np.float32(1.5 + w["beta"]))[0]
    np.testing.assert_array_equal(np.sort(rows), want)


def test_arrow_mcm_blocks_equal_the_dense_regime():
    w, x0 = arrow_case(nd=25)
    oa, _ = make(w, x0)
    od, _ = make(w, x0, mm_HI=1)
    for o in (oa, od):
        o.check_dist()
        o.mob()
    Ma, Md = oa.mob_dense(), od.mob_dense()
    nm3 = 3 * 50
    # rotation (rows nm3+3..nm3+5) and stresslet (nm3+6..nm3+10) rows over the monomer columns
    np.testing.assert_array_equal(Ma[nm3 + 3: nm3 + 11, :nm3], Md[nm3 + 3: nm3 + 11, :nm3])
    np.testing.assert_array_equal(Ma[:nm3, nm3 + 3: nm3 + 11], Md[:nm3, nm3 + 3: nm3 + 11])
    assert not np.any(Ma[nm3: nm3 + 3, :nm3])          # translation rows of M^cm are zero
    np.testing.assert_array_equal(Ma[:nm3, :nm3], np.eye(nm3) / w["beta"])
    # before any sync M^cc is the colloid self block
    np.testing.assert_array_equal(Ma[nm3:, nm3:], Md[nm3:, nm3:])


def schur_expected(M, nm3, beta):
    Mcm = M[nm3:, :nm3]
    base = M[nm3:, nm3:]
    S = base - beta * Mcm @ Mcm.T
    lam, Q = np.linalg.eigh(S)
    delta = np.where(lam < 0.1, 0.1 - lam, 0.0)
    return base + Q @ np.diag(delta) @ Q.T, lam


@pytest.mark.parametrize("beta,clearance,floored", [(0.001, 1.5, False), (0.5, 1.05, True)])
def test_arrow_schur_floor(beta, clearance, floored):
    w, x0 = arrow_case(nd=60, beta=beta, clearance=clearance, L=4.0)
    o, _ = make(w, x0)
    o.check_dist()
    o.mob()
    M0 = o.mob_dense()
    nm3 = 3 * 120
    want, lam = schur_expected(M0, nm3, beta)
    assert (lam.min() < 0.1) == floored
    o.schur_sync()
    mcc, Q, laml = o.schur()
    np.testing.assert_allclose(mcc, want, rtol=0, atol=1e-12 * np.abs(want).max())
    if not floored:
        np.testing.assert_array_equal(mcc, M0[nm3:, nm3:])
    # the floored Schur complement has no eigenvalue below 0.1, and Q diag(lam) Q^T reproduces it
    Mcm = M0[nm3:, :nm3]
    Sf = mcc - beta * Mcm @ Mcm.T
    assert np.linalg.eigvalsh(Sf).min() > 0.1 - 1e-10
    np.testing.assert_allclose(Q @ np.diag(laml) @ Q.T, Sf, atol=1e-11 * np.abs(Sf).max())
    # the apply uses the floored block
    M1 = o.mob_dense()
    np.testing.assert_array_equal(M1[nm3:, nm3:], mcc)


@pytest.mark.parametrize("beta,clearance", [(0.1, 1.5), (0.5, 1.05)])
def test_arrow_noise_is_a_cholesky_factor(beta, clearance):
    w, x0 = arrow_case(nd=20, beta=beta, clearance=clearance, L=4.0)
    o, _ = make(w, x0)
    o.check_dist()
    o.mob()
    o.schur_sync()
    M = o.mob_dense()
    n = M.shape[0]
    Lc = np.stack([o.sqrt_apply(np.eye(n)[k])[0] for k in range(n)], 1)
    np.testing.assert_allclose(Lc @ Lc.T, M, rtol=0, atol=1e-11 * np.abs(M).max())
    # lower block-triangular: the monomer part depends on the monomer noise only
    nm3 = n - 11
    np.testing.assert_array_equal(Lc[:nm3, nm3:], 0.0)
    np.testing.assert_allclose(Lc[:nm3, :nm3], np.eye(nm3) / np.sqrt(beta), rtol=1e-15)


def test_arrow_apply_and_saddle_against_dense():
    w, x0 = arrow_case(nd=30, beta=0.5, clearance=1.05, L=4.0)
    o, _ = make(w, x0)
    o.check_dist()
    o.mob()
    o.schur_sync()
    M = o.mob_dense()
    n = M.shape[0]
    nu = n - 5
    rng = np.random.default_rng(3)
    f = rng.normal(size=n)
    np.testing.assert_allclose(o.mob_mult(f), M @ f, rtol=1e-13, atol=1e-13 * np.abs(M @ f).max())
    A = np.zeros((n + nu, n + nu))
    A[:n, :n] = M
    A[:nu, n:] = -np.eye(nu)
    A[n:, :nu] = -np.eye(nu)
    xs = rng.normal(size=n + nu)
    np.testing.assert_allclose(o.saddle_mult(xs), A @ xs, rtol=1e-13, atol=1e-13 * np.abs(A @ xs).max())
    d = o.saddle_diag()
    np.testing.assert_array_equal(d[:n], np.diag(M))
    np.testing.assert_array_equal(d[n:], 1.0)


@pytest.mark.parametrize("solver", [0, 1])
def test_arrow_solves_against_direct(solver):
    w, x0 = arrow_case(nd=30, beta=0.5, clearance=1.05, L=4.0)
    o, _ = make(w, x0, ksp_rtol=1e-12, solver=solver)
    o.check_dist()
    o.mob()
    o.schur_sync()
    M = o.mob_dense()
    n = M.shape[0]
    nu = n - 5
    A = np.zeros((n + nu, n + nu))
    A[:n, :n] = M
    A[:nu, n:] = -np.eye(nu)
    A[n:, :nu] = -np.eye(nu)
    b = np.random.default_rng(4).normal(size=n + nu)
    x, info, rc = o.saddle_solve(b)
    assert rc == 0
    want = np.linalg.solve(A, b)
    assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-9


def test_arrow_cell_list_binning():
    for L, periodic in ((6.0, True), (0.5, False)):
        w, x0 = arrow_case(nd=40, beta=0.1, L=6.0, mono_ev=1)
        if L != 6.0:
            w["L"] = L       # a box below 3 cells per axis: one cell
        o, _ = make(w, x0)
        cell, dims, size, org, per = o.cell_list()
        assert per == periodic
        rcell = float(np.float32(0.05 + 2 * 0.1))
        if periodic:
            assert list(dims) == [int(np.floor(L / rcell))] * 3
            x = x0.reshape(-1, 3)[:80]
            c = np.clip(np.floor((x - org) / size).astype(int), 0, dims - 1)
            np.testing.assert_array_equal(cell, c[:, 0] + dims[0] * (c[:, 1] + dims[1] * c[:, 2]))
        else:
            assert list(dims) == [1, 1, 1] and np.all(cell == 0)
    # unbounded: the monomers' bounding box
    w, x0 = arrow_case(nd=40, beta=0.1, L=None, mono_ev=1)
    o, _ = make(w, x0)
    cell, dims, size, org, per = o.cell_list()
    x = x0.reshape(-1, 3)[:80]
    np.testing.assert_array_equal(org, x.min(0))
    assert not per


@pytest.mark.parametrize("L", [6.0, None, 0.5])
def test_arrow_wca_against_brute_force(L):
    # a dense cloud so that many monomer pairs fall inside the WCA cutoff
    nd, beta = 150, 0.1
    w, x0 = arrow_case(nd=nd, beta=beta, L=6.0, mono_ev=1, clearance=1.2)
    rng = np.random.default_rng(8)
    x = x0.reshape(-1, 3).copy()
    x[: 2 * nd] = rng.uniform(-1.2, 1.2, size=(2 * nd, 3)) + np.array([2.4, 0, 0])
    x0 = x.reshape(-1)
    if L != 6.0:
        w["L"] = L
    on, _ = make(w, x0)
    off, _ = make(w, x0, mono_ev=0)
    for o in (on, off):
        o.check_dist()
    diff = on.rhs() - off.rhs()
    n = 3 * 2 * nd + 11
    f = diff[n: n + 6 * nd].reshape(-1, 3)
    sig = float(np.float32(2 * beta))
    rcut = float(np.float32(2 ** (1 / 6) * sig))
    eps = w["epsilon"]
    want = np.zeros((2 * nd, 3))
    xm = x[: 2 * nd]
    Lb = w["L"]
    npairs = 0
    for i in range(2 * nd):
        d = xm[i] - xm[i + 1:]
        if Lb:
            d -= Lb * np.rint(d / Lb)
        r = np.sqrt((d * d).sum(1))
        sel = r < rcut
        npairs += sel.sum()
        rr = r[sel]
        fm = 24 * eps / rr * (2 * (sig / rr) ** 12 - (sig / rr) ** 6)
        fv = fm[:, None] * d[sel] / rr[:, None]
        want[i] -= fv.sum(0)
        want[i + 1:][sel] += fv
    assert npairs > 50
    np.testing.assert_allclose(f, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_arrow_run_is_deterministic_and_moves():
    w, x0 = arrow_case(nd=20, beta=0.1, L=6.0)
    xs = []
    for _ in range(2):
        o, _ = make(w, x0)
        o.run(3)
        xs.append(o.state()[0])
    np.testing.assert_array_equal(xs[0], xs[1])
    assert np.abs(xs[0] - x0).max() > 1e-6
