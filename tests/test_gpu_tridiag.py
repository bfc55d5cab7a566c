"""The device eigensolver of the Lanczos square root (k_lz_ritz: the two-wave QL of tql2.h) on one
tridiagonal at a time, through ps_tridiag_sqrt, against the oracle's tridiag_sqrt (pinned against
numpy's eigh in test_oracle.py) and against eigh itself: random, indefinite, Lanczos-built and widely
graded tridiagonals at every order from 1 to 64, and the invariant-subspace cut."""
import numpy as np
import pytest

from oracle.pyoracle import tridiag_sqrt as oracle_tridiag_sqrt
from polystokes_brownian_amd.polystokes import PsError, tridiag_sqrt


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def eigh_weights(a, b, beta0):
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    w, U = np.linalg.eigh(T)
    return beta0 * (U * np.sqrt(np.maximum(w, 0.0))) @ U[0], int((w < 0).sum())


def lanczos_tridiagonal(n, m, seed, indefinite):
    """alpha, beta of m Lanczos steps (full reorthogonalisation) on a random symmetric n x n matrix
    whose eigenvalues spread over four decades (both signs when indefinite)."""
    rng = np.random.default_rng(seed)
    Qm, _ = np.linalg.qr(rng.normal(size=(n, n)))
    lam = np.logspace(-2, 2, n)
    if indefinite:
        lam *= np.where(rng.random(n) < 0.3, -1.0, 1.0)
    A = (Qm * lam) @ Qm.T
    q = rng.normal(size=n)
    Q = [q / np.linalg.norm(q)]
    al, be = [], []
    for j in range(m):
        v = A @ Q[-1]
        for _ in range(2):
            v -= np.array(Q).T @ (np.array(Q) @ v)
        al.append(float(Q[-1] @ (A @ Q[-1])))
        bj = float(np.linalg.norm(v))
        if j < m - 1:
            be.append(bj)
            Q.append(v / bj)
    return np.array(al), np.array(be)


@pytest.mark.gpu
@pytest.mark.parametrize("m", [1, 2, 3, 7, 16, 31, 32, 33, 48, 63, 64])
def test_random_tridiagonals(m):
    rng = np.random.default_rng(100 + m)
    for indefinite in (False, True):
        a = rng.uniform(-3.0, 3.0, m) if indefinite else rng.uniform(0.5, 3.0, m)
        b = rng.uniform(0.05, 0.5, m - 1)
        g, mu, nneg = tridiag_sqrt(a, b, 2.5)
        go, nno = oracle_tridiag_sqrt(a, b, 2.5)
        ge, nne = eigh_weights(a, b, 2.5)
        print("m", m, "indefinite", indefinite, "vs oracle", rel(g, go), "vs eigh", rel(g, ge), flush=True)
        assert mu == m and nneg == nno == nne
        assert rel(g, go) < 1e-12 and rel(g, ge) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("m,indefinite", [(32, False), (32, True), (64, False), (64, True)])
def test_lanczos_tridiagonals(m, indefinite):
    a, b = lanczos_tridiagonal(300, m, 7 + m, indefinite)
    g, mu, nneg = tridiag_sqrt(a, b, 1.7)
    go, nno = oracle_tridiag_sqrt(a, b, 1.7)
    ge, nne = eigh_weights(a, b, 1.7)
    print("lanczos m", m, "indefinite", indefinite, "vs oracle", rel(g, go), "vs eigh", rel(g, ge), flush=True)
    assert mu == m and nneg == nno == nne
    assert rel(g, go) < 1e-11 and rel(g, ge) < 1e-11


@pytest.mark.gpu
def test_graded_tridiagonal():
    # diagonal over twelve decades with small couplings: deflation at very different scales
    m = 40
    a = np.logspace(-6, 6, m)
    b = 1e-3 * np.sqrt(a[:-1] * a[1:])
    g, mu, nneg = tridiag_sqrt(a, b, 1.0)
    ge, nne = eigh_weights(a, b, 1.0)
    go, _ = oracle_tridiag_sqrt(a, b, 1.0)
    assert mu == m and nneg == nne == 0
    assert rel(g, ge) < 1e-11 and rel(g, go) < 1e-11


@pytest.mark.gpu
def test_invariant_subspace_cut():
    # beta_j <= 1e-12 beta0 ends the recurrence: only the leading j + 1 block is solved
    rng = np.random.default_rng(3)
    a = rng.uniform(0.5, 3.0, 20)
    b = rng.uniform(0.05, 0.5, 19)
    b[8] = 0.0
    g, mu, nneg = tridiag_sqrt(a, b, 2.0)
    go, _ = oracle_tridiag_sqrt(a[:9], b[:8], 2.0)
    assert mu == 9 and nneg == 0
    assert rel(g[:9], go) < 1e-12 and not g[9:].any()


def test_argument_checks():
    # rejected on the host, before any device work
    for m in (0, 65):
        with pytest.raises(PsError):
            tridiag_sqrt(np.ones(m), np.ones(max(m - 1, 0)), 1.0)
    with pytest.raises(PsError):
        tridiag_sqrt(np.ones(3), np.ones(2), -1.0)
