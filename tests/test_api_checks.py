"""The Python front end rejects arrays of the wrong length before any call into the C ABI.

The library trusts the sizes it is given (it reads n, ns or 3 Np doubles, and some entry points
write ns or nrhs*n doubles back), so a short array would be read or written past its end.  These
tests run without a GPU: the wrappers are exercised on objects whose C handle is never used, and a
stand-in for the loaded library fails the test if any entry point is reached.
"""
import numpy as np
import pytest

from polystokes_brownian_amd import polystokes as P


class _NoLib:
    def __getattr__(self, name):
        raise AssertionError(f"{name} was called with an array of the wrong size")


@pytest.fixture
def no_lib(monkeypatch):
    monkeypatch.setattr(P, "_lib", _NoLib())
    monkeypatch.setattr(P, "lib", lambda: P._lib)


def _ctx(Nm=40, Nc=3):
    c = object.__new__(P.Context)
    c.h = None
    c.Np, c.Nm, c.Nc = Nm + Nc, Nm, Nc
    c.n = 3 * Nm + 11 * Nc
    c.nu = 3 * Nm + 6 * Nc
    c.ns = c.n + c.nu
    return c


@pytest.mark.parametrize("delta", [-1, 1, -30])
def test_context_rejects_wrong_lengths(no_lib, delta):
    c = _ctx()
    cases = [
        (c.mob_mult, c.n),
        (c.saddle_mult, c.ns),
        (c.saddle_solve, c.ns),
        (c.sqrt_apply, c.n),
        (c.set_positions, 3 * c.Np),
        (c.initial_configuration, 3 * c.Np),
        (c.lub_apply, 6 * c.Nc),
    ]
    for fn, n in cases:
        with pytest.raises(ValueError):
            fn(np.zeros(n + delta))
    # the warm-start vector of a solve, the quaternions and the strain are checked too
    with pytest.raises(ValueError):
        c.saddle_solve(np.zeros(c.ns), x0=np.zeros(c.ns + delta))
    with pytest.raises(ValueError):
        c.initial_configuration(np.zeros(3 * c.Np), np.zeros(4 * c.Nc + delta))
    with pytest.raises(ValueError):
        c.lub_apply(np.zeros(6 * c.Nc), einf=np.zeros(5 + delta))


def test_dense_mult_shape_checked(no_lib):
    c = _ctx()
    for shape in [(1, c.n - 1), (2, c.n + 1), (17, c.n), (0, c.n)]:
        with pytest.raises(ValueError):
            c.dense_mult(np.zeros(shape))


def _batch(sizes):
    b = object.__new__(P.Batch)
    b.h = None
    b.R = len(sizes)
    b.sizes = []
    for Nc, Nm in sizes:
        b.sizes.append(dict(Np=Nm + Nc, Nc=Nc, Nm=Nm, n=3 * Nm + 11 * Nc, nu=3 * Nm + 6 * Nc,
                            ns=6 * Nm + 17 * Nc))
    return b


def test_batch_rejects_wrong_lengths(no_lib):
    b = _batch([(2, 20), (5, 60)])
    with pytest.raises(ValueError):
        b.set_state(1, np.zeros(3 * 65 - 3))
    with pytest.raises(ValueError):
        b.set_state(0, np.zeros(3 * 22), q0=np.zeros(7))
    with pytest.raises(ValueError):
        b.set_state(2, np.zeros(3 * 22))          # no such replica
    with pytest.raises(ValueError):
        b.op(0, [np.zeros(b.sizes[0]["n"]), np.zeros(b.sizes[1]["n"] - 1)])
    with pytest.raises(ValueError):
        b.op(2, [np.zeros(b.sizes[0]["ns"])])     # one vector per replica
    with pytest.raises(ValueError):
        b.op(9, None)


def test_checked_vectors_pass_through():
    v = P._vec([1, 2, 3], 3, "v")
    assert v.dtype == np.float64 and v.flags["C_CONTIGUOUS"] and v.tolist() == [1.0, 2.0, 3.0]
    assert P._vec(np.ones((2, 3)), 6, "v").shape == (6,)
