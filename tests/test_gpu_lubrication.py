"""GPU parity of the pair lubrication operator (csrc/lub.hip, through the C ABI) against the oracle's
dense restatement (oracle/lub.py) on the same colloid positions.

Tolerance: FP64, 1e-12 relative (both sides interpolate the same tables with the same monotone
cubic; the oracle through scipy, the kernel through its own Hermite evaluation).
"""
import numpy as np
import pytest

from polystokes_brownian_amd import configs
from polystokes_brownian_amd.polystokes import Context, PsError

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def clustered_probes(Nc, L, seed, lo=1.97, hi=4.3):
    """Colloid centres built by attaching each new one to an earlier one at a distance in [lo, hi):
    near-contact pairs, overlaps below the first tabulated distance, pairs past the cutoff."""
    rng = np.random.default_rng(seed)
    p = [rng.uniform(-1, 1, 3)]
    while len(p) < Nc:
        v = rng.normal(size=3)
        c = p[rng.integers(len(p))] + v / np.linalg.norm(v) * rng.uniform(lo, hi)
        d = np.array(p) - c
        if L is not None:
            d -= L * np.rint(d / L)
        if np.all(np.linalg.norm(d, axis=1) > 1.9):
            p.append(c)
    p = np.array(p)
    if L is not None:
        p = p - L * np.floor((p + L / 2) / L)      # into [-L/2, L/2): some pairs now meet across the box
    return p


CASES = {
    # name: (Nc, Npoly, Nmono, L, seed)
    "box16": (24, 2, 5, 16.0, 3),
    "box_tight": (30, 1, 4, 9.0, 5),
    "nobox": (12, 2, 4, None, 7),
}


def build(name, rcut=4.0):
    Nc, Npoly, Nmono, L, seed = CASES[name]
    probes = clustered_probes(Nc, L, seed)
    Lgen = L if L is not None else 40.0
    w = configs.small(Nc, Npoly, Nmono, Lgen, seed=seed, probes=probes)
    if L is None:
        w["L"] = None
    x0 = configs.make_configuration(Nc, Npoly, Nmono, Lgen, w["r0"], seed, probe_positions=probes)
    pd = configs.params_dict(w)
    g = Context(pd)
    g.initial_configuration(x0)
    g.check_dist()
    g.mob()
    g.lub_set_tables()
    npairs = g.lub_build(rcut)
    xc = x0.reshape(-1, 3)[Npoly * Nmono:]
    box = None if L is None else [L] * 3
    return g, xc, box, npairs


@pytest.mark.parametrize("name", list(CASES))
def test_lub_apply_matches_oracle(name):
    from oracle import lub as olub
    g, xc, box, npairs = build(name)
    Nc = len(xc)
    R = olub.resistance(xc, 4.0, box=box)
    # pair count: pairs with a nonzero coupling block
    want_pairs = sum(1 for i in range(Nc) for j in range(i + 1, Nc)
                     if np.any(R[3 * i: 3 * i + 3, 3 * j: 3 * j + 3]))
    assert npairs == want_pairs and npairs > Nc // 2
    rng = np.random.default_rng(11)
    for k in range(3):
        uw = rng.normal(size=6 * Nc)
        ev = None if k == 0 else rng.normal(size=5)
        got = g.lub_apply(uw, ev)
        want = R @ np.concatenate([uw, np.zeros(5) if ev is None else ev])
        assert rel(got, want) < 1e-12, (name, k, rel(got, want))
    # the columns one by one (every block, including the strain columns)
    cols = rng.choice(6 * Nc, size=min(12, 6 * Nc), replace=False)
    for c in cols:
        uw = np.zeros(6 * Nc)
        uw[c] = 1.0
        np.testing.assert_allclose(g.lub_apply(uw), R[:, c], rtol=1e-12, atol=1e-12 * np.abs(R).max())
    for c in range(5):
        ev = np.zeros(5)
        ev[c] = 1.0
        np.testing.assert_allclose(g.lub_apply(np.zeros(6 * Nc), ev), R[:, 6 * Nc + c], rtol=1e-12,
                                   atol=1e-12 * np.abs(R).max())
    g.close()


def test_lub_cutoff_and_rebuild():
    from oracle import lub as olub
    g, xc, box, npairs4 = build("box16", rcut=4.0)
    Nc = len(xc)
    uw = np.random.default_rng(2).normal(size=6 * Nc)
    ev = np.random.default_rng(3).normal(size=5)
    n3 = g.lub_build(3.0)
    assert 0 < n3 < npairs4
    want = olub.apply(xc, uw, ev, rcut=3.0, box=box)
    assert rel(g.lub_apply(uw, ev), want) < 1e-12
    # back to 4.0 on the same context
    assert g.lub_build(4.0) == npairs4
    assert rel(g.lub_apply(uw, ev), olub.apply(xc, uw, ev, rcut=4.0, box=box)) < 1e-12
    g.close()


def test_lub_errors():
    Nc, Npoly, Nmono, L, seed = CASES["box16"]
    w = configs.small(Nc, Npoly, Nmono, L, seed=seed, probes=clustered_probes(Nc, L, seed))
    g = Context(configs.params_dict(w))
    g.initial_configuration(configs.positions(w))
    with pytest.raises(PsError):
        g.lub_apply(np.zeros(6 * Nc))           # not built
    g.check_dist()
    g.mob()
    with pytest.raises(PsError):
        g.lub_build(4.0)                        # no tables
    g.lub_set_tables()
    with pytest.raises(PsError):
        g.lub_build(4.5)                        # beyond the last node
    with pytest.raises(PsError):
        g.lub_build(0.0)
    assert g.lub_build(4.0) >= 0
    g.close()
