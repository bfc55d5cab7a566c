"""Lubrication tables, their interpolant and the oracle's pair resistance (CPU).

The reference ships lubdat/*.dat but no code that reads them, so nothing in it fixes what the columns
mean.  These tests pin the meaning from the mathematics instead:
  - X11C, X12C against the exact bispherical solution for two spheres rotating about their line of
    centres (Jeffery 1915): sinh^3(a) sum_n (-+1)^(n+1) csch^3(n a), cosh a = r/2;
  - X11A, Y11A and X11A - X12A against the near-contact forms of Jeffrey & Onishi (1984);
  - the sign convention of the odd functions (X11G from the exact derivative table r2bghd.dat, Y11B,
    Y12B) and the units against the singular terms of the same asymptotics;
  - r2bm.dat as M11 + M12 from its 1/xi singularity.
"Far field" below is the inverse of the two-sphere Rotne-Prager-Yamakawa FTS mobility: the oracle's
mobility of two colloids, mapped to Kim & Karrila's variables, with the rotation-force coupling made odd
in the separation (the reference's colloid-colloid W-F and U-T blocks are even in it; its M[W2, F1]
block is the physical one, M[W1, F2] is mirrored).
"""
import numpy as np
import pytest
from scipy.interpolate import PchipInterpolator

from oracle import lub as olub
from oracle.pyoracle import Oracle
from polystokes_brownian_amd import configs
from polystokes_brownian_amd import lubrication as plub

TABLES = olub.load_tables()
ABC, GH, MM = TABLES
GHD = np.loadtxt(olub.LUBDAT + "/r2bghd.dat")


# ------------------------------------------------------------------------------------- far field
def _far_field_x(r):
    """22 x 22 far-field resistance of two spheres separated by (r, 0, 0), Kim & Karrila's variables."""
    pd = dict(configs.params_dict(configs.small(2, 0, 1, 30.0)), Nm=0, Nc=2, Np=2, Npoly=0, box=None)
    o = Oracle(pd)
    o.initial_configuration(np.array([0.0, 0.0, 0.0, r, 0.0, 0.0]))
    o.check_dist()
    o.mob()
    M = o.mob_dense().copy()
    W1, W2, F1, F2 = slice(6, 9), slice(9, 12), slice(0, 3), slice(3, 6)
    M[W1, F2] = -M[W2, F1]
    M[F2, W1] = M[W1, F2].T
    R = np.linalg.inv(M)
    D = np.ones(22)
    D[12:] = -1.0           # the reference's stresslet / strain variables carry the opposite sign
    return D[:, None] * R * D[None, :]


def _eflat(E):
    return np.array([E[0, 0] - E[2, 2], 2 * E[0, 1], 2 * E[0, 2], 2 * E[1, 2], E[1, 1] - E[2, 2]])


def _smat(s):
    return np.array([[s[0], s[1], s[2]], [s[1], s[4], s[3]], [s[2], s[3], -s[0] - s[4]]])


def _sflat(S):
    return np.array([S[0, 0], S[0, 1], S[0, 2], S[1, 2], S[1, 1]])


def rotation_to(e):
    """Proper rotation Q with Q (1, 0, 0) = e."""
    x = np.array([1.0, 0.0, 0.0])
    v, c = np.cross(x, e), float(x @ e)
    if np.linalg.norm(v) < 1e-12:
        return np.eye(3) if c > 0 else np.diag([-1.0, -1.0, 1.0])
    K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + K + K @ K / (1 + c)


def far_field(sep):
    """22 x 22 two-sphere far-field resistance in Kim & Karrila's variables (6 pi eta = 1): rows
    [F1 F2 T1 T2 S1 S2] (S as SV), columns [U1 U2 W1 W2 -E_1 -E_2] (E as EV), i.e. (F, T, S) on the
    fluid for (U, W) and minus the strain at each sphere.  Built along +x, where the reference's
    colloid-colloid blocks are isotropic once the rotation-force coupling is made odd (its rotation-
    rotation cross block carries (x12c + y12c) e_i e_j off the diagonal, which vanishes on an axis),
    then rotated to the separation."""
    sep = np.asarray(sep, dtype=float)
    r = np.linalg.norm(sep)
    Q = rotation_to(sep / r)
    TE = np.stack([_eflat(Q @ olub.EBASIS[c] @ Q.T) for c in range(5)], 1)
    TS = np.stack([_sflat(Q @ _smat(np.eye(5)[c]) @ Q.T) for c in range(5)], 1)
    Tin = np.zeros((22, 22))
    Tout = np.zeros((22, 22))
    for k in range(4):
        Tin[3 * k: 3 * k + 3, 3 * k: 3 * k + 3] = Tout[3 * k: 3 * k + 3, 3 * k: 3 * k + 3] = Q
    for k in range(2):
        Tin[12 + 5 * k: 17 + 5 * k, 12 + 5 * k: 17 + 5 * k] = TE
        Tout[12 + 5 * k: 17 + 5 * k, 12 + 5 * k: 17 + 5 * k] = TS
    return Tout @ _far_field_x(r) @ np.linalg.inv(Tin)


def far_funcs(r):
    """Far-field scalar functions for e = +x from sphere 1 to sphere 2 (6 pi eta = 1 units)."""
    R = far_field([r, 0.0, 0.0])
    f = dict(X11A=R[0, 0], X12A=R[0, 3], Y11A=R[1, 1], Y12A=R[1, 4], Y11B=R[7, 2], Y12B=R[7, 5],
             X11C=R[6, 6], X12C=R[6, 9], Y11C=R[7, 7], Y12C=R[7, 10],
             X11G=R[12, 0] / (2 / 3), X12G=R[12, 3] / (2 / 3))
    # common-strain stresslet response M11 + M12 on the d0 basis: S11 for E = (2/3, -1/3, -1/3)
    # diag -- EV = (1, 0, 0, 0, 0) -- is (2/3) XM for e = +x
    f["XM_sum"] = (R[12, 12] + R[12, 17]) / (2 / 3)
    return f


def jeffery_c(r):
    """Exact X11C, X12C of two equal spheres (normalised by 8 pi eta a^3)."""
    a = np.arccosh(r / 2)
    n = np.arange(1, min(600, int(300 / a)) + 1)     # later terms are below 1e-300
    c = (np.sinh(a) / np.sinh(n * a)) ** 3
    co, counter = np.sum(c * (-1.0) ** (n + 1)), np.sum(c)
    return 0.5 * (co + counter), 0.5 * (co - counter)


def col(table, name):
    names = {0: olub.NAMES[:10], 1: olub.NAMES[10:16], 2: olub.NAMES[16:]}
    for k, t in enumerate(TABLES):
        if t is table:
            return 1 + names[k].index(name)
    raise KeyError(name)


# ----------------------------------------------------------------------------- table semantics
def test_tables_shape_and_grid():
    assert ABC.shape == (39, 11) and GH.shape == (47, 7) and MM.shape == (47, 4) and GHD.shape == (47, 7)
    for t in TABLES + [GHD]:
        assert t[0, 0] == pytest.approx(2.1) and t[-1, 0] == pytest.approx(4.0)
        assert np.all(np.diff(t[:, 0]) > 0)
    np.testing.assert_array_equal(GH[:, 0], GHD[:, 0])


def test_c_functions_are_exact_minus_far_field():
    """Table = (4/3) (Jeffery exact - far field) at every node: fixes 'exact minus far-field inverse',
    the units (6 pi eta a^3, not Kim & Karrila's 8 pi eta a^3) and the X^C columns."""
    for row in ABC:
        r = row[0]
        x11, x12 = jeffery_c(r)
        f = far_funcs(r)
        a, b = 4 / 3 * x11 - f["X11C"], 4 / 3 * x12 - f["X12C"]
        # the tables carry about 1e-5 absolute (six printed digits of values up to 0.02)
        assert row[7] == pytest.approx(a, abs=1.5e-5) and row[8] == pytest.approx(b, abs=1.5e-6)
        if r <= 2.5:
            assert row[7] == pytest.approx(a, rel=1e-3) and row[8] == pytest.approx(b, rel=1e-3)


@pytest.mark.parametrize("i,tol", [(0, 0.01), (1, 0.02), (2, 0.02)])
def test_a_functions_near_contact(i, tol):
    """Table + far field = Jeffrey & Onishi's near-contact forms (xi = r - 2; the O(xi ln xi) terms
    they leave out reach 1-2% at xi = 0.2)."""
    r = ABC[i, 0]
    xi = r - 2
    L = np.log(1 / xi)
    f = far_funcs(r)
    x11 = 1 / (4 * xi) + 9 / 40 * L + 0.995419
    x12 = -1 / (4 * xi) - 9 / 40 * L - 0.350153
    y11 = L / 6 + 0.998317
    assert ABC[i, 1] + f["X11A"] == pytest.approx(x11, rel=tol)
    assert ABC[i, 2] + f["X12A"] == pytest.approx(x12, rel=tol)
    assert ABC[i, 3] + f["Y11A"] == pytest.approx(y11, rel=tol)


def test_g_functions_sign_and_units_from_exact_derivatives():
    """r2bghd (exact d/dr of r2bgh) = (2/3) * JO's X11G' - far field' near contact, e from sphere 1
    to sphere 2: X11G = 3/(8 xi) + 27/80 ln(1/xi) + ... in units of 4 pi eta a^2."""
    h = 1e-4
    for i in range(4):          # r <= 2.13: the omitted xi ln xi terms stay below 3%
        r = GH[i, 0]
        xi = r - 2
        dinf = (far_funcs(r + h)["X11G"] - far_funcs(r - h)["X11G"]) / (2 * h)
        expect = 2 / 3 * (-(3 / 8) / xi**2 - (27 / 80) / xi) - dinf
        assert GHD[i, 1] == pytest.approx(expect, rel=0.03)
        # X12G mirrors it near contact
        assert GHD[i, 2] == pytest.approx(-GHD[i, 1], rel=0.01)


@pytest.mark.parametrize("name,u6,g2", [("Y11B", 2 / 3, -1 / 4), ("Y12B", 2 / 3, 1 / 4), ("Y11A", 1.0, 1 / 6)])
def test_log_singular_odd_functions_sign(name, u6, g2):
    """(table + far field)' follows the JO log term -u6 g2 / xi in sign and size (the secant of the
    0.05-spaced table), and would not with the opposite convention of e."""
    c = col(ABC, name)
    for i in range(2):
        r0, r1 = ABC[i, 0], ABC[i + 1, 0]
        xm = 0.5 * (r0 + r1) - 2
        tp = (ABC[i + 1, c] - ABC[i, c]) / (r1 - r0)
        ip = (far_funcs(r1)[name] - far_funcs(r0)[name]) / (r1 - r0)
        lead = -u6 * g2 / xm
        assert 0.6 < (tp + ip) / lead < 1.1
        assert not 0.3 < (-tp + ip) / lead < 1.5


def test_m_table_is_pair_sum():
    """XM near contact carries the 1/xi singularity of X11M + X12M, 2 * (6/5)/8 / xi in units of
    (20/3) pi eta a^3, i.e. (1/3)/xi in 6 pi eta a^3 (X11M alone would carry half)."""
    for i in range(5):
        r0, r1 = MM[i, 0], MM[i + 1, 0]
        xm = 0.5 * (r0 + r1) - 2
        tp = (MM[i + 1, 1] - MM[i, 1]) / (r1 - r0)
        ip = (far_funcs(r1)["XM_sum"] - far_funcs(r0)["XM_sum"]) / (r1 - r0)
        ratio = (tp + ip) / (-(1 / 3) / xm**2)
        assert 1.0 < ratio < 1.15


# -------------------------------------------------------------------------------- interpolant
@pytest.mark.parametrize("k", [0, 1, 2])
def test_interpolant_matches_scipy_pchip(k):
    t = TABLES[k]
    rng = np.random.default_rng(k)
    rs = np.concatenate([t[:, 0], rng.uniform(t[0, 0], t[-1, 0], 300)])
    for r in rs:
        got = plub.interp(t, r)
        want = np.array([PchipInterpolator(t[:, 0], t[:, c])(r) for c in range(1, t.shape[1])])
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15 * np.abs(t[:, 1:]).max())
    # nodes exactly, and held at the first row below it
    for row in t:
        np.testing.assert_array_equal(plub.interp(t, row[0]), row[1:])
    np.testing.assert_array_equal(plub.interp(t, 2.0), t[0, 1:])
    np.testing.assert_array_equal(plub.interp(t, 1.5), t[0, 1:])


def test_interpolant_slopes_track_exact_derivatives():
    """Slopes of the monotone interpolant against r2bghd.dat (the exact derivatives): within 2% of
    each column's largest derivative, except Y11H (column 4), which turns at the first node where the
    monotone end slope is 0 by construction (10%); at interior nodes of uniform spacing the steep
    X11G, X12G slopes are within 1% of the local derivative (at r = 2.2, where the spacing changes from
    0.01 to 0.05, the weighted harmonic mean is within 8%)."""
    s = plub.slopes(GH)
    d = GHD[:, 1:]
    scale = np.abs(d).max(axis=0)
    tol = np.array([0.02, 0.02, 0.02, 0.02, 0.10, 0.02])
    assert np.all(np.abs(s - d).max(axis=0) <= tol * scale)
    h = np.diff(GH[:, 0])
    uniform = np.concatenate([[False], np.isclose(h[1:], h[:-1]), [False]])
    for c in (0, 1):
        err = np.abs(s[:, c] - d[:, c])
        assert np.all(err[uniform] <= 0.01 * np.abs(d[uniform, c]) + 1e-3)
        assert np.all(err[1:-1] <= 0.08 * np.abs(d[1:-1, c]) + 1e-3)


def test_interpolant_accuracy_on_held_out_nodes():
    """Interpolating every other node of r2bgh / r2bm reproduces the nodes in between."""
    for t in (GH, MM):
        sub = t[::2]
        for row in t[1::2]:
            got = plub.interp(sub, row[0])
            scale = np.abs(t[:, 1:]).max(axis=0)
            assert np.all(np.abs(got - row[1:]) <= 0.01 * scale)


def test_interpolant_monotone_between_nodes():
    for t in TABLES:
        rs = np.linspace(t[0, 0], t[-1, 0], 2000)
        v = np.array([plub.interp(t, r) for r in rs])
        for c in range(t.shape[1] - 1):
            for k in range(t.shape[0] - 1):
                lo, hi = sorted((t[k, c + 1], t[k + 1, c + 1]))
                sel = (rs >= t[k, 0]) & (rs <= t[k + 1, 0])
                assert np.all(v[sel, c] >= lo - 1e-12) and np.all(v[sel, c] <= hi + 1e-12)


# ------------------------------------------------------------------------- oracle pair operator
def random_pair_system(seed, n=6, L=None):
    rng = np.random.default_rng(seed)
    xc = [rng.normal(size=3)]
    while len(xc) < n:
        base = xc[rng.integers(len(xc))]
        v = rng.normal(size=3)
        p = base + v / np.linalg.norm(v) * rng.uniform(1.97, 4.3)
        xc.append(p)
    return np.array(xc)


def test_oracle_is_symmetric_and_reciprocal():
    xc = random_pair_system(3, 7)
    Nc = len(xc)
    R = olub.resistance(xc, 4.0)
    V = R[: 6 * Nc, : 6 * Nc]
    np.testing.assert_allclose(V, V.T, atol=1e-12 * np.abs(V).max())
    # (F, T) from a shared strain are the transpose of the summed stresslet responses
    P = R[: 6 * Nc, 6 * Nc:]
    Qsum = sum(R[6 * Nc + 5 * i: 6 * Nc + 5 * i + 5, : 6 * Nc] for i in range(Nc))
    np.testing.assert_allclose(P, Qsum.T, atol=1e-12 * np.abs(P).max())
    for i in range(Nc):
        K = R[6 * Nc + 5 * i: 6 * Nc + 5 * i + 5, 6 * Nc:]
        np.testing.assert_allclose(K, K.T, atol=1e-12 * np.abs(K).max())


def rot(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_oracle_rotation_covariance():
    xc = random_pair_system(5, 6)
    Nc = len(xc)
    rng = np.random.default_rng(9)
    uw = rng.normal(size=6 * Nc)
    E = rng.normal(size=(3, 3))
    E = E + E.T
    E -= np.trace(E) / 3 * np.eye(3)
    ev = np.array([E[0, 0] - E[2, 2], 2 * E[0, 1], 2 * E[0, 2], 2 * E[1, 2], E[1, 1] - E[2, 2]])
    out = olub.apply(xc, uw, ev)
    Q = rot(11)
    E2 = Q @ E @ Q.T
    ev2 = np.array([E2[0, 0] - E2[2, 2], 2 * E2[0, 1], 2 * E2[0, 2], 2 * E2[1, 2], E2[1, 1] - E2[2, 2]])
    uw2 = np.concatenate([(uw[: 3 * Nc].reshape(-1, 3) @ Q.T).ravel(), (uw[3 * Nc:].reshape(-1, 3) @ Q.T).ravel()])
    out2 = olub.apply(xc @ Q.T, uw2, ev2)
    np.testing.assert_allclose(out2[: 3 * Nc].reshape(-1, 3), out[: 3 * Nc].reshape(-1, 3) @ Q.T, atol=1e-11)
    np.testing.assert_allclose(out2[3 * Nc: 6 * Nc].reshape(-1, 3), out[3 * Nc: 6 * Nc].reshape(-1, 3) @ Q.T,
                               atol=1e-11)
    for i in range(Nc):
        s, s2 = out[6 * Nc + 5 * i: 6 * Nc + 5 * i + 5], out2[6 * Nc + 5 * i: 6 * Nc + 5 * i + 5]
        S = np.array([[s[0], s[1], s[2]], [s[1], s[4], s[3]], [s[2], s[3], -s[0] - s[4]]])
        S2 = np.array([[s2[0], s2[1], s2[2]], [s2[1], s2[4], s2[3]], [s2[2], s2[3], -s2[0] - s2[4]]])
        np.testing.assert_allclose(S2, Q @ S @ Q.T, atol=1e-11)


def pair_out(r, e, uw, einf=None):
    xc = np.array([[0.2, -0.1, 0.4], [0.2, -0.1, 0.4] + r * e])
    return olub.apply(xc, uw, einf)


def unit(v):
    v = np.asarray(v, dtype=float)
    return v / np.linalg.norm(v)


def test_oracle_axial_rotation_is_jeffery():
    """Torque on sphere 1 rotating about an arbitrary line of centres: lubrication + far field =
    the exact (4/3) X11C (and (4/3) X12C for the partner's rotation)."""
    e = unit([0.3, -0.8, 0.5])
    for r in (2.1, 2.37, 3.1):
        x11, x12 = jeffery_c(r)
        R = far_field(r * e)
        for w1, w2, want in ((1.0, 0.0, x11), (0.0, 1.0, x12)):
            uw = np.concatenate([np.zeros(6), w1 * e, w2 * e])
            t1 = pair_out(r, e, uw)[6:9] + R[6:9, 6:12] @ uw[6:12]
            np.testing.assert_allclose(t1, 4 / 3 * want * e, atol=2e-5)


def test_oracle_squeeze_is_jeffrey_onishi():
    """Force on two spheres approaching along an arbitrary axis: (X11A - X12A) near contact."""
    e = unit([-0.6, 0.2, 0.77])
    for r, tol in ((2.1, 0.01), (2.15, 0.01), (2.2, 0.02)):
        xi = r - 2
        L = np.log(1 / xi)
        jo = 1 / (2 * xi) + 9 / 20 * L + 0.995419 + 0.350153
        uw = np.concatenate([e, -e, np.zeros(6)])
        f1 = pair_out(r, e, uw)[0:3] + far_field(r * e)[0:3, 0:6] @ uw[0:6]
        np.testing.assert_allclose(f1, jo * e, rtol=tol)


def test_oracle_g_block_sign_through_operator():
    """Stresslet of sphere 1 approaching sphere 2 (U1 = e) through the operator, plus the far field:
    its r-derivative follows (2/3) JO X11G' near contact (a flipped e would reverse it)."""
    e = unit([0.1, 0.9, -0.42])
    P = np.outer(e, e) - np.eye(3) / 3
    uw = np.concatenate([e, np.zeros(9)])

    def x11g_total(r):
        s = pair_out(r, e, uw)[12:17]
        S = np.array([[s[0], s[1], s[2]], [s[1], s[4], s[3]], [s[2], s[3], -s[0] - s[4]]])
        R = far_field(r * e)
        s_inf = R[12:17, 0:3] @ e
        S_inf = np.array([[s_inf[0], s_inf[1], s_inf[2]], [s_inf[1], s_inf[4], s_inf[3]],
                          [s_inf[2], s_inf[3], -s_inf[0] - s_inf[4]]])
        tot = S + S_inf
        # S = X11G (ee - I/3) for U1 = e (the Y term vanishes); project on (ee - I/3)
        assert np.allclose(tot, np.sum(tot * P) / np.sum(P * P) * P, atol=1e-8)
        return np.sum(tot * P) / np.sum(P * P)

    for r in (2.105, 2.125, 2.145):
        h = 0.002
        d = (x11g_total(r + h) - x11g_total(r - h)) / (2 * h)
        xi = r - 2
        assert d == pytest.approx(2 / 3 * (-(3 / 8) / xi**2 - (27 / 80) / xi), rel=0.04)


@pytest.mark.parametrize("r", [2.05, 2.1, 2.23, 2.8, 3.6])
def test_oracle_exact_two_sphere_resistance_is_positive_definite(r):
    """Lubrication + far field (the exact two-sphere translation/rotation resistance) is SPD."""
    for seed in range(3):
        e = unit(np.random.default_rng(seed).normal(size=3))
        R = olub.resistance(np.array([[0.0, 0.0, 0.0], r * e]), 4.0)[:12, :12]
        Rf = far_field(r * e)[:12, :12]
        w = np.linalg.eigvalsh(0.5 * (R + Rf + (R + Rf).T))
        assert w.min() > 0


def test_oracle_cutoff_and_periodic_image():
    L = [10.0, 11.0, 12.0]
    # colloids 0 and 1 couple through the x image (separation (2.6, 0.3, 0)); colloid 2 is far from both
    xc = np.array([[4.6, 0.0, 0.0], [-2.8, 0.3, 0.0], [0.0, 5.0, 5.0]])
    d = olub.min_image(xc[1] - xc[0], L)
    r = np.linalg.norm(d)
    assert 2.5 < r < 2.7
    R = olub.resistance(xc, 4.0, box=L)
    e = d / r
    fn = olub.Tables()(r)
    np.testing.assert_allclose(R[0:3, 3:6], fn["X12A"] * np.outer(e, e) + fn["Y12A"] * (np.eye(3) - np.outer(e, e)))
    # colloid 2 is beyond the cutoff from both: its rows are zero
    Nc = 3
    assert not np.any(R[6:9]) and not np.any(R[3 * Nc + 6: 3 * Nc + 9]) and not np.any(R[6 * Nc + 10: 6 * Nc + 15])
    # with rcut below the pair distance nothing couples
    assert not np.any(olub.resistance(xc, 2.5, box=L))


def test_oracle_textbook_slots():
    """Pair along +x: every function lands in its Kim & Karrila component (e.g. S12 = Y11H W_z,
    S13 = -Y11H W_y, S23 = ZM E23 with the flattened 2 E23 = 1, T_y = Y11B U_z)."""
    r = 2.37
    fn = olub.Tables()(r)
    e = np.array([1.0, 0.0, 0.0])
    ex, ey, ez = np.eye(3)

    def out(u1=None, u2=None, w1=None, w2=None, ev=None):
        z = np.zeros(3)
        uw = np.concatenate([z if u1 is None else u1, z if u2 is None else u2,
                             z if w1 is None else w1, z if w2 is None else w2])
        o = pair_out(r, e, uw, ev)
        return o[0:3], o[6:9], o[12:17]       # F1, T1, S1 (SV)

    F, T, S = out(u1=ex)
    np.testing.assert_allclose(F, fn["X11A"] * ex, atol=1e-14)
    np.testing.assert_allclose(S, [2 / 3 * fn["X11G"], 0, 0, 0, -fn["X11G"] / 3], atol=1e-14)
    F, T, S = out(u1=ey)
    np.testing.assert_allclose(F, fn["Y11A"] * ey, atol=1e-14)
    np.testing.assert_allclose(T, -fn["Y11B"] * ez, atol=1e-14)
    np.testing.assert_allclose(S, [0, fn["Y11G"], 0, 0, 0], atol=1e-14)
    F, T, S = out(u1=ez)
    np.testing.assert_allclose(T, fn["Y11B"] * ey, atol=1e-14)
    F, T, S = out(u2=ex)
    np.testing.assert_allclose(F, fn["X12A"] * ex, atol=1e-14)
    np.testing.assert_allclose(S, [2 / 3 * fn["X12G"], 0, 0, 0, -fn["X12G"] / 3], atol=1e-14)
    F, T, S = out(u2=ez)
    np.testing.assert_allclose(T, fn["Y12B"] * ey, atol=1e-14)
    F, T, S = out(w1=ex)
    np.testing.assert_allclose(T, fn["X11C"] * ex, atol=1e-14)
    np.testing.assert_allclose(S, 0, atol=1e-14)
    F, T, S = out(w1=ez)
    np.testing.assert_allclose(F, -fn["Y11B"] * ey, atol=1e-14)      # ~B11: (e x W) Y11B
    np.testing.assert_allclose(T, fn["Y11C"] * ez, atol=1e-14)
    np.testing.assert_allclose(S, [0, fn["Y11H"], 0, 0, 0], atol=1e-14)
    F, T, S = out(w1=ey)
    np.testing.assert_allclose(S, [0, 0, -fn["Y11H"], 0, 0], atol=1e-14)
    F, T, S = out(w2=ez)
    np.testing.assert_allclose(F, fn["Y12B"] * ey, atol=1e-14)       # ~B12 = -B12^T
    np.testing.assert_allclose(T, fn["Y12C"] * ez, atol=1e-14)
    np.testing.assert_allclose(S, [0, fn["Y12H"], 0, 0, 0], atol=1e-14)
    # strain (flattened EV): E = diag(2/3, -1/3, -1/3); E12 = 1/2; E23 = 1/2
    F, T, S = out(ev=[1, 0, 0, 0, 0])
    np.testing.assert_allclose(F, 2 / 3 * (fn["X11G"] - fn["X12G"]) * ex, atol=1e-14)
    np.testing.assert_allclose(S, [2 / 3 * fn["XM"], 0, 0, 0, -fn["XM"] / 3], atol=1e-14)
    F, T, S = out(ev=[0, 1, 0, 0, 0])
    np.testing.assert_allclose(F, (fn["Y11G"] - fn["Y12G"]) * ey, atol=1e-14)
    np.testing.assert_allclose(T, (fn["Y11H"] + fn["Y12H"]) * ez, atol=1e-14)
    np.testing.assert_allclose(S, [0, fn["YM"] / 2, 0, 0, 0], atol=1e-14)
    F, T, S = out(ev=[0, 0, 0, 1, 0])
    np.testing.assert_allclose(np.concatenate([F, T]), 0, atol=1e-14)
    np.testing.assert_allclose(S, [0, 0, 0, fn["ZM"] / 2, 0], atol=1e-14)
