"""CPU restatement of the pair lubrication resistance (FP64, numpy), the checker of
polystokes_brownian_amd/csrc/lub.hip.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker, never by the product path.

The reference has no lubrication code (src/saddle.cpp:53 is a TODO; src/check_dist.cpp:73-85 and
src/init.cpp:10 are commented out); it only distributes the tables lubdat/r2babc.dat, r2bgh.dat,
r2bm.dat.  What this module restates is the physics those tables encode, written independently of the
kernel: every pair tensor is assembled explicitly (a Levi-Civita array, einsum contractions, the
tilde-transposes taken as index permutations of the partner's block) and the matrix is dense.

Table semantics (each pinned in tests/test_lubrication.py, not assumed):
  value = exact two-sphere resistance - inverse of the far-field two-sphere mobility, units
  6 pi eta = 1 and radius 1, Kim & Karrila (1991, ch. 11) convention: e is the unit vector from
  sphere 1 (whose F, T, S is given) to sphere 2; r2bm.dat is M11 + M12.
Interpolation: monotone piecewise-cubic Hermite (Fritsch & Carlson 1980) as implemented by
scipy.interpolate.PchipInterpolator, held at the first row below r = 2.1, zero at r >= rcut.

The operator maps (U - U_inf, W - W_inf, -E_inf) -> (F, T, S) exerted on the fluid, with
uw = [U (3Nc); W (3Nc)], the strain flattened as the reference's EV (E11 - E33, 2E12, 2E13, 2E23,
E22 - E33; src/mob.cpp:300-305) and S as SV (S11, S12, S13, S23, S22).
"""
import os

import numpy as np
from scipy.interpolate import PchipInterpolator

LUBDAT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "polystokes_brownian_amd", "lubdat")
FILES = ("r2babc.dat", "r2bgh.dat", "r2bm.dat")
NAMES = ("X11A", "X12A", "Y11A", "Y12A", "Y11B", "Y12B", "X11C", "X12C", "Y11C", "Y12C",
         "X11G", "X12G", "Y11G", "Y12G", "Y11H", "Y12H", "XM", "YM", "ZM")

EPS = np.zeros((3, 3, 3))
for _a, _b, _c in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
    EPS[_a, _b, _c], EPS[_a, _c, _b] = 1.0, -1.0
DELTA = np.eye(3)

# strain flattening: E (3x3, symmetric, traceless) = sum_c EV_c * EBASIS[c]
EBASIS = np.zeros((5, 3, 3))
EBASIS[0] = np.diag([2.0, -1.0, -1.0]) / 3.0      # E11 - E33 = 1, E22 - E33 = 0
EBASIS[4] = np.diag([-1.0, 2.0, -1.0]) / 3.0      # E22 - E33 = 1, E11 - E33 = 0
for _c, (_a, _b) in ((1, (0, 1)), (2, (0, 2)), (3, (1, 2))):
    EBASIS[_c, _a, _b] = EBASIS[_c, _b, _a] = 0.5  # 2 E_ab = 1
SV_INDEX = ((0, 0), (0, 1), (0, 2), (1, 2), (1, 1))


def load_tables(directory=LUBDAT):
    return [np.loadtxt(os.path.join(directory, f), ndmin=2) for f in FILES]


class Tables:
    """The 19 functions of r, each a scipy PCHIP of its column."""

    def __init__(self, tables=None):
        tables = load_tables() if tables is None else tables
        self.r0 = max(t[0, 0] for t in tables)
        self.rmax = min(t[-1, 0] for t in tables)
        self.f = []
        for t in tables:
            for c in range(1, t.shape[1]):
                self.f.append(PchipInterpolator(t[:, 0], t[:, c], extrapolate=False))

    def __call__(self, r):
        rr = max(float(r), self.r0)
        return dict(zip(NAMES, (float(f(rr)) for f in self.f)))


def pair_tensors(e, fn):
    """The K&K pair tensors of one sphere (1) and its partner (2) for unit vector e from 1 to 2."""
    ee = np.outer(e, e)
    eps_e = np.einsum("abk,k->ab", EPS, e)
    A = lambda X, Y: X * ee + Y * (DELTA - ee)
    Gt = lambda X, Y: (X * np.einsum("ab,k->abk", ee - DELTA / 3, e)
                       + Y * (np.einsum("a,bk->abk", e, DELTA) + np.einsum("b,ak->abk", e, DELTA)
                              - 2 * np.einsum("a,b,k->abk", e, e, e)))
    Ht = lambda Y: Y * (np.einsum("a,bkl,l->abk", e, EPS, e) + np.einsum("b,akl,l->abk", e, EPS, e))
    d0 = 1.5 * np.einsum("ij,kl->ijkl", ee - DELTA / 3, ee - DELTA / 3)
    d1 = 0.5 * (np.einsum("i,jl,k->ijkl", e, DELTA, e) + np.einsum("j,il,k->ijkl", e, DELTA, e)
                + np.einsum("i,jk,l->ijkl", e, DELTA, e) + np.einsum("j,ik,l->ijkl", e, DELTA, e)
                - 4 * np.einsum("ij,kl->ijkl", ee, ee))
    d2 = 0.5 * (np.einsum("ik,jl->ijkl", DELTA, DELTA) + np.einsum("jk,il->ijkl", DELTA, DELTA)
                - np.einsum("ij,kl->ijkl", DELTA, DELTA) + np.einsum("ij,kl->ijkl", ee, DELTA)
                + np.einsum("ij,kl->ijkl", DELTA, ee) - np.einsum("i,jl,k->ijkl", e, DELTA, e)
                - np.einsum("j,il,k->ijkl", e, DELTA, e) - np.einsum("i,jk,l->ijkl", e, DELTA, e)
                - np.einsum("j,ik,l->ijkl", e, DELTA, e) + np.einsum("ij,kl->ijkl", ee, ee))
    return {
        "A11": A(fn["X11A"], fn["Y11A"]), "A12": A(fn["X12A"], fn["Y12A"]),
        "B11": fn["Y11B"] * eps_e, "B12": fn["Y12B"] * eps_e,
        "C11": A(fn["X11C"], fn["Y11C"]), "C12": A(fn["X12C"], fn["Y12C"]),
        "G11": Gt(fn["X11G"], fn["Y11G"]), "G12": Gt(fn["X12G"], fn["Y12G"]),
        "H11": Ht(fn["Y11H"]), "H12": Ht(fn["Y12H"]),
        "M": fn["XM"] * d0 + fn["YM"] * d1 + fn["ZM"] * d2,
    }


def min_image(d, L):
    if L is None:
        return d
    L = np.asarray(L, dtype=float)
    return d - L * np.rint(d / L)


def resistance(xc, rcut, box=None, tables=None):
    """Dense (11 Nc) x (6 Nc + 5) matrix of the pair lubrication, colloid positions xc (Nc, 3)."""
    T = tables if isinstance(tables, Tables) else Tables(tables)
    if not (0 < rcut <= T.rmax):
        raise ValueError("0 < rcut <= last tabulated distance")
    xc = np.asarray(xc, dtype=float).reshape(-1, 3)
    Nc = xc.shape[0]
    R = np.zeros((11 * Nc, 6 * Nc + 5))
    F = lambda i: slice(3 * i, 3 * i + 3)
    Tq = lambda i: slice(3 * Nc + 3 * i, 3 * Nc + 3 * i + 3)
    U = F
    W = Tq
    for i in range(Nc):
        S_E = np.zeros((3, 3, 3, 3))     # S_i,ab per E_kl
        F_E = np.zeros((3, 3, 3))        # F_i,a per E_kl
        T_E = np.zeros((3, 3, 3))
        S_U = {}                         # partner -> (S_ab per U_k, S_ab per W_k)
        for j in range(Nc):
            if j == i:
                continue
            d = min_image(xc[j] - xc[i], box)
            r = np.sqrt(d @ d)
            if r >= rcut or r == 0.0:      # coincident centres have no axis: not paired
                continue
            e = d / r
            t = pair_tensors(e, T(r))
            R[F(i), U(i)] += t["A11"]
            R[F(i), U(j)] += t["A12"]
            # ~B11_ab = B11_ba; ~B12 (force on i from rotation of j) = B21_ba = -B12_ba (odd in e)
            R[F(i), W(i)] += t["B11"].T
            R[F(i), W(j)] += -t["B12"].T
            R[Tq(i), U(i)] += t["B11"]
            R[Tq(i), U(j)] += t["B12"]
            R[Tq(i), W(i)] += t["C11"]
            R[Tq(i), W(j)] += t["C12"]
            s_u_i, s_w_i = S_U.get(i, (np.zeros((3, 3, 3)), np.zeros((3, 3, 3))))
            S_U[i] = (s_u_i + t["G11"], s_w_i + t["H11"])
            s_u_j, s_w_j = S_U.get(j, (np.zeros((3, 3, 3)), np.zeros((3, 3, 3))))
            S_U[j] = (s_u_j + t["G12"], s_w_j + t["H12"])
            # ~G_akl = G_kla; the partner's G21 = -G12, H21 = +H12
            F_E += np.transpose(t["G11"] - t["G12"], (2, 0, 1))
            T_E += np.transpose(t["H11"] + t["H12"], (2, 0, 1))
            S_E += t["M"]
        for j, (gu, hw) in S_U.items():
            for row, (a, b) in enumerate(SV_INDEX):
                R[6 * Nc + 5 * i + row, U(j)] = gu[a, b]
                R[6 * Nc + 5 * i + row, W(j)] = hw[a, b]
        for c in range(5):
            Eb = EBASIS[c]
            R[F(i), 6 * Nc + c] = np.einsum("akl,kl->a", F_E, Eb)
            R[Tq(i), 6 * Nc + c] = np.einsum("akl,kl->a", T_E, Eb)
            S = np.einsum("abkl,kl->ab", S_E, Eb)
            for row, (a, b) in enumerate(SV_INDEX):
                R[6 * Nc + 5 * i + row, 6 * Nc + c] = S[a, b]
    return R


def apply(xc, uw, einf=None, rcut=4.0, box=None, tables=None):
    """fts = R [uw; einf] (einf None: no strain)."""
    R = resistance(xc, rcut, box, tables)
    v = np.concatenate([np.asarray(uw, dtype=float), np.zeros(5) if einf is None else np.asarray(einf, float)])
    return R @ v
