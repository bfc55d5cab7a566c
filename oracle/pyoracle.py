"""ctypes binding of the CPU FP64 oracle (oracle/libpsoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker.  The product path never imports this module.
"""
import ctypes as C
import os

import numpy as np

from polystokes_brownian_amd._abi import PsParams, PsSolveInfo, PsSqrtInfo, params_from_dict

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsoracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        vp = C.c_void_p
        L.or_params_default.argtypes = [C.POINTER(PsParams)]
        L.or_create.restype = vp
        L.or_create.argtypes = [C.POINTER(PsParams), C.c_char_p, C.c_int]
        L.or_destroy.argtypes = [vp]
        L.or_set_dense.argtypes = [vp, C.c_int]
        L.or_sizes.argtypes = [vp, C.POINTER(C.c_int64)] * 1 + [C.POINTER(C.c_int64)] * 2
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        L.or_initial_configuration.argtypes = [vp, dp, dp]
        L.or_set_positions.argtypes = [vp, dp]
        L.or_get_state.argtypes = [vp, dp, dp, dp]
        L.or_get_vectors.argtypes = [vp] + [dp] * 7
        L.or_check_dist.restype = C.c_int64
        L.or_check_dist.argtypes = [vp]
        L.or_get_vlist.restype = C.c_int64
        L.or_get_vlist.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.c_int64]
        L.or_mob.argtypes = [vp]
        L.or_stale_tile.argtypes = [vp, dp]
        L.or_mob_dense.argtypes = [vp, dp]
        L.or_mob_mult.argtypes = [vp, dp, dp]
        L.or_saddle_mult.argtypes = [vp, dp, dp]
        L.or_saddle_diag.argtypes = [vp, dp]
        L.or_saddle_solve.argtypes = [vp, dp, dp, C.c_int, C.POINTER(PsSolveInfo)]
        L.or_sqrt_apply.argtypes = [vp, dp, dp, C.POINTER(PsSqrtInfo)]
        L.or_rhs.argtypes = [vp, C.c_int, dp]
        L.or_run.argtypes = [vp, C.c_int]
        L.or_draw_normals.argtypes = [vp, dp, C.c_int64]
        L.or_tridiag_sqrt.argtypes = [C.c_int, dp, dp, C.c_double, dp]
        L.or_counters.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int)]
        L.or_mt19937_64_nth.restype = C.c_uint64
        L.or_mt19937_64_nth.argtypes = [C.c_uint64, C.c_int64]
        L.or_normals.argtypes = [C.c_uint64, C.c_int64, dp]
        L.or_schur_sync.argtypes = [vp]
        L.or_schur.restype = C.c_int64
        L.or_schur.argtypes = [vp, dp, dp, dp]
        L.or_cell_list.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_params(**kw):
    p = PsParams()
    lib().or_params_default(C.byref(p))
    return params_from_dict(kw, p)


class Oracle:
    """One FP64 CPU simulation (mm_HI=True regime)."""

    def __init__(self, params, dense=None):
        if isinstance(params, dict):
            params = default_params(**params)
        self.params = params
        err = C.create_string_buffer(512)
        h = lib().or_create(C.byref(params), err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self.h = h
        n, nu, ns = C.c_int64(), C.c_int64(), C.c_int64()
        lib().or_sizes(h, C.byref(n), C.byref(nu), C.byref(ns))
        self.n, self.nu, self.ns = n.value, nu.value, ns.value
        self.Np, self.Nc, self.Nm = params.Np, params.Nc, params.Nm
        if dense is not None:
            lib().or_set_dense(h, int(dense))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def initial_configuration(self, x0, q0=None):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        q = None if q0 is None else _p(np.ascontiguousarray(q0, dtype=np.float64))
        lib().or_initial_configuration(self.h, _p(x0), q)

    def set_positions(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        lib().or_set_positions(self.h, _p(x))

    def state(self):
        x = np.zeros(3 * self.Np)
        q = np.zeros(max(4 * self.Nc, 1))
        t = np.zeros(1)
        lib().or_get_state(self.h, _p(x), _p(q), _p(t))
        return x, q[: 4 * self.Nc], float(t[0])

    def vectors(self):
        out = {k: np.zeros(self.nu) for k in ("up", "v_det_n", "v_det_c", "v_brown", "udiff", "fext")}
        out["sb"] = np.zeros(max(5 * self.Nc, 1))
        lib().or_get_vectors(self.h, _p(out["up"]), _p(out["v_det_n"]), _p(out["v_det_c"]), _p(out["v_brown"]),
                             _p(out["udiff"]), _p(out["sb"]), _p(out["fext"]))
        out["sb"] = out["sb"][: 5 * self.Nc]
        return out

    def check_dist(self):
        return lib().or_check_dist(self.h)

    def vlist(self):
        cnt = lib().or_get_vlist(self.h, None, None, None, 0)
        rows = np.zeros(max(cnt, 1), np.int32)
        parts = np.zeros(max(cnt, 1), np.int32)
        pidx = np.zeros(max(cnt, 1), np.int64)
        lib().or_get_vlist(self.h, rows.ctypes.data_as(C.POINTER(C.c_int32)),
                           parts.ctypes.data_as(C.POINTER(C.c_int32)),
                           pidx.ctypes.data_as(C.POINTER(C.c_int64)), cnt)
        return rows[:cnt], parts[:cnt], pidx[:cnt]

    def mob(self):
        rc = lib().or_mob(self.h)
        if rc:
            raise RuntimeError("mob before check_dist")

    def stale_tile(self):
        t = np.zeros(9)
        lib().or_stale_tile(self.h, _p(t))
        return t.reshape(3, 3)

    def mob_dense(self):
        M = np.zeros((self.n, self.n))
        lib().or_mob_dense(self.h, _p(M))
        return M

    def schur_sync(self):
        """sync_mcc_schur_correction of the mm_HI=False regime."""
        if lib().or_schur_sync(self.h):
            raise RuntimeError("schur_sync needs mm_HI=False and a mob()")

    def schur(self):
        """(floored M^cc, eigenvectors Q (columns), floored spectrum) of the last sync."""
        m = 11 * self.Nc
        mcc, q, lam = np.zeros((m, m)), np.zeros((m, m)), np.zeros(m)
        if lib().or_schur(self.h, _p(mcc), _p(q), _p(lam)) < 0:
            raise RuntimeError("schur needs mm_HI=False")
        return mcc, q, lam

    def cell_list(self):
        """Monomer cell list at the current positions: (cell of each monomer, dims, size, origin,
        periodic)."""
        cell = np.zeros(self.Nm, dtype=np.int32)
        dims = np.zeros(3, dtype=np.int32)
        geom = np.zeros(7)
        if lib().or_cell_list(self.h, cell.ctypes.data_as(C.POINTER(C.c_int32)),
                              dims.ctypes.data_as(C.POINTER(C.c_int32)), _p(geom)):
            raise RuntimeError("cell_list needs mm_HI=False")
        return cell, dims, geom[:3].copy(), geom[3:6].copy(), bool(geom[6])

    def mob_mult(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        u = np.zeros(self.n)
        lib().or_mob_mult(self.h, _p(f), _p(u))
        return u

    def saddle_mult(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.ns)
        lib().or_saddle_mult(self.h, _p(x), _p(y))
        return y

    def saddle_diag(self):
        d = np.zeros(self.ns)
        lib().or_saddle_diag(self.h, _p(d))
        return d

    def saddle_solve(self, b, x0=None):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.ns) if x0 is None else np.array(x0, dtype=np.float64)
        info = PsSolveInfo()
        rc = lib().or_saddle_solve(self.h, _p(b), _p(x), int(x0 is not None), C.byref(info))
        return x, info, rc

    def sqrt_apply(self, xi):
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        w = np.zeros(self.n)
        info = PsSqrtInfo()
        lib().or_sqrt_apply(self.h, _p(xi), _p(w), C.byref(info))
        return w, info

    def rhs(self, drift=False):
        r = np.zeros(self.ns)
        rc = lib().or_rhs(self.h, int(drift), _p(r))
        if rc:
            raise RuntimeError("rhs before check_dist")
        return r

    def run(self, nsteps):
        return lib().or_run(self.h, int(nsteps))

    def draw_normals(self, n):
        out = np.zeros(n)
        lib().or_draw_normals(self.h, _p(out), n)
        return out

    def counters(self):
        a, m = C.c_int64(), C.c_int64()
        lm, ls = C.c_int(), C.c_int()
        lib().or_counters(self.h, C.byref(a), C.byref(m), C.byref(lm), C.byref(ls))
        return dict(applies=a.value, minres_total=m.value, last_minres=lm.value, last_sqrt=ls.value)


def set_threads(n):
    lib().or_set_threads(int(n))


def get_threads():
    return lib().or_get_threads()


def tridiag_sqrt(alpha, betad, beta0):
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    betad = np.ascontiguousarray(betad if len(betad) else np.zeros(1), dtype=np.float64)
    g = np.zeros(len(alpha))
    nneg = lib().or_tridiag_sqrt(len(alpha), _p(alpha), _p(betad), float(beta0), _p(g))
    return g, nneg


def mt19937_64_nth(seed, nth):
    return lib().or_mt19937_64_nth(seed, nth)


def normals(seed, n):
    out = np.zeros(n)
    lib().or_normals(seed, n, _p(out))
    return out
