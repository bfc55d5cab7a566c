"""Python front end of libpsgpu.so: the reference's pybind11 `PolyStokes` / `Box` API
(src/bindings.cpp:41-119) over the C ABI, plus `Context`, a thin wrapper of every entry point.

The HIP library is the only compute path: if libpsgpu.so is missing or no GPU is visible, every
call raises -- there is no CPU fallback.
"""
import ctypes as C
import math
import os

import numpy as np

from ._abi import (ERR_NAMES, PS_ERR_NOCONV, PS_OK, PS_PREC_FP64, PS_PREC_MIXED, PsCounters, PsParams,
                   PsProfile, PsSolveInfo, PsSqrtInfo, params_from_dict)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsgpu.so")
_lib = None


class PsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


def lib():
    """Load libpsgpu.so (built by __graft_entry__.build() / `make -C polystokes_brownian_amd/csrc`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C polystokes_brownian_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        dp, vp, ip = C.POINTER(C.c_double), C.c_void_p, C.c_int
        L.ps_params_default.argtypes = [C.POINTER(PsParams)]
        L.ps_params_default.restype = None
        L.ps_create.argtypes = [C.POINTER(PsParams), C.POINTER(vp)]
        L.ps_destroy.argtypes = [vp]
        L.ps_last_error.argtypes = [vp]
        L.ps_last_error.restype = C.c_char_p
        L.ps_device_info.argtypes = [vp, C.c_char_p, ip, C.POINTER(ip)]
        L.ps_initial_configuration.argtypes = [vp, dp, dp]
        L.ps_set_positions.argtypes = [vp, dp]
        L.ps_get_state.argtypes = [vp, dp, dp, dp]
        L.ps_get_vectors.argtypes = [vp] + [dp] * 7
        L.ps_get_counters.argtypes = [vp, C.POINTER(PsCounters)]
        L.ps_check_dist.argtypes = [vp, C.POINTER(C.c_int64)]
        L.ps_get_vlist.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                   C.c_int64, C.POINTER(C.c_int64)]
        L.ps_mob.argtypes = [vp]
        L.ps_stale_tile.argtypes = [vp, dp]
        for name in ("ps_mob_mult", "ps_mob_mult_dev", "ps_saddle_mult", "ps_saddle_mult_dev"):
            getattr(L, name).argtypes = [vp, C.c_void_p, C.c_void_p]
        L.ps_mob_mult_noalloc.argtypes = [C.POINTER(PsParams), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ps_tridiag_sqrt.argtypes = [C.c_int, dp, dp, C.c_double, dp, C.POINTER(ip), C.POINTER(ip)]
        L.ps_saddle_diag.argtypes = [vp, dp]
        L.ps_saddle_solve.argtypes = [vp, C.c_void_p, C.c_void_p, ip, C.POINTER(PsSolveInfo)]
        L.ps_saddle_solve_dev.argtypes = [vp, C.c_void_p, C.c_void_p, ip, C.POINTER(PsSolveInfo)]
        L.ps_sqrt_apply.argtypes = [vp, C.c_void_p, C.c_void_p, C.POINTER(PsSqrtInfo)]
        L.ps_sqrt_apply_dev.argtypes = [vp, C.c_void_p, C.c_void_p, C.POINTER(PsSqrtInfo)]
        L.ps_rhs.argtypes = [vp, ip, dp]
        L.ps_run.argtypes = [vp, ip]
        L.ps_draw_normals.argtypes = [vp, dp, C.c_int64]
        L.ps_profile_begin.argtypes = [vp]
        L.ps_profile_end.argtypes = [vp, C.POINTER(PsProfile)]
        L.ps_write_outputs.argtypes = [vp, C.c_char_p]
        L.ps_run_output.argtypes = [vp, ip, C.c_char_p]
        L.ps_format_time.argtypes = [C.c_double, C.c_char_p, ip]
        L.ps_dense_build.argtypes = [vp, ip]
        L.ps_dense_mult.argtypes = [vp, C.c_void_p, C.c_void_p, ip]
        L.ps_dense_mult_dev.argtypes = [vp, C.c_void_p, C.c_void_p, ip]
        L.ps_lub_set_table.argtypes = [vp, ip, dp, ip, ip]
        L.ps_lub_build.argtypes = [vp, C.c_double, C.POINTER(C.c_int64)]
        for name in ("ps_lub_apply", "ps_lub_apply_dev"):
            getattr(L, name).argtypes = [vp, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ps_sync_mcc.argtypes = [vp]
        L.ps_get_schur.argtypes = [vp, dp, dp, dp]
        L.ps_get_cell_list.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), dp]
        L.ps_lub_interp.argtypes = [dp, ip, ip, C.c_double, dp]
        L.ps_lub_slopes.argtypes = [dp, ip, ip, dp]
        L.ps_batch_create.argtypes = [C.POINTER(PsParams), ip, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip),
                                      C.POINTER(C.c_uint64), C.POINTER(vp)]
        L.ps_batch_destroy.argtypes = [vp]
        L.ps_batch_last_error.argtypes = [vp]
        L.ps_batch_last_error.restype = C.c_char_p
        L.ps_batch_sizes.argtypes = [vp, ip, C.POINTER(C.c_int64)]
        L.ps_batch_set_state.argtypes = [vp, ip, dp, dp]
        L.ps_batch_get_state.argtypes = [vp, ip, dp, dp, dp]
        L.ps_batch_get_vectors.argtypes = [vp, ip, dp, dp, dp, dp, dp]
        L.ps_batch_run.argtypes = [vp, ip]
        L.ps_batch_stats.argtypes = [vp, C.POINTER(ip)]
        L.ps_batch_op.argtypes = [vp, ip, dp, dp, C.POINTER(ip)]
        L.ps_batch_times.argtypes = [vp, dp, C.POINTER(ip)]
        L.ps_batch_config.argtypes = [vp, ip, ip, ip]
        L.ps_kernel_shape.argtypes = [vp, C.POINTER(ip), C.POINTER(ip)]
        L.ps_batch_graph_info.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(ip)]
        L.ps_batch_phase_times.argtypes = [vp, dp]
        L.ps_batch_workgroups.argtypes = [vp, C.POINTER(ip), C.POINTER(ip)]
        L.ps_kernel_choice.argtypes = [vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(C.c_float),
                                       C.POINTER(C.c_float)]
        L.ps_batch_plan.argtypes = [ip, dp, C.POINTER(ip), ip, C.POINTER(ip), C.POINTER(ip), dp, dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _vec(a, n, name):
    """`a` as a contiguous float64 vector of exactly n entries; ValueError otherwise.  Every array the
    C ABI reads or writes is checked here, before the call: the library trusts the sizes it is given."""
    v = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if v.size != n:
        raise ValueError(f"{name} must have {n} entries, got {v.size}")
    return v


def default_params(**kw):
    p = PsParams()
    lib().ps_params_default(C.byref(p))
    return params_from_dict(kw, p)


class Context:
    """One device-resident simulation: a 1:1 wrapper of the C ABI (host numpy arrays in and out)."""

    def __init__(self, params):
        if isinstance(params, dict):
            params = default_params(**params)
        self.params = params
        h = C.c_void_p()
        rc = lib().ps_create(C.byref(params), C.byref(h))
        if rc != PS_OK:
            raise PsError(rc, lib().ps_last_error(None).decode())
        self.h = h
        self.Np, self.Nm, self.Nc = params.Np, params.Nm, params.Nc
        self.n = 3 * self.Nm + 11 * self.Nc
        self.nu = 3 * self.Nm + 6 * self.Nc
        self.ns = self.n + self.nu
        if params.lub:
            # lubrication in the saddle reads the bundled two-sphere tables
            self.lub_set_tables()

    def _check(self, rc, allow_noconv=False):
        if rc == PS_OK or (allow_noconv and rc == PS_ERR_NOCONV):
            return rc
        raise PsError(rc, lib().ps_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().ps_destroy(self.h)
            self.h = None

    __del__ = close

    def device_info(self):
        name = C.create_string_buffer(256)
        ncu = C.c_int()
        self._check(lib().ps_device_info(self.h, name, 256, C.byref(ncu)))
        return name.value.decode(), ncu.value

    def initial_configuration(self, x0, q0=None):
        x0 = _vec(x0, 3 * self.Np, "init_x0 (3 Np positions)")
        q = None if q0 is None else _vec(q0, 4 * self.Nc, "init_q (4 Nc quaternion entries)")
        self._check(lib().ps_initial_configuration(self.h, _p(x0), None if q is None else _p(q)))

    def set_positions(self, x):
        x = _vec(x, 3 * self.Np, "x (3 Np positions)")
        self._check(lib().ps_set_positions(self.h, _p(x)))

    def state(self):
        x = np.zeros(3 * self.Np)
        q = np.zeros(max(4 * self.Nc, 1))
        t = np.zeros(1)
        self._check(lib().ps_get_state(self.h, _p(x), _p(q), _p(t)))
        return x, q[: 4 * self.Nc], float(t[0])

    def vectors(self):
        out = {k: np.zeros(self.nu) for k in ("up", "v_det_n", "v_det_c", "v_brown", "udiff", "fext")}
        out["sb"] = np.zeros(max(5 * self.Nc, 1))
        self._check(lib().ps_get_vectors(self.h, _p(out["up"]), _p(out["v_det_n"]), _p(out["v_det_c"]),
                                         _p(out["v_brown"]), _p(out["udiff"]), _p(out["sb"]), _p(out["fext"])))
        out["sb"] = out["sb"][: 5 * self.Nc]
        return out

    def counters(self):
        c = PsCounters()
        self._check(lib().ps_get_counters(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in PsCounters._fields_}

    def kernel_choice(self):
        """ps_kernel_choice: {"fp64": "ordered"|"symmetric", "fp32": ... or None, "ms64", "ms32"} --
        the monomer-monomer kernel each apply precision runs and, under aa_kernel "auto", the
        chosen kernel's measured milliseconds per apply (0 where nothing was timed)."""
        k64, k32 = C.c_int(), C.c_int()
        t64, t32 = C.c_float(), C.c_float()
        self._check(lib().ps_kernel_choice(self.h, C.byref(k64), C.byref(k32), C.byref(t64), C.byref(t32)))
        s64, s32 = C.c_int(), C.c_int()
        self._check(lib().ps_kernel_shape(self.h, C.byref(s64), C.byref(s32)))
        name = {1: "ordered", 2: "symmetric"}
        return {"fp64": name[k64.value], "fp32": name.get(k32.value), "ms64": t64.value, "ms32": t32.value,
                "split64": s64.value, "split32": s32.value}

    def check_dist(self):
        n = C.c_int64()
        self._check(lib().ps_check_dist(self.h, C.byref(n)))
        return n.value

    def vlist(self):
        n = C.c_int64()
        self._check(lib().ps_get_vlist(self.h, None, None, None, 0, C.byref(n)))
        cnt = n.value
        rows = np.zeros(max(cnt, 1), np.int32)
        parts = np.zeros(max(cnt, 1), np.int32)
        pidx = np.zeros(max(cnt, 1), np.int64)
        self._check(lib().ps_get_vlist(self.h, rows.ctypes.data_as(C.POINTER(C.c_int32)),
                                       parts.ctypes.data_as(C.POINTER(C.c_int32)),
                                       pidx.ctypes.data_as(C.POINTER(C.c_int64)), cnt, C.byref(n)))
        return rows[:cnt], parts[:cnt], pidx[:cnt]

    def mob(self):
        self._check(lib().ps_mob(self.h))

    def stale_tile(self):
        t = np.zeros(9)
        self._check(lib().ps_stale_tile(self.h, _p(t)))
        return t.reshape(3, 3)

    def mob_mult(self, f):
        f = _vec(f, self.n, "f (n = 3 Nm + 11 Nc)")
        u = np.zeros(self.n)
        self._check(lib().ps_mob_mult(self.h, f.ctypes.data, u.ctypes.data))
        return u

    def saddle_mult(self, x):
        x = _vec(x, self.ns, "x (ns = 6 Nm + 17 Nc)")
        y = np.zeros(self.ns)
        self._check(lib().ps_saddle_mult(self.h, x.ctypes.data, y.ctypes.data))
        return y

    def saddle_diag(self):
        d = np.zeros(self.ns)
        self._check(lib().ps_saddle_diag(self.h, _p(d)))
        return d

    def saddle_solve(self, b, x0=None):
        b = _vec(b, self.ns, "b (ns = 6 Nm + 17 Nc)")
        x = np.zeros(self.ns) if x0 is None else _vec(x0, self.ns, "x0 (ns = 6 Nm + 17 Nc)").copy()
        info = PsSolveInfo()
        rc = self._check(lib().ps_saddle_solve(self.h, b.ctypes.data, x.ctypes.data, int(x0 is not None),
                                               C.byref(info)), allow_noconv=True)
        return x, info, rc

    def sqrt_apply(self, xi):
        xi = _vec(xi, self.n, "xi (n = 3 Nm + 11 Nc)")
        w = np.zeros(self.n)
        info = PsSqrtInfo()
        self._check(lib().ps_sqrt_apply(self.h, xi.ctypes.data, w.ctypes.data, C.byref(info)))
        return w, info

    def rhs(self, drift=False):
        r = np.zeros(self.ns)
        self._check(lib().ps_rhs(self.h, int(drift), _p(r)))
        return r

    def run(self, nsteps):
        self._check(lib().ps_run(self.h, int(nsteps)))

    def draw_normals(self, n):
        out = np.zeros(n)
        self._check(lib().ps_draw_normals(self.h, _p(out), n))
        return out

    def write_outputs(self, output_dir):
        """Frame files of the current state (ps_write_outputs, src/write_configuration.cpp)."""
        self._check(lib().ps_write_outputs(self.h, os.fsencode(output_dir)))

    def run_output(self, nsteps, output_dir):
        """nsteps steps with the reference's output schedule (ps_run_output, src/run.cpp:148-160)."""
        self._check(lib().ps_run_output(self.h, int(nsteps), os.fsencode(output_dir)))

    def dense_build(self, storage="bf16"):
        """Store the M of the last mob() in 16 bits in MFMA operand order (ps_dense_build)."""
        self._check(lib().ps_dense_build(self.h, {"bf16": 1, "f16": 2}.get(storage, storage)))

    def dense_mult(self, X):
        """Y = M X for X of shape (nrhs, n), nrhs <= 16, on the stored-dense M (ps_dense_mult)."""
        X = np.ascontiguousarray(np.atleast_2d(X), dtype=np.float64)
        if X.ndim != 2 or X.shape[1] != self.n or not 1 <= X.shape[0] <= 16:
            raise ValueError(f"X must have shape (nrhs, n) with 1 <= nrhs <= 16 and n = {self.n}, got {X.shape}")
        Y = np.zeros_like(X)
        self._check(lib().ps_dense_mult(self.h, X.ctypes.data, Y.ctypes.data, X.shape[0]))
        return Y

    def sync_mcc(self):
        """sync_mcc_schur_correction of the mm_HI=False regime (ps_sync_mcc)."""
        self._check(lib().ps_sync_mcc(self.h))

    def schur(self):
        """(floored M^cc, eigenvectors Q as columns, floored spectrum) of the last ps_sync_mcc."""
        mcc, q, lam = np.zeros((11, 11)), np.zeros((11, 11)), np.zeros(11)
        self._check(lib().ps_get_schur(self.h, _p(mcc), _p(q), _p(lam)))
        return mcc, q, lam

    def cell_list(self):
        """Monomer cell list of the last WCA evaluation: (cell of each monomer, dims, size, origin,
        periodic)."""
        cell = np.zeros(max(self.Nm, 1), dtype=np.int32)
        dims = np.zeros(3, dtype=np.int32)
        geom = np.zeros(7)
        self._check(lib().ps_get_cell_list(self.h, cell.ctypes.data_as(C.POINTER(C.c_int32)),
                                           dims.ctypes.data_as(C.POINTER(C.c_int32)), _p(geom)))
        return cell[: self.Nm], dims, geom[:3].copy(), geom[3:6].copy(), bool(geom[6])

    def lub_set_tables(self, tables=None):
        """Load the three two-sphere tables (rows with r first; default: the bundled lubdat files)."""
        from .lubrication import load_tables
        tables = load_tables() if tables is None else tables
        for which, rows in enumerate(tables):
            rows = np.ascontiguousarray(rows, dtype=np.float64)
            self._check(lib().ps_lub_set_table(self.h, which, _p(rows), rows.shape[0], rows.shape[1]))

    def lub_build(self, rcut=4.0):
        """Colloid pairs closer than rcut at the last mob() positions; returns the pair count."""
        n = C.c_int64()
        self._check(lib().ps_lub_build(self.h, float(rcut), C.byref(n)))
        return n.value

    def lub_apply(self, uw, einf=None):
        """[F_c; T_c; S_c] (11 Nc) of the pair lubrication for uw = [U_c; W_c] (6 Nc) and strain einf (5)."""
        uw = _vec(uw, 6 * self.Nc, "uw (6 Nc)")
        out = np.zeros(11 * self.Nc)
        e = None
        if einf is not None:
            e = _vec(einf, 5, "einf")
        self._check(lib().ps_lub_apply(self.h, uw.ctypes.data, None if e is None else e.ctypes.data,
                                       out.ctypes.data))
        return out

    def profile_begin(self):
        self._check(lib().ps_profile_begin(self.h))

    def profile_end(self):
        p = PsProfile()
        self._check(lib().ps_profile_end(self.h, C.byref(p)))
        return {k: getattr(p, k) for k, _ in PsProfile._fields_}


def mob_mult_noalloc(params, x_ptr, f_ptr, u_ptr, stream=0):
    """ps_mob_mult_noalloc: u = M f at positions x with no context and no allocation; the arguments
    are device addresses (ints) of 3 Np, n and n doubles and a hipStream_t handle (0: null stream)."""
    if isinstance(params, dict):
        params = default_params(**params)
    rc = lib().ps_mob_mult_noalloc(C.byref(params), x_ptr, f_ptr, u_ptr, stream or None)
    if rc != PS_OK:
        raise PsError(rc, lib().ps_last_error(None).decode())


def tridiag_sqrt(alpha, beta, beta0):
    """ps_tridiag_sqrt: the square-root weights beta0 U sqrt(max(theta, 0)) U^T e1 of the tridiagonal
    (alpha, beta) on the device.  Returns (g, m_used, nneg)."""
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta if len(beta) else np.zeros(1), dtype=np.float64)
    g = np.zeros(len(a))
    mu, nn = C.c_int(0), C.c_int(0)
    rc = lib().ps_tridiag_sqrt(len(a), _p(a), _p(b), float(beta0), _p(g), C.byref(mu), C.byref(nn))
    if rc != PS_OK:
        raise PsError(rc, lib().ps_last_error(None).decode())
    return g, mu.value, nn.value


class Box:
    """Periodic box helper with the reference's semantics (src/box.cpp, src/bindings.cpp:96-118)."""

    def __init__(self):
        self._L = None

    def configure(self, lengths):
        lengths = list(lengths)
        if not lengths:
            self._L = None
            return
        if len(lengths) != 3:
            raise ValueError("Box::configure expects an empty list (no PBC) or exactly 3 lengths [Lx, Ly, Lz]")
        if not all(v > 0.0 for v in lengths):
            raise ValueError("Box lengths must be positive")
        self._L = [float(v) for v in lengths]

    def active(self):
        return self._L is not None

    def minimum_image(self, dx, dy, dz):
        if self._L is None:
            return dx, dy, dz
        out = []
        for v, L in zip((dx, dy, dz), self._L):
            out.append(v - L * np.rint(v * (1.0 / L)))
        return tuple(float(v) for v in out)

    def wrap(self, x, y, z):
        return self.minimum_image(x, y, z)


class PolyStokes:
    """The reference's Python class (src/bindings.cpp:41-95), driving the MI355X library.

    Same constructor keywords and defaults, same particle_info / trap_info / initial_configuration /
    run sequence.  Two keyword-only additions select the MI355X forms and solver tolerances through
    set_solver_options(); by default the reference's settings apply (MINRES rtol 1e-6, FP64)."""

    def __init__(self, dt, samplerate, tmax, output_dir, mm_HI=True, chain_HI=False, fene=True,
                 record_forces=True, tether=False, mono_ev=False, box=None, t=0.0, seed=12345, restart=0.0):
        self._p = default_params()
        b = Box()
        b.configure([] if box is None else box)
        self._pd = dict(dt=dt, samplerate=samplerate, tmax=tmax, mm_HI=int(mm_HI), chain_HI=int(chain_HI),
                        fene=int(fene), record_forces=int(record_forces), tether=int(tether),
                        mono_ev=int(mono_ev), box=None if box is None else list(box), t=t, seed=seed,
                        restart=restart)
        self.output_dir = output_dir
        self._x0 = None
        self._q0 = None
        self._ctx = None
        self.warn_neg_eig = True
        # nsteps = floor((tmax - restart)/dt) (src/Stokes.cpp:25)
        self.nsteps = max(0, int(math.floor((tmax - restart) / dt)))

    def particle_info(self, kT, epsilon, Np, Nc, Nm, Npoly, Nmono_per_chain, beta, kbond, r0, Lmax, tau):
        self._pd.update(kT=kT, epsilon=epsilon, Np=Np, Nc=Nc, Nm=Nm, Npoly=Npoly,
                        Nmono_per_chain=Nmono_per_chain, beta=beta, kbond=kbond, r0=r0, Lmax=Lmax, tau=tau)

    def trap_info(self, ktrap, tstart, trun, weaken_trap=-1):
        self._pd.update(ktrap=ktrap, tstart=tstart, trun=trun, weaken_trap=weaken_trap)

    def set_solver_options(self, precision=None, ksp_rtol=None, sqrt_rtol=None, sqrt_kmax=None, ir_rtol=None,
                           solver=None, aa_kernel=None):
        """MI355X forms (no reference counterpart): precision "fp64" | "mixed"; solver "minres" (the
        reference's KSPMINRES + PCJACOBI) | "block" (exact block elimination); aa_kernel "auto" |
        "ordered" | "symmetric"; tolerances as in ps_params."""
        for k, v in dict(ksp_rtol=ksp_rtol, sqrt_rtol=sqrt_rtol, sqrt_kmax=sqrt_kmax, ir_rtol=ir_rtol).items():
            if v is not None:
                self._pd[k] = v
        if precision is not None:
            self._pd["precision"] = {"fp64": PS_PREC_FP64, "mixed": PS_PREC_MIXED}.get(precision, precision)
        if solver is not None:
            self._pd["solver"] = {"minres": 0, "block": 1}.get(solver, solver)
        if aa_kernel is not None:
            self._pd["aa_kernel"] = {"auto": 0, "ordered": 1, "symmetric": 2}.get(aa_kernel, aa_kernel)

    def set_warn_neg_eig(self, on=True):
        self.warn_neg_eig = bool(on)

    def initial_configuration(self, init_x0, init_q=None):
        self._x0 = np.ascontiguousarray(init_x0, dtype=np.float64).reshape(-1)
        self._q0 = None if init_q is None else np.ascontiguousarray(init_q, dtype=np.float64).reshape(-1)
        if self._q0 is not None and self._q0.size != 4 * self._pd.get("Nc", 0):
            self._q0 = None   # src/Stokes.cpp:298-301: size mismatch falls back to identity

    def _context(self):
        if self._ctx is None:
            if self._x0 is None:
                raise RuntimeError("initial_configuration must be called before run()")
            self._ctx = Context(params_from_dict(self._pd, default_params()))
            self._ctx.initial_configuration(self._x0, self._q0)
        return self._ctx

    def run(self):
        """PolyStokes::run (src/run.cpp:55-175): nsteps steps on the device, writing the frame files
        (config_/quats_/force_/brownian_stresslet_<t>_.txt) to output_dir on the reference's
        schedule when output_dir is given; returns the final positions."""
        ctx = self._context()
        if self.output_dir:
            os.makedirs(self.output_dir, exist_ok=True)
            ctx.run_output(self.nsteps, self.output_dir)
        else:
            ctx.run(self.nsteps)
        return ctx.state()[0]

    @property
    def context(self):
        return self._context()


def batch_plan(cost, cap, W):
    """ps_batch_plan: (team sizes, workgroup of each replica's first rank, per-workgroup loads,
    makespan target) for lone-workgroup costs and team caps on W workgroups (host only)."""
    cost = np.ascontiguousarray(cost, dtype=np.float64)
    R = len(cost)
    capa = (C.c_int * R)(*[int(c) for c in cap])
    team, wg = (C.c_int * R)(), (C.c_int * R)()
    load, T = np.zeros(W), np.zeros(1)
    rc = lib().ps_batch_plan(R, _p(cost), capa, int(W), team, wg, _p(load), _p(T))
    if rc != PS_OK:
        raise PsError(rc, "ps_batch_plan")
    return np.array(list(team)), np.array(list(wg)), load, float(T[0])


class Batch:
    """Batched replicas (ps_batch_*): R independent systems sharing the parameters `params` (a dict
    of ps_params fields; unbounded, full monomer hydrodynamics), each given by a dict with Nc,
    Npoly, Nmono, seed and, optionally, x0 / q0.  Each replica reproduces a single Context run with
    its own seed."""

    def __init__(self, params, replicas):
        p = default_params()
        params_from_dict(params, p)
        self.R = len(replicas)
        ia = lambda key: (C.c_int * self.R)(*[int(r[key]) for r in replicas])
        seeds = (C.c_uint64 * self.R)(*[int(r["seed"]) for r in replicas])
        h = C.c_void_p()
        rc = lib().ps_batch_create(C.byref(p), self.R, ia("Nc"), ia("Npoly"), ia("Nmono"), seeds, C.byref(h))
        if rc != 0:
            raise PsError(rc, lib().ps_batch_last_error(None).decode())
        self.h = h
        self.sizes = []
        for r in range(self.R):
            v = (C.c_int64 * 6)()
            self._check(lib().ps_batch_sizes(self.h, r, v))
            self.sizes.append(dict(zip(("Np", "Nc", "Nm", "n", "nu", "ns"), list(v))))
        for r, rep in enumerate(replicas):
            if rep.get("x0") is not None:
                self.set_state(r, rep["x0"], rep.get("q0"))

    def _check(self, rc):
        if rc != 0:
            raise PsError(rc, lib().ps_batch_last_error(self.h).decode())

    def _replica(self, r):
        if not 0 <= int(r) < self.R:
            raise ValueError(f"replica index {r} outside [0, {self.R})")
        return self.sizes[int(r)]

    def set_state(self, r, x0, q0=None):
        s = self._replica(r)
        x0 = _vec(x0, 3 * s["Np"], f"x0 of replica {r} (3 Np)")
        q0 = None if q0 is None else _vec(q0, 4 * s["Nc"], f"q0 of replica {r} (4 Nc)")
        self._check(lib().ps_batch_set_state(self.h, r, _p(x0), None if q0 is None else _p(q0)))

    def state(self, r):
        s = self._replica(r)
        x, q, t = np.zeros(3 * s["Np"]), np.zeros(max(4 * s["Nc"], 1)), C.c_double()
        self._check(lib().ps_batch_get_state(self.h, r, _p(x), _p(q), C.byref(t)))
        return x, q[: 4 * s["Nc"]], t.value

    def vectors(self, r):
        nu = self._replica(r)["nu"]
        out = {k: np.zeros(nu) for k in ("up", "v_det_n", "v_det_c", "v_brown", "udiff")}
        self._check(lib().ps_batch_get_vectors(self.h, r, *[_p(out[k]) for k in
                                                             ("up", "v_det_n", "v_det_c", "v_brown", "udiff")]))
        return out

    def run(self, nsteps):
        self._check(lib().ps_batch_run(self.h, int(nsteps)))

    def stats(self):
        out = (C.c_int * (4 * self.R))()
        self._check(lib().ps_batch_stats(self.h, out))
        return np.array(list(out), dtype=np.int64).reshape(self.R, 4)

    def times(self):
        """ps_batch_times: (per-replica wall-clock microseconds of the last step, team sizes)."""
        us = np.zeros(self.R)
        team = (C.c_int * self.R)()
        self._check(lib().ps_batch_times(self.h, _p(us), team))
        return us, np.array(list(team), dtype=np.int64)

    def workgroups(self):
        """ps_batch_workgroups: (workgroups of the last step's launch, co-resident maximum)."""
        a, m = C.c_int(), C.c_int()
        self._check(lib().ps_batch_workgroups(self.h, C.byref(a), C.byref(m)))
        return a.value, m.value

    def phase_times(self):
        """ps_batch_phase_times: (R, 5) microseconds in forces, mob(), saddle solves, square root,
        and the team leader's time waiting in joins."""
        out = np.zeros(5 * self.R)
        self._check(lib().ps_batch_phase_times(self.h, _p(out)))
        return out.reshape(self.R, 5)

    def config(self, team_max=16, min_blocks=2, autotune=True):
        """ps_batch_config: the step schedule's team limits and autotuning."""
        self._check(lib().ps_batch_config(self.h, int(team_max), int(min_blocks), int(bool(autotune))))

    def graph_info(self):
        """ps_batch_graph_info: {"replays", "captures", "failed"} of the step graphs."""
        rep, cap, fail = C.c_int64(), C.c_int64(), C.c_int()
        self._check(lib().ps_batch_graph_info(self.h, C.byref(rep), C.byref(cap), C.byref(fail)))
        return {"replays": rep.value, "captures": cap.value, "failed": bool(fail.value)}

    def op(self, op, inputs=None):
        """ps_batch_op: a list of per-replica input vectors in, per-replica outputs (and the
        iteration counts) out."""
        key = "ns" if op in (2, 5) else "n"
        lens = [s[key] for s in self.sizes]
        offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        tot = int(offs[-1])
        if op not in (0, 1, 2, 3, 4, 5, 6):
            raise ValueError(f"unknown batch op {op}")
        xin = np.zeros(max(tot, 1))
        if inputs is not None:
            if len(inputs) != self.R:
                raise ValueError(f"op {op} takes one input vector per replica ({self.R}), got {len(inputs)}")
            for r, v in enumerate(inputs):
                xin[offs[r]:offs[r + 1]] = _vec(v, lens[r], f"input of replica {r} ({key})")
        xout = np.zeros(max(tot, 1))
        its = (C.c_int * self.R)()
        self._check(lib().ps_batch_op(self.h, int(op), _p(xin), _p(xout), its))
        if op == 4:
            return None, list(its)
        return [xout[offs[r]:offs[r + 1]].copy() for r in range(self.R)], list(its)

    def close(self):
        if getattr(self, "h", None):
            lib().ps_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

