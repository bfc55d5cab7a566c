"""ctypes mirror of include/polystokes_gpu.h (structs shared by the product wrapper and the tests)."""
import ctypes as C

PS_OK, PS_ERR_ARG, PS_ERR_HIP, PS_ERR_STATE, PS_ERR_NOCONV, PS_ERR_ALLOC, PS_ERR_NODEV = range(7)
PS_PREC_FP64, PS_PREC_MIXED = 0, 1
PS_SOLVER_MINRES, PS_SOLVER_BLOCK = 0, 1
PS_FORCE_HS, PS_FORCE_HS_EXP = 0, 1

ERR_NAMES = {
    PS_ERR_ARG: "PS_ERR_ARG",
    PS_ERR_HIP: "PS_ERR_HIP",
    PS_ERR_STATE: "PS_ERR_STATE",
    PS_ERR_NOCONV: "PS_ERR_NOCONV",
    PS_ERR_ALLOC: "PS_ERR_ALLOC",
    PS_ERR_NODEV: "PS_ERR_NODEV",
}


class PsParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("samplerate", C.c_int), ("tmax", C.c_double),
        ("mm_HI", C.c_int), ("chain_HI", C.c_int), ("fene", C.c_int), ("record_forces", C.c_int),
        ("tether", C.c_int), ("mono_ev", C.c_int),
        ("box_active", C.c_int), ("box", C.c_double * 3),
        ("t", C.c_double), ("seed", C.c_uint64), ("restart", C.c_double),
        ("kT", C.c_double), ("epsilon", C.c_double),
        ("Np", C.c_int), ("Nc", C.c_int), ("Nm", C.c_int), ("Npoly", C.c_int), ("Nmono_per_chain", C.c_int),
        ("beta", C.c_double), ("kbond", C.c_double), ("r0", C.c_double), ("Lmax", C.c_double), ("tau", C.c_double),
        ("ktrap", C.c_double), ("tstart", C.c_double), ("trun", C.c_double), ("weaken_trap", C.c_double),
        ("ksp_rtol", C.c_double), ("ksp_abstol", C.c_double), ("ksp_maxit", C.c_int),
        ("sqrt_rtol", C.c_double), ("sqrt_kmax", C.c_int),
        ("force_sel", C.c_int), ("hs_kpref", C.c_double), ("hs_kmode", C.c_int), ("hs_exp_k", C.c_double),
        ("precision", C.c_int), ("ir_rtol", C.c_double), ("ir_maxit", C.c_int),
        ("solver", C.c_int), ("aa_kernel", C.c_int), ("graph", C.c_int),
        ("lub", C.c_int), ("lub_rcut", C.c_double),
    ]


class PsSolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("reason", C.c_int), ("rnorm", C.c_double),
                ("rnorm0", C.c_double), ("refinements", C.c_int), ("true_rel_residual", C.c_double)]


class PsSqrtInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("nneg", C.c_int), ("beta0", C.c_double), ("rel_change", C.c_double)]


class PsCounters(C.Structure):
    _fields_ = [("mob_calls", C.c_int64), ("applies", C.c_int64), ("minres_iterations", C.c_int64),
                ("lanczos_iterations", C.c_int64), ("saddle_solves", C.c_int64), ("sqrt_calls", C.c_int64),
                ("nvlist", C.c_int64), ("lanczos_nneg_events", C.c_int64), ("lanczos_nneg_count", C.c_int64),
                ("stresslet_iterations", C.c_int64), ("stresslet_capped", C.c_int64)]


def params_from_dict(d, base=None):
    """Fill a PsParams from a dict of field -> value (box given as 3-sequence or None)."""
    p = PsParams() if base is None else base
    for k, v in d.items():
        if k == "box":
            if v is None:
                p.box_active = 0
            else:
                p.box_active = 1
                for i in range(3):
                    p.box[i] = float(v[i])
            continue
        if not hasattr(p, k):
            raise KeyError(f"unknown parameter {k!r}")
        setattr(p, k, type(getattr(p, k))(v))
    return p


def sizes(p):
    """(n, nu, ns): mobility, constraint and saddle sizes (src/Stokes.cpp:85-94)."""
    n = 3 * p.Nm + 11 * p.Nc
    nu = 3 * p.Nm + 6 * p.Nc
    return n, nu, n + nu


PS_DENSE_BF16, PS_DENSE_F16 = 1, 2


class PsProfile(C.Structure):
    _fields_ = [("aa_launches", C.c_int64), ("aa_ms", C.c_double), ("apply_launches", C.c_int64),
                ("apply_ms", C.c_double), ("aa_pairs", C.c_int64), ("dense_launches", C.c_int64),
                ("dense_ms", C.c_double), ("arrow_launches", C.c_int64), ("arrow_ms", C.c_double),
                ("arsolve_launches", C.c_int64), ("arsolve_ms", C.c_double)]
