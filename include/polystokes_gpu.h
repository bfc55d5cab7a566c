/*
 * polystokes_gpu.h -- C ABI of the MI355X (gfx950) per-timestep hydrodynamic solve of PolyStokes.
 *
 * The library (polystokes_brownian_amd/libpsgpu.so) holds one simulation context per handle, keeps
 * every per-step array resident in HBM, and runs hand-written HIP kernels for the FTS grand-mobility
 * tiles, the saddle-point operator, MINRES, the Lanczos square root, the Verlet pairs, the forces and
 * the integrator.  Every entry point returns 0 on success and a positive ps_err code otherwise; the
 * message of the last failure is read with ps_last_error().  The library never aborts the process
 * (the reference installs PetscAbortErrorHandler, src/init.cpp:367; this is a deliberate difference).
 *
 * Pointers are host pointers unless the name ends in _dev, in which case they are device pointers
 * ordered on the context's stream.  One context must not be used from two host threads at once.
 *
 * Vector layouts follow the reference exactly (src/Stokes.cpp:67-98, src/mob.cpp:586-591):
 *   mobility vector (n  = 3Nm + 11Nc):  [ f_m (3Nm) ; F_c (3Nc) ; T_c (3Nc) ; S_c (5Nc) ]
 *   saddle vector   (ns = 6Nm + 17Nc):  [ f (n) ; u (3Nm + 6Nc) ]
 */
#ifndef POLYSTOKES_GPU_H
#define POLYSTOKES_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int ps_err;
enum {
    PS_OK = 0,
    PS_ERR_ARG = 1,      /* invalid argument or size */
    PS_ERR_HIP = 2,      /* a HIP runtime call failed */
    PS_ERR_STATE = 3,    /* call out of order (e.g. solve before mob) */
    PS_ERR_NOCONV = 4,   /* a Krylov solve hit its iteration cap (result still written) */
    PS_ERR_ALLOC = 5,    /* device allocation failed */
    PS_ERR_NODEV = 6     /* no gfx950 device visible */
};

/* Arithmetic forms of the mobility apply. */
enum {
    PS_PREC_FP64 = 0,    /* tiles and applies in FP64 (parity form) */
    PS_PREC_MIXED = 1    /* FP32 tiles/applies, FP64 recurrences, FP64 iterative refinement */
};

/* Excluded-volume force law, src/rhs.cpp:194-199 (POLYSTOKES_FORCE). */
enum { PS_FORCE_HS = 0, PS_FORCE_HS_EXP = 1 };

/*
 * Parameters.  The first three groups are the reference's pybind11 constructor
 * (src/bindings.cpp:44-69), particle_info (src/Stokes.cpp:171) and trap_info
 * (src/Stokes.cpp:241), with the same meaning and defaults (ps_params_default).
 */
typedef struct ps_params {
    /* PolyStokes(dt, samplerate, tmax, output_dir, mm_HI, chain_HI, fene, record_forces, tether,
     * mono_ev, box, t, seed, restart) */
    double dt;
    int samplerate;
    double tmax;
    int mm_HI, chain_HI, fene, record_forces, tether, mono_ev;
    int box_active;               /* 0: unbounded; 1: orthorhombic periodic box */
    double box[3];
    double t;
    uint64_t seed;
    double restart;
    /* particle_info(kT, epsilon, Np, Nc, Nm, Npoly, Nmono_per_chain, beta, kbond, r0, Lmax, tau) */
    double kT, epsilon;
    int Np, Nc, Nm, Npoly, Nmono_per_chain;
    double beta, kbond, r0, Lmax, tau;
    /* trap_info(ktrap, tstart, trun, weaken_trap) */
    double ktrap, tstart, trun, weaken_trap;
    /* Saddle solve (KSPMINRES + PCJACOBI, src/init.cpp:40-64).  The reference hard-codes
     * rtol = 1e-6 after KSPSetFromOptions; here it is a parameter (default 1e-6). */
    double ksp_rtol, ksp_abstol;
    int ksp_maxit;
    /* Brownian square root M^{1/2} xi by Lanczos (replaces SLEPc MFNSolve, src/slip_vel.cpp:27-49):
     * stop when the relative change of the estimate drops below sqrt_rtol, at sqrt_kmax vectors,
     * or at an invariant subspace (beta_j <= 1e-12 beta_0, src/slip_vel.cpp:417). */
    double sqrt_rtol;
    int sqrt_kmax;
    /* Environment knobs of the reference: POLYSTOKES_FORCE, HS_KPREF, HS_KMODE, HS_EXP_K
     * (src/rhs.cpp:194, src/pair_interaction.cpp:108-110,227). */
    int force_sel;
    double hs_kpref;
    int hs_kmode;
    double hs_exp_k;
    /* MI355X forms (no reference counterpart). */
    int precision;                /* PS_PREC_FP64 or PS_PREC_MIXED */
    double ir_rtol;               /* mixed form: FP64 refinement target on the saddle residual */
    int ir_maxit;                 /* mixed form: cap on refinement sweeps; with PS_SOLVER_BLOCK, 0 means
                                     no refinement (the solve's one apply runs on FP32 tiles) */
    int solver;                   /* PS_SOLVER_MINRES (reference algorithm) or PS_SOLVER_BLOCK */
    int aa_kernel;                /* monomer-monomer kernel: PS_AA_AUTO, PS_AA_ORDERED, PS_AA_SYMMETRIC */
    int graph;                    /* ps_run: 1 (default) replays each step as a HIP graph where the step
                                     has no host round trip: PS_SOLVER_BLOCK, kT > 0, profiling off,
                                     lubrication off, and either mm_HI (FP64, or the mixed form with
                                     ir_maxit = 0) or mm_HI=False with one colloid, no mono_ev and no
                                     record_forces; 0 launches every kernel */
    /* Lubrication in the saddle (no reference counterpart; src/saddle.cpp:53 leaves it as a TODO).
     * lub = 1 puts -R_lub, the pair lubrication resistance between colloid velocities and colloid
     * forces and torques (the ps_lub_* operator with no strain), in the saddle's lower-right block:
     * the constraint rows become -f_u - R_lub u = b_lambda.  Off by default (the reference's system).
     * Needs mm_HI and the three tables (ps_lub_set_table) before the first ps_mob; every ps_mob
     * rebuilds the colloid pairs closer than lub_rcut (default 4.0, at most the last table node). */
    int lub;
    double lub_rcut;
} ps_params;

/* Monomer-monomer tile kernel.  ORDERED evaluates every ordered pair (target-stationary, source
 * splits); SYMMETRIC evaluates every unordered pair once and applies the tile both ways (cubic or
 * no box, chain_HI off; otherwise ORDERED is used).  AUTO, where both apply, times each on
 * synthetic positions of the system's size at ps_create and keeps the faster (the outcome is
 * cached per size class for the life of the process; ps_kernel_choice reports it).  Both give the
 * same operator to rounding. */
enum { PS_AA_AUTO = 0, PS_AA_ORDERED = 1, PS_AA_SYMMETRIC = 2 };

/* Saddle-solve algorithms.  Both solve the same system A x = b.
 *   PS_SOLVER_MINRES: preconditioned MINRES on the full saddle system, as KSPSolve does.
 *   PS_SOLVER_BLOCK:  exact block elimination of the constraint B = [-I; 0]: f_u = -b_u',
 *                     M_ss f_s = b_s - M_su f_u (MINRES on the 5Nc x 5Nc stresslet block),
 *                     u = (M f)_u - b_u -- two mobility applies per solve instead of one per
 *                     MINRES iteration. */
enum { PS_SOLVER_MINRES = 0, PS_SOLVER_BLOCK = 1 };

typedef struct ps_solve_info {
    int iterations;      /* MINRES iterations (summed over refinement sweeps in the mixed form) */
    int reason;          /* >0 converged (2 = rtol, 3 = abstol), <0 diverged (-3 = maxit), KSPConvergedReason signs */
    double rnorm;        /* final preconditioned residual estimate */
    double rnorm0;       /* initial preconditioned residual norm */
    int refinements;     /* FP64 refinement sweeps (mixed form), 0 in the FP64 form */
    double true_rel_residual; /* ||b - A x|| / ||b|| in FP64 (mixed form; -1 when not computed) */
} ps_solve_info;

typedef struct ps_sqrt_info {
    int iterations;      /* Lanczos vectors used */
    int nneg;            /* negative Ritz values floored to zero (src/slip_vel.cpp:450-459) */
    double beta0;        /* ||xi|| */
    double rel_change;   /* last relative change of the estimate (convergence measure) */
} ps_sqrt_info;

typedef struct ps_counters {
    int64_t mob_calls, applies, minres_iterations, lanczos_iterations, saddle_solves, sqrt_calls;
    int64_t nvlist;      /* Verlet pairs found by the last check_dist */
    int64_t lanczos_nneg_events, lanczos_nneg_count;
    int64_t stresslet_iterations;  /* PS_SOLVER_BLOCK: MINRES iterations of the 5Nc x 5Nc stresslet solves
                                      (with lubrication: of the 17Nc colloid saddle solves) */
    int64_t stresslet_capped;      /* PS_SOLVER_BLOCK: those solves stopped at their cap (10000) */
} ps_counters;

/* Fill *p with the reference defaults (src/bindings.cpp:56-69; record_forces=true as bound). */
void ps_params_default(ps_params* p);

/* Create a context: alok_arrays + set_vars + init_solver + init_random (src/init.cpp:112-419),
 * ~ PolyStokes ctor + particle_info + trap_info.  Allocates all device buffers. */
ps_err ps_create(const ps_params* p, void** out_ctx);
/* MATOP_DESTROY + cleanup (src/matfree_A.cpp:104-115, src/cleanup.cpp). */
ps_err ps_destroy(void* ctx);
const char* ps_last_error(const void* ctx);
/* Device name and gfx arch of the context's GPU. */
ps_err ps_device_info(void* ctx, char* name, int name_len, int* n_cu);

/* initial_configuration (src/Stokes.cpp:254-317): x0 has 3Np entries; q0 has 4Nc entries or is NULL
 * (identity quaternions).  Trap centres are the colloids' initial positions when restart == 0. */
ps_err ps_initial_configuration(void* ctx, const double* x0, const double* q0);
ps_err ps_set_positions(void* ctx, const double* x);
/* Copy state back: x (3Np), q (4Nc), t; any pointer may be NULL. */
ps_err ps_get_state(void* ctx, double* x, double* q, double* t);
/* Copy per-step vectors back (any may be NULL): up, v_det_n, v_det_c, v_brown, udiff (3Nm+6Nc each),
 * sb (5Nc), fext (3Nm+6Nc). */
ps_err ps_get_vectors(void* ctx, double* up, double* v_det_n, double* v_det_c, double* v_brown,
                      double* udiff, double* sb, double* fext);
ps_err ps_get_counters(void* ctx, ps_counters* c);

/* check_dist (src/check_dist.cpp:12-87): rebuild the Verlet list on the device with a cell list.
 * Pairs (k, j), k < j, in the reference's enumeration (AA, AB, BB; src/init.cpp:225-291) enter
 * vlist[k] iff dr < rverls[type] (float-rounded cutoffs, src/init.cpp:332).  Returns the count. */
ps_err ps_check_dist(void* ctx, int64_t* n_pairs);
/* Copy the Verlet list back as (row k, partner j, reference pair index) triples sorted by (k, j),
 * i.e. exactly the order of vlist[k] in the reference.  cap is the capacity of each array. */
ps_err ps_get_vlist(void* ctx, int32_t* rows, int32_t* partners, int64_t* pair_index, int64_t cap,
                    int64_t* n_pairs);

/* mob() (src/mob.cpp:635-1213): define the grand mobility M at the current positions.  The tiles
 * are generated on the fly by every apply; this call snapshots the positions and the one 3x3 tile
 * the reference carries over between pair loops (see ps_stale_tile). */
ps_err ps_mob(void* ctx);
/* The monomer-colloid translation block the reference writes for every AB pair: the mob_a left by
 * the last processed AA pair of this mob() call (mobility_AB does not write mob_a,
 * src/mob.cpp:407-453; initialize_rank2_array zeroes a copy, src/multi_arrays_initialize.cpp:5). */
ps_err ps_stale_tile(void* ctx, double* tile9);

/* u = M f (n entries), FP64 or the context's precision.  ~ the mobility block of MatMult(A). */
ps_err ps_mob_mult(void* ctx, const double* f, double* u);
ps_err ps_mob_mult_dev(void* ctx, const double* f_dev, double* u_dev);
/* u = M f with no context and no allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  M is
 * the FP64 grand mobility mob() would define at the positions x (3 Np, device), with the parameters *p
 * (mm_HI = 1; box, chain_HI, beta, Nm, Nc, Nmono_per_chain are read); f and u (n = 3 Nm + 11 Nc,
 * device) are the caller's; the work is ordered on `stream` (a hipStream_t, NULL = the null stream).
 * Partial sums, the carried tile and the chain ids live in one 16 MiB device array per GPU, held until
 * the process ends: systems that need more (about Nm > 300,000) are rejected with PS_ERR_ARG.  Calls on
 * one GPU share that array, so calls that overlap in time on one GPU race; errors are read with
 * ps_last_error(NULL). */
ps_err ps_mob_mult_noalloc(const ps_params* p, const double* x_dev, const double* f_dev, double* u_dev,
                           void* stream);
/* MATOP_MULT of the saddle operator A = [[M, B], [B^T, 0]], B = [-I_{3Nm+6Nc}; 0]
 * (src/matfree_A.cpp:38-80, src/arrays.cpp:349-408): y = A x, ns entries. */
ps_err ps_saddle_mult(void* ctx, const double* x, double* y);
ps_err ps_saddle_mult_dev(void* ctx, const double* x_dev, double* y_dev);
/* MATOP_GET_DIAGONAL with PCJacobiSetFixDiagonal (src/matfree_A.cpp:84-102, src/init.cpp:58-59):
 * d = [diag(M) ; 1 ... 1]. */
ps_err ps_saddle_diag(void* ctx, double* d);
/* KSPSolve(ksp, b, x) with KSPMINRES + PCJACOBI (src/vel.cpp:41-51).  warm != 0 uses x as the
 * initial guess (KSPSetInitialGuessNonzero).  Returns PS_ERR_NOCONV at the iteration cap. */
ps_err ps_saddle_solve(void* ctx, const double* b, double* x, int warm, ps_solve_info* info);
ps_err ps_saddle_solve_dev(void* ctx, const double* b_dev, double* x_dev, int warm, ps_solve_info* info);
/* MFNSolve(mfn, xi, w) with FNSQRT on M (src/slip_vel.cpp:27-49): w = M^{1/2} xi by Lanczos with
 * two-pass full reorthogonalisation and negative Ritz values floored (src/slip_vel.cpp:351-477). */
ps_err ps_sqrt_apply(void* ctx, const double* xi, double* w, ps_sqrt_info* info);
ps_err ps_sqrt_apply_dev(void* ctx, const double* xi_dev, double* w_dev, ps_sqrt_info* info);
/* The square-root weights g = beta0 U diag(sqrt(max(theta, 0))) U^T e1 of one symmetric tridiagonal
 * T = U diag(theta) U^T (diagonal alpha[m], off-diagonal beta[m - 1]; 1 <= m <= 64), computed on the
 * device by the eigensolver ps_sqrt_apply runs after its recurrence, from host arrays (the current
 * device; a blocking call).  ~ the LAPACK dsyev step of schur_sqrt_lanczos (src/slip_vel.cpp:427-459).
 * As in the recurrence, beta enters squared (its sign is dropped) and T is cut after the first
 * beta_j <= 1e-12 beta0: *m_used is the order solved, g[m_used .. m) are zero, *nneg counts the
 * negative eigenvalues floored to zero. */
ps_err ps_tridiag_sqrt(int m, const double* alpha, const double* beta, double beta0, double* g, int* m_used,
                       int* nneg);
/* RHS(drift) (src/rhs.cpp:160-260): rebuild the saddle right-hand side on the device from the
 * current Verlet list; rhs (ns entries) is copied back when non-NULL. */
ps_err ps_rhs(void* ctx, int drift, double* rhs);
/* PolyStokes::run() step loop (src/run.cpp:70-162) for nsteps steps, all device-resident.  Output
 * writing is not performed. */
ps_err ps_run(void* ctx, int nsteps);
/* Kernel timing with HIP events on the context's stream (for roofline figures): between begin and
 * end every monomer-monomer tile launch (the dominant kernel) and every whole mobility apply is
 * timed; end synchronises and returns the summed durations.  The symmetric tile kernel's events are
 * recorded by its own dispatch (hipExtLaunchKernel), so they measure the kernel alone; under
 * rocprofv3 (ROCPROF_OUTPUT_PATH set), whose dispatch interception crashes together with such
 * events, it is bracketed by event records like the other kernels (a few microseconds longer). */
typedef struct ps_profile {
    int64_t aa_launches;  double aa_ms;      /* monomer-monomer tile kernel */
    int64_t apply_launches; double apply_ms; /* whole apply (all tile kernels + finalize) */
    int64_t aa_pairs;                        /* target x source monomer pairs per AA launch */
    int64_t dense_launches; double dense_ms; /* stored-dense multi-vector apply kernel */
    int64_t arrow_launches; double arrow_ms; /* mm_HI=False: the DOF pass of the arrowhead apply */
    int64_t arsolve_launches; double arsolve_ms; /* mm_HI=False: a block-eliminated saddle solve, first
                                                    dispatch to last (its two DOF passes) */
} ps_profile;
ps_err ps_profile_begin(void* ctx);
ps_err ps_profile_end(void* ctx, ps_profile* out);
/* Stored-dense grand mobility on MFMA (an MI355X form).  The reference's counterpart is the dense
 * assembly of mob() (MatSetValues INSERT into A, src/mob.cpp:635-1213) followed by PETSc's dense
 * MatMult.  ps_dense_build stores the M of the last ps_mob, rounded once to 16 bits, in the
 * operand order of v_mfma_f32_16x16x32_{bf16,f16} (n rounded up to 32 in both dimensions:
 * 2 bytes per entry).  ps_dense_mult applies it to nrhs <= 16 vectors at once with FP32
 * accumulation: X and Y hold nrhs vectors of length n one after the other (X[j n + i]).  One call
 * reads the stored M once, whatever nrhs.  Timing: ps_profile's dense_* fields. */
enum { PS_DENSE_BF16 = 1, PS_DENSE_F16 = 2 };
ps_err ps_dense_build(void* ctx, int storage);
ps_err ps_dense_mult(void* ctx, const double* X, double* Y, int nrhs);        /* host pointers */
ps_err ps_dense_mult_dev(void* ctx, const double* X, double* Y, int nrhs);    /* device pointers */

/* Output frames (src/write_configuration.cpp:10-166): config_<t>_.txt (3Np positions),
 * quats_<t>_.txt (4Nc), force_<t>_.txt (fext: 3Np forces then 3Nc torques; only with
 * record_forces) and brownian_stresslet_<t>_.txt (5Nc), "%.10f" tab-separated, one particle per
 * line, <t> printed by a default std::ostream.  ps_run_output is run() with the reference's
 * schedule (src/run.cpp:148-160): frames after loop index i == 0 and every i % samplerate == 0.
 * ps_format_time writes the <t> text of a file name. */
ps_err ps_write_outputs(void* ctx, const char* output_dir);
ps_err ps_run_output(void* ctx, int nsteps, const char* output_dir);
ps_err ps_format_time(double t, char* buf, int len);

/* Near-field lubrication between colloids (an addition: the reference distributes the two-sphere
 * tables lubdat/r2babc.dat, r2bgh.dat, r2bm.dat but no source reads them; src/saddle.cpp:53 leaves
 * it as a TODO).  Each table value is the exact two-sphere resistance minus the inverse of the
 * far-field two-sphere mobility, 6 pi eta = 1 and radius 1, Kim & Karrila's convention (e from the
 * sphere whose F, T, S is given to its partner; r2bm holds M11 + M12).  Each table is passed as its
 * rows, r first (which 0: r2babc, 11 columns; 1: r2bgh, 7; 2: r2bm, 4) and interpolated by monotone
 * cubic Hermite (Fritsch-Carlson) in r; below the first node the first row is used.  ps_lub_build
 * lists the colloid pairs with 0 < r < rcut (rcut at most the last node) at the positions of the
 * last ps_mob, through a cell list over the colloids.  ps_lub_apply maps uw = [U_c (3Nc); W_c (3Nc)] and a strain einf shared by all colloids
 * (5 values flattened as the reference's EV, src/mob.cpp:300-305; NULL for none) to
 * fts = [F_c (3Nc); T_c (3Nc); S_c (5Nc, flattened as SV)] in FP64: (U - U_inf, W - W_inf, -E_inf)
 * -> (F, T, S) exerted on the fluid.  ps_lub_interp / ps_lub_slopes evaluate the interpolant and
 * its node slopes on the host without a context (ncols - 1 outputs). */
enum { PS_LUB_ABC = 0, PS_LUB_GH = 1, PS_LUB_M = 2 };
ps_err ps_lub_set_table(void* ctx, int which, const double* rows, int nrows, int ncols);
ps_err ps_lub_build(void* ctx, double rcut, int64_t* npairs);
ps_err ps_lub_apply(void* ctx, const double* uw, const double* einf, double* fts);    /* host pointers */
ps_err ps_lub_apply_dev(void* ctx, const double* uw, const double* einf, double* fts);/* device pointers */
ps_err ps_lub_interp(const double* rows, int nrows, int ncols, double r, double* out);
ps_err ps_lub_slopes(const double* rows, int nrows, int ncols, double* slopes);

/* The mm_HI=False (arrowhead) regime, at most one colloid (src/matfree_A.cpp:1-16): ps_mob builds
 * M^cm; ps_sync_mcc is sync_mcc_schur_correction (src/slip_vel.cpp:81-203), the eigenvalue floor of
 * the colloid Schur complement that ps_run applies after the predictor's and the corrector's mob
 * (src/run.cpp:92, :129).  ps_get_schur returns the floored M^cc (11 x 11, row-major), its
 * eigenvectors (columns of q) and the floored spectrum; ps_get_cell_list the monomer cell list of
 * the last WCA force evaluation (mono_ev): the cell of each monomer, the grid dimensions and
 * geom = [cell size (3), origin (3), periodic] (src/cell_list.cpp:23-96). */
ps_err ps_sync_mcc(void* ctx);
ps_err ps_get_schur(void* ctx, double* mcc, double* q, double* lam);
ps_err ps_get_cell_list(void* ctx, int32_t* cell_of, int32_t* dims, double* geom);

/* Normal draws of the context's noise stream (std::mt19937_64 + std::normal_distribution<double>,
 * src/init.cpp:105-110) -- the draws the next step would consume; advances the stream. */
ps_err ps_draw_normals(void* ctx, double* out, int64_t n);
/* The same stream without a context (seeded with `seed`, drawn in fills of `chunk` values): the host
 * generator the library uses, for checking it against std::normal_distribution.  `margin` > 0
 * replaces the number of polar-method attempts a fill first takes per pair still missing (1.35 by
 * default; acceptance is pi/4), so a small value makes each fill take several further ranges. */
ps_err ps_normal_stream(uint64_t seed, double* out, int64_t n, int64_t chunk, double margin);

/* Batched replicas (an MI355X form with no reference counterpart): R independent systems, each the
 * reference's simulation of its own probes and chains (unbounded, full monomer hydrodynamics) with
 * its own noise seed, advanced together by one persistent kernel per step (one workgroup per
 * replica; the step of src/run.cpp:70-162 with block-eliminated saddle solves and a Lanczos square
 * root of sqrt_kmax iterations).  Replica r has Nc[r] colloids and Npoly[r] chains of Nmono[r]
 * monomers (monomers first, chain by chain, then the colloids, as in a single context); the shared
 * parameters come from *p (mm_HI = 1 and no box required; precision PS_PREC_MIXED runs the square
 * root's monomer tiles in FP32).  Each replica's noise is its own std::mt19937_64 +
 * normal_distribution stream seeded seeds[r] + llround(restart / dt), drawn in a single context's
 * order, so every replica reproduces a single-context run. */
ps_err ps_batch_create(const ps_params* p, int R, const int* Nc, const int* Npoly, const int* Nmono,
                       const uint64_t* seeds, void** out_batch);
ps_err ps_batch_destroy(void* batch);
const char* ps_batch_last_error(const void* batch);
/* out6 = {Np, Nc, Nm, n, nu, ns} of replica r. */
ps_err ps_batch_sizes(void* batch, int r, int64_t* out6);
/* Initial configuration of replica r: 3Np positions, 4Nc quaternions (null: identity); the trap
 * centres are the colloids' positions on a fresh start. */
ps_err ps_batch_set_state(void* batch, int r, const double* x0, const double* q0);
ps_err ps_batch_get_state(void* batch, int r, double* x, double* q, double* t);
ps_err ps_batch_get_vectors(void* batch, int r, double* up, double* v_det_n, double* v_det_c, double* v_brown,
                            double* udiff);
ps_err ps_batch_run(void* batch, int nsteps);
/* Per replica (4 R ints) for the last step: stresslet MINRES iterations, Lanczos iterations,
 * negative Ritz values floored, saddle solves whose stresslet MINRES stopped at its cap (1000
 * iterations; the solve continues unconverged, as the reference continues after KSPSolve's cap). */
ps_err ps_batch_stats(void* batch, int* out);
/* One operation on every replica at its current positions, vectors concatenated over the
 * replicas: 0 / 1: y = M f (FP64 / FP32 monomer tiles; n each), 2: the saddle solve (ns each),
 * 3 / 6: W = M^{1/2} xi (n each; FP64 / FP32 monomer tiles), 4: mob() (no vectors; run it before
 * 0-3 and 6), 5: the force rhs (ns each, no input).  iters (R, may be null): op 2 MINRES, op 3 / 6
 * Lanczos iterations. */
ps_err ps_batch_op(void* batch, int op, const double* in, double* out, int* iters);
/* The last step's wall-clock time of each replica in microseconds and the team size (workgroups)
 * it ran with; either output (R entries) may be null.  No reference counterpart: the reference
 * runs one system per process. */
ps_err ps_batch_times(void* batch, double* us, int* team);
/* The monomer-monomer kernel a context runs (PS_AA_ORDERED or PS_AA_SYMMETRIC) for its FP64
 * applies (kernel64) and its FP32 applies (kernel32, 0 when the context has none), and, where
 * PS_AA_AUTO timed them, the chosen kernel's measured milliseconds per apply (ms64, ms32; else 0).
 * Any output may be null.  No reference counterpart. */
ps_err ps_kernel_choice(void* ctx, int* kernel64, int* kernel32, float* ms64, float* ms32);
/* The tile shape PS_AA_AUTO chose for the ordered kernel: its source split in beads (a multiple of the
 * 256-bead tile; each block applies one 256-target tile to one split), for the FP64 and the FP32
 * applies; 0 where the symmetric kernel (256-bead tile pairs) runs or the context has no FP32 applies.
 * Either output may be null.  No reference counterpart. */
ps_err ps_kernel_shape(void* ctx, int* split64, int* split32);
/* Workgroups of the last step's launch (nwg_last) and the co-resident maximum the schedule packs
 * replicas onto (nwg_max); either pointer may be null.  No reference counterpart. */
ps_err ps_batch_workgroups(void* batch, int* nwg_last, int* nwg_max);
/* The last step's time of each replica per phase in microseconds (5 R: forces, mob(), saddle
 * solves, square root, and within those the team leader's time waiting in joins).  No reference
 * counterpart. */
ps_err ps_batch_phase_times(void* batch, double* us5);
/* Step schedule: replicas far above the mean work run on a team of up to team_max (1..16)
 * workgroups, each rank with at least min_blocks (>= 1) work units (blocks of 64 monomer rows or
 * 16 colloid rows); with autotune
 * the packing is redone before every step from the measured per-replica times. */
ps_err ps_batch_config(void* batch, int team_max, int min_blocks, int autotune);
/* With ps_params.graph = 1 (the default) each step of ps_batch_run is a HIP graph per noise slot --
 * the noise copy, the team reset and the cooperative step kernel -- captured once and replayed (the
 * packing lives in device memory, so a re-tuned schedule keeps the graph; a new trap stiffness
 * recaptures it).  Out (each may be null): graph launches so far, captures, and 1 when a capture
 * failed and the steps run eagerly.  No reference counterpart. */
ps_err ps_batch_graph_info(void* batch, int64_t* replays, int64_t* captures, int* failed);
/* The step schedule's planner, without a GPU: lone-workgroup costs and team caps (>= 1) of R
 * replicas on W co-resident workgroups.  Out (each may be null): the team size of each replica, the
 * workgroup of its first rank, each workgroup's planned load (W) and the makespan target T.  No
 * reference counterpart. */
ps_err ps_batch_plan(int R, const double* cost, const int* cap, int W, int* team, int* wg, double* load, double* T);

#ifdef __cplusplus
}
#endif

#endif /* POLYSTOKES_GPU_H */
