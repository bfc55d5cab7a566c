// ps_common.h -- context, error plumbing and device helpers shared by the gfx950 translation units.
#pragma once
#include "normals.h"
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>
#include "../../include/polystokes_gpu.h"

#define PS_BLOCK 256

// Return-on-error plumbing: every HIP call is checked; the message lands in ctx->err.
#define PS_HIP(ctx, call)                                                                         \
    do {                                                                                          \
        hipError_t _e = (call);                                                                   \
        if (_e != hipSuccess) {                                                                   \
            (ctx)->set_error(std::string(#call) + ": " + hipGetErrorString(_e));                  \
            return PS_ERR_HIP;                                                                    \
        }                                                                                         \
    } while (0)
#define PS_TRY(expr)                                                                              \
    do {                                                                                          \
        ps_err _r = (expr);                                                                       \
        if (_r != PS_OK) return _r;                                                               \
    } while (0)
#define PS_LAUNCH(ctx)                                                                            \
    do {                                                                                          \
        hipError_t _e = hipGetLastError();                                                        \
        if (_e != hipSuccess) {                                                                   \
            (ctx)->set_error(std::string("kernel launch: ") + hipGetErrorString(_e));             \
            return PS_ERR_HIP;                                                                    \
        }                                                                                         \
    } while (0)

// Constants of one system, passed by value to the kernels.
struct SysConst {
    int Np, Nm, Nc;
    int chain_HI;
    int box_active;
    double L[3], Linv[3];
    double beta, beta2, beta_inv, two_beta;
    int64_t n, nu, ns, nm3, nc3;
};

// Device-side counters and flags (Ctx::dcnt): accumulated by the kernels that decide them, so that
// a step needs no host round trip; ps_get_counters adds them to the host-side counts.
enum {
    DC_LANCZOS_ITERS = 0,   // Lanczos vectors used by the square roots
    DC_NNEG_EVENTS,         // square roots that floored a negative Ritz value
    DC_NNEG_COUNT,          // negative Ritz values floored
    DC_MSS_ITERS,           // stresslet MINRES iterations of the block solves
    DC_MSS_CAPPED,          // stresslet solves that stopped at their iteration cap
    DC_VLIST_OVERFLOW,      // a Verlet row or the whole list exceeded its fixed capacity (sticky flag)
    DC_BARRIER_TIMEOUT,     // a grid-wide wait ran out of time (sticky flag)
    DC_N = 8
};

// Device-resident MINRES scalars (one per solve).
struct MinresState {
    double beta1, tol, oldb, beta, dbar, epsln, phibar, cs, sn;
    double alfa, oldeps, delta, gamma, phi;
    int its, done, reason, pad;
};

// Apply workspace (partial sums) sized at create time.
struct ApplyWork {
    // symmetric AA kernel: 256-bead tiles, upper-triangle tile pairs (the pairs that involve the
    // partial last tile first, then the full ones), scaled positions [Nm][4]
    bool sym = false;
    int ntile = 0, npair_full = 0, npair_pad = 0;
    int2* pairs = nullptr;
    void* xs = nullptr;
    double inv_L = 1.0;
    int nsplit_aa = 1, nchunk_c = 1, nchunk_m = 1;
    int per_split_aa = 0, per_chunk_c = 0, per_chunk_m = 0;
    void* part_aa = nullptr;     // [nsplit][Nm][3]  (T): a monomer row r = 3 i + c at [sp][r]
    void* part_abm = nullptr;    // [nchunk_m][Nm][3] (T), colloid-source chunks
    void* part_c = nullptr;      // [nchunk][Nc][12] (T) (11 rows + monomer-force sum slot reuse)
    void* part_fs = nullptr;     // [nchunk][3]      (T) sum of monomer forces per chunk
    double* bb = nullptr;        // [Nc][11]
    size_t bytes_aa = 0, bytes_abm = 0, bytes_c = 0, bytes_fs = 0;
};

struct Ctx {
    ps_params p{};
    SysConst sc{};
    int Np = 0, Nm = 0, Nc = 0;
    int64_t n = 0, nu = 0, ns = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    int n_cu = 0;
    std::string err;
    void set_error(const std::string& m) { err = m; }

    // ---- positions and per-step vectors (device)
    double *x = nullptr, *xpd = nullptr, *xm = nullptr, *xpos0 = nullptr, *q = nullptr, *x_n = nullptr;
    float* xm32 = nullptr;          // FP32 snapshot of xm for the mixed form
    int* chain = nullptr;
    double *up = nullptr, *udiff = nullptr, *v_det_n = nullptr, *v_det_c = nullptr, *v_brown = nullptr;
    double *fext = nullptr, *drift_v = nullptr, *sb = nullptr;
    double *rhs = nullptr, *X = nullptr, *Xdet = nullptr, *Xd = nullptr, *W = nullptr, *xi = nullptr;
    double* stale = nullptr;        // [18]: stale tile (9) + tile carried to the next mob() (9)
    double* trap = nullptr;         // [2]: current ktrap, time t (device copies for the force kernel)
    int* bonds = nullptr;           // CSR: offsets [Np+1], then endpoints k [bond_stride], j [bond_stride]
    int nbonds = 0;
    size_t bond_stride = 1;
    // ---- host mirrors
    int stale_i = -1, stale_j = -1; // last processed AA pair (fixed by chain ids)
    double ktrap_cur = 0.0, t = 0.0;
    bool have_pd = false, have_mob = false, have_x = false;
    // ---- apply
    ApplyWork aw64, aw32;
    float aa_tuned_ms[2] = {0.f, 0.f};   // PS_AA_AUTO: the measured ms per apply of the chosen kernel (FP64, FP32)
    // ---- Krylov work
    double *kr = nullptr;           // 9 * ns scratch vectors for MINRES
    double *dinv = nullptr;         // ns
    double *red = nullptr;          // reduction partials
    int red_cap = 0;
    MinresState* mstate = nullptr;  // device
    MinresState* mstate_h = nullptr;// pinned host
    double* Q = nullptr;            // Lanczos basis kmax x n
    double* lz = nullptr;           // Lanczos scalars on device: alpha [64], beta^2 [64], then scalar slots
    double* lz_h = nullptr;         // pinned host mirror
    double *ir_r = nullptr, *ir_d = nullptr;   // refinement residual / correction (ns)
    int kmax = 0;
    double* lzred = nullptr;        // Lanczos Gram-Schmidt partials: [64][nb] per pass (two), then [nb] |w|^2
    double* lzg = nullptr;          // Lanczos square root weights of every m (64 x 64), then per m the
                                    // negative Ritz value count (64)
    double* lzinfo = nullptr;       // [4]: m, negative Ritz values, relative change, beta0 of the last one
    // ---- device-side counters and flags (DC_*), read by ps_get_counters and after ps_run
    int64_t* dcnt = nullptr;
    int64_t* dcnt_h = nullptr;      // pinned
    unsigned* gbar = nullptr;       // grid barriers of the persistent MINRES kernels: [0] count, [1] generation,
                                    // [2] error; per wide-MINRES system s, a flag per block at 128 (1 + s)
    // ---- block-elimination solver (PS_SOLVER_BLOCK)
    double* Mss = nullptr;          // 5Nc x 5Nc stresslet block, rebuilt by mob()
    double* bs_y = nullptr;         // n
    double* bs_r = nullptr;         // 5Nc
    // ---- Verlet list (check_dist)
    int ncell[3] = {0, 0, 0};
    int cell_mode = 0;              // 0: brute force; 1: periodic cell list
    double cell_size[3] = {0, 0, 0}, cell_org[3] = {0, 0, 0};
    int* cell_of = nullptr;         // [Np]
    int* cell_start = nullptr;      // [ncells+2]: exclusive scan, total, max
    int* cell_count = nullptr;      // [ncells]
    int* cell_perm = nullptr;       // [Np]
    double* cell_xs = nullptr;      // [3 Np] positions in cell order
    int* nb_count = nullptr;        // [Np] pairs per row
    int* nb_off = nullptr;          // [Np+2]: row offsets, total, longest row
    int* nb_part = nullptr;         // partners (sorted per row)
    int64_t nb_cap = 0;
    int64_t nvlist = 0;
    int* dflag = nullptr;           // overflow flag etc. [4]
    int* hflag = nullptr;           // pinned
    // ---- host RNG: the stream of std::mt19937_64 + normal_distribution (src/init.cpp:105-110)
    NormalStream normals;
    double* noise_h = nullptr;      // pinned, 2 slots of n + nu (ps_run's noise producer)
    hipEvent_t noise_ev[2] = {nullptr, nullptr};
    // ---- event timing (ps_profile_begin/end)
    bool prof_on = false;
    bool prof_dispatch_events = true;   // false under rocprofv3 (ROCPROF_OUTPUT_PATH set)
    std::vector<hipEvent_t> ev_pool;
    int ev_used = 0;
    std::vector<int> ev_kind;       // 0: AA kernel, 1: whole apply, 2: dense apply, 3: arrowhead apply pass
    ps_profile prof{};
    ps_err prof_record(int kind, hipEvent_t* a, hipEvent_t* b);   // reserve a start/stop pair
    ps_err prof_harvest();
    // ---- counters
    ps_counters cnt{};
    int last_minres_its = 0;
    // ---- scratch used by the host-pointer entry points
    double* io_a = nullptr;
    double* io_b = nullptr;
    // ---- stored-dense M (dense.hip)
    void* dense_M = nullptr;        // np x np 16-bit, MFMA operand order
    void* dense_X = nullptr;        // vectors in B-operand order
    double* dense_io = nullptr;     // 2 x 16 x n, host-pointer entry point
    size_t dense_bytes = 0;
    int dense_storage = 0;
    // ---- mm_HI == false regime (arrow.hip)
    bool arrow = false;
    double* ar_mcm = nullptr;       // [8][3Nm]: the rotation and stresslet rows of M^cm
    double* ar_part = nullptr;      // per-block partials
    double* ar_gpart = nullptr;     // per-block Gram partials of the last M^cm (for the Schur floor)
    bool ar_gram_next = false;      // the next mob() is followed by a Schur floor: store its Gram partials
    bool ar_gram_ready = false;     // ar_gpart holds the current M^cm's Gram partials
    hipGraphExec_t step_graph[2] = {nullptr, nullptr};   // ps_run's step replays, one per noise slot
    ps_counters step_graph_delta{};                      // counter increments of one step
    const double* ar_last_part = nullptr;   // the Gram partials of the last Schur floor, and their count
    int ar_last_nb = 0;
    double* ar_mcc = nullptr;       // floored M^cc (121), eigenvectors Q (121), floored spectrum (11);
                                    // at 256: Cholesky factor of its stresslet block (25, 1/diag)
    double* ar_bbox = nullptr;      // monomer bounding box (unbounded WCA grid)
    double* ar_host = nullptr;      // pinned
    int* ar_flag = nullptr;         // [Nm+2] monomer-colloid Verlet flags
    int* ar_pos = nullptr;          // [Nm+2] their exclusive scan within each block
    int* ar_bscan = nullptr;        // [2 nb]: per-block flag counts, then their exclusive scan
    int* ar_counter = nullptr;      // [3] last-block counters (apply, noise, Verlet list), reset by their
                                    // last block
    int ar_nb = 0;
    bool ar_synced = false;         // a Schur floor has run
    bool ar_diag_frozen = false;    // the Jacobi diagonal is fixed at the first MINRES solve
    int* wc_cell = nullptr;         // WCA cell list: cell of each monomer
    int* wc_count = nullptr;
    int* wc_start = nullptr;        // [ncells+2]
    int* wc_mem = nullptr;          // monomers by cell, ascending within a cell
    int64_t wc_cap = 0;
    int wc_dims[3] = {0, 0, 0};
    double wc_geom[7] = {0, 0, 0, 0, 0, 0, 0};   // size (3), origin (3), periodic
    bool wc_built = false;
    // ---- lubrication (lub.hip)
    double* lub_tab[3] = {nullptr, nullptr, nullptr};   // r[n], f[n k], slopes[n k] per table
    int lub_rows[3] = {0, 0, 0};
    double lub_rlast[3] = {0, 0, 0};   // last node of each table
    int* lub_off = nullptr;         // [Nc+1] partner offsets
    int* lub_part = nullptr;        // partners (both directions)
    int lub_cap = 0;
    double lub_rcut = 0.0;
    bool lub_built = false;
    void* lub_grid = nullptr;       // cell-list geometry (device)
    int* lub_cell_of = nullptr;     // [Nc] cell of each colloid
    int* lub_cell_mem = nullptr;    // [Nc] colloids by cell
    int* lub_cells = nullptr;       // counts, starts, cursors of the cells
    // lubricated block solve: the colloid saddle (17 Nc unknowns) and its work vectors
    double* cs_buf = nullptr;
    // ---- joint step (one_step_joint): the mob() states at x, x + eps w and x - eps w held at once.
    //      Slot 0 is the context's own buffers; slots 1 and 2 serve the two RFD probes.  The current
    //      slot's buffers are the ones in xm, xm32, aw64.xs, aw32.xs, stale and Mss.
    struct MobSlot {
        double* xm = nullptr;
        float* xm32 = nullptr;
        void* xs64 = nullptr;
        void* xs32 = nullptr;
        double* stale = nullptr;
        double* Mss = nullptr;
    };
    MobSlot mslot[3];
    int nmslot = 0;                 // 0: one mob() state (no joint step); 3: slots in use
    double* stale_prev = nullptr;   // the carried tile of the last mob() call (the next one reads it)
    hipStream_t side = nullptr;     // second stream of the joint step (default priority)
    hipEvent_t jev[2] = {nullptr, nullptr};   // fork and join of the side stream
    double* rhs_s = nullptr;        // ns: the slip solve's right-hand side
    double* rhs_d = nullptr;        // ns: the RFD probes' right-hand side
    double* Xd2 = nullptr;          // ns: the second probe's solution
    double* bsr4 = nullptr;         // [4][5Nc]: stresslet right-hand sides of the four joint systems
};

// ----------------------------------------------------------------------------- device helpers
__device__ __forceinline__ double ps_rsqrt(double x) {
    // Hardware reciprocal square root refined by two Newton-Raphson steps (full FP64 accuracy).
    double y = __builtin_amdgcn_rsq(x);
    double h = x * y;
    double r = __builtin_fma(-h, y, 1.0);
    y = __builtin_fma(0.5 * y, r, y);
    h = x * y;
    r = __builtin_fma(-h, y, 1.0);
    y = __builtin_fma(0.5 * y, r, y);
    return y;
}
__device__ __forceinline__ float ps_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }

template <class T>
__device__ __forceinline__ T ps_rint(T v);
template <>
__device__ __forceinline__ double ps_rint<double>(double v) { return __builtin_rint(v); }
template <>
__device__ __forceinline__ float ps_rint<float>(float v) { return __builtin_rintf(v); }

// Host entry points implemented in the translation units.
ps_err mob_prepare(Ctx* c);                                        // stale tile + snapshots
ps_err mob_slots_alloc(Ctx* c);                                    // slots 1, 2 of the joint step
void mob_slots_free(Ctx* c);
void mob_use_slot(Ctx* c, int k);                                  // make slot k the current mob() state
ps_err block_stage1(Ctx* c, const double* b, double* x, double* rs);
ps_err block_stage3(Ctx* c, const double* b, double* x, int prec, double* uout);
bool mss_joint_fits(const Ctx* c, int nsys);
ps_err mss_joint(Ctx* c, int nsys, const double* const* A, const double* const* rs, double* const* x);
ps_err sqrt_recurrence_dev(Ctx* c, const double* xi);
ps_err sqrt_finish_dev(Ctx* c, double* out, ps_sqrt_info* info);
ps_err mob_apply(Ctx* c, const double* f, double* u, int prec);    // u = M f (device pointers)
// The Lanczos iteration's apply: y = M (f / beta), q = f / beta, beta^2 = the sum of the nb partials in
// red3 (stored to *beta2_out); with red3 null, y = M f and q is not written.  Then the partial dots
// red1[k nb + block] of Q_k . y for k < m (Q_{m-1} = q), nb = ceil(n / 256).
ps_err mob_apply_lanczos(Ctx* c, const double* f, double* y, double* qout, const double* red3, int nb,
                         double* beta2_out, int m, double* red1, int prec);
ps_err saddle_apply(Ctx* c, const double* x, double* y, int prec); // y = A x (device pointers)
ps_err mob_apply_stresslet_rows(Ctx* c, const double* f, double* u); // u[nu:n] = (M f)[nu:n], FP64
ps_err mob_apply_colloid_rows(Ctx* c, const double* f, double* u);   // u[nm3:n] = (M f)[nm3:n], FP64
ps_err apply_alloc(Ctx* c);
void apply_free(Ctx* c);
ps_err krylov_alloc(Ctx* c);
void krylov_free(Ctx* c);
ps_err saddle_diag_dev(Ctx* c, double* d);
ps_err saddle_solve_dev(Ctx* c, const double* b, double* x, int warm, ps_solve_info* info,
                        double* uout = nullptr);
ps_err sqrt_apply_dev(Ctx* c, const double* xi, double* w, ps_sqrt_info* info);
ps_err pairs_alloc(Ctx* c);
void pairs_free(Ctx* c);
// check_dist on the device; with xsave, the positions it reads are also copied there
ps_err check_dist_dev(Ctx* c, double* xsave = nullptr);
ps_err probe_geometry_dev(Ctx* c);
ps_err rhs_dev(Ctx* c, int drift);
ps_err rhs_drift_dev(Ctx* c, double* out);
ps_err vlist_copy(Ctx* c, int32_t* rows, int32_t* partners, int64_t* pidx, int64_t cap, int64_t* n_pairs);
void ps_set_context_free_error(const std::string& msg);           // for ps_last_error(NULL)
ps_err vlist_count(Ctx* c, int64_t* n_pairs);                      // pair count of the last check_dist (syncs)
ps_err run_steps(Ctx* c, int nsteps);
ps_err dense_build(Ctx* c, int storage);
ps_err dense_mult_dev(Ctx* c, const double* X, double* Y, int nrhs);   // Y_j = M X_j, vector-major
void dense_free(Ctx* c);
ps_err arrow_alloc(Ctx* c);
void arrow_free(Ctx* c);
ps_err arrow_check_dist(Ctx* c);                                   // Verlet list, AB + BB pairs only
ps_err arrow_count_pairs(Ctx* c);                                  // host copy of the pair count (syncs)
ps_err arrow_block_solve(Ctx* c, const double* b, double* x, double* uout);
ps_err arrow_mss_solve(Ctx* c, const double* b, const double* y, double* x);   // 5 x 5 stresslet system
ps_err arrow_mob(Ctx* c);                                          // snapshot + M^cm
ps_err arrow_schur(Ctx* c, int lazy);                              // sync_mcc_schur_correction
ps_err arrow_schur_redo(Ctx* c);                                   // the last one again, in full
ps_err arrow_apply(Ctx* c, const double* x, double* y, int saddle);
ps_err arrow_noise(Ctx* c, const double* xi, double* out);         // block-Cholesky slip noise
ps_err arrow_diag(Ctx* c, double* d, int inverse);
ps_err arrow_wca(Ctx* c);                                          // monomer WCA into the RHS
ps_err scan_ints(Ctx* c, const int* in, int* out, int n);          // out[n] = total, out[n+1] = max
ps_err lub_set_table(Ctx* c, int which, const double* rows, int nrows, int ncols);
ps_err lub_build(Ctx* c, double rcut, int64_t* npairs);
ps_err lub_pairs(Ctx* c, double rcut, int64_t* npairs);   // the pair list, without the mob() check
ps_err lub_alloc(Ctx* c);                                  // work buffers of the lubricated saddle
ps_err colloid_saddle_apply(Ctx* c, const double* v, double* y);   // the lubricated colloid saddle, 17 Nc
ps_err lub_saddle_term(Ctx* c, const double* x, double* y);   // y[n+nm3 : n+nu] -= R_lub x[n+nm3 : n+nu]
ps_err lub_apply_dev(Ctx* c, const double* uw, const double* einf, double* fts);
void lub_free(Ctx* c);
ps_err lub_interp_host(const double* rows, int nrows, int ncols, double r, double* out);
ps_err lub_slopes_host(const double* rows, int nrows, int ncols, double* slopes);
ps_err vec_axpby(Ctx* c, double* y, double a, const double* x, double b, int64_t n);  // y = a x + b y
ps_err vec_copy(Ctx* c, double* dst, const double* src, int64_t n);
ps_err vec_dot(Ctx* c, const double* a, const double* b, int64_t n, double* out);      // sync
