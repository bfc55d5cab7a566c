// batch.h -- layout shared by the batched-replica kernels (batch.hip) and their host side.
#pragma once
#include <cstdint>

#include "../../include/polystokes_gpu.h"

#include <hip/hip_runtime.h>

constexpr int BATCH_THREADS = 1024;   // one workgroup per replica
constexpr int BATCH_KMAX = 48;        // Lanczos iterations held on chip (LDS tridiagonal and U)
constexpr int BATCH_TEAM_MAX = 16;    // workgroups sharing one replica's applies

// Per-replica arrays in the FP64 pool (BO_NZ indexes the noise buffer instead).
enum BatchOff {
    BO_X, BO_XN, BO_XPD, BO_XM,         // positions, x_n, check_dist snapshot, mob() snapshot (3Np)
    BO_Q, BO_XP0,                       // quaternions (4Nc), trap centres (3Nc)
    BO_RHS, BO_XS, BO_XDET, BO_XD,      // saddle vectors (ns)
    BO_UP, BO_VDN, BO_VDC, BO_VB, BO_UD, BO_DV,   // velocities (nu): up, v_det_n, v_det_c, v_brown, udiff, drift
    BO_W, BO_Y, BO_LW,                  // n: square-root output, apply output, Lanczos work
    BO_QB,                              // Lanczos basis (kmax n)
    BO_SB,                              // Brownian stresslet (5Nc)
    BO_M5,                              // stresslet MINRES work (9 x 5Nc)
    BO_PART,                            // monomer-monomer partial sums per partner block (3 nblk Nm)
    BO_ST,                              // carried tile (18)
    BO_NZ,                              // this step's noise (n + nu), in the noise buffer
    BO_NOFF
};

struct BRep {
    int Np, Nm, Nc, Nmono;   // monomers in Nm / Nmono chains of Nmono consecutive beads, then colloids
    int si, sj;              // mob()'s last monomer-monomer pair (the carried tile), -1 without one
    int64_t n, nu, ns, nm3, nc3;
    int64_t off[BO_NOFF];
};

struct BConst {
    double beta, beta2, beta_inv, two_beta, dt, kT;
    double kbond, r0, Lmax, ktrap, k_hs, hs_exp_k;
    double sig[3], rverl[3];
    int fene, force_sel, mono_ev, chain_HI, trap_on;
    int kmax;      // Lanczos iterations (an invariant subspace ends it earlier)
    int mixed;     // FP32 monomer tiles in the square root's applies
    double ms_rtol;
    int ms_maxit;   // cap on stresslet MINRES iterations (past it the solve continues unconverged, as the
                    // reference does after KSPSolve's cap, and the replica's step counts it)
};

// Fork-join block of a team (one per team, 64 bytes): the leader publishes an operation and bumps
// seq; each member runs its slice and bumps done; err stops every later wait of the team.
struct BTeamCtl {
    int seq, done, err, op;
    const double* in;
    double* out;
    const double* mat;
    int64_t n;
    int64_t pad[3];
};

// Step schedule (device ints): per workgroup w, sched[5w..5w+4] = {team or -1, rank in the team,
// team size, first and end index of its solo replicas in sched[solo_off..]}; the replica of team k
// is sched[team_off + k].  A team's workgroups run that replica first (rank 0 leads), then each
// runs its own solo replicas.
struct BSchedDev {
    const int* sched;
    int team_off, solo_off;
    BTeamCtl* ctl;
    int* err;          // set when a team wait timed out
    double* times;     // per replica: wall-clock ticks of its last step (leader's view)
    double* phases;    // per replica: ticks in forces, mob(), saddle solves, square root, team joins
};

ps_err batch_launch_step(hipStream_t st, const BConst& B, const BRep* reps, const BSchedDev& sd, int nwg,
                         bool cooperative, double* pool, const double* noise, double ktrap, int* stats);
int batch_step_occupancy();   // co-resident step workgroups per CU
ps_err batch_launch_op(hipStream_t st, const BConst& B, const BRep* reps, int R, double* pool, int op,
                       const double* in, double* out, const int64_t* io_off, int* stats, int* err);
