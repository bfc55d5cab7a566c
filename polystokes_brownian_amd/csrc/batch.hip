// batch.hip -- batched replicas: many independent PolyStokes systems (uneven bead counts, each with
// its own noise seed) advanced by one persistent kernel per step.
//
// Each replica is an unbounded system of probes and chains with full monomer hydrodynamics, run
// exactly as a single context runs it (the reference's run() step, src/run.cpp:70-162: predictor,
// slip noise, RFD drift, corrector) with block-eliminated saddle solves and the Lanczos square
// root.  A workgroup of 1024 threads runs a replica's whole step (forces, mob()'s snapshot and
// carried tile, the grand-mobility applies, the stresslet MINRES, the Lanczos recurrence and its
// tridiagonal square root, the RFD probes and the integrator) with no host round trip; every
// block of the mobility is applied from positions, none is stored.  Replicas far above the mean
// work run on a team of workgroups that shares the applies, the stresslet products and the pair
// passes; the host packs the other replicas into the remaining workgroups (makespan schedule,
// re-tuned from measured per-replica times).
//
// The noise is drawn on the host from each replica's own stream (std::mt19937_64 + normal
// distribution, seeded as a single context is), so every replica reproduces an independent
// single-system run.  Lanczos runs a fixed number of iterations (sqrt_kmax; an invariant subspace
// ends it early), with two-pass classical Gram-Schmidt reorthogonalisation.
//
// Layout: one FP64 pool; each replica's arrays sit at its own offsets (BRep::off).  Positions are
// interleaved xyz; saddle vectors [f (n) ; u (nu)] as in a single context.
#include "ps_common.h"
#include "tiles.h"
#include "batch.h"
#include "tql2.h"
#include "dpp.h"

#pragma clang fp contract(off)

namespace {

constexpr int BW = BATCH_THREADS;   // threads per replica workgroup
constexpr int NW = BW / 64;
constexpr int BCCH = 512;   // colloids staged in LDS per pass of the monomer rows
constexpr int FTILE = 1024;   // force partners per LDS tile

// Wave sum, the same bits in every lane (DPP row sums and four lane reads); every lane active.
__device__ __forceinline__ double wsum(double v) { return wave_sum64(v); }
// Workgroup sum; every thread gets the same value (fixed order).  red: LDS, NW doubles.
__device__ double wg_sum(double v, double* red) {
    v = wsum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < NW; k++) s += red[k];
    return s;
}
__device__ double wg_dot(const double* a, const double* b, int64_t n, double* red) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += BW) s += a[i] * b[i];
    return wg_sum(s, red);
}

// ---------------------------------------------------------------------------------- teams
// A replica far above the mean work is run by a team of workgroups: the leader (rank 0) runs the
// step and forks each apply pass and stresslet product (M_ss v) to the members through its
// BTeamCtl; every rank runs the same slice function on its share and the leader joins.  The step
// kernel is a cooperative launch, so all workgroups are co-resident and a member always reaches
// its next wait.  Every wait is bounded (wall clock): a timeout sets the team's err, which ends
// every later wait of the team at once, and the batch's error flag, which the host reports.
enum TeamOp { TOP_EXIT = 1, TOP_APPLY_D, TOP_APPLY_DS, TOP_APPLY_F, TOP_MSS, TOP_ROWS_D, TOP_ROWS_F };
constexpr uint64_t TEAM_WAIT_TICKS = 1000000000ull;   // 10 s of the 100 MHz wall clock

struct TeamCtx {
    BTeamCtl* ctl;   // null for a workgroup alone
    int* gerr;
    int rank, g, nfork;
    uint64_t wait_ticks;   // the leader's thread 0: wall-clock ticks spent in joins
};

__device__ bool team_wait_ge(int* p, int target, int* err) {
    const uint64_t t0 = wall_clock64();
    for (;;) {
        if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) return true;
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
        if (wall_clock64() - t0 > TEAM_WAIT_TICKS) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(8);
    }
}
// Release point of a workgroup: every wave waits for its own memory operations, then the
// workgroup meets; one thread then publishes with an agent-scope release (a single L2 write-back
// for the workgroup instead of one per wave).
__device__ __forceinline__ void wg_release_point() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}
// Leader: publish op after every thread's earlier stores, then bump seq (release).
__device__ void team_fork(TeamCtx& tc, int op, const double* in, double* out, const double* mat, int64_t n) {
    wg_release_point();
    if (threadIdx.x == 0) {
        BTeamCtl* c = tc.ctl;
        c->op = op;
        c->in = in;
        c->out = out;
        c->mat = mat;
        c->n = n;
        __hip_atomic_fetch_add(&c->seq, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (op != TOP_EXIT) tc.nfork++;
}
// Leader: wait for every member's slice of the last fork; one agent-scope acquire by the waiting
// thread before the workgroup proceeds.
__device__ void team_join(TeamCtx& tc) {
    if (threadIdx.x == 0) {
        const uint64_t t0 = wall_clock64();
        if (!team_wait_ge(&tc.ctl->done, tc.nfork * (tc.g - 1), &tc.ctl->err))
            __hip_atomic_store(tc.gerr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        tc.wait_ticks += wall_clock64() - t0;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------- geometry
// Unbounded systems: d = x_k - x_j, dr, e = d / dr (the reference's pair_geom without a box).
__device__ __forceinline__ void bgeom(const double* xs, int k, int j, double e[3], double& dr, double& dr_inv) {
    const double dx = xs[3 * k] - xs[3 * j], dy = xs[3 * k + 1] - xs[3 * j + 1], dz = xs[3 * k + 2] - xs[3 * j + 2];
    dr = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
    dr_inv = 1 / dr;
    e[0] = dx * dr_inv;
    e[1] = dy * dr_inv;
    e[2] = dz * dr_inv;
}
__device__ __forceinline__ int btype(int Nm, int k, int j) { return (j < Nm) ? 0 : (k < Nm ? 1 : 2); }

// ---------------------------------------------------------------------------------- forces
// rhs = [0 (n) ; fint (3Np) ; 0 (3Nc)]: excluded volume over the pairs a Verlet list would hold
// (dr below the pair type's cutoff, at the snapshot positions), bonds along each chain, the static
// trap on the colloids (src/rhs.cpp:160-260, src/pair_interaction.h:13-72).
// force partners' positions, one LDS tile (the force pass is not inlined into the step, so its
// loops get registers of their own; the tile is referenced directly so it stays LDS)
__shared__ double g_ftile[3 * FTILE];

__device__ __noinline__ void wg_forces(const BConst& B, const BRep& R, const double* xpd, const double* x,
                                       const double* xpos0, double ktrap, double* rhs) {
    double* tile = g_ftile;
    const int Np = R.Np, Nm = R.Nm, t = threadIdx.x;
    for (int64_t i = t; i < R.ns; i += BW)
        if (i < R.n || i >= R.n + 3 * (int64_t)Np) rhs[i] = 0.0;
    for (int p0 = 0; p0 < Np; p0 += BW) {
        const int p = p0 + t;
        const bool act = p < Np;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        const double xp0 = act ? xpd[3 * p] : 0.0, xp1 = act ? xpd[3 * p + 1] : 0.0, xp2 = act ? xpd[3 * p + 2] : 0.0;
        // without monomer excluded volume a monomer's partners are the colloids (same order)
        const int qs = (!B.mono_ev && p < Nm) ? Nm : 0;
        // per pair type (registers, not an indexed kernel argument)
        const double rv0 = B.rverl[0], rv1 = B.rverl[1], rv2 = B.rverl[2];
        const double sg0 = B.sig[0], sg1 = B.sig[1], sg2 = B.sig[2];
        // partners q in increasing order, their positions staged in LDS tiles of FTILE
        for (int q0 = 0; q0 < Np; q0 += FTILE) {
            const int qn = min(FTILE, Np - q0);
            __syncthreads();
            for (int k = t; k < 3 * qn; k += BW) tile[k] = xpd[3 * (int64_t)q0 + k];
            __syncthreads();
            if (!act) continue;
            for (int q = max(q0, qs); q < q0 + qn; q++) {
                if (q == p) continue;
                const int k = min(p, q), j = max(p, q);
                const int type = btype(Nm, k, j);
                if (type == 0 && !B.mono_ev) continue;
                const double* xq = tile + 3 * (q - q0);
                // d = x_k - x_j as the reference's pair_geom
                const double sgn = (p == k) ? 1.0 : -1.0;
                const double dx = sgn * (xp0 - xq[0]), dy = sgn * (xp1 - xq[1]), dz = sgn * (xp2 - xq[2]);
                const double r2 = dx * dx + dy * dy + dz * dz;
                // far pairs: r^2 above rverl^2 (1 + 1e-12) has sqrt(r^2) > rverl, so the exact
                // test below would reject it too
                const double rv = type == 0 ? rv0 : type == 1 ? rv1 : rv2;
                if (r2 > rv * rv * (1.0 + 1e-12)) continue;
                const double dr = __builtin_sqrt(r2), dr_inv = 1 / dr;
                const double e[3] = {dx * dr_inv, dy * dr_inv, dz * dr_inv};
                if (!(dr < rv)) continue;
                const double sig_eff = (type == 0 ? sg0 : type == 1 ? sg1 : sg2) * (1.0 + 0.01);
                double fmag;
                if (B.force_sel == PS_FORCE_HS_EXP) {
                    const double ex_ = exp(-B.hs_exp_k * (dr - sig_eff));
                    fmag = ex_ / (1 - ex_);
                } else {
                    fmag = (dr >= sig_eff) ? 0.0 : B.k_hs * (sig_eff - dr);
                }
                if (fmag == 0.0) continue;
                const double gx = fmag * e[0], gy = fmag * e[1], gz = fmag * e[2];
                if (p == k) {
                    fx -= gx;
                    fy -= gy;
                    fz -= gz;
                } else {
                    fx += gx;
                    fy += gy;
                    fz += gz;
                }
            }
        }
        if (!act) continue;
        if (p < Nm) {   // bonds (bond_forces, src/rhs.cpp:102-158): p - 1 and p + 1 in the chain
            const int c0 = (p / R.Nmono) * R.Nmono;
            for (int side = 0; side < 2; side++) {
                const int k = side == 0 ? p - 1 : p, j = side == 0 ? p : p + 1;
                if (k < c0 || j >= c0 + R.Nmono) continue;
                const double r0_ = x[3 * k] - x[3 * j], r1_ = x[3 * k + 1] - x[3 * j + 1], r2_ = x[3 * k + 2] - x[3 * j + 2];
                const double r = __builtin_sqrt(r0_ * r0_ + r1_ * r1_ + r2_ * r2_);
                double f0 = -B.kbond * (1 - (B.r0 / r)) * r0_;
                double f1 = -B.kbond * (1 - (B.r0 / r)) * r1_;
                double f2 = -B.kbond * (1 - (B.r0 / r)) * r2_;
                if (B.fene) {
                    const double ext2 = 1 / (1 - (r / B.Lmax) * (r / B.Lmax));
                    f0 *= ext2;
                    f1 *= ext2;
                    f2 *= ext2;
                }
                if (p == k) {
                    fx -= f0;
                    fy -= f1;
                    fz -= f2;
                } else {
                    fx += f0;
                    fy += f1;
                    fz += f2;
                }
            }
        } else if (B.trap_on) {   // trapping_forces (src/rhs.cpp:46-100)
            const int c = p - Nm;
            fx += ktrap * (x[3 * p] - xpos0[3 * c]);
            fy += ktrap * (x[3 * p + 1] - xpos0[3 * c + 1]);
            fz += ktrap * (x[3 * p + 2] - xpos0[3 * c + 2]);
        }
        rhs[R.n + 3 * (int64_t)p] = fx;
        rhs[R.n + 3 * (int64_t)p + 1] = fy;
        rhs[R.n + 3 * (int64_t)p + 2] = fz;
    }
}

// ---------------------------------------------------------------------------------- mob()
// The carried tile (the mob_a every AB pair sees, src/mob.cpp:926-958): the tile of the last AA
// pair of this mob() (si, sj), else the previous call's leftover; st[9..17] keeps what the next call
// inherits (the last BB pair's tile, or this one).  Thread 0.
__device__ void b_stale(const BConst& B, const BRep& R, const double* xm, double* st) {
    double cur[9];
    if (R.si >= 0) {
        double e[3], dr, dr_inv;
        bgeom(xm, R.si, R.sj, e, dr, dr_inv);
        const double dr_inv2 = dr_inv * dr_inv, dr_inv3 = dr_inv2 * dr_inv;
        double x12a, y12a;
        if (dr > B.two_beta) {
            x12a = C3D8 * dr_inv - 2.0 * dr_inv3;
            y12a = C3D4 * dr_inv - C1D2 * dr_inv3;
        } else {
            x12a = C3D8 * dr - 1.0;
            y12a = 1.0 - C9D32 * dr;
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const double ee = e[i] * e[j];
                cur[3 * i + j] = (i == j) ? x12a * ee + y12a * (1. - ee) : x12a * ee - y12a * ee;
            }
    } else {
        for (int k = 0; k < 9; k++) cur[k] = st[9 + k];
    }
    for (int k = 0; k < 9; k++) st[k] = cur[k];
    if (R.Nc >= 2) {
        double e[3], dr, dr_inv;
        bgeom(xm, R.Np - 2, R.Np - 1, e, dr, dr_inv);
        const double dr_inv2 = dr_inv * dr_inv, dr_inv3 = dr_inv2 * dr_inv;
        const double x12a = C3D2 * dr_inv - dr_inv3, y12a = C3D4 * dr_inv + C1D2 * dr_inv3;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const double ee = e[i] * e[j];
                st[9 + 3 * i + j] = (i == j) ? x12a * ee + y12a * (1. - ee) : x12a * ee - y12a * ee;
            }
    } else {
        for (int k = 0; k < 9; k++) st[9 + k] = cur[k];
    }
}
// ---------------------------------------------------------------------------------- apply
// Fused multiply-add, written out where the tiles want it (this file is built without contraction,
// which the pair geometry and the integrator need).
__device__ __forceinline__ float b_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double b_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Wave rotation (DPP wave_rol:1): lane t receives lane t + 1's value.
__device__ __forceinline__ float b_rol1(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x134, 0xF, 0xF, false));
}
__device__ __forceinline__ double b_rol1(double v) {
    const long long q = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(q & 0xffffffff), 0x134, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(q >> 32), 0x134, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// The AA tile's coefficients at r^2 (lengths in monomer radii): ya = 3/(4r) - 1/(2r^3) and
// cc = -3/(8r^3) - 3/(2r^5), or the overlap forms 1 - 9r/32 and (21r/32 - 2)/r^2 for r <= 2 (the
// overlap branch is taken only when some lane of the wave needs it).
template <class T>
__device__ __forceinline__ void aa_coef(T r2, T tb2, T& ya, T& cc) {
    const T rinv = ps_rsqrt(r2);
    const T rinv2 = rinv * rinv;
    ya = rinv * b_fma(-(T)C1D2, rinv2, (T)C3D4);
    cc = (rinv * rinv2) * b_fma(-(T)C3D2, rinv2, -(T)C3D8);
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(r2 <= tb2) != 0, 0)) {
        if (r2 <= tb2) {
            const T r = r2 * rinv;
            ya = b_fma(-(T)C9D32, r, (T)1);
            cc = b_fma((T)(21. / 32.), r, -(T)2) * rinv2;
        }
    }
}

// Monomer-monomer sums of u = M f (mobility(..., AA=true), src/mob.cpp:122-131: y f + (x - y)/r^2
// (d.f) d) by 64 x 64 blocks, one block pair per wave (pairs rank 16 + w, + 16 g, ...):
// part[(3K + c) Nm + i] = the sum over the sources of block K for target i.  The pair tile is
// symmetric, so a pair I < J is evaluated once: lanes hold the targets of I, and at step m lane l
// takes source (l + m) mod 64 of J from the wave's LDS window and adds the tile to u_i; the
// source-side sum u_j travels with the sources by a wave rotation, so after 64 steps lane l holds
// it for source (l + 63) mod 64.  Diagonal blocks (and every block pair when chain_HI skips same-
// chain pairs) run one way.  Padding beads sit far apart with zero force.  stage: LDS, 16 windows
// of 64 sources.
template <class T>
__device__ __forceinline__ void aa_pairs(const BConst& B, const BRep& R, const double* xm, const double* f,
                                         double* part, double* stage, int rank, int g) {
    const int Nm = R.Nm, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nblk = (Nm + 63) >> 6;
    constexpr int SW = sizeof(T) == 8 ? 6 : 8;
    T* win = reinterpret_cast<T*>(stage) + wv * 64 * SW;
    const T tb2 = (T)(B.two_beta * B.two_beta), FAR = (T)1e10;
    const bool chain = B.chain_HI != 0;
    const int64_t nitem = chain ? (int64_t)nblk * nblk : (int64_t)nblk * (nblk + 1) / 2;
    auto first = [nblk](int64_t I) { return I * nblk - I * (I - 1) / 2; };   // items before row I
    auto decode = [&](int64_t it, int& I, int& J) {
        if (chain) {
            I = (int)(it / nblk);
            J = (int)(it - (int64_t)I * nblk);
        } else {
            const double m2 = 2.0 * nblk + 1.0;
            int64_t r = (int64_t)((m2 - __builtin_sqrt(fmax(m2 * m2 - 8.0 * (double)it, 0.0))) * 0.5);
            if (r < 0) r = 0;
            while (r > 0 && first(r) > it) r--;
            while (r + 1 < nblk && first(r + 1) <= it) r++;
            I = (int)r;
            J = (int)(I + (it - first(r)));
        }
    };
    // this lane's target and source beads of an item, loaded one item ahead (the loads of the next
    // item are in flight while the current one runs)
    double pt[6], ps[6];
    auto load = [&](int I, int J) {
        const int i = (I << 6) + lane, j = (J << 6) + lane;
        const bool ai = i < Nm, aj = j < Nm;
        for (int q = 0; q < 3; q++) {
            pt[q] = ai ? xm[3 * (int64_t)i + q] : (double)FAR;
            pt[3 + q] = ai ? f[3 * (int64_t)i + q] : 0.0;
            ps[q] = aj ? xm[3 * (int64_t)j + q] : -(double)FAR;
            ps[3 + q] = aj ? f[3 * (int64_t)j + q] : 0.0;
        }
    };
    int64_t it = (int64_t)rank * NW + wv;
    int In = 0, Jn = 0;
    if (it < nitem) {
        decode(it, In, Jn);
        load(In, Jn);
    }
    for (; it < nitem; it += (int64_t)g * NW) {
        const int I = In, J = Jn;
        const int i = (I << 6) + lane;
        const bool ai = i < Nm;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");   // the previous item's window reads
        {
            T* w = win + SW * lane;
            for (int q = 0; q < 6; q++) w[q] = (T)ps[q];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const T xi = (T)pt[0], yi = (T)pt[1], zi = (T)pt[2];
        const T fxi = (T)pt[3], fyi = (T)pt[4], fzi = (T)pt[5];
        const int64_t itn = it + (int64_t)g * NW;
        if (itn < nitem) {
            decode(itn, In, Jn);
            load(In, Jn);
        }
        T ux = 0, uy = 0, uz = 0;
        if (I == J || chain) {
            const int ci = i / R.Nmono;
            for (int m = 0; m < 64; m++) {
                const T* sj = win + SW * m;
                const int jj = (J << 6) + m;
                bool skip = jj == i || jj >= Nm;
                if (chain) skip = skip || jj / R.Nmono == ci;
                const T dx = xi - sj[0], dy = yi - sj[1], dz = zi - sj[2];
                T r2 = b_fma(dz, dz, b_fma(dy, dy, dx * dx));
                r2 = skip ? (T)1 : r2;
                T ya, cc;
                aa_coef(r2, tb2, ya, cc);
                ya = skip ? (T)0 : ya;
                cc = skip ? (T)0 : cc;
                const T fx = sj[3], fy = sj[4], fz = sj[5];
                const T cs = cc * b_fma(dz, fz, b_fma(dy, fy, dx * fx));
                ux = b_fma(cs, dx, b_fma(ya, fx, ux));
                uy = b_fma(cs, dy, b_fma(ya, fy, uy));
                uz = b_fma(cs, dz, b_fma(ya, fz, uz));
            }
        } else {
            T Rx = 0, Ry = 0, Rz = 0;
            for (int m = 0; m < 64; m++) {
                const T* sj = win + SW * ((lane + m) & 63);
                const T dx = xi - sj[0], dy = yi - sj[1], dz = zi - sj[2];
                const T r2 = b_fma(dz, dz, b_fma(dy, dy, dx * dx));
                T ya, cc;
                aa_coef(r2, tb2, ya, cc);
                const T fx = sj[3], fy = sj[4], fz = sj[5];
                const T cs = cc * b_fma(dz, fz, b_fma(dy, fy, dx * fx));
                ux = b_fma(cs, dx, b_fma(ya, fx, ux));
                uy = b_fma(cs, dy, b_fma(ya, fy, uy));
                uz = b_fma(cs, dz, b_fma(ya, fz, uz));
                // u_j += ya f_i + (d.f_i)/r^2 d (the same tile: d -> -d twice)
                const T cj = cc * b_fma(dz, fzi, b_fma(dy, fyi, dx * fxi));
                Rx = b_rol1(Rx) + b_fma(cj, dx, ya * fxi);
                Ry = b_rol1(Ry) + b_fma(cj, dy, ya * fyi);
                Rz = b_rol1(Rz) + b_fma(cj, dz, ya * fzi);
            }
            const int jo = (J << 6) + ((lane + 63) & 63);
            if (jo < Nm) {
                part[(3 * (int64_t)I + 0) * Nm + jo] = (double)Rx;
                part[(3 * (int64_t)I + 1) * Nm + jo] = (double)Ry;
                part[(3 * (int64_t)I + 2) * Nm + jo] = (double)Rz;
            }
        }
        if (ai) {
            part[(3 * (int64_t)J + 0) * Nm + i] = (double)ux;
            part[(3 * (int64_t)J + 1) * Nm + i] = (double)uy;
            part[(3 * (int64_t)J + 2) * Nm + i] = (double)uz;
        }
    }
}

// u = M f for the grand mobility of src/mob.cpp:635-1213 (the layout of a single context; the
// reference's quirks: the carried translation tile, the '+' of mob_c, the 2-1 gt sign).  T is the
// arithmetic of the monomer-monomer and monomer-colloid tiles (float in the mixed form's square
// root); colloid-colloid tiles are FP64.  `srows`: only the stresslet rows (nu..n).
//
// Two passes (a team joins between them).  sel 0: the monomer-monomer block pairs (aa_pairs) and
// the rows of colloids c = rank, rank + g, ... (one wave per colloid: the monomer sources from LDS
// tiles over the lanes, then the colloid sources over the lanes, coupled in contracted form).
// sel 1: the monomer rows in blocks of 64 targets b = rank, rank + g, ...: the colloid sources
// (wave w taking colloids w, w + 16, ... from LDS), the partner blocks' partial sums, the carried
// tile and the self term.  stage: LDS, 6 BW doubles; cstage: 11 BCCH doubles.
template <class T>
__device__ __forceinline__ void apply_slice(const BConst& B, const BRep& R, double* pool, const double* f, double* u,
                                            bool srows, int sel, double* stage, double* cstage, int rank, int g) {
    const double* xm = pool + R.off[BO_XM];
    const double* st = pool + R.off[BO_ST];
    const int Nm = R.Nm, Nc = R.Nc, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int64_t nm3 = R.nm3, nu = R.nu;
    const T gA = (T)(C9D20 * (3 + 5 * B.beta2));
    const T gB = (T)(C9D10 * (C1D2 + C5D6 * B.beta2));
    double* part = pool + R.off[BO_PART];
    if (!srows && Nm > 0 && sel == 0) aa_pairs<T>(B, R, xm, f, part, stage, rank, g);
    if (!srows && Nm > 0 && sel == 1) {
        // the carried tile applied to the sum of colloid forces (the same in every monomer row):
        // a workgroup sum in a fixed order
        double F0 = 0.0, F1 = 0.0, F2 = 0.0;
        for (int c = t; c < Nc; c += BW) {
            F0 += f[nm3 + 3 * c];
            F1 += f[nm3 + 3 * c + 1];
            F2 += f[nm3 + 3 * c + 2];
        }
        F0 = wsum(F0);
        F1 = wsum(F1);
        F2 = wsum(F2);
        __syncthreads();
        if (lane == 0) {
            stage[wv] = F0;
            stage[NW + wv] = F1;
            stage[2 * NW + wv] = F2;
        }
        __syncthreads();
        F0 = F1 = F2 = 0.0;
        for (int w = 0; w < NW; w++) {
            F0 += stage[w];
            F1 += stage[NW + w];
            F2 += stage[2 * NW + w];
        }
        const double car0 = st[0] * F0 + st[1] * F1 + st[2] * F2;
        const double car1 = st[3] * F0 + st[4] * F1 + st[5] * F2;
        const double car2 = st[6] * F0 + st[7] * F1 + st[8] * F2;
        const int nblk = (Nm + 63) >> 6;
        for (int b = rank; b < nblk; b += g) {
            const int i = (b << 6) + lane;
            const bool act = i < Nm;
            T xi = 0, yi = 0, zi = 0;
            if (act) {
                xi = (T)xm[3 * i];
                yi = (T)xm[3 * i + 1];
                zi = (T)xm[3 * i + 2];
            }
            // colloid sources (mobility_AB, src/mob.cpp:362-537): rotlet and stresslet couplings,
            // [x T S] of up to BCCH colloids staged in LDS per pass, wave w taking w, w + 16, ...
            T vx = 0, vy = 0, vz = 0;
            T* cs = reinterpret_cast<T*>(cstage);
            for (int c0 = 0; c0 < Nc; c0 += BCCH) {
                const int cn = min(BCCH, Nc - c0);
                // with every colloid in one chunk, the first block's staging serves the others
                const bool stage_now = Nc > BCCH || b == rank;
                if (stage_now) __syncthreads();
                if (stage_now)
                    for (int idx = t; idx < 11 * cn; idx += BW) {
                        const int cc = idx / 11, q = idx - 11 * (idx / 11), c = c0 + cc;
                        const double v = q < 3 ? xm[3 * (int64_t)(Nm + c) + q]
                                               : q < 6 ? f[nm3 + 3 * (int64_t)Nc + 3 * c + (q - 3)]
                                                       : f[nu + 5 * (int64_t)c + (q - 6)];
                        cs[idx] = (T)v;
                    }
                if (stage_now) __syncthreads();
                if (!act) continue;
                for (int cc = wv; cc < cn; cc += NW) {
                    const T* cv = cs + 11 * cc;
                    const T dx = xi - cv[0], dy = yi - cv[1], dz = zi - cv[2];
                    const T r2 = b_fma(dz, dz, b_fma(dy, dy, dx * dx));
                    const T rinv = ps_rsqrt(r2);
                    const T ex = dx * rinv, ey = dy * rinv, ez = dz * rinv;
                    const T rinv2 = rinv * rinv, rinv4 = rinv2 * rinv2;
                    const T y12b = -(T)C3D4 * rinv2;
                    const T x12g = b_fma(-gA, rinv4, (T)C9D4 * rinv2);
                    const T y12g = gB * rinv4;
                    const T Tx = cv[3], Ty = cv[4], Tz = cv[5];
                    const T S0 = cv[6], S1 = cv[7], S2 = cv[8], S3 = cv[9], S4 = cv[10];
                    const T sex = b_fma(S2, ez, b_fma(S1, ey, S0 * ex));
                    const T sey = b_fma(S3, ez, b_fma(S4, ey, S1 * ex));
                    const T sez = b_fma(-(S0 + S4), ez, b_fma(S3, ey, S2 * ex));
                    const T ese = b_fma(ez, sez, b_fma(ey, sey, ex * sex));
                    const T a = -b_fma((T)2, y12g, x12g) * ese;
                    const T b2 = (T)2 * y12g;
                    vx = b_fma(b2, sex, b_fma(a, ex, b_fma(y12b, b_fma(ey, Tz, -(ez * Ty)), vx)));
                    vy = b_fma(b2, sey, b_fma(a, ey, b_fma(y12b, b_fma(ez, Tx, -(ex * Tz)), vy)));
                    vz = b_fma(b2, sez, b_fma(a, ez, b_fma(y12b, b_fma(ex, Ty, -(ey * Tx)), vz)));
                }
            }
            // wave w adds the monomer partial sums of partner blocks w, w + 16, ...; then the 16
            // sums of each target, added in wave order
            double p0 = (double)vx, p1 = (double)vy, p2 = (double)vz;
            if (act)
                for (int K = wv; K < nblk; K += NW) {
                    p0 += part[(3 * (int64_t)K + 0) * Nm + i];
                    p1 += part[(3 * (int64_t)K + 1) * Nm + i];
                    p2 += part[(3 * (int64_t)K + 2) * Nm + i];
                }
            __syncthreads();
            double* pr = stage;
            pr[(0 * NW + wv) * 64 + lane] = p0;
            pr[(1 * NW + wv) * 64 + lane] = p1;
            pr[(2 * NW + wv) * 64 + lane] = p2;
            __syncthreads();
            if (wv == 0 && act) {
                double s0 = 0.0, s1 = 0.0, s2 = 0.0;
                for (int w = 0; w < NW; w++) {
                    s0 += pr[(0 * NW + w) * 64 + lane];
                    s1 += pr[(1 * NW + w) * 64 + lane];
                    s2 += pr[(2 * NW + w) * 64 + lane];
                }
                u[3 * i] = B.beta_inv * f[3 * i] + (s0 + car0);
                u[3 * i + 1] = B.beta_inv * f[3 * i + 1] + (s1 + car1);
                u[3 * i + 2] = B.beta_inv * f[3 * i + 2] + (s2 + car2);
            }
        }
    }
    if (sel == 1) return;
    // colloid rows, 16 at a time (wave w: this rank's colloid r0 + w).  Monomer sources (AB: rotation
    // and stresslet rows, and the sum of monomer forces for the transposed carried tile) from LDS
    // tiles of 1024 monomers shared by the 16 waves, over the lanes; then the colloid sources over
    // the lanes (the contracted colloid-colloid coupling, FP64).
    const int ncol = Nc > rank ? (Nc - rank + g - 1) / g : 0;
    T* sx = reinterpret_cast<T*>(stage);
    for (int r0 = 0; r0 < ncol; r0 += NW) {
        const int kc = r0 + wv;
        const bool cact = kc < ncol;
        const int c = rank + kc * g;
        const int gc = Nm + (cact ? c : 0);
        const T xc = (T)xm[3 * gc], yc = (T)xm[3 * gc + 1], zc = (T)xm[3 * gc + 2];
        T acc[11];
        for (int k = 0; k < 11; k++) acc[k] = 0;
        T fsx = 0, fsy = 0, fsz = 0;
        for (int m0 = 0; m0 < Nm; m0 += BW) {
            const int mn = min(BW, Nm - m0);
            __syncthreads();
            if (t < mn) {
                const int64_t m = m0 + t;
                sx[t] = (T)xm[3 * m];
                sx[BW + t] = (T)xm[3 * m + 1];
                sx[2 * BW + t] = (T)xm[3 * m + 2];
                sx[3 * BW + t] = (T)f[3 * m];
                sx[4 * BW + t] = (T)f[3 * m + 1];
                sx[5 * BW + t] = (T)f[3 * m + 2];
            }
            __syncthreads();
            if (!cact) continue;
            for (int mm = lane; mm < mn; mm += 64) {
                const T dx = sx[mm] - xc, dy = sx[BW + mm] - yc, dz = sx[2 * BW + mm] - zc;
                const T r2 = b_fma(dz, dz, b_fma(dy, dy, dx * dx));
                const T rinv = ps_rsqrt(r2);
                const T ex = dx * rinv, ey = dy * rinv, ez = dz * rinv;
                const T rinv2 = rinv * rinv, rinv4 = rinv2 * rinv2;
                const T y12b = -(T)C3D4 * rinv2;
                const T x12g = b_fma(-gA, rinv4, (T)C9D4 * rinv2);
                const T y12g = gB * rinv4, y2 = (T)2 * y12g;
                const T fx = sx[3 * BW + mm], fy = sx[4 * BW + mm], fz = sx[5 * BW + mm];
                fsx += fx;
                fsy += fy;
                fsz += fz;
                acc[3] = b_fma(y12b, b_fma(fy, ez, -(fz * ey)), acc[3]);
                acc[4] = b_fma(y12b, b_fma(fz, ex, -(fx * ez)), acc[4]);
                acc[5] = b_fma(y12b, b_fma(fx, ey, -(fy * ex)), acc[5]);
                const T ef = b_fma(ez, fz, b_fma(ey, fy, ex * fx));
                const T p = -b_fma((T)2, y12g, x12g) * ef;
                const T pex = p * ex, pey = p * ey;
                const T G00 = b_fma(pex, ex, y2 * (ex * fx));
                const T G11 = b_fma(pey, ey, y2 * (ey * fy));
                const T G22 = b_fma(p * ez, ez, y2 * (ez * fz));
                const T G01 = b_fma(pex, ey, y12g * b_fma(ex, fy, ey * fx));
                const T G02 = b_fma(pex, ez, y12g * b_fma(ex, fz, ez * fx));
                const T G12 = b_fma(pey, ez, y12g * b_fma(ey, fz, ez * fy));
                acc[6] += G00 - G22;
                acc[7] = b_fma((T)2, G01, acc[7]);
                acc[8] = b_fma((T)2, G02, acc[8]);
                acc[9] = b_fma((T)2, G12, acc[9]);
                acc[10] += G11 - G22;
            }
        }
        if (!cact) continue;
        double av[11];
        for (int k = 3; k < 11; k++) av[k] = wsum((double)acc[k]);
        const double fs[3] = {wsum((double)fsx), wsum((double)fsy), wsum((double)fsz)};
        double gm[11];
        {
            double U[3] = {0.0, 0.0, 0.0}, Wr[3] = {0.0, 0.0, 0.0}, Sr[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            const double xc0 = xm[3 * gc], xc1 = xm[3 * gc + 1], xc2 = xm[3 * gc + 2];
            for (int s2 = lane; s2 < Nc; s2 += 64) {
                if (s2 == c) continue;
                const int gs = Nm + s2;
                // e = (x_lo - x_hi) / r
                const double sgn = (c < s2) ? 1.0 : -1.0;
                const double dx = sgn * (xc0 - xm[3 * gs]), dy = sgn * (xc1 - xm[3 * gs + 1]),
                             dz = sgn * (xc2 - xm[3 * gs + 2]);
                const double dr = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
                const double ri = 1 / dr;
                const double e[3] = {dx * ri, dy * ri, dz * ri};
                const double* Fs = f + nm3 + 3 * (int64_t)s2;
                const double* Ts = f + nm3 + 3 * (int64_t)Nc + 3 * (int64_t)s2;
                const double* Ss = f + nu + 5 * (int64_t)s2;
                const double Fv[3] = {Fs[0], Fs[1], Fs[2]}, Tv[3] = {Ts[0], Ts[1], Ts[2]};
                const double Sv[5] = {Ss[0], Ss[1], Ss[2], Ss[3], Ss[4]};
                if (srows)
                    cc_couple<true>(e, ri, sgn, Fv, Tv, Sv, U, Wr, Sr);
                else
                    cc_couple<false>(e, ri, sgn, Fv, Tv, Sv, U, Wr, Sr);
            }
            if (!srows)
                for (int q = 0; q < 3; q++) {
                    gm[q] = wsum(U[q]);
                    gm[3 + q] = wsum(Wr[q]);
                }
            for (int q = 0; q < 5; q++) gm[6 + q] = wsum(Sr[q]);
        }
        if (lane != 0) continue;
        if (!srows) {
            for (int comp = 0; comp < 3; comp++) {
                const int64_t r = nm3 + 3 * (int64_t)c + comp;
                const double carried = st[0 * 3 + comp] * fs[0] + st[1 * 3 + comp] * fs[1] + st[2 * 3 + comp] * fs[2];
                u[r] = 1.0 * f[r] + carried + gm[comp];
            }
            for (int comp = 0; comp < 3; comp++) {
                const int64_t r = nm3 + 3 * (int64_t)Nc + 3 * (int64_t)c + comp;
                u[r] = (3. / 4.) * f[r] + av[3 + comp] + gm[3 + comp];
            }
        }
        const double* S = f + nu + 5 * (int64_t)c;
        for (int comp = 0; comp < 5; comp++) {
            double self = (9. / 5.) * S[comp];
            if (comp == 0) self += (9. / 10.) * S[4];
            if (comp == 4) self += (9. / 10.) * S[0];
            u[nu + 5 * (int64_t)c + comp] = self + av[6 + comp] + gm[6 + comp];
        }
    }
}

// y = M_ss v without storing M_ss (5Nc x 5Nc: the stresslet self blocks and every colloid pair's
// m tile, placement of src/mob.cpp:1129-1143): colloids c = rank 16 + w, + 16 g, ... one wave
// per colloid, the source colloids over the lanes, m applied in contracted form (cc_mS, even in e).
__device__ __forceinline__ void mss_slice(const BRep& R, const double* P, const double* v, double* y, int rank, int g) {
    const int Nc = R.Nc, Nm = R.Nm, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const double* xm = P + R.off[BO_XM];
    for (int c = rank * NW + wv; c < Nc; c += g * NW) {
        const int gc = Nm + c;
        const double xc0 = xm[3 * gc], xc1 = xm[3 * gc + 1], xc2 = xm[3 * gc + 2];
        double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int s2 = lane; s2 < Nc; s2 += 64) {
            if (s2 == c) continue;
            const int gs = Nm + s2;
            const double dx = xc0 - xm[3 * gs], dy = xc1 - xm[3 * gs + 1], dz = xc2 - xm[3 * gs + 2];
            const double dr = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
            const double ri = 1 / dr;
            const double e[3] = {dx * ri, dy * ri, dz * ri};
            const double ri3 = ri * ri * ri, ri5 = ri3 * ri * ri;
            const double* vs = v + 5 * (int64_t)s2;
            const double Sv[5] = {vs[0], vs[1], vs[2], vs[3], vs[4]};
            double m[5];
            cc_mS(e, ri3, ri5, Sv, m);
            for (int q = 0; q < 5; q++) acc[q] += m[q];
        }
        for (int q = 0; q < 5; q++) acc[q] = wsum(acc[q]);
        if (lane == 0) {
            const double* vc = v + 5 * (int64_t)c;
            double* yc = y + 5 * (int64_t)c;
            yc[0] = ((9. / 5.) * vc[0] + (9. / 10.) * vc[4]) + acc[0];
            yc[1] = (9. / 5.) * vc[1] + acc[1];
            yc[2] = (9. / 5.) * vc[2] + acc[2];
            yc[3] = (9. / 5.) * vc[3] + acc[3];
            yc[4] = ((9. / 5.) * vc[4] + (9. / 10.) * vc[0]) + acc[4];
        }
    }
}

// One team operation's slice (leaders, members and lone workgroups all run it; not inlined, so the
// step kernel holds one copy of each slice).
__device__ __noinline__ void run_slice(const BConst& B, const BRep& R, double* P, int op, const double* in,
                                       double* out, const double* mat, int64_t n, int rank, int g) {
    __shared__ double stage[6 * BW];
    __shared__ double cstage[11 * BCCH];
    if (op == TOP_APPLY_D)
        apply_slice<double>(B, R, P, in, out, false, 0, stage, cstage, rank, g);
    else if (op == TOP_ROWS_D)
        apply_slice<double>(B, R, P, in, out, false, 1, stage, cstage, rank, g);
    else if (op == TOP_APPLY_DS)
        apply_slice<double>(B, R, P, in, out, true, 0, stage, cstage, rank, g);
    else if (op == TOP_APPLY_F)
        apply_slice<float>(B, R, P, in, out, false, 0, stage, cstage, rank, g);
    else if (op == TOP_ROWS_F)
        apply_slice<float>(B, R, P, in, out, false, 1, stage, cstage, rank, g);
    else if (op == TOP_MSS)
        mss_slice(R, P, in, out, rank, g);
}

// One split operation on this workgroup (alone) or its team (as the leader).
__device__ void team_op(const BConst& B, const BRep& R, double* P, TeamCtx& tc, int op, const double* in,
                        double* out, const double* mat, int64_t n) {
    if (tc.g > 1) {
        team_fork(tc, op, in, out, mat, n);
        run_slice(B, R, P, op, in, out, mat, n, 0, tc.g);
        team_join(tc);
    } else {
        run_slice(B, R, P, op, in, out, mat, n, 0, 1);
        __syncthreads();
    }
}

template <class T>
__device__ void wg_apply(const BConst& B, const BRep& R, double* P, const double* f, double* u, bool srows,
                         TeamCtx& tc) {
    if (srows) {
        team_op(B, R, P, tc, TOP_APPLY_DS, f, u, nullptr, 0);
        return;
    }
    team_op(B, R, P, tc, sizeof(T) == 4 ? TOP_APPLY_F : TOP_APPLY_D, f, u, nullptr, 0);
    team_op(B, R, P, tc, sizeof(T) == 4 ? TOP_ROWS_F : TOP_ROWS_D, f, u, nullptr, 0);
}

// mob(): snapshot xm = xpd and the carried tile (every other block of the grand mobility is applied
// from positions).
__device__ void wg_mob(const BConst& B, const BRep& R, double* P, const double* xpd, TeamCtx& tc) {
    double* xm = P + R.off[BO_XM];
    double* st = P + R.off[BO_ST];
    for (int64_t i = threadIdx.x; i < 3 * (int64_t)R.Np; i += BW) xm[i] = xpd[i];
    __syncthreads();
    if (threadIdx.x == 0) b_stale(B, R, xm, st);
    __syncthreads();
}

// ---------------------------------------------------------------------------------- solves
// MINRES (Paige & Saunders) on M_ss x = r, preconditioned by the inverse of the stresslet self block
// (every diagonal 5 x 5 block of M_ss), to rtol in the preconditioned residual norm.  Work: 8
// vectors of n5 at w.  Returns the iterations.
__device__ __forceinline__ void self_inv5(const double* v, double* out) {
    // inverse of [[9/5, 0, 0, 0, 9/10], [0, 9/5, ...], ..., [9/10, 0, 0, 0, 9/5]]
    const double a = 9. / 5., b = 9. / 10., det = a * a - b * b;
    out[0] = (a * v[0] - b * v[4]) / det;
    out[4] = (a * v[4] - b * v[0]) / det;
    out[1] = v[1] / a;
    out[2] = v[2] / a;
    out[3] = v[3] / a;
}
__device__ int wg_minres_ss(const BConst& B, const BRep& R, double* P, TeamCtx& tc, int64_t n5, const double* rhs,
                            double* x, double* w, double* red) {
    TeamCtx solo{nullptr, tc.gerr, 0, 1, 0, 0};
    TeamCtx& gtc = n5 >= 512 ? tc : solo;   // M_ss v splits over the team when it is large
    double *r1 = w, *r2 = w + n5, *y = w + 2 * n5, *v = w + 3 * n5, *ww = w + 4 * n5, *w1 = w + 5 * n5,
           *w2 = w + 6 * n5;
    const int t = threadIdx.x;
    for (int64_t i = t; i < n5; i += BW) {
        x[i] = 0.0;
        r1[i] = rhs[i];
        r2[i] = rhs[i];
        ww[i] = 0.0;
        w2[i] = 0.0;
    }
    __syncthreads();
    for (int64_t b = t; b < n5 / 5; b += BW) self_inv5(r1 + 5 * b, y + 5 * b);
    __syncthreads();
    const double beta1 = sqrt(fmax(wg_dot(r1, y, n5, red), 0.0));
    if (beta1 == 0.0) return 0;
    double oldb = 0.0, beta = beta1, dbar = 0.0, epsln = 0.0, phibar = beta1, cs = -1.0, sn = 0.0;
    int its = 0;
    for (int itn = 1; itn <= B.ms_maxit; itn++) {
        its = itn;
        const double s = 1.0 / beta;
        for (int64_t i = t; i < n5; i += BW) v[i] = s * y[i];
        __syncthreads();
        team_op(B, R, P, gtc, TOP_MSS, v, y, nullptr, n5);   // y = M_ss v
        if (itn >= 2)
            for (int64_t i = t; i < n5; i += BW) y[i] -= (beta / oldb) * r1[i];
        __syncthreads();
        const double alfa = wg_dot(v, y, n5, red);
        for (int64_t i = t; i < n5; i += BW) {
            y[i] -= (alfa / beta) * r2[i];
            r1[i] = r2[i];
            r2[i] = y[i];
        }
        __syncthreads();
        for (int64_t b = t; b < n5 / 5; b += BW) self_inv5(r2 + 5 * b, y + 5 * b);
        __syncthreads();
        oldb = beta;
        beta = sqrt(fmax(wg_dot(r2, y, n5, red), 0.0));
        const double oldeps = epsln;
        const double delta = cs * dbar + sn * alfa;
        const double gbar = sn * dbar - cs * alfa;
        epsln = sn * beta;
        dbar = -cs * beta;
        double gamma = sqrt(gbar * gbar + beta * beta);
        gamma = fmax(gamma, 2.2e-16);
        cs = gbar / gamma;
        sn = beta / gamma;
        const double phi = cs * phibar;
        phibar = sn * phibar;
        const double denom = 1.0 / gamma;
        for (int64_t i = t; i < n5; i += BW) {
            const double a1 = w2[i];
            w1[i] = a1;
            w2[i] = ww[i];
            ww[i] = (v[i] - oldeps * a1 - delta * ww[i]) * denom;
            x[i] += phi * ww[i];
        }
        __syncthreads();
        if (phibar <= B.ms_rtol * beta1 || beta == 0.0) break;
    }
    return its;
}

// The saddle system [[M, B], [B^T, 0]] [f; u] = b with B = [-I_nu; 0] by exact block elimination:
// f_u = -b_l;  M_ss f_s = b_s - (M [f_u; 0])_s;  u = (M f)_u - b_u.  `uout` gets u too.
__device__ int wg_block_solve(const BConst& B, const BRep& R, double* pool, const double* b, double* x, double* uout,
                              double* red, TeamCtx& tc) {
    const int64_t n = R.n, nu = R.nu, n5 = 5 * (int64_t)R.Nc;
    double* y = pool + R.off[BO_Y];
    double* w5 = pool + R.off[BO_M5];
    for (int64_t i = threadIdx.x; i < n; i += BW) x[i] = (i < nu) ? -b[n + i] : 0.0;
    __syncthreads();
    int its = 0;
    if (n5 > 0) {
        wg_apply<double>(B, R, pool, x, y, true, tc);
        double* rs = w5 + 8 * n5;
        for (int64_t k = threadIdx.x; k < n5; k += BW) rs[k] = b[nu + k] - y[nu + k];
        __syncthreads();
        its = wg_minres_ss(B, R, pool, tc, n5, rs, x + nu, w5, red);
    }
    wg_apply<double>(B, R, pool, x, y, false, tc);
    for (int64_t i = threadIdx.x; i < nu; i += BW) {
        const double u = y[i] - b[i];
        x[n + i] = u;
        if (uout) uout[i] = u;
    }
    __syncthreads();
    return its;
}

// ---------------------------------------------------------------------------------- square root
// out = M^{1/2} xi (n) by Lanczos (schur_sqrt_lanczos, src/slip_vel.cpp:351-477, on the full grand
// mobility): kmax iterations, or fewer on an invariant subspace (beta_j <= 1e-12 beta_0); two-pass
// classical Gram-Schmidt against the stored basis; T_m = U diag(theta) U^T by QL; negative Ritz
// values floored at 0; out = beta_0 Q U diag(sqrt(theta+)) U^T e_1.  Applies in T.  Returns m.
template <class T>
__device__ int wg_sqrt(const BConst& B, const BRep& R, double* pool, const double* xi, double* out, double* red,
                       double* lz, int* nneg_out, TeamCtx& tc) {
    const int64_t n = R.n;
    double* Q = pool + R.off[BO_QB];
    double* w = pool + R.off[BO_LW];
    double* alpha = lz;              // [BATCH_KMAX]
    double* betad = lz + BATCH_KMAX;  // [BATCH_KMAX]
    double* h = lz + 2 * BATCH_KMAX;  // [BATCH_KMAX]
    double* g = lz + 3 * BATCH_KMAX;  // [BATCH_KMAX]
    double* Z = lz + 4 * BATCH_KMAX;  // [BATCH_KMAX^2]
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const double beta0 = sqrt(wg_dot(xi, xi, n, red));
    for (int64_t i = t; i < n; i += BW) out[i] = 0.0;
    if (nneg_out && t == 0) *nneg_out = 0;
    __syncthreads();
    if (beta0 == 0.0) return 0;
    int kmax = B.kmax;
    if (kmax > n) kmax = (int)n;
    if (kmax < 1) kmax = 1;
    for (int64_t i = t; i < n; i += BW) Q[i] = xi[i] * (1.0 / beta0);
    __syncthreads();
    int m = 0;
    for (int j = 0; j < kmax; j++) {
        const double* qj = Q + (int64_t)j * n;
        wg_apply<T>(B, R, pool, qj, w, false, tc);
        if (j > 0) {
            const double bj1 = betad[j - 1];
            const double* qp = Q + (int64_t)(j - 1) * n;
            for (int64_t i = t; i < n; i += BW) w[i] -= bj1 * qp[i];
            __syncthreads();
        }
        const double aj = wg_dot(qj, w, n, red);
        if (t == 0) alpha[j] = aj;
        for (int64_t i = t; i < n; i += BW) w[i] -= aj * qj[i];
        __syncthreads();
        for (int pass = 0; pass < 2; pass++) {
            // h_i = Q_i . w for i <= j, 8 basis vectors at a time over every thread's elements; the
            // per-wave sums meet in Z (free until the tridiagonal solve) and add in wave order
            for (int i0 = 0; i0 <= j; i0 += 8) {
                const int ni = min(8, j + 1 - i0);
                double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                const double* q0 = Q + (int64_t)i0 * n;
                for (int64_t k = t; k < n; k += BW) {
                    const double wk = w[k];
#pragma unroll
                    for (int qq = 0; qq < 8; qq++)
                        if (qq < ni) acc[qq] += q0[(int64_t)qq * n + k] * wk;
                }
#pragma unroll
                for (int qq = 0; qq < 8; qq++) {
                    const double v = wsum(acc[qq]);
                    if (lane == 0) Z[wv * 8 + qq] = v;
                }
                __syncthreads();
                if (t < ni) {
                    double sum = 0.0;
                    for (int ww = 0; ww < NW; ww++) sum += Z[ww * 8 + t];
                    h[i0 + t] = sum;
                }
                __syncthreads();
            }
            for (int64_t k = t; k < n; k += BW) {
                double acc = w[k];
                for (int i = 0; i <= j; i++) acc -= h[i] * Q[(int64_t)i * n + k];
                w[k] = acc;
            }
            __syncthreads();
        }
        m = j + 1;
        const double bj = sqrt(wg_dot(w, w, n, red));
        if (bj <= 1e-12 * beta0 || j == kmax - 1) break;
        if (t == 0) betad[j] = bj;
        double* qn = Q + (int64_t)(j + 1) * n;
        for (int64_t i = t; i < n; i += BW) qn[i] = w[i] * (1.0 / bj);
        __syncthreads();
    }
    __syncthreads();
    if (wv == 0) {
        // g = beta0 U diag(sqrt(max(theta, 0))) U^T e1 (src/slip_vel.cpp:427-459); the QL keeps d and e
        // in the lanes' registers, the spectrum then goes to h for every lane to read
        double dv = lane < m ? alpha[lane] : 0.0;
        double ev = (lane < m - 1) ? betad[lane] : 0.0;
        tql2_wave(m, dv, ev, Z);
        double* dd = h;
        if (lane < m) dd[lane] = dv;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        double gi = 0.0;
        if (lane < m)
            for (int jj = 0; jj < m; jj++) {
                double th = dd[jj];
                if (th < 0.0) th = 0.0;
                gi += Z[lane + jj * m] * sqrt(th) * Z[0 + jj * m];
            }
        int nneg = 0;
        for (int jj = 0; jj < m; jj++)
            if (dd[jj] < 0.0) nneg++;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane < m) g[lane] = beta0 * gi;
        if (lane == 0 && nneg_out) *nneg_out = nneg;
    }
    __syncthreads();
    for (int64_t k = t; k < n; k += BW) {
        double acc = 0.0;
        for (int i = 0; i < m; i++) acc += g[i] * Q[(int64_t)i * n + k];
        out[k] = acc;
    }
    __syncthreads();
    return m;
}

// ---------------------------------------------------------------------------------- the step
struct BStepArgs {
    BConst B;
    const BRep* reps;
    BSchedDev sd;
    double* pool;
    const double* noise;   // device copy of this step's noise: replica r at reps[r].off[BO_NZ] - noise base
    double ktrap;
    int* stats;            // per replica: [minres iterations, lanczos m, negative Ritz values, unconverged]
};

// A team member: run the leader's forks on this rank's share until it exits.
__device__ void team_member(const BStepArgs& a, const BRep& R, TeamCtx& tc, int* cmd) {
    int last = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            const bool ok = team_wait_ge(&tc.ctl->seq, last + 1, &tc.ctl->err);
            if (!ok) __hip_atomic_store(tc.gerr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // the fork's arguments and inputs
            *cmd = ok;
        }
        __syncthreads();
        const int ok = *cmd;
        __syncthreads();
        if (!ok) return;
        last++;
        const BTeamCtl* c = tc.ctl;
        const int op = c->op;
        if (op == TOP_EXIT) return;
        run_slice(a.B, R, a.pool, op, c->in, c->out, c->mat, c->n, tc.rank, tc.g);
        wg_release_point();   // this slice's outputs
        if (threadIdx.x == 0) __hip_atomic_fetch_add(&tc.ctl->done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The step of replica r by this workgroup alone or as its team's leader.
__device__ __forceinline__ void step_replica(const BStepArgs& a, int r, TeamCtx& tc, double* red, double* lz,
                                             int& nneg_sh) {
    const BConst& B = a.B;
    const BRep R = a.reps[r];
    const uint64_t t_start = wall_clock64();
    double* P = a.pool;
    const int64_t n = R.n, nu = R.nu, n3 = 3 * (int64_t)R.Np, nc5 = 5 * (int64_t)R.Nc;
    double *x = P + R.off[BO_X], *xn = P + R.off[BO_XN], *xpd = P + R.off[BO_XPD];
    double *rhs = P + R.off[BO_RHS], *X = P + R.off[BO_XS], *Xdet = P + R.off[BO_XDET], *Xd = P + R.off[BO_XD];
    double *up = P + R.off[BO_UP], *vdn = P + R.off[BO_VDN], *vdc = P + R.off[BO_VDC], *vb = P + R.off[BO_VB];
    double *ud = P + R.off[BO_UD], *dv = P + R.off[BO_DV], *W = P + R.off[BO_W], *sb = P + R.off[BO_SB];
    double* q = P + R.off[BO_Q];
    const double* xpos0 = P + R.off[BO_XP0];
    const double* noise = a.noise + R.off[BO_NZ];   // n slip normals, then nu RFD normals
    const int t = threadIdx.x;
    const double dt = B.dt, kT = B.kT, eps = 1.0e-3;
    int its = 0, unconverged = 0;
    // wall-clock ticks per phase (thread 0): forces, mob(), saddle solves, square root
    uint64_t ph_last = t_start, ph[4] = {0, 0, 0, 0};
    auto mark = [&](int k) {
        if (t == 0) {
            const uint64_t now = wall_clock64();
            ph[k] += now - ph_last;
            ph_last = now;
        }
    };
    auto snapshot = [&]() {
        for (int64_t i = t; i < n3; i += BW) xpd[i] = x[i];
        __syncthreads();
    };
    auto solve = [&](const double* b, double* Xo, double* uo) {
        const int k = wg_block_solve(B, R, P, b, Xo, uo, red, tc);
        its += k;
        unconverged += k >= B.ms_maxit;
        mark(2);
    };
    for (int64_t i = t; i < n3; i += BW) xn[i] = x[i];
    // ---- predictor
    snapshot();
    wg_forces(B, R, xpd, x, xpos0, a.ktrap, rhs);
    __syncthreads();
    mark(0);
    wg_mob(B, R, P, xpd, tc);
    mark(1);
    solve(rhs, Xdet, vdn);
    int lm = 0;
    if (kT > 0) {
        lm = B.mixed ? wg_sqrt<float>(B, R, P, noise, W, red, lz, &nneg_sh, tc)
                     : wg_sqrt<double>(B, R, P, noise, W, red, lz, &nneg_sh, tc);
        mark(3);
        const double sc = sqrt(2 * kT / dt);
        for (int64_t i = t; i < n; i += BW) rhs[i] += sc * W[i];
        __syncthreads();
        solve(rhs, X, up);
        for (int64_t i = t; i < nu; i += BW) vb[i] = up[i] - vdn[i];
        // drift(): RFD probes at x +/- eps W (src/drift.cpp:17-62)
        const double* fext = noise + n;
        for (int64_t i = t; i < n3; i += BW) x[i] += eps * fext[i];
        __syncthreads();
        snapshot();
        wg_mob(B, R, P, xpd, tc);
        mark(1);
        for (int64_t i = t; i < R.ns; i += BW) rhs[i] = (i < n) ? 0.0 : fext[i - n];
        __syncthreads();
        solve(rhs, Xd, dv);
        for (int64_t i = t; i < nc5; i += BW) sb[i] = Xd[nu + i];
        for (int64_t i = t; i < n3; i += BW) x[i] += -2.0 * eps * fext[i];
        __syncthreads();
        snapshot();
        wg_mob(B, R, P, xpd, tc);
        mark(1);
        for (int64_t i = t; i < R.ns; i += BW) rhs[i] = (i < n) ? 0.0 : fext[i - n];
        __syncthreads();
        solve(rhs, Xd, ud);
        for (int64_t i = t; i < nu; i += BW) ud[i] = ud[i] - dv[i];
        for (int64_t i = t; i < nc5; i += BW) sb[i] = (sb[i] - Xd[nu + i]) / (2 * eps);
        __syncthreads();
        // step() (src/traj.cpp:9-35): undo the probe, add the drift and the velocity
        const double drift_scale = kT / (2.0 * eps);
        for (int64_t i = t; i < n3; i += BW) {
            double v = x[i];
            v += eps * fext[i];
            v += drift_scale * ud[i] * dt;
            v += up[i] * dt;
            x[i] = v;
        }
    } else {
        solve(rhs, X, up);
        for (int64_t i = t; i < nu; i += BW) vb[i] = 0.0;
        for (int64_t i = t; i < n3; i += BW) x[i] = x[i] + up[i] * dt;
    }
    __syncthreads();
    // ---- corrector
    snapshot();
    wg_forces(B, R, xpd, x, xpos0, a.ktrap, rhs);
    __syncthreads();
    mark(0);
    wg_mob(B, R, P, xpd, tc);
    mark(1);
    solve(rhs, Xdet, vdc);
    for (int64_t i = t; i < nu; i += BW) up[i] = 0.5 * (vdn[i] + vdc[i]) + vb[i];
    __syncthreads();
    for (int64_t i = t; i < n3; i += BW)
        x[i] = (kT > 0) ? xn[i] + (kT / (2.0 * eps)) * ud[i] * dt + up[i] * dt : xn[i] + up[i] * dt;
    // step_quaternion (src/traj.cpp:64-94)
    for (int c = t; c < R.Nc; c += BW) {
        double* qi = q + 4 * (int64_t)c;
        const double* om = up + R.nm3 + R.nc3 + 3 * (int64_t)c;
        const double w1 = qi[0], x1 = qi[1], y1 = qi[2], z1 = qi[3];
        const double w2 = 0.0, x2 = om[0], y2 = om[1], z2 = om[2];
        const double dq[4] = {w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                              w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2};
        double qn[4];
        for (int k = 0; k < 4; k++) qn[k] = qi[k] + 0.5 * dq[k] * dt;
        const double norm = __builtin_sqrt(qn[0] * qn[0] + qn[1] * qn[1] + qn[2] * qn[2] + qn[3] * qn[3]);
        for (int k = 0; k < 4; k++) qi[k] = qn[k] / norm;
    }
    __syncthreads();
    if (t == 0) {
        a.stats[4 * r] = its;
        a.stats[4 * r + 1] = lm;
        a.stats[4 * r + 2] = kT > 0 ? nneg_sh : 0;
        a.stats[4 * r + 3] = unconverged;
        a.sd.times[r] = (double)(wall_clock64() - t_start);
        for (int k = 0; k < 4; k++) a.sd.phases[5 * r + k] = (double)ph[k];
        a.sd.phases[5 * r + 4] = (double)tc.wait_ticks;
    }
}

// One step of every replica.  Workgroup w first takes its team role (sched[5w] >= 0: leader or
// member of that team's replica), then its solo replicas in order.
__global__ __launch_bounds__(BW) void k_batch_step(BStepArgs a) {
    __shared__ double red[NW];
    __shared__ double lz[4 * BATCH_KMAX + BATCH_KMAX * BATCH_KMAX];
    __shared__ int nneg_sh, cmd_sh;
    const int* wi = a.sd.sched + 5 * blockIdx.x;
    const int team = wi[0];
    for (int k = team >= 0 ? -1 : wi[3]; k < wi[4]; k = (k < 0) ? wi[3] : k + 1) {
        TeamCtx tc{nullptr, a.sd.err, 0, 1, 0, 0};
        int r;
        if (k < 0) {
            tc.ctl = a.sd.ctl + team;
            tc.rank = wi[1];
            tc.g = wi[2];
            r = a.sd.sched[a.sd.team_off + team];
            if (tc.rank > 0) {
                const BRep R = a.reps[r];
                team_member(a, R, tc, &cmd_sh);
                continue;
            }
        } else {
            r = a.sd.sched[a.sd.solo_off + k];
        }
        step_replica(a, r, tc, red, lz, nneg_sh);
        if (tc.g > 1) team_fork(tc, TOP_EXIT, nullptr, nullptr, nullptr, 0);
        __syncthreads();
    }
}

// Single-operation kernels (one replica per workgroup, for the checks against the oracle):
// op 0: y = M f (FP64), op 1: y = M f with float tiles, op 2: saddle solve b -> x,
// op 3: W = M^{1/2} xi (FP64 Lanczos), op 4: mob() at the current positions (snapshot + tiles),
// op 5: rhs of the forces at the current positions.
__global__ __launch_bounds__(BW) void k_batch_op(BConst B, const BRep* reps, double* pool, int op, const double* in,
                                                 double* out, const int64_t* io_off, int* stats, int* err) {
    __shared__ double red[NW];
    __shared__ double lz[4 * BATCH_KMAX + BATCH_KMAX * BATCH_KMAX];
    __shared__ int nneg_sh;
    const int r = blockIdx.x;
    const BRep R = reps[r];
    double* P = pool;
    const double* inr = in + io_off[r];
    double* outr = out + io_off[r];
    double *x = P + R.off[BO_X], *xpd = P + R.off[BO_XPD];
    TeamCtx tc{nullptr, err, 0, 1, 0, 0};
    int s0 = 0;
    if (op == 0) {
        wg_apply<double>(B, R, P, inr, outr, false, tc);
    } else if (op == 1) {
        wg_apply<float>(B, R, P, inr, outr, false, tc);
    } else if (op == 2) {
        s0 = wg_block_solve(B, R, P, inr, outr, nullptr, red, tc);
    } else if (op == 3) {
        s0 = wg_sqrt<double>(B, R, P, inr, outr, red, lz, &nneg_sh, tc);
    } else if (op == 6) {
        s0 = wg_sqrt<float>(B, R, P, inr, outr, red, lz, &nneg_sh, tc);
    } else if (op == 4) {
        for (int64_t i = threadIdx.x; i < 3 * (int64_t)R.Np; i += BW) xpd[i] = x[i];
        __syncthreads();
        wg_mob(B, R, P, xpd, tc);
    } else if (op == 5) {
        for (int64_t i = threadIdx.x; i < 3 * (int64_t)R.Np; i += BW) xpd[i] = x[i];
        __syncthreads();
        wg_forces(B, R, xpd, x, P + R.off[BO_XP0], B.ktrap, outr);
    }
    if (threadIdx.x == 0 && stats) stats[r] = s0;
}

}  // namespace

// ---------------------------------------------------------------------------------- host
ps_err batch_launch_step(hipStream_t st, const BConst& B, const BRep* reps, const BSchedDev& sd, int nwg,
                         bool cooperative, double* pool, const double* noise, double ktrap, int* stats) {
    BStepArgs a{B, reps, sd, pool, noise, ktrap, stats};
    if (cooperative) {
        void* args[] = {&a};
        return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_batch_step), dim3(nwg), dim3(BW), args, 0,
                                          st) == hipSuccess
                   ? PS_OK
                   : PS_ERR_HIP;
    }
    k_batch_step<<<nwg, BW, 0, st>>>(a);
    return hipGetLastError() == hipSuccess ? PS_OK : PS_ERR_HIP;
}
namespace {
// Clears the team fork-join blocks before a step (a kernel rather than a memset node, so that a captured
// step replays it exactly).
__global__ void k_ctl_reset(BTeamCtl* __restrict__ ctl, int n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int64_t W = sizeof(BTeamCtl) / sizeof(int);
    if (i < n * W) reinterpret_cast<int*>(ctl)[i] = 0;
}
ps_err batch_launch_reset(hipStream_t st, BTeamCtl* ctl, int n) {
    const int64_t words = (int64_t)n * (int64_t)(sizeof(BTeamCtl) / sizeof(int));
    k_ctl_reset<<<(unsigned)((words + 255) / 256), 256, 0, st>>>(ctl, n);
    return hipGetLastError() == hipSuccess ? PS_OK : PS_ERR_HIP;
}
// Marks the step kernel's node of a captured graph cooperative (co-resident workgroups, which the team
// waits rely on) where the capture did not already.
hipError_t batch_mark_cooperative(hipGraph_t g) {
    size_t nn = 0;
    hipError_t e = hipGraphGetNodes(g, nullptr, &nn);
    if (e != hipSuccess) return e;
    std::vector<hipGraphNode_t> nodes(nn);
    if ((e = hipGraphGetNodes(g, nodes.data(), &nn)) != hipSuccess) return e;
    int found = 0;
    for (hipGraphNode_t node : nodes) {
        hipGraphNodeType t;
        if ((e = hipGraphNodeGetType(node, &t)) != hipSuccess) return e;
        if (t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams kp{};
        if ((e = hipGraphKernelNodeGetParams(node, &kp)) != hipSuccess) return e;
        if (kp.func != reinterpret_cast<void*>(k_batch_step)) continue;
        found++;
        hipKernelNodeAttrValue v{};
        if ((e = hipGraphKernelNodeGetAttribute(node, hipKernelNodeAttributeCooperative, &v)) != hipSuccess) return e;
        if (!v.cooperative) {
            v.cooperative = 1;
            if ((e = hipGraphKernelNodeSetAttribute(node, hipKernelNodeAttributeCooperative, &v)) != hipSuccess)
                return e;
        }
    }
    return found == 1 ? hipSuccess : hipErrorInvalidValue;
}
}  // namespace
int batch_step_occupancy() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch_step, BW, 0) != hipSuccess) return 0;
    return nb;
}
ps_err batch_launch_op(hipStream_t st, const BConst& B, const BRep* reps, int R, double* pool, int op,
                       const double* in, double* out, const int64_t* io_off, int* stats, int* err) {
    k_batch_op<<<R, BW, 0, st>>>(B, reps, pool, op, in, out, io_off, stats, err);
    return hipGetLastError() == hipSuccess ? PS_OK : PS_ERR_HIP;
}

// ---------------------------------------------------------------------------------- C ABI
#include "normals.h"

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_batch_create_error;

struct Batch {
    int R = 0;
    ps_params p{};
    BConst B{};
    std::vector<BRep> reps;
    BRep* d_reps = nullptr;
    int* d_stats = nullptr;
    // step schedule: teams for the replicas far above the mean work, longest-first packing of the
    // rest; re-packed after every step from each replica's measured time (autotune)
    int nwg_max = 1;             // co-resident step workgroups (CUs x occupancy)
    bool coop = false;           // cooperative launches available (teams need them)
    int team_max = BATCH_TEAM_MAX, min_blocks = 2, autotune = 1;
    double tick_us = 0.01;       // wall-clock tick in microseconds
    std::vector<double> cost;    // per replica: work estimate (microseconds on one workgroup)
    std::vector<int> team_of;    // per replica: team size in the current schedule
    std::vector<int> sched;      // 5 nwg ints, team replicas, solo replicas
    int nwg = 1, nteams = 0, team_off = 0, solo_off = 0;
    bool sched_dirty = true, resched = true;
    int* d_sched = nullptr;
    BTeamCtl* d_ctl = nullptr;
    int* d_err = nullptr;
    double* d_times = nullptr;
    double* d_phases = nullptr;
    int64_t steps_run = 0;
    std::thread prefetch;        // draws the next run's first step of noise into slot 0
    double* pool = nullptr;
    int64_t pool_n = 0;
    double* noise_d = nullptr;
    double* noise_h[2] = {nullptr, nullptr};
    int64_t noise_n = 0;
    hipEvent_t copied[2] = {nullptr, nullptr};
    // the step as a HIP graph per noise slot (noise copy, team reset, the step kernel), captured once
    // and replayed; recaptured when the trap stiffness the kernel takes changes
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    double gktrap[2] = {0.0, 0.0};
    bool graph_failed = false;
    int64_t graph_replays = 0, graph_captures = 0;
    std::vector<std::unique_ptr<NormalStream>> rng;
    hipStream_t stream = nullptr;
    double t = 0.0, ktrap_cur = 0.0;
    std::string err;
    int nthreads = 1;

    void set_error(const std::string& m) { err = m; }
    // draw this step's noise of every replica (n slip normals, then nu RFD normals) into slot k,
    // replicas split over host threads; each replica's stream advances as a single context's does
    void fill(int k) {
        auto work = [this, k](int w) {
            for (int r = w; r < R; r += nthreads)
                rng[r]->fill(noise_h[k] + reps[r].off[BO_NZ], reps[r].n + reps[r].nu);
        };
        if (nthreads <= 1) {
            work(0);
            return;
        }
        std::vector<std::thread> th;
        for (int w = 1; w < nthreads; w++) th.emplace_back(work, w);
        work(0);
        for (auto& x : th) x.join();
    }
};

#define BPS_HIP(b, call)                                                                          \
    do {                                                                                          \
        hipError_t _e = (call);                                                                   \
        if (_e != hipSuccess) {                                                                   \
            (b)->set_error(std::string(#call) + ": " + hipGetErrorString(_e));                    \
            return PS_ERR_HIP;                                                                    \
        }                                                                                         \
    } while (0)

void batch_free(Batch* b) {
    if (!b) return;
    if (b->prefetch.joinable()) b->prefetch.join();
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (void* q : {(void*)b->d_reps, (void*)b->d_stats, (void*)b->pool, (void*)b->noise_d, (void*)b->d_sched,
                    (void*)b->d_ctl, (void*)b->d_err, (void*)b->d_times, (void*)b->d_phases})
        if (q) (void)hipFree(q);
    for (double* h : b->noise_h)
        if (h) (void)hipHostFree(h);
    for (hipEvent_t e : b->copied)
        if (e) (void)hipEventDestroy(e);
    for (hipGraphExec_t g : b->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int64_t pad8(int64_t v) { return (v + 7) / 8 * 8; }

// Team time model: a replica of lone-workgroup time t1 on a team of g takes
// t1 (TEAM_SERIAL + (1 - TEAM_SERIAL) / g) TEAM_OVER (forces, Lanczos orthogonalisation and the
// stresslet MINRES vector work stay on the leader; TEAM_OVER covers the fork-joins).
constexpr double TEAM_SERIAL = 0.08, TEAM_OVER = 1.05;
double team_time(double t1, int g) { return g <= 1 ? t1 : t1 * (TEAM_SERIAL + (1.0 - TEAM_SERIAL) / g) * TEAM_OVER; }

// The step schedule from the lone-workgroup time estimates (cost): the smallest makespan T (by
// bisection) for which every replica with cost > T gets the smallest team (at most team_max, and at
// most one rank per min_blocks work units -- blocks of 64 monomer rows or 16 colloid rows) that
// finishes within T, the teams fit in the co-resident workgroups, and the other replicas, longest
// first, each fit into the workgroup with the most time left before T (team workgroups after their
// team phase).  Without cooperative launches every replica runs alone, packed the same way.
// The planner behind build_schedule (no GPU state): costs and team caps of R replicas, W
// workgroups.  Returns the makespan target T; g = team size per replica, lists = the solo replicas
// of each workgroup (team workgroups first, in decreasing cost of their replica), ord = replicas by
// decreasing cost.
double plan_schedule(const std::vector<double>& cost, const std::vector<int>& cap, int W, std::vector<int>& g,
                     std::vector<std::vector<int>>& lists, std::vector<int>& ord) {
    const int R = (int)cost.size();
    ord.resize(R);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    g.assign(R, 1);
    // feasibility of makespan T: fills g and lists (per workgroup: its solo replicas)
    auto feasible = [&](double T) -> bool {
        int used = 0;
        std::vector<double> left;
        for (int r : ord) {
            g[r] = 1;
            if (cost[r] <= T) continue;
            int gr = 2;
            while (gr <= cap[r] && team_time(cost[r], gr) > T) gr++;
            if (gr > cap[r]) return false;
            g[r] = gr;
            for (int k = 0; k < gr; k++) left.push_back(T - team_time(cost[r], gr));
            used += gr;
            if (used > W) return false;
        }
        const int nteam_wg = used;
        int nsolo = 0;
        for (int r : ord) nsolo += g[r] == 1;
        const int nw = std::max(1, std::min(W, nteam_wg + nsolo));
        while ((int)left.size() < nw) left.push_back(T);
        lists.assign(left.size(), {});
        for (int r : ord) {
            if (g[r] > 1) continue;
            const int w = (int)(std::max_element(left.begin(), left.end()) - left.begin());
            if (left[w] < cost[r]) return false;
            lists[w].push_back(r);
            left[w] -= cost[r];
        }
        return true;
    };
    double total = 0.0, lo = 0.0;
    for (int r = 0; r < R; r++) {
        total += cost[r];
        lo = std::max(lo, team_time(cost[r], cap[r]));
    }
    lo = std::max(lo, total / W);
    double hi = std::max(lo, 1e-30) * 2.0;
    int guard = 0;
    while (!feasible(hi) && guard++ < 60) hi *= 2.0;
    for (int it = 0; it < 40; it++) {
        const double mid = 0.5 * (lo + hi);
        if (feasible(mid))
            hi = mid;
        else
            lo = mid;
    }
    feasible(hi);
    return hi;
}

void build_schedule(Batch* b) {
    const int R = b->R, W = b->nwg_max;
    const bool teams = b->coop && b->team_max > 1;
    std::vector<int> cap(R, 1);
    if (teams)
        for (int r = 0; r < R; r++) {
            const int units = (b->reps[r].Nm + 63) / 64 + (b->reps[r].Nc + 15) / 16;
            cap[r] = std::min(b->team_max, std::max(1, units / std::max(1, b->min_blocks)));
        }
    std::vector<int> g, ord;
    std::vector<std::vector<int>> lists;
    plan_schedule(b->cost, cap, W, g, lists, ord);
    b->team_of = g;
    b->nteams = 0;
    for (int r : ord) b->nteams += g[r] > 1;
    b->nwg = std::min((int)lists.size(), W);
    // every one of the W launched workgroups has an entry (the unused ones an empty solo range) and the
    // team and solo lists sit at fixed offsets, so the launch's arguments do not depend on the packing
    // (a captured step stays valid when the schedule is re-tuned)
    std::vector<int>& S = b->sched;
    S.assign(5 * (size_t)W + 2 * (size_t)R, 0);
    b->team_off = 5 * W;
    b->solo_off = b->team_off + R;
    for (int i = 0; i < W; i++) S[5 * i] = -1;
    int w = 0, k = 0;
    for (int r : ord) {
        if (g[r] <= 1) continue;
        S[b->team_off + k] = r;
        for (int j = 0; j < g[r]; j++, w++) {
            S[5 * w] = k;
            S[5 * w + 1] = j;
            S[5 * w + 2] = g[r];
        }
        k++;
    }
    int pos = 0;
    for (int i = 0; i < W; i++) {
        S[5 * i + 3] = pos;
        if (i < b->nwg)
            for (int r : lists[i]) S[b->solo_off + pos++] = r;
        S[5 * i + 4] = pos;
    }
    S.resize(b->solo_off + pos);
    b->sched_dirty = true;
    b->resched = false;
}
}  // namespace

extern "C" {

ps_err ps_batch_create(const ps_params* p, int R, const int* Nc, const int* Npoly, const int* Nmono,
                       const uint64_t* seeds, void** out) {
    if (!p || !out || R <= 0 || !Nc || !Npoly || !Nmono || !seeds) return PS_ERR_ARG;
    *out = nullptr;
    auto bad = [](const std::string& m) {
        g_batch_create_error = m;
        return PS_ERR_ARG;
    };
    if (!p->mm_HI) return bad("batched replicas: full monomer hydrodynamics (mm_HI) only");
    if (p->box_active) return bad("batched replicas: unbounded systems only (no periodic box)");
    if (p->lub) return bad("batched replicas: no lubrication (a single context's saddle offers it)");
    if (p->sqrt_kmax < 1 || p->sqrt_kmax > BATCH_KMAX) return bad("batched replicas: sqrt_kmax must be in 1..48");
    Batch* b = new Batch();
    b->R = R;
    b->p = *p;
    BConst& B = b->B;
    B.beta = p->beta;
    B.beta2 = p->beta * p->beta;
    B.beta_inv = 1.0 / p->beta;
    B.two_beta = 2.0 * p->beta;
    B.dt = p->dt;
    B.kT = p->kT;
    B.kbond = p->kbond;
    B.r0 = p->r0;
    B.Lmax = p->Lmax;
    B.ktrap = p->ktrap;
    B.k_hs = (p->hs_kmode == 1) ? p->hs_kpref / (p->dt * p->beta) : p->hs_kpref / p->dt;
    B.hs_exp_k = p->hs_exp_k;
    const float sg[3] = {(float)(2. * p->beta), (float)(1. + p->beta), (float)2.0};
    const float rv[3] = {(float)(0.05 + 2. * p->beta), (float)(1.5 + p->beta), (float)2.5};
    for (int k = 0; k < 3; k++) {
        B.sig[k] = sg[k];
        B.rverl[k] = rv[k];
    }
    B.fene = p->fene;
    B.force_sel = p->force_sel;
    B.mono_ev = p->mono_ev;
    B.chain_HI = p->chain_HI;
    B.trap_on = !(std::fabs(p->ktrap) < 1e-12);
    B.kmax = p->sqrt_kmax;
    B.mixed = p->precision == PS_PREC_MIXED;
    B.ms_rtol = 1e-12;
    B.ms_maxit = 1000;
    b->ktrap_cur = p->ktrap;
    b->t = p->restart;
    // per-replica sizes and pool offsets
    int64_t pos = 0, npos = 0;
    b->reps.resize(R);
    for (int r = 0; r < R; r++) {
        if (Nc[r] < 0 || Npoly[r] < 0 || Nmono[r] < 0 || (Npoly[r] > 0 && Nmono[r] < 1)) {
            batch_free(b);
            return bad("batched replicas: negative sizes");
        }
        BRep& Q = b->reps[r];
        Q.Nm = Npoly[r] * Nmono[r];
        Q.Nc = Nc[r];
        Q.Np = Q.Nm + Q.Nc;
        Q.Nmono = std::max(Nmono[r], 1);
        if (Q.Np < 2) {
            batch_free(b);
            return bad("batched replicas: every replica needs at least two beads");
        }
        Q.nm3 = 3 * (int64_t)Q.Nm;
        Q.nc3 = 3 * (int64_t)Q.Nc;
        Q.n = 3 * (int64_t)Q.Nm + 11 * (int64_t)Q.Nc;
        Q.nu = 3 * (int64_t)Q.Nm + 6 * (int64_t)Q.Nc;
        Q.ns = Q.n + Q.nu;
        // mob()'s last monomer-monomer pair (src/mob.cpp:845-897)
        Q.si = Q.sj = -1;
        for (int i = Q.Nm - 2; i >= 0 && Q.si < 0; i--)
            for (int j = Q.Nm - 1; j > i; j--) {
                if (p->chain_HI && i / Q.Nmono == j / Q.Nmono) continue;
                Q.si = i;
                Q.sj = j;
                break;
            }
        const int64_t n3 = 3 * (int64_t)Q.Np, n5 = 5 * (int64_t)Q.Nc;
        const int64_t sizes[BO_NOFF] = {n3, n3, n3, n3, 4 * (int64_t)Q.Nc, 3 * (int64_t)Q.Nc, Q.ns, Q.ns, Q.ns, Q.ns,
                                        Q.nu, Q.nu, Q.nu, Q.nu, Q.nu, Q.nu, Q.n, Q.n, Q.n, (int64_t)p->sqrt_kmax * Q.n,
                                        n5, 9 * n5, 3 * (int64_t)((Q.Nm + 63) / 64) * Q.Nm, 18, 0};
        for (int k = 0; k < BO_NOFF; k++) {
            if (k == BO_NZ) continue;
            Q.off[k] = pos;
            pos += pad8(std::max<int64_t>(sizes[k], 1));
        }
        Q.off[BO_NZ] = npos;
        npos += pad8(Q.n + Q.nu);
    }
    b->pool_n = pos;
    b->noise_n = npos;
    // first cost estimate: the monomer tiles grow as Np^2 and the colloid pair couplings as Nc^2,
    // over a floor of vector work that every step pays (the Lanczos recurrence, the MINRES
    // iterations; fitted to lone-workgroup step times at Np = 156 ... 2980); measured step times
    // replace it after the first step
    b->cost.resize(R);
    for (int r = 0; r < R; r++) {
        const double np = b->reps[r].Np, nc = b->reps[r].Nc;
        b->cost[r] = 1e-3 * (np * np + 30.0 * nc * nc) + 0.5 * np + 400.0;
    }
    auto fail_hip = [&](hipError_t e) {
        g_batch_create_error = std::string("batched replicas: ") + hipGetErrorString(e);
        batch_free(b);
        return PS_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return fail_hip(e);
    if ((e = hipMalloc(&b->pool, sizeof(double) * (size_t)b->pool_n)) != hipSuccess) return fail_hip(e);
    if ((e = hipMemset(b->pool, 0, sizeof(double) * (size_t)b->pool_n)) != hipSuccess) return fail_hip(e);
    if ((e = hipMalloc(&b->noise_d, sizeof(double) * (size_t)std::max<int64_t>(b->noise_n, 1))) != hipSuccess)
        return fail_hip(e);
    for (int k = 0; k < 2; k++) {
        if ((e = hipHostMalloc(&b->noise_h[k], sizeof(double) * (size_t)std::max<int64_t>(b->noise_n, 1),
                               hipHostMallocDefault)) != hipSuccess)
            return fail_hip(e);
        if ((e = hipEventCreateWithFlags(&b->copied[k], hipEventDisableTiming)) != hipSuccess) return fail_hip(e);
    }
    if ((e = hipMalloc(&b->d_reps, sizeof(BRep) * R)) != hipSuccess) return fail_hip(e);
    {
        int dev = 0, ncu = 1, coop = 0, khz = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return fail_hip(e);
        if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
            return fail_hip(e);
        if ((e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev)) != hipSuccess) coop = 0;
        if ((e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev)) != hipSuccess || khz <= 0)
            khz = 100000;
        const int occ = batch_step_occupancy();
        b->nwg_max = std::max(1, ncu * std::max(1, occ));
        b->coop = coop != 0 && occ >= 1;
        b->tick_us = 1000.0 / khz;
    }
    if ((e = hipMalloc(&b->d_sched, sizeof(int) * (5 * (size_t)b->nwg_max + 2 * (size_t)R))) != hipSuccess)
        return fail_hip(e);
    if ((e = hipMalloc(&b->d_ctl, sizeof(BTeamCtl) * (size_t)b->nwg_max)) != hipSuccess) return fail_hip(e);
    if ((e = hipMalloc(&b->d_err, sizeof(int))) != hipSuccess) return fail_hip(e);
    if ((e = hipMemset(b->d_err, 0, sizeof(int))) != hipSuccess) return fail_hip(e);
    if ((e = hipMalloc(&b->d_times, sizeof(double) * R)) != hipSuccess) return fail_hip(e);
    if ((e = hipMemset(b->d_times, 0, sizeof(double) * R)) != hipSuccess) return fail_hip(e);
    if ((e = hipMalloc(&b->d_phases, sizeof(double) * 5 * (size_t)R)) != hipSuccess) return fail_hip(e);
    if ((e = hipMemset(b->d_phases, 0, sizeof(double) * 5 * (size_t)R)) != hipSuccess) return fail_hip(e);
    if ((e = hipMalloc(&b->d_stats, sizeof(int) * 4 * R)) != hipSuccess) return fail_hip(e);
    if ((e = hipMemcpy(b->d_reps, b->reps.data(), sizeof(BRep) * R, hipMemcpyHostToDevice)) != hipSuccess)
        return fail_hip(e);
    if ((e = hipMemset(b->d_stats, 0, sizeof(int) * 4 * R)) != hipSuccess) return fail_hip(e);
    // initial state: identity quaternions, the carried tile after fill_self (the identity)
    for (int r = 0; r < R; r++) {
        const BRep& Q = b->reps[r];
        std::vector<double> q(4 * (size_t)Q.Nc, 0.0);
        for (int i = 0; i < Q.Nc; i++) q[4 * i] = 1.0;
        const double st0[18] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0, 0, 1};
        if (Q.Nc > 0 &&
            (e = hipMemcpy(b->pool + Q.off[BO_Q], q.data(), sizeof(double) * q.size(), hipMemcpyHostToDevice)) !=
                hipSuccess)
            return fail_hip(e);
        if ((e = hipMemcpy(b->pool + Q.off[BO_ST], st0, sizeof(st0), hipMemcpyHostToDevice)) != hipSuccess)
            return fail_hip(e);
    }
    // each replica's noise stream, seeded as a single context's (seed + llround(restart / dt))
    b->rng.resize(R);
    for (int r = 0; r < R; r++) {
        b->rng[r].reset(new NormalStream());
        b->rng[r]->set_max_parts(1);
        b->rng[r]->seed(seeds[r] + (uint64_t)std::llround(p->restart / p->dt));
    }
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    b->nthreads = (int)std::min<unsigned>(16, std::max(1u, hw / 2));
    *out = b;
    return PS_OK;
}

const char* ps_batch_last_error(const void* h) {
    return h ? static_cast<const Batch*>(h)->err.c_str() : g_batch_create_error.c_str();
}

ps_err ps_batch_destroy(void* h) {
    batch_free(static_cast<Batch*>(h));
    return PS_OK;
}

ps_err ps_batch_sizes(void* h, int r, int64_t* out6) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || !out6 || r < 0 || r >= b->R) return PS_ERR_ARG;
    const BRep& Q = b->reps[r];
    const int64_t v[6] = {Q.Np, Q.Nc, Q.Nm, Q.n, Q.nu, Q.ns};
    std::copy(v, v + 6, out6);
    return PS_OK;
}

ps_err ps_batch_set_state(void* h, int r, const double* x0, const double* q0) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || !x0 || r < 0 || r >= b->R) return PS_ERR_ARG;
    const BRep& Q = b->reps[r];
    BPS_HIP(b, hipStreamSynchronize(b->stream));
    BPS_HIP(b, hipMemcpy(b->pool + Q.off[BO_X], x0, sizeof(double) * 3 * (size_t)Q.Np, hipMemcpyHostToDevice));
    if (Q.Nc > 0) {
        // trap centres = initial colloid positions on a fresh start (src/Stokes.cpp:279-284)
        std::vector<double> c0(3 * (size_t)Q.Nc, 0.0);
        if (b->p.restart == 0.0) std::copy(x0 + 3 * (size_t)Q.Nm, x0 + 3 * (size_t)Q.Np, c0.begin());
        BPS_HIP(b, hipMemcpy(b->pool + Q.off[BO_XP0], c0.data(), sizeof(double) * c0.size(), hipMemcpyHostToDevice));
        std::vector<double> q(4 * (size_t)Q.Nc, 0.0);
        for (int i = 0; i < Q.Nc; i++)
            for (int k = 0; k < 4; k++) q[4 * i + k] = q0 ? q0[4 * i + k] : (k == 0 ? 1.0 : 0.0);
        BPS_HIP(b, hipMemcpy(b->pool + Q.off[BO_Q], q.data(), sizeof(double) * q.size(), hipMemcpyHostToDevice));
    }
    return PS_OK;
}

ps_err ps_batch_get_state(void* h, int r, double* x, double* q, double* t) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || r < 0 || r >= b->R) return PS_ERR_ARG;
    const BRep& Q = b->reps[r];
    BPS_HIP(b, hipStreamSynchronize(b->stream));
    if (x) BPS_HIP(b, hipMemcpy(x, b->pool + Q.off[BO_X], sizeof(double) * 3 * (size_t)Q.Np, hipMemcpyDeviceToHost));
    if (q && Q.Nc > 0)
        BPS_HIP(b, hipMemcpy(q, b->pool + Q.off[BO_Q], sizeof(double) * 4 * (size_t)Q.Nc, hipMemcpyDeviceToHost));
    if (t) *t = b->t;
    return PS_OK;
}

// The last step's velocity vectors of replica r (nu each; any may be null).
ps_err ps_batch_get_vectors(void* h, int r, double* up, double* v_det_n, double* v_det_c, double* v_brown,
                            double* udiff) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || r < 0 || r >= b->R) return PS_ERR_ARG;
    const BRep& Q = b->reps[r];
    BPS_HIP(b, hipStreamSynchronize(b->stream));
    const std::pair<double*, int> outs[5] = {{up, BO_UP}, {v_det_n, BO_VDN}, {v_det_c, BO_VDC}, {v_brown, BO_VB},
                                             {udiff, BO_UD}};
    for (const auto& o : outs)
        if (o.first)
            BPS_HIP(b, hipMemcpy(o.first, b->pool + Q.off[o.second], sizeof(double) * (size_t)Q.nu,
                                 hipMemcpyDeviceToHost));
    return PS_OK;
}

namespace {
// The team error flag of the steps so far (a team wait that timed out).
ps_err batch_check_err(Batch* b) {
    int err = 0;
    BPS_HIP(b, hipMemcpy(&err, b->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        b->set_error("batched step: a team wait timed out (results of this run are invalid)");
        return PS_ERR_HIP;
    }
    return PS_OK;
}
// Before a step: with `measured` (a step has run) and autotune on, wait for it, turn its
// per-replica times into cost estimates and re-pack; upload the schedule when it changed.
ps_err batch_schedule_step(Batch* b, bool measured) {
    if (measured && b->autotune) {
        BPS_HIP(b, hipStreamSynchronize(b->stream));
        std::vector<double> ticks(b->R);
        BPS_HIP(b, hipMemcpy(ticks.data(), b->d_times, sizeof(double) * b->R, hipMemcpyDeviceToHost));
        for (int r = 0; r < b->R; r++) {
            const double us = ticks[r] * b->tick_us;
            const int g = b->team_of.empty() ? 1 : b->team_of[r];
            if (us > 0) b->cost[r] = us / team_time(1.0, g);   // the lone-workgroup time it implies
        }
        build_schedule(b);
    } else if (b->resched) {
        build_schedule(b);
    }
    if (b->sched_dirty) {
        BPS_HIP(b, hipStreamSynchronize(b->stream));
        BPS_HIP(b, hipMemcpy(b->d_sched, b->sched.data(), sizeof(int) * b->sched.size(), hipMemcpyHostToDevice));
        b->sched_dirty = false;
    }
    return PS_OK;
}
}  // namespace

// Captures step graph k (the noise copy from pinned slot k, the team reset and the step kernel, with
// the launch arguments of this step: none of them depends on the packing, which lives in d_sched).
// A capture that fails leaves graphs off for this batch; its steps then launch eagerly.
static void batch_capture_step(Batch* b, int k, const BSchedDev& sd) {
    if (b->gexec[k]) {
        (void)hipGraphExecDestroy(b->gexec[k]);
        b->gexec[k] = nullptr;
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        bool ok = true;
        if (b->p.kT > 0)
            ok = hipMemcpyAsync(b->noise_d, b->noise_h[k], sizeof(double) * (size_t)b->noise_n, hipMemcpyHostToDevice,
                                b->stream) == hipSuccess;
        if (ok) ok = batch_launch_reset(b->stream, b->d_ctl, b->nwg_max) == PS_OK;
        if (ok)
            ok = batch_launch_step(b->stream, b->B, b->d_reps, sd, b->nwg_max, b->coop, b->pool, b->noise_d,
                                   b->ktrap_cur, b->d_stats) == PS_OK;
        e = hipStreamEndCapture(b->stream, &g);
        if (e == hipSuccess && !ok) e = hipErrorUnknown;
    }
    if (e == hipSuccess && b->coop) e = batch_mark_cooperative(g);
    if (e == hipSuccess) e = hipGraphInstantiate(&b->gexec[k], g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b->gexec[k] = nullptr;
        b->graph_failed = true;
        return;
    }
    b->gktrap[k] = b->ktrap_cur;
    b->graph_captures++;
}

// nsteps of every replica: one persistent kernel per step; a host thread draws step k + 1's noise
// while the GPU runs step k (two pinned slots; a slot is refilled only after its copy completed).
ps_err ps_batch_run(void* h, int nsteps) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || nsteps < 0) return PS_ERR_ARG;
    if (nsteps == 0) return PS_OK;
    const bool noisy = b->p.kT > 0;
    std::mutex mu;
    std::condition_variable cv;
    int filled = 0, consumed = 0;
    bool abort = false;
    std::thread producer;
    int first = 0;   // steps whose noise is already drawn (the previous run's prefetch)
    if (b->prefetch.joinable()) {
        b->prefetch.join();
        first = 1;
        filled = 1;
    }
    if (noisy) {
        producer = std::thread([&, first] {
            for (int s = first; s < nsteps; s++) {
                const int k = s & 1;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return abort || consumed >= s - 1; });
                    if (abort) return;
                }
                if (s >= 2) (void)hipEventSynchronize(b->copied[k]);
                b->fill(k);
                {
                    std::lock_guard<std::mutex> lk(mu);
                    filled = s + 1;
                }
                cv.notify_all();
            }
        });
    }
    ps_err rc = PS_OK;
    const bool use_graph = b->p.graph != 0;
    for (int s = 0; s < nsteps && rc == PS_OK; s++) {
        const int k = noisy ? (s & 1) : 0;
        b->t += b->p.dt;
        // trapping_forces: weaken_trap replaces ktrap permanently after tstart + trun
        if (b->B.trap_on && b->t > b->p.tstart + b->p.trun && b->p.weaken_trap >= 0) b->ktrap_cur = b->p.weaken_trap;
        if (noisy) {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return filled > s; });
        }
        rc = batch_schedule_step(b, s > 0 || b->steps_run > 0);
        const BSchedDev sd{b->d_sched, b->team_off, b->solo_off, b->d_ctl, b->d_err, b->d_times, b->d_phases};
        if (rc == PS_OK && use_graph && !b->graph_failed && (!b->gexec[k] || b->gktrap[k] != b->ktrap_cur))
            batch_capture_step(b, k, sd);
        if (rc == PS_OK && use_graph && !b->graph_failed && b->gexec[k]) {
            const hipError_t e = hipGraphLaunch(b->gexec[k], b->stream);
            if (e != hipSuccess) {
                b->set_error(std::string("batched step graph: ") + hipGetErrorString(e));
                rc = PS_ERR_HIP;
            } else {
                b->graph_replays++;
            }
        } else if (rc == PS_OK) {
            if (noisy) {
                const hipError_t e = hipMemcpyAsync(b->noise_d, b->noise_h[k], sizeof(double) * (size_t)b->noise_n,
                                                    hipMemcpyHostToDevice, b->stream);
                if (e != hipSuccess) {
                    b->set_error(std::string("noise copy: ") + hipGetErrorString(e));
                    rc = PS_ERR_HIP;
                }
            }
            if (rc == PS_OK && b->nteams > 0) {
                rc = batch_launch_reset(b->stream, b->d_ctl, b->nwg_max);
                if (rc != PS_OK) b->set_error("team reset: kernel launch failed");
            }
            if (rc == PS_OK) {
                rc = batch_launch_step(b->stream, b->B, b->d_reps, sd, b->nwg_max, b->nteams > 0, b->pool,
                                       b->noise_d, b->ktrap_cur, b->d_stats);
                if (rc != PS_OK) b->set_error("batched step: kernel launch failed");
            }
        }
        if (noisy) {
            // the slot's copy is part of this step's work; the producer refills it only after it
            if (rc == PS_OK) {
                const hipError_t e = hipEventRecord(b->copied[k], b->stream);
                if (e != hipSuccess) {
                    b->set_error(std::string("noise event: ") + hipGetErrorString(e));
                    rc = PS_ERR_HIP;
                }
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                consumed = s + 1;
            }
            cv.notify_all();
        }
        b->steps_run++;
    }
    if (noisy) {
        {
            std::lock_guard<std::mutex> lk(mu);
            abort = true;
        }
        cv.notify_all();
        producer.join();
    }
    if (rc != PS_OK) return rc;
    BPS_HIP(b, hipStreamSynchronize(b->stream));
    // draw the next run's first step of noise while the caller works (every copy has completed)
    if (noisy) b->prefetch = std::thread([b] { b->fill(0); });
    return batch_check_err(b);
}

// Per replica: [stresslet MINRES iterations of the last step, Lanczos iterations, negative Ritz
// values, solves whose stresslet MINRES stopped at its cap] (4 R ints).
ps_err ps_batch_stats(void* h, int* out) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || !out) return PS_ERR_ARG;
    BPS_HIP(b, hipStreamSynchronize(b->stream));
    BPS_HIP(b, hipMemcpy(out, b->d_stats, sizeof(int) * 4 * (size_t)b->R, hipMemcpyDeviceToHost));
    return PS_OK;
}

// The last step's wall-clock time of each replica in microseconds (a team's replica: its leader's
// view) and the team size it ran with (1: a workgroup alone); either output may be null.
ps_err ps_batch_times(void* h, double* us, int* team) {
    Batch* b = static_cast<Batch*>(h);
    if (!b) return PS_ERR_ARG;
    BPS_HIP(b, hipStreamSynchronize(b->stream));
    if (us) {
        std::vector<double> ticks(b->R);
        BPS_HIP(b, hipMemcpy(ticks.data(), b->d_times, sizeof(double) * b->R, hipMemcpyDeviceToHost));
        for (int r = 0; r < b->R; r++) us[r] = ticks[r] * b->tick_us;
    }
    if (team)
        for (int r = 0; r < b->R; r++) team[r] = b->team_of.empty() ? 1 : b->team_of[r];
    return PS_OK;
}

// Workgroups of the last step's launch and the most that can be co-resident.
ps_err ps_batch_workgroups(void* h, int* nwg_last, int* nwg_max) {
    Batch* b = static_cast<Batch*>(h);
    if (!b) return PS_ERR_ARG;
    if (nwg_last) *nwg_last = b->nwg;
    if (nwg_max) *nwg_max = b->nwg_max;
    return PS_OK;
}

// The last step's time of each replica per phase, in microseconds (R x 5: forces, mob(), saddle
// solves, square root -- vector work between phases counts toward the next phase -- and, within
// those, the team leader's time waiting in joins).
ps_err ps_batch_phase_times(void* h, double* us5) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || !us5) return PS_ERR_ARG;
    BPS_HIP(b, hipStreamSynchronize(b->stream));
    BPS_HIP(b, hipMemcpy(us5, b->d_phases, sizeof(double) * 5 * (size_t)b->R, hipMemcpyDeviceToHost));
    for (int k = 0; k < 5 * b->R; k++) us5[k] *= b->tick_us;
    return PS_OK;
}

// Schedule knobs: team_max (1..16; 1 turns teams off), min_blocks (>= 1: blocks of 64 monomer rows
// per team rank at least), autotune (0: keep the size-based cost model).  Takes effect at the next
// step.
ps_err ps_batch_config(void* h, int team_max, int min_blocks, int autotune) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || team_max < 1 || team_max > BATCH_TEAM_MAX || min_blocks < 1) return PS_ERR_ARG;
    b->team_max = team_max;
    b->min_blocks = min_blocks;
    b->autotune = autotune != 0;
    b->resched = true;   // re-pack before the next step
    return PS_OK;
}

ps_err ps_batch_graph_info(void* h, int64_t* replays, int64_t* captures, int* failed) {
    Batch* b = static_cast<Batch*>(h);
    if (!b) return PS_ERR_ARG;
    if (replays) *replays = b->graph_replays;
    if (captures) *captures = b->graph_captures;
    if (failed) *failed = b->graph_failed ? 1 : 0;
    return PS_OK;
}

// The step schedule's planner without a GPU: lone-workgroup costs and team caps of R replicas on W
// co-resident workgroups.  Out (each may be null): team size per replica (R), the workgroup of
// each replica's first (or only) rank (R), each workgroup's planned load (W; unused ones 0) and
// the makespan target T.
ps_err ps_batch_plan(int R, const double* cost, const int* cap, int W, int* team, int* wg, double* load, double* T) {
    if (R <= 0 || W <= 0 || !cost || !cap) return PS_ERR_ARG;
    for (int r = 0; r < R; r++)
        if (!(cost[r] >= 0.0) || cap[r] < 1) return PS_ERR_ARG;
    std::vector<double> c(cost, cost + R);
    std::vector<int> cp(cap, cap + R), g, ord;
    std::vector<std::vector<int>> lists;
    const double Tm = plan_schedule(c, cp, W, g, lists, ord);
    std::vector<int> first(R, -1);
    std::vector<double> ld(W, 0.0);
    int w = 0;
    for (int r : ord)
        if (g[r] > 1) {
            first[r] = w;
            for (int j = 0; j < g[r] && w < W; j++, w++) ld[w] = team_time(c[r], g[r]);
        }
    for (int i = 0; i < (int)lists.size() && i < W; i++)
        for (int r : lists[i]) {
            first[r] = i;
            ld[i] += c[r];
        }
    if (team)
        for (int r = 0; r < R; r++) team[r] = g[r];
    if (wg)
        for (int r = 0; r < R; r++) wg[r] = first[r];
    if (load)
        for (int i = 0; i < W; i++) load[i] = ld[i];
    if (T) *T = Tm;
    return PS_OK;
}

// One operation on every replica at its current positions (for checks), inputs and outputs
// concatenated over the replicas: op 0 / 1 y = M f (FP64 / float monomer tiles; n each, after
// op 4), op 2 the saddle solve (b, x: ns each, after op 4), op 3 / 6 W = M^{1/2} xi (n each, FP64 /
// float monomer tiles, after op 4), op 4 mob() at the current positions (no input or output), op 5
// the force rhs (ns).  iters (R ints, may be null): op 2 stresslet MINRES iterations, op 3 / 6
// Lanczos iterations.
ps_err ps_batch_op(void* h, int op, const double* in, double* out, int* iters) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || op < 0 || op > 6) return PS_ERR_ARG;
    std::vector<int64_t> io(b->R, 0);
    int64_t tot = 0;
    for (int r = 0; r < b->R; r++) {
        io[r] = tot;
        const BRep& Q = b->reps[r];
        tot += (op == 2 || op == 5) ? Q.ns : (op == 4 ? 0 : Q.n);
    }
    if (op != 4 && (!out || (op != 5 && !in))) return PS_ERR_ARG;
    double *din = nullptr, *dout = nullptr;
    int64_t* dio = nullptr;
    int* dit = nullptr;
    auto cleanup = [&]() {
        for (void* q : {(void*)din, (void*)dout, (void*)dio, (void*)dit})
            if (q) (void)hipFree(q);
    };
    ps_err rc = PS_OK;
    do {
        const size_t bytes = sizeof(double) * (size_t)std::max<int64_t>(tot, 1);
        if (hipMalloc(&din, bytes) != hipSuccess || hipMalloc(&dout, bytes) != hipSuccess ||
            hipMalloc(&dio, sizeof(int64_t) * b->R) != hipSuccess || hipMalloc(&dit, sizeof(int) * b->R) != hipSuccess) {
            b->set_error("batch op: allocation failed");
            rc = PS_ERR_HIP;
            break;
        }
        if ((in && tot > 0 && hipMemcpy(din, in, bytes, hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(dio, io.data(), sizeof(int64_t) * b->R, hipMemcpyHostToDevice) != hipSuccess) {
            b->set_error("batch op: copy failed");
            rc = PS_ERR_HIP;
            break;
        }
        BConst B = b->B;
        B.ktrap = b->ktrap_cur;
        rc = batch_launch_op(b->stream, B, b->d_reps, b->R, b->pool, op, din, dout, dio, dit, b->d_err);
        if (rc != PS_OK) {
            b->set_error("batch op: kernel launch failed");
            break;
        }
        if (hipStreamSynchronize(b->stream) != hipSuccess ||
            (out && tot > 0 && hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess) ||
            (iters && hipMemcpy(iters, dit, sizeof(int) * b->R, hipMemcpyDeviceToHost) != hipSuccess)) {
            b->set_error("batch op: execution failed");
            rc = PS_ERR_HIP;
        }
    } while (false);
    cleanup();
    return rc;
}

}  // extern "C"
