// arrow.hip -- the mm_HI == false regime: the O(N) "arrowhead" grand mobility
//     M = [[ beta_inv I , M^mc ],
//          [ M^cm       , M^cc ]]          (src/matfree_A.cpp:1-16, :38-80)
// for monomers without monomer-monomer hydrodynamics and at most one colloid, its eigenvalue-floored
// colloid block (src/slip_vel.cpp:81-203), its block-Cholesky noise (src/slip_vel.cpp:205-349), the
// Verlet list of monomer-colloid pairs only (src/init.cpp:263-290) and the monomer cell list of the
// WCA forces (src/cell_list.cpp:23-172).
//
// HBM layout: M^cm is stored as its 8 non-zero rows (colloid rotation 3, stresslet 5) over the 3 Nm
// monomer DOFs, row-major [8][3Nm], so that a thread per monomer DOF reads every row coalesced.  The
// three translation rows are zero by construction: mobility_AB never writes its translation tile
// (src/mob.cpp:407-453 are commented out), so the reference fills them from a value-initialised
// buffer.  M^cc, its eigenvectors and floored spectrum live in `ar_mcc` [121 + 121 + 11].
//
// Every operation is a DOF-parallel pass plus one small finishing block: the apply and the noise
// are two launches each, the Schur floor a Gram reduction plus one workgroup that runs a parallel
// Jacobi eigensolver on the 11 x 11 Schur complement -- no host round trip.
#include "ps_common.h"
#include "tiles.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>

namespace {

constexpr int AB = 256;             // threads per DOF-parallel block
constexpr int NR = 8;               // stored rows of M^cm (rotation 3 + stresslet 5)
constexpr int NSYM = NR * (NR + 1) / 2;

unsigned dof_blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + AB - 1) / AB); }
// The passes that end in a grid-wide reduction (apply, block solve, noise) give each thread ARD DOFs:
// fewer blocks store partials and take the finishing counter, which shortens the last-block chain.
constexpr int ARD = 4;
unsigned red_blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + AB * ARD - 1) / (AB * ARD)); }

// ------------------------------------------------------------------------------------ Verlet list
// Monomer k pairs with the colloid when their separation is below the AB Verlet cutoff (a float,
// compared after promotion, src/check_dist.cpp:49-62).  Rows (CSR, both directions): a monomer's
// single partner is the colloid; the colloid's row lists its monomers ascending; nb_off[k] = pos[k]
// (the exclusive scan of the flags) up to the colloid, 2 * total after it, and nb_off[Np + 1] = the
// longest row.  Two launches, no host round trip:
//   pass 1: flag, the flag's exclusive scan within the block (ballot + popcount, then the 4 wave
//           totals), the block's count; the last block to finish scans the block counts and writes
//           the colloid rows' offsets.  The same pass snapshots the positions.
//   pass 2: every monomer's offset = its block's offset + its scan within the block, and the rows.
__device__ __forceinline__ bool last_block(int* __restrict__ counter);
__global__ __launch_bounds__(AB) void k_ar_vl1(SysConst s, const double* __restrict__ x, double* __restrict__ xpd,
                                               double rverl_ab, int* __restrict__ flag, int* __restrict__ lpos,
                                               int* __restrict__ bscan, int* __restrict__ nb_off,
                                               int* __restrict__ counter) {
    __shared__ int wsum[AB / 64];
    __shared__ int csum[AB / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int k = blockIdx.x * AB + t;
    if (k < s.Np) {
        xpd[3 * k] = x[3 * k];
        xpd[3 * k + 1] = x[3 * k + 1];
        xpd[3 * k + 2] = x[3 * k + 2];
    }
    int f = 0;
    if (k < s.Nm && s.Nc == 1) {
        double e[3], dr, dr_inv;
        pair_geom_dev(s, x, k, s.Nm, e, dr, dr_inv);
        f = dr < rverl_ab;
    }
    const unsigned long long m = __ballot(f);
    const int below = __popcll(m & ((1ull << lane) - 1ull));   // flags of lower lanes
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int base = 0, cnt = 0;
    for (int q = 0; q < AB / 64; q++) {
        base += q < w ? wsum[q] : 0;
        cnt += wsum[q];
    }
    if (k < s.Nm) {
        flag[k] = f;
        lpos[k] = base + below;
    }
    if (t == 0) bscan[blockIdx.x] = cnt;
    if (!last_block(counter)) return;
    // exclusive scan of the nb block counts: each thread sums a contiguous chunk, the chunk sums are
    // scanned by shuffles within a wave and across the waves
    const int nb = gridDim.x, per = (nb + AB - 1) / AB;
    const int b0 = min(t * per, nb), b1 = min(b0 + per, nb);
    int cs = 0;
    for (int b = b0; b < b1; b++) cs += bscan[b];
    int inc = cs;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) csum[w] = inc;
    __syncthreads();
    int cbase = 0, total = 0;
    for (int q = 0; q < AB / 64; q++) {
        cbase += q < w ? csum[q] : 0;
        total += csum[q];
    }
    int run = cbase + inc - cs;
    for (int b = b0; b < b1; b++) {
        const int v = bscan[b];
        bscan[nb + b] = run;
        run += v;
    }
    if (t == 0) {
        nb_off[s.Nm] = total;
        for (int q = s.Nm + 1; q <= s.Np; q++) nb_off[q] = 2 * total;
        nb_off[s.Np + 1] = s.Nc == 1 ? (total > 0 ? total : 1) : 0;
    }
}
__global__ __launch_bounds__(AB) void k_ar_vl2(int Nm, int nb1, const int* __restrict__ flag,
                                               const int* __restrict__ lpos, const int* __restrict__ bscan,
                                               int* __restrict__ nb_off, int* __restrict__ part) {
    const int k = blockIdx.x * AB + threadIdx.x;
    if (k >= Nm) return;
    const int pos = bscan[nb1 + blockIdx.x] + lpos[k];   // pass 1 had nb1 blocks of the same size
    nb_off[k] = pos;
    if (flag[k]) {
        part[pos] = Nm;                 // the monomer's row: the colloid
        part[nb_off[Nm] + pos] = k;     // the colloid's row, ascending
    }
}

// ------------------------------------------------------------------------------------ M^cm
// Column block of monomer k: rotation rows take mobility_AB's b, stresslet rows its flattened gt
// (placement of src/mob.cpp:1013-1020).  The same pass snapshots the positions of this mob().  With
// `gram`, the block also stores its partials of the Gram block M8 M8^T (upper triangle, NSYM per
// block) for the Schur floor that follows.
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double* __restrict__ part);
__global__ __launch_bounds__(AB) void k_ar_mcm(SysConst s, const double* __restrict__ xpd, double* __restrict__ xm,
                                               double beta2, double* __restrict__ mcm, double* __restrict__ gram) {
    const int k = blockIdx.x * AB + threadIdx.x;
    double acc[NSYM];
#pragma unroll
    for (int q = 0; q < NSYM; q++) acc[q] = 0.0;
    if (k < s.Np) {
        xm[3 * k] = xpd[3 * k];
        xm[3 * k + 1] = xpd[3 * k + 1];
        xm[3 * k + 2] = xpd[3 * k + 2];
    }
    if (k < s.Nm && s.Nc == 1) {
        const int64_t nm3 = s.nm3;
        double e[3], dr, dr_inv, b[3][3], bt[3][3], gt[3][5];
        pair_geom_dev(s, xpd, k, s.Nm, e, dr, dr_inv);
        tile_ab_dev(dr_inv, e, beta2, b, bt, gt);
        for (int ii = 0; ii < 3; ii++) {
            for (int jj = 0; jj < 3; jj++) mcm[ii * nm3 + 3 * k + jj] = b[ii][jj];
            for (int jj = 0; jj < 5; jj++) mcm[(3 + jj) * nm3 + 3 * k + ii] = gt[ii][jj];
        }
        if (gram) {
            // DOF 3k + d holds row r's value b[r][d] (r < 3) or gt[d][r - 3] (r >= 3)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                double v[NR];
#pragma unroll
                for (int r = 0; r < NR; r++) v[r] = r < 3 ? b[r][d] : gt[d][r - 3];
                int q = 0;
#pragma unroll
                for (int a = 0; a < NR; a++)
#pragma unroll
                    for (int c = a; c < NR; c++) acc[q++] += v[a] * v[c];
            }
        }
    }
    if (gram) block_reduce_store<NSYM>(acc, gram);   // uniform across the block
}

// Block reduction of NV doubles per thread into part[blockIdx.x * NV ...] (fixed order).
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double* __restrict__ part) {
    __shared__ double sh[AB / 64][NV];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // offsets outermost: the NV shuffle chains are independent and overlap
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int q = 0; q < NV; q++) v[q] += __shfl_down(v[q], off, 64);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) sh[wv][q] = v[q];
    __syncthreads();
    for (int q = threadIdx.x; q < NV; q += AB) {
        double a = 0.0;
        for (int w = 0; w < AB / 64; w++) a += sh[w][q];
        part[(size_t)blockIdx.x * NV + q] = a;
    }
}

// After every block has stored its partials: true in the one block that arrives last (which then
// resets the counter for the next launch).  Partials must be stored before the call.
__device__ __forceinline__ bool last_block(int* __restrict__ counter) {
    __shared__ int is_last;
    __threadfence();   // release: this block's partials are visible device-wide before it counts
    __syncthreads();
    if (threadIdx.x == 0) {
        is_last = atomicAdd(counter, 1) == (int)gridDim.x - 1;
        if (is_last) *counter = 0;
    }
    __syncthreads();
    if (is_last) __threadfence();   // acquire: later loads see every block's partials
    return is_last;
}
// In the last block: the NR row sums of all blocks' partials, 32 lanes per row, combined in a fixed
// order (deterministic), into rs[NR] (LDS).  Ordinary loads after last_block's acquire fence.
__device__ __forceinline__ void sum_rows(const double* part, int nb, double* rs) {
    const int r = threadIdx.x >> 5, l = threadIdx.x & 31;
    double a = 0.0;
    if (r < NR)
        for (int b = l; b < nb; b += 32) a += part[(size_t)b * NR + r];
    for (int off = 16; off > 0; off >>= 1) a += __shfl_down(a, off, 32);
    if (r < NR && l == 0) rs[r] = a;
    __syncthreads();
}

// Gram partials of the stored rows: P = M8 M8^T, upper triangle (NSYM entries) per block.
__global__ __launch_bounds__(AB) void k_ar_gram(int64_t nm3, const double* __restrict__ mcm, double* __restrict__ part) {
    double acc[NSYM];
#pragma unroll
    for (int q = 0; q < NSYM; q++) acc[q] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * AB + threadIdx.x; i < nm3; i += (int64_t)gridDim.x * AB) {
        double v[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) v[r] = mcm[r * nm3 + i];
        int q = 0;
#pragma unroll
        for (int a = 0; a < NR; a++)
#pragma unroll
            for (int b = a; b < NR; b++) acc[q++] += v[a] * v[b];
    }
    block_reduce_store<NSYM>(acc, part);
}

// The colloid self block (fill_self, src/mob.cpp:575-627) for one colloid: [F 3 | T 3 | S 5].
__device__ __forceinline__ double mcc_base(int a, int b) {
    if (a != b) return ((a == 6 && b == 10) || (a == 10 && b == 6)) ? 9. / 10. : 0.0;
    return a < 3 ? 1.0 : a < 6 ? 3. / 4. : 9. / 5.;
}

// Cholesky factor of the 5 x 5 stresslet block m (SPD: the stresslet block of the floored M^cc):
// L row-major, lower triangle, with the reciprocals of its diagonal on the diagonal.
__host__ __device__ inline void mss_factor(const double* m, double* L) {
    for (int q = 0; q < 25; q++) L[q] = 0.0;
    for (int j = 0; j < 5; j++) {
        double d = m[j * 5 + j];
        for (int k = 0; k < j; k++) d -= L[j * 5 + k] * L[j * 5 + k];
        const double ljj = sqrt(d);
        L[j * 5 + j] = 1.0 / ljj;
        for (int i = j + 1; i < 5; i++) {
            double v = m[i * 5 + j];
            for (int k = 0; k < j; k++) v -= L[i * 5 + k] * L[j * 5 + k];
            L[i * 5 + j] = v / ljj;
        }
    }
}
// Solves L L^T f = r in place with a factor from mss_factor.
__device__ __forceinline__ void mss_factor_solve(const double* L, double (&r)[5]) {
    for (int i = 0; i < 5; i++) {
        for (int k = 0; k < i; k++) r[i] -= L[i * 5 + k] * r[k];
        r[i] *= L[i * 5 + i];
    }
    for (int i = 4; i >= 0; i--) {
        for (int k = i + 1; k < 5; k++) r[i] -= L[k * 5 + i] * r[k];
        r[i] *= L[i * 5 + i];
    }
}

// sync_mcc_schur_correction (src/slip_vel.cpp:81-203) in one wave: S = M^cc_base - beta P, its
// eigen-decomposition, eigenvalues below 0.1 raised to 0.1, and M^cc = base + Q diag(delta) Q^T.
// Writes M^cc, Q (row-major, eigenvectors as columns), the floored spectrum, the 5 x 5 stresslet
// block of M^cc (mss) and its Cholesky factor (at out + 256) for the block solver, and at out + 300
// a flag: 1 when Q and the spectrum were left from an earlier sync (lazy mode, below).
//
// Only the rotation and stresslet rows (3..10) of S differ from the base: its translation block is
// the identity and decoupled (the stored rows of M^cm have no translation rows).  So the translation
// eigenpairs are (1, e_0..e_2) and the eigensolver works on the 8 x 8 block: the parallel cyclic
// Jacobi method, warm-started from the previous sync's eigenvectors Q0 (S moves little between
// syncs, so Q0^T S Q0 is nearly diagonal and a few sweeps converge it); Q0 is first
// re-orthonormalised by one Newton-Schulz step, Q0 <- (3/2) Q0 - (1/2) Q0 Q0^T Q0, so rounding does
// not accumulate over syncs.  Each round of the round-robin ordering rotates 4 disjoint index pairs
// at once (Rutishauser's angle formulas) as one two-sided update A' = J^T A J: each of 16 lanes
// updates the 2 x 2 block of A on one pair's rows and another's columns (B' = J1^T B J2); the rotated
// pairs' off-diagonal entries are set to zero and their diagonal entries become a_pp - t a_pq and
// a_qq + t a_pq.  Sweeps stop when the off-diagonal mass is below 1e-30 of the diagonal's.  The
// floored matrix and the square root Q diag(sqrt(lam)) Q^T do not depend on the eigenvector choice,
// so the warm start changes them only by rounding.
//
// Lazy mode (a sync whose Q and spectrum nothing reads before the next sync): when Gershgorin's discs
// put every eigenvalue of S above 0.1, no eigenvalue is floored and M^cc is the base block exactly,
// so the eigensolver is skipped.
constexpr int JN = 8;
// Round-robin ordering (circle method) of round r: seat 0 holds index 0, seat j >= 1 index
// 1 + (j - 1 + r) mod (JN - 1); pair k joins seats k and JN - 1 - k.
__device__ __forceinline__ int jseat(int j, int r) { return j == 0 ? 0 : 1 + (j - 1 + r) % (JN - 1); }
// the indices p < q of pair k in round r
__device__ __forceinline__ void jpair_of(int k, int r, int& p, int& q) {
    const int a = jseat(k, r), b = jseat(JN - 1 - k, r);
    p = min(a, b);
    q = max(a, b);
}
// C = X^T Y (tx) or X Y (!tx) for JN x JN LDS matrices; lane t computes entry t.
template <bool TX>
__device__ __forceinline__ void jgemm(const double (*X)[JN + 1], const double (*Y)[JN + 1], double (*C)[JN + 1]) {
    const int t = threadIdx.x, a = t / JN, b = t % JN;
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < JN; k++) r += (TX ? X[k][a] : X[a][k]) * Y[k][b];
    __syncthreads();
    C[a][b] = r;
    __syncthreads();
}
__global__ __launch_bounds__(64) void k_ar_schur(const double* __restrict__ part, int nb, double beta,
                                                 double* __restrict__ out, double* __restrict__ mss, int lazy) {
    static_assert(JN * JN == 64, "one lane per entry of the 8 x 8 block");
    __shared__ double A[JN][JN + 1], V[JN][JN + 1], Q0[JN][JN + 1], W[JN][JN + 1];
    __shared__ double P[NSYM], m5[25], lam8[JN], delta[JN];
    __shared__ int skip;
    const int t = threadIdx.x, a = t / JN, b = t % JN;
    if (t < NSYM) {
        double acc = 0.0;
#pragma unroll 16
        for (int q = 0; q < nb; q++) acc += part[(size_t)q * NSYM + t];
        P[t] = acc;
    }
    __syncthreads();
    {   // S on rows / columns 3..10
        const int i = min(a, b), j = max(a, b);   // upper-triangle index of (i, j)
        A[a][b] = mcc_base(a + 3, b + 3) - beta * P[i * NR - i * (i - 1) / 2 + (j - i)];
        Q0[a][b] = out[121 + (a + 3) * 11 + (b + 3)];
    }
    __syncthreads();
    if (lazy) {
        if (t < JN) {   // Gershgorin: row t's disc lies above 0.1
            double r = 0.0;
            for (int k = 0; k < JN; k++)
                if (k != t) r += fabs(A[t][k]);
            W[0][t] = (A[t][t] - r > 0.1) ? 1.0 : 0.0;
        }
        __syncthreads();
        if (t == 0) {
            int all = 1;
            for (int k = 0; k < JN; k++) all &= W[0][k] != 0.0;
            skip = all;
        }
        __syncthreads();
        if (skip) {
            for (int q = t; q < 121; q += 64) out[q] = mcc_base(q / 11, q % 11);
            if (t < 25) {
                const double m = mcc_base(6 + t / 5, 6 + t % 5);
                if (mss) mss[t] = m;
                out[256 + t] = out[320 + t];   // the base block's factor, stored at allocation
            }
            if (t == 0) out[300] = 1.0;
            return;
        }
    }
    // Newton-Schulz: Q0 <- 1.5 Q0 - 0.5 Q0 (Q0^T Q0)
    jgemm<true>(Q0, Q0, W);    // W = Q0^T Q0
    jgemm<false>(Q0, W, V);    // V = Q0 W
    Q0[a][b] = 1.5 * Q0[a][b] - 0.5 * V[a][b];
    __syncthreads();
    jgemm<false>(A, Q0, W);    // W = S Q0
    jgemm<true>(Q0, W, V);     // V = Q0^T S Q0
    A[a][b] = 0.5 * (V[a][b] + V[b][a]);
    __syncthreads();
    V[a][b] = (a == b) ? 1.0 : 0.0;
    __syncthreads();
    for (int sweep = 0; sweep < 60; sweep++) {
        // off-diagonal and diagonal mass (lane t: entry (a, b))
        const double v = A[a][b];
        double off = (a < b) ? v * v : 0.0, dia = (a == b) ? v * v : 0.0;
        for (int o = 32; o > 0; o >>= 1) {
            off += __shfl_xor(off, o, 64);
            dia += __shfl_xor(dia, o, 64);
        }
        if (off <= 1e-30 * dia || off == 0.0) break;   // the same value in every lane
        for (int round = 0; round < JN - 1; round++) {
            // Lane t < 16 owns the 2 x 2 block of A on rows of pair t / 4 and columns of pair t % 4
            // (B' = J1^T B J2); lane t < 32 updates row t / 4 of V on pair t % 4; lanes 0..3 compute
            // the 4 rotations, which the others take by shuffle.
            const int k1 = min(t / (JN / 2), JN / 2 - 1), k2 = t % (JN / 2);
            int p1, q1, p2, q2;
            jpair_of(k1, round, p1, q1);
            jpair_of(k2, round, p2, q2);
            double b00 = 0.0, b01 = 0.0, b10 = 0.0, b11 = 0.0;
            if (t < 16) {
                b00 = A[p1][p2];
                b01 = A[p1][q2];
                b10 = A[q1][p2];
                b11 = A[q1][q2];
            }
            const int va = min(t / (JN / 2), JN - 1);   // V task: row va, pair k2
            const double v0 = V[va][p2], v1 = V[va][q2];
            double c = 1.0, sv = 0.0, tt = 0.0;
            if (t < JN / 2) {
                int p, q;
                jpair_of(t, round, p, q);
                const double apq = A[p][q];
                if (apq != 0.0) {
                    // t = sgn(d) h / (|d| + sqrt(d^2 + h^2)), h = 2 a_pq, d = a_qq - a_pp: Rutishauser's
                    // tan of the rotation angle, the smaller root
                    const double d = A[q][q] - A[p][p], h = 2.0 * apq;
                    const double rr = sqrt(d * d + h * h);
                    tt = (d >= 0 ? h : -h) / (fabs(d) + rr);
                    c = 1.0 / sqrt(tt * tt + 1.0);
                    sv = tt * c;
                }
            }
            const double c1 = __shfl(c, k1, 64), s1 = __shfl(sv, k1, 64), t1 = __shfl(tt, k1, 64);
            const double c2 = __shfl(c, k2, 64), s2 = __shfl(sv, k2, 64);
            __syncthreads();   // every old value read
            if (t < 16) {
                if (k1 == k2) {
                    // the rotated pair: b00 = a_pp, b01 = a_pq, b11 = a_qq
                    A[p1][p1] = b00 - t1 * b01;
                    A[q1][q1] = b11 + t1 * b01;
                    A[p1][q1] = 0.0;
                    A[q1][p1] = 0.0;
                } else {
                    // M = B J2 (columns p2, q2), then J1^T M (rows p1, q1)
                    const double m00 = b00 * c2 - b01 * s2, m01 = b00 * s2 + b01 * c2;
                    const double m10 = b10 * c2 - b11 * s2, m11 = b10 * s2 + b11 * c2;
                    A[p1][p2] = c1 * m00 - s1 * m10;
                    A[p1][q2] = c1 * m01 - s1 * m11;
                    A[q1][p2] = s1 * m00 + c1 * m10;
                    A[q1][q2] = s1 * m01 + c1 * m11;
                }
            }
            if (t < 32) {
                V[va][p2] = c2 * v0 - s2 * v1;
                V[va][q2] = s2 * v0 + c2 * v1;
            }
            __syncthreads();
        }
    }
    jgemm<false>(Q0, V, W);    // Q = Q0 V (the 8 x 8 block)
    if (t < JN) {
        const double l = A[t][t];
        delta[t] = l < 0.1 ? 0.1 - l : 0.0;
        lam8[t] = l + delta[t];
    }
    __syncthreads();
    for (int q = t; q < 121; q += 64) {
        const int i = q / 11, j = q % 11;
        double corr = 0.0;
        if (i >= 3 && j >= 3)
            for (int k = 0; k < JN; k++)
                if (delta[k] != 0.0) corr += W[i - 3][k] * delta[k] * W[j - 3][k];
        const double m = mcc_base(i, j) + corr;
        out[q] = m;
        out[121 + q] = (i >= 3 && j >= 3) ? W[i - 3][j - 3] : (i == j ? 1.0 : 0.0);
        if (i >= 6 && j >= 6) {
            m5[(i - 6) * 5 + (j - 6)] = m;
            if (mss) mss[(i - 6) * 5 + (j - 6)] = m;
        }
    }
    if (t < 11) out[242 + t] = t < 3 ? 1.0 : lam8[t - 3];
    __syncthreads();
    if (t == 0) {   // the block solver's factor of the stresslet block, at out + 256
        double L[25];
        mss_factor(m5, L);
        for (int q = 0; q < 25; q++) out[256 + q] = L[q];
        out[300] = 0.0;
    }
}

// ------------------------------------------------------------------------------------ apply
// y = M x, or the saddle A x, in one launch (src/matfree_A.cpp:54-78):
//   monomer rows  y_m = beta_inv x_m + M^mc x_c  [saddle: - l_m, and y_l,m = -x_m]
//   colloid rows  y_c = M^cm x_m + M^cc x_c      [saddle: - l_c, and y_l,c = -x_c on the F, T rows]
// Each block stores its partials of M^cm x_m; the last block to finish sums them (fixed order) and
// writes the colloid rows.  xr: the colloid's rotation and stresslet entries of x (x + nm3 + 3),
// null without a colloid (then no colloid rows).
__global__ __launch_bounds__(AB) void k_ar_apply(int64_t nm3, int64_t n, int64_t nu, double binv,
                                                 const double* __restrict__ mcm, const double* __restrict__ mcc,
                                                 const double* __restrict__ x, const double* __restrict__ xr,
                                                 double* __restrict__ y, double* __restrict__ part,
                                                 int* __restrict__ counter, int saddle) {
    double p[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) p[r] = 0.0;
#pragma unroll
    for (int d = 0; d < ARD; d++) {
        const int64_t i = ((int64_t)blockIdx.x * ARD + d) * AB + threadIdx.x;
        if (i < nm3) {
            const double xi = x[i];
            double acc = 0.0;
            if (xr) {
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    const double v = mcm[r * nm3 + i];
                    acc += v * xr[r];
                    p[r] += v * xi;
                }
            }
            double yi = acc + binv * xi;
            if (saddle) {
                yi -= x[n + i];
                y[n + i] = -xi;
            }
            y[i] = yi;
        }
    }
    if (!xr) return;
    block_reduce_store<NR>(p, part);
    if (!last_block(counter)) return;
    __shared__ double rs[NR];
    sum_rows(part, gridDim.x, rs);
    const int r = threadIdx.x;
    if (r >= 11) return;
    double a2 = 0.0;
    for (int q = 0; q < 11; q++) a2 += mcc[r * 11 + q] * x[nm3 + q];
    double yr = (r >= 3 ? rs[r - 3] : 0.0) + a2;
    if (saddle && nm3 + r < nu) {
        yr -= x[n + nm3 + r];
        y[n + nm3 + r] = -x[nm3 + r];
    }
    y[nm3 + r] = yr;
}
// The colloid rows alone (no monomers): y_c = M^cc x_c (+ B terms).
__global__ void k_ar_apply_c(int64_t nm3, int64_t n, int64_t nu, const double* __restrict__ mcc,
                             const double* __restrict__ x, double* __restrict__ y, int saddle) {
    const int r = threadIdx.x;
    if (r >= 11) return;
    double yr = 0.0;
    for (int q = 0; q < 11; q++) yr += mcc[r * 11 + q] * x[nm3 + q];
    if (saddle && nm3 + r < nu) {
        yr -= x[n + nm3 + r];
        y[n + nm3 + r] = -x[nm3 + r];
    }
    y[nm3 + r] = yr;
}

// Solves M_ss f = r in place for the 5 x 5 stresslet block (SPD: the stresslet block of the floored
// M^cc) by Cholesky, in one thread.
__device__ __forceinline__ void mss_cholesky_solve(const double* __restrict__ mss, double (&r)[5]) {
    double Lm[5][5] = {};
    for (int j = 0; j < 5; j++) {
        double d = mss[j * 5 + j];
        for (int k = 0; k < j; k++) d -= Lm[j][k] * Lm[j][k];
        Lm[j][j] = sqrt(d);
        for (int i = j + 1; i < 5; i++) {
            double v = mss[i * 5 + j];
            for (int k = 0; k < j; k++) v -= Lm[i][k] * Lm[j][k];
            Lm[i][j] = v / Lm[j][j];
        }
    }
    for (int i = 0; i < 5; i++) {   // L z = r
        for (int k = 0; k < i; k++) r[i] -= Lm[i][k] * r[k];
        r[i] /= Lm[i][i];
    }
    for (int i = 4; i >= 0; i--) {   // L^T f = z
        for (int k = i + 1; k < 5; k++) r[i] -= Lm[k][i] * r[k];
        r[i] /= Lm[i][i];
    }
}

// The 5 x 5 stresslet system of the block elimination, M_ss f_s = b_s - y_s.
__global__ void k_ar_mss_solve(const double* __restrict__ mss, const double* __restrict__ b,
                               const double* __restrict__ y, int64_t nu, double* __restrict__ x) {
    if (threadIdx.x != 0) return;
    double r[5];
    for (int k = 0; k < 5; k++) r[k] = b[nu + k] - y[nu + k];
    mss_cholesky_solve(mss, r);
    for (int k = 0; k < 5; k++) x[nu + k] = r[k];
}

// ------------------------------------------------------------------------------------ block solve
// The saddle system [[M, B], [B^T, 0]] [f; u] = [b_f; b_l] with B = [-I_nu; 0] by exact block
// elimination (f_u = -b_l; M_ss f_s = b_s - (M [f_u; 0])_s; u = (M f)_u - b_u) for one colloid,
// in two DOF passes.  The only sum over monomers is R = M^cm f_m (8 rows), which both the stresslet
// system and the colloid's velocity rows need, so:
//   pass 1: f_m = -b_l,m and the partials of R; the last block to finish sums them (fixed order),
//           sets the colloid's F, T from b_l, solves M_ss f_s = b_s - R_s - M^cc_{s,FT} f_FT by
//           Cholesky and writes the colloid's velocity rows u_c = R + M^cc f_c - b_u,c;
//   pass 2: the monomer velocity rows u_m = beta_inv f_m + M^mc f_c - b_u,m.
// `uout` (optional) receives a copy of u.  Same arithmetic as two applies of k_ar_apply around the
// 5 x 5 solve, with the reduction done once.
__global__ __launch_bounds__(AB) void k_ar_bs1(int64_t nm3, int64_t n, const double* __restrict__ mcm,
                                               const double* __restrict__ mcc, const double* __restrict__ b,
                                               double* __restrict__ x,
                                               double* __restrict__ uout, double* __restrict__ part,
                                               int* __restrict__ counter) {
    double p[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) p[r] = 0.0;
#pragma unroll
    for (int d = 0; d < ARD; d++) {
        const int64_t i = ((int64_t)blockIdx.x * ARD + d) * AB + threadIdx.x;
        if (i < nm3) {
            const double f = -b[n + i];
            x[i] = f;
#pragma unroll
            for (int r = 0; r < NR; r++) p[r] += mcm[r * nm3 + i] * f;
        }
    }
    block_reduce_store<NR>(p, part);
    if (!last_block(counter)) return;
    // the last block: stage M^cc, the factor and the colloid's entries of b in LDS, sum the partials
    __shared__ double rs[NR];
    __shared__ double fc[11], bc[11], mc[121], lf[25];
    const int t = threadIdx.x;
    if (t < 121) mc[t] = mcc[t];
    else if (t < 146) lf[t - 121] = mcc[256 + (t - 121)];
    else if (t < 157) bc[t - 146] = b[nm3 + (t - 146)];
    else if (t < 163) fc[t - 157] = -b[n + nm3 + (t - 157)];
    sum_rows(part, gridDim.x, rs);   // ends with a barrier
    if (t == 0) {
        double r[5];
        for (int k = 0; k < 5; k++) {
            double a2 = 0.0;
            for (int q = 0; q < 6; q++) a2 += mc[(6 + k) * 11 + q] * fc[q];
            r[k] = bc[6 + k] - (rs[3 + k] + a2);
        }
        mss_factor_solve(lf, r);
        for (int k = 0; k < 5; k++) fc[6 + k] = r[k];
    }
    __syncthreads();
    if (t < 11) x[nm3 + t] = fc[t];
    if (t < 6) {
        double a2 = 0.0;
        for (int q = 0; q < 11; q++) a2 += mc[t * 11 + q] * fc[q];
        const double u = ((t >= 3 ? rs[t - 3] : 0.0) + a2) - bc[t];
        x[n + nm3 + t] = u;
        if (uout) uout[nm3 + t] = u;
    }
}
__global__ __launch_bounds__(AB) void k_ar_bs2(int64_t nm3, int64_t n, double binv, const double* __restrict__ mcm,
                                               const double* __restrict__ b, double* __restrict__ x,
                                               double* __restrict__ uout) {
    const int64_t i = (int64_t)blockIdx.x * AB + threadIdx.x;
    if (i >= nm3) return;
    const double* xr = x + nm3 + 3;   // the colloid's torque and stresslet entries of f
    const double f = x[i];
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < NR; r++) acc += mcm[r * nm3 + i] * xr[r];
    const double u = (acc + binv * f) - b[i];
    x[n + i] = u;
    if (uout) uout[i] = u;
}

// ------------------------------------------------------------------------------------ noise
// build_slip_vel_schur in one launch: u_m = sqrt(beta_inv) xi_m; the last block adds
// u_c = sqrt(beta) M^cm xi_m + Q diag(sqrt(lam)) Q^T xi_c.
__global__ __launch_bounds__(AB) void k_ar_noise(int64_t nm3, double sq_binv, double sq_beta,
                                                 const double* __restrict__ mcm, const double* __restrict__ mcc_q_lam,
                                                 const double* __restrict__ xi, double* __restrict__ out,
                                                 double* __restrict__ part, int* __restrict__ counter, int colloid) {
    double p[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) p[r] = 0.0;
#pragma unroll
    for (int d = 0; d < ARD; d++) {
        const int64_t i = ((int64_t)blockIdx.x * ARD + d) * AB + threadIdx.x;
        if (i < nm3) {
            const double v = xi[i];
#pragma unroll
            for (int r = 0; r < NR; r++) p[r] += mcm[r * nm3 + i] * v;
            out[i] = sq_binv * v;
        }
    }
    if (!colloid) return;
    block_reduce_store<NR>(p, part);
    if (!last_block(counter)) return;
    __shared__ double rs[NR];
    __shared__ double yv[11];
    sum_rows(part, gridDim.x, rs);
    const double* Q = mcc_q_lam + 121;
    const double* lam = mcc_q_lam + 242;
    const int r = threadIdx.x;
    if (r < 11) {
        double yj = 0.0;
        for (int k = 0; k < 11; k++) yj += Q[k * 11 + r] * xi[nm3 + k];
        yv[r] = sqrt(lam[r]) * yj;
    }
    __syncthreads();
    if (r >= 11) return;
    const double uc = (r >= 3 ? rs[r - 3] : 0.0) * sq_beta;
    double si = 0.0;
    for (int j = 0; j < 11; j++) si += Q[r * 11 + j] * yv[j];
    out[nm3 + r] = uc + si;
}

// ------------------------------------------------------------------------------------ diagonal
// ArrowheadGetDiagonal + fix-diagonal (src/matfree_A.cpp:84-102): [beta_inv; diag(M^cc); 1]; as
// the inverse when `inverse` is set.
__global__ void k_ar_diag(int64_t nm3, int64_t n, int64_t ns, double beta, const double* __restrict__ mcc,
                          double* __restrict__ d, int inverse) {
    const int64_t i = (int64_t)blockIdx.x * AB + threadIdx.x;
    if (i >= ns) return;
    double v = (i < nm3) ? 1.0 / beta : (i < n) ? mcc[(i - nm3) * 12] : 1.0;
    d[i] = inverse ? 1.0 / v : v;
}

// ------------------------------------------------------------------------------------ WCA cell list
struct WcaGrid {
    int nc[3];
    double size[3], org[3];
    int periodic;
};
__device__ __forceinline__ int wca_cell_coord(const WcaGrid& g, double x, int d) {
    int ci = (int)floor((x - g.org[d]) / g.size[d]);
    if (ci < 0) ci = 0;
    if (ci >= g.nc[d]) ci = g.nc[d] - 1;
    return ci;
}
__global__ void k_wc_bin(int Nm, WcaGrid g, const double* __restrict__ x, int* __restrict__ cell_of,
                         int* __restrict__ count) {
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= Nm) return;
    const int cx = wca_cell_coord(g, x[3 * i], 0), cy = wca_cell_coord(g, x[3 * i + 1], 1),
              cz = wca_cell_coord(g, x[3 * i + 2], 2);
    const int c = cx + g.nc[0] * (cy + g.nc[1] * cz);
    cell_of[i] = c;
    atomicAdd(&count[c], 1);
}
__global__ void k_wc_scatter(int Nm, const int* __restrict__ cell_of, const int* __restrict__ start,
                             int* __restrict__ cursor, int* __restrict__ mem) {
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= Nm) return;
    const int c = cell_of[i];
    mem[start[c] + atomicAdd(&cursor[c], 1)] = i;
}
// Members of each cell in ascending index (a deterministic summation order).
__global__ void k_wc_sort(int ncells, const int* __restrict__ start, int* __restrict__ mem) {
    const int c = blockIdx.x * AB + threadIdx.x;
    if (c >= ncells) return;
    const int a = start[c], b = start[c + 1];
    for (int i = a + 1; i < b; i++) {
        const int v = mem[i];
        int j = i - 1;
        while (j >= a && mem[j] > v) {
            mem[j + 1] = mem[j];
            j--;
        }
        mem[j + 1] = v;
    }
}
// Force on monomer i from every monomer j within the WCA cutoff in the 27 surrounding cells
// (monomer_wca, src/cell_list.cpp:98-172, gathered per monomer; the reference visits each pair once
// and applies it to both).  Added to the force slot of the RHS, which holds minus the force.
__global__ void k_wc_force(SysConst s, WcaGrid g, const double* __restrict__ x, const int* __restrict__ cell_of,
                           const int* __restrict__ start, const int* __restrict__ mem, double sig, double rcut,
                           double eps, double* __restrict__ rhs) {
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= s.Nm) return;
    const int c = cell_of[i];
    const int cx = c % g.nc[0], cy = (c / g.nc[0]) % g.nc[1], cz = c / (g.nc[0] * g.nc[1]);
    const double rcut2 = rcut * rcut;
    const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int oz = -1; oz <= 1; oz++)
        for (int oy = -1; oy <= 1; oy++)
            for (int ox = -1; ox <= 1; ox++) {
                int nx = cx + ox, ny = cy + oy, nz = cz + oz;
                if (g.periodic) {
                    nx = (nx % g.nc[0] + g.nc[0]) % g.nc[0];
                    ny = (ny % g.nc[1] + g.nc[1]) % g.nc[1];
                    nz = (nz % g.nc[2] + g.nc[2]) % g.nc[2];
                } else if (nx < 0 || nx >= g.nc[0] || ny < 0 || ny >= g.nc[1] || nz < 0 || nz >= g.nc[2]) {
                    continue;
                }
                const int nb = nx + g.nc[0] * (ny + g.nc[1] * nz);
                for (int m = start[nb]; m < start[nb + 1]; m++) {
                    const int j = mem[m];
                    if (j == i) continue;
                    double dx = xi0 - x[3 * j], dy = xi1 - x[3 * j + 1], dz = xi2 - x[3 * j + 2];
                    if (s.box_active) {
                        dx -= s.L[0] * rint(dx * s.Linv[0]);
                        dy -= s.L[1] * rint(dy * s.Linv[1]);
                        dz -= s.L[2] * rint(dz * s.Linv[2]);
                    }
                    const double dr2 = dx * dx + dy * dy + dz * dz;
                    if (dr2 >= rcut2 || dr2 == 0.0) continue;
                    const double dr = sqrt(dr2), dr_inv = 1.0 / dr;
                    if (dr >= rcut) continue;
                    const double sr = sig * dr_inv, s2 = sr * sr, s6 = s2 * s2 * s2, s12 = s6 * s6;
                    const double fmag = 24.0 * eps * dr_inv * (2.0 * s12 - s6);
                    if (fmag == 0.0) continue;
                    f0 += fmag * dx * dr_inv;
                    f1 += fmag * dy * dr_inv;
                    f2 += fmag * dz * dr_inv;
                }
            }
    rhs[s.n + 3 * i] -= f0;
    rhs[s.n + 3 * i + 1] -= f1;
    rhs[s.n + 3 * i + 2] -= f2;
}
// Bounding box of the monomers (the unbounded cell grid), one block.
__global__ void k_wc_bbox(int Nm, const double* __restrict__ x, double* __restrict__ out) {
    __shared__ double lo[3][256], hi[3][256];
    const int t = threadIdx.x;
    double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
    for (int i = t; i < Nm; i += blockDim.x)
        for (int d = 0; d < 3; d++) {
            l[d] = fmin(l[d], x[3 * i + d]);
            h[d] = fmax(h[d], x[3 * i + d]);
        }
    for (int d = 0; d < 3; d++) {
        lo[d][t] = l[d];
        hi[d][t] = h[d];
    }
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int d = 0; d < 3; d++) {
                lo[d][t] = fmin(lo[d][t], lo[d][t + w]);
                hi[d][t] = fmax(hi[d][t], hi[d][t + w]);
            }
        __syncthreads();
    }
    if (t < 3) {
        out[t] = lo[t][0];
        out[3 + t] = hi[t][0];
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------- host side
ps_err arrow_alloc(Ctx* c) {
    const int64_t nm3 = c->sc.nm3;
    const int nb = (int)dof_blocks(nm3);
    c->ar_nb = nb;
    PS_HIP(c, hipMalloc(&c->ar_mcm, sizeof(double) * (size_t)std::max<int64_t>(NR * nm3, 1)));
    PS_HIP(c, hipMemset(c->ar_mcm, 0, sizeof(double) * (size_t)std::max<int64_t>(NR * nm3, 1)));
    PS_HIP(c, hipMalloc(&c->ar_part, sizeof(double) * (size_t)nb * NSYM));
    PS_HIP(c, hipMalloc(&c->ar_gpart, sizeof(double) * (size_t)dof_blocks(c->Np) * NSYM));
    PS_HIP(c, hipMalloc(&c->ar_mcc, sizeof(double) * 512));
    PS_HIP(c, hipMalloc(&c->ar_flag, sizeof(int) * (size_t)(c->Nm + 2)));
    PS_HIP(c, hipMalloc(&c->ar_pos, sizeof(int) * (size_t)(c->Nm + 2)));
    PS_HIP(c, hipHostMalloc(&c->ar_host, sizeof(double) * 8, hipHostMallocDefault));
    PS_HIP(c, hipMalloc(&c->ar_bscan, sizeof(int) * 2 * (size_t)dof_blocks(c->Np)));
    PS_HIP(c, hipMalloc(&c->ar_counter, sizeof(int) * 3));
    PS_HIP(c, hipMemset(c->ar_counter, 0, sizeof(int) * 3));
    // Verlet partners: at most 2 Nm entries (each monomer with the colloid, both directions)
    if (c->nb_cap < 2 * (int64_t)c->Nm + 2) {
        if (c->nb_part) (void)hipFree(c->nb_part);
        c->nb_part = nullptr;
        c->nb_cap = 2 * (int64_t)c->Nm + 2;
        PS_HIP(c, hipMalloc(&c->nb_part, sizeof(int) * (size_t)c->nb_cap));
    }
    // M^cc starts as the colloid self block, Q as the identity (src/mob.cpp:629-632)
    double h[512] = {};
    for (int a = 0; a < 11; a++) {
        h[a * 11 + a] = a < 3 ? 1.0 : a < 6 ? 0.75 : 1.8;
        h[121 + a * 11 + a] = 1.0;
    }
    h[6 * 11 + 10] = h[10 * 11 + 6] = 0.9;
    {
        double m5[25];
        for (int a = 0; a < 5; a++)
            for (int b = 0; b < 5; b++) m5[a * 5 + b] = h[(6 + a) * 11 + 6 + b];
        mss_factor(m5, h + 256);
        mss_factor(m5, h + 320);   // kept: the factor of the base block, for lazy syncs
    }
    PS_HIP(c, hipMemcpy(c->ar_mcc, h, sizeof(double) * 512, hipMemcpyHostToDevice));
    if (c->Mss) {
        double m5[25] = {};
        for (int a = 0; a < 5; a++)
            for (int b = 0; b < 5; b++) m5[a * 5 + b] = h[(6 + a) * 11 + 6 + b];
        PS_HIP(c, hipMemcpy(c->Mss, m5, sizeof(m5), hipMemcpyHostToDevice));
    }
    return PS_OK;
}

void arrow_free(Ctx* c) {
    for (void* p : {(void*)c->ar_mcm, (void*)c->ar_part, (void*)c->ar_mcc, (void*)c->ar_flag, (void*)c->ar_pos,
                    (void*)c->ar_bscan, (void*)c->ar_counter, (void*)c->ar_gpart,
                    (void*)c->wc_cell, (void*)c->wc_count, (void*)c->wc_start, (void*)c->wc_mem,
                    (void*)c->ar_bbox})
        if (p) (void)hipFree(p);
    if (c->ar_host) (void)hipHostFree(c->ar_host);
    c->ar_mcm = c->ar_part = c->ar_gpart = c->ar_mcc = c->ar_bbox = c->ar_host = nullptr;
    c->ar_flag = c->ar_pos = c->ar_bscan = c->wc_cell = c->wc_count = c->wc_start = c->wc_mem = c->ar_counter = nullptr;
}

// The Verlet list with no host round trip: a monomer has at most one partner (the colloid), so the
// partner array's capacity is fixed at allocation (2 Nm); the pair count is read only by
// ps_check_dist.
ps_err arrow_check_dist(Ctx* c) {
    hipStream_t st = c->stream;
    const int Nm = c->Nm, Np = c->Np;
    const double rverl_ab = (double)(float)(1.5 + c->p.beta);
    const unsigned nb1 = dof_blocks(Np);
    k_ar_vl1<<<nb1, AB, 0, st>>>(c->sc, c->x, c->xpd, rverl_ab, c->ar_flag, c->ar_pos, c->ar_bscan, c->nb_off,
                                 c->ar_counter + 2);
    PS_LAUNCH(c);
    if (Nm > 0) {
        k_ar_vl2<<<dof_blocks(Nm), AB, 0, st>>>(Nm, (int)nb1, c->ar_flag, c->ar_pos, c->ar_bscan, c->nb_off,
                                                c->nb_part);
        PS_LAUNCH(c);
    }
    c->nvlist = -1;   // counted on demand (arrow_count_pairs)
    c->have_pd = true;
    return PS_OK;
}

ps_err arrow_count_pairs(Ctx* c) {
    PS_HIP(c, hipMemcpyAsync(c->hflag, c->nb_off + c->Np, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    c->nvlist = c->hflag[0] / 2;
    c->cnt.nvlist = c->nvlist;
    return PS_OK;
}

// M^cm of this mob(); when a Schur floor follows (ar_gram_next), the same launch stores the Gram
// partials it needs.
ps_err arrow_mob(Ctx* c) {
    const bool gram = c->ar_gram_next && c->Nc == 1 && c->Nm > 0;
    k_ar_mcm<<<dof_blocks(c->Np), AB, 0, c->stream>>>(c->sc, c->xpd, c->xm, c->p.beta * c->p.beta, c->ar_mcm,
                                                      gram ? c->ar_gpart : nullptr);
    PS_LAUNCH(c);
    c->ar_gram_next = false;
    c->ar_gram_ready = gram;
    c->have_mob = true;
    return PS_OK;
}

// The Schur floor of the current M^cm.  `lazy`: the caller reads only the floored M^cc before the
// next sync (k_ar_schur may then skip the eigensolver and flag Q as left over; ps_get_schur redoes
// the sync from the same Gram partials).
ps_err arrow_schur(Ctx* c, int lazy) {
    if (c->Nc != 1) return PS_OK;
    hipStream_t st = c->stream;
    const int64_t nm3 = c->sc.nm3;
    if (c->ar_gram_ready) {   // the Gram partials of the current M^cm, from arrow_mob
        c->ar_last_part = c->ar_gpart;
        c->ar_last_nb = (int)dof_blocks(c->Np);
    } else {
        const int nb = std::min(c->ar_nb, 64);   // grid-stride: few partials for the one-block finish
        if (nm3 > 0) {
            k_ar_gram<<<nb, AB, 0, st>>>(nm3, c->ar_mcm, c->ar_part);
            PS_LAUNCH(c);
        }
        c->ar_last_part = c->ar_part;
        c->ar_last_nb = nm3 > 0 ? nb : 0;
        lazy = 0;   // ar_part is reused by the applies: no later redo from it
    }
    k_ar_schur<<<1, 64, 0, st>>>(c->ar_last_part, c->ar_last_nb, c->p.beta, c->ar_mcc, c->Mss, lazy);
    PS_LAUNCH(c);
    c->ar_gram_ready = false;
    c->ar_synced = true;
    return PS_OK;
}

// The last Schur floor again, in full, from its Gram partials (after a lazy one skipped Q).
ps_err arrow_schur_redo(Ctx* c) {
    if (!c->ar_last_part) return PS_OK;
    k_ar_schur<<<1, 64, 0, c->stream>>>(c->ar_last_part, c->ar_last_nb, c->p.beta, c->ar_mcc, c->Mss, 0);
    PS_LAUNCH(c);
    return PS_OK;
}

ps_err arrow_apply(Ctx* c, const double* x, double* y, int saddle) {
    hipStream_t st = c->stream;
    const int64_t nm3 = c->sc.nm3, n = c->n, nu = c->nu;
    const unsigned nb = red_blocks(nm3);
    if (nm3 > 0) {
        // timing (ps_profile_begin/end): the launch records its own start/stop events at dispatch, or
        // is bracketed by event records under rocprofv3
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (c->prof_on) PS_TRY(c->prof_record(3, &e0, &e1));
        const bool dispatch_ev = c->prof_on && c->prof_dispatch_events;
        if (c->prof_on && !dispatch_ev) PS_HIP(c, hipEventRecord(e0, st));
        hipExtLaunchKernelGGL(k_ar_apply, dim3(nb), dim3(AB), 0, st, dispatch_ev ? e0 : nullptr,
                              dispatch_ev ? e1 : nullptr, 0, nm3, n, nu, 1.0 / c->p.beta, (const double*)c->ar_mcm,
                              (const double*)c->ar_mcc, x, c->Nc == 1 ? x + nm3 + 3 : (const double*)nullptr, y,
                              c->ar_part, c->ar_counter, saddle);
        PS_LAUNCH(c);
        if (c->prof_on && !dispatch_ev) PS_HIP(c, hipEventRecord(e1, st));
    } else if (c->Nc == 1) {
        k_ar_apply_c<<<1, 64, 0, st>>>(nm3, n, nu, c->ar_mcc, x, y, saddle);
        PS_LAUNCH(c);
    }
    c->cnt.applies++;
    return PS_OK;
}

// One block-eliminated saddle solve (Nc == 1, Nm > 0): k_ar_bs1 + k_ar_bs2.  Timing kind 4 spans the
// first dispatch to the second's end.
ps_err arrow_block_solve(Ctx* c, const double* b, double* x, double* uout) {
    hipStream_t st = c->stream;
    const int64_t nm3 = c->sc.nm3, n = c->n;
    const unsigned nb1 = red_blocks(nm3), nb2 = dof_blocks(nm3);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->prof_on) PS_TRY(c->prof_record(4, &e0, &e1));
    const bool dispatch_ev = c->prof_on && c->prof_dispatch_events;
    if (c->prof_on && !dispatch_ev) PS_HIP(c, hipEventRecord(e0, st));
    if (dispatch_ev) {
        hipExtLaunchKernelGGL(k_ar_bs1, dim3(nb1), dim3(AB), 0, st, e0, nullptr, 0, nm3, n, (const double*)c->ar_mcm,
                              (const double*)c->ar_mcc, b, x, uout, c->ar_part, c->ar_counter);
        PS_LAUNCH(c);
        hipExtLaunchKernelGGL(k_ar_bs2, dim3(nb2), dim3(AB), 0, st, nullptr, e1, 0, nm3, n, 1.0 / c->p.beta,
                              (const double*)c->ar_mcm, b, x, uout);
        PS_LAUNCH(c);
    } else {   // plain launches (also what a step graph captures)
        k_ar_bs1<<<nb1, AB, 0, st>>>(nm3, n, c->ar_mcm, c->ar_mcc, b, x, uout, c->ar_part, c->ar_counter);
        PS_LAUNCH(c);
        k_ar_bs2<<<nb2, AB, 0, st>>>(nm3, n, 1.0 / c->p.beta, c->ar_mcm, b, x, uout);
        PS_LAUNCH(c);
    }
    if (c->prof_on && !dispatch_ev) PS_HIP(c, hipEventRecord(e1, st));
    c->cnt.applies += 2;
    return PS_OK;
}

ps_err arrow_mss_solve(Ctx* c, const double* b, const double* y, double* x) {
    k_ar_mss_solve<<<1, 64, 0, c->stream>>>(c->Mss, b, y, c->nu, x);
    PS_LAUNCH(c);
    return PS_OK;
}

ps_err arrow_noise(Ctx* c, const double* xi, double* out) {
    hipStream_t st = c->stream;
    const int64_t nm3 = c->sc.nm3;
    if (c->Nc == 1 && !c->ar_synced) PS_TRY(arrow_schur(c, 0));
    k_ar_noise<<<red_blocks(nm3), AB, 0, st>>>(nm3, std::sqrt(1.0 / c->p.beta), std::sqrt(c->p.beta), c->ar_mcm,
                                               c->ar_mcc, xi, out, c->ar_part, c->ar_counter + 1, c->Nc == 1);
    PS_LAUNCH(c);
    c->cnt.sqrt_calls++;
    return PS_OK;
}

ps_err arrow_diag(Ctx* c, double* d, int inverse) {
    k_ar_diag<<<dof_blocks(c->ns), AB, 0, c->stream>>>(c->sc.nm3, c->n, c->ns, c->p.beta, c->ar_mcc, d, inverse);
    PS_LAUNCH(c);
    return PS_OK;
}

// The monomer cell list of the WCA forces: the reference's grid (src/cell_list.cpp:23-96) with cells
// of the AA Verlet cutoff, periodic when every axis has >= 3 cells, else one cell; unbounded: the
// monomers' bounding box.
ps_err arrow_wca(Ctx* c) {
    hipStream_t st = c->stream;
    const int Nm = c->Nm;
    if (Nm == 0) return PS_OK;
    const double rcell = (double)(float)(0.05 + 2 * c->p.beta);
    WcaGrid g{};
    if (c->sc.box_active) {
        int nmin = 1 << 30;
        for (int d = 0; d < 3; d++) {
            g.nc[d] = (int)std::floor(c->sc.L[d] / rcell);
            nmin = std::min(nmin, g.nc[d]);
        }
        if (nmin >= 3) {
            for (int d = 0; d < 3; d++) {
                g.size[d] = c->sc.L[d] / g.nc[d];
                g.org[d] = -0.5 * c->sc.L[d];
            }
            g.periodic = 1;
        } else {
            for (int d = 0; d < 3; d++) {
                g.nc[d] = 1;
                g.size[d] = c->sc.L[d] > 0.0 ? c->sc.L[d] : 1.0;
                g.org[d] = -0.5 * c->sc.L[d];
            }
        }
    } else {
        if (!c->ar_bbox) PS_HIP(c, hipMalloc(&c->ar_bbox, sizeof(double) * 6));
        k_wc_bbox<<<1, 256, 0, st>>>(Nm, c->x, c->ar_bbox);
        PS_LAUNCH(c);
        PS_HIP(c, hipMemcpyAsync(c->ar_host, c->ar_bbox, sizeof(double) * 6, hipMemcpyDeviceToHost, st));
        PS_HIP(c, hipStreamSynchronize(st));
        for (int d = 0; d < 3; d++) {
            const double ext = c->ar_host[3 + d] - c->ar_host[d];
            g.nc[d] = std::max(1, (int)std::floor(ext / rcell));
            g.size[d] = ext > 0.0 ? ext / g.nc[d] : rcell;
            g.org[d] = c->ar_host[d];
        }
    }
    const int64_t ncells = (int64_t)g.nc[0] * g.nc[1] * g.nc[2];
    if (ncells > (int64_t)1 << 30) {
        c->set_error("WCA cell list: more than 2^30 cells");
        return PS_ERR_ARG;
    }
    if (ncells > c->wc_cap) {
        for (int** p : {&c->wc_count, &c->wc_start})
            if (*p) (void)hipFree(*p);
        c->wc_cap = ncells;
        PS_HIP(c, hipMalloc(&c->wc_count, sizeof(int) * (size_t)ncells));
        PS_HIP(c, hipMalloc(&c->wc_start, sizeof(int) * (size_t)(ncells + 2)));
    }
    if (!c->wc_cell) {
        PS_HIP(c, hipMalloc(&c->wc_cell, sizeof(int) * (size_t)Nm));
        PS_HIP(c, hipMalloc(&c->wc_mem, sizeof(int) * (size_t)Nm));
    }
    for (int d = 0; d < 3; d++) {
        c->wc_dims[d] = g.nc[d];
        c->wc_geom[d] = g.size[d];
        c->wc_geom[3 + d] = g.org[d];
    }
    c->wc_geom[6] = g.periodic;
    PS_HIP(c, hipMemsetAsync(c->wc_count, 0, sizeof(int) * (size_t)ncells, st));
    k_wc_bin<<<dof_blocks(Nm), AB, 0, st>>>(Nm, g, c->x, c->wc_cell, c->wc_count);
    PS_LAUNCH(c);
    PS_TRY(scan_ints(c, c->wc_count, c->wc_start, (int)ncells));
    PS_HIP(c, hipMemsetAsync(c->wc_count, 0, sizeof(int) * (size_t)ncells, st));
    k_wc_scatter<<<dof_blocks(Nm), AB, 0, st>>>(Nm, c->wc_cell, c->wc_start, c->wc_count, c->wc_mem);
    PS_LAUNCH(c);
    k_wc_sort<<<dof_blocks(ncells), AB, 0, st>>>((int)ncells, c->wc_start, c->wc_mem);
    PS_LAUNCH(c);
    const double sig = (double)(float)(2. * c->p.beta);
    const double rcut = (double)(float)(std::pow(2, 1.0 / 6.0) * sig);
    k_wc_force<<<dof_blocks(Nm), AB, 0, st>>>(c->sc, g, c->x, c->wc_cell, c->wc_start, c->wc_mem, sig, rcut,
                                              c->p.epsilon, c->rhs);
    PS_LAUNCH(c);
    c->wc_built = true;
    return PS_OK;
}
