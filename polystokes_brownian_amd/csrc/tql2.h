// tql2.h -- symmetric tridiagonal eigensolver on one wavefront (device code).
//
// Implicit QL with shifts (EISPACK tql2: Bowdler, Martin, Reinsch & Wilkinson, Numer. Math. 11
// (1968) 293-306), eigenvectors accumulated.  It serves the Lanczos square root, whose tridiagonal
// T_m = U diag(theta) U^T gives M^{1/2} xi ~ beta0 Q U diag(sqrt(max(theta, 0))) U^T e1
// (src/slip_vel.cpp:427-459, where LAPACK dsyev plays this part).
//
// Layout: lane k holds d[k] and e[k] in registers and row k of Z (column-major m x m in LDS,
// Z[k + col m]).  The QL sweep is one sequential chain of plane rotations; every lane runs it on
// scalars read from the owning lanes (readlane, no LDS round trip and no fence per rotation), and
// lane i keeps rotation i's cosine and sine.  After the sweep each lane applies the sweep's rotations
// to its own row of Z in the same order, so the chain never waits for LDS and the row updates'
// loads do not depend on it.  A rotation's radius comes from a reciprocal square root refined by two
// Newton steps (one chain of multiply-adds instead of a square root and two divisions).  m <= 64.
#pragma once
#include <hip/hip_runtime.h>

__device__ __forceinline__ double tql2_rl(double v, int lane) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// 1/sqrt(t) for t > 0 to FP64 accuracy: hardware estimate, two Newton-Raphson steps.
__device__ __forceinline__ double tql2_rsqrt(double t) {
    double y = __builtin_amdgcn_rsq(t);
    double h = t * y;
    double r = __builtin_fma(-h, y, 1.0);
    y = __builtin_fma(0.5 * y, r, y);
    h = t * y;
    r = __builtin_fma(-h, y, 1.0);
    y = __builtin_fma(0.5 * y, r, y);
    return y;
}

// Eigenvalues of T (diagonal dv, off-diagonal ev of lane k; ev of lane m-1 unused) into dv,
// eigenvectors into the columns of Z.  Returns false when an eigenvalue needs more than 200 QL
// iterations (T is then left partly reduced).  LDZ = 0: Z is m x m (column stride m), lanes >= m
// idle in the row updates.  LDZ = 64: column stride 64 (Z holds 64 m doubles), every lane updates its
// own row and rows >= m carry values no one reads, so the row updates need no lane mask.
template <int LDZ = 0>
__device__ inline bool tql2_wave(int m, double& dv, double& ev, double* Z) {
    static_assert(LDZ == 0 || LDZ == 64, "column stride m or 64");
    const int lane = threadIdx.x & 63;
    const int ld = LDZ ? LDZ : m;
    const bool own = LDZ ? true : lane < m;
    if (LDZ) {
        for (int col = 0; col < m; col++) Z[lane + col * ld] = lane == col ? 1.0 : 0.0;
    } else {
        for (int k = lane; k < m * m; k += 64) Z[k] = 0.0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane < m) Z[lane + lane * m] = 1.0;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane >= m - 1) ev = 0.0;
    for (int l = 0; l < m; l++) {
        int iter = 0, mm;
        do {
            // mm: the first index >= l with a negligible e[mm], else m - 1 (all lanes test at once)
            const double dnext = __shfl_down(dv, 1, 64);
            const bool small = lane >= l && lane < m - 1 &&
                               fabs(ev) <= 1e-300 + 2.2e-16 * (fabs(dv) + fabs(dnext));
            const unsigned long long mask = __ballot(small);
            mm = mask ? (int)__builtin_ctzll(mask) : m - 1;
            if (mm != l) {
                if (iter++ == 200) return false;
                const double dl = tql2_rl(dv, l), el = tql2_rl(ev, l);
                double g0 = (tql2_rl(dv, l + 1) - dl) / (2.0 * el);
                const double r0 = sqrt(g0 * g0 + 1.0);
                const double dmm = tql2_rl(dv, mm);   // d[mm] before this sweep changes it
                g0 = dmm - dl + el / (g0 + (g0 >= 0 ? fabs(r0) : -fabs(r0)));
                const double dv0 = dv, ev0 = ev;
                double cr = 1.0, sr = 0.0;   // lane i: rotation i of this sweep
                double p = 0.0, g = g0;
                int ilast = l;
                bool underflow = false;
                // The sweep without the test for a vanishing rotation (t = 0), which only records
                // whether one occurred: then d and e are restored and the sweep is run again with the
                // test, deflating at the first such rotation.  d[i] and e[i] are read from lane i,
                // which this sweep writes only at rotation i - 1.
                {
                    double s = 1.0, c = 1.0, dip1 = dmm;
                    bool zero = false;
                    for (int i = mm - 1; i >= l; i--) {
                        const double ei = tql2_rl(ev, i), di = tql2_rl(dv, i);
                        const double f = s * ei, b = c * ei;
                        const double t = __builtin_fma(f, f, g * g);
                        zero = zero || t == 0.0;
                        const double y = tql2_rsqrt(t);
                        const double r = t * y;
                        s = f * y;
                        c = g * y;
                        g = dip1 - p;
                        const double rr = (di - g) * s + 2.0 * c * b;
                        p = s * rr;
                        if (lane == i + 1) {
                            ev = r;
                            dv = g + p;
                        }
                        if (lane == i) {
                            cr = c;
                            sr = s;
                        }
                        g = c * rr - b;
                        dip1 = di;
                    }
                    // t is the same in every lane: a uniform branch
                    if (__builtin_amdgcn_readfirstlane((int)zero)) {
                        dv = dv0;
                        ev = ev0;
                        s = 1.0;
                        c = 1.0;
                        p = 0.0;
                        g = g0;
                        dip1 = dmm;
                        ilast = mm;
                        for (int i = mm - 1; i >= l; i--) {
                            const double ei = tql2_rl(ev, i), di = tql2_rl(dv, i);
                            const double f = s * ei, b = c * ei;
                            const double t = __builtin_fma(f, f, g * g);
                            if (__builtin_amdgcn_readfirstlane((int)(t == 0.0))) {
                                // both zero: the rotation is the identity, deflate here
                                if (lane == i + 1) {
                                    ev = 0.0;
                                    dv = dip1 - p;
                                }
                                if (lane == mm) ev = 0.0;
                                underflow = true;
                                break;
                            }
                            ilast = i;
                            const double y = tql2_rsqrt(t);
                            const double r = t * y;
                            s = f * y;
                            c = g * y;
                            g = dip1 - p;
                            const double rr = (di - g) * s + 2.0 * c * b;
                            p = s * rr;
                            if (lane == i + 1) {
                                ev = r;
                                dv = g + p;
                            }
                            if (lane == i) {
                                cr = c;
                                sr = s;
                            }
                            g = c * rr - b;
                            dip1 = di;
                        }
                    }
                }
                // the rotations mm - 1 .. ilast on each lane's row of Z: column i + 1 is carried in zc
                // from one rotation to the next (each reads column i, not yet written by this pass);
                // column i - 1 is loaded one rotation ahead, into za and zb in turn
                {
                    double zc = own ? Z[lane + mm * ld] : 0.0;
                    double za = (own && ilast < mm) ? Z[lane + (mm - 1) * ld] : 0.0, zb;
                    int i = mm - 1;
                    while (i >= ilast) {
                        zb = own ? Z[lane + max(i - 1, ilast) * ld] : 0.0;
                        {
                            const double ci = tql2_rl(cr, i), si = tql2_rl(sr, i);
                            if (own) Z[lane + (i + 1) * ld] = si * za + ci * zc;
                            zc = ci * za - si * zc;
                        }
                        if (--i < ilast) break;
                        za = own ? Z[lane + max(i - 1, ilast) * ld] : 0.0;
                        {
                            const double ci = tql2_rl(cr, i), si = tql2_rl(sr, i);
                            if (own) Z[lane + (i + 1) * ld] = si * zb + ci * zc;
                            zc = ci * zb - si * zc;
                        }
                        --i;
                    }
                    if (own) Z[lane + ilast * ld] = zc;   // the column the last rotation left in zc
                }
                if (underflow) continue;
                if (lane == l) {
                    dv -= p;
                    ev = g;
                }
                if (lane == mm) ev = 0.0;
            }
        } while (mm != l);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    return true;
}

// 1/sqrt(t) for t > 0: hardware estimate (relative error about 2^-22), one third-order step
// y (1 + r/2 + 3r^2/8), r = 1 - t y^2, whose error (5/16) r^3 is below double rounding.
__device__ __forceinline__ double tql2_rsqrt3(double t) {
    const double y = __builtin_amdgcn_rsq(t);
    const double h = t * y;
    const double r = __builtin_fma(-h, y, 1.0);
    const double q = __builtin_fma(r, 0.375, 0.5);
    return __builtin_fma(y * r, q, y);
}

// 1/x: hardware estimate and two Newton steps.
__device__ __forceinline__ double tql2_rcp(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(y, e, y);
}

// The same QL on a block of two waves.  Wave 0 runs the sweeps with d and e in its lanes' registers
// (lane k: d[k], e[k]); a sweep reads d[i], e[i] from a copy taken before it (rotation i - 1 is the
// first to change them), one rotation ahead, so the reads wait for nothing in the chain.  After
// each sweep its rotations' cosine and sine go to LDS, and wave 1 applies them to its rows of Z
// (column stride 64, every lane a row) while wave 0 runs the next sweep.  Phase k: wave 0 runs at
// most one sweep and publishes it in sh.rot[k & 1] with sh.desc[k & 1] = {mm, ilast, has a sweep,
// state}; wave 1 applies the sweep published in phase k - 1; one block barrier ends the phase.  The
// loop ends after the phase in which wave 0 has finished (state 1, or 2 when an eigenvalue needed
// more than 200 iterations) and published nothing, read by both waves from the same LDS words after
// the same barrier.  The shift's divisions and square root, and each rotation's radius, use the
// refined hardware estimates above (the rotations agree with tql2_wave's to rounding).  A vanishing
// rotation (t = 0) makes the sweep's p and g NaN (0 times the infinite estimate): the sweep is then
// run again from the copy with the test, deflating there.  On entry wave 0's lane k holds d[k] and
// e[k] (e of lane m - 1 and above ignored); on return sh.de[k].x holds the eigenvalues and the
// columns of Z the eigenvectors.
struct Tql2Pair {
    double2 de[64];       // {d, e} on return
    double2 rot[2][64];   // {cos, sin} of rotation i of the sweep published in each phase parity
    int desc[2][4];
};

__device__ inline bool tql2_pair(int m, double dv, double ev, Tql2Pair& sh, double* Z) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: scalar loop state
    if (wave == 0) {
        if (lane >= m) dv = 0.0;
        if (lane >= m - 1) ev = 0.0;
    } else {
        for (int col = 0; col < m; col++) Z[lane + col * 64] = lane == col ? 1.0 : 0.0;
    }
    int l = 0, iter = 0, state = 0;   // wave 0's QL position
    bool ok = true;
    for (int ph = 0;; ph++) {
        const int b = ph & 1;
        if (wave == 0) {
            int mm = 0, ilast = 0, has = 0;
            while (state == 0) {
                if (l >= m) {
                    state = 1;
                    break;
                }
                const double dnext = __shfl_down(dv, 1, 64);
                const bool small = lane >= l && lane < m - 1 &&
                                   fabs(ev) <= 1e-300 + 2.2e-16 * (fabs(dv) + fabs(dnext));
                const unsigned long long mask = __ballot(small);
                mm = mask ? (int)__builtin_ctzll(mask) : m - 1;
                if (mm == l) {
                    l++;
                    iter = 0;
                    continue;
                }
                if (iter++ == 200) {
                    state = 2;
                    break;
                }
                const double d0 = dv, e0 = ev;   // d and e before this sweep
                const double dl = tql2_rl(d0, l), el = tql2_rl(e0, l);
                const double dmm = tql2_rl(d0, mm);
                double g0 = (tql2_rl(d0, l + 1) - dl) * (0.5 * tql2_rcp(el));
                const double t0 = __builtin_fma(g0, g0, 1.0);
                const double r0 = t0 * tql2_rsqrt3(t0);
                g0 = dmm - dl + el * tql2_rcp(g0 + (g0 >= 0 ? r0 : -r0));
                double s = 1.0, c = 1.0, p = 0.0, g = g0, dip1 = dmm;
                double cr = 0.0, sr = 0.0;   // lane i: rotation i of this sweep
                bool underflow = false;
                ilast = l;
                double nd = tql2_rl(d0, mm - 1), ne = tql2_rl(e0, mm - 1);
                for (int i = mm - 1; i >= l; i--) {
                    const double ei = ne, di = nd;
                    const int ip = max(i - 1, l);
                    nd = tql2_rl(d0, ip);
                    ne = tql2_rl(e0, ip);
                    const double f = s * ei, bb = c * ei;
                    const double t = __builtin_fma(f, f, g * g);
                    const double y = tql2_rsqrt3(t);
                    const double r = t * y;
                    s = f * y;
                    c = g * y;
                    g = dip1 - p;
                    const double rr = (di - g) * s + 2.0 * c * bb;
                    p = s * rr;
                    if (lane == i + 1) {
                        dv = g + p;
                        ev = r;
                    }
                    if (lane == i) {
                        cr = c;
                        sr = s;
                    }
                    g = c * rr - bb;
                    dip1 = di;
                }
                // p and g are the same in every lane: a uniform branch
                if (__builtin_amdgcn_readfirstlane((int)!(fabs(g) + fabs(p) <= 1.79e308))) {
                    dv = d0;
                    ev = e0;
                    s = 1.0;
                    c = 1.0;
                    p = 0.0;
                    g = g0;
                    dip1 = dmm;
                    ilast = mm;
                    for (int i = mm - 1; i >= l; i--) {
                        const double ei = tql2_rl(e0, i), di = tql2_rl(d0, i);
                        const double f = s * ei, bb = c * ei;
                        const double t = __builtin_fma(f, f, g * g);
                        if (__builtin_amdgcn_readfirstlane((int)(t == 0.0))) {
                            // both zero: the rotation is the identity, deflate here
                            if (lane == i + 1) {
                                dv = dip1 - p;
                                ev = 0.0;
                            }
                            underflow = true;
                            break;
                        }
                        ilast = i;
                        const double y = tql2_rsqrt3(t);
                        const double r = t * y;
                        s = f * y;
                        c = g * y;
                        g = dip1 - p;
                        const double rr = (di - g) * s + 2.0 * c * bb;
                        p = s * rr;
                        if (lane == i + 1) {
                            dv = g + p;
                            ev = r;
                        }
                        if (lane == i) {
                            cr = c;
                            sr = s;
                        }
                        g = c * rr - bb;
                        dip1 = di;
                    }
                }
                if (lane >= ilast && lane < mm) sh.rot[b][lane] = make_double2(cr, sr);
                if (!underflow && lane == l) {
                    dv = dl - p;
                    ev = g;
                }
                if (lane == mm) ev = 0.0;
                has = 1;
                break;
            }
            if (state != 0) sh.de[lane] = make_double2(dv, ev);
            if (lane == 0) {
                sh.desc[b][0] = mm;
                sh.desc[b][1] = ilast;
                sh.desc[b][2] = has;
                sh.desc[b][3] = state;
            }
        } else if (ph > 0 && sh.desc[b ^ 1][2]) {
            // rotations mm - 1 .. ilast of the previous phase's sweep: column i + 1 carried in zc
            const int mm = sh.desc[b ^ 1][0], ilast = sh.desc[b ^ 1][1];
            const double2* R = sh.rot[b ^ 1];
            double zc = Z[lane + mm * 64];
            double za = Z[lane + max(mm - 1, ilast) * 64];
            for (int i = mm - 1; i >= ilast; i--) {
                const double zn = Z[lane + max(i - 1, ilast) * 64];
                const double2 cs = R[i];
                Z[lane + (i + 1) * 64] = cs.y * za + cs.x * zc;
                zc = cs.x * za - cs.y * zc;
                za = zn;
            }
            Z[lane + ilast * 64] = zc;
        }
        __syncthreads();
        if (sh.desc[b][3] != 0 && sh.desc[b][2] == 0) {
            ok = sh.desc[b][3] == 1;
            break;
        }
    }
    return ok;
}
