// tiles.h -- FP64 tile algebra of the grand mobility shared by the apply kernels and the dense
// build: the coefficients (src/Stokes.cpp:43-65), the Levi-Civita / Kronecker helpers and index
// fix-up of src/mob.cpp, the colloid-colloid FTS tile, the monomer-colloid couplings and the
// reference's pair geometry.
#pragma once
#include "ps_common.h"

namespace {

constexpr double C1D2 = 1. / 2., C1D3 = 1. / 3., C3D2 = 3. / 2., C3D4 = 3. / 4., C3D8 = 3. / 8.,
                 C5D6 = 5. / 6., C9D4 = 9. / 4., C9D5 = 9. / 5., C6D5 = 6. / 5., C9D8 = 9. / 8.,
                 C9D2 = 9. / 2., C9D10 = 9. / 10., C9D20 = 9. / 20., C9D32 = 9. / 32.,
                 C18D5 = 18. / 5., C36D5 = 36. / 5., C54D5 = 54. / 5.;

__device__ __forceinline__ double eps3(int i, int j, int k) {
    if (i == j || j == k || i == k) return 0.0;
    if ((i == 0 && j == 1 && k == 2) || (i == 1 && j == 2 && k == 0) || (i == 2 && j == 0 && k == 1))
        return 1.0;
    return -1.0;
}
__device__ __forceinline__ double del(int i, int j) { return i == j ? 1.0 : 0.0; }
__device__ __forceinline__ int third(int a, int b) {
    int k = 3 - a - b;
    if (k == -1) k = 2;
    if (k == 3) k = 0;
    return k;
}

struct BBT {
    double a[3][3], b[3][3], bt[3][3], c[3][3], gt[3][5], ht[3][5], m[5][5];
};

// The full colloid-colloid FTS tile of mobility(..., AB=false, AA=false), src/mob.cpp:96-356.
__device__ void tile_bb(double dr_inv, const double e[3], BBT& T) {
    double ee[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) ee[i][j] = e[i] * e[j];
    const double dr_inv2 = dr_inv * dr_inv, dr_inv3 = dr_inv2 * dr_inv, dr_inv4 = dr_inv3 * dr_inv,
                 dr_inv5 = dr_inv4 * dr_inv;
    const double x12a = C3D2 * dr_inv - dr_inv3, y12a = C3D4 * dr_inv + C1D2 * dr_inv3;
    const double y12b = -C3D4 * dr_inv2;
    const double x12g = C9D4 * dr_inv2 - C18D5 * dr_inv4, y12g = C6D5 * dr_inv4;
    const double x12c = C3D4 * dr_inv3, y12c = -C3D8 * dr_inv3;
    const double y12h = -C9D8 * dr_inv3;
    const double x12m = -C9D2 * dr_inv3 + C54D5 * dr_inv5, y12m = C9D4 * dr_inv3 - C36D5 * dr_inv5,
                 z12m = C9D5 * dr_inv5;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            T.a[i][j] = (i == j) ? x12a * ee[i][i] + y12a * (1. - ee[i][i]) : x12a * ee[i][j] - y12a * ee[i][j];
            T.c[i][j] = (i == j) ? x12c * ee[i][i] + y12c * (1. - ee[i][i]) : x12c * ee[i][j] + y12c * ee[i][j];
            const int k = third(i, j);
            T.b[i][j] = y12b * eps3(i, j, k) * e[k];
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) T.bt[j][i] = T.b[i][j];
    double gt[3][3][3], ht[3][3][3];
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                gt[k][i][j] = -x12g * (ee[i][j] - C1D3 * del(i, j)) * e[k] +
                              y12g * (e[i] * del(j, k) + e[j] * del(i, k) - 2. * ee[i][j] * e[k]);
                const int ll = third(j, k), mm = third(i, k);
                ht[k][i][j] = y12h * (ee[i][ll] * eps3(j, k, ll) + ee[j][mm] * eps3(i, k, mm));
            }
    for (int i = 0; i < 3; i++) {
        T.gt[i][0] = gt[i][0][0] - gt[i][2][2];
        T.gt[i][1] = 2. * gt[i][0][1];
        T.gt[i][2] = 2. * gt[i][0][2];
        T.gt[i][3] = 2. * gt[i][1][2];
        T.gt[i][4] = gt[i][1][1] - gt[i][2][2];
        T.ht[i][0] = ht[i][0][0] - ht[i][2][2];
        T.ht[i][1] = 2. * ht[i][0][1];
        T.ht[i][2] = 2. * ht[i][0][2];
        T.ht[i][3] = 2. * ht[i][1][2];
        T.ht[i][4] = ht[i][1][1] - ht[i][2][2];
    }
    // m_ijkl contracted through mesid (src/mob.cpp:278-355).
    const int MA[5] = {0, 0, 0, 1, 1}, MB[5] = {2, 1, 2, 2, 2};
    auto mt = [&](int i, int j, int k, int l) {
        return C3D2 * x12m * (ee[i][j] - C1D3 * del(i, j)) * (ee[k][l] - C1D3 * del(k, l)) +
               C1D2 * y12m *
                   (ee[i][k] * del(j, l) + ee[j][k] * del(i, l) + ee[i][l] * del(j, k) +
                    ee[j][l] * del(i, k) - 4. * ee[i][j] * ee[k][l]) +
               C1D2 * z12m *
                   (del(i, k) * del(j, l) + del(j, k) * del(i, l) - del(i, j) * del(k, l) +
                    ee[i][j] * del(k, l) + ee[k][l] * del(i, j) - ee[i][k] * del(j, l) -
                    ee[j][k] * del(i, l) - ee[i][l] * del(j, k) - ee[j][l] * del(i, k) + ee[i][j] * ee[k][l]);
    };
    for (int r = 0; r < 5; r++) {
        const int a = MA[r], b = MB[r];
        if (r == 0 || r == 4) {
            T.m[r][0] = mt(a, a, 0, 0) - mt(a, a, 2, 2) - (mt(b, b, 0, 0) - mt(b, b, 2, 2));
            T.m[r][1] = 2. * (mt(a, a, 0, 1) - mt(b, b, 0, 1));
            T.m[r][2] = 2. * (mt(a, a, 0, 2) - mt(b, b, 0, 2));
            T.m[r][3] = 2. * (mt(a, a, 1, 2) - mt(b, b, 1, 2));
            T.m[r][4] = mt(a, a, 1, 1) - mt(a, a, 2, 2) - (mt(b, b, 1, 1) - mt(b, b, 2, 2));
        } else {
            T.m[r][0] = 2. * (mt(a, b, 0, 0) - mt(a, b, 2, 2));
            T.m[r][1] = 4. * mt(a, b, 0, 1);
            T.m[r][2] = 4. * mt(a, b, 0, 2);
            T.m[r][3] = 4. * mt(a, b, 1, 2);
            T.m[r][4] = 2. * (mt(a, b, 1, 1) - mt(a, b, 2, 2));
        }
    }
}

__device__ __forceinline__ void pair_geom_dev(const SysConst& s, const double* xs, int k, int j, double e[3],
                                              double& dr, double& dr_inv) {
#pragma clang fp contract(off)
    double dx = xs[3 * k] - xs[3 * j];
    double dy = xs[3 * k + 1] - xs[3 * j + 1];
    double dz = xs[3 * k + 2] - xs[3 * j + 2];
    if (s.box_active) {
        dx -= s.L[0] * __builtin_rint(dx * s.Linv[0]);
        dy -= s.L[1] * __builtin_rint(dy * s.Linv[1]);
        dz -= s.L[2] * __builtin_rint(dz * s.Linv[2]);
    }
    dr = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
    dr_inv = 1 / dr;
    e[0] = dx * dr_inv;
    e[1] = dy * dr_inv;
    e[2] = dz * dr_inv;
}


// Monomer-colloid couplings of mobility_AB (src/mob.cpp:362-537; mob_a is not written there, the
// translation block is the carried tile): b_ij = y12b eps_ijk e_k, bt = b^T, and the stresslet
// coupling gt (rank 3) flattened onto 3 x 5 (src/mob.cpp:198-227).  beta2 = beta^2.
__device__ void tile_ab_dev(double dr_inv, const double e[3], double beta2, double b[3][3], double bt[3][3],
                            double gtf[3][5]) {
    double ee[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) ee[i][j] = e[i] * e[j];
    const double dr_inv2 = dr_inv * dr_inv, dr_inv4 = dr_inv2 * dr_inv2;
    const double y12b = -C3D4 * dr_inv2;
    const double x12g = C9D4 * dr_inv2 - C9D20 * dr_inv4 * (3 + 5 * beta2);
    const double y12g = C9D10 * dr_inv4 * (C1D2 + C5D6 * beta2);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const int k = third(i, j);
            b[i][j] = y12b * eps3(i, j, k) * e[k];
            bt[j][i] = b[i][j];
        }
    double gt[3][3][3];
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                gt[k][i][j] = -x12g * (ee[i][j] - C1D3 * del(i, j)) * e[k] +
                              y12g * (e[i] * del(j, k) + e[j] * del(i, k) - 2. * ee[i][j] * e[k]);
    for (int i = 0; i < 3; i++) {
        gtf[i][0] = gt[i][0][0] - gt[i][2][2];
        gtf[i][1] = 2. * gt[i][0][1];
        gtf[i][2] = 2. * gt[i][0][2];
        gtf[i][3] = 2. * gt[i][1][2];
        gtf[i][4] = gt[i][1][1] - gt[i][2][2];
    }
}

// Monomer-monomer translation tile of mobility(..., AA=true) (src/mob.cpp:96-151), from the pair
// geometry: x ee + y (I - ee) with the reference's non-standard coefficients and near branch.
__device__ void tile_aa_dev(double dr, double dr_inv, const double e[3], double beta, double a[3][3]) {
    const double dr_inv3 = dr_inv * dr_inv * dr_inv;
    double x12a, y12a;
    if (dr > 2.0 * beta) {
        x12a = C3D8 * dr_inv - 2.0 * dr_inv3;
        y12a = C3D4 * dr_inv - C1D2 * dr_inv3;
    } else {
        x12a = C3D8 * dr - 1.0;
        y12a = 1.0 - C9D32 * dr;
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const double ee = e[i] * e[j];
            a[i][j] = (i == j) ? x12a * ee + y12a * (1. - ee) : x12a * ee - y12a * ee;
        }
}

// Colloid-colloid coupling in contracted form: the FTS tile of mobility(..., AB=false, AA=false)
// (src/mob.cpp:96-356) applied to one source's [F T S] without forming the 11 x 11 tile, for
// e = (x_lo - x_hi) / r of the pair's lower and higher colloid index.  With S the symmetric
// traceless matrix of the stresslet (S_22 = -S_0 - S_4), a = S e, s = e.S e and w = T x e, the
// role expressions of src/mob.cpp:1042-1143 reduce to (sg = +1 when the target is the lower index,
// -1 when it is the higher; flat(X) = [X00 - X22, 2X01, 2X02, 2X12, X11 - X22]):
//   U += (x12a - y12a)(e.F) e + y12a F - y12b w + sg [(-x12g - 2 y12g) s e + 2 y12g a]
//   W += y12b (F x e) + (x12c + y12c)(e.T) e + y12c (1 - 2 e_i^2) T_i + 2 y12h (e x a)
//   Sr += -sg flat(p e e^T + y12g (e F^T + F e^T)) + flat(y12h (e w^T + w e^T))
//         + flat(A e e^T + B (e a^T + a e^T) + z12m S),
// p = -(x12g + 2 y12g)(e.F), A = s (3/2 x12m - 2 y12m + 1/2 z12m), B = y12m - z12m (the '+' of
// mob_c is the (x12c + y12c) off-diagonal term).
__device__ __forceinline__ void cc_mS(const double e[3], double ri3, double ri5, const double S[5], double out[5]) {
    const double x12m = -C9D2 * ri3 + C54D5 * ri5, y12m = C9D4 * ri3 - C36D5 * ri5, z12m = C9D5 * ri5;
    const double S22 = -S[0] - S[4];
    const double a0 = S[0] * e[0] + S[1] * e[1] + S[2] * e[2];
    const double a1 = S[1] * e[0] + S[4] * e[1] + S[3] * e[2];
    const double a2 = S[2] * e[0] + S[3] * e[1] + S22 * e[2];
    const double s = e[0] * a0 + e[1] * a1 + e[2] * a2;
    const double A = s * (C3D2 * x12m - 2.0 * y12m + C1D2 * z12m), Bc = y12m - z12m;
    out[0] = A * (e[0] * e[0] - e[2] * e[2]) + 2.0 * Bc * (e[0] * a0 - e[2] * a2) + z12m * (S[0] - S22);
    out[1] = 2.0 * (A * e[0] * e[1] + Bc * (e[0] * a1 + e[1] * a0) + z12m * S[1]);
    out[2] = 2.0 * (A * e[0] * e[2] + Bc * (e[0] * a2 + e[2] * a0) + z12m * S[2]);
    out[3] = 2.0 * (A * e[1] * e[2] + Bc * (e[1] * a2 + e[2] * a1) + z12m * S[3]);
    out[4] = A * (e[1] * e[1] - e[2] * e[2]) + 2.0 * Bc * (e[1] * a1 - e[2] * a2) + z12m * (S[4] - S22);
}
template <bool SROWS>
__device__ __forceinline__ void cc_couple(const double e[3], double ri, double sg, const double F[3], const double Tq[3],
                                          const double S[5], double U[3], double W[3], double Sr[5]) {
    const double ri2 = ri * ri, ri3 = ri2 * ri, ri4 = ri2 * ri2, ri5 = ri4 * ri;
    const double y12g = C6D5 * ri4, x12g = C9D4 * ri2 - C18D5 * ri4, y12h = -C9D8 * ri3, y12b = -C3D4 * ri2;
    const double S22 = -S[0] - S[4];
    const double a0 = S[0] * e[0] + S[1] * e[1] + S[2] * e[2];
    const double a1 = S[1] * e[0] + S[4] * e[1] + S[3] * e[2];
    const double a2 = S[2] * e[0] + S[3] * e[1] + S22 * e[2];
    const double w0 = Tq[1] * e[2] - Tq[2] * e[1], w1 = Tq[2] * e[0] - Tq[0] * e[2], w2 = Tq[0] * e[1] - Tq[1] * e[0];
    const double eF = e[0] * F[0] + e[1] * F[1] + e[2] * F[2];
    if (!SROWS) {
        const double x12a = C3D2 * ri - ri3, y12a = C3D4 * ri + C1D2 * ri3;
        const double x12c = C3D4 * ri3, y12c = -C3D8 * ri3;
        const double s = e[0] * a0 + e[1] * a1 + e[2] * a2;
        const double gs = sg * (-x12g * s - 2.0 * y12g * s), g2 = sg * 2.0 * y12g;
        const double ca = (x12a - y12a) * eF;
        U[0] += ca * e[0] + y12a * F[0] - y12b * w0 + (gs * e[0] + g2 * a0);
        U[1] += ca * e[1] + y12a * F[1] - y12b * w1 + (gs * e[1] + g2 * a1);
        U[2] += ca * e[2] + y12a * F[2] - y12b * w2 + (gs * e[2] + g2 * a2);
        const double eT = e[0] * Tq[0] + e[1] * Tq[1] + e[2] * Tq[2];
        const double cc = (x12c + y12c) * eT, h2 = 2.0 * y12h;
        W[0] += y12b * (F[1] * e[2] - F[2] * e[1]) + cc * e[0] + y12c * (1.0 - 2.0 * e[0] * e[0]) * Tq[0] +
                h2 * (e[1] * a2 - e[2] * a1);
        W[1] += y12b * (F[2] * e[0] - F[0] * e[2]) + cc * e[1] + y12c * (1.0 - 2.0 * e[1] * e[1]) * Tq[1] +
                h2 * (e[2] * a0 - e[0] * a2);
        W[2] += y12b * (F[0] * e[1] - F[1] * e[0]) + cc * e[2] + y12c * (1.0 - 2.0 * e[2] * e[2]) * Tq[2] +
                h2 * (e[0] * a1 - e[1] * a0);
    }
    const double p = -(x12g + 2.0 * y12g) * eF, ms = -sg, h = 2.0 * y12h;
    double m[5];
    cc_mS(e, ri3, ri5, S, m);
    Sr[0] += ms * (p * (e[0] * e[0] - e[2] * e[2]) + 2.0 * y12g * (e[0] * F[0] - e[2] * F[2])) +
             h * (e[0] * w0 - e[2] * w2) + m[0];
    Sr[1] += ms * (2.0 * (p * e[0] * e[1] + y12g * (e[0] * F[1] + e[1] * F[0]))) + h * (e[0] * w1 + e[1] * w0) + m[1];
    Sr[2] += ms * (2.0 * (p * e[0] * e[2] + y12g * (e[0] * F[2] + e[2] * F[0]))) + h * (e[0] * w2 + e[2] * w0) + m[2];
    Sr[3] += ms * (2.0 * (p * e[1] * e[2] + y12g * (e[1] * F[2] + e[2] * F[1]))) + h * (e[1] * w2 + e[2] * w1) + m[3];
    Sr[4] += ms * (p * (e[1] * e[1] - e[2] * e[2]) + 2.0 * y12g * (e[1] * F[1] - e[2] * F[2])) +
             h * (e[1] * w1 - e[2] * w2) + m[4];
}

}  // namespace
