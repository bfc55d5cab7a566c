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

}  // namespace
