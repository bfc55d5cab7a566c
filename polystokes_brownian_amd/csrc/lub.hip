// lub.hip -- near-field (lubrication) resistance between colloid pairs from the two-sphere tables.
//
// New physics with no reference counterpart: the reference distributes the tables (lubdat/*.dat) but
// never reads them (src/saddle.cpp:53 leaves "add lubrication" as a TODO).  Each table value is the
// exact two-sphere resistance of equal spheres minus the inverse of the far-field (Rotne-Prager-Yamakawa
// FTS) two-sphere mobility, in units 6 pi eta = 1 and radius 1, in the convention of Kim & Karrila
// (1991, ch. 11): e is the unit vector from sphere 1 (the one whose force/torque/stresslet is given)
// to sphere 2, and r2bm.dat holds M11 + M12, the stresslet response to a strain shared by both spheres.
// Columns: r2babc X11A X12A Y11A Y12A Y11B Y12B X11C X12C Y11C Y12C; r2bgh X11G X12G Y11G Y12G Y11H
// Y12H; r2bm XM YM ZM.
//
// For every colloid pair closer than a cutoff the 19 functions are interpolated at the centre
// distance r by monotone piecewise-cubic Hermite interpolation (Fritsch & Carlson 1980) and the pair
// resistance is applied in FP64 with e = (x_j - x_i)/r for colloid i and partner j:
//   F_i = sum_j A11 U_i + A12 U_j + ~B11 W_i + ~B12 W_j + (~G11 - ~G12) : E,
//   T_i = sum_j B11 U_i + B12 U_j + C11 W_i + C12 W_j + (~H11 + ~H12) : E,
//   S_i = sum_j G11 U_i + G12 U_j + H11 W_i + H12 W_j + M : E,
// with A = X ee + Y (I - ee), B_ab = Y eps_abk e_k, C like A, G_abk = X (e_a e_b - d_ab/3) e_k +
// Y (e_a d_bk + e_b d_ak - 2 e_a e_b e_k), H_abk = Y (e_a eps_bkl e_l + e_b eps_akl e_l), M the
// d0/d1/d2 combination, E a strain rate shared by all colloids, and ~ the transposes that make the
// resistance symmetric (B21 = -B12 and G21 = -G12 are odd in e, H21 = H12 even).  This is the map
// (U - U_inf, W - W_inf, -E_inf) -> (F, T, S) exerted by the spheres on the fluid, or equally
// (U_inf - U, W_inf - W, E_inf) -> the hydrodynamic (F, T, S) on the spheres.  Below the first node
// (r < 2.1) the functions are held at their r = 2.1 values; at and beyond the cutoff (at most the last
// node, 4.0) they are zero; colloids with coincident centres (r = 0, no axis) are not paired.
//
// Vector layouts: uw = [U_c (3Nc); W_c (3Nc)]; E as the reference flattens a strain rate
// (E11 - E33, 2E12, 2E13, 2E23, E22 - E33, src/mob.cpp:300-305); fts = [F_c (3Nc); T_c (3Nc);
// S_c (5Nc)] with S flattened as (S11, S12, S13, S23, S22), the dual pairing of that flattening.
#include "ps_common.h"

#include <algorithm>
#include <cmath>

namespace {

constexpr int LT = PS_BLOCK;

// Monotone cubic Hermite slopes of one column (Fritsch & Carlson 1980; the end slopes by the
// shape-preserving three-point formula).
void pchip_slopes(int n, const double* x, const double* y, int stride, double* s) {
    if (n < 2) {
        if (n == 1) s[0] = 0.0;
        return;
    }
    std::vector<double> h(n - 1), d(n - 1);
    for (int k = 0; k + 1 < n; k++) {
        h[k] = x[k + 1] - x[k];
        d[k] = (y[(k + 1) * stride] - y[k * stride]) / h[k];
    }
    if (n == 2) {
        s[0] = s[stride] = d[0];
        return;
    }
    for (int k = 1; k + 1 < n; k++) {
        if (d[k - 1] * d[k] <= 0.0) {
            s[k * stride] = 0.0;
        } else {
            const double w1 = 2 * h[k] + h[k - 1], w2 = h[k] + 2 * h[k - 1];
            s[k * stride] = (w1 + w2) / (w1 / d[k - 1] + w2 / d[k]);
        }
    }
    auto endslope = [](double h0, double h1, double d0, double d1) {
        double v = ((2 * h0 + h1) * d0 - h0 * d1) / (h0 + h1);
        if (v * d0 <= 0.0) v = 0.0;
        else if (d0 * d1 <= 0.0 && std::fabs(v) > std::fabs(3 * d0)) v = 3 * d0;
        return v;
    };
    s[0] = endslope(h[0], h[1], d[0], d[1]);
    s[(n - 1) * stride] = endslope(h[n - 2], h[n - 3], d[n - 2], d[n - 3]);
}

}  // namespace

// One table on the device: n nodes r[], k columns f[n k] (row-major) and slopes s[n k].
struct LubTab {
    const double* r;
    const double* f;
    const double* s;
    int n, k;
};
struct LubSet {
    LubTab t[3];
    double rcut;
};

// Hermite evaluation of all columns of a table at r (clamped below the first node).
__host__ __device__ inline void lub_eval(const LubTab& T, double r, double* out) {
    if (T.n <= 0) {
        for (int c = 0; c < T.k; c++) out[c] = 0.0;
        return;
    }
    if (r <= T.r[0]) {
        for (int c = 0; c < T.k; c++) out[c] = T.f[c];
        return;
    }
    if (r >= T.r[T.n - 1]) {
        for (int c = 0; c < T.k; c++) out[c] = T.f[(T.n - 1) * T.k + c];
        return;
    }
    int lo = 0, hi = T.n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (T.r[mid] <= r) lo = mid;
        else hi = mid;
    }
    const double h = T.r[hi] - T.r[lo], t = (r - T.r[lo]) / h;
    const double t2 = t * t, t3 = t2 * t;
    const double h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t, h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
    for (int c = 0; c < T.k; c++)
        out[c] = h00 * T.f[lo * T.k + c] + h10 * h * T.s[lo * T.k + c] + h01 * T.f[hi * T.k + c] +
                 h11 * h * T.s[hi * T.k + c];
}

namespace {

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// Pair geometry with the reference's minimum image: d = x_j - x_i (from i to its partner j).
__device__ __forceinline__ double lub_geom(const SysConst& s, const double* x, int i, int j, double e[3]) {
    double d[3];
    for (int k = 0; k < 3; k++) {
        d[k] = x[3 * j + k] - x[3 * i + k];
        if (s.box_active) d[k] -= s.L[k] * rint(d[k] * s.Linv[k]);
    }
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; k++) e[k] = d[k] / r;
    return r;
}

// Colloid cell list.  Cells are at least rcut wide, so every partner of a colloid lies in its own
// cell or one of the 26 around it.  Periodic box: floor(L / rcut) cells per axis, wrapped (an axis of
// one or two cells visits each of its cells once).  No box: the colloids' bounding box, cells of side
// rcut.  Either way the cell side grows until the grid has at most `cap` cells.
struct LubGrid {
    double org[3], h[3];
    int dims[3];
    int periodic;
};

__global__ __launch_bounds__(LT) void k_lub_grid(SysConst s, const double* __restrict__ xm, double rcut, int cap,
                                                  LubGrid* __restrict__ g) {
    __shared__ double lo[3][LT], hi[3][LT];
    const int t = threadIdx.x;
    for (int k = 0; k < 3; k++) {
        lo[k][t] = 1e300;
        hi[k][t] = -1e300;
    }
    if (!s.box_active)
        for (int i = t; i < s.Nc; i += LT)
            for (int k = 0; k < 3; k++) {
                const double v = xm[3 * (s.Nm + i) + k];
                lo[k][t] = fmin(lo[k][t], v);
                hi[k][t] = fmax(hi[k][t], v);
            }
    __syncthreads();
    for (int w = LT / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 3; k++) {
                lo[k][t] = fmin(lo[k][t], lo[k][t + w]);
                hi[k][t] = fmax(hi[k][t], hi[k][t + w]);
            }
        __syncthreads();
    }
    if (t != 0) return;
    LubGrid G{};
    G.periodic = s.box_active;
    double ext[3];
    for (int k = 0; k < 3; k++) {
        ext[k] = s.box_active ? s.L[k] : fmax(hi[k][0] - lo[k][0], 0.0);
        G.org[k] = s.box_active ? 0.0 : lo[k][0];
        if (!(ext[k] < 1e300)) ext[k] = 0.0;   // non-finite positions: one cell on this axis
    }
    double side = rcut;
    for (;;) {
        long long tot = 1;
        for (int k = 0; k < 3; k++) {
            // periodic: whole cells of at least `side`; open: enough cells of `side` to cover the extent
            G.dims[k] = s.box_active ? max(1, (int)floor(ext[k] / side)) : (int)floor(ext[k] / side) + 1;
            tot *= G.dims[k];
        }
        if (tot <= cap) break;
        side *= 1.25;
    }
    for (int k = 0; k < 3; k++) G.h[k] = s.box_active ? ext[k] / G.dims[k] : side;
    *g = G;
}

__device__ __forceinline__ int lub_cell_axis(const LubGrid& G, const double* x, int k) {
    if (G.periodic) {
        const double f = x[k] / (G.h[k] * G.dims[k]);
        const int c = (int)floor((f - floor(f)) * G.dims[k]);
        return min(max(c, 0), G.dims[k] - 1);
    }
    return min(max((int)floor((x[k] - G.org[k]) / G.h[k]), 0), G.dims[k] - 1);
}

__global__ void k_lub_bin(SysConst s, const double* __restrict__ xm, const LubGrid* __restrict__ gp,
                          int* __restrict__ cell_of, int* __restrict__ cell_cnt) {
    const int i = blockIdx.x * LT + threadIdx.x;
    if (i >= s.Nc) return;
    const LubGrid G = *gp;
    const double* x = xm + 3 * (s.Nm + i);
    const int c = (lub_cell_axis(G, x, 2) * G.dims[1] + lub_cell_axis(G, x, 1)) * G.dims[0] + lub_cell_axis(G, x, 0);
    cell_of[i] = c;
    atomicAdd(&cell_cnt[c], 1);
}

// Exclusive scan of the cell counts into start[0..ncell] (one block), and a copy into the fill cursors.
__global__ __launch_bounds__(LT) void k_lub_scan(const LubGrid* __restrict__ gp, const int* __restrict__ cnt,
                                                 int* __restrict__ start, int* __restrict__ cursor) {
    __shared__ int part[LT];
    const LubGrid G = *gp;
    const int ncell = G.dims[0] * G.dims[1] * G.dims[2];
    const int t = threadIdx.x;
    const int per = (ncell + LT - 1) / LT;
    const int a = min(t * per, ncell), b = min(a + per, ncell);
    int sum = 0;
    for (int k = a; k < b; k++) sum += cnt[k];
    part[t] = sum;
    __syncthreads();
    if (t == 0) {
        int run = 0;
        for (int k = 0; k < LT; k++) {
            const int v = part[k];
            part[k] = run;
            run += v;
        }
        start[ncell] = run;
    }
    __syncthreads();
    int run = part[t];
    for (int k = a; k < b; k++) {
        start[k] = run;
        cursor[k] = run;
        run += cnt[k];
    }
}

__global__ void k_zero_cells(int* __restrict__ a, int n) {
    const int i = blockIdx.x * LT + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ void k_lub_scatter(int Nc, const int* __restrict__ cell_of, int* __restrict__ cursor,
                              int* __restrict__ mem) {
    const int i = blockIdx.x * LT + threadIdx.x;
    if (i >= Nc) return;
    mem[atomicAdd(&cursor[cell_of[i]], 1)] = i;   // order within a cell is irrelevant: partners are sorted
}

// Partners of colloid i (both directions) from the 27 cells around its own: count, then fill and sort
// ascending (the same list, in the same order, as a scan over every colloid).
template <bool FILL>
__global__ void k_lub_pairs(SysConst s, const double* __restrict__ xm, double rcut, const LubGrid* __restrict__ gp,
                            const int* __restrict__ cell_of, const int* __restrict__ start,
                            const int* __restrict__ mem, int* __restrict__ cnt_or_off, int* __restrict__ part) {
    const int i = blockIdx.x * LT + threadIdx.x;
    if (i >= s.Nc) return;
    const LubGrid G = *gp;
    const int ci = cell_of[i];
    const int cx = ci % G.dims[0], cy = (ci / G.dims[0]) % G.dims[1], cz = ci / (G.dims[0] * G.dims[1]);
    const int c0[3] = {cx, cy, cz};
    int lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        // periodic axes of 1 or 2 cells visit each cell once; open axes stop at the grid's edge
        if (G.periodic) {
            lo[k] = G.dims[k] >= 3 ? -1 : 0;
            hi[k] = G.dims[k] >= 3 ? 1 : G.dims[k] - 1;
        } else {
            lo[k] = c0[k] > 0 ? -1 : 0;
            hi[k] = c0[k] < G.dims[k] - 1 ? 1 : 0;
        }
    }
    int o = FILL ? cnt_or_off[i] : 0;
    const int o0 = o;
    double e[3];
    for (int dz = lo[2]; dz <= hi[2]; dz++)
        for (int dy = lo[1]; dy <= hi[1]; dy++)
            for (int dx = lo[0]; dx <= hi[0]; dx++) {
                int q[3] = {c0[0] + dx, c0[1] + dy, c0[2] + dz};
                if (G.periodic)
                    for (int k = 0; k < 3; k++) q[k] = (q[k] + G.dims[k]) % G.dims[k];
                const int cq = (q[2] * G.dims[1] + q[1]) * G.dims[0] + q[0];
                for (int t = start[cq]; t < start[cq + 1]; t++) {
                    const int j = mem[t];
                    if (j == i) continue;
                    const double r = lub_geom(s, xm, s.Nm + i, s.Nm + j, e);
                    if (r > 0.0 && r < rcut) {
                        if (FILL) part[o] = j;
                        o++;
                    }
                }
            }
    if (!FILL) {
        cnt_or_off[i] = o;
        return;
    }
    for (int a = o0 + 1; a < o; a++) {   // insertion sort of the (short) partner list
        const int v = part[a];
        int b = a - 1;
        while (b >= o0 && part[b] > v) {
            part[b + 1] = part[b];
            b--;
        }
        part[b + 1] = v;
    }
}

// Full symmetric tensor from the flattened strain (E11 - E33, 2E12, 2E13, 2E23, E22 - E33), traceless.
__device__ __forceinline__ void unflatten_E(const double* ef, double E[3][3]) {
    const double a = ef[0], b = ef[4];
    E[2][2] = -(a + b) / 3.0;
    E[0][0] = a + E[2][2];
    E[1][1] = b + E[2][2];
    E[0][1] = E[1][0] = 0.5 * ef[1];
    E[0][2] = E[2][0] = 0.5 * ef[2];
    E[1][2] = E[2][1] = 0.5 * ef[3];
}

// One thread per colloid: its rows (F_i, T_i, S_i) from its partners.
__global__ void k_lub_apply(SysConst s, LubSet L, const double* __restrict__ xm, const int* __restrict__ off,
                            const int* __restrict__ part, const double* __restrict__ uw,
                            const double* __restrict__ einf, double* __restrict__ fts) {
    const int i = blockIdx.x * LT + threadIdx.x;
    const int Nc = s.Nc;
    if (i >= Nc) return;
    double E[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const bool hasE = einf != nullptr;
    if (hasE) unflatten_E(einf, E);
    const double Ui[3] = {uw[3 * i], uw[3 * i + 1], uw[3 * i + 2]};
    const double Wi[3] = {uw[3 * Nc + 3 * i], uw[3 * Nc + 3 * i + 1], uw[3 * Nc + 3 * i + 2]};
    double F[3] = {0, 0, 0}, T[3] = {0, 0, 0}, S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int t = off[i]; t < off[i + 1]; t++) {
        const int j = part[t];
        double e[3];
        const double r = lub_geom(s, xm, s.Nm + i, s.Nm + j, e);
        if (!(r > 0.0 && r < L.rcut)) continue;
        double fa[10], fg[6], fm[3];
        lub_eval(L.t[0], r, fa);
        lub_eval(L.t[1], r, fg);
        lub_eval(L.t[2], r, fm);
        const double X11A = fa[0], X12A = fa[1], Y11A = fa[2], Y12A = fa[3], Y11B = fa[4], Y12B = fa[5];
        const double X11C = fa[6], X12C = fa[7], Y11C = fa[8], Y12C = fa[9];
        const double X11G = fg[0], X12G = fg[1], Y11G = fg[2], Y12G = fg[3], Y11H = fg[4], Y12H = fg[5];
        const double XM = fm[0], YM = fm[1], ZM = fm[2];
        const double Uj[3] = {uw[3 * j], uw[3 * j + 1], uw[3 * j + 2]};
        const double Wj[3] = {uw[3 * Nc + 3 * j], uw[3 * Nc + 3 * j + 1], uw[3 * Nc + 3 * j + 2]};
        const double eUi = e[0] * Ui[0] + e[1] * Ui[1] + e[2] * Ui[2];
        const double eUj = e[0] * Uj[0] + e[1] * Uj[1] + e[2] * Uj[2];
        const double eWi = e[0] * Wi[0] + e[1] * Wi[1] + e[2] * Wi[2];
        const double eWj = e[0] * Wj[0] + e[1] * Wj[1] + e[2] * Wj[2];
        double exWi[3], exWj[3], Uixe[3], Ujxe[3], Wixe[3], Wjxe[3];
        cross3(e, Wi, exWi);
        cross3(e, Wj, exWj);
        cross3(Ui, e, Uixe);
        cross3(Uj, e, Ujxe);
        cross3(Wi, e, Wixe);
        cross3(Wj, e, Wjxe);
        for (int a = 0; a < 3; a++) {
            // A U (both spheres), ~B W: (~B11 W)_a = Y11B (e x W)_a, (~B12 W)_a = -Y12B (e x W)_a
            F[a] += Y11A * Ui[a] + (X11A - Y11A) * eUi * e[a] + Y12A * Uj[a] + (X12A - Y12A) * eUj * e[a] +
                    Y11B * exWi[a] - Y12B * exWj[a];
            // B U: (B U)_a = Y (U x e)_a; C W
            T[a] += Y11B * Uixe[a] + Y12B * Ujxe[a] + Y11C * Wi[a] + (X11C - Y11C) * eWi * e[a] + Y12C * Wj[a] +
                    (X12C - Y12C) * eWj * e[a];
        }
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                const double eab = e[a] * e[b] - (a == b ? 1.0 / 3.0 : 0.0);
                S[a][b] += X11G * eab * eUi + Y11G * (e[a] * Ui[b] + e[b] * Ui[a] - 2 * e[a] * e[b] * eUi) +
                           X12G * eab * eUj + Y12G * (e[a] * Uj[b] + e[b] * Uj[a] - 2 * e[a] * e[b] * eUj) +
                           Y11H * (e[a] * Wixe[b] + e[b] * Wixe[a]) + Y12H * (e[a] * Wjxe[b] + e[b] * Wjxe[a]);
            }
        if (hasE) {
            double Ee[3];
            for (int a = 0; a < 3; a++) Ee[a] = E[a][0] * e[0] + E[a][1] * e[1] + E[a][2] * e[2];
            const double eEe = e[0] * Ee[0] + e[1] * Ee[1] + e[2] * Ee[2];
            double exEe[3];
            cross3(e, Ee, exEe);
            // (G:E)_a = X eEe e_a + Y (2 (Ee)_a - 2 eEe e_a); F gets (G11 - G12):E, T gets (H11 + H12):E
            for (int a = 0; a < 3; a++) {
                F[a] += (X11G - X12G) * eEe * e[a] + (Y11G - Y12G) * (2 * Ee[a] - 2 * eEe * e[a]);
                T[a] += 2 * (Y11H + Y12H) * exEe[a];
            }
            // M:E with d0, d1, d2
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    const double dab = a == b ? 1.0 : 0.0;
                    const double d0 = 1.5 * (e[a] * e[b] - dab / 3.0) * eEe;
                    const double d1 = e[a] * Ee[b] + e[b] * Ee[a] - 2 * e[a] * e[b] * eEe;
                    const double d2 = E[a][b] + 0.5 * dab * eEe - e[a] * Ee[b] - e[b] * Ee[a] + 0.5 * e[a] * e[b] * eEe;
                    S[a][b] += XM * d0 + YM * d1 + ZM * d2;
                }
        }
    }
    for (int a = 0; a < 3; a++) {
        fts[3 * i + a] = F[a];
        fts[3 * Nc + 3 * i + a] = T[a];
    }
    double* so = fts + 6 * Nc + 5 * i;
    so[0] = S[0][0];
    so[1] = S[0][1];
    so[2] = S[0][2];
    so[3] = S[1][2];
    so[4] = S[1][1];
}

}  // namespace

ps_err lub_set_table(Ctx* c, int which, const double* rows, int nrows, int ncols) {
    static const int want[3] = {10, 6, 3};
    if (which < 0 || which > 2 || !rows || nrows < 2 || ncols != want[which] + 1) {
        c->set_error("lub_set_table: which in 0..2 with 11, 7 or 4 columns (r first), at least 2 rows");
        return PS_ERR_ARG;
    }
    for (int k = 1; k < nrows; k++)
        if (!(rows[k * ncols] > rows[(k - 1) * ncols])) {
            c->set_error("lub_set_table: r must increase");
            return PS_ERR_ARG;
        }
    const int k = want[which];
    std::vector<double> buf((size_t)nrows * (1 + 2 * k));
    double* r = buf.data();
    double* f = r + nrows;
    double* s = f + (size_t)nrows * k;
    for (int i = 0; i < nrows; i++) {
        r[i] = rows[i * ncols];
        for (int q = 0; q < k; q++) f[i * k + q] = rows[i * ncols + 1 + q];
    }
    for (int q = 0; q < k; q++) pchip_slopes(nrows, r, f + q, k, s + q);
    if (c->lub_tab[which]) (void)hipFree(c->lub_tab[which]);
    c->lub_tab[which] = nullptr;
    PS_HIP(c, hipMalloc(&c->lub_tab[which], sizeof(double) * buf.size()));
    PS_HIP(c, hipMemcpy(c->lub_tab[which], buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
    c->lub_rows[which] = nrows;
    c->lub_rlast[which] = r[nrows - 1];
    return PS_OK;
}

// The colloid pairs closer than rcut at the positions of the last mob(), through the cell list.  One
// host round trip (the partner counts size the list).
ps_err lub_pairs(Ctx* c, double rcut, int64_t* npairs) {
    for (int w = 0; w < 3; w++)
        if (!c->lub_tab[w]) {
            c->set_error("lubrication: load the three tables with ps_lub_set_table first");
            return PS_ERR_STATE;
        }
    const double rmax = std::min(c->lub_rlast[0], std::min(c->lub_rlast[1], c->lub_rlast[2]));
    if (!(rcut > 0.0 && rcut <= rmax)) {
        c->set_error("lubrication: 0 < rcut <= the last tabulated distance");
        return PS_ERR_ARG;
    }
    const int Nc = c->Nc;
    const int cap = 2 * Nc + 64;   // cells
    hipStream_t st = c->stream;
    if (!c->lub_off) {
        PS_HIP(c, hipMalloc(&c->lub_off, sizeof(int) * (Nc + 1)));
        PS_HIP(c, hipMalloc(&c->lub_grid, sizeof(LubGrid)));
        PS_HIP(c, hipMalloc(&c->lub_cell_of, sizeof(int) * std::max(Nc, 1)));
        PS_HIP(c, hipMalloc(&c->lub_cell_mem, sizeof(int) * std::max(Nc, 1)));
        PS_HIP(c, hipMalloc(&c->lub_cells, sizeof(int) * (3 * (size_t)cap + 1)));
    }
    int* cell_cnt = c->lub_cells;
    int* cell_start = c->lub_cells + cap;
    int* cell_cur = c->lub_cells + 2 * cap + 1;
    std::vector<int> cnt(std::max(Nc, 1)), off(Nc + 1, 0);
    const unsigned g = (unsigned)((Nc + LT - 1) / LT);
    if (Nc > 0) {
        LubGrid* G = static_cast<LubGrid*>(c->lub_grid);
        k_lub_grid<<<1, LT, 0, st>>>(c->sc, c->xm, rcut, cap, G);
        PS_LAUNCH(c);
        k_zero_cells<<<(cap + LT - 1) / LT, LT, 0, st>>>(cell_cnt, cap);
        PS_LAUNCH(c);
        k_lub_bin<<<g, LT, 0, st>>>(c->sc, c->xm, G, c->lub_cell_of, cell_cnt);
        PS_LAUNCH(c);
        k_lub_scan<<<1, LT, 0, st>>>(G, cell_cnt, cell_start, cell_cur);
        PS_LAUNCH(c);
        k_lub_scatter<<<g, LT, 0, st>>>(Nc, c->lub_cell_of, cell_cur, c->lub_cell_mem);
        PS_LAUNCH(c);
        k_lub_pairs<false><<<g, LT, 0, st>>>(c->sc, c->xm, rcut, G, c->lub_cell_of, cell_start, c->lub_cell_mem,
                                               c->lub_off, nullptr);
        PS_LAUNCH(c);
        PS_HIP(c, hipMemcpyAsync(cnt.data(), c->lub_off, sizeof(int) * Nc, hipMemcpyDeviceToHost, st));
        PS_HIP(c, hipStreamSynchronize(st));
    }
    for (int i = 0; i < Nc; i++) off[i + 1] = off[i] + cnt[i];
    const int total = off[Nc];
    if (total > c->lub_cap) {
        if (c->lub_part) (void)hipFree(c->lub_part);
        c->lub_part = nullptr;
        c->lub_cap = std::max(total, 16);
        PS_HIP(c, hipMalloc(&c->lub_part, sizeof(int) * c->lub_cap));
    }
    PS_HIP(c, hipMemcpyAsync(c->lub_off, off.data(), sizeof(int) * (Nc + 1), hipMemcpyHostToDevice, st));
    if (total > 0) {
        k_lub_pairs<true><<<g, LT, 0, st>>>(c->sc, c->xm, rcut, static_cast<LubGrid*>(c->lub_grid), c->lub_cell_of,
                                              cell_start, c->lub_cell_mem, c->lub_off, c->lub_part);
        PS_LAUNCH(c);
    }
    // off.data() is read by the copy above before the vector goes out of scope
    PS_HIP(c, hipStreamSynchronize(st));
    c->lub_rcut = rcut;
    c->lub_built = true;
    if (npairs) *npairs = total / 2;
    return PS_OK;
}

ps_err lub_build(Ctx* c, double rcut, int64_t* npairs) {
    if (!c->have_mob) {
        c->set_error("lub_build: call ps_mob first (the pairs use the mob() positions)");
        return PS_ERR_STATE;
    }
    return lub_pairs(c, rcut, npairs);
}

ps_err lub_apply_dev(Ctx* c, const double* uw, const double* einf, double* fts) {
    if (!c->lub_built) {
        c->set_error("lub_apply: call ps_lub_build first");
        return PS_ERR_STATE;
    }
    if (c->Nc == 0) return PS_OK;
    LubSet L{};
    static const int k[3] = {10, 6, 3};
    for (int w = 0; w < 3; w++) {
        const int n = c->lub_rows[w];
        const double* b = c->lub_tab[w];
        L.t[w] = LubTab{b, b + n, b + n + (size_t)n * k[w], n, k[w]};
    }
    L.rcut = c->lub_rcut;
    const unsigned g = (unsigned)((c->Nc + LT - 1) / LT);
    k_lub_apply<<<g, LT, 0, c->stream>>>(c->sc, L, c->xm, c->lub_off, c->lub_part, uw, einf, fts);
    PS_LAUNCH(c);
    return PS_OK;
}

ps_err lub_alloc(Ctx* c) {
    // colloid saddle of the lubricated block solve: b, x, inverse diagonal (17 Nc each), the lubrication
    // output and the colloid-source partial rows (11 Nc each)
    const size_t nc = (size_t)std::max(c->Nc, 1);
    PS_HIP(c, hipMalloc(&c->cs_buf, sizeof(double) * 73 * nc));
    return PS_OK;
}

namespace {
// y[n + nm3 + k] -= fts[k], k < 6 Nc: the lower-right block -R_lub acting on the colloid velocities
__global__ void k_lub_sub(int64_t nc6, const double* __restrict__ fts, double* __restrict__ y) {
    const int64_t k = (int64_t)blockIdx.x * LT + threadIdx.x;
    if (k < nc6) y[k] -= fts[k];
}
}  // namespace

ps_err lub_saddle_term(Ctx* c, const double* x, double* y) {
    if (!c->lub_built) {
        c->set_error("lubrication: no pair list (ps_mob builds it)");
        return PS_ERR_STATE;
    }
    const int64_t off = c->n + c->sc.nm3, nc6 = 6 * (int64_t)c->Nc;
    double* fts = c->cs_buf + 51 * (size_t)std::max(c->Nc, 1);
    PS_TRY(lub_apply_dev(c, x + off, nullptr, fts));
    if (nc6 > 0) {
        k_lub_sub<<<(unsigned)((nc6 + LT - 1) / LT), LT, 0, c->stream>>>(nc6, fts, y + off);
        PS_LAUNCH(c);
    }
    return PS_OK;
}

void lub_free(Ctx* c) {
    for (int w = 0; w < 3; w++) {
        if (c->lub_tab[w]) (void)hipFree(c->lub_tab[w]);
        c->lub_tab[w] = nullptr;
    }
    for (void* q : {(void*)c->lub_off, (void*)c->lub_part, c->lub_grid, (void*)c->lub_cell_of, (void*)c->lub_cell_mem,
                    (void*)c->lub_cells, (void*)c->cs_buf})
        if (q) (void)hipFree(q);
    c->lub_off = c->lub_part = c->lub_cell_of = c->lub_cell_mem = c->lub_cells = nullptr;
    c->lub_grid = nullptr;
    c->cs_buf = nullptr;
    c->lub_built = false;
}

// Context-free evaluation of one table (rows with r first) at r: the interpolant the device uses.
ps_err lub_interp_host(const double* rows, int nrows, int ncols, double r, double* out) {
    if (!rows || !out || nrows < 2 || ncols < 2) return PS_ERR_ARG;
    const int k = ncols - 1;
    std::vector<double> rr(nrows), f((size_t)nrows * k), s((size_t)nrows * k);
    for (int i = 0; i < nrows; i++) {
        rr[i] = rows[i * ncols];
        for (int q = 0; q < k; q++) f[i * k + q] = rows[i * ncols + 1 + q];
    }
    for (int q = 0; q < k; q++) pchip_slopes(nrows, rr.data(), f.data() + q, k, s.data() + q);
    const LubTab T{rr.data(), f.data(), s.data(), nrows, k};
    lub_eval(T, r, out);
    return PS_OK;
}

ps_err lub_slopes_host(const double* rows, int nrows, int ncols, double* slopes) {
    if (!rows || !slopes || nrows < 2 || ncols < 2) return PS_ERR_ARG;
    const int k = ncols - 1;
    std::vector<double> rr(nrows), f((size_t)nrows * k);
    for (int i = 0; i < nrows; i++) {
        rr[i] = rows[i * ncols];
        for (int q = 0; q < k; q++) f[i * k + q] = rows[i * ncols + 1 + q];
    }
    for (int q = 0; q < k; q++) pchip_slopes(nrows, rr.data(), f.data() + q, k, slopes + q);
    return PS_OK;
}
