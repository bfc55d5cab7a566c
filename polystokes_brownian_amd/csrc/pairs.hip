// pairs.hip -- check_dist (Verlet pairs) and RHS (forces) on the device.
//
// check_dist (src/check_dist.cpp:12-87) visits all Np(Np-1)/2 pairs; here a periodic cell list of
// edge >= max(rverls) yields the candidates, and every candidate is tested with exactly the
// reference's FP expression sequence (d = x_k - x_j with k < j, minimum image by nearbyint,
// dr = sqrt(dx*dx + dy*dy + dz*dz) without contraction, dr < (float)rverl), so the pair set is
// bit-identical to the O(N^2) scan.  The list is kept in both directions (row p holds every partner
// q != p) sorted by partner, so forces reduce per particle with no atomics; the reference's
// vlist[k] is the q > k half, in the same order (pair index order, src/init.cpp:225-262).
// Without a periodic box, or with fewer than 3 cells per axis, the candidates are all particles.
#include "ps_common.h"

#include <algorithm>
#include <cmath>

#pragma clang fp contract(off)

namespace {

constexpr int TB = PS_BLOCK;

struct PairConst {
    SysConst s;
    double rverl[3];
    double sig[3];
    int ncell[3];
    double cell_size[3], cell_org[3];
    int mode;   // 0: all-pairs candidates, 1: periodic cell list
};

__device__ __forceinline__ int ptype(int Nm, int k, int j) {
    if (k < Nm && j < Nm) return 0;
    if (k < Nm) return 1;
    return 2;
}

// The reference pair geometry with k < j (src/check_dist.cpp:49-71).
__device__ __forceinline__ void geom(const SysConst& s, const double* xs, int k, int j, double& dx, double& dy,
                                     double& dz, double& dr) {
    dx = xs[3 * k] - xs[3 * j];
    dy = xs[3 * k + 1] - xs[3 * j + 1];
    dz = xs[3 * k + 2] - xs[3 * j + 2];
    if (s.box_active) {
        dx -= s.L[0] * __builtin_rint(dx * s.Linv[0]);
        dy -= s.L[1] * __builtin_rint(dy * s.Linv[1]);
        dz -= s.L[2] * __builtin_rint(dz * s.Linv[2]);
    }
    dr = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
}

__device__ __forceinline__ int cell_coord(const PairConst& pc, double x, int d) {
    // wrap into [org, org + L) for binning only; the pair test uses the raw positions
    const double L = pc.s.L[d];
    double xw = x - L * floor((x - pc.cell_org[d]) / L);
    int ci = (int)floor((xw - pc.cell_org[d]) / pc.cell_size[d]);
    if (ci < 0) ci = 0;
    if (ci >= pc.ncell[d]) ci = pc.ncell[d] - 1;
    return ci;
}

// Bin particle p by the positions x, which become this check_dist's geometry snapshot xpd.
__global__ void k_bin(PairConst pc, const double* __restrict__ x, double* __restrict__ xpd,
                      double* __restrict__ xsave, int* __restrict__ cell_of, int* __restrict__ cell_count) {
    const int p = blockIdx.x * TB + threadIdx.x;
    if (p >= pc.s.Np) return;
    const double px = x[3 * p], py = x[3 * p + 1], pz = x[3 * p + 2];
    xpd[3 * p] = px;
    xpd[3 * p + 1] = py;
    xpd[3 * p + 2] = pz;
    if (xsave) {
        xsave[3 * p] = px;
        xsave[3 * p + 1] = py;
        xsave[3 * p + 2] = pz;
    }
    const int cx = cell_coord(pc, px, 0), cy = cell_coord(pc, py, 1), cz = cell_coord(pc, pz, 2);
    const int cell = cx + pc.ncell[0] * (cy + pc.ncell[1] * cz);
    cell_of[p] = cell;
    atomicAdd(&cell_count[cell], 1);
}

// Scatter particles into cell order: perm[slot] = p and a cell-ordered copy of the positions, so a
// cell's candidates are read contiguously.  The cell counts are the cursors, counted down, so they
// are all zero again afterwards, ready for the next binning.  The in-cell order depends on atomics;
// rows are sorted by partner afterwards, so the pair lists do not.
__global__ void k_scatter(int Np, const double* __restrict__ xs, const int* __restrict__ cell_of,
                          const int* __restrict__ cell_start, int* __restrict__ cursor, int* __restrict__ perm,
                          double* __restrict__ xsort) {
    const int p = blockIdx.x * TB + threadIdx.x;
    if (p >= Np) return;
    const int c = cell_of[p];
    const int slot = cell_start[c] + atomicSub(&cursor[c], 1) - 1;
    perm[slot] = p;
    xsort[3 * slot] = xs[3 * p];
    xsort[3 * slot + 1] = xs[3 * p + 1];
    xsort[3 * slot + 2] = xs[3 * p + 2];
}

// Single-block exclusive scan of n ints: out[0..n] with out[n] = total, and out[n + 1] = max(in).
// The 1024 threads walk the array in tiles of 4096: each thread loads four consecutive ints (one
// 16-byte load, coalesced across the block), the thread sums are scanned within each wave by
// shuffles and across the 16 waves through LDS, and the next tile's load is in flight meanwhile.
// With `clear` the input is zeroed once read (the cell counts double as the scatter's cursors).
constexpr int SCAN_T = 1024, SCAN_TILE = 4 * SCAN_T;

__device__ __forceinline__ int4 scan_load4(const int* __restrict__ in, int n, int i0) {
    if (i0 + 3 < n) return *reinterpret_cast<const int4*>(in + i0);
    int4 v = {0, 0, 0, 0};
    if (i0 < n) v.x = in[i0];
    if (i0 + 1 < n) v.y = in[i0 + 1];
    if (i0 + 2 < n) v.z = in[i0 + 2];
    return v;
}

__global__ __launch_bounds__(SCAN_T) void k_scan(int* __restrict__ in, int* __restrict__ out, int n, int clear) {
    __shared__ int wsum[16], wmax[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int carry = 0, mx = 0;
    int4 cur = scan_load4(in, n, 4 * t);
    for (int base = 0; base < n; base += SCAN_TILE) {
        const int i0 = base + 4 * t;
        const int4 nxt = base + SCAN_TILE < n ? scan_load4(in, n, i0 + SCAN_TILE) : int4{0, 0, 0, 0};
        const int s = (cur.x + cur.y) + (cur.z + cur.w);
        mx = max(mx, max(max(cur.x, cur.y), max(cur.z, cur.w)));
        int inc = s;   // inclusive scan of the thread sums within the wave
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        int before = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int v = wsum[k];
            before += k < w ? v : 0;
            tot += v;
        }
        int4 e;
        e.x = carry + before + inc - s;
        e.y = e.x + cur.x;
        e.z = e.y + cur.y;
        e.w = e.z + cur.z;
        if (i0 + 3 < n) {
            *reinterpret_cast<int4*>(out + i0) = e;
            if (clear) *reinterpret_cast<int4*>(in + i0) = int4{0, 0, 0, 0};
        } else {
            const int ev[4] = {e.x, e.y, e.z, e.w};
            for (int k = 0; k < 4; k++)
                if (i0 + k < n) {
                    out[i0 + k] = ev[k];
                    if (clear) in[i0 + k] = 0;
                }
        }
        carry += tot;
        __syncthreads();   // wsum is rewritten by the next tile
        cur = nxt;
    }
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) wmax[w] = mx;
    __syncthreads();
    if (t == 0) {
        out[n] = carry;
        int m = 0;
        for (int k = 0; k < 16; k++) m = max(m, wmax[k]);
        out[n + 1] = m;
    }
}

// A group of GL lanes serves one particle.  In cell mode lane l < 27 finds the range of stencil cell
// l in the cell-sorted order (the 27 cells are distinct when every axis has >= 3 cells); the group
// then walks the concatenation of the 27 ranges, candidate k by lane k mod GL, so that the lanes
// share the work however the particles crowd into cells (chains fill a few cells densely).
// Otherwise the lanes stride over all particles.  Every lane of the group calls it.
constexpr int GL = 32;

template <class F>
__device__ __forceinline__ void visit_candidates(const PairConst& pc, const double* xs, const int* cell_of,
                                                 const int* cell_start, const int* perm, const double* xsort,
                                                 int p, int lane, F&& fn) {
    if (pc.mode == 0) {
        for (int q = lane; q < pc.s.Np; q += GL)
            if (q != p) fn(q, xs + 3 * q);
        return;
    }
    // per group: each stencil cell's first sorted index and the inclusive prefix of the counts
    __shared__ int rng[TB / GL][2][GL];
    const int g = (threadIdx.x % TB) / GL;
    int st = 0, cnt = 0;
    if (lane < 27) {
        const int c = cell_of[p];
        const int nx = pc.ncell[0], ny = pc.ncell[1], nz = pc.ncell[2];
        const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
        const int ox = lane % 3 - 1, oy = (lane / 3) % 3 - 1, oz = lane / 9 - 1;
        const int ax = (cx + ox + nx) % nx, ay = (cy + oy + ny) % ny, az = (cz + oz + nz) % nz;
        const int cc = ax + nx * (ay + ny * az);
        st = cell_start[cc];
        cnt = cell_start[cc + 1] - st;
    }
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < GL; o <<= 1) {
        const int v = __shfl_up(incl, o, GL);
        if (lane >= o) incl += v;
    }
    rng[g][0][lane] = st;
    rng[g][1][lane] = incl;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int total = rng[g][1][GL - 1];
    int r = 0;   // the stencil cell of this lane's current candidate (candidates come in cell order)
    for (int k = lane; k < total; k += GL) {
        while (rng[g][1][r] <= k) r++;
        const int t = rng[g][0][r] + k - (r ? rng[g][1][r - 1] : 0);
        const int q = perm[t];
        if (q != p) fn(q, xsort + 3 * t);
    }
}

// The reference's pair test for particles p, q with positions xp, xq (the pair's k = min(p, q)).
__device__ __forceinline__ bool within_pq(const PairConst& pc, int p, const double* xp, int q, const double* xq) {
    const bool pk = p < q;
    const double* xk = pk ? xp : xq;
    const double* xj = pk ? xq : xp;
    double dx = xk[0] - xj[0];
    double dy = xk[1] - xj[1];
    double dz = xk[2] - xj[2];
    if (pc.s.box_active) {
        dx -= pc.s.L[0] * __builtin_rint(dx * pc.s.Linv[0]);
        dy -= pc.s.L[1] * __builtin_rint(dy * pc.s.Linv[1]);
        dz -= pc.s.L[2] * __builtin_rint(dz * pc.s.Linv[2]);
    }
    const double dr = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
    return dr < pc.rverl[ptype(pc.s.Nm, pk ? p : q, pk ? q : p)];
}

__device__ __forceinline__ int group_sum(int v) {
    for (int off = GL / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, GL);
    return v;
}

__global__ __launch_bounds__(TB) void k_nb_count(PairConst pc, const double* __restrict__ xs,
                                                 const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                                                 const int* __restrict__ perm, const double* __restrict__ xsort,
                                                 int* __restrict__ count) {
    const int p = (blockIdx.x * TB + threadIdx.x) / GL;
    const int lane = threadIdx.x % GL;
    if (p >= pc.s.Np) return;   // whole groups leave together
    const double xp[3] = {xs[3 * p], xs[3 * p + 1], xs[3 * p + 2]};
    int cnt = 0;
    visit_candidates(pc, xs, cell_of, cell_start, perm, xsort, p, lane, [&](int q, const double* xq) {
        if (within_pq(pc, p, xp, q, xq)) cnt++;
    });
    cnt = group_sum(cnt);
    if (lane == 0) count[p] = cnt;
}

// Fill row p with its partners in increasing order.  The hits are collected in LDS (cap entries per
// group, cap >= the longest row), then each entry goes to its rank in the row (partners are distinct).
__global__ __launch_bounds__(TB) void k_nb_fill(PairConst pc, const double* __restrict__ xs,
                                                const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                                                const int* __restrict__ perm, const double* __restrict__ xsort,
                                                const int* __restrict__ off, int cap, int* __restrict__ part,
                                                int64_t part_cap, int64_t* __restrict__ overflow) {
    extern __shared__ int lds[];
    const int g = threadIdx.x / GL;
    const int ngroups = blockDim.x / GL;
    int* cursor = lds;                       // [ngroups]
    int* buf = lds + ngroups + g * cap;      // [cap] for this group
    const int p = (blockIdx.x * blockDim.x + threadIdx.x) / GL;
    const int lane = threadIdx.x % GL;
    if (lane == 0) cursor[g] = 0;
    __syncthreads();
    if (p < pc.s.Np) {
        const double xp[3] = {xs[3 * p], xs[3 * p + 1], xs[3 * p + 2]};
        visit_candidates(pc, xs, cell_of, cell_start, perm, xsort, p, lane, [&](int q, const double* xq) {
            if (within_pq(pc, p, xp, q, xq)) {
                const int k = atomicAdd(&cursor[g], 1);
                if (k < cap) buf[k] = q;
            }
        });
    }
    __syncthreads();
    if (p >= pc.s.Np) return;
    const int o = off[p];
    const int n = min(cursor[g], cap);
    // a row longer than the group's LDS row or a list longer than `part` is cut and flagged
    if (lane == 0 && (cursor[g] > cap || (int64_t)off[pc.s.Np] > part_cap)) *overflow = 1;
    for (int e = lane; e < n; e += GL) {
        const int v = buf[e];
        int rank = 0;
        for (int a = 0; a < n; a++) rank += buf[a] < v;
        if ((int64_t)o + rank < part_cap) part[o + rank] = v;
    }
}

// ---------------------------------------------------------------------------- forces
struct ForceConst {
    SysConst s;
    double sig[3];
    int mono_ev, force_sel, fene;
    double k_hs, hs_exp_k, kbond, r0, Lmax, ktrap;
    int trap_on;
};

// Excluded-volume force of the Verlet pair (p, q) on p (pair_interaction_hs / _hs_exp), added to f.
__device__ __forceinline__ void pair_force_add(const ForceConst& fc, const double* __restrict__ xpd, int p, int q,
                                               double& fx, double& fy, double& fz) {
    const SysConst& s = fc.s;
    const int k = min(p, q), j = max(p, q);
    const int type = ptype(s.Nm, k, j);
    if (type == 0 && !fc.mono_ev) return;
    double dx, dy, dz, dr;
    geom(s, xpd, k, j, dx, dy, dz, dr);
    const double dr_inv = 1 / dr;
    const double ex = dx * dr_inv, ey = dy * dr_inv, ez = dz * dr_inv;
    const double sig_eff = fc.sig[type] * (1.0 + 0.01);
    double fmag;
    if (fc.force_sel == PS_FORCE_HS_EXP) {
        const double ex_ = exp(-fc.hs_exp_k * (dr - sig_eff));
        fmag = ex_ / (1 - ex_);
    } else {
        fmag = (dr >= sig_eff) ? 0.0 : fc.k_hs * (sig_eff - dr);
    }
    if (fmag == 0.0) return;
    const double gx = fmag * ex, gy = fmag * ey, gz = fmag * ez;
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

// Bonds (bond_forces, src/rhs.cpp:102-158) and the static harmonic trap on the colloids
// (trapping_forces, src/rhs.cpp:46-100) of particle p, added to f.
__device__ __forceinline__ void bond_trap_add(const ForceConst& fc, const double* __restrict__ x,
                                              const double* __restrict__ xpos0, const int* __restrict__ boff,
                                              const int* __restrict__ bond_k, const int* __restrict__ bond_j, int p,
                                              double& fx, double& fy, double& fz) {
    const SysConst& s = fc.s;
    for (int t = boff[p]; t < boff[p + 1]; t++) {
        const int k = bond_k[t], j = bond_j[t];
        double r0_ = x[3 * k] - x[3 * j], r1_ = x[3 * k + 1] - x[3 * j + 1], r2_ = x[3 * k + 2] - x[3 * j + 2];
        if (s.box_active) {
            r0_ -= s.L[0] * __builtin_rint(r0_ * s.Linv[0]);
            r1_ -= s.L[1] * __builtin_rint(r1_ * s.Linv[1]);
            r2_ -= s.L[2] * __builtin_rint(r2_ * s.Linv[2]);
        }
        const double r = __builtin_sqrt(r0_ * r0_ + r1_ * r1_ + r2_ * r2_);
        double f0 = -fc.kbond * (1 - (fc.r0 / r)) * r0_;
        double f1 = -fc.kbond * (1 - (fc.r0 / r)) * r1_;
        double f2 = -fc.kbond * (1 - (fc.r0 / r)) * r2_;
        if (fc.fene) {
            const double ext2 = 1 / (1 - (r / fc.Lmax) * (r / fc.Lmax));
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
    if (fc.trap_on && p >= s.Nm) {
        const int ic3 = 3 * (p - s.Nm);
        double dx = x[3 * p] - xpos0[ic3], dy = x[3 * p + 1] - xpos0[ic3 + 1], dz = x[3 * p + 2] - xpos0[ic3 + 2];
        if (s.box_active) {
            dx -= s.L[0] * __builtin_rint(dx * s.Linv[0]);
            dy -= s.L[1] * __builtin_rint(dy * s.Linv[1]);
            dz -= s.L[2] * __builtin_rint(dz * s.Linv[2]);
        }
        fx += fc.ktrap * dx;
        fy += fc.ktrap * dy;
        fz += fc.ktrap * dz;
    }
}

// rhs = [0 (n) ; fint (3Np) ; 0 (3Nc)] with fint = pairs + bonds + trap (src/rhs.cpp:160-260).
// Blocks below `mono_blocks` give each monomer a thread; the blocks after them give each colloid a
// wave, whose lanes stride over its partner list (a colloid can have thousands of Verlet partners),
// combined by a fixed shuffle tree, so the result is deterministic.
constexpr int FW = TB / 64;   // colloids per colloid block
__global__ __launch_bounds__(TB) void k_forces(ForceConst fc, const double* __restrict__ xpd,
                                               const double* __restrict__ x, const double* __restrict__ xpos0,
                                               const int* __restrict__ off, const int* __restrict__ part,
                                               const int* __restrict__ boff, const int* __restrict__ bond_k,
                                               const int* __restrict__ bond_j, int mono_blocks,
                                               double* __restrict__ rhs) {
    const SysConst& s = fc.s;
    {   // the rows this launch does not compute: rhs[0:n] and rhs[n + 3Np:] are zero
        const int64_t n3 = 3 * (int64_t)s.Np, nclear = s.ns - n3;
        const int64_t stride = (int64_t)gridDim.x * TB;
        for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < nclear; i += stride)
            rhs[i < s.n ? i : i + n3] = 0.0;
    }
    double fx = 0.0, fy = 0.0, fz = 0.0;
    int p;
    if ((int)blockIdx.x < mono_blocks) {
        p = blockIdx.x * TB + threadIdx.x;
        if (p >= s.Nm) return;
        for (int t = off[p]; t < off[p + 1]; t++) pair_force_add(fc, xpd, p, part[t], fx, fy, fz);
    } else {
        const int lane = threadIdx.x & 63;
        const int ci = ((int)blockIdx.x - mono_blocks) * FW + (int)(threadIdx.x >> 6);
        if (ci >= s.Nc) return;   // whole wave
        p = s.Nm + ci;
        for (int t = off[p] + lane; t < off[p + 1]; t += 64) pair_force_add(fc, xpd, p, part[t], fx, fy, fz);
        for (int o = 32; o > 0; o >>= 1) {
            fx += __shfl_down(fx, o, 64);
            fy += __shfl_down(fy, o, 64);
            fz += __shfl_down(fz, o, 64);
        }
        if (lane != 0) return;
    }
    bond_trap_add(fc, x, xpos0, boff, bond_k, bond_j, p, fx, fy, fz);
    rhs[s.n + 3 * (int64_t)p] = fx;
    rhs[s.n + 3 * (int64_t)p + 1] = fy;
    rhs[s.n + 3 * (int64_t)p + 2] = fz;
}

__global__ void k_rhs_drift(SysConst s, const double* __restrict__ fext, double* __restrict__ rhs) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= s.ns) return;
    rhs[i] = (i < s.n) ? 0.0 : fext[i - s.n];
}

PairConst make_pc(Ctx* c) {
    PairConst pc{};
    pc.s = c->sc;
    const float rv[3] = {(float)(0.05 + 2 * c->p.beta), (float)(1.5 + c->p.beta), (float)2.5};
    const float sg[3] = {(float)(2. * c->p.beta), (float)(1. + c->p.beta), (float)2.0};
    for (int k = 0; k < 3; k++) {
        pc.rverl[k] = rv[k];
        pc.sig[k] = sg[k];
        pc.ncell[k] = c->ncell[k];
        pc.cell_size[k] = c->cell_size[k];
        pc.cell_org[k] = c->cell_org[k];
    }
    pc.mode = c->cell_mode;
    return pc;
}

}  // namespace

ps_err pairs_alloc(Ctx* c) {
    const int Np = c->Np;
    // Cell grid: edge >= the largest Verlet cutoff (the BB one, 2.5 colloid radii).
    const double rmax = std::max({(double)(float)(0.05 + 2 * c->p.beta), (double)(float)(1.5 + c->p.beta), 2.5});
    c->cell_mode = 0;
    int ncells = 1;
    if (c->p.box_active) {
        int nc[3];
        bool ok = true;
        for (int d = 0; d < 3; d++) {
            nc[d] = (int)std::floor(c->p.box[d] / (rmax * (1.0 + 1e-9)));
            if (nc[d] < 3) ok = false;
        }
        if (ok) {
            c->cell_mode = 1;
            for (int d = 0; d < 3; d++) {
                c->ncell[d] = nc[d];
                c->cell_size[d] = c->p.box[d] / nc[d];
                c->cell_org[d] = -0.5 * c->p.box[d];
            }
            ncells = nc[0] * nc[1] * nc[2];
        }
    }
    if (c->cell_mode == 0) {
        for (int d = 0; d < 3; d++) {
            c->ncell[d] = 1;
            c->cell_size[d] = 1.0;
            c->cell_org[d] = 0.0;
        }
    }
    PS_HIP(c, hipMalloc(&c->cell_of, sizeof(int) * std::max(Np, 1)));
    PS_HIP(c, hipMalloc(&c->cell_start, sizeof(int) * (ncells + 2)));
    PS_HIP(c, hipMalloc(&c->cell_count, sizeof(int) * ncells));
    // zero here once: every check_dist leaves the counts at zero (the scatter counts them down)
    PS_HIP(c, hipMemset(c->cell_count, 0, sizeof(int) * ncells));
    PS_HIP(c, hipMalloc(&c->cell_perm, sizeof(int) * std::max(Np, 1)));
    PS_HIP(c, hipMalloc(&c->cell_xs, sizeof(double) * 3 * std::max(Np, 1)));
    PS_HIP(c, hipMalloc(&c->nb_count, sizeof(int) * std::max(Np, 1)));
    PS_HIP(c, hipMalloc(&c->nb_off, sizeof(int) * (Np + 2)));
    c->nb_cap = std::max<int64_t>(1 << 16, 64 * (int64_t)Np);
    PS_HIP(c, hipMalloc(&c->nb_part, sizeof(int) * (size_t)c->nb_cap));
    PS_HIP(c, hipMalloc(&c->dflag, sizeof(int) * 4));
    PS_HIP(c, hipHostMalloc(&c->hflag, sizeof(int) * 4, hipHostMallocDefault));
    // per-particle bond lists (CSR)
    std::vector<std::vector<std::pair<int, int>>> per(Np);
    std::vector<int> b0, b1;
    const ps_params& p = c->p;
    for (int ii = 0; ii < p.Npoly; ii++)
        for (int jj = 0; jj < p.Nmono_per_chain - 1; jj++) {
            const int k = ii * p.Nmono_per_chain + jj;
            if (k + 1 < Np) {
                b0.push_back(k);
                b1.push_back(k + 1);
            }
        }
    if (p.tether && c->Nc > 0) {
        const int cpc = p.Npoly / c->Nc;
        for (int ii = 0; ii < p.Npoly; ii++) {
            int host = c->Nm + (cpc > 0 ? ii / cpc : 0);
            if (host >= Np) host = Np - 1;
            b0.push_back(ii * p.Nmono_per_chain);
            b1.push_back(host);
        }
    }
    c->nbonds = (int)b0.size();
    for (size_t b = 0; b < b0.size(); b++) {
        per[b0[b]].push_back({b0[b], b1[b]});
        per[b1[b]].push_back({b0[b], b1[b]});
    }
    std::vector<int> boff(Np + 1, 0), bk, bj;
    for (int q = 0; q < Np; q++) {
        boff[q + 1] = boff[q] + (int)per[q].size();
        for (auto& pr : per[q]) {
            bk.push_back(pr.first);
            bj.push_back(pr.second);
        }
    }
    const size_t nb = std::max<size_t>(bk.size(), 1);
    PS_HIP(c, hipMalloc(&c->bonds, sizeof(int) * (Np + 1 + 2 * nb)));
    PS_HIP(c, hipMemcpy(c->bonds, boff.data(), sizeof(int) * (Np + 1), hipMemcpyHostToDevice));
    if (!bk.empty()) {
        PS_HIP(c, hipMemcpy(c->bonds + Np + 1, bk.data(), sizeof(int) * bk.size(), hipMemcpyHostToDevice));
        PS_HIP(c, hipMemcpy(c->bonds + Np + 1 + nb, bj.data(), sizeof(int) * bj.size(), hipMemcpyHostToDevice));
    }
    c->bond_stride = nb;
    return PS_OK;
}

void pairs_free(Ctx* c) {
    for (void* p : {(void*)c->cell_of, (void*)c->cell_start, (void*)c->cell_count, (void*)c->cell_perm,
                    (void*)c->cell_xs, (void*)c->nb_count, (void*)c->nb_off, (void*)c->nb_part, (void*)c->dflag,
                    (void*)c->bonds})
        if (p) (void)hipFree(p);
    c->cell_xs = nullptr;
    if (c->hflag) (void)hipHostFree(c->hflag);
    c->cell_of = c->cell_start = c->cell_count = c->cell_perm = c->nb_count = c->nb_off = c->nb_part = nullptr;
    c->dflag = c->bonds = nullptr;
    c->hflag = nullptr;
}

ps_err scan_ints(Ctx* c, const int* in, int* out, int n) {
    k_scan<<<1, SCAN_T, 0, c->stream>>>(const_cast<int*>(in), out, n, 0);
    PS_LAUNCH(c);
    return PS_OK;
}

// The Verlet list with no host round trip: every row holds at most VL_ROW_CAP partners and the
// whole list at most nb_cap entries (64 per particle); a longer row or list sets the sticky
// DC_VLIST_OVERFLOW flag, which the entry points report as an error.
constexpr int VL_ROW_CAP = 511;

ps_err check_dist_dev(Ctx* c, double* xsave) {
    if (c->arrow) {
        if (xsave) PS_TRY(vec_copy(c, xsave, c->x, 3 * (int64_t)c->Np));
        return arrow_check_dist(c);
    }
    const int Np = c->Np;
    hipStream_t st = c->stream;
    PairConst pc = make_pc(c);
    // geometry of this check_dist = current positions (copied by k_bin in cell mode)
    const unsigned gp = (unsigned)((Np + TB - 1) / TB);
    const double* xsort = c->xpd;
    if (c->cell_mode != 1) {
        PS_HIP(c, hipMemcpyAsync(c->xpd, c->x, sizeof(double) * 3 * (size_t)Np, hipMemcpyDeviceToDevice, st));
        if (xsave) PS_TRY(vec_copy(c, xsave, c->x, 3 * (int64_t)Np));
    }
    if (c->cell_mode == 1) {
        const int ncells = c->ncell[0] * c->ncell[1] * c->ncell[2];
        // the counts are zero on entry (k_scatter leaves them so)
        k_bin<<<gp, TB, 0, st>>>(pc, c->x, c->xpd, xsave, c->cell_of, c->cell_count);
        PS_LAUNCH(c);
        k_scan<<<1, SCAN_T, 0, st>>>(c->cell_count, c->cell_start, ncells, 0);
        PS_LAUNCH(c);
        k_scatter<<<gp, TB, 0, st>>>(Np, c->xpd, c->cell_of, c->cell_start, c->cell_count, c->cell_perm, c->cell_xs);
        PS_LAUNCH(c);
        xsort = c->cell_xs;
    }
    const unsigned gg = (unsigned)(((int64_t)Np * GL + TB - 1) / TB);
    k_nb_count<<<gg, TB, 0, st>>>(pc, c->xpd, c->cell_of, c->cell_start, c->cell_perm, xsort, c->nb_count);
    PS_LAUNCH(c);
    k_scan<<<1, SCAN_T, 0, st>>>(c->nb_count, c->nb_off, Np, 0);
    PS_LAUNCH(c);
    // LDS per block: a cursor and VL_ROW_CAP entries for each of its 8 groups
    const int gpb = TB / GL;
    const size_t lds = sizeof(int) * (size_t)gpb * (VL_ROW_CAP + 1);
    k_nb_fill<<<(unsigned)((Np + gpb - 1) / gpb), gpb * GL, lds, st>>>(pc, c->xpd, c->cell_of, c->cell_start,
                                                                        c->cell_perm, xsort, c->nb_off, VL_ROW_CAP,
                                                                        c->nb_part, c->nb_cap,
                                                                        c->dcnt + DC_VLIST_OVERFLOW);
    PS_LAUNCH(c);
    c->have_pd = true;
    return PS_OK;
}

// check_dist at an RFD probe point (src/drift.cpp:38,51): mob() there reads only the pair geometry,
// i.e. the positions, and RHS(true) uses no Verlet pairs, so the list built at a probe is never read
// (the corrector's check_dist replaces it before the next force evaluation).  Only the geometry
// snapshot is taken; the arrowhead regime keeps its full check_dist, whose AB list its mob() reads.
ps_err probe_geometry_dev(Ctx* c) {
    if (c->arrow) return check_dist_dev(c);
    PS_HIP(c, hipMemcpyAsync(c->xpd, c->x, sizeof(double) * 3 * (size_t)c->Np, hipMemcpyDeviceToDevice, c->stream));
    c->have_pd = true;
    return PS_OK;
}

// The pair count of the last check_dist (a host round trip); an overflowed list is an error.
ps_err vlist_count(Ctx* c, int64_t* n_pairs) {
    PS_HIP(c, hipMemcpyAsync(c->hflag, c->nb_off + c->Np, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PS_HIP(c, hipMemcpyAsync(c->dcnt_h, c->dcnt, sizeof(int64_t) * DC_N, hipMemcpyDeviceToHost, c->stream));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    if (c->dcnt_h[DC_VLIST_OVERFLOW]) {
        c->set_error("check_dist: a Verlet row exceeded " + std::to_string(VL_ROW_CAP) +
                     " partners or the list exceeded " + std::to_string(c->nb_cap) + " entries");
        return PS_ERR_STATE;
    }
    c->nvlist = c->hflag[0] / 2;
    c->cnt.nvlist = c->nvlist;
    if (n_pairs) *n_pairs = c->nvlist;
    return PS_OK;
}

// RHS(true) (src/rhs.cpp:175-184) into `out` (ns): zero force rows, the RFD noise in the constraint rows.
ps_err rhs_drift_dev(Ctx* c, double* out) {
    k_rhs_drift<<<(unsigned)((c->sc.ns + TB - 1) / TB), TB, 0, c->stream>>>(c->sc, c->fext, out);
    PS_LAUNCH(c);
    return PS_OK;
}

ps_err rhs_dev(Ctx* c, int drift) {
    hipStream_t st = c->stream;
    const SysConst& s = c->sc;
    const unsigned ge = (unsigned)((s.ns + TB - 1) / TB);
    if (drift) {
        k_rhs_drift<<<ge, TB, 0, st>>>(s, c->fext, c->rhs);
        PS_LAUNCH(c);
        return PS_OK;
    }
    ForceConst fc{};
    fc.s = s;
    const float sg[3] = {(float)(2. * c->p.beta), (float)(1. + c->p.beta), (float)2.0};
    for (int k = 0; k < 3; k++) fc.sig[k] = sg[k];
    fc.mono_ev = c->p.mono_ev;
    fc.force_sel = c->p.force_sel;
    fc.fene = c->p.fene;
    fc.k_hs = (c->p.hs_kmode == 1) ? c->p.hs_kpref / (c->p.dt * c->p.beta) : c->p.hs_kpref / c->p.dt;
    fc.hs_exp_k = c->p.hs_exp_k;
    fc.kbond = c->p.kbond;
    fc.r0 = c->p.r0;
    fc.Lmax = c->p.Lmax;
    // trapping_forces: |ktrap| < 1e-12 disables; weaken_trap replaces ktrap permanently after tstop
    fc.trap_on = !(std::fabs(c->ktrap_cur) < 1e-12);
    if (fc.trap_on && c->t > c->p.tstart + c->p.trun && c->p.weaken_trap >= 0) c->ktrap_cur = c->p.weaken_trap;
    fc.ktrap = c->ktrap_cur;
    const int Np = c->Np;
    const size_t nb = c->bond_stride;
    const int* boff = c->bonds;
    const int* bk = c->bonds + Np + 1;
    const int* bj = c->bonds + Np + 1 + nb;
    const int mono_blocks = (c->Nm + TB - 1) / TB, coll_blocks = (c->Nc + FW - 1) / FW;
    k_forces<<<(unsigned)std::max(1, mono_blocks + coll_blocks), TB, 0, st>>>(fc, c->xpd, c->x, c->xpos0, c->nb_off,
                                                                             c->nb_part, boff, bk, bj, mono_blocks,
                                                                             c->rhs);
    PS_LAUNCH(c);
    // mm_HI off: monomer-monomer excluded volume through the monomer cell list (src/rhs.cpp:200-205)
    if (c->arrow && c->p.mono_ev) PS_TRY(arrow_wca(c));
    return PS_OK;
}

// Copy back the reference-ordered Verlet list (q > k half of each row).
ps_err vlist_copy(Ctx* c, int32_t* rows, int32_t* partners, int64_t* pidx, int64_t cap, int64_t* n_pairs) {
    const int Np = c->Np;
    if (!c->arrow) PS_TRY(vlist_count(c, nullptr));
    std::vector<int> off(Np + 1);
    PS_HIP(c, hipMemcpyAsync(off.data(), c->nb_off, sizeof(int) * (Np + 1), hipMemcpyDeviceToHost, c->stream));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<int> part(std::max(off[Np], 1));
    if (off[Np] > 0)
        PS_HIP(c, hipMemcpy(part.data(), c->nb_part, sizeof(int) * off[Np], hipMemcpyDeviceToHost));
    const int64_t Nm = c->Nm, Nc = c->Nc;
    // with mm_HI off no monomer-monomer pair is stored: AB pairs are numbered from 0 (src/init.cpp:263-290)
    const int64_t npAA = c->arrow ? 0 : Nm * (Nm - 1) / 2, npAB = Nm * Nc;
    int64_t cnt = 0;
    for (int k = 0; k < Np; k++)
        for (int t = off[k]; t < off[k + 1]; t++) {
            const int j = part[t];
            if (j <= k) continue;
            if (cnt < cap) {
                if (rows) rows[cnt] = k;
                if (partners) partners[cnt] = j;
                if (pidx) {
                    int64_t v;
                    if (j < Nm) v = (int64_t)k * (2 * Nm - k - 1) / 2 + (j - k - 1);
                    else if (k < Nm) v = npAA + (int64_t)k * Nc + (j - Nm);
                    else {
                        const int64_t kc = k - Nm, jc = j - Nm;
                        v = npAA + npAB + kc * (2 * Nc - kc - 1) / 2 + (jc - kc - 1);
                    }
                    pidx[cnt] = v;
                }
            }
            cnt++;
        }
    if (n_pairs) *n_pairs = cnt;
    return PS_OK;
}
