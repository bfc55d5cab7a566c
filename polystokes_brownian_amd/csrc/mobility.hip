// mobility.hip -- u = M f with the FTS grand-mobility tiles generated on the fly (gfx950).
//
// M is never stored.  One apply launches four independent kernels whose partial sums are combined
// by a deterministic finalize (no atomics, so every apply is bit-reproducible):
//   k_aa       monomer rows x monomer sources: the 3x3 tile of mobility(..., AA=true)
//              (src/mob.cpp:96-151), source positions/forces staged through LDS in 256-bead tiles,
//              grid = (target blocks, source splits) to fill the 256 CUs;
//   k_ab_mono  monomer rows x colloid sources: rotation and stresslet couplings of mobility_AB
//              (src/mob.cpp:362-537) plus the carried-over translation tile;
//   k_ab_coll  colloid rows x monomer sources (the transposed AB blocks, src/mob.cpp:926-972);
//   k_bb       colloid rows x colloid sources: the full 11x11 FTS tile (src/mob.cpp:133-356,
//              placement :1026-1144), always FP64;
//   k_finalize self blocks (fill_self, src/mob.cpp:539-633) + the partial sums, and for the saddle
//              operator the constraint coupling B = [-I; 0] (src/arrays.cpp:349-408).
// The tile algebra is contracted analytically: the AA tile x ee + y (I - ee) applied to f is
// y f + (x - y) (e.f) e; the flattened stresslet couplings act through the traceless symmetric
// matrix S^ = [[S0,S1,S2],[S1,S4,S3],[S2,S3,-S0-S4]] (the inverse of the flattening at
// src/mob.cpp:220-227), so no 3x5 or 5x5 block is ever materialised for AB pairs.
#include "ps_common.h"
#include "tiles.h"
#include "lzdots.h"
#include "dpp.h"

#include <hip/hip_ext.h>

#include <type_traits>
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

namespace {



template <class T>
__device__ __forceinline__ void min_img(const SysConst& s, T& dx, T& dy, T& dz) {
    dx -= (T)s.L[0] * ps_rint<T>(dx * (T)s.Linv[0]);
    dy -= (T)s.L[1] * ps_rint<T>(dy * (T)s.Linv[1]);
    dz -= (T)s.L[2] * ps_rint<T>(dz * (T)s.Linv[2]);
}


// Block-wide sum of NV values per thread (256 threads = 4 waves); result valid in thread 0.
template <class T, int NV>
__device__ __forceinline__ void block_sum(T (&v)[NV], T (*sh)[NV]) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum64(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[wid][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] = sh[0][k] + sh[1][k] + sh[2][k] + sh[3][k];
}

// ---------------------------------------------------------------------------- AA (monomer rows)
template <class T, bool BOX, bool CHAIN>
__global__ __launch_bounds__(PS_BLOCK) void k_aa(SysConst s, const T* __restrict__ xm,
                                                 const double* __restrict__ f,
                                                 const int* __restrict__ chain, int per_split,
                                                 T* __restrict__ part) {
    __shared__ T sx[PS_BLOCK], sy[PS_BLOCK], sz[PS_BLOCK], sfx[PS_BLOCK], sfy[PS_BLOCK], sfz[PS_BLOCK];
    __shared__ int sc[PS_BLOCK];
    const int Nm = s.Nm;
    const int i = blockIdx.x * PS_BLOCK + threadIdx.x;
    const bool act = i < Nm;
    T xi = 0, yi = 0, zi = 0;
    int ci = -1;
    if (act) {
        xi = xm[3 * i];
        yi = xm[3 * i + 1];
        zi = xm[3 * i + 2];
        if (CHAIN) ci = chain[i];
    }
    // mobility(..., AA=true), src/mob.cpp:122-131, applied as y f + (x - y)/r^2 (d.f) d:
    //   far  (r > 2 beta): y = r^-1 (3/4 - 1/2 r^-2),  (x - y)/r^2 = r^-3 (-3/8 - 3/2 r^-2)
    //   near (r <= 2 beta): y = 1 - 9/32 r,          (x - y)/r^2 = (21/32 r - 2) r^-2
    const T tb2 = (T)(s.two_beta * s.two_beta);
    T ux = 0, uy = 0, uz = 0;
    const int s0 = blockIdx.y * per_split;
    const int s1 = min(Nm, s0 + per_split);
    for (int t0 = s0; t0 < s1; t0 += PS_BLOCK) {
        __syncthreads();
        const int j = t0 + threadIdx.x;
        if (j < s1) {
            sx[threadIdx.x] = xm[3 * j];
            sy[threadIdx.x] = xm[3 * j + 1];
            sz[threadIdx.x] = xm[3 * j + 2];
            sfx[threadIdx.x] = (T)f[3 * j];
            sfy[threadIdx.x] = (T)f[3 * j + 1];
            sfz[threadIdx.x] = (T)f[3 * j + 2];
            if (CHAIN) sc[threadIdx.x] = chain[j];
        }
        __syncthreads();
        const int cnt = min(PS_BLOCK, s1 - t0);
        // Source tiles are aligned to 256, so only the tile holding this block's own beads can
        // contain i == j: the masked loop runs there (and under chain_HI), the lean loop elsewhere.
        const bool masked = CHAIN || (t0 == (int)(blockIdx.x * PS_BLOCK));
        if (masked) {
            for (int k = 0; k < cnt; k++) {
                T dx = xi - sx[k], dy = yi - sy[k], dz = zi - sz[k];
                if (BOX) min_img<T>(s, dx, dy, dz);
                T r2 = dx * dx + dy * dy + dz * dz;
                bool skip = (t0 + k == i);
                if (CHAIN) skip |= (sc[k] == ci);
                r2 = skip ? (T)1 : r2;
                const T rinv = ps_rsqrt(r2);
                const T rinv2 = rinv * rinv;
                T ya = rinv * ((T)C3D4 - (T)C1D2 * rinv2);
                T cc = rinv * rinv2 * (-(T)C3D8 - (T)C3D2 * rinv2);
                if (r2 <= tb2) {
                    const T r = r2 * rinv;
                    ya = (T)1 - (T)C9D32 * r;
                    cc = ((T)(21. / 32.) * r - (T)2) * rinv2;
                }
                ya = skip ? (T)0 : ya;
                cc = skip ? (T)0 : cc;
                const T fx = sfx[k], fy = sfy[k], fz = sfz[k];
                const T cs = cc * (dx * fx + dy * fy + dz * fz);
                ux += ya * fx + cs * dx;
                uy += ya * fy + cs * dy;
                uz += ya * fz + cs * dz;
            }
        } else {
#pragma unroll 4
            for (int k = 0; k < cnt; k++) {
                T dx = xi - sx[k], dy = yi - sy[k], dz = zi - sz[k];
                if (BOX) min_img<T>(s, dx, dy, dz);
                const T r2 = dx * dx + dy * dy + dz * dz;
                const T rinv = ps_rsqrt(r2);
                const T rinv2 = rinv * rinv;
                T ya = rinv * ((T)C3D4 - (T)C1D2 * rinv2);
                T cc = rinv * rinv2 * (-(T)C3D8 - (T)C3D2 * rinv2);
                if (__builtin_expect(r2 <= tb2, 0)) {
                    const T r = r2 * rinv;
                    ya = (T)1 - (T)C9D32 * r;
                    cc = ((T)(21. / 32.) * r - (T)2) * rinv2;
                }
                const T fx = sfx[k], fy = sfy[k], fz = sfz[k];
                const T cs = cc * (dx * fx + dy * fy + dz * fz);
                ux += ya * fx + cs * dx;
                uy += ya * fy + cs * dy;
                uz += ya * fz + cs * dz;
            }
        }
    }
    if (act) {
        part[((int64_t)blockIdx.y * Nm + i) * 3 + 0] = ux;
        part[((int64_t)blockIdx.y * Nm + i) * 3 + 1] = uy;
        part[((int64_t)blockIdx.y * Nm + i) * 3 + 2] = uz;
    }
}

// ---------------------------------------------------------------------------- AA, symmetric form
// The AA tile is symmetric in the pair (the same e e^T for d and -d), so one evaluation of the
// pair geometry serves both u_i += A f_j and u_j += A f_i.  Block (I, J), I < J, of the
// upper triangle of 256-bead tiles: lane t of wave w owns target i = 256 I + 64 w + t in registers;
// the source tile J sits in LDS in four windows of 64, each stored twice (entries 128 q + m and
// 128 q + m + 64 hold source 64 q + m), so at step k lane t reads entry 128 q + t + k: a systolic
// sweep in which every lane meets a different source, with no index arithmetic.  The u_j
// contributions accumulate in registers that rotate with the sources and are added to an LDS
// accumulator per source at the end of each window; the waves visit the windows in staggered
// phases (window (w + p) mod 4 in phase p, a barrier between phases), so no two waves touch one
// window at once and every accumulator sums in a fixed order.
// Diagonal blocks (I == J) run the ordered loop.  Output: part[(J Nm + i) 3 + c] is the
// contribution of source tile J to target i, written exactly once per (J, i).
//
// Positions.  FP64, and FP32 without a box: x / L in the compute type (L = 1 without a box), so the
// minimum image is d - rint(d).  FP32 in a cubic box: fixed point, x / L * 2^32 as a 32-bit
// integer, so the wrap-around of the integer difference IS the minimum image (one integer
// subtraction and one conversion per axis instead of three floating-point operations), and the
// separation carries 32 bits of the box for every pair -- close pairs keep their relative
// precision (6e-8) instead of losing it to the rounding of two large coordinates.
// In either case the unit u of d (L, or L / 2^32) folds into the constants:
//   far:  ya = a (c1 - c2 a^2),  cs = a^3 (c3 + c4 a^2)   with a = |d|^-1, u += ya f + cs (d.f) d
//   near: r = u |d|^2 a:  ya = 1 - 9/32 r,  cs = (21/32 r - 2) a^2
// c1 = 3/(4u), c2 = 1/(2u^3), c3 = -3/(8u), c4 = -3/(2u^3).
struct AAConst {
    double c1, c2, c3, c4, L, tb2s, clamp;
};

template <class T>
struct Vec4;
template <>
struct Vec4<float> {
    using type = float4;
};
template <>
struct Vec4<double> {
    using type = double4;
};

// The stored position type of the symmetric kernel: FIX = FP32 in a box (fixed point).
template <class T, bool BOX>
struct AAPos {
    static constexpr bool FIX = sizeof(T) == sizeof(float) && BOX;
    using P = typename std::conditional<FIX, int, T>::type;
    using V = typename std::conditional<FIX, int4, typename Vec4<T>::type>::type;
};

template <class T>
__device__ __forceinline__ T ps_fma(T a, T b, T c);
template <>
__device__ __forceinline__ float ps_fma<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <>
__device__ __forceinline__ double ps_fma<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

// One separation component in the kernel's unit: the minimum image of x_i - x_j.
template <class T, bool BOX, bool FIX, class P>
__device__ __forceinline__ T aa_sep(P xi, P xj) {
    if constexpr (FIX) {
        return (T)(int)((unsigned)xi - (unsigned)xj);
    } else {
        T d = xi - xj;
        if (BOX) d -= ps_rint<T>(d);
        return d;
    }
}

// lane t receives the value of lane (t + 1) mod 64 (DPP wave_rol:1; a double moves as two halves).
// Every lane reads a valid lane, so no old value is needed.
__device__ __forceinline__ float wave_rol1(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x134, 0xF, 0xF, false));
}
__device__ __forceinline__ double wave_rol1(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), 0x134, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x134, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

template <class T, bool BOX>
struct AASmem {
    typename AAPos<T, BOX>::V S[512];
    typename Vec4<T>::type F[512];
    T A[3][PS_BLOCK];
};

// One tile pair (I, J); PAD: the pair involves the partial last tile (ghost beads beyond Nm, at the
// origin with f = 0).
template <class T, bool BOX, bool PAD>
__device__ __forceinline__ void aa_sym_tile(int Nm, int I, int J, const AAConst& k,
                                            const typename AAPos<T, BOX>::P* __restrict__ xs,
                                            const double* __restrict__ f, T* __restrict__ part, AASmem<T, BOX>& sm) {
    using PT = AAPos<T, BOX>;
    constexpr bool FIX = PT::FIX;
    using P = typename PT::P;
    using PV = typename PT::V;
    using V = typename Vec4<T>::type;
    PV* S = sm.S;
    V* F = sm.F;
    T(*A)[PS_BLOCK] = sm.A;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    {
        const int j = J * PS_BLOCK + t;
        PV sv = {0, 0, 0, 0};
        V fv = {0, 0, 0, 0};
        if (!PAD || j < Nm) {
            sv.x = xs[4 * j];
            sv.y = xs[4 * j + 1];
            sv.z = xs[4 * j + 2];
            fv.x = (T)f[3 * j];
            fv.y = (T)f[3 * j + 1];
            fv.z = (T)f[3 * j + 2];
        }
        const int e = (w << 7) + l;
        S[e] = sv;
        S[e + 64] = sv;
        F[e] = fv;
        F[e + 64] = fv;
#pragma unroll
        for (int c = 0; c < 3; c++) A[c][t] = 0;
    }
    const int i = I * PS_BLOCK + t;
    P xi = 0, yi = 0, zi = 0;
    T fxi = 0, fyi = 0, fzi = 0;
    if (!PAD || i < Nm) {
        xi = xs[4 * i];
        yi = xs[4 * i + 1];
        zi = xs[4 * i + 2];
        fxi = (T)f[3 * i];
        fyi = (T)f[3 * i + 1];
        fzi = (T)f[3 * i + 2];
    }
    const T c1 = (T)k.c1, c2 = (T)k.c2, c3 = (T)k.c3, c4 = (T)k.c4, Lb = (T)k.L, tb2s = (T)k.tb2s;
    T ux = 0, uy = 0, uz = 0;
    __syncthreads();
    if (I == J) {
        // ordered loop over the 256 sources of the own tile, skipping i itself (and ghosts)
        for (int q = 0; q < 4; q++) {
            for (int m = 0; m < 64; m++) {
                const int jt = (q << 6) + m;
                const PV xj = S[(q << 7) + m];
                const V fj = F[(q << 7) + m];
                const T dx = aa_sep<T, BOX, FIX>(xi, xj.x), dy = aa_sep<T, BOX, FIX>(yi, xj.y),
                        dz = aa_sep<T, BOX, FIX>(zi, xj.z);
                T r2 = dx * dx + dy * dy + dz * dz;
                bool skip = jt == t;
                if (PAD) skip |= (J * PS_BLOCK + jt >= Nm);
                r2 = skip ? (T)1 : r2;
                const T a = ps_rsqrt(r2);
                const T a2 = a * a;
                T ya = a * (c1 - c2 * a2);
                T cs = a * a2 * (c3 + c4 * a2);
                if (r2 <= tb2s) {
                    const T r = Lb * r2 * a;
                    ya = (T)1 - (T)C9D32 * r;
                    cs = ((T)(21. / 32.) * r - (T)2) * a2;
                }
                ya = skip ? (T)0 : ya;
                cs = skip ? (T)0 : cs;
                const T ci = cs * (dx * fj.x + dy * fj.y + dz * fj.z);
                ux = ps_fma(ya, fj.x, ux);
                uy = ps_fma(ya, fj.y, uy);
                uz = ps_fma(ya, fj.z, uz);
                ux = ps_fma(ci, dx, ux);
                uy = ps_fma(ci, dy, uy);
                uz = ps_fma(ci, dz, uz);
            }
        }
        if (!PAD || i < Nm) {
            part[((int64_t)J * Nm + i) * 3 + 0] = ux;
            part[((int64_t)J * Nm + i) * 3 + 1] = uy;
            part[((int64_t)J * Nm + i) * 3 + 2] = uz;
        }
        return;
    }
    const T clamp = (T)k.clamp;
    // The u_j partial sums travel with their sources in registers: at step m lane t holds the sum
    // for source (t + m - 1) mod 64 of the window; a wave rotation (DPP wave_rol:1, lane t reads lane
    // t + 1) hands it on to the lane that meets that source's successor, and the step's two
    // contributions go into it with fused multiply-adds.  After 64 steps every lane has met every
    // source.  (LDS atomics for the same sums measured several times slower.)
    constexpr int UNR = sizeof(T) == sizeof(float) ? 8 : 4;
    for (int p = 0; p < 4; p++) {
        const int q = (w + p) & 3;
        const int base = (q << 7) + l;
        const PV* Sq = S + base;
        const V* Fq = F + base;
        T Rx = 0, Ry = 0, Rz = 0;
#pragma unroll UNR
        for (int m = 0; m < 64; m++) {
            const PV xj = Sq[m];
            const V fj = Fq[m];
            const T dx = aa_sep<T, BOX, FIX>(xi, xj.x), dy = aa_sep<T, BOX, FIX>(yi, xj.y),
                    dz = aa_sep<T, BOX, FIX>(zi, xj.z);
            T r2 = ps_fma(dz, dz, ps_fma(dy, dy, dx * dx));
            if (PAD) r2 = r2 > clamp ? r2 : clamp;   // ghost beads (f = 0) stay finite
            const T a = ps_rsqrt(r2);
            const T a2 = a * a;
            T ya = a * ps_fma(-c2, a2, c1);
            T cs = (a * a2) * ps_fma(c4, a2, c3);
            if (__builtin_expect(r2 <= tb2s, 0)) {
                const T r = Lb * r2 * a;
                ya = (T)1 - (T)C9D32 * r;
                cs = ((T)(21. / 32.) * r - (T)2) * a2;
            }
            // u_i += ya f_j + cs (d.f_j) d
            const T ci = cs * ps_fma(dz, fj.z, ps_fma(dy, fj.y, dx * fj.x));
            ux = ps_fma(ya, fj.x, ux);
            uy = ps_fma(ya, fj.y, uy);
            uz = ps_fma(ya, fj.z, uz);
            ux = ps_fma(ci, dx, ux);
            uy = ps_fma(ci, dy, uy);
            uz = ps_fma(ci, dz, uz);
            // u_j += ya f_i + cs (d.f_i) d
            const T cj = cs * ps_fma(dz, fzi, ps_fma(dy, fyi, dx * fxi));
            Rx = ps_fma(cj, dx, ps_fma(ya, fxi, wave_rol1(Rx)));
            Ry = ps_fma(cj, dy, ps_fma(ya, fyi, wave_rol1(Ry)));
            Rz = ps_fma(cj, dz, ps_fma(ya, fzi, wave_rol1(Rz)));
        }
        // lane t now holds the complete sum of the wave for source (t + 63) mod 64; the waves
        // visit a window one at a time, so a plain LDS update is race-free and ordered
        const int e = (q << 6) + ((l + 63) & 63);
        A[0][e] += Rx;
        A[1][e] += Ry;
        A[2][e] += Rz;
        __syncthreads();
    }
    if (!PAD || i < Nm) {
        part[((int64_t)J * Nm + i) * 3 + 0] = ux;
        part[((int64_t)J * Nm + i) * 3 + 1] = uy;
        part[((int64_t)J * Nm + i) * 3 + 2] = uz;
    }
    const int j = J * PS_BLOCK + t;
    if (!PAD || j < Nm) {
#pragma unroll
        for (int c = 0; c < 3; c++) part[((int64_t)I * Nm + j) * 3 + c] = A[c][t];
    }
}

// Scaled monomer positions for the symmetric tile pairs: [Nm][4] (x, y, z, 0) * inv_L.
template <class T>
__global__ void k_aa_pos(int Nm, const double* __restrict__ xm, double inv_L, T* __restrict__ xs) {
    const int i = blockIdx.x * PS_BLOCK + threadIdx.x;
    if (i >= Nm) return;
    xs[4 * i] = (T)(xm[3 * i] * inv_L);
    xs[4 * i + 1] = (T)(xm[3 * i + 1] * inv_L);
    xs[4 * i + 2] = (T)(xm[3 * i + 2] * inv_L);
    xs[4 * i + 3] = (T)0;
}

// Fixed-point monomer positions for the FP32 symmetric tile pairs in a box: [Nm][4] (x, y, z, 0) with
// x / L wrapped into [-1/2, 1/2] and scaled by 2^32 (the ends meet: 2^31 wraps to -2^31).
__global__ void k_aa_pos_fix(int Nm, const double* __restrict__ xm, double inv_L, int* __restrict__ xs) {
    const int i = blockIdx.x * PS_BLOCK + threadIdx.x;
    if (i >= Nm) return;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double t = xm[3 * i + d] * inv_L;
        t -= __builtin_rint(t);
        const long long v = __builtin_llrint(t * 4294967296.0);
        xs[4 * i + d] = (int)(unsigned)(unsigned long long)v;
    }
    xs[4 * i + 3] = 0;
}

// ---------------------------------------------------------------------------- AB, monomer rows
// u_m += stale . sum_c F_c + y12b (e x T_c) - x12g (e.S^e) e + 2 y12g (S^e - (e.S^e) e)
template <class T>
struct ABMonoSmem {
    T cx[PS_BLOCK], cy[PS_BLOCK], cz[PS_BLOCK];
    T cv[11][PS_BLOCK];
};

// One block of monomer rows (bx) against one chunk of colloid sources (by).
template <class T, bool BOX>
__device__ __forceinline__ void ab_mono_block(int bx, int by, const SysConst& s, const T* __restrict__ xm,
                                              const double* __restrict__ f, const double* __restrict__ stale,
                                              int per_cchunk, T* __restrict__ part, ABMonoSmem<T>& sm) {
    T* cx = sm.cx;
    T* cy = sm.cy;
    T* cz = sm.cz;
    T(*cv)[PS_BLOCK] = sm.cv;
    const int Nm = s.Nm;
    // colloid sources [cb, ce) of this block's chunk
    const int cb = by * per_cchunk;
    const int ce = min(s.Nc, cb + per_cchunk);
    const int i = bx * PS_BLOCK + threadIdx.x;
    const bool act = i < Nm;
    T xi = 0, yi = 0, zi = 0;
    if (act) {
        xi = xm[3 * i];
        yi = xm[3 * i + 1];
        zi = xm[3 * i + 2];
    }
    const T gA = (T)(C9D20 * (3 + 5 * s.beta2));
    const T gB = (T)(C9D10 * (C1D2 + C5D6 * s.beta2));
    T ux = 0, uy = 0, uz = 0, Fx = 0, Fy = 0, Fz = 0;
    const int64_t nm3 = s.nm3, nc3 = s.nc3, nu = s.nu;
    for (int c0 = cb; c0 < ce; c0 += PS_BLOCK) {
        __syncthreads();
        const int c = c0 + threadIdx.x;
        if (c < ce) {
            const int g = Nm + c;
            cx[threadIdx.x] = xm[3 * g];
            cy[threadIdx.x] = xm[3 * g + 1];
            cz[threadIdx.x] = xm[3 * g + 2];
            for (int k = 0; k < 3; k++) cv[k][threadIdx.x] = (T)f[nm3 + 3 * c + k];
            for (int k = 0; k < 3; k++) cv[3 + k][threadIdx.x] = (T)f[nm3 + nc3 + 3 * c + k];
            for (int k = 0; k < 5; k++) cv[6 + k][threadIdx.x] = (T)f[nu + 5 * c + k];
        }
        __syncthreads();
        const int cnt = min(PS_BLOCK, ce - c0);
        for (int k = 0; k < cnt; k++) {
            T dx = xi - cx[k], dy = yi - cy[k], dz = zi - cz[k];
            if (BOX) min_img<T>(s, dx, dy, dz);
            const T r2 = dx * dx + dy * dy + dz * dz;
            const T rinv = ps_rsqrt(r2);
            const T ex = dx * rinv, ey = dy * rinv, ez = dz * rinv;
            const T rinv2 = rinv * rinv, rinv4 = rinv2 * rinv2;
            const T y12b = -(T)C3D4 * rinv2;
            const T x12g = (T)C9D4 * rinv2 - gA * rinv4;
            const T y12g = gB * rinv4;
            const T Tx = cv[3][k], Ty = cv[4][k], Tz = cv[5][k];
            const T S0 = cv[6][k], S1 = cv[7][k], S2 = cv[8][k], S3 = cv[9][k], S4 = cv[10][k];
            // S^ e
            const T sex = S0 * ex + S1 * ey + S2 * ez;
            const T sey = S1 * ex + S4 * ey + S3 * ez;
            const T sez = S2 * ex + S3 * ey - (S0 + S4) * ez;
            const T ese = ex * sex + ey * sey + ez * sez;
            const T a = -x12g * ese - (T)2 * y12g * ese;
            const T b2 = (T)2 * y12g;
            ux += y12b * (ey * Tz - ez * Ty) + a * ex + b2 * sex;
            uy += y12b * (ez * Tx - ex * Tz) + a * ey + b2 * sey;
            uz += y12b * (ex * Ty - ey * Tx) + a * ez + b2 * sez;
            Fx += cv[0][k];
            Fy += cv[1][k];
            Fz += cv[2][k];
        }
    }
    if (act) {
        // the translation block every AB pair carries (stale tile), applied to sum_c F_c
        ux += (T)stale[0] * Fx + (T)stale[1] * Fy + (T)stale[2] * Fz;
        uy += (T)stale[3] * Fx + (T)stale[4] * Fy + (T)stale[5] * Fz;
        uz += (T)stale[6] * Fx + (T)stale[7] * Fy + (T)stale[8] * Fz;
        part[((int64_t)by * Nm + i) * 3 + 0] = ux;
        part[((int64_t)by * Nm + i) * 3 + 1] = uy;
        part[((int64_t)by * Nm + i) * 3 + 2] = uz;
    }
}

template <class T, bool BOX>
__global__ __launch_bounds__(PS_BLOCK) void k_ab_mono(SysConst s, const T* __restrict__ xm,
                                                      const double* __restrict__ f,
                                                      const double* __restrict__ stale, int per_cchunk,
                                                      T* __restrict__ part) {
    __shared__ ABMonoSmem<T> sm;
    ab_mono_block<T, BOX>(blockIdx.x, blockIdx.y, s, xm, f, stale, per_cchunk, part, sm);
}

// ---------------------------------------------------------------------------- AB, colloid rows
// Omega_c += y12b (f_m x e);  E_c += flatten(G), G_ij = sum_k f_k gt_kij;  also sum_m f_m.
// One colloid row set (c) against one chunk of monomer sources (ch).
template <class T, bool BOX>
__device__ __forceinline__ void ab_coll_block(int c, int ch, const SysConst& s, const T* __restrict__ xm,
                                              const double* __restrict__ f, int per_chunk, T* __restrict__ part_c,
                                              T* __restrict__ part_fs, T (*sh)[11]) {
    const int Nm = s.Nm, Nc = s.Nc;
    const int g = Nm + c;
    const T xc = xm[3 * g], yc = xm[3 * g + 1], zc = xm[3 * g + 2];
    const T gA = (T)(C9D20 * (3 + 5 * s.beta2));
    const T gB = (T)(C9D10 * (C1D2 + C5D6 * s.beta2));
    T acc[11];
#pragma unroll
    for (int k = 0; k < 11; k++) acc[k] = 0;
    T fsx = 0, fsy = 0, fsz = 0;
    const int m0 = ch * per_chunk, m1 = min(Nm, m0 + per_chunk);
    for (int m = m0 + threadIdx.x; m < m1; m += PS_BLOCK) {
        T dx = xm[3 * m] - xc, dy = xm[3 * m + 1] - yc, dz = xm[3 * m + 2] - zc;
        if (BOX) min_img<T>(s, dx, dy, dz);
        const T r2 = dx * dx + dy * dy + dz * dz;
        const T rinv = ps_rsqrt(r2);
        const T ex = dx * rinv, ey = dy * rinv, ez = dz * rinv;
        const T rinv2 = rinv * rinv, rinv4 = rinv2 * rinv2;
        const T y12b = -(T)C3D4 * rinv2;
        const T x12g = (T)C9D4 * rinv2 - gA * rinv4;
        const T y12g = gB * rinv4;
        const T fx = (T)f[3 * m], fy = (T)f[3 * m + 1], fz = (T)f[3 * m + 2];
        fsx += fx;
        fsy += fy;
        fsz += fz;
        // rotation rows: y12b (f x e)
        acc[3] += y12b * (fy * ez - fz * ey);
        acc[4] += y12b * (fz * ex - fx * ez);
        acc[5] += y12b * (fx * ey - fy * ex);
        // stresslet rows
        const T ef = ex * fx + ey * fy + ez * fz;
        const T p = -x12g * ef - (T)2 * y12g * ef;   // coefficient of e_i e_j
        // G_ij = p e_i e_j + y12g (e_i f_j + e_j f_i) + (x12g ef / 3) delta_ij
        const T G00 = p * ex * ex + (T)2 * y12g * ex * fx;
        const T G11 = p * ey * ey + (T)2 * y12g * ey * fy;
        const T G22 = p * ez * ez + (T)2 * y12g * ez * fz;
        const T G01 = p * ex * ey + y12g * (ex * fy + ey * fx);
        const T G02 = p * ex * ez + y12g * (ex * fz + ez * fx);
        const T G12 = p * ey * ez + y12g * (ey * fz + ez * fy);
        acc[6] += G00 - G22;
        acc[7] += (T)2 * G01;
        acc[8] += (T)2 * G02;
        acc[9] += (T)2 * G12;
        acc[10] += G11 - G22;
    }
    acc[0] = fsx;
    acc[1] = fsy;
    acc[2] = fsz;
    block_sum<T, 11>(acc, sh);
    if (threadIdx.x == 0) {
        T* out = part_c + ((int64_t)ch * Nc + c) * 12;
        out[0] = 0;
        out[1] = 0;
        out[2] = 0;
#pragma unroll
        for (int k = 3; k < 11; k++) out[k] = acc[k];
        if (c == 0) {
            part_fs[3 * ch] = acc[0];
            part_fs[3 * ch + 1] = acc[1];
            part_fs[3 * ch + 2] = acc[2];
        }
    }
}

template <class T, bool BOX>
__global__ __launch_bounds__(PS_BLOCK) void k_ab_coll(SysConst s, const T* __restrict__ xm,
                                                      const double* __restrict__ f, int per_chunk,
                                                      T* __restrict__ part_c, T* __restrict__ part_fs) {
    __shared__ T sh[4][11];
    ab_coll_block<T, BOX>(blockIdx.x, blockIdx.y, s, xm, f, per_chunk, part_c, part_fs, sh);
}

// The monomer-monomer tile pairs and, behind them in the same grid, the monomer rows against the
// colloid sources and the colloid rows against the monomer sources: the small colloid blocks fill
// the tail of the tile-pair blocks instead of waiting in launches of their own.  Blocks [0, nbb)
// are the colloid-pair rows (one per colloid), then nsym tile pairs (the ghost-bead pairs first),
// then nmono = gx * nchm monomer-row blocks, then Nc * nchc colloid-row blocks.
struct ApplyGrid {
    int nsym, npad, gx, nmono, ncoll, nbb;
    int per_cchunk, per_chunk;
};

// Colloid rows from colloid sources for one target colloid c (the k_bb result), with the pair tile
// applied in contracted form (cc_couple) instead of forming the 11 x 11 tile.  The translation,
// rotation and stresslet rows are summed in three passes over the partners (the pair geometry is
// recomputed in each), which keeps the registers below those of the tile pairs sharing the launch.
// fc is the colloid segment of the force vector (F 3Nc, T 3Nc, S 5Nc); the FP64 positions give the
// pair geometry.
template <int PART>
__device__ __forceinline__ void bb_pass(int c, const SysConst& s, const double* __restrict__ xm,
                                        const double* __restrict__ fc, double (&acc)[11]) {
    const int Nm = s.Nm, Nc = s.Nc;
    const int64_t nc3 = s.nc3;
    for (int o = threadIdx.x; o < Nc; o += PS_BLOCK) {
        if (o == c) continue;
        // e = (x_lo - x_hi) / r of the pair's lower and higher index; sg = +1 when c is the lower
        const bool lo = c < o;
        double e[3], dr, dr_inv;
        pair_geom_dev(s, xm, Nm + (lo ? c : o), Nm + (lo ? o : c), e, dr, dr_inv);
        const double F[3] = {fc[3 * o], fc[3 * o + 1], fc[3 * o + 2]};
        const double Tq[3] = {fc[nc3 + 3 * o], fc[nc3 + 3 * o + 1], fc[nc3 + 3 * o + 2]};
        const double S[5] = {fc[2 * nc3 + 5 * o], fc[2 * nc3 + 5 * o + 1], fc[2 * nc3 + 5 * o + 2],
                             fc[2 * nc3 + 5 * o + 3], fc[2 * nc3 + 5 * o + 4]};
        double U[3] = {0.0, 0.0, 0.0}, W[3] = {0.0, 0.0, 0.0}, Sr[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        cc_couple<PART == 2>(e, dr_inv, lo ? 1.0 : -1.0, F, Tq, S, U, W, Sr);
        if (PART == 0)
            for (int k = 0; k < 3; k++) acc[k] += U[k];
        if (PART == 1)
            for (int k = 0; k < 3; k++) acc[3 + k] += W[k];
        if (PART == 2)
            for (int k = 0; k < 5; k++) acc[6 + k] += Sr[k];
    }
}
__device__ __forceinline__ void bb_block(int c, const SysConst& s, const double* __restrict__ xm,
                                         const double* __restrict__ fc, double* __restrict__ bb, double (*sh)[11]) {
    double acc[11];
    for (int k = 0; k < 11; k++) acc[k] = 0.0;
    bb_pass<0>(c, s, xm, fc, acc);
    bb_pass<1>(c, s, xm, fc, acc);
    bb_pass<2>(c, s, xm, fc, acc);
    block_sum<double, 11>(acc, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 11; k++) bb[(int64_t)c * 11 + k] = acc[k];
}

template <class T, bool BOX>
__global__ __launch_bounds__(PS_BLOCK) void k_apply_sym(SysConst s, ApplyGrid G, const int2* __restrict__ pairs,
                                                        AAConst k, const typename AAPos<T, BOX>::P* __restrict__ xs,
                                                        const T* __restrict__ xm, const double* __restrict__ f,
                                                        const double* __restrict__ stale, T* __restrict__ part_aa,
                                                        T* __restrict__ part_abm, T* __restrict__ part_c,
                                                        T* __restrict__ part_fs, const double* __restrict__ xm64,
                                                        double* __restrict__ bb) {
    constexpr size_t SZ_AA = sizeof(AASmem<T, BOX>), SZ_M = sizeof(ABMonoSmem<T>);
    __shared__ __align__(16) unsigned char smem[SZ_AA > SZ_M ? SZ_AA : SZ_M];
    // the colloid-pair blocks come first: short and latency-bound, they run beside the first tile pairs
    if ((int)blockIdx.x < G.nbb) {
        bb_block(blockIdx.x, s, xm64, f + s.nm3, bb, reinterpret_cast<double(*)[11]>(smem));
        return;
    }
    const int b = blockIdx.x - G.nbb;
    if (b < G.nsym) {
        const int2 pr = pairs[b];
        AASmem<T, BOX>& sm = *reinterpret_cast<AASmem<T, BOX>*>(smem);
        if (b < G.npad)
            aa_sym_tile<T, BOX, true>(s.Nm, pr.x, pr.y, k, xs, f, part_aa, sm);
        else
            aa_sym_tile<T, BOX, false>(s.Nm, pr.x, pr.y, k, xs, f, part_aa, sm);
    } else if (b < G.nsym + G.nmono) {
        const int q = b - G.nsym;
        ab_mono_block<T, BOX>(q % G.gx, q / G.gx, s, xm, f, stale, G.per_cchunk, part_abm,
                              *reinterpret_cast<ABMonoSmem<T>*>(smem));
    } else {
        const int q = b - G.nsym - G.nmono;
        ab_coll_block<T, BOX>(q % s.Nc, q / s.Nc, s, xm, f, G.per_chunk, part_c, part_fs,
                              reinterpret_cast<T(*)[11]>(smem));
    }
}

// ---------------------------------------------------------------------------- BB (FP64)
// Colloid rows from colloid sources: one block per colloid, threads over partners.  fc is the colloid
// segment of a force vector: F (3Nc), T (3Nc), S (5Nc).
__global__ __launch_bounds__(PS_BLOCK) void k_bb(SysConst s, const double* __restrict__ xm,
                                                 const double* __restrict__ fc, double* __restrict__ bb) {
    __shared__ double sh[4][11];
    const int Nm = s.Nm, Nc = s.Nc;
    const int c = blockIdx.x;
    double acc[11];
    for (int k = 0; k < 11; k++) acc[k] = 0.0;
    const int64_t nc3 = s.nc3;
    for (int o = threadIdx.x; o < Nc; o += PS_BLOCK) {
        if (o == c) continue;
        const bool role1 = c < o;   // pair (min, max); c is particle "1" when it is the smaller index
        const int c1 = role1 ? c : o, c2 = role1 ? o : c;
        double e[3], dr, dr_inv;
        pair_geom_dev(s, xm, Nm + c1, Nm + c2, e, dr, dr_inv);
        BBT T;
        tile_bb(dr_inv, e, T);
        double Fo[3], To[3], So[5];
        for (int k = 0; k < 3; k++) Fo[k] = fc[3 * o + k];
        for (int k = 0; k < 3; k++) To[k] = fc[nc3 + 3 * o + k];
        for (int k = 0; k < 5; k++) So[k] = fc[2 * nc3 + 5 * o + k];
        if (role1) {
            for (int i = 0; i < 3; i++) {
                double u = 0.0, w = 0.0;
                for (int j = 0; j < 3; j++) {
                    u += T.a[i][j] * Fo[j] - T.bt[j][i] * To[j];
                    w += T.b[i][j] * Fo[j] + T.c[i][j] * To[j];
                }
                for (int j = 0; j < 5; j++) {
                    u += T.gt[i][j] * So[j];
                    w += T.ht[i][j] * So[j];
                }
                acc[i] += u;
                acc[3 + i] += w;
            }
            for (int j = 0; j < 5; j++) {
                double v = 0.0;
                for (int i = 0; i < 3; i++) v += -T.gt[i][j] * Fo[i] + T.ht[i][j] * To[i];
                for (int k = 0; k < 5; k++) v += T.m[j][k] * So[k];
                acc[6 + j] += v;
            }
        } else {
            for (int i = 0; i < 3; i++) {
                double u = 0.0, w = 0.0;
                for (int j = 0; j < 3; j++) {
                    u += T.a[j][i] * Fo[j] + T.bt[i][j] * To[j];
                    w += -T.b[j][i] * Fo[j] + T.c[j][i] * To[j];
                }
                for (int j = 0; j < 5; j++) {
                    u += -T.gt[i][j] * So[j];
                    w += T.ht[i][j] * So[j];
                }
                acc[i] += u;
                acc[3 + i] += w;
            }
            for (int j = 0; j < 5; j++) {
                double v = 0.0;
                for (int i = 0; i < 3; i++) v += T.gt[i][j] * Fo[i] + T.ht[i][j] * To[i];
                for (int k = 0; k < 5; k++) v += T.m[k][j] * So[k];
                acc[6 + j] += v;
            }
        }
    }
    block_sum<double, 11>(acc, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 11; k++) bb[(int64_t)c * 11 + k] = acc[k];
}

// ---------------------------------------------------------------------------- finalize
// Row r of u = M f from the self blocks (fill_self, src/mob.cpp:539-633) and the partial sums the
// apply kernels left: the monomer tile partials (nsplit per target), the monomer-colloid chunks, the
// colloid rows' chunks and the colloid pairs (bb).  Every partial is added in a fixed order.
template <class T>
struct FinArgs {
    const T* part_aa;
    int nsplit;
    const T* part_abm;
    int nchunk_m;
    const T* part_c;
    const T* part_fs;
    int nchunk;
    const double* bb;
    const double* stale;
};

template <class T>
__device__ __forceinline__ double finalize_row(const SysConst& s, const double* __restrict__ f, const FinArgs<T>& a,
                                               int64_t r) {
    const int64_t nm3 = s.nm3, nc3 = s.nc3, nu = s.nu;
    const int Nm = s.Nm, Nc = s.Nc;
    if (r < nm3) {
        const int i = (int)(r / 3), comp = (int)(r - 3 * (int64_t)i);
        double acc = 0.0;
        // the loads are independent: unrolled so that several are in flight (the sum order is fixed)
#pragma unroll 8
        for (int sp = 0; sp < a.nsplit; sp++) acc += (double)a.part_aa[((int64_t)sp * Nm + i) * 3 + comp];
        for (int cm = 0; cm < a.nchunk_m; cm++) acc += (double)a.part_abm[((int64_t)cm * Nm + i) * 3 + comp];
        return s.beta_inv * f[r] + acc;
    } else if (r < nm3 + nc3) {
        const int64_t q = r - nm3;
        const int c = (int)(q / 3), comp = (int)(q - 3 * (int64_t)c);
        double fs[3] = {0.0, 0.0, 0.0};
        double acc = 0.0;
        for (int ch = 0; ch < a.nchunk; ch++) {
            fs[0] += (double)a.part_fs[3 * ch];
            fs[1] += (double)a.part_fs[3 * ch + 1];
            fs[2] += (double)a.part_fs[3 * ch + 2];
        }
        for (int ch = 0; ch < a.nchunk; ch++) acc += (double)a.part_c[((int64_t)ch * Nc + c) * 12 + comp];
        // transposed carried tile: U_c[comp] += sum_jj stale[jj][comp] * sum_m f_m[jj]
        acc += a.stale[0 * 3 + comp] * fs[0] + a.stale[1 * 3 + comp] * fs[1] + a.stale[2 * 3 + comp] * fs[2];
        return 1.0 * f[r] + acc + a.bb[(int64_t)c * 11 + comp];
    } else if (r < nu) {
        const int64_t q = r - nm3 - nc3;
        const int c = (int)(q / 3), comp = (int)(q - 3 * (int64_t)c);
        double acc = 0.0;
        for (int ch = 0; ch < a.nchunk; ch++) acc += (double)a.part_c[((int64_t)ch * Nc + c) * 12 + 3 + comp];
        return (3. / 4.) * f[r] + acc + a.bb[(int64_t)c * 11 + 3 + comp];
    } else {
        const int64_t q = r - nu;
        const int c = (int)(q / 5), comp = (int)(q - 5 * (int64_t)c);
        const double* S = f + nu + 5 * (int64_t)c;
        double self = (9. / 5.) * S[comp];
        if (comp == 0) self += (9. / 10.) * S[4];
        if (comp == 4) self += (9. / 10.) * S[0];
        double acc = 0.0;
        for (int ch = 0; ch < a.nchunk; ch++) acc += (double)a.part_c[((int64_t)ch * Nc + c) * 12 + 6 + comp];
        return self + acc + a.bb[(int64_t)c * 11 + 6 + comp];
    }
}

// mode 0: u = M f;  mode 1: saddle y = A x with x = [f ; lambda] (lambda = f + n);
// rows start at r_begin (the stresslet-rows-only apply starts at nu).
template <class T>
__global__ __launch_bounds__(PS_BLOCK) void k_finalize(SysConst s, const double* __restrict__ f,
                                                       double* __restrict__ u, FinArgs<T> a, int saddle,
                                                       int64_t r_begin) {
    const int64_t r = r_begin + (int64_t)blockIdx.x * PS_BLOCK + threadIdx.x;
    const int64_t n = s.n, nu = s.nu;
    if (r >= n) return;
    double v = finalize_row<T>(s, f, a, r);
    if (saddle) {
        if (r < nu) {
            v -= f[n + r];
            u[n + r] = -f[r];
        }
    }
    u[r] = v;
}

// The Lanczos iteration's finalize: f is the previous iteration's orthogonalised vector w, not yet
// normalised.  beta^2 = the sum of the nb block partials of |w|^2 in red3 (every block alike, block 0
// stores it); y = (M f) / beta = M q and q = f / beta is stored as the basis vector.  With red3 null,
// f is already q and nothing is scaled.  The first Gram-Schmidt pass's partial dots follow in the same
// launch: red1[k * nb + block] = the block's sum of Q_k[r] y[r] for k < m, with Q_{m-1} = q.
template <class T>
__global__ __launch_bounds__(LZB) void k_finalize_lz(SysConst s, const double* __restrict__ f,
                                                     double* __restrict__ y, FinArgs<T> a,
                                                     const double* __restrict__ red3, int nb,
                                                     double* __restrict__ beta2_out, double* __restrict__ qout,
                                                     const double* __restrict__ Q, int m,
                                                     double* __restrict__ red1) {
    __shared__ double sh[64][LZW];
    __shared__ double b2s;
    const int64_t n = s.n;
    double sc = 1.0;
    if (red3) {
        // wave 0 of every block forms the same sum of the nb partials
        if (threadIdx.x < 64) {
            double v = 0.0;
            for (int i = threadIdx.x; i < nb; i += 64) v += red3[i];
            v = wave_sum64(v);
            if (threadIdx.x == 0) b2s = v;
        }
        __syncthreads();
        const double b2 = b2s;
        if (blockIdx.x == 0 && threadIdx.x == 0 && beta2_out) *beta2_out = b2;
        sc = 1.0 / sqrt(fmax(b2, 0.0));
    }
    const int64_t r = (int64_t)blockIdx.x * LZB + threadIdx.x;
    const bool in = r < n;
    double yr = 0.0, qr = 0.0;
    if (in) {
        yr = sc * finalize_row<T>(s, f, a, r);
        y[r] = yr;
        qr = sc * f[r];
        if (qout) qout[r] = qr;
    }
    const int64_t rc = in ? r : 0;
    // Q_{m-1} is q, which this thread has just stored for its row; rows past n read row 0 of Q (a
    // finite value) times yr = 0.  No branch around the load, so the eight of a batch are in flight
    // together.
    lzd_block_dots(
        m, [&](int q) { return Q[(int64_t)q * n + rc] * yr; }, sh, red1, nb);
}

// ---------------------------------------------------------------------------- carried tile
// stale[0..8]: the mob_a the AB pairs of this mob() call see; stale[9..17]: what is left for the
// next call (the last BB pair's mob_a, or this call's tile when there is no BB pair).
// `prev` is the previous mob() call's tile pair (st itself when the state has one slot).
__device__ void stale_tile(const SysConst& s, const double* __restrict__ xm, int si, int sj, const double* prev,
                           double* st) {
    double cur[9];
    if (si >= 0) {
        double e[3], dr, dr_inv;
        pair_geom_dev(s, xm, si, sj, e, dr, dr_inv);
        const double dr_inv2 = dr_inv * dr_inv, dr_inv3 = dr_inv2 * dr_inv;
        double x12a, y12a;
        if (dr > 2.0 * s.beta) {
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
        for (int k = 0; k < 9; k++) cur[k] = prev[9 + k];
    }
    for (int k = 0; k < 9; k++) st[k] = cur[k];
    if (s.Nc >= 2) {
        double e[3], dr, dr_inv;
        pair_geom_dev(s, xm, s.Np - 2, s.Np - 1, e, dr, dr_inv);
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

__global__ void k_stale(SysConst s, const double* __restrict__ xm, int si, int sj, const double* prev,
                        double* st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    stale_tile(s, xm, si, sj, prev, st);
}

// mob()'s position snapshot in one launch: xm = xpd (3 Np), its FP32 copy (mixed form), the scaled
// monomer positions of the FP64 and FP32 symmetric tile pairs (as k_aa_pos / k_aa_pos_fix), and, in
// thread 0, the carried tile (as k_stale, from the same positions).
__global__ __launch_bounds__(PS_BLOCK) void k_mob_pos(SysConst s, const double* __restrict__ xpd, double* __restrict__ xm,
                                                      float* __restrict__ xm32, double* __restrict__ xs64,
                                                      double inv_L64, void* __restrict__ xs32, int xs32_fix,
                                                      double inv_L32, int si, int sj, const double* prev,
                                                      double* st) {
    const int64_t g = (int64_t)blockIdx.x * PS_BLOCK + threadIdx.x;
    const int64_t n3 = 3 * (int64_t)s.Np;
    if (g < n3) {
        const double v = xpd[g];
        xm[g] = v;
        if (xm32) xm32[g] = (float)v;
    }
    if (g < s.Nm) {
        const int i = (int)g;
        const double p[3] = {xpd[3 * g], xpd[3 * g + 1], xpd[3 * g + 2]};
        if (xs64) {
#pragma unroll
            for (int d = 0; d < 3; d++) xs64[4 * g + d] = p[d] * inv_L64;
            xs64[4 * g + 3] = 0.0;
        }
        if (xs32 && xs32_fix) {
            int* xf = static_cast<int*>(xs32);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                double t = p[d] * inv_L32;
                t -= __builtin_rint(t);
                const long long q = __builtin_llrint(t * 4294967296.0);
                xf[4 * i + d] = (int)(unsigned)(unsigned long long)q;
            }
            xf[4 * i + 3] = 0;
        } else if (xs32) {
            float* xf = static_cast<float*>(xs32);
#pragma unroll
            for (int d = 0; d < 3; d++) xf[4 * i + d] = (float)(p[d] * inv_L32);
            xf[4 * i + 3] = 0.0f;
        }
    }
    if (g == 0) stale_tile(s, xpd, si, sj, prev, st);
}

// Dense stresslet-stresslet block M_ss (5Nc x 5Nc, row-major): self blocks (fill_self) and the
// m tiles of every colloid pair in the placement of src/mob.cpp:1129-1143.
__global__ __launch_bounds__(PS_BLOCK) void k_mss(SysConst s, const double* __restrict__ xm,
                                                  double* __restrict__ Mss) {
    const int Nm = s.Nm, Nc = s.Nc;
    const int c = blockIdx.x;
    const int64_t n5 = 5 * (int64_t)Nc;
    for (int o = threadIdx.x; o < Nc; o += PS_BLOCK) {
        double blk[5][5];
        if (o == c) {
            for (int i = 0; i < 5; i++)
                for (int j = 0; j < 5; j++)
                    blk[i][j] = (i == j) ? 9. / 5. : ((i == 0 && j == 4) || (i == 4 && j == 0)) ? 9. / 10. : 0.0;
        } else {
            const bool role1 = c < o;
            double e[3], dr, dr_inv;
            pair_geom_dev(s, xm, Nm + (role1 ? c : o), Nm + (role1 ? o : c), e, dr, dr_inv);
            BBT T;
            tile_bb(dr_inv, e, T);
            for (int i = 0; i < 5; i++)
                for (int j = 0; j < 5; j++) blk[i][j] = role1 ? T.m[i][j] : T.m[j][i];
        }
        for (int i = 0; i < 5; i++)
            for (int j = 0; j < 5; j++) Mss[(5 * (int64_t)c + i) * n5 + 5 * (int64_t)o + j] = blk[i][j];
    }
}

__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * PS_BLOCK + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

template <class T>
FinArgs<T> fin_args(const Ctx* c, const ApplyWork& aw) {
    const SysConst& s = c->sc;
    return FinArgs<T>{(const T*)aw.part_aa, s.Nm > 0 ? aw.nsplit_aa : 0, (const T*)aw.part_abm,
                      (s.Nm > 0 && s.Nc > 0) ? aw.nchunk_m : 0, (const T*)aw.part_c, (const T*)aw.part_fs,
                      (s.Nm > 0 && s.Nc > 0) ? aw.nchunk_c : 0, aw.bb, c->stale};
}

// saddle: 0 = u = M f, 1 = saddle operator, 2 = only the stresslet rows [nu, n) of M f, 3 = only the
// colloid rows [nm3, n) of M f, 4 = the apply kernels only (the partial sums; the caller finalizes).
template <class T>
ps_err launch_apply(Ctx* c, const T* xm, ApplyWork& aw, const double* f, double* u, int saddle) {
    const SysConst& s = c->sc;
    hipStream_t st = c->stream;
    hipEvent_t ea0 = nullptr, ea1 = nullptr, ek0 = nullptr, ek1 = nullptr;
    const bool srows = saddle == 2 || saddle == 3;
    if (c->prof_on && !srows) {
        PS_TRY(c->prof_record(1, &ea0, &ea1));
        PS_HIP(c, hipEventRecord(ea0, st));
    }
    const bool box = s.box_active, ch = s.chain_HI;
    // colloid pairs (FP64, every form): blocks of the fused apply launch, else a launch of their own
    const bool fused_bb = !srows && s.Nm > 0 && aw.sym;
    if (s.Nc > 0 && !fused_bb) {
        k_bb<<<s.Nc, PS_BLOCK, 0, st>>>(s, c->xm, f + s.nm3, aw.bb);
        PS_LAUNCH(c);
    }
    if (srows) {
        // the stresslet rows only: the colloid rows against the monomer sources
        if (s.Nc > 0 && s.Nm > 0) {
            dim3 g3(s.Nc, aw.nchunk_c);
            if (box)
                k_ab_coll<T, true><<<g3, PS_BLOCK, 0, st>>>(s, xm, f, aw.per_chunk_c, (T*)aw.part_c, (T*)aw.part_fs);
            else
                k_ab_coll<T, false><<<g3, PS_BLOCK, 0, st>>>(s, xm, f, aw.per_chunk_c, (T*)aw.part_c, (T*)aw.part_fs);
            PS_LAUNCH(c);
        }
    } else if (s.Nm > 0 && aw.sym) {
        // one launch: monomer tile pairs, then the monomer-colloid blocks in its tail.  Profiling:
        // the launch is bracketed by event records on the library's stream.
        if (c->prof_on) {
            PS_TRY(c->prof_record(0, &ek0, &ek1));
            PS_HIP(c, hipEventRecord(ek0, st));
        }
        AAConst k{};
        constexpr bool f32 = sizeof(T) == sizeof(float);
        // the unit of the separations: box units, or box / 2^32 for the fixed-point FP32 form
        const double L = box ? (f32 ? s.L[0] / 4294967296.0 : s.L[0]) : 1.0;
        k.c1 = 3.0 / (4.0 * L);
        k.c2 = 1.0 / (2.0 * L * L * L);
        k.c3 = -3.0 / (8.0 * L);
        k.c4 = -3.0 / (2.0 * L * L * L);
        k.L = L;
        k.tb2s = s.two_beta * s.two_beta / (L * L);
        // ghost beads sit at the origin with f = 0: the clamp keeps every coefficient finite
        k.clamp = f32 ? (box ? 1.0 : 1e-30) : 1e-200;
        ApplyGrid G{};
        G.nsym = aw.npair_full + aw.npair_pad;
        G.npad = aw.npair_pad;
        G.gx = (s.Nm + PS_BLOCK - 1) / PS_BLOCK;
        G.nmono = s.Nc > 0 ? G.gx * aw.nchunk_m : 0;
        G.ncoll = s.Nc > 0 ? s.Nc * aw.nchunk_c : 0;
        G.nbb = s.Nc > 0 ? s.Nc : 0;
        G.per_cchunk = aw.per_chunk_m;
        G.per_chunk = aw.per_chunk_c;
        const unsigned nblk = (unsigned)(G.nsym + G.nmono + G.ncoll + G.nbb);
        T* pa = (T*)aw.part_aa;
        T* pm = (T*)aw.part_abm;
        T* pc = (T*)aw.part_c;
        T* pf = (T*)aw.part_fs;
        using PB = const typename AAPos<T, true>::P*;
        using PN = const typename AAPos<T, false>::P*;
        if (box)
            k_apply_sym<T, true><<<nblk, PS_BLOCK, 0, st>>>(s, G, aw.pairs, k, (PB)aw.xs, xm, f, c->stale, pa, pm, pc,
                                                            pf, c->xm, aw.bb);
        else
            k_apply_sym<T, false><<<nblk, PS_BLOCK, 0, st>>>(s, G, aw.pairs, k, (PN)aw.xs, xm, f, c->stale, pa, pm,
                                                             pc, pf, c->xm, aw.bb);
        PS_LAUNCH(c);
        if (c->prof_on) PS_HIP(c, hipEventRecord(ek1, st));
    } else {
        // the ordered monomer kernel (chain_HI, non-cubic boxes, small systems), then the
        // monomer-colloid kernels
        if (s.Nm > 0) {
            if (c->prof_on) PS_TRY(c->prof_record(0, &ek0, &ek1));
            if (c->prof_on) PS_HIP(c, hipEventRecord(ek0, st));
            dim3 g1((s.Nm + PS_BLOCK - 1) / PS_BLOCK, aw.nsplit_aa);
            if (box && ch)
                k_aa<T, true, true><<<g1, PS_BLOCK, 0, st>>>(s, xm, f, c->chain, aw.per_split_aa, (T*)aw.part_aa);
            else if (box)
                k_aa<T, true, false><<<g1, PS_BLOCK, 0, st>>>(s, xm, f, c->chain, aw.per_split_aa, (T*)aw.part_aa);
            else if (ch)
                k_aa<T, false, true><<<g1, PS_BLOCK, 0, st>>>(s, xm, f, c->chain, aw.per_split_aa, (T*)aw.part_aa);
            else
                k_aa<T, false, false><<<g1, PS_BLOCK, 0, st>>>(s, xm, f, c->chain, aw.per_split_aa, (T*)aw.part_aa);
            PS_LAUNCH(c);
            if (c->prof_on) PS_HIP(c, hipEventRecord(ek1, st));
            if (s.Nc > 0) {
                dim3 g2((s.Nm + PS_BLOCK - 1) / PS_BLOCK, aw.nchunk_m);
                if (box)
                    k_ab_mono<T, true><<<g2, PS_BLOCK, 0, st>>>(s, xm, f, c->stale, aw.per_chunk_m, (T*)aw.part_abm);
                else
                    k_ab_mono<T, false><<<g2, PS_BLOCK, 0, st>>>(s, xm, f, c->stale, aw.per_chunk_m, (T*)aw.part_abm);
                PS_LAUNCH(c);
                dim3 g3(s.Nc, aw.nchunk_c);
                if (box)
                    k_ab_coll<T, true><<<g3, PS_BLOCK, 0, st>>>(s, xm, f, aw.per_chunk_c, (T*)aw.part_c,
                                                                (T*)aw.part_fs);
                else
                    k_ab_coll<T, false><<<g3, PS_BLOCK, 0, st>>>(s, xm, f, aw.per_chunk_c, (T*)aw.part_c,
                                                                 (T*)aw.part_fs);
                PS_LAUNCH(c);
            }
        }
    }
    if (saddle == 4) {
        if (c->prof_on) PS_HIP(c, hipEventRecord(ea1, st));
        c->cnt.applies++;
        return PS_OK;
    }
    const int64_t r_begin = saddle == 2 ? s.nu : saddle == 3 ? s.nm3 : 0;
    if (s.n - r_begin <= 0) return PS_OK;
    dim3 g4((unsigned)((s.n - r_begin + PS_BLOCK - 1) / PS_BLOCK));
    k_finalize<T><<<g4, PS_BLOCK, 0, st>>>(s, f, u, fin_args<T>(c, aw), srows ? 0 : saddle, r_begin);
    PS_LAUNCH(c);
    if (srows) return PS_OK;
    if (c->prof_on) PS_HIP(c, hipEventRecord(ea1, st));
    c->cnt.applies++;
    return PS_OK;
}

// The ordered kernel's default source split, in 256-bead tiles: about 4 blocks per CU.
int default_split_tiles(const Ctx* c) {
    const SysConst& s = c->sc;
    const int ntb = (s.Nm + PS_BLOCK - 1) / PS_BLOCK;
    const int target = 4 * (c->n_cu > 0 ? c->n_cu : 256);
    int nsplit = ntb > 0 ? (target + ntb - 1) / ntb : 1;
    nsplit = std::max(1, std::min(nsplit, std::max(1, s.Nm / 64)));
    const int per = s.Nm > 0 ? (s.Nm + nsplit - 1) / nsplit : 0;
    return std::max(1, (per + PS_BLOCK - 1) / PS_BLOCK);
}

// split_tiles: the ordered kernel's source split in whole 256-bead tiles (0: the default).
ps_err alloc_work(Ctx* c, ApplyWork& aw, size_t elt, bool sym, int split_tiles = 0) {
    const SysConst& s = c->sc;
    const int ntb = (s.Nm + PS_BLOCK - 1) / PS_BLOCK;
    // source splits are whole 256-bead tiles (k_aa relies on tile-aligned sources)
    const int target = 4 * (c->n_cu > 0 ? c->n_cu : 256);
    const int tiles = std::max(1, std::min(split_tiles > 0 ? split_tiles : default_split_tiles(c), std::max(ntb, 1)));
    aw.per_split_aa = s.Nm > 0 ? tiles * PS_BLOCK : 0;
    aw.nsplit_aa = s.Nm > 0 ? (s.Nm + aw.per_split_aa - 1) / aw.per_split_aa : 1;
    int nchunk = s.Nc > 0 ? (2 * (c->n_cu > 0 ? c->n_cu : 256) + s.Nc - 1) / s.Nc : 1;
    nchunk = std::max(1, std::min(nchunk, std::max(1, s.Nm / PS_BLOCK)));
    aw.per_chunk_c = s.Nm > 0 ? (s.Nm + nchunk - 1) / nchunk : 0;
    aw.nchunk_c = s.Nm > 0 ? (s.Nm + aw.per_chunk_c - 1) / aw.per_chunk_c : 1;
    aw.sym = sym;
    if (aw.sym) {
        aw.ntile = ntb;
        aw.nsplit_aa = ntb;
        aw.inv_L = s.box_active ? 1.0 / s.L[0] : 1.0;
        const bool partial = (s.Nm % PS_BLOCK) != 0;
        std::vector<int2> full, pad;
        for (int I = 0; I < ntb; I++)
            for (int J = I; J < ntb; J++) {
                const bool last = partial && (J == ntb - 1);
                (last ? pad : full).push_back(make_int2(I, J));
            }
        aw.npair_full = (int)full.size();
        aw.npair_pad = (int)pad.size();
        pad.insert(pad.end(), full.begin(), full.end());   // ghost-bead pairs first
        PS_HIP(c, hipMalloc(&aw.pairs, sizeof(int2) * pad.size()));
        PS_HIP(c, hipMemcpy(aw.pairs, pad.data(), sizeof(int2) * pad.size(), hipMemcpyHostToDevice));
        PS_HIP(c, hipMalloc(&aw.xs, 4 * elt * (size_t)s.Nm));
    }
    aw.bytes_aa = (size_t)aw.nsplit_aa * 3 * std::max(s.Nm, 1) * elt;
    // monomer rows from colloid sources: chunks of >= 32 colloids on blockIdx.y
    int nchm = std::max(1, std::min((s.Nc + 31) / 32, ntb > 0 ? (target + ntb - 1) / ntb : 1));
    aw.per_chunk_m = s.Nc > 0 ? (s.Nc + nchm - 1) / nchm : 1;
    aw.nchunk_m = s.Nc > 0 ? (s.Nc + aw.per_chunk_m - 1) / aw.per_chunk_m : 1;
    aw.bytes_abm = (size_t)aw.nchunk_m * 3 * std::max(s.Nm, 1) * elt;
    aw.bytes_c = (size_t)aw.nchunk_c * std::max(s.Nc, 1) * 12 * elt;
    aw.bytes_fs = (size_t)aw.nchunk_c * 3 * elt;
    PS_HIP(c, hipMalloc(&aw.part_aa, aw.bytes_aa));
    PS_HIP(c, hipMalloc(&aw.part_abm, aw.bytes_abm));
    PS_HIP(c, hipMalloc(&aw.part_c, aw.bytes_c));
    PS_HIP(c, hipMalloc(&aw.part_fs, aw.bytes_fs));
    PS_HIP(c, hipMalloc(&aw.bb, sizeof(double) * 11 * std::max(s.Nc, 1)));
    PS_HIP(c, hipMemsetAsync(aw.part_c, 0, aw.bytes_c, c->stream));
    PS_HIP(c, hipMemsetAsync(aw.part_fs, 0, aw.bytes_fs, c->stream));
    PS_HIP(c, hipMemsetAsync(aw.bb, 0, sizeof(double) * 11 * std::max(s.Nc, 1), c->stream));
    return PS_OK;
}

void free_work(ApplyWork& aw) {
    for (void* p : {aw.part_aa, aw.part_abm, aw.part_c, aw.part_fs, (void*)aw.bb, (void*)aw.pairs, aw.xs})
        if (p) (void)hipFree(p);
    aw = ApplyWork();
}

}  // namespace

// ---------------------------------------------------------------------------- kernel choice
// PS_AA_AUTO: where both monomer kernels apply (a cubic box or none, no chain_HI) the choice is
// measured, not ruled: at create, each kernel runs a few applies on synthetic positions of the
// system's size and density (uniform in the box, or in a cube of the box's density without one) and
// the faster is kept.  The outcome is cached per size class (monomers, colloids, precision, box, CUs)
// for the life of the process, so later contexts of the same size do not time again.
namespace {

struct AATuneKey {
    int Nm, Nc, elt, box, ncu;
    bool operator<(const AATuneKey& o) const {
        return std::tie(Nm, Nc, elt, box, ncu) < std::tie(o.Nm, o.Nc, o.elt, o.box, o.ncu);
    }
};
struct AATuneVal {
    bool sym;          // the symmetric kernel is the faster
    int split_tiles;   // else the ordered kernel's fastest source split (256-bead tiles)
    float ms;          // the chosen layout's ms per apply
};
std::map<AATuneKey, AATuneVal> g_aa_tune;
std::mutex g_aa_tune_mu;

// synthetic monomer and colloid positions of the system's size, seeded (the apply times do not depend
// on where the beads are, only on how many pairs fall in the short-range branch)
std::vector<double> tune_positions(const Ctx* c) {
    const SysConst& s = c->sc;
    // no box: a cube with about one bead per (2 radii)^3
    const double side = s.box_active ? s.L[0] : 2.0 * std::cbrt((double)std::max(1, s.Np));
    std::vector<double> x(3 * (size_t)s.Np);
    uint64_t st = 0x9e3779b97f4a7c15ull;
    for (auto& v : x) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        v = ((double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5) * side;
    }
    return x;
}

template <class T>
ps_err time_apply(Ctx* c, const T* xm, ApplyWork& aw, int reps, float* ms) {
    const unsigned gm = (unsigned)((c->Nm + PS_BLOCK - 1) / PS_BLOCK);
    if (aw.sym && sizeof(T) == sizeof(float) && c->sc.box_active)
        k_aa_pos_fix<<<gm, PS_BLOCK, 0, c->stream>>>(c->Nm, c->xm, aw.inv_L, (int*)aw.xs);
    else if (aw.sym)
        k_aa_pos<T><<<gm, PS_BLOCK, 0, c->stream>>>(c->Nm, c->xm, aw.inv_L, (T*)aw.xs);
    PS_LAUNCH(c);
    hipEvent_t e0, e1;
    PS_HIP(c, hipEventCreate(&e0));
    PS_HIP(c, hipEventCreate(&e1));
    ps_err rc = launch_apply<T>(c, xm, aw, c->io_a, c->io_b, 0);   // warmup
    if (rc == PS_OK) rc = hipEventRecord(e0, c->stream) == hipSuccess ? PS_OK : PS_ERR_HIP;
    for (int k = 0; k < reps && rc == PS_OK; k++) rc = launch_apply<T>(c, xm, aw, c->io_a, c->io_b, 0);
    if (rc == PS_OK) rc = hipEventRecord(e1, c->stream) == hipSuccess ? PS_OK : PS_ERR_HIP;
    if (rc == PS_OK) rc = hipEventSynchronize(e1) == hipSuccess ? PS_OK : PS_ERR_HIP;
    float t = 0.f;
    if (rc == PS_OK) rc = hipEventElapsedTime(&t, e0, e1) == hipSuccess ? PS_OK : PS_ERR_HIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == PS_ERR_HIP) c->set_error("apply timing for the kernel choice failed");
    *ms = t / reps;
    return rc;
}

// aw <- the fastest layout for precision T: the symmetric kernel (when `sym_ok`) and the ordered kernel
// at source splits of half, one and two times its default length (the tile shape of its blocks), each
// allocated and timed on the synthetic system, the slower ones freed.  The choice is cached per size
// class for the life of the process.
template <class T>
ps_err choose_kernel(Ctx* c, ApplyWork& aw, const T* xm, bool sym_ok) {
    const AATuneKey key{c->Nm, c->Nc, (int)sizeof(T), c->sc.box_active | (sym_ok ? 0 : 2), c->n_cu};
    {
        std::lock_guard<std::mutex> lk(g_aa_tune_mu);
        auto it = g_aa_tune.find(key);
        if (it != g_aa_tune.end()) {
            c->aa_tuned_ms[sizeof(T) == sizeof(float)] = it->second.ms;
            return alloc_work(c, aw, sizeof(T), it->second.sym, it->second.split_tiles);
        }
    }
    const int ntb = std::max(1, (c->Nm + PS_BLOCK - 1) / PS_BLOCK);
    const int d = default_split_tiles(c);
    std::vector<int> splits;
    for (int t : {d, std::max(1, d / 2), std::min(ntb, 2 * d)})
        if (std::find(splits.begin(), splits.end(), t) == splits.end()) splits.push_back(t);
    ApplyWork best{};
    AATuneVal bv{false, 0, 0.f};
    bool have = false;
    ps_err rc = PS_OK;
    const int64_t cnt0 = c->cnt.applies;
    auto consider = [&](bool sym, int split) {
        if (rc != PS_OK) return;
        ApplyWork w{};
        rc = alloc_work(c, w, sizeof(T), sym, split);
        float t = 0.f;
        if (rc == PS_OK) rc = time_apply<T>(c, xm, w, 3, &t);
        if (rc != PS_OK) {
            free_work(w);
            return;
        }
        if (!have || t < bv.ms) {
            if (have) free_work(best);
            best = w;
            bv = AATuneVal{sym, sym ? 0 : split, t};
            have = true;
        } else {
            free_work(w);
        }
    };
    for (int t : splits) consider(false, t);
    if (sym_ok) consider(true, 0);
    c->cnt.applies = cnt0;
    if (rc != PS_OK) {
        if (have) free_work(best);
        return rc;
    }
    aw = best;
    c->aa_tuned_ms[sizeof(T) == sizeof(float)] = bv.ms;
    std::lock_guard<std::mutex> lk(g_aa_tune_mu);
    g_aa_tune[key] = bv;
    return PS_OK;
}

}  // namespace

ps_err apply_alloc(Ctx* c) {
    const SysConst& s = c->sc;
    const bool cubic = !s.box_active || (s.L[0] == s.L[1] && s.L[1] == s.L[2]);
    // the symmetric kernel needs a cubic box (or none) and no same-chain skip
    const bool eligible = s.Nm > 0 && cubic && !s.chain_HI;
    const int sel = c->p.aa_kernel;
    const bool mixed = c->p.precision == PS_PREC_MIXED;
    if (mixed) PS_HIP(c, hipMalloc(&c->xm32, sizeof(float) * 3 * (size_t)c->Np));
    if (c->arrow || s.Nm <= 0 || sel != PS_AA_AUTO) {   // the arrowhead regime applies no tiles
        const bool sym = eligible && sel == PS_AA_SYMMETRIC;
        PS_TRY(alloc_work(c, c->aw64, sizeof(double), sym));
        if (mixed) PS_TRY(alloc_work(c, c->aw32, sizeof(float), sym));
        return PS_OK;
    }
    // measured choice on synthetic positions (c->xm is overwritten by the first mob())
    const std::vector<double> xt = tune_positions(c);
    PS_HIP(c, hipMemcpy(c->xm, xt.data(), sizeof(double) * xt.size(), hipMemcpyHostToDevice));
    std::vector<double> ft(c->n);
    uint64_t st = 12345;
    for (auto& v : ft) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        v = (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
    PS_HIP(c, hipMemcpy(c->io_a, ft.data(), sizeof(double) * ft.size(), hipMemcpyHostToDevice));
    PS_TRY(choose_kernel<double>(c, c->aw64, c->xm, eligible));
    if (mixed) {
        const int64_t n3 = 3 * (int64_t)c->Np;
        k_to_f32<<<(unsigned)((n3 + PS_BLOCK - 1) / PS_BLOCK), PS_BLOCK, 0, c->stream>>>(c->xm, c->xm32, n3);
        PS_LAUNCH(c);
        PS_TRY(choose_kernel<float>(c, c->aw32, c->xm32, eligible));
    }
    PS_HIP(c, hipStreamSynchronize(c->stream));
    return PS_OK;
}

void apply_free(Ctx* c) {
    free_work(c->aw64);
    free_work(c->aw32);
    if (c->xm32) (void)hipFree(c->xm32);
    c->xm32 = nullptr;
}

// mob(): snapshot the positions of the last check_dist and compute the carried translation tile.
ps_err mob_prepare(Ctx* c) {
    if (c->arrow) {
        c->cnt.mob_calls++;
        return arrow_mob(c);
    }
    {
        // xm, its FP32 copy, the tile pairs' scaled positions and the carried tile in one launch
        const int64_t n3 = 3 * (int64_t)c->Np;
        const unsigned gp = (unsigned)((n3 + PS_BLOCK - 1) / PS_BLOCK);
        const bool fix32 = c->aw32.sym && c->sc.box_active;
        k_mob_pos<<<gp, PS_BLOCK, 0, c->stream>>>(c->sc, c->xpd, c->xm, c->xm32,
                                                  c->aw64.sym ? (double*)c->aw64.xs : nullptr, c->aw64.inv_L,
                                                  c->aw32.sym ? c->aw32.xs : nullptr, fix32 ? 1 : 0, c->aw32.inv_L,
                                                  c->stale_i, c->stale_j, c->stale_prev ? c->stale_prev : c->stale,
                                                  c->stale);
        PS_LAUNCH(c);
    }
    c->stale_prev = c->stale;
    if (c->Mss && c->Nc > 0 && !c->p.lub) {   // the lubricated block solve applies M_cc on the fly
        k_mss<<<c->Nc, PS_BLOCK, 0, c->stream>>>(c->sc, c->xm, c->Mss);
        PS_LAUNCH(c);
    }
    // lubrication: the colloid pairs at these positions (one host round trip)
    if (c->p.lub) PS_TRY(lub_pairs(c, c->p.lub_rcut, nullptr));
    c->cnt.mob_calls++;
    return PS_OK;
}

// ---------------------------------------------------------------------------- mob() slots
// The joint step keeps the mob() states of the predictor point and of both RFD probe points alive
// at once.  Slot 0 is the context's own buffers; slots 1 and 2 get their own copies of everything a
// mob() call writes: the positions snapshot (FP64 and FP32), the kernel-layout positions, the carried
// tile pair and the stresslet block.
ps_err mob_slots_alloc(Ctx* c) {
    Ctx::MobSlot& s0 = c->mslot[0];
    s0.xm = c->xm;
    s0.xm32 = c->xm32;
    s0.xs64 = c->aw64.sym ? c->aw64.xs : nullptr;
    s0.xs32 = c->aw32.sym ? c->aw32.xs : nullptr;
    s0.stale = c->stale;
    s0.Mss = c->Mss;
    const size_t n3 = 3 * (size_t)c->Np, nm4 = 4 * (size_t)std::max(c->Nm, 1), n5 = 5 * (size_t)std::max(c->Nc, 1);
    for (int k = 1; k < 3; k++) {
        Ctx::MobSlot& m = c->mslot[k];
        PS_HIP(c, hipMalloc(&m.xm, sizeof(double) * n3));
        if (s0.xm32) PS_HIP(c, hipMalloc(&m.xm32, sizeof(float) * n3));
        if (s0.xs64) PS_HIP(c, hipMalloc(&m.xs64, sizeof(double) * nm4));
        if (s0.xs32) PS_HIP(c, hipMalloc(&m.xs32, sizeof(float) * nm4));
        PS_HIP(c, hipMalloc(&m.stale, sizeof(double) * 18));
        PS_HIP(c, hipMemcpy(m.stale, s0.stale, sizeof(double) * 18, hipMemcpyDeviceToDevice));
        if (s0.Mss) PS_HIP(c, hipMalloc(&m.Mss, sizeof(double) * n5 * n5));
    }
    c->nmslot = 3;
    return PS_OK;
}

void mob_use_slot(Ctx* c, int k) {
    if (k < 0 || k >= c->nmslot) return;
    const Ctx::MobSlot& m = c->mslot[k];
    c->xm = m.xm;
    c->xm32 = m.xm32;
    if (c->aw64.sym) c->aw64.xs = m.xs64;
    if (c->aw32.sym) c->aw32.xs = m.xs32;
    c->stale = m.stale;
    c->Mss = m.Mss;
}

void mob_slots_free(Ctx* c) {
    if (c->nmslot > 0) mob_use_slot(c, 0);
    for (int k = 1; k < 3; k++) {
        Ctx::MobSlot& m = c->mslot[k];
        for (void* p : {(void*)m.xm, (void*)m.xm32, m.xs64, m.xs32, (void*)m.stale, (void*)m.Mss})
            if (p) (void)hipFree(p);
        m = Ctx::MobSlot();
    }
    c->nmslot = 0;
    c->stale_prev = c->stale;
}

ps_err mob_apply(Ctx* c, const double* f, double* u, int prec) {
    if (c->arrow) return arrow_apply(c, f, u, 0);
    if (prec == PS_PREC_MIXED && c->xm32) return launch_apply<float>(c, c->xm32, c->aw32, f, u, 0);
    return launch_apply<double>(c, c->xm, c->aw64, f, u, 0);
}

// The Lanczos iteration's apply: y = M (f / beta) and q = f / beta with beta^2 the sum of the nb
// partials in red3 (stored to *beta2_out), or y = M f when red3 is null, then the first Gram-Schmidt
// pass's partial dots against Q_0..Q_{m-1} into red1 (k_finalize_lz; nb = ceil(n / LZB)).
template <class T>
static ps_err apply_lanczos(Ctx* c, const T* xm, ApplyWork& aw, const double* f, double* y, double* qout,
                            const double* red3, int nb, double* beta2_out, int m, double* red1) {
    PS_TRY(launch_apply<T>(c, xm, aw, f, nullptr, 4));
    const unsigned g = (unsigned)lzd_blocks(c->n);
    k_finalize_lz<T><<<g, LZB, 0, c->stream>>>(c->sc, f, y, fin_args<T>(c, aw), red3, nb, beta2_out, qout, c->Q, m,
                                               red1);
    PS_LAUNCH(c);
    return PS_OK;
}
ps_err mob_apply_lanczos(Ctx* c, const double* f, double* y, double* qout, const double* red3, int nb,
                         double* beta2_out, int m, double* red1, int prec) {
    if (prec == PS_PREC_MIXED && c->xm32)
        return apply_lanczos<float>(c, c->xm32, c->aw32, f, y, qout, red3, nb, beta2_out, m, red1);
    return apply_lanczos<double>(c, c->xm, c->aw64, f, y, qout, red3, nb, beta2_out, m, red1);
}

ps_err mob_apply_stresslet_rows(Ctx* c, const double* f, double* u) {
    if (c->arrow) return arrow_apply(c, f, u, 0);   // every row: O(N) anyway
    return launch_apply<double>(c, c->xm, c->aw64, f, u, 2);
}

ps_err mob_apply_colloid_rows(Ctx* c, const double* f, double* u) {
    return launch_apply<double>(c, c->xm, c->aw64, f, u, 3);
}

ps_err saddle_apply(Ctx* c, const double* x, double* y, int prec) {
    if (c->arrow) return arrow_apply(c, x, y, 1);
    if (prec == PS_PREC_MIXED && c->xm32) PS_TRY(launch_apply<float>(c, c->xm32, c->aw32, x, y, 1));
    else PS_TRY(launch_apply<double>(c, c->xm, c->aw64, x, y, 1));
    // lubrication: the lower-right block -R_lub on the colloid velocities (FP64 in every form)
    if (c->p.lub) PS_TRY(lub_saddle_term(c, x, y));
    return PS_OK;
}

namespace {
// The colloid saddle of the lubricated block solve, v = [z (11Nc: F, T, S); w (6Nc: U, W)]:
//   y_z = M_cc z - [w; 0],  y_w = -z_FT - R_lub w,
// with M_cc the colloid-colloid block of M (self blocks of fill_self, src/mob.cpp:600-612, and the
// BB tiles in bb) and fts = R_lub w.
__global__ void k_cs_combine(int Nc, const double* __restrict__ v, const double* __restrict__ bb,
                             const double* __restrict__ fts, double* __restrict__ y) {
    const int64_t r = (int64_t)blockIdx.x * PS_BLOCK + threadIdx.x;
    const int64_t nc3 = 3 * (int64_t)Nc, nc6 = 6 * (int64_t)Nc, nc11 = 11 * (int64_t)Nc;
    if (r >= nc11 + nc6) return;
    const double* w = v + nc11;
    double out;
    if (r < nc3) {
        const int c = (int)(r / 3), comp = (int)(r % 3);
        out = v[r] + bb[11 * (int64_t)c + comp] - w[r];
    } else if (r < nc6) {
        const int c = (int)((r - nc3) / 3), comp = (int)((r - nc3) % 3);
        out = 0.75 * v[r] + bb[11 * (int64_t)c + 3 + comp] - w[r];
    } else if (r < nc11) {
        const int c = (int)((r - nc6) / 5), comp = (int)((r - nc6) % 5);
        const double* S = v + nc6 + 5 * (int64_t)c;
        double self = (9. / 5.) * S[comp];
        if (comp == 0) self += (9. / 10.) * S[4];
        if (comp == 4) self += (9. / 10.) * S[0];
        out = self + bb[11 * (int64_t)c + 6 + comp];
    } else {
        const int64_t k = r - nc11;
        out = -v[k] - fts[k];
    }
    y[r] = out;
}
}  // namespace

ps_err colloid_saddle_apply(Ctx* c, const double* v, double* y) {
    const int Nc = c->Nc;
    const size_t ncb = (size_t)std::max(Nc, 1);
    double* fts = c->cs_buf + 51 * ncb;
    double* bb = c->cs_buf + 62 * ncb;
    hipStream_t st = c->stream;
    k_bb<<<Nc, PS_BLOCK, 0, st>>>(c->sc, c->xm, v, bb);
    PS_LAUNCH(c);
    PS_TRY(lub_apply_dev(c, v + 11 * (int64_t)Nc, nullptr, fts));
    const int64_t N = 17 * (int64_t)Nc;
    k_cs_combine<<<(unsigned)((N + PS_BLOCK - 1) / PS_BLOCK), PS_BLOCK, 0, st>>>(Nc, v, bb, fts, y);
    PS_LAUNCH(c);
    return PS_OK;
}

// ---------------------------------------------------------------------------- no-allocation apply
// u = M f at positions x with no context and no allocation: the partial sums, the carried tile and
// the chain ids live in one file-scope device array (each GPU the library runs on holds its own copy
// until the process ends).  Host code never uses the array's name as an address: its device address
// is looked up once per device and kept.
namespace {

constexpr size_t NOALLOC_DOUBLES = 2u << 20;   // 16 MiB
__device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEV = 64;
double* g_scratch_ptr[NOALLOC_MAX_DEV] = {};
std::mutex g_scratch_mu;

// stale[9..17] starts as the identity (the mob_a fill_self leaves, src/mob.cpp:65-86) for a system
// with no processed monomer pair; the colloid-row partials start at zero; chain ids for chain_HI.
__global__ void k_noalloc_init(double* __restrict__ st, int* __restrict__ chain, int Nm, int nmono,
                               double* __restrict__ zero, int64_t nzero) {
    const int64_t t = (int64_t)blockIdx.x * PS_BLOCK + threadIdx.x;
    if (t < 18) st[t] = (t >= 9 && ((t - 9) % 4 == 0)) ? 1.0 : 0.0;
    if (t < Nm) chain[t] = nmono > 0 ? (int)(t / nmono) : 0;
    if (t < nzero) zero[t] = 0.0;
}

}  // namespace

extern "C" ps_err ps_mob_mult_noalloc(const ps_params* p, const double* x, const double* f, double* u,
                                      void* stream) {
    if (!p || !x || !f || !u) {
        ps_set_context_free_error("ps_mob_mult_noalloc: null argument");
        return PS_ERR_ARG;
    }
    if (!p->mm_HI || p->Np != p->Nm + p->Nc || p->Np < 2 || p->Nm < 0 || p->Nc < 0 || !(p->beta > 0)) {
        ps_set_context_free_error("ps_mob_mult_noalloc: needs mm_HI = 1, Np = Nm + Nc >= 2 and beta > 0");
        return PS_ERR_ARG;
    }
    if (p->box_active && (p->box[0] <= 0 || p->box[1] <= 0 || p->box[2] <= 0)) {
        ps_set_context_free_error("ps_mob_mult_noalloc: box lengths must be positive");
        return PS_ERR_ARG;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEV) {
        ps_set_context_free_error("ps_mob_mult_noalloc: no usable current device");
        return PS_ERR_NODEV;
    }
    double* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {   // first call on this device, with it current
            void* ptr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr) {
                ps_set_context_free_error(std::string("ps_mob_mult_noalloc: hipGetSymbolAddress: ") +
                                          hipGetErrorString(e));
                return PS_ERR_HIP;
            }
            g_scratch_ptr[dev] = static_cast<double*>(ptr);
        }
        base = g_scratch_ptr[dev];
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipDeviceProp_t prop;
    int ncu = 256;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) ncu = prop.multiProcessorCount;
    SysConst s{};
    s.Np = p->Np;
    s.Nm = p->Nm;
    s.Nc = p->Nc;
    s.chain_HI = p->chain_HI;
    s.box_active = p->box_active;
    for (int d = 0; d < 3; d++) {
        s.L[d] = p->box_active ? p->box[d] : 0.0;
        s.Linv[d] = p->box_active ? 1.0 / p->box[d] : 0.0;
    }
    s.beta = p->beta;
    s.beta2 = p->beta * p->beta;
    s.beta_inv = 1.0 / p->beta;
    s.two_beta = 2.0 * p->beta;
    s.n = 3 * (int64_t)s.Nm + 11 * (int64_t)s.Nc;
    s.nu = 3 * (int64_t)s.Nm + 6 * (int64_t)s.Nc;
    s.ns = s.n + s.nu;
    s.nm3 = 3 * (int64_t)s.Nm;
    s.nc3 = 3 * (int64_t)s.Nc;
    const int Nm = s.Nm, Nc = s.Nc;
    // the last processed monomer pair of mob() (src/mob.cpp:845-897): chain ids i / Nmono_per_chain
    const int nmono = p->Nmono_per_chain > 0 ? p->Nmono_per_chain : Nm;
    int si = -1, sj = -1;
    for (int i = Nm - 2; i >= 0 && si < 0; i--)
        for (int j = Nm - 1; j > i; j--) {
            if (p->chain_HI && i / nmono == j / nmono) continue;
            si = i;
            sj = j;
            break;
        }
    // splits: the ordered monomer kernel over source splits of whole 256-bead tiles, the colloid
    // chunks as in a context; all partials in the scratch array
    const int ntb = (Nm + PS_BLOCK - 1) / PS_BLOCK;
    const int target = 4 * ncu;
    int nchm = Nc > 0 ? std::max(1, std::min((Nc + 31) / 32, ntb > 0 ? (target + ntb - 1) / ntb : 1)) : 1;
    const int per_chunk_m = Nc > 0 ? (Nc + nchm - 1) / nchm : 1;
    nchm = Nc > 0 ? (Nc + per_chunk_m - 1) / per_chunk_m : 1;
    int nchc = Nc > 0 ? (2 * ncu + Nc - 1) / Nc : 1;
    nchc = std::max(1, std::min(nchc, std::max(1, Nm / PS_BLOCK)));
    const int per_chunk_c = Nm > 0 ? (Nm + nchc - 1) / nchc : 0;
    nchc = Nm > 0 ? (Nm + per_chunk_c - 1) / per_chunk_c : 1;
    auto al = [](size_t v) { return (v + 31) / 32 * 32; };
    const size_t off_chain = 32, n_chain = al(((size_t)std::max(Nm, 1) + 1) / 2);
    const size_t off_c = off_chain + n_chain, n_c = al((size_t)nchc * std::max(Nc, 1) * 12);
    const size_t off_fs = off_c + n_c, n_fs = al((size_t)nchc * 3);
    const size_t off_bb = off_fs + n_fs, n_bb = al((size_t)11 * std::max(Nc, 1));
    const size_t off_abm = off_bb + n_bb, n_abm = al((size_t)nchm * 3 * std::max(Nm, 1));
    const size_t off_aa = off_abm + n_abm;
    if (off_aa + 3 * (size_t)std::max(Nm, 1) > NOALLOC_DOUBLES) {
        ps_set_context_free_error("ps_mob_mult_noalloc: the system needs more than the " +
                                  std::to_string(NOALLOC_DOUBLES * 8 >> 20) + " MiB scratch array (Nm = " +
                                  std::to_string(Nm) + ", Nc = " + std::to_string(Nc) + ")");
        return PS_ERR_ARG;
    }
    const size_t room = (NOALLOC_DOUBLES - off_aa) / (3 * (size_t)std::max(Nm, 1));
    int nsplit = ntb > 0 ? (target + ntb - 1) / ntb : 1;
    nsplit = std::max(1, std::min({nsplit, std::max(1, Nm / 64), (int)std::min<size_t>(room, 1 << 20)}));
    int per_split = Nm > 0 ? (Nm + nsplit - 1) / nsplit : 0;
    per_split = ((per_split + PS_BLOCK - 1) / PS_BLOCK) * PS_BLOCK;
    nsplit = Nm > 0 ? (Nm + per_split - 1) / per_split : 1;
    double* stale = base;
    int* chain = reinterpret_cast<int*>(base + off_chain);
    double* part_c = base + off_c;
    double* part_fs = base + off_fs;
    double* bb = base + off_bb;
    double* part_abm = base + off_abm;
    double* part_aa = base + off_aa;
    const int64_t nzero = (int64_t)(off_abm - off_c);   // colloid partials and bb
    const int64_t ninit = std::max<int64_t>({18, (int64_t)Nm, nzero});
    k_noalloc_init<<<(unsigned)((ninit + PS_BLOCK - 1) / PS_BLOCK), PS_BLOCK, 0, st>>>(stale, chain, Nm, nmono,
                                                                                      part_c, nzero);
    k_stale<<<1, 64, 0, st>>>(s, x, si, sj, stale, stale);
    if (Nc > 0) k_bb<<<Nc, PS_BLOCK, 0, st>>>(s, x, f + s.nm3, bb);
    const bool box = s.box_active, ch = s.chain_HI;
    if (Nm > 0) {
        dim3 g1(ntb, nsplit);
        if (box && ch)
            k_aa<double, true, true><<<g1, PS_BLOCK, 0, st>>>(s, x, f, chain, per_split, part_aa);
        else if (box)
            k_aa<double, true, false><<<g1, PS_BLOCK, 0, st>>>(s, x, f, chain, per_split, part_aa);
        else if (ch)
            k_aa<double, false, true><<<g1, PS_BLOCK, 0, st>>>(s, x, f, chain, per_split, part_aa);
        else
            k_aa<double, false, false><<<g1, PS_BLOCK, 0, st>>>(s, x, f, chain, per_split, part_aa);
        if (Nc > 0) {
            dim3 g2(ntb, nchm);
            if (box)
                k_ab_mono<double, true><<<g2, PS_BLOCK, 0, st>>>(s, x, f, stale, per_chunk_m, part_abm);
            else
                k_ab_mono<double, false><<<g2, PS_BLOCK, 0, st>>>(s, x, f, stale, per_chunk_m, part_abm);
            dim3 g3(Nc, nchc);
            if (box)
                k_ab_coll<double, true><<<g3, PS_BLOCK, 0, st>>>(s, x, f, per_chunk_c, part_c, part_fs);
            else
                k_ab_coll<double, false><<<g3, PS_BLOCK, 0, st>>>(s, x, f, per_chunk_c, part_c, part_fs);
        }
    }
    dim3 g4((unsigned)((s.n + PS_BLOCK - 1) / PS_BLOCK));
    const FinArgs<double> fa{part_aa, Nm > 0 ? nsplit : 0, part_abm, (Nm > 0 && Nc > 0) ? nchm : 0, part_c, part_fs,
                             (Nm > 0 && Nc > 0) ? nchc : 0, bb, stale};
    k_finalize<double><<<g4, PS_BLOCK, 0, st>>>(s, f, u, fa, 0, 0);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ps_set_context_free_error(std::string("ps_mob_mult_noalloc: kernel launch: ") + hipGetErrorString(e));
        return PS_ERR_HIP;
    }
    return PS_OK;
}
