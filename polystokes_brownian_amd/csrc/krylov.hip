// krylov.hip -- MINRES (saddle solve) and Lanczos (M^{1/2} xi) on the device, FP64 recurrences.
//
// MINRES (replaces PETSc KSPSolve with KSPMINRES + PCJACOBI, src/init.cpp:40-64, src/vel.cpp:7-66):
// Paige & Saunders' preconditioned MINRES with the Jacobi diagonal of ArrowheadGetDiagonal +
// PCJacobiSetFixDiagonal.  The scalar recurrences never leave the GPU: every vector kernel reduces
// the previous kernel's block partials itself (same order in every block, so deterministic) and
// the Givens update is recomputed redundantly per block from a double-buffered state (read slot A,
// block 0 writes slot B).  Once converged the kernels turn into no-ops, so the host only polls the
// state every few iterations and the result equals stopping exactly at convergence.
//
// Lanczos (replaces SLEPc MFNSolve/FNSQRT on the mobility, src/slip_vel.cpp:27-49, following
// schur_sqrt_lanczos src/slip_vel.cpp:351-477): FP64 basis in HBM, two full reorthogonalisation
// passes done as classical Gram-Schmidt (one block GEMV each) instead of the reference's sequential
// modified Gram-Schmidt, tridiagonal eigendecomposition on the host (m <= 64), negative Ritz
// values floored, relative-change stopping test.
#include "ps_common.h"
#include "lzdots.h"
#include "tql2.h"

#include <algorithm>
#include <cmath>

namespace {

constexpr int VB = PS_BLOCK;   // threads per vector block
constexpr int VPT = 4;         // elements per thread
constexpr int VCHUNK = VB * VPT;
constexpr int MS_NSYS = 4;     // stresslet systems one launch of the wide MINRES can carry
constexpr int GB_SYS = 128;    // unsigned words per system's barrier flags (a flag per block, nbs <= 128)

inline int nblocks(int64_t n) { return (int)std::max<int64_t>(1, (n + VCHUNK - 1) / VCHUNK); }

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Deterministic block sum (fixed tree); result broadcast to all threads.
__device__ __forceinline__ double block_reduce(double v, double* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    return r;
}

// Sum of nb partials, identical in every block.
__device__ __forceinline__ double sum_partials(const double* __restrict__ red, int nb, double* sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nb; i += VB) v += red[i];
    return block_reduce(v, sh);
}

// ---------------------------------------------------------------------------- generic vector ops
__global__ void k_axpby(double* __restrict__ y, double a, const double* __restrict__ x, double b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}
__global__ void k_copy(double* __restrict__ d, const double* __restrict__ s, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i < n) d[i] = s[i];
}
__global__ void k_dot_part(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                           double* __restrict__ red) {
    __shared__ double sh[4];
    const int64_t base = (int64_t)blockIdx.x * VCHUNK;
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const int64_t i = base + k * VB + threadIdx.x;
        if (i < n) v += a[i] * b[i];
    }
    v = block_reduce(v, sh);
    if (threadIdx.x == 0) red[blockIdx.x] = v;
}
__global__ void k_sum_to(const double* __restrict__ red, int nb, double* __restrict__ out) {
    __shared__ double sh[4];
    const double v = sum_partials(red, nb, sh);
    if (threadIdx.x == 0) *out = v;
}

// ---------------------------------------------------------------------------- MINRES kernels
// init: r1 = b - (warm ? y : 0), x = warm ? x : 0, y = dinv r1; partials r1.y and b.dinv.b.
__global__ void k_mr_init(int64_t n, const double* __restrict__ b, double* __restrict__ x,
                          double* __restrict__ r1, double* __restrict__ r2, double* __restrict__ y,
                          const double* __restrict__ dinv, int warm, double* __restrict__ red) {
    __shared__ double sh[4];
    const int64_t base = (int64_t)blockIdx.x * VCHUNK;
    double v1 = 0.0, v2 = 0.0;
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const int64_t i = base + k * VB + threadIdx.x;
        if (i < n) {
            const double bi = b[i];
            double ri;
            if (warm) {
                ri = bi - y[i];
            } else {
                ri = bi;
                x[i] = 0.0;
            }
            const double zi = dinv[i] * ri;
            r1[i] = ri;
            r2[i] = ri;
            y[i] = zi;
            v1 += ri * zi;
            v2 += bi * dinv[i] * bi;
        }
    }
    v1 = block_reduce(v1, sh);
    v2 = block_reduce(v2, sh);
    if (threadIdx.x == 0) {
        red[blockIdx.x] = v1;
        red[gridDim.x + blockIdx.x] = v2;
    }
}

__global__ void k_mr_start(const double* __restrict__ red, int nb, double rtol, double abstol,
                           MinresState* __restrict__ st) {
    __shared__ double sh[4];
    const double b1 = sum_partials(red, nb, sh);
    const double bn = sum_partials(red + nb, nb, sh);
    if (threadIdx.x == 0) {
        MinresState s{};
        s.beta1 = b1 > 0.0 ? sqrt(b1) : 0.0;
        s.tol = fmax(rtol * sqrt(fmax(bn, 0.0)), abstol);
        s.oldb = 0.0;
        s.beta = s.beta1;
        s.dbar = 0.0;
        s.epsln = 0.0;
        s.phibar = s.beta1;
        s.cs = -1.0;
        s.sn = 0.0;
        s.its = 0;
        s.done = 0;
        s.reason = 0;
        if (b1 <= 0.0) {
            s.done = 1;
            s.reason = 3;
        } else if (s.phibar <= s.tol) {
            s.done = 1;
            s.reason = 2;
        }
        *st = s;
    }
}

__global__ void k_mr_v0(int64_t n, const MinresState* __restrict__ st, const double* __restrict__ y,
                        double* __restrict__ v) {
    if (st->done) return;
    const double sc = 1.0 / st->beta;
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i < n) v[i] = sc * y[i];
}

// a: y -= (beta/oldb) r1 (from the 2nd iteration); partial v.y
__global__ void k_mr_a(int64_t n, const MinresState* __restrict__ st, double* __restrict__ y,
                       const double* __restrict__ r1, const double* __restrict__ v, double* __restrict__ red) {
    __shared__ double sh[4];
    if (st->done) return;
    const bool sub = st->its >= 1;
    const double f = sub ? st->beta / st->oldb : 0.0;
    const int64_t base = (int64_t)blockIdx.x * VCHUNK;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const int64_t i = base + k * VB + threadIdx.x;
        if (i < n) {
            double yi = y[i];
            if (sub) yi -= f * r1[i];
            y[i] = yi;
            acc += v[i] * yi;
        }
    }
    acc = block_reduce(acc, sh);
    if (threadIdx.x == 0) red[blockIdx.x] = acc;
}

// b: alfa = sum(red); y -= (alfa/beta) r2; z = dinv y (into zbuf); partial y.z
__global__ void k_mr_b(int64_t n, const MinresState* __restrict__ st, const double* __restrict__ red, int nb,
                       double* __restrict__ y, const double* __restrict__ r2, const double* __restrict__ dinv,
                       double* __restrict__ zbuf, double* __restrict__ red2) {
    __shared__ double sh[4];
    if (st->done) return;
    const double alfa = sum_partials(red, nb, sh);
    const double f = alfa / st->beta;
    const int64_t base = (int64_t)blockIdx.x * VCHUNK;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const int64_t i = base + k * VB + threadIdx.x;
        if (i < n) {
            const double yi = y[i] - f * r2[i];
            const double zi = dinv[i] * yi;
            y[i] = yi;
            zbuf[i] = zi;
            acc += yi * zi;
        }
    }
    acc = block_reduce(acc, sh);
    if (threadIdx.x == 0) red2[blockIdx.x] = acc;
}

// c: scalar Givens update (redundant per block), w rotation, x += phi w, v <- z / beta_new.
__global__ void k_mr_c(int64_t n, const MinresState* __restrict__ st, MinresState* __restrict__ st_next,
                       const double* __restrict__ red_a, const double* __restrict__ red_b, int nb,
                       double* __restrict__ v, const double* __restrict__ z, const double* __restrict__ w_old,
                       const double* __restrict__ w2_old, double* __restrict__ w_new, double* __restrict__ x,
                       int maxit, double abstol) {
    __shared__ double sh[4];
    if (st->done) {
        // the host alternates the read/write slots every iteration: carry the final state over
        if (blockIdx.x == 0 && threadIdx.x == 0) *st_next = *st;
        return;
    }
    const double alfa = sum_partials(red_a, nb, sh);
    const double bb = sum_partials(red_b, nb, sh);
    MinresState s = *st;
    s.alfa = alfa;
    s.oldb = s.beta;
    s.beta = sqrt(fmax(bb, 0.0));
    const double oldeps = s.epsln;
    const double delta = s.cs * s.dbar + s.sn * alfa;
    const double gbar = s.sn * s.dbar - s.cs * alfa;
    s.epsln = s.sn * s.beta;
    s.dbar = -s.cs * s.beta;
    double gamma = sqrt(gbar * gbar + s.beta * s.beta);
    gamma = fmax(gamma, 1e-300);
    s.cs = gbar / gamma;
    s.sn = s.beta / gamma;
    const double phi = s.cs * s.phibar;
    s.phibar = s.sn * s.phibar;
    s.oldeps = oldeps;
    s.delta = delta;
    s.gamma = gamma;
    s.phi = phi;
    s.its += 1;
    if (s.phibar <= s.tol) {
        s.done = 1;
        s.reason = (s.phibar <= abstol) ? 3 : 2;
    } else if (s.beta == 0.0) {
        s.done = 1;
        s.reason = 2;
    } else if (s.its >= maxit) {
        s.done = 1;
        s.reason = -3;
    }
    const double denom = 1.0 / gamma;
    const double vs = (s.beta > 0.0) ? 1.0 / s.beta : 0.0;
    const int64_t base = (int64_t)blockIdx.x * VCHUNK;
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const int64_t i = base + k * VB + threadIdx.x;
        if (i < n) {
            // w1_new = w2_old (held in w_old's rotation), w2_new = w_old
            const double wn = (v[i] - oldeps * w2_old[i] - delta * w_old[i]) * denom;
            w_new[i] = wn;
            x[i] += phi * wn;
            v[i] = vs * z[i];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *st_next = s;
}


__global__ void k_diag(SysConst s, double* __restrict__ d) {
    const int64_t r = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (r >= s.ns) return;
    double v;
    if (r < s.nm3) v = s.beta_inv;
    else if (r < s.nm3 + s.nc3) v = 1.0;
    else if (r < s.nu) v = 3. / 4.;
    else if (r < s.n) v = 9. / 5.;
    else v = 1.0;   // zero constraint diagonal fixed to 1 (PCJacobiSetFixDiagonal)
    d[r] = v;
}
__global__ void k_recip(double* __restrict__ d, int64_t n) {
    const int64_t r = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (r < n) d[r] = 1.0 / d[r];
}

// ---------------------------------------------------------------------------- Lanczos kernels
// The two Gram-Schmidt passes of a Lanczos iteration over rows of LZB (block b owns rows LZB b ..),
// with the block partials of the Lanczos finalize (red1) or of the previous launch (red2):
//   k_lz_sub_dots: h = the sums of red1 (every block alike); alpha_j = h_{m-1}; w -= sum_k h_k Q_k;
//                  then the second pass's partial dots Q_k . w into red2;
//   k_lz_sub_norm: h from red2; w -= sum_k h_k Q_k; the block partials of |w|^2 into red3, which the
//                  next iteration's finalize sums into beta^2.
__global__ __launch_bounds__(LZB) void k_lz_sub_dots(int64_t n, const double* __restrict__ Q, int m,
                                                     double* __restrict__ w, const double* __restrict__ red1, int nb,
                                                     double* __restrict__ alpha_out, double* __restrict__ red2) {
    __shared__ double h[64];
    __shared__ double sh[64][LZW];
    lzd_sum(red1, nb, m, h);
    if (blockIdx.x == 0 && threadIdx.x == 0) *alpha_out = h[m - 1];
    const int64_t r = (int64_t)blockIdx.x * LZB + threadIdx.x;
    const bool in = r < n;
    const int64_t rc = in ? r : 0;
    double wi = in ? w[r] : 0.0;
#pragma unroll 8
    for (int q = 0; q < m; q++) wi -= h[q] * Q[(int64_t)q * n + rc];
    if (in) w[r] = wi;
    // rows past n read row 0 of Q (finite) times zero; no branch around the load, so the eight of a
    // batch are in flight together
    const double wz = in ? wi : 0.0;
    lzd_block_dots(
        m, [&](int q) { return Q[(int64_t)q * n + rc] * wz; }, sh, red2, nb);
}
__global__ __launch_bounds__(LZB) void k_lz_sub_norm(int64_t n, const double* __restrict__ Q, int m,
                                                     double* __restrict__ w, const double* __restrict__ red2, int nb,
                                                     double* __restrict__ red3) {
    __shared__ double h[64];
    __shared__ double sh[1][LZW];
    lzd_sum(red2, nb, m, h);
    const int64_t r = (int64_t)blockIdx.x * LZB + threadIdx.x;
    const bool in = r < n;
    const int64_t rc = in ? r : 0;
    double wi = in ? w[r] : 0.0;
#pragma unroll 8
    for (int q = 0; q < m; q++) wi -= h[q] * Q[(int64_t)q * n + rc];
    if (in) w[r] = wi;
    const double wz = in ? wi : 0.0;   // rows past n hold -sum_q h_q Q_q[0], not part of w
    lzd_block_dots(1, [&](int) { return wz * wz; }, sh, red3, nb);
}
inline unsigned eblocks(int64_t n) { return (unsigned)((n + VB - 1) / VB); }

}  // namespace

// ---------------------------------------------------------------------------- host side
ps_err krylov_alloc(Ctx* c) {
    const int64_t ns = c->ns, n = c->n;
    PS_HIP(c, hipMalloc(&c->kr, sizeof(double) * 9 * (size_t)ns));
    PS_HIP(c, hipMalloc(&c->dinv, sizeof(double) * (size_t)ns));
    c->red_cap = std::max(8192, 2 * nblocks(ns) * 66);
    PS_HIP(c, hipMalloc(&c->red, sizeof(double) * (size_t)c->red_cap));
    PS_HIP(c, hipMalloc(&c->mstate, sizeof(MinresState) * MS_NSYS));
    PS_HIP(c, hipHostMalloc(&c->mstate_h, sizeof(MinresState) * 2, hipHostMallocDefault));
    c->kmax = std::max(1, std::min<int>(c->p.sqrt_kmax > 0 ? c->p.sqrt_kmax : 64, 64));
    if ((int64_t)c->kmax > n) c->kmax = (int)std::max<int64_t>(1, n);
    PS_HIP(c, hipMalloc(&c->Q, sizeof(double) * (size_t)c->kmax * (size_t)n));
    PS_HIP(c, hipMalloc(&c->lz, sizeof(double) * (3 * 64 + 8)));
    PS_HIP(c, hipMemset(c->lz, 0, sizeof(double) * (3 * 64 + 8)));
    PS_HIP(c, hipHostMalloc(&c->lz_h, sizeof(double) * (3 * 64 + 8), hipHostMallocDefault));
    PS_HIP(c, hipMalloc(&c->lzg, sizeof(double) * 65 * 64));
    PS_HIP(c, hipMalloc(&c->lzred, sizeof(double) * 129 * (size_t)lzd_blocks(n)));
    PS_HIP(c, hipMalloc(&c->lzinfo, sizeof(double) * 4));
    PS_HIP(c, hipMemset(c->lzinfo, 0, sizeof(double) * 4));
    PS_HIP(c, hipMalloc(&c->dcnt, sizeof(int64_t) * DC_N));
    PS_HIP(c, hipMemset(c->dcnt, 0, sizeof(int64_t) * DC_N));
    PS_HIP(c, hipHostMalloc(&c->dcnt_h, sizeof(int64_t) * DC_N, hipHostMallocDefault));
    PS_HIP(c, hipMalloc(&c->gbar, sizeof(unsigned) * GB_SYS * (1 + MS_NSYS)));
    PS_HIP(c, hipMemset(c->gbar, 0, sizeof(unsigned) * GB_SYS * (1 + MS_NSYS)));
    PS_HIP(c, hipMalloc(&c->ir_r, sizeof(double) * (size_t)ns));
    PS_HIP(c, hipMalloc(&c->ir_d, sizeof(double) * (size_t)ns));
    if (!c->arrow) {   // the arrowhead diagonal is taken at the first MINRES solve (arrow.hip)
        PS_TRY(saddle_diag_dev(c, c->dinv));
        k_recip<<<eblocks(ns), VB, 0, c->stream>>>(c->dinv, ns);
        PS_LAUNCH(c);
    }
    if (c->p.solver == PS_SOLVER_BLOCK) {
        const size_t n5 = 5 * (size_t)std::max(c->Nc, 1);
        PS_HIP(c, hipMalloc(&c->Mss, sizeof(double) * n5 * n5));
        PS_HIP(c, hipMalloc(&c->bs_y, sizeof(double) * (size_t)n));
        PS_HIP(c, hipMalloc(&c->bs_r, sizeof(double) * n5));
    }
    return PS_OK;
}

void krylov_free(Ctx* c) {
    for (void* p : {(void*)c->kr, (void*)c->dinv, (void*)c->red, (void*)c->mstate, (void*)c->Q, (void*)c->lz,
                    (void*)c->lzg, (void*)c->lzinfo, (void*)c->dcnt, (void*)c->gbar, (void*)c->ir_r, (void*)c->ir_d,
                    (void*)c->Mss, (void*)c->bs_y, (void*)c->bs_r, (void*)c->lzred})
        if (p) (void)hipFree(p);
    c->lzred = nullptr;
    c->Mss = c->bs_y = c->bs_r = nullptr;
    c->lzg = c->lzinfo = nullptr;
    c->dcnt = nullptr;
    c->gbar = nullptr;
    if (c->mstate_h) (void)hipHostFree(c->mstate_h);
    if (c->lz_h) (void)hipHostFree(c->lz_h);
    if (c->dcnt_h) (void)hipHostFree(c->dcnt_h);
    c->dcnt_h = nullptr;
    c->kr = c->dinv = c->red = c->Q = c->lz = c->ir_r = c->ir_d = nullptr;
    c->mstate = nullptr;
    c->mstate_h = nullptr;
    c->lz_h = nullptr;
}

ps_err saddle_diag_dev(Ctx* c, double* d) {
    if (c->arrow) return arrow_diag(c, d, 0);
    k_diag<<<eblocks(c->ns), VB, 0, c->stream>>>(c->sc, d);
    PS_LAUNCH(c);
    return PS_OK;
}

ps_err vec_axpby(Ctx* c, double* y, double a, const double* x, double b, int64_t n) {
    if (n <= 0) return PS_OK;
    k_axpby<<<eblocks(n), VB, 0, c->stream>>>(y, a, x, b, n);
    PS_LAUNCH(c);
    return PS_OK;
}
ps_err vec_copy(Ctx* c, double* d, const double* s, int64_t n) {
    if (n <= 0) return PS_OK;
    PS_HIP(c, hipMemcpyAsync(d, s, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
    return PS_OK;
}
ps_err vec_dot(Ctx* c, const double* a, const double* b, int64_t n, double* out) {
    const int nb = nblocks(n);
    k_dot_part<<<nb, VB, 0, c->stream>>>(a, b, n, c->red);
    PS_LAUNCH(c);
    k_sum_to<<<1, VB, 0, c->stream>>>(c->red, nb, c->lz + 3 * 64);
    PS_LAUNCH(c);
    PS_HIP(c, hipMemcpyAsync(c->lz_h + 3 * 64, c->lz + 3 * 64, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    *out = c->lz_h[3 * 64];
    return PS_OK;
}

// Preconditioned MINRES for a symmetric operator `op` of size N with Jacobi inverse diagonal dinv.
template <class Op>
static ps_err minres_core(Ctx* c, int64_t N, const double* dinv, Op&& op, const double* b, double* x, int warm,
                          double rtol, int maxit, ps_solve_info* info) {
    const int nb = nblocks(N);
    double* R1 = c->kr;
    double* R2 = c->kr + N;
    double* Y = c->kr + 2 * N;
    double* V = c->kr + 3 * N;
    double* Wb = c->kr + 4 * N;
    double* W1 = c->kr + 5 * N;
    double* W2 = c->kr + 6 * N;
    double* red_a = c->red;
    double* red_b = c->red + nb;
    hipStream_t st = c->stream;
    MinresState* sA = c->mstate;
    MinresState* sB = c->mstate + 1;
    if (warm) PS_TRY(op(x, Y));
    k_mr_init<<<nb, VB, 0, st>>>(N, b, x, R1, R2, Y, dinv, warm, c->red);
    PS_LAUNCH(c);
    k_mr_start<<<1, VB, 0, st>>>(c->red, nb, rtol, c->p.ksp_abstol, sA);
    PS_LAUNCH(c);
    PS_HIP(c, hipMemsetAsync(Wb, 0, sizeof(double) * 3 * (size_t)N, st));   // w, w1, w2 = 0
    k_mr_v0<<<eblocks(N), VB, 0, st>>>(N, sA, Y, V);
    PS_LAUNCH(c);
    // r2 = r1 was written by init; rotation state (r1, r2, y) and (w, w1, w2)
    double *r1 = R1, *r2 = R2, *y = Y;
    double *w = Wb, *w1 = W1, *w2 = W2;
    int it = 0;
    const int poll = 4;
    bool done = false;
    int reason = 0;
    int its = 0;
    double phibar = 0.0, beta1 = 0.0;
    PS_HIP(c, hipMemcpyAsync(c->mstate_h, sA, sizeof(MinresState), hipMemcpyDeviceToHost, st));
    PS_HIP(c, hipStreamSynchronize(st));
    if (c->mstate_h[0].done) {
        done = true;
        reason = c->mstate_h[0].reason;
        phibar = c->mstate_h[0].phibar;
        beta1 = c->mstate_h[0].beta1;
    }
    while (!done) {
        // y = A v
        PS_TRY(op(V, y));
        k_mr_a<<<nb, VB, 0, st>>>(N, sA, y, r1, V, red_a);
        PS_LAUNCH(c);
        // z into r1's buffer (old r1 is dead after k_mr_a)
        k_mr_b<<<nb, VB, 0, st>>>(N, sA, red_a, nb, y, r2, dinv, r1, red_b);
        PS_LAUNCH(c);
        // rotation: r1 <- r2, r2 <- y, y <- z(r1 buffer)
        double* z = r1;
        r1 = r2;
        r2 = y;
        y = z;
        // w_new written into w1's buffer; w1 <- w2, w2 <- w
        k_mr_c<<<nb, VB, 0, st>>>(N, sA, sB, red_a, red_b, nb, V, y, w, w2, w1, x, maxit, c->p.ksp_abstol);
        PS_LAUNCH(c);
        double* wn = w1;
        w1 = w2;
        w2 = w;
        w = wn;
        // the next iteration reads the state k_mr_c just wrote
        std::swap(sA, sB);
        it++;
        if (it % poll == 0 || it >= maxit) {
            PS_HIP(c, hipMemcpyAsync(c->mstate_h, sA, sizeof(MinresState), hipMemcpyDeviceToHost, st));
            PS_HIP(c, hipStreamSynchronize(st));
            if (c->mstate_h[0].done || it >= maxit) {
                done = true;
                reason = c->mstate_h[0].done ? c->mstate_h[0].reason : -3;
                its = c->mstate_h[0].its;
                phibar = c->mstate_h[0].phibar;
                beta1 = c->mstate_h[0].beta1;
            }
        }
    }
    if (info) {
        info->iterations = its;
        info->reason = reason;
        info->rnorm = phibar;
        info->rnorm0 = beta1;
        info->refinements = 0;
        info->true_rel_residual = -1.0;
    }
    return reason > 0 ? PS_OK : PS_ERR_NOCONV;
}

// Preconditioned MINRES on the saddle operator with the apply precision `prec`.
static ps_err minres_dev(Ctx* c, const double* b, double* x, int warm, double rtol, int maxit, int prec,
                         ps_solve_info* info) {
    ps_solve_info ii{};
    if (c->arrow && !c->ar_diag_frozen) {
        // mm_HI off: the operator is a PETSc MatShell whose state never changes, so PCJacobi takes the
        // diagonal at the first KSPSolve and keeps it for the run (PCSetUp skips an unchanged operator)
        PS_TRY(arrow_diag(c, c->dinv, 1));
        c->ar_diag_frozen = true;
    }
    auto op = [c, prec](const double* in, double* out) { return saddle_apply(c, in, out, prec); };
    ps_err r = minres_core(c, c->ns, c->dinv, op, b, x, warm, rtol, maxit, &ii);
    c->last_minres_its = ii.iterations;
    c->cnt.minres_iterations += ii.iterations;
    c->cnt.saddle_solves++;
    if (info) *info = ii;
    return r;
}

// ---------------------------------------------------------------------------- small dense MINRES
// The stresslet system M_ss f_s = r (5Nc unknowns, dense, FP64) is solved by the same preconditioned
// MINRES recurrence in one persistent launch: every block keeps an identical copy of the scalar
// state, does the O(5Nc) vector work of an iteration redundantly over the whole vectors (same loops,
// same reduction tree, hence the same scalars everywhere), keeps v_{k+1} in LDS, writes its own row
// range of the next vector set and its rows of y_{k+1} = A v_{k+1}; a grid-wide barrier separates
// the iterations.  The vectors are double-buffered by iteration parity, so no block reads in an
// iteration what another block writes in it.  All blocks see the same state, so they leave the loop
// at the same iteration: at convergence or at the cap, with no host round trip.
// Preconditioner: block Jacobi with the exact inverse of the 5x5 diagonal blocks.  Every diagonal
// block of M_ss is the stresslet self block of fill_self (src/mob.cpp:600-612), 9/5 I with 9/10 at
// (0,4) and (4,0), so its inverse couples entries 0 and 4 of a colloid by [[a, b], [b, a]],
// a = (9/5)/((9/5)^2 - (9/10)^2), b = -(9/10)/((9/5)^2 - (9/10)^2), and scales entries 1..3 by 5/9.
constexpr uint64_t GRID_WAIT_TICKS = 1000000000ull;   // 10 s of the 100 MHz wall clock

struct MsSet {
    double *r1, *r2, *y, *v, *wa, *wb;   // wa = w_{k-1}, wb = w_{k-2}; y = A v
};

// Barrier over nblocks co-resident blocks (the grid is at most one block per CU): the whole grid, or
// the blocks of one system.  bw[0] counts arrivals, bw[1] is the generation, *err the sticky error
// word shared by every barrier of the context.  Each wave waits for its own stores, the block meets,
// one thread arrives with an agent-scope release and waits for the generation to move on, bounded by
// the wall clock; a timeout sets the error word, which ends every other block's wait at once.
// Returns false after a timeout.
__device__ bool grid_sync(unsigned* bw, unsigned* err, unsigned nblocks, unsigned& gen) {
    __shared__ int ok_sh;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned target = gen + 1;
        if (__hip_atomic_fetch_add(&bw[0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) == nblocks - 1) {
            __hip_atomic_store(&bw[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bw[1], target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = wall_clock64();
            while ((int)(__hip_atomic_load(&bw[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    ok = 0;
                    break;
                }
                if (wall_clock64() - t0 > GRID_WAIT_TICKS) {
                    __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok_sh = ok;
    }
    __syncthreads();
    gen++;
    return ok_sh != 0;
}

// Barrier among the nblocks (<= GB_SYS) blocks of one system: block `me` publishes the generation it
// has reached in its own word flags[me] (agent-scope release), and wave 0 waits until every word of the
// system has reached it, one lane per word, bounded by the wall clock; a timeout sets the sticky error
// word, which ends every other block's wait at once.  No read-modify-write: no block waits on another's
// atomic.  Returns false after a timeout.
__device__ bool flag_sync(unsigned* flags, unsigned* err, int nblocks, int me, unsigned& gen) {
    __shared__ int ok_sh;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const unsigned target = gen + 1;
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        if (lane == 0) __hip_atomic_store(&flags[me], target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        const uint64_t t0 = wall_clock64();
        while (true) {
            bool reached = true;
            for (int b = lane; b < nblocks; b += 64)
                reached = reached &&
                          (int)(__hip_atomic_load(&flags[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) >= 0;
            if (__all(reached)) break;
            if (__any(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
                ok = 0;
                break;
            }
            if (wall_clock64() - t0 > GRID_WAIT_TICKS) {
                if (lane == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (lane == 0) ok_sh = ok;
    }
    __syncthreads();
    gen++;
    return ok_sh != 0;
}

__device__ __forceinline__ void ms_prec5(const double (&r)[5], double (&z)[5]) {
    const double pa = 1.8 / 2.43, pb = -0.9 / 2.43;
    z[0] = pa * r[0] + pb * r[4];
    z[1] = r[1] * (5.0 / 9.0);
    z[2] = r[2] * (5.0 / 9.0);
    z[3] = r[3] * (5.0 / 9.0);
    z[4] = pb * r[0] + pa * r[4];
}

// Rows [rb, re) of y = A V, one wave per row, four independent accumulators per lane.
__device__ __forceinline__ void ms_gemv_rows(int64_t N, const double* __restrict__ A, const double* V, int64_t rb,
                                             int64_t re, double* __restrict__ y) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int64_t row = rb + wid; row < re; row += VB / 64) {
        const double* a = A + row * N;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int64_t j = lane;
        for (; j + 192 < N; j += 256) {
            s0 += a[j] * V[j];
            s1 += a[j + 64] * V[j + 64];
            s2 += a[j + 128] * V[j + 128];
            s3 += a[j + 192] * V[j + 192];
        }
        for (; j < N; j += 64) s0 += a[j] * V[j];
        const double t = wave_sum((s0 + s1) + (s2 + s3));
        if (lane == 0) y[row] = t;
    }
}

__global__ __launch_bounds__(VB) void k_ms_solve(int64_t N, const double* __restrict__ A, const double* __restrict__ b,
                                                 double* __restrict__ x, MsSet set0, MsSet set1, int rows,
                                                 double rtol, double abstol, int maxit, unsigned* gb,
                                                 MinresState* __restrict__ st_out, int64_t* __restrict__ dcnt) {
    extern __shared__ double V[];   // [N]: z, then v_{k+1}
    __shared__ double sh[4];
    if (__hip_atomic_load(&gb[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    unsigned gen = __hip_atomic_load(&gb[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned nbk = gridDim.x;
    const int64_t rb = (int64_t)blockIdx.x * rows, re = min(N, rb + rows);
    const int64_t ncb = N / 5;
    // start (cold): x = 0, r1 = r2 = b, z = P^-1 b, beta1^2 = b . z (also the b-norm of the test)
    double acc = 0.0;
    for (int64_t cb = threadIdx.x; cb < ncb; cb += VB) {
        double r[5], z[5];
#pragma unroll
        for (int k = 0; k < 5; k++) r[k] = b[5 * cb + k];
        ms_prec5(r, z);
#pragma unroll
        for (int k = 0; k < 5; k++) {
            V[5 * cb + k] = z[k];
            acc += r[k] * z[k];
        }
    }
    const double b1 = block_reduce(acc, sh);
    MinresState s{};
    s.beta1 = b1 > 0.0 ? sqrt(b1) : 0.0;
    s.tol = fmax(rtol * sqrt(fmax(b1, 0.0)), abstol);
    s.beta = s.beta1;
    s.phibar = s.beta1;
    s.cs = -1.0;
    if (b1 <= 0.0) {
        s.done = 1;
        s.reason = 3;
    } else if (s.phibar <= s.tol) {
        s.done = 1;
        s.reason = 2;
    }
    const double sc0 = s.done ? 0.0 : 1.0 / s.beta;
    for (int64_t cb = threadIdx.x; cb < ncb; cb += VB)
#pragma unroll
        for (int k = 0; k < 5; k++) V[5 * cb + k] *= sc0;
    __syncthreads();
    for (int64_t i = rb + threadIdx.x; i < re; i += VB) {
        x[i] = 0.0;
        set0.r1[i] = b[i];
        set0.r2[i] = b[i];
        set0.wa[i] = 0.0;
        set0.wb[i] = 0.0;
        set0.v[i] = V[i];
    }
    bool ok = true;
    if (!s.done) {
        ms_gemv_rows(N, A, V, rb, re, set0.y);
        ok = grid_sync(gb, gb + 2, nbk, gen);
    }
    int par = 0;
    while (ok && !s.done) {
        const MsSet cur = par ? set1 : set0;
        const MsSet nxt = par ? set0 : set1;
        // y -= (beta/oldb) r1 (from the 2nd iteration); alfa = v . y
        const bool sub = s.its >= 1;
        const double f = sub ? s.beta / s.oldb : 0.0;
        acc = 0.0;
        for (int64_t i = threadIdx.x; i < N; i += VB) {
            double ya = cur.y[i];
            if (sub) ya -= f * cur.r1[i];
            acc += cur.v[i] * ya;
        }
        const double alfa = block_reduce(acc, sh);
        // r2_new = y - (alfa/beta) r2; z = P^-1 r2_new per colloid block; beta_new^2 = r2_new . z
        const double f2 = alfa / s.beta;
        acc = 0.0;
        for (int64_t cb = threadIdx.x; cb < ncb; cb += VB) {
            double r[5], z[5];
#pragma unroll
            for (int k = 0; k < 5; k++) {
                const int64_t i = 5 * cb + k;
                double ya = cur.y[i];
                if (sub) ya -= f * cur.r1[i];
                r[k] = ya - f2 * cur.r2[i];
            }
            ms_prec5(r, z);
#pragma unroll
            for (int k = 0; k < 5; k++) {
                V[5 * cb + k] = z[k];
                acc += r[k] * z[k];
            }
        }
        const double bb = block_reduce(acc, sh);
        s.alfa = alfa;
        s.oldb = s.beta;
        s.beta = sqrt(fmax(bb, 0.0));
        const double oldeps = s.epsln;
        const double delta = s.cs * s.dbar + s.sn * alfa;
        const double gbar = s.sn * s.dbar - s.cs * alfa;
        s.epsln = s.sn * s.beta;
        s.dbar = -s.cs * s.beta;
        const double gamma = fmax(sqrt(gbar * gbar + s.beta * s.beta), 1e-300);
        s.cs = gbar / gamma;
        s.sn = s.beta / gamma;
        const double phi = s.cs * s.phibar;
        s.phibar = s.sn * s.phibar;
        s.oldeps = oldeps;
        s.delta = delta;
        s.gamma = gamma;
        s.phi = phi;
        s.its += 1;
        if (s.phibar <= s.tol) {
            s.done = 1;
            s.reason = (s.phibar <= abstol) ? 3 : 2;
        } else if (s.beta == 0.0) {
            s.done = 1;
            s.reason = 2;
        } else if (s.its >= maxit) {
            s.done = 1;
            s.reason = -3;
        }
        const double denom = 1.0 / gamma;
        const double vs = (s.beta > 0.0) ? 1.0 / s.beta : 0.0;
        // V holds z; each thread scales the entries it wrote
        for (int64_t cb = threadIdx.x; cb < ncb; cb += VB)
#pragma unroll
            for (int k = 0; k < 5; k++) V[5 * cb + k] *= vs;
        __syncthreads();
        // own rows: w_k = (v - oldeps w_{k-2} - delta w_{k-1}) / gamma, x += phi w_k, shift the vectors
        for (int64_t i = rb + threadIdx.x; i < re; i += VB) {
            double ya = cur.y[i];
            if (sub) ya -= f * cur.r1[i];
            const double w1 = cur.wa[i];
            const double wn = (cur.v[i] - oldeps * cur.wb[i] - delta * w1) * denom;
            nxt.wb[i] = w1;
            nxt.wa[i] = wn;
            x[i] += phi * wn;
            nxt.r1[i] = cur.r2[i];
            nxt.r2[i] = ya - f2 * cur.r2[i];
            nxt.v[i] = V[i];
        }
        par ^= 1;
        if (s.done) break;
        ms_gemv_rows(N, A, V, rb, re, nxt.y);
        ok = grid_sync(gb, gb + 2, nbk, gen);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (!ok) {
            s.done = 1;
            s.reason = -5;   // KSP_DIVERGED_BREAKDOWN's sign: the wait ran out of time
            dcnt[DC_BARRIER_TIMEOUT] = 1;
        }
        *st_out = s;
        dcnt[DC_MSS_ITERS] += s.its;
        if (s.reason < 0) dcnt[DC_MSS_CAPPED] += 1;
    }
}

// The same recurrence for blocks of up to 2048 unknowns, several independent systems in one launch
// (MS_NSYS at most).  Each system gets nbs blocks of 1024 threads; a block keeps r1, r2, v and y of
// its system's whole vector in LDS and z = P^-1 r2 in registers (computed redundantly, so identical in
// every block of the system) and R = 1024 / TPR of its rows of A on chip (registers, the last columns
// in LDS): TPR threads per row, 32 columns a thread.  The only data crossing blocks is y = A v: each
// block writes its rows and, after the barrier among its system's blocks (flag_sync over the system's
// GB_SYS flag words from gb + GB_SYS (1 + system)), reads its system's whole vector once.  Every block
// of a system runs the same
// recurrence on the same data, so all of them see convergence at the same iteration and leave
// together; a system does not wait for the others.  With one system this is the single solve; with
// several, one launch costs the slowest system's iterations instead of their sum.
constexpr int MW = 1024, MW_COLS = 32;
// Of a thread's MW_COLS columns of A, the last ms_ldsc(TPR) live in LDS, which leaves the registers
// room for the recurrence without spills (the LDS then holds at most 3 N + TPR MW_COLS + that many
// times MW doubles: 136 KiB at N = 2048).
constexpr int ms_ldsc(int tpr) { return tpr == 32 ? 10 : 8; }

struct MsSys {
    const double* A;    // N x N, row-major
    const double* b;    // N
    double* x;          // N
    double* ybuf;       // [2][N]: y = A v, double-buffered across the barrier
    MinresState* st;    // final state
};
struct MsBatch {
    MsSys s[MS_NSYS];
};

// Deterministic sum over the 16 waves of a 1024-thread block (fixed tree), broadcast to all threads.
// sh holds two slots of 16; consecutive calls alternate slots (`slot`), so one block barrier per call
// suffices: a slot is written again only after the next call's barrier, which every reader of its
// previous contents has passed.
__device__ __forceinline__ double block_reduce_w(double v, double (*sh)[16], int slot) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    v = wave_sum64(v);
    if (lane == 0) sh[slot][wid] = v;
    __syncthreads();
    double r[8];
#pragma unroll
    for (int k = 0; k < 8; k++) r[k] = sh[slot][2 * k] + sh[slot][2 * k + 1];
    return ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
}

// Entry k = i mod 5 of the stresslet self-block preconditioner z = P^-1 r (ms_prec5), one entry per
// thread: rp is the entry it couples to (i + 4 for k = 0, i - 4 for k = 4; unused otherwise).
__device__ __forceinline__ double ms_prec_el(int k, double r, double rp) {
    const double pa = 1.8 / 2.43, pb = -0.9 / 2.43;
    if (k == 0) return pa * r + pb * rp;
    if (k == 4) return pb * rp + pa * r;
    return r * (5.0 / 9.0);
}

// Sum over each group of TPR consecutive lanes (32: the two halves of the wave; 64: the wave), the
// same bits in every lane of the group: quad and row sums by DPP, then the rows' sums by lane reads.
template <int TPR>
__device__ __forceinline__ double group_sum(double a) {
    a += dpp_mov<0xB1>(a);    // quad_perm [1, 0, 3, 2]
    a += dpp_mov<0x4E>(a);    // quad_perm [2, 3, 0, 1]
    a += dpp_mov<0x124>(a);   // row_ror:4
    a += dpp_mov<0x128>(a);   // row_ror:8
    const double s0 = lane_read(a, 0), s1 = lane_read(a, 16), s2 = lane_read(a, 32), s3 = lane_read(a, 48);
    if constexpr (TPR == 64) return (s0 + s1) + (s2 + s3);
    return (threadIdx.x & 32) ? s2 + s3 : s0 + s1;
}

template <int TPR>
__global__ __launch_bounds__(MW) void k_ms_wide(int N, MsBatch bt, int nbs, double rtol, double abstol, int maxit,
                                                unsigned* gb, int64_t* __restrict__ dcnt) {
    constexpr int R = MW / TPR;             // rows per block
    constexpr int NV = TPR * MW_COLS;       // columns the gemv threads cover (v is zero beyond N)
    constexpr int NE = NV / MW;             // vector entries per thread (t, t + MW, ...)
    extern __shared__ double lds[];
    // residuals r1 (previous) and r2 (current) live in two buffers that trade places every iteration:
    // the new residual is written over r1, which the iteration has read by then
    double* r1 = lds;           // [N]
    double* r2 = lds + N;       // [N]
    double* y = lds + 2 * N;    // [N]: A v_k - f r1
    double* v = lds + 3 * N;    // [NV]: v_k
    double* Al = v + NV;        // [ms_ldsc(TPR)][MW]: the last columns of each thread's row of A
    __shared__ double sh[2][16];
    // the recurrence's scalars, the same in every thread: two slots, the current one read by all
    // threads and the next one written by thread 0 (every thread computes it; only registers for the
    // span of the update)
    __shared__ MinresState ss[2];
    if (__hip_atomic_load(&gb[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    const int sys = blockIdx.x / nbs, bs = blockIdx.x - sys * nbs;
    unsigned* flags = gb + GB_SYS * (1 + sys);
    unsigned gen = __hip_atomic_load(&flags[bs], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const MsSys S = bt.s[sys];
    const int t = threadIdx.x, q = t / TPR, tl = t % TPR;
    const int rb = bs * R;
    const int nr = max(0, min(R, N - rb));
    // this block's rows of A, kept on chip for the whole solve (thread (q, tl): row rb + q, columns
    // tl + k TPR; k < MW_COLS - ms_ldsc(TPR) in registers, the rest in LDS)
    constexpr int KR = MW_COLS - ms_ldsc(TPR);
    double Ac[KR];
#pragma unroll
    for (int k = 0; k < MW_COLS; k++) {
        const int j = tl + k * TPR;
        const double a = (q < nr && j < N) ? S.A[(int64_t)(rb + q) * N + j] : 0.0;
        if (k < KR)
            Ac[k] = a;
        else
            Al[(k - KR) * MW + t] = a;
    }
    // y rows = A v: no bounds test per column (v is zero past N, as are the padded entries of A)
    auto gemv = [&](double* __restrict__ yout) {
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < KR; k++) a = __builtin_fma(Ac[k], v[tl + k * TPR], a);
#pragma unroll
        for (int k = KR; k < MW_COLS; k++) a = __builtin_fma(Al[(k - KR) * MW + t], v[tl + k * TPR], a);
        a = group_sum<TPR>(a);
        if (tl == 0 && q < nr) yout[rb + q] = a;
    };
    // each thread owns entries i = t + e MW (e < NE, i < N) of the vectors; the partner of i is the
    // entry its preconditioner row couples to
    auto partner = [](int i) {
        const int k = i % 5;
        return k == 0 ? i + 4 : (k == 4 ? i - 4 : i);
    };
    // start (cold): x = 0, r1 = r2 = b, z = P^-1 b, beta1^2 = b . z
    double acc = 0.0;
    double zs[NE];
#pragma unroll
    for (int e = 0; e < NE; e++) {
        const int i = t + e * MW;
        zs[e] = 0.0;
        if (i < N) {
            const double r = S.b[i];
            zs[e] = ms_prec_el(i % 5, r, S.b[partner(i)]);
            r1[i] = r;
            r2[i] = r;
            acc += r * zs[e];
        } else {
            v[i] = 0.0;
        }
    }
    int slot = 0;
    const double b1 = block_reduce_w(acc, sh, slot);
    slot ^= 1;
    int sp = 0;
    bool done;
    {
        MinresState s{};
        s.beta1 = b1 > 0.0 ? sqrt(b1) : 0.0;
        s.tol = fmax(rtol * sqrt(fmax(b1, 0.0)), abstol);
        s.beta = s.beta1;
        s.phibar = s.beta1;
        s.cs = -1.0;
        if (b1 <= 0.0) {
            s.done = 1;
            s.reason = 3;
        } else if (s.phibar <= s.tol) {
            s.done = 1;
            s.reason = 2;
        }
        done = s.done != 0;
        const double sc0 = done ? 0.0 : 1.0 / s.beta;
#pragma unroll
        for (int e = 0; e < NE; e++) {
            const int i = t + e * MW;
            if (i < N) v[i] = zs[e] * sc0;
        }
        if (t == 0) ss[0] = s;
    }
    // own rows: w_{k-1}, w_{k-2}, x (thread t < nr holds row rb + t)
    double wa = 0.0, wb = 0.0, xo = 0.0;
    __syncthreads();
    int par = 0;
    bool ok = true;
    if (!done) {
        gemv(S.ybuf);
        ok = flag_sync(flags, gb + 2, nbs, bs, gen);
    }
    while (ok && !done) {
        const MinresState& cur = ss[sp];
        const double* yk = S.ybuf + (par ? N : 0);
        double* ynext = S.ybuf + (par ? 0 : N);
        // y -= (beta/oldb) r1 (from the 2nd iteration); alfa = v . y; this thread's row of v_k kept for
        // the w update
        const bool sub = cur.its >= 1;
        const double f = sub ? cur.beta / cur.oldb : 0.0;
        const double vown = t < nr ? v[rb + t] : 0.0;
        acc = 0.0;
#pragma unroll
        for (int e = 0; e < NE; e++) {
            const int i = t + e * MW;
            if (i < N) {
                double ya = yk[i];
                if (sub) ya -= f * r1[i];
                y[i] = ya;
                acc += v[i] * ya;
            }
        }
        const double alfa = block_reduce_w(acc, sh, slot);
        slot ^= 1;
        // r2_new = y - (alfa/beta) r2 (over r1); z = P^-1 r2_new; beta_new^2 = r2_new . z
        const double f2 = alfa / cur.beta;
        acc = 0.0;
#pragma unroll
        for (int e = 0; e < NE; e++) {
            const int i = t + e * MW;
            if (i < N) {
                const int pi = partner(i);
                const double r = y[i] - f2 * r2[i];
                zs[e] = ms_prec_el(i % 5, r, y[pi] - f2 * r2[pi]);
                r1[i] = r;
                acc += r * zs[e];
            }
        }
        const double bb = block_reduce_w(acc, sh, slot);
        slot ^= 1;
        {
            double* tmp = r1;   // r1 <- r2, r2 <- r2_new
            r1 = r2;
            r2 = tmp;
        }
        MinresState s = cur;
        s.alfa = alfa;
        s.oldb = s.beta;
        s.beta = sqrt(fmax(bb, 0.0));
        const double oldeps = s.epsln;
        const double delta = s.cs * s.dbar + s.sn * alfa;
        const double gbar = s.sn * s.dbar - s.cs * alfa;
        s.epsln = s.sn * s.beta;
        s.dbar = -s.cs * s.beta;
        const double gamma = fmax(sqrt(gbar * gbar + s.beta * s.beta), 1e-300);
        s.cs = gbar / gamma;
        s.sn = s.beta / gamma;
        const double phi = s.cs * s.phibar;
        s.phibar = s.sn * s.phibar;
        s.oldeps = oldeps;
        s.delta = delta;
        s.gamma = gamma;
        s.phi = phi;
        s.its += 1;
        if (s.phibar <= s.tol) {
            s.done = 1;
            s.reason = (s.phibar <= abstol) ? 3 : 2;
        } else if (s.beta == 0.0) {
            s.done = 1;
            s.reason = 2;
        } else if (s.its >= maxit) {
            s.done = 1;
            s.reason = -3;
        }
        done = s.done != 0;
        // the other slot was last read before this iteration's first reduction barrier; it is read
        // next after the barrier ahead of the gemv (or, once done, by this thread at the end)
        if (t == 0) ss[sp ^ 1] = s;
        sp ^= 1;
        const double denom = 1.0 / gamma;
        const double vs = (s.beta > 0.0) ? 1.0 / s.beta : 0.0;
        // own rows: w_k = (v - oldeps w_{k-2} - delta w_{k-1}) / gamma, x += phi w_k
        if (t < nr) {
            const double wn = (vown - oldeps * wb - delta * wa) * denom;
            wb = wa;
            wa = wn;
            xo += phi * wn;
        }
        par ^= 1;
        if (!done) {
            // v_{k+1} = z / beta_new: every read of v_k was before the last reduction's barrier
#pragma unroll
            for (int e = 0; e < NE; e++) {
                const int i = t + e * MW;
                if (i < N) v[i] = zs[e] * vs;
            }
            __syncthreads();
            gemv(ynext);
            ok = flag_sync(flags, gb + 2, nbs, bs, gen);
        }
    }
    if (t < nr) S.x[rb + t] = xo;
    if (bs == 0 && t == 0) {
        MinresState s = ss[sp];
        if (!ok) {
            s.done = 1;
            s.reason = -5;   // KSP_DIVERGED_BREAKDOWN's sign: the wait ran out of time
            dcnt[DC_BARRIER_TIMEOUT] = 1;
        }
        *S.st = s;
        atomicAdd((unsigned long long*)&dcnt[DC_MSS_ITERS], (unsigned long long)s.its);
        if (s.reason < 0) atomicAdd((unsigned long long*)&dcnt[DC_MSS_CAPPED], 1ull);
    }
}

// Launch shape of k_ms_wide for N unknowns: threads per row (32 or 64), rows per block, blocks per
// system; false when a system of this size does not fit (more than 2048 unknowns).
static_assert(2 * MW / (MW / 64) <= GB_SYS, "a system's blocks must fit its barrier flags");
inline bool ms_wide_shape(int64_t N, int* tpr, int* nbs) {
    if (N <= 0 || N > 2 * MW) return false;
    *tpr = N <= MW ? 32 : 64;
    const int R = MW / *tpr;
    *nbs = (int)((N + R - 1) / R);
    return true;
}

// Launch k_ms_wide over `ns` systems of N unknowns each (at most one block per CU in all, so the whole
// grid is co-resident for its barriers).
static ps_err ms_wide_launch(Ctx* c, int64_t N, const MsBatch& bt, int ns, double rtol, int maxit) {
    int tpr = 0, nbs = 0;
    if (ns < 1 || ns > MS_NSYS || !ms_wide_shape(N, &tpr, &nbs) || (int64_t)nbs * ns > std::max(1, c->n_cu)) {
        c->set_error("stresslet MINRES: " + std::to_string(ns) + " systems of " + std::to_string(N) +
                     " unknowns do not fit one block per CU");
        return PS_ERR_ARG;
    }
    const size_t lds = sizeof(double) * (3 * (size_t)N + (size_t)tpr * MW_COLS + (size_t)ms_ldsc(tpr) * MW);
    const unsigned nb = (unsigned)(nbs * ns);
    if (tpr == 32)
        k_ms_wide<32><<<nb, MW, lds, c->stream>>>((int)N, bt, nbs, rtol, c->p.ksp_abstol, maxit, c->gbar, c->dcnt);
    else
        k_ms_wide<64><<<nb, MW, lds, c->stream>>>((int)N, bt, nbs, rtol, c->p.ksp_abstol, maxit, c->gbar, c->dcnt);
    PS_LAUNCH(c);
    return PS_OK;
}

static ps_err minres_small_info(Ctx* c, ps_solve_info* info);
// Solve A x = b for the small dense block A (N x N, row-major) by the recurrence above in one launch.
// With `info` the final state is read back (a host round trip); without, nothing is.
static ps_err minres_small(Ctx* c, int64_t N, const double* A, const double* b, double* x, double rtol, int maxit,
                           ps_solve_info* info) {
    hipStream_t st = c->stream;
    if (N > 8192 || N % 5 != 0) {   // v_{k+1} lives in 64 KiB of LDS; the rows come in colloid blocks
        c->set_error("block solver: 5 Nc = " + std::to_string(N) + " stresslet unknowns exceed 8192");
        return PS_ERR_ARG;
    }
    MsSet set[2];
    for (int k = 0; k < 2; k++) {
        double* base = c->kr + (size_t)k * 6 * N;
        set[k] = MsSet{base, base + N, base + 2 * N, base + 3 * N, base + 4 * N, base + 5 * N};
    }
    // at most one block per CU, so the whole grid is co-resident for its barriers
    const int ncu = std::max(1, c->n_cu);
    int tpr = 0, nbs = 0;
    if (ms_wide_shape(N, &tpr, &nbs) && nbs <= ncu) {
        MsBatch bt{};
        bt.s[0] = MsSys{A, b, x, c->kr, c->mstate};
        PS_TRY(ms_wide_launch(c, N, bt, 1, rtol, maxit));
        return minres_small_info(c, info);
    }
    int rows = (int)std::max<int64_t>(8, (N + 255) / 256);
    while ((N + rows - 1) / rows > ncu) rows *= 2;
    const unsigned nb = (unsigned)((N + rows - 1) / rows);
    const size_t lds = sizeof(double) * (size_t)N;
    k_ms_solve<<<nb, VB, lds, st>>>(N, A, b, x, set[0], set[1], rows, rtol, c->p.ksp_abstol, maxit, c->gbar,
                                     c->mstate, c->dcnt);
    PS_LAUNCH(c);
    return minres_small_info(c, info);
}

// The stresslet solve's outcome, read back only when asked for (a host round trip).
static ps_err minres_small_info(Ctx* c, ps_solve_info* info) {
    hipStream_t st = c->stream;
    if (!info) return PS_OK;
    PS_HIP(c, hipMemcpyAsync(c->mstate_h, c->mstate, sizeof(MinresState), hipMemcpyDeviceToHost, st));
    PS_HIP(c, hipStreamSynchronize(st));
    const MinresState& s = c->mstate_h[0];
    if (s.reason == -5) {
        c->set_error("stresslet MINRES: a grid-wide wait ran out of time");
        return PS_ERR_HIP;
    }
    *info = ps_solve_info{};
    info->iterations = s.its;
    info->reason = s.reason;
    info->rnorm = s.phibar;
    info->rnorm0 = s.beta1;
    info->true_rel_residual = -1.0;
    return s.reason > 0 ? PS_OK : PS_ERR_NOCONV;
}

// x = [-b_lambda (nu) ; 0 (5Nc)]  (f_u from the constraint rows B^T f = b_lambda); entries from
// `upto` on are zero (the lubricated solve fixes only the monomer forces, upto = 3Nm)
__global__ void k_bs_init(double* __restrict__ x, const double* __restrict__ b, int64_t n, int64_t upto) {
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= n) return;
    x[i] = (i < upto) ? -b[n + i] : 0.0;
}

// Colloid saddle of the lubricated block solve: its right-hand side [b_c - y_c (11Nc); b_lambda,c (6Nc)]
// and Jacobi inverse diagonal (self blocks 1, 4/3, 5/9; 1 on the velocity rows, whose diagonal is
// -R_ii: any positive diagonal preconditions MINRES)
__global__ void k_cs_rhs(const double* __restrict__ b, const double* __restrict__ y, int64_t n, int64_t nm3,
                         int Nc, double* __restrict__ rhs, double* __restrict__ dinv) {
    const int64_t k = (int64_t)blockIdx.x * VB + threadIdx.x;
    const int64_t nc3 = 3 * (int64_t)Nc, nc6 = 6 * (int64_t)Nc, nc11 = 11 * (int64_t)Nc;
    if (k >= nc11 + nc6) return;
    if (k < nc11) {
        rhs[k] = b[nm3 + k] - y[nm3 + k];
        dinv[k] = k < nc3 ? 1.0 : k < nc6 ? 4.0 / 3.0 : 5.0 / 9.0;
    } else {
        rhs[k] = b[n + nm3 + (k - nc11)];
        dinv[k] = 1.0;
    }
}
// the colloid forces of the solution into x
__global__ void k_cs_scatter(const double* __restrict__ v, double* __restrict__ x, int64_t nm3, int64_t nc11) {
    const int64_t k = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (k < nc11) x[nm3 + k] = v[k];
}
// rs = b_s - y_s
__global__ void k_bs_rs(double* __restrict__ rs, const double* __restrict__ b, const double* __restrict__ y,
                        int64_t nu, int64_t n5) {
    const int64_t k = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (k < n5) rs[k] = b[nu + k] - y[nu + k];
}
// u = (M f)_u - b_u, into x[n:] and, when given, uout
__global__ void k_bs_u(double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ b,
                       int64_t n, int64_t nu, double* __restrict__ uout) {
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= nu) return;
    const double u = y[i] - b[i];
    x[n + i] = u;
    if (uout) uout[i] = u;
}

// Exact block elimination of A x = b with B = [-I_nu; 0]:
//   f_u = -b[n:n+nu];  M_ss f_s = b_s - (M [f_u; 0])_s;  u = (M f)_u - b_u.
// With one colloid and monomers, the arrowhead regime runs the whole elimination in two DOF passes
// (arrow_block_solve), which also write u to `uout` when given (*u_done is then set).
static ps_err block_solve_once(Ctx* c, const double* b, double* x, int prec, ps_solve_info* mss_info,
                               double* uout = nullptr, bool* u_done = nullptr) {
    const int64_t n = c->n, nu = c->nu, n5 = 5 * (int64_t)c->Nc;
    hipStream_t st = c->stream;
    if (c->arrow && c->Nc == 1 && c->sc.nm3 > 0) {
        PS_TRY(arrow_block_solve(c, b, x, uout));
        if (u_done) *u_done = uout != nullptr;
        c->cnt.saddle_solves++;
        return PS_OK;
    }
    if (c->p.lub && c->Nc > 0) {
        // Lubrication couples the colloid forces to the colloid velocities (-f_u - R u = b_lambda), so
        // only the monomer forces are fixed: f_m = -b_lambda,m.  With y = M [f_m; 0], the colloid
        // unknowns z = (F, T, S) and w = (U, W) solve the symmetric colloid saddle
        //   [[M_cc, -E], [-E^T, -R]] [z; w] = [b_c - y_c; b_lambda,c]
        // (E: the velocity rows of z), by MINRES on the operator (colloid pairs on the fly); then
        // u = (M f)_u - b_u as without lubrication.
        const int64_t nm3 = c->sc.nm3, nc11 = 11 * (int64_t)c->Nc, N = 17 * (int64_t)c->Nc;
        const size_t ncb = (size_t)std::max(c->Nc, 1);
        double* rhs = c->cs_buf;
        double* v = c->cs_buf + 17 * ncb;
        double* dinv = c->cs_buf + 34 * ncb;
        k_bs_init<<<eblocks(n), VB, 0, st>>>(x, b, n, nm3);
        PS_LAUNCH(c);
        PS_TRY(mob_apply_colloid_rows(c, x, c->bs_y));
        k_cs_rhs<<<eblocks(N), VB, 0, st>>>(b, c->bs_y, n, nm3, c->Nc, rhs, dinv);
        PS_LAUNCH(c);
        auto op = [c](const double* in, double* out) { return colloid_saddle_apply(c, in, out); };
        ps_solve_info si{};
        ps_err r = minres_core(c, N, dinv, op, rhs, v, 0, 1e-12, 10000, &si);
        if (r != PS_OK && r != PS_ERR_NOCONV) return r;
        c->cnt.stresslet_iterations += si.iterations;
        if (r == PS_ERR_NOCONV) c->cnt.stresslet_capped++;
        if (mss_info) mss_info->iterations += si.iterations;
        k_cs_scatter<<<eblocks(nc11), VB, 0, st>>>(v, x, nm3, nc11);
        PS_LAUNCH(c);
        PS_TRY(mob_apply(c, x, c->bs_y, prec));
        k_bs_u<<<eblocks(nu), VB, 0, st>>>(x, c->bs_y, b, n, nu, nullptr);
        PS_LAUNCH(c);
        c->cnt.saddle_solves++;
        return PS_OK;
    }
    k_bs_init<<<eblocks(n), VB, 0, st>>>(x, b, n, nu);
    PS_LAUNCH(c);
    if (n5 > 0 && c->arrow) {
        // the arrowhead's stresslet block is the 5 x 5 floored M^cc block: one apply, then Cholesky
        PS_TRY(mob_apply_stresslet_rows(c, x, c->bs_y));
        PS_TRY(arrow_mss_solve(c, b, c->bs_y, x));
    } else if (n5 > 0) {
        // only the 5Nc stresslet rows of M [f_u; 0] are needed: colloid-row kernels, O(Nm Nc)
        PS_TRY(mob_apply_stresslet_rows(c, x, c->bs_y));
        k_bs_rs<<<eblocks(n5), VB, 0, st>>>(c->bs_r, b, c->bs_y, nu, n5);
        PS_LAUNCH(c);
        ps_solve_info si{};
        ps_err r = minres_small(c, n5, c->Mss, c->bs_r, x + nu, 1e-12, 10000, mss_info ? &si : nullptr);
        if (r != PS_OK && r != PS_ERR_NOCONV) return r;
        if (mss_info) mss_info->iterations += si.iterations;
    }
    PS_TRY(mob_apply(c, x, c->bs_y, prec));
    k_bs_u<<<eblocks(nu), VB, 0, st>>>(x, c->bs_y, b, n, nu, nullptr);
    PS_LAUNCH(c);
    c->cnt.saddle_solves++;
    return PS_OK;
}

// The block elimination in stages, for solves that share one launch of the stresslet MINRES (the
// joint step).  Stage 1 at the current mob() state: x = [-b_lambda; 0], rs = b_s - (M [f_u; 0])_s.
ps_err block_stage1(Ctx* c, const double* b, double* x, double* rs) {
    const int64_t n = c->n, nu = c->nu, n5 = 5 * (int64_t)c->Nc;
    hipStream_t st = c->stream;
    k_bs_init<<<eblocks(n), VB, 0, st>>>(x, b, n, nu);
    PS_LAUNCH(c);
    PS_TRY(mob_apply_stresslet_rows(c, x, c->bs_y));
    k_bs_rs<<<eblocks(n5), VB, 0, st>>>(rs, b, c->bs_y, nu, n5);
    PS_LAUNCH(c);
    return PS_OK;
}

// Stage 3 at the current mob() state, once x's stresslet forces are solved: u = (M f)_u - b_u into
// x[n:] and, when given, uout.
ps_err block_stage3(Ctx* c, const double* b, double* x, int prec, double* uout) {
    const int64_t n = c->n, nu = c->nu;
    PS_TRY(mob_apply(c, x, c->bs_y, prec));
    k_bs_u<<<eblocks(nu), VB, 0, c->stream>>>(x, c->bs_y, b, n, nu, uout);
    PS_LAUNCH(c);
    c->cnt.saddle_solves++;
    return PS_OK;
}

// Whether `nsys` stresslet systems fit one launch of k_ms_wide on this device.
bool mss_joint_fits(const Ctx* c, int nsys) {
    int tpr = 0, nbs = 0;
    const int64_t N = 5 * (int64_t)c->Nc;
    return c->Nc > 0 && nsys >= 1 && nsys <= MS_NSYS && ms_wide_shape(N, &tpr, &nbs) &&
           (int64_t)nbs * nsys <= std::max(1, c->n_cu) && 2 * N * nsys <= 9 * c->ns;
}

// The stresslet solves M_k x_k = rs_k of `nsys` block solves in one launch (rtol 1e-12, cap 10000).
ps_err mss_joint(Ctx* c, int nsys, const double* const* A, const double* const* rs, double* const* x) {
    const int64_t N = 5 * (int64_t)c->Nc;
    if (!mss_joint_fits(c, nsys)) {
        c->set_error("joint stresslet MINRES: " + std::to_string(nsys) + " systems of " + std::to_string(N) +
                     " unknowns do not fit one launch");
        return PS_ERR_ARG;
    }
    MsBatch bt{};
    for (int k = 0; k < nsys; k++) bt.s[k] = MsSys{A[k], rs[k], x[k], c->kr + 2 * N * k, c->mstate + k};
    return ms_wide_launch(c, N, bt, nsys, 1e-12, 10000);
}


static ps_err saddle_solve_into(Ctx* c, const double* b, double* x, int warm, ps_solve_info* info, double* uout,
                                bool* u_done);
// The saddle solve; `uout` (optional) receives u = x[n : n + nu].
ps_err saddle_solve_dev(Ctx* c, const double* b, double* x, int warm, ps_solve_info* info, double* uout) {
    bool u_done = false;
    ps_err rc = saddle_solve_into(c, b, x, warm, info, uout, &u_done);
    if (uout && !u_done && (rc == PS_OK || rc == PS_ERR_NOCONV)) {
        const ps_err r2 = vec_copy(c, uout, x + c->n, c->nu);
        if (r2 != PS_OK) return r2;
    }
    return rc;
}

static ps_err saddle_solve_into(Ctx* c, const double* b, double* x, int warm, ps_solve_info* info, double* uout,
                                bool* u_done) {
    const bool block = c->p.solver == PS_SOLVER_BLOCK;
    const bool mixed = c->p.precision == PS_PREC_MIXED;
    // Block elimination costs one full apply per solve.  In the mixed form that apply runs on FP32
    // tiles; without refinement (ir_maxit = 0) the solve is exact up to that apply's rounding, and
    // with it every sweep adds an FP64 residual apply.
    if (!mixed || (block && c->p.ir_maxit <= 0)) {
        if (!block) return minres_dev(c, b, x, warm, c->p.ksp_rtol, c->p.ksp_maxit, PS_PREC_FP64, info);
        // the stresslet solve's iteration count is read back only when asked for (a host round trip)
        ps_solve_info si{};
        PS_TRY(block_solve_once(c, b, x, c->p.precision, info ? &si : nullptr, uout, u_done));
        if (info) {
            *info = ps_solve_info{};
            info->iterations = si.iterations;
            info->reason = 2;
            info->true_rel_residual = -1.0;
        }
        return PS_OK;
    }
    // Mixed form: FP32-apply corrections, FP64 residual r = b - A x (iterative refinement).
    const int64_t N = c->ns;
    double bn2 = 0.0;
    PS_TRY(vec_dot(c, b, b, N, &bn2));
    const double bn = std::sqrt(bn2);
    int sweeps = 0, its = 0;
    double rel;
    double rn2 = 0.0;
    ps_solve_info si{};
    if (block && !warm) {
        // first sweep straight from b
        PS_TRY(block_solve_once(c, b, x, PS_PREC_MIXED, &si));
        sweeps = 1;
    } else if (!warm) {
        PS_HIP(c, hipMemsetAsync(x, 0, sizeof(double) * (size_t)N, c->stream));
    }
    PS_TRY(saddle_apply(c, x, c->ir_r, PS_PREC_FP64));
    PS_TRY(vec_axpby(c, c->ir_r, 1.0, b, -1.0, N));   // r = b - A x
    PS_TRY(vec_dot(c, c->ir_r, c->ir_r, N, &rn2));
    rel = bn > 0 ? std::sqrt(rn2) / bn : 0.0;
    ps_err rc = PS_OK;
    // FP32 applies carry ~1e-7 relative noise: the inner MINRES stops well above that floor and the
    // FP64 residual of each sweep recovers the accuracy.
    const double inner = std::max(c->p.ksp_rtol, 1e-4);
    while (rel > c->p.ir_rtol && sweeps < std::max(1, c->p.ir_maxit)) {
        if (block) {
            PS_TRY(block_solve_once(c, c->ir_r, c->ir_d, PS_PREC_MIXED, &si));
        } else {
            ps_solve_info ii{};
            ps_err r = minres_dev(c, c->ir_r, c->ir_d, 0, inner, c->p.ksp_maxit, PS_PREC_MIXED, &ii);
            its += ii.iterations;
            if (r != PS_OK && r != PS_ERR_NOCONV) return r;
        }
        PS_TRY(vec_axpby(c, x, 1.0, c->ir_d, 1.0, N));
        PS_TRY(saddle_apply(c, x, c->ir_r, PS_PREC_FP64));
        PS_TRY(vec_axpby(c, c->ir_r, 1.0, b, -1.0, N));
        PS_TRY(vec_dot(c, c->ir_r, c->ir_r, N, &rn2));
        rel = bn > 0 ? std::sqrt(rn2) / bn : 0.0;
        sweeps++;
    }
    if (rel > c->p.ir_rtol) rc = PS_ERR_NOCONV;
    its += si.iterations;
    c->last_minres_its = its;
    if (info) {
        info->iterations = its;
        info->reason = rc == PS_OK ? 2 : -3;
        info->rnorm = rel * bn;
        info->rnorm0 = bn;
        info->refinements = sweeps;
        info->true_rel_residual = rel;
    }
    return rc;
}

// Lanczos scalars in c->lz: alpha_j at [j], beta_j^2 at [LZ_BETA2 + j], |xi|^2 at [LZ_XI2].
constexpr int LZ_BETA2 = 64, LZ_XI2 = 3 * 64 + 1;
constexpr int LZ_G = 64;   // row stride of the weights table c->lzg (one row per m)

// q_0 = xi / |xi| (zero when xi = 0).
__global__ void k_lz_start(int64_t n, const double* __restrict__ xi, double* __restrict__ q0,
                           const double* __restrict__ lz) {
    const double b2 = lz[LZ_XI2];
    const double sc = b2 > 0.0 ? 1.0 / sqrt(b2) : 0.0;
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i < n) q0[i] = xi[i] * sc;
}

// The number of Lanczos vectors before the stopping test: the first j with beta_j <= 1e-12 beta_0
// (an invariant subspace, src/slip_vel.cpp:417) gives m = j + 1, else m = kmax (<= 64).  Lane j
// tests beta_j; every lane of the calling wave must be active.
__device__ __forceinline__ int lz_mlim(const double* __restrict__ lz, int kmax, double beta0) {
    const int j = threadIdx.x & 63;
    const bool stop = j < kmax - 1 && sqrt(fmax(lz[LZ_BETA2 + j], 0.0)) <= 1e-12 * beta0;
    const unsigned long long mask = __ballot(stop);
    return mask ? (int)__builtin_ctzll(mask) + 1 : kmax;
}

// The square-root weights g_m = beta0 U diag(sqrt(max(theta, 0))) U^T e1 of T_m = U diag(theta) U^T
// (src/slip_vel.cpp:427-459) for m = blockIdx.x + 1, one block of two waves each (tql2_pair: the QL
// sweeps on one wave, the eigenvector rows on the other), all m in parallel: the recurrence ran
// ahead to kmax, and the stopping test (k_lz_final) needs the weights of every m up to the one where
// it stops.  Without the test only the last m is solved.
__global__ __launch_bounds__(128) void k_lz_ritz(const double* __restrict__ lz, int kmax, int every_m,
                                                double* __restrict__ G, double* __restrict__ negs) {
    __shared__ Tql2Pair sh;
    __shared__ double Z[64 * 64];
    const int lane = threadIdx.x & 63;
    const int m = blockIdx.x + 1;
    const double beta0 = sqrt(fmax(lz[LZ_XI2], 0.0));
    if (!(beta0 > 0.0)) return;
    const int mlim = lz_mlim(lz, kmax, beta0);
    if (m > mlim || (!every_m && m != mlim)) return;
    const bool bad = (lane < m && !isfinite(lz[lane])) || (lane < m - 1 && !isfinite(lz[LZ_BETA2 + lane]));
    if (__ballot(bad)) {   // a broken recurrence: no QL sweeps over NaNs, the weights say so instead
        if (threadIdx.x < m) G[(m - 1) * LZ_G + lane] = __builtin_nan("");
        if (threadIdx.x == 0) negs[m - 1] = 0.0;
        return;
    }
    const double dv = lane < m ? lz[lane] : 0.0;
    const double ev = (lane < m - 1) ? sqrt(fmax(lz[LZ_BETA2 + lane], 0.0)) : 0.0;
    tql2_pair(m, dv, ev, sh, Z);
    if (threadIdx.x >= 64) return;
    double gi = 0.0;
    if (lane < m)
        for (int jj = 0; jj < m; jj++) {
            double th = sh.de[jj].x;
            if (th < 0.0) th = 0.0;
            gi += Z[lane + jj * 64] * sqrt(th) * Z[0 + jj * 64];
        }
    int nneg = 0;
    for (int jj = 0; jj < m; jj++)
        if (sh.de[jj].x < 0.0) nneg++;
    if (lane < m) G[(m - 1) * LZ_G + lane] = beta0 * gi;
    if (lane == 0) negs[m - 1] = (double)nneg;
}

// Stopping test and recombination.  m* is the first m >= 2 whose relative change
// |g_m - [g_{m-1}; 0]| / |g_m| is at most rtol (rtol > 0), else the m of lz_mlim; out = Q g_{m*}.
// Every block finds m* itself (the weights staged in LDS, one lane per m, the first m that passes
// by ballot); block 0 records it, the negative Ritz values and the last relative change in info and
// adds them to the device counters.
__global__ __launch_bounds__(VB) void k_lz_final(int64_t n, const double* __restrict__ Q, const double* __restrict__ lz,
                                                 int kmax, double rtol, const double* __restrict__ G,
                                                 const double* __restrict__ negs, double* __restrict__ out,
                                                 double* __restrict__ info, int64_t* __restrict__ dcnt) {
    __shared__ double sg[64];
    __shared__ double sG[64 * 65];   // rows m - 1 of G, stride 65 (lanes reading one row each spread over the banks)
    __shared__ int sm;
    const double beta0 = sqrt(fmax(lz[LZ_XI2], 0.0));
    const int mlim = beta0 > 0.0 ? lz_mlim(lz, kmax, beta0) : 0;
    const bool test = rtol > 0.0 && mlim >= 2;
    if (test) {
        for (int k = threadIdx.x; k < mlim * LZ_G; k += VB) sG[(k / LZ_G) * 65 + k % LZ_G] = G[k];
        __syncthreads();
    }
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int mstar = mlim;
        double relc = 0.0;
        if (test) {
            // lane j: the relative change at m = j + 2; m* is the first m that passes
            const int m = lane + 2;
            double dn = 0.0, gn = 0.0;
            if (m <= mlim)
                for (int k = 0; k < m; k++) {
                    const double gi = sG[(m - 1) * 65 + k];
                    const double gp = k < m - 1 ? sG[(m - 2) * 65 + k] : 0.0;
                    dn += (gi - gp) * (gi - gp);
                    gn += gi * gi;
                }
            const double rc = m <= mlim ? sqrt(dn / gn) : 0.0;
            const unsigned long long pass = __ballot(m <= mlim && rc <= rtol);
            const int j = pass ? (int)__builtin_ctzll(pass) : mlim - 2;
            mstar = j + 2;
            relc = __shfl(rc, j, 64);
        }
        sg[lane] = (lane < mstar) ? G[(mstar - 1) * LZ_G + lane] : 0.0;
        if (lane == 0) {
            sm = mstar;
            if (blockIdx.x == 0) {
                const double nneg = mstar > 0 ? negs[mstar - 1] : 0.0;
                info[0] = (double)mstar;
                info[1] = nneg;
                info[2] = relc;
                info[3] = beta0;
                dcnt[DC_LANCZOS_ITERS] += mstar;
                if (nneg > 0.0) {
                    dcnt[DC_NNEG_EVENTS] += 1;
                    dcnt[DC_NNEG_COUNT] += (int64_t)nneg;
                }
            }
        }
    }
    __syncthreads();
    const int m = sm;
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int q = 0; q < m; q++) acc += sg[q] * Q[(int64_t)q * n + i];
    out[i] = acc;
}

// One tridiagonal's weights on the device from host arrays: k_lz_ritz on a recurrence record that
// holds only alpha, beta^2 and beta0^2 (the solve the Lanczos square root runs after its recurrence).
extern "C" ps_err ps_tridiag_sqrt(int m, const double* alpha, const double* beta, double beta0, double* g,
                                  int* m_used, int* nneg) {
    if (m < 1 || m > 64 || !alpha || !g || !m_used || !nneg || (m > 1 && !beta) || !(beta0 >= 0.0)) {
        ps_set_context_free_error("ps_tridiag_sqrt: m must be in 1..64, alpha, beta (m - 1 entries), g, m_used "
                                  "and nneg given, beta0 >= 0");
        return PS_ERR_ARG;
    }
    double lz[3 * 64 + 8] = {};
    for (int i = 0; i < m; i++) lz[i] = alpha[i];
    for (int i = 0; i + 1 < m; i++) lz[LZ_BETA2 + i] = beta[i] * beta[i];
    lz[LZ_XI2] = beta0 * beta0;
    int ml = m;   // lz_mlim's cut, on the host
    for (int j = 0; j + 1 < m; j++)
        if (sqrt(fmax(lz[LZ_BETA2 + j], 0.0)) <= 1e-12 * beta0) {
            ml = j + 1;
            break;
        }
    double *dlz = nullptr, *dg = nullptr;
    if (hipMalloc(&dlz, sizeof(lz)) != hipSuccess || hipMalloc(&dg, sizeof(double) * (LZ_G + 1) * 64) != hipSuccess) {
        if (dlz) (void)hipFree(dlz);
        ps_set_context_free_error("ps_tridiag_sqrt: device allocation failed");
        return PS_ERR_ALLOC;
    }
    double row[64], nn = 0.0;
    hipError_t e = hipMemcpy(dlz, lz, sizeof(lz), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(dg, 0, sizeof(double) * (LZ_G + 1) * 64);
    if (e == hipSuccess) {
        k_lz_ritz<<<m, 128>>>(dlz, m, 0, dg, dg + LZ_G * 64);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(row, dg + (ml - 1) * LZ_G, sizeof(double) * 64, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&nn, dg + LZ_G * 64 + (ml - 1), sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(dlz);
    (void)hipFree(dg);
    if (e != hipSuccess) {
        ps_set_context_free_error(std::string("ps_tridiag_sqrt: ") + hipGetErrorString(e));
        return PS_ERR_HIP;
    }
    for (int i = 0; i < m; i++) g[i] = i < ml ? row[i] : 0.0;
    *m_used = ml;
    *nneg = (int)nn;
    return PS_OK;
}

// w = M^{1/2} xi by Lanczos (schur_sqrt_lanczos, src/slip_vel.cpp:351-477, applied to the whole grand
// mobility) with no host round trip: the recurrence runs kmax iterations on the device (two
// classical Gram-Schmidt passes each), the weights of every m come from two-wave QL solves, and the
// stopping test picks the m the sequential test would have stopped at, so the result equals
// stopping there; vectors built past it are never used.  With `info` the outcome is read back.
ps_err sqrt_apply_dev(Ctx* c, const double* xi, double* out, ps_sqrt_info* info) {
    if (c->arrow) {
        if (info) *info = ps_sqrt_info{};
        return arrow_noise(c, xi, out);
    }
    PS_TRY(sqrt_recurrence_dev(c, xi));
    return sqrt_finish_dev(c, out, info);
}

// The Lanczos recurrence of sqrt_apply_dev: kmax vectors into Q, the scalars into lz.
ps_err sqrt_recurrence_dev(Ctx* c, const double* xi) {
    const int64_t n = c->n;
    const unsigned ne = eblocks(n);
    const int nbd = nblocks(n);
    hipStream_t st = c->stream;
    const int kmax = c->kmax;
    const int prec = c->p.precision;
    double* w = c->kr;   // MINRES scratch is free during the square root
    c->cnt.sqrt_calls++;
    k_dot_part<<<nbd, VB, 0, st>>>(xi, xi, n, c->red);
    PS_LAUNCH(c);
    k_sum_to<<<1, VB, 0, st>>>(c->red, nbd, c->lz + LZ_XI2);
    PS_LAUNCH(c);
    k_lz_start<<<ne, VB, 0, st>>>(n, xi, c->Q, c->lz);
    PS_LAUNCH(c);
    // w alternates between two buffers: iteration j applies M to the previous iteration's w, which is
    // orthogonalised but not yet normalised.  Its finalize sums |w|^2 into beta_{j-1}^2, stores
    // q_j = w / beta_{j-1}, scales M w by 1 / beta_{j-1} (M is linear) and leaves the first
    // Gram-Schmidt pass's partial dots; the two subtract launches leave the second pass's dots and the
    // partials of |w|^2.  The last iteration's beta is never needed.
    const int nb = (int)lzd_blocks(n);
    double* red1 = c->lzred;
    double* red2 = c->lzred + 64 * (size_t)nb;
    double* red3 = c->lzred + 128 * (size_t)nb;
    double* wbuf[2] = {c->kr, c->kr + n};
    for (int j = 0; j < kmax; j++) {
        w = wbuf[j & 1];
        if (j == 0)
            PS_TRY(mob_apply_lanczos(c, c->Q, w, nullptr, nullptr, nb, nullptr, 1, red1, prec));
        else
            PS_TRY(mob_apply_lanczos(c, wbuf[(j - 1) & 1], w, c->Q + (int64_t)j * n, red3, nb,
                                     c->lz + LZ_BETA2 + j - 1, j + 1, red1, prec));
        k_lz_sub_dots<<<nb, LZB, 0, st>>>(n, c->Q, j + 1, w, red1, nb, c->lz + j, red2);
        PS_LAUNCH(c);
        k_lz_sub_norm<<<nb, LZB, 0, st>>>(n, c->Q, j + 1, w, red2, nb, red3);
        PS_LAUNCH(c);
    }
    return PS_OK;
}

// The square-root weights of every m, the stopping test and out = Q g (reads only Q and lz).
ps_err sqrt_finish_dev(Ctx* c, double* out, ps_sqrt_info* info) {
    const int64_t n = c->n;
    const unsigned ne = eblocks(n);
    hipStream_t st = c->stream;
    const int kmax = c->kmax;
    const double rtol = c->p.sqrt_rtol;
    double* negs = c->lzg + LZ_G * 64;
    k_lz_ritz<<<kmax, 128, 0, st>>>(c->lz, kmax, rtol > 0.0 ? 1 : 0, c->lzg, negs);
    PS_LAUNCH(c);
    k_lz_final<<<ne, VB, 0, st>>>(n, c->Q, c->lz, kmax, rtol, c->lzg, negs, out, c->lzinfo, c->dcnt);
    PS_LAUNCH(c);
    if (info) {
        double h[4];
        PS_HIP(c, hipMemcpyAsync(h, c->lzinfo, sizeof(h), hipMemcpyDeviceToHost, st));
        PS_HIP(c, hipStreamSynchronize(st));
        info->iterations = (int)h[0];
        info->nneg = (int)h[1];
        info->rel_change = h[2];
        info->beta0 = h[3];
    }
    return PS_OK;
}
