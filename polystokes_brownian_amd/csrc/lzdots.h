// lzdots.h -- block partials of many dot products for the Lanczos Gram-Schmidt passes (device code),
// shared by the Lanczos finalize (mobility.hip) and the subtract kernels (krylov.hip).
//
// The Lanczos vector kernels run blocks of LZB threads, one row of the vectors per thread, so a vector
// of n rows leaves nb = ceil(n / LZB) block partials per dot product.  lzd_block_dots forms one partial
// per dot and block in a fixed order (DPP sums within each wave, then the wave sums in a fixed tree);
// lzd_sum gives every block the same h_q = sum over the nb partials.
#pragma once
#include "ps_common.h"
#include "dpp.h"

#include <hip/hip_runtime.h>

// PS_LZB: 256, 512 or 1024.  256 measured fastest at N = 20,000 (per Lanczos iteration, finalize plus
// both subtract kernels: 34.5 us at 256, 36.2 us at 512, 48 us at 1024): with fewer, larger blocks the
// finalize's reads of the apply partials have too few rows in flight.
#ifndef PS_LZB
#define PS_LZB 256
#endif
constexpr int LZB = PS_LZB;      // threads (rows) per block of the Lanczos vector kernels
constexpr int LZW = LZB / 64;    // waves per block
static_assert(LZW == 4 || LZW == 8 || LZW == 16, "256, 512 or 1024 threads per block");
__host__ __device__ inline int64_t lzd_blocks(int64_t n) { return (n + LZB - 1) / LZB; }

// The block partials of m <= 64 dot products: contrib(q) is this thread's row's term of dot q (zero
// for rows past n).  Per q every wave sums its 64 rows and lane 0 leaves the sum in sh[q][wave]; after
// a block barrier thread q adds the LZW wave sums in a fixed tree into red[q * nb + block].  Every
// thread of the block calls it.  sh is not written again by this call after the barrier.
template <class F>
__device__ __forceinline__ void lzd_block_dots(int m, F&& contrib, double (*sh)[LZW], double* __restrict__ red,
                                               int nb) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // eight terms at a time, all loaded before any is summed (the index clamped to m - 1, so that the
    // loads need no branch; the padded terms are not summed), the next eight loaded while these are
    // summed
    double c[8], nx[8];
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = contrib(min(k, m - 1));
    for (int q0 = 0; q0 < m; q0 += 8) {
        if (q0 + 8 < m) {
#pragma unroll
            for (int k = 0; k < 8; k++) nx[k] = contrib(min(q0 + 8 + k, m - 1));
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (q0 + k < m) {
                const double s = wave_sum64(c[k]);
                if (lane == 0) sh[q0 + k][wid] = s;
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) c[k] = nx[k];
    }
    __syncthreads();
    if ((int)threadIdx.x < m) {
        double r[LZW];
#pragma unroll
        for (int k = 0; k < LZW; k++) r[k] = sh[threadIdx.x][k];
#pragma unroll
        for (int w = LZW / 2; w > 0; w >>= 1)
#pragma unroll
            for (int k = 0; k < w; k++) r[k] = r[2 * k] + r[2 * k + 1];
        red[(int64_t)threadIdx.x * nb + blockIdx.x] = r[0];
    }
}

// h[q] = sum over b < nb of red[q * nb + b] for q < m (h in LDS, the same bits in every block): the
// 16 threads of one DPP row per q (LZB / 16 q at a time), each with four interleaved accumulators (so
// that its loads are in flight together), combined within the row.  Every thread of the block calls it; ends with a block
// barrier.
__device__ __forceinline__ void lzd_sum(const double* __restrict__ red, int nb, int m, double* h) {
    const int t = threadIdx.x, s = t & 15;
    for (int q0 = 0; q0 < m; q0 += LZB / 16) {
        const int q = q0 + (t >> 4);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        if (q < m) {
            const double* rq = red + (int64_t)q * nb;
            int b = s;
            for (; b + 48 < nb; b += 64) {
                a0 += rq[b];
                a1 += rq[b + 16];
                a2 += rq[b + 32];
                a3 += rq[b + 48];
            }
            if (b < nb) a0 += rq[b];
            if (b + 16 < nb) a1 += rq[b + 16];
            if (b + 32 < nb) a2 += rq[b + 32];
        }
        const double v = row_sum16((a0 + a1) + (a2 + a3));
        if (s == 0 && q < m) h[q] = v;
    }
    __syncthreads();
}
