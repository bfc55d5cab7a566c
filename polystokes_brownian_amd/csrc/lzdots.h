// lzdots.h -- block partials of many dot products for the Lanczos Gram-Schmidt passes (device code),
// shared by the Lanczos finalize (mobility.hip) and the subtract kernels (krylov.hip).
//
// A block of 256 threads owns 256 rows of the vectors.  Each thread leaves its row's contributions
// c_q (q < m, at most LZD_CH at a time) in LDS (lds[q * 256 + thread]); lzd_store reduces them to one
// partial per q and block, red[q * nb + block], with a fixed order (the four 64-row quarters, then the
// wave by row DPP moves and lane reads); lzd_sum gives every block the same h_q = sum over the nb
// partials.
#pragma once
#include "ps_common.h"

#include <hip/hip_runtime.h>

constexpr int LZD_T = 256;     // threads (rows) per block
constexpr int LZD_CH = 32;     // basis vectors per pass over the LDS rows (32 x 256 doubles = 64 KiB)
static_assert(LZD_T == PS_BLOCK, "the Lanczos dot kernels run one row per thread in blocks of PS_BLOCK");

template <int CTRL>
__device__ __forceinline__ double lzd_dpp(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lzd_lane(double v, int lane) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
// Sum over the 64 lanes, the same bits in every lane: quad and row sums by DPP (every lane of a row of
// 16 ends with a sum of the row), then the sums held by lanes 0, 16, 32 and 48 in a fixed order.
__device__ __forceinline__ double lzd_wave_sum(double v) {
    v += lzd_dpp<0xB1>(v);    // quad_perm [1, 0, 3, 2]
    v += lzd_dpp<0x4E>(v);    // quad_perm [2, 3, 0, 1]
    v += lzd_dpp<0x124>(v);   // row_ror:4
    v += lzd_dpp<0x128>(v);   // row_ror:8
    return (lzd_lane(v, 0) + lzd_lane(v, 16)) + (lzd_lane(v, 32) + lzd_lane(v, 48));
}

// red[q * nb + blockIdx.x] = the block's sum of lds[q * 256 + t] over its 256 threads, q < m.  The
// caller has filled lds and synchronised the block.
__device__ __forceinline__ void lzd_store(const double* lds, int m, double* __restrict__ red, int nb) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int q = wid; q < m; q += LZD_T / 64) {
        const double* row = lds + q * LZD_T;
        const double v = lzd_wave_sum((row[lane] + row[lane + 64]) + (row[lane + 128] + row[lane + 192]));
        if (lane == 0) red[(int64_t)q * nb + blockIdx.x] = v;
    }
}

// h[q] = sum over b < nb of red[q * nb + b] for q < m (h in LDS, the same bits in every block): four
// threads per q, each with four interleaved accumulators (so that the loads are in flight together),
// combined in a fixed order.  Ends with a block barrier.
__device__ __forceinline__ void lzd_sum(const double* __restrict__ red, int nb, int m, double* h) {
    const int t = threadIdx.x, q = t >> 2, s = t & 3;
    double v = 0.0;
    if (q < m) {
        const double* rq = red + (int64_t)q * nb;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int b = s;
        for (; b + 12 < nb; b += 16) {
            a0 += rq[b];
            a1 += rq[b + 4];
            a2 += rq[b + 8];
            a3 += rq[b + 12];
        }
        for (; b < nb; b += 4) a0 += rq[b];
        v = (a0 + a1) + (a2 + a3);
    }
    // the four threads of a q are consecutive lanes: (v0 + v1) + (v2 + v3) in each of them
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    if (s == 0 && q < m) h[q] = v;
    __syncthreads();
}
