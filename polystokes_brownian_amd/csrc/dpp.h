// dpp.h -- wave-level sums by DPP lane moves (device code), shared by the Krylov, Lanczos and
// batched-replica kernels.
//
// A 64-bit value moves as its two 32-bit halves (v_mov_b32_dpp with every row and bank enabled).
// Every lane of the wave must be active and hold a value (zero where it contributes nothing).
#pragma once
#include <hip/hip_runtime.h>

template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ double lane_read(double v, int lane) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float lane_read(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
// Sum over each row of 16 lanes, left in every lane of the row: quad sums, then the row's quads.
template <class T>
__device__ __forceinline__ T row_sum16(T v) {
    v += dpp_mov<0xB1>(v);    // quad_perm [1, 0, 3, 2]
    v += dpp_mov<0x4E>(v);    // quad_perm [2, 3, 0, 1]
    v += dpp_mov<0x124>(v);   // row_ror:4
    v += dpp_mov<0x128>(v);   // row_ror:8
    return v;
}
// Sum over the 64 lanes, the same bits in every lane: row sums by DPP, then the sums held by lanes 0,
// 16, 32 and 48 in a fixed order.
template <class T>
__device__ __forceinline__ T wave_sum64(T v) {
    v = row_sum16(v);
    return (lane_read(v, 0) + lane_read(v, 16)) + (lane_read(v, 32) + lane_read(v, 48));
}
