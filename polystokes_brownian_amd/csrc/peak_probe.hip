// peak_probe.hip -- VALU issue-rate probe of the GPU the library runs on (measurement only; no part
// of the solve).  Each kernel runs independent chains of one instruction form over 16 blocks of 256
// threads per CU; the host times it with events and reads the shader clock (s_memtime against the
// 100 MHz wall clock) inside the kernel.  The result is wave64 instructions issued per CU per shader
// cycle: 1.0 is one instruction per cycle per CU (64 lanes), 2.0 is the dual-issue ceiling that the
// 157.3 TFLOP/s FP32 figure of MI355X assumes.
//
// Forms (the instruction under test, written in inline asm so that the compiler neither folds nor
// reorders it; every operand a VGPR):
//   0  v_fmac_f32 (VOP2 fused multiply-add, the form of the tile loops' accumulations)
//   1  v_fma_f32  (VOP3 fused multiply-add)
//   2  v_fmac_f32 and v_mul_f32 alternating
//   3  v_fma_f64
//   4  v_mov_b32_dpp wave_rol:1 (the tile loops' source-side rotation)
//   5  v_rsq_f32
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int PP_BLOCK = 256;
constexpr int PP_STEPS = 8;   // unrolled steps per loop trip

template <int FORM>
__global__ __launch_bounds__(PP_BLOCK) void k_peak(int iters, float s0, float* __restrict__ out,
                                                   long long* __restrict__ clk) {
    const float s = s0 + threadIdx.x * 1e-9f, c = 0.5f - threadIdx.x * 1e-9f;
    const double ds = s, dc = c;
    float a0 = threadIdx.x * 1e-3f, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6,
          a7 = a0 + 7;
    double d0 = a0, d1 = a1, d2 = a2, d3 = a3, d4 = a4, d5 = a5, d6 = a6, d7 = a7;
    const long long c0 = clock64(), w0 = wall_clock64();
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int u = 0; u < PP_STEPS; u++) {
#define PP8(OP) OP(a0); OP(a1); OP(a2); OP(a3); OP(a4); OP(a5); OP(a6); OP(a7)
#define PP8D(OP) OP(d0); OP(d1); OP(d2); OP(d3); OP(d4); OP(d5); OP(d6); OP(d7)
#define PP_FMAC(a) asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(a) : "v"(s), "v"(c))
#define PP_FMA(a) asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(a) : "v"(s), "v"(c))
#define PP_MUL(a) asm volatile("v_mul_f32 %0, %0, %1" : "+v"(a) : "v"(s))
#define PP_FMA64(a) asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(a) : "v"(ds), "v"(dc))
#define PP_DPP(a) asm volatile("v_mov_b32_dpp %0, %0 wave_rol:1 row_mask:0xf bank_mask:0xf" : "+v"(a))
#define PP_RSQ(a) asm volatile("v_rsq_f32 %0, %0" : "+v"(a))
            if constexpr (FORM == 0) { PP8(PP_FMAC); }
            if constexpr (FORM == 1) { PP8(PP_FMA); }
            if constexpr (FORM == 2) {
                PP_FMAC(a0); PP_MUL(a1); PP_FMAC(a2); PP_MUL(a3); PP_FMAC(a4); PP_MUL(a5); PP_FMAC(a6); PP_MUL(a7);
            }
            if constexpr (FORM == 3) { PP8D(PP_FMA64); }
            if constexpr (FORM == 4) { PP8(PP_DPP); }
            if constexpr (FORM == 5) { PP8(PP_RSQ); }
        }
    }
    const long long c1 = clock64(), w1 = wall_clock64();
    const int g = blockIdx.x * PP_BLOCK + threadIdx.x;
    out[g] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7 + (float)(d0 + d1 + d2 + d3 + d4 + d5 + d6 + d7);
    if (g == 0) {
        clk[0] = c1 - c0;
        clk[1] = w1 - w0;
    }
}

template <int FORM>
hipError_t run_form(int nblk, int iters, float* out, long long* clk, hipEvent_t e0, hipEvent_t e1, float* ms) {
    k_peak<FORM><<<nblk, PP_BLOCK>>>(iters, 0.999f, out, clk);   // warm-up
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = hipEventRecord(e0)) != hipSuccess) return e;
    k_peak<FORM><<<nblk, PP_BLOCK>>>(iters, 0.999f, out, clk);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipEventRecord(e1)) != hipSuccess) return e;
    if ((e = hipEventSynchronize(e1)) != hipSuccess) return e;
    return hipEventElapsedTime(ms, e0, e1);
}

}  // namespace

extern "C" {

// Issue rate of one instruction form on the current device.  Out: ms (kernel time), ghz (shader clock
// from s_memtime / wall clock), ipc (wave64 instructions per CU per shader cycle).  Returns 0, or a
// hipError_t code.
int psp_valu_rate(int form, double* ms, double* ghz, double* ipc) {
    int dev = 0, ncu = 0, wclk_khz = 0;
    hipError_t e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return (int)e;
    if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return (int)e;
    if ((e = hipDeviceGetAttribute(&wclk_khz, hipDeviceAttributeWallClockRate, dev)) != hipSuccess) return (int)e;
    const int nblk = 16 * ncu, iters = form == 5 ? 250 : 1000;
    float* out = nullptr;
    long long* clk = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float t = 0.0f;
    long long h[2] = {0, 0};
    if ((e = hipMalloc(&out, sizeof(float) * (size_t)nblk * PP_BLOCK)) != hipSuccess) return (int)e;
    if ((e = hipMalloc(&clk, sizeof(long long) * 2)) != hipSuccess) goto done;
    if ((e = hipEventCreate(&e0)) != hipSuccess) goto done;
    if ((e = hipEventCreate(&e1)) != hipSuccess) goto done;
    switch (form) {
        case 0: e = run_form<0>(nblk, iters, out, clk, e0, e1, &t); break;
        case 1: e = run_form<1>(nblk, iters, out, clk, e0, e1, &t); break;
        case 2: e = run_form<2>(nblk, iters, out, clk, e0, e1, &t); break;
        case 3: e = run_form<3>(nblk, iters, out, clk, e0, e1, &t); break;
        case 4: e = run_form<4>(nblk, iters, out, clk, e0, e1, &t); break;
        case 5: e = run_form<5>(nblk, iters, out, clk, e0, e1, &t); break;
        default: e = hipErrorInvalidValue;
    }
    if (e != hipSuccess) goto done;
    if ((e = hipMemcpy(h, clk, sizeof(h), hipMemcpyDeviceToHost)) != hipSuccess) goto done;
    {
        const double f = (double)h[0] / ((double)h[1] / (wclk_khz * 1e3));
        const double instr = (double)nblk * (PP_BLOCK / 64) * iters * PP_STEPS * 8;
        *ms = t;
        *ghz = f * 1e-9;
        *ipc = instr / (t * 1e-3) / ncu / f;
    }
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (clk) (void)hipFree(clk);
    if (out) (void)hipFree(out);
    return (int)e;
}

}  // extern "C"
