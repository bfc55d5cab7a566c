// dense.hip -- the stored-dense grand mobility in 16-bit storage and its multi-RHS apply on MFMA.
//
// The reference assembles M densely in mob() (MatSetValues INSERT, src/mob.cpp:635-1213) and applies
// it with PETSc's dense MatMult.  Here M (n x n, FP64 tile algebra) is rounded once to BF16 or FP16
// and stored in the operand order of v_mfma_f32_16x16x32_{bf16,f16}, so the apply of up to 16
// vectors streams every 1 KiB tile with one coalesced load per lane and accumulates in FP32.  The
// apply reads n^2 x 2 bytes per call whatever the number of vectors: HBM-bound.
//
// Storage (np = n rounded up to 32): tile (rb, kb) holds rows 16 rb .. +16 and columns 32 kb .. +32
// at element offset (rb * nkb + kb) * 512; element (i, k) of a tile sits in lane (i % 16) +
// 16 ((k % 32) / 8), slot k % 8 -- the A-operand layout (lane l: row l % 16, k = 8 (l / 16) .. +8).
// The vectors use the B-operand layout per 32-row k-block: lane l holds vector l % 16, rows
// 8 (l / 16) .. +8.  The result lane l holds rows 4 (l / 16) .. +4 of vector l % 16.
#include "ps_common.h"
#include "tiles.h"

typedef __bf16 ps_bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 ps_f16x8 __attribute__((ext_vector_type(8)));
typedef float ps_f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int DT = PS_BLOCK;

// Degree of freedom r -> particle, kind (0 monomer translation, 1 colloid force, 2 colloid torque,
// 3 colloid stresslet) and component, in the layout [f_m; F_c; T_c; S_c].
struct Dof {
    int p, kind, comp;
};
__device__ __forceinline__ Dof dof_of(const SysConst& s, int64_t r) {
    if (r < s.nm3) return Dof{(int)(r / 3), 0, (int)(r % 3)};
    if (r < s.nm3 + s.nc3) {
        const int64_t q = r - s.nm3;
        return Dof{s.Nm + (int)(q / 3), 1, (int)(q % 3)};
    }
    if (r < s.nu) {
        const int64_t q = r - s.nm3 - s.nc3;
        return Dof{s.Nm + (int)(q / 3), 2, (int)(q % 3)};
    }
    const int64_t q = r - s.nu;
    return Dof{s.Nm + (int)(q / 5), 3, (int)(q % 5)};
}

// M[r][c] with the placement of the dense assembly (fill_self, src/mob.cpp:539-633; AA pairs
// :843-897; AB :900-974 with the carried translation tile `stale` (3x3, row-major); BB :1026-1144).
__device__ double dense_entry(const SysConst& s, double beta, const double* __restrict__ xm,
                              const int* __restrict__ chain, const double* __restrict__ stale, int64_t r,
                              int64_t c) {
    const Dof R = dof_of(s, r), C = dof_of(s, c);
    const int a = R.comp, b = C.comp;
    if (R.p == C.p) {
        if (R.kind != C.kind) return 0.0;
        switch (R.kind) {
            case 0: return a == b ? s.beta_inv : 0.0;
            case 1: return a == b ? 1.0 : 0.0;
            case 2: return a == b ? 3. / 4. : 0.0;
            default: return a == b ? 9. / 5. : (((a == 0 && b == 4) || (a == 4 && b == 0)) ? 9. / 10. : 0.0);
        }
    }
    double e[3], dr, dr_inv;
    if (R.kind == 0 && C.kind == 0) {
        if (s.chain_HI && chain[R.p] == chain[C.p]) return 0.0;   // never inserted
        pair_geom_dev(s, xm, min(R.p, C.p), max(R.p, C.p), e, dr, dr_inv);
        double t[3][3];
        tile_aa_dev(dr, dr_inv, e, beta, t);
        return t[a][b];
    }
    if (R.kind == 0 || C.kind == 0) {
        // monomer-colloid: geometry with the monomer first (src/mob.cpp:904)
        const bool mrow = R.kind == 0;
        const int mono = mrow ? R.p : C.p, col = mrow ? C.p : R.p;
        const int ck = mrow ? C.kind : R.kind;
        if (ck == 1) return mrow ? stale[3 * a + b] : stale[3 * b + a];
        pair_geom_dev(s, xm, mono, col, e, dr, dr_inv);
        double bb[3][3], bt[3][3], gt[3][5];
        tile_ab_dev(dr_inv, e, s.beta2, bb, bt, gt);
        if (ck == 2) return mrow ? bt[a][b] : bb[a][b];
        return mrow ? gt[a][b] : gt[b][a];
    }
    // colloid-colloid, tile of the pair (c1 < c2)
    const int c1 = min(R.p, C.p), c2 = max(R.p, C.p);
    pair_geom_dev(s, xm, c1, c2, e, dr, dr_inv);
    BBT T;
    tile_bb(dr_inv, e, T);
    const int kr = R.kind, kc = C.kind;
    if (R.p == c1) {
        if (kr == 1 && kc == 1) return T.a[a][b];
        if (kr == 2 && kc == 1) return T.b[a][b];
        if (kr == 1 && kc == 2) return -T.bt[b][a];
        if (kr == 2 && kc == 2) return T.c[a][b];
        if (kr == 1 && kc == 3) return T.gt[a][b];
        if (kr == 2 && kc == 3) return T.ht[a][b];
        if (kr == 3 && kc == 1) return -T.gt[b][a];
        if (kr == 3 && kc == 2) return T.ht[b][a];
        return T.m[a][b];
    }
    if (kr == 1 && kc == 1) return T.a[b][a];
    if (kr == 1 && kc == 2) return T.bt[a][b];
    if (kr == 2 && kc == 1) return -T.b[b][a];
    if (kr == 2 && kc == 2) return T.c[b][a];
    if (kr == 1 && kc == 3) return -T.gt[a][b];
    if (kr == 3 && kc == 1) return T.gt[b][a];
    if (kr == 2 && kc == 3) return T.ht[a][b];
    if (kr == 3 && kc == 2) return T.ht[b][a];
    return T.m[b][a];
}

template <class H>
__device__ __forceinline__ H to_h(double v);
template <>
__device__ __forceinline__ __bf16 to_h<__bf16>(double v) { return (__bf16)(float)v; }
template <>
__device__ __forceinline__ _Float16 to_h<_Float16>(double v) { return (_Float16)(float)v; }

// One thread per lane slot group: 8 entries of one row of one tile.
template <class H>
__global__ __launch_bounds__(DT) void k_dense_build(SysConst s, double beta, const double* __restrict__ xm,
                                                    const int* __restrict__ chain, const double* __restrict__ stale,
                                                    int64_t nkb, int64_t ntiles, H* __restrict__ Mb) {
    const int64_t g = (int64_t)blockIdx.x * DT + threadIdx.x;   // = tile * 64 + lane
    if (g >= ntiles * 64) return;
    const int64_t tile = g >> 6;
    const int lane = (int)(g & 63);
    const int64_t rb = tile / nkb, kb = tile - rb * nkb;
    const int64_t row = 16 * rb + (lane & 15);
    const int64_t col0 = 32 * kb + 8 * (lane >> 4);
    H* out = Mb + g * 8;
    for (int e = 0; e < 8; e++) {
        const int64_t col = col0 + e;
        const double v = (row < s.n && col < s.n) ? dense_entry(s, beta, xm, chain, stale, row, col) : 0.0;
        out[e] = to_h<H>(v);
    }
}

// Vectors (vector-major, X[j n + i]) -> B-operand layout [nkb][64][8].
template <class H>
__global__ void k_dense_x(int64_t n, int nrhs, int64_t nkb, const double* __restrict__ X, H* __restrict__ Xb) {
    const int64_t g = (int64_t)blockIdx.x * DT + threadIdx.x;
    if (g >= nkb * 64) return;
    const int64_t kb = g >> 6;
    const int lane = (int)(g & 63);
    const int j = lane & 15;
    const int64_t k0 = 32 * kb + 8 * (lane >> 4);
    for (int e = 0; e < 8; e++) {
        const int64_t k = k0 + e;
        Xb[g * 8 + e] = to_h<H>((j < nrhs && k < n) ? X[(int64_t)j * n + k] : 0.0);
    }
}

template <class H>
struct MfmaOp;
template <>
struct MfmaOp<__bf16> {
    using V = ps_bf16x8;
    static __device__ __forceinline__ ps_f32x4 run(V a, V b, ps_f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct MfmaOp<_Float16> {
    using V = ps_f16x8;
    static __device__ __forceinline__ ps_f32x4 run(V a, V b, ps_f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// Y[j n + i] = sum_k M[i][k] X[j n + k]: one wave per 16-row block, the k-blocks streamed in order,
// two accumulators to shorten the MFMA dependency chain.  The stored M is read once per call with
// non-temporal loads, so it does not evict the vector tiles every wave re-reads from L2 (measured:
// 5.36 against 5.04 TB/s at n = 24,640; splitting k over the waves of a block gained nothing).
template <class H>
__global__ __launch_bounds__(DT) void k_dense_apply(int64_t n, int nrhs, int64_t nrb, int64_t nkb,
                                                    const H* __restrict__ Mb, const H* __restrict__ Xb,
                                                    double* __restrict__ Y) {
    using V = typename MfmaOp<H>::V;
    const int lane = threadIdx.x & 63;
    const int64_t rb = (int64_t)blockIdx.x * (DT / 64) + (threadIdx.x >> 6);
    if (rb >= nrb) return;
    const V* A = reinterpret_cast<const V*>(Mb) + rb * nkb * 64 + lane;
    const V* B = reinterpret_cast<const V*>(Xb) + lane;
    ps_f32x4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
    int64_t kb = 0;
    for (; kb + 8 <= nkb; kb += 8) {
        V a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            a[u] = __builtin_nontemporal_load(&A[(kb + u) * 64]);
            b[u] = B[(kb + u) * 64];
        }
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
            acc0 = MfmaOp<H>::run(a[u], b[u], acc0);
            acc1 = MfmaOp<H>::run(a[u + 1], b[u + 1], acc1);
        }
    }
    for (; kb < nkb; kb++) acc0 = MfmaOp<H>::run(A[kb * 64], B[kb * 64], acc0);
    const int j = lane & 15;
    if (j >= nrhs) return;
    for (int r = 0; r < 4; r++) {
        const int64_t row = 16 * rb + 4 * (lane >> 4) + r;
        if (row < n) Y[(int64_t)j * n + row] = (double)(acc0[r] + acc1[r]);
    }
}

}  // namespace

ps_err dense_build(Ctx* c, int storage) {
    if (!c->have_mob) {
        c->set_error("dense_build: call ps_mob first");
        return PS_ERR_STATE;
    }
    if (storage != PS_DENSE_BF16 && storage != PS_DENSE_F16) {
        c->set_error("dense_build: storage must be PS_DENSE_BF16 or PS_DENSE_F16");
        return PS_ERR_ARG;
    }
    const int64_t n = c->n;
    const int64_t np = (n + 31) / 32 * 32;
    const int64_t nrb = np / 16, nkb = np / 32;
    const size_t bytes = (size_t)np * (size_t)np * 2;
    if (!c->dense_M || c->dense_bytes < bytes) {
        if (c->dense_M) (void)hipFree(c->dense_M);
        c->dense_M = nullptr;
        if (hipMalloc(&c->dense_M, bytes) != hipSuccess) {
            (void)hipGetLastError();
            c->set_error("dense_build: cannot allocate " + std::to_string(bytes) + " bytes for the stored M");
            return PS_ERR_ALLOC;
        }
        c->dense_bytes = bytes;
        if (c->dense_X) (void)hipFree(c->dense_X);
        PS_HIP(c, hipMalloc(&c->dense_X, (size_t)nkb * 64 * 8 * 2));
    }
    const int64_t ntiles = nrb * nkb;
    const unsigned gb = (unsigned)((ntiles * 64 + DT - 1) / DT);
    if (storage == PS_DENSE_BF16)
        k_dense_build<__bf16><<<gb, DT, 0, c->stream>>>(c->sc, c->p.beta, c->xm, c->chain, c->stale, nkb, ntiles,
                                                         (__bf16*)c->dense_M);
    else
        k_dense_build<_Float16><<<gb, DT, 0, c->stream>>>(c->sc, c->p.beta, c->xm, c->chain, c->stale, nkb,
                                                           ntiles, (_Float16*)c->dense_M);
    PS_LAUNCH(c);
    c->dense_storage = storage;
    return PS_OK;
}

ps_err dense_mult_dev(Ctx* c, const double* X, double* Y, int nrhs) {
    if (!c->dense_M || c->dense_storage == 0) {
        c->set_error("dense_mult: call ps_dense_build first");
        return PS_ERR_STATE;
    }
    if (nrhs < 1 || nrhs > 16) {
        c->set_error("dense_mult: 1 <= nrhs <= 16");
        return PS_ERR_ARG;
    }
    const int64_t n = c->n;
    const int64_t np = (n + 31) / 32 * 32;
    const int64_t nrb = np / 16, nkb = np / 32;
    hipStream_t st = c->stream;
    const unsigned gx = (unsigned)((nkb * 64 + DT - 1) / DT);
    const unsigned ga = (unsigned)((nrb + DT / 64 - 1) / (DT / 64));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->prof_on) PS_TRY(c->prof_record(2, &e0, &e1));
    if (c->dense_storage == PS_DENSE_BF16) {
        k_dense_x<__bf16><<<gx, DT, 0, st>>>(n, nrhs, nkb, X, (__bf16*)c->dense_X);
        PS_LAUNCH(c);
        if (e0) PS_HIP(c, hipEventRecord(e0, st));
        k_dense_apply<__bf16><<<ga, DT, 0, st>>>(n, nrhs, nrb, nkb, (const __bf16*)c->dense_M,
                                                  (const __bf16*)c->dense_X, Y);
    } else {
        k_dense_x<_Float16><<<gx, DT, 0, st>>>(n, nrhs, nkb, X, (_Float16*)c->dense_X);
        PS_LAUNCH(c);
        if (e0) PS_HIP(c, hipEventRecord(e0, st));
        k_dense_apply<_Float16><<<ga, DT, 0, st>>>(n, nrhs, nrb, nkb, (const _Float16*)c->dense_M,
                                                    (const _Float16*)c->dense_X, Y);
    }
    PS_LAUNCH(c);
    if (e1) PS_HIP(c, hipEventRecord(e1, st));
    return PS_OK;
}

void dense_free(Ctx* c) {
    if (c->dense_M) (void)hipFree(c->dense_M);
    if (c->dense_X) (void)hipFree(c->dense_X);
    c->dense_M = nullptr;
    c->dense_X = nullptr;
    c->dense_bytes = 0;
    c->dense_storage = 0;
}
