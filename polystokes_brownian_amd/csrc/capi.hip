// capi.hip -- the C ABI (include/polystokes_gpu.h), the integrator kernels and the run() loop.
//
// The step loop restates PolyStokes::run (src/run.cpp:70-162) with every array resident in HBM:
// predictor (check_dist, RHS, mob, deterministic solve), Brownian slip (M^{1/2} xi, second solve),
// RFD thermal drift (two probe solves at x +/- eps W, src/drift.cpp:17-62), Euler probe step
// (src/traj.cpp:9-35), corrector (check_dist, RHS, mob, solve), trapezoidal update from x_n
// (src/traj.cpp:37-62) and the quaternion update (src/traj.cpp:64-94, src/quat.cpp:40-54).
// The noise stream is the reference's: std::mt19937_64 seeded with seed + llround(restart/dt) and
// std::normal_distribution<double>, n draws for the slip then 3Nm+6Nc for the RFD per step
// (src/slip_vel.cpp:11-17, src/drift.cpp:9-15), drawn on the host into pinned memory.
#include "ps_common.h"

#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <sstream>
#include <cmath>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>

#pragma clang fp contract(off)

namespace {

constexpr int TB = PS_BLOCK;
inline unsigned eb(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TB - 1) / TB); }

// step(): undo the RFD probe, add the drift and the velocity, wrap (src/traj.cpp:9-35).
__global__ void k_step(SysConst s, double* __restrict__ x, const double* __restrict__ fext,
                       const double* __restrict__ udiff, const double* __restrict__ up, double dt, double kT) {
    const int p = blockIdx.x * TB + threadIdx.x;
    if (p >= s.Np) return;
    const double eps = 1.0e-3;
    double v[3];
    for (int d = 0; d < 3; d++) {
        const int64_t i = 3 * (int64_t)p + d;
        double xi = x[i];
        if (kT > 0) {
            const double drift_scale = kT / (2.0 * eps);
            xi += eps * fext[i];
            xi += drift_scale * udiff[i] * dt;
        }
        xi += up[i] * dt;
        v[d] = xi;
    }
    if (s.box_active)
        for (int d = 0; d < 3; d++) v[d] -= s.L[d] * __builtin_rint(v[d] * s.Linv[d]);
    for (int d = 0; d < 3; d++) x[3 * (int64_t)p + d] = v[d];
}

// step_from(x_base) (src/traj.cpp:37-62).
__global__ void k_step_from(SysConst s, double* __restrict__ x, const double* __restrict__ xb,
                            const double* __restrict__ udiff, const double* __restrict__ up, double dt, double kT) {
    const int p = blockIdx.x * TB + threadIdx.x;
    if (p >= s.Np) return;
    const double eps = 1.0e-3;
    double v[3];
    for (int d = 0; d < 3; d++) {
        const int64_t i = 3 * (int64_t)p + d;
        if (kT > 0) {
            const double drift_scale = kT / (2.0 * eps);
            v[d] = xb[i] + drift_scale * udiff[i] * dt + up[i] * dt;
        } else {
            v[d] = xb[i] + up[i] * dt;
        }
    }
    if (s.box_active)
        for (int d = 0; d < 3; d++) v[d] -= s.L[d] * __builtin_rint(v[d] * s.Linv[d]);
    for (int d = 0; d < 3; d++) x[3 * (int64_t)p + d] = v[d];
}

// x += a * w over the first n3 entries (the RFD probe displacements, src/drift.cpp:33,48).
// x += a w; with xpd, the probe's geometry snapshot (xpd = x) in the same pass
__global__ void k_probe(double* __restrict__ x, const double* __restrict__ w, double a, int64_t n3,
                        double* __restrict__ xpd) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n3) return;
    const double v = x[i] + a * w[i];
    x[i] = v;
    if (xpd) xpd[i] = v;
}

// Move to an RFD probe point and take the geometry its mob() reads (the arrowhead regime runs its
// full check_dist there instead).
ps_err probe_move(Ctx* c, double a) {
    const int64_t n3 = 3 * (int64_t)c->Np;
    k_probe<<<eb(n3), TB, 0, c->stream>>>(c->x, c->fext, a, n3, c->arrow ? nullptr : c->xpd);
    PS_LAUNCH(c);
    if (c->arrow) return probe_geometry_dev(c);
    c->have_pd = true;
    return PS_OK;
}

// step_quaternion (src/traj.cpp:64-94, src/quat.cpp:4-54): q <- normalize(q + dt/2 q (x) (0, omega)).
__global__ void k_quat(int Nc, int64_t nm3nc3, double* __restrict__ q, const double* __restrict__ up, double dt) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= Nc) return;
    double* qi = q + 4 * (int64_t)i;
    const double* om = up + nm3nc3 + 3 * (int64_t)i;
    const double w1 = qi[0], x1 = qi[1], y1 = qi[2], z1 = qi[3];
    const double w2 = 0.0, x2 = om[0], y2 = om[1], z2 = om[2];
    const double dq[4] = {w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                          w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2};
    double qn[4];
    for (int k = 0; k < 4; k++) qn[k] = qi[k] + 0.5 * dq[k] * dt;
    const double norm = __builtin_sqrt(qn[0] * qn[0] + qn[1] * qn[1] + qn[2] * qn[2] + qn[3] * qn[3]);
    for (int k = 0; k < 4; k++) qi[k] = qn[k] / norm;
}

// out[i] = X[off + i]
__global__ void k_extract(double* __restrict__ out, const double* __restrict__ X, int64_t off, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i < n) out[i] = X[off + i];
}
// c = a - b
__global__ void k_sub(double* __restrict__ c, const double* __restrict__ a, const double* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i < n) c[i] = a[i] - b[i];
}
// up = 0.5 (v_det_n + v_det_c) + v_brown
__global__ void k_trapz(double* __restrict__ up, const double* __restrict__ a, const double* __restrict__ b,
                        const double* __restrict__ vb, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i < n) up[i] = 0.5 * (a[i] + b[i]) + vb[i];
}
// rhs[0:n] += a * w
__global__ void k_add_slip(double* __restrict__ rhs, const double* __restrict__ w, double a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i < n) rhs[i] += a * w[i];
}
// sb = X[nu : nu+nc5] (first probe) / (sb - X[...]) / (2 eps) (second probe)
__global__ void k_sb(double* __restrict__ sb, const double* __restrict__ X, int64_t nu, int64_t nc5, int second) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= nc5) return;
    const double eps = 1.0e-3;
    if (second) sb[i] = (sb[i] - X[nu + i]) / (2 * eps);
    else sb[i] = X[nu + i];
}

// Noise producer for ps_run: a host thread draws each step's n + (3Nm+6Nc) normals from the
// reference's stream, in the reference's order (slip first, then RFD), into one of two pinned slots
// while the GPU works on the previous step.  It draws exactly nsteps blocks, so the stream position
// after ps_run(k) equals k sequential steps.  A slot is refilled only after the copies that read it
// have completed (event).
struct NoisePipe {
    Ctx* c = nullptr;
    int64_t len = 0;
    int nsteps = 0;
    double* slot[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    std::mutex mu;
    std::condition_variable cv;
    int filled = 0, consumed = 0;
    bool abort = false;
    std::thread th;

    void produce() {
        for (int s = 0; s < nsteps; s++) {
            const int k = s & 1;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return abort || consumed >= s - 1; });
                if (abort) return;
            }
            if (s >= 2) (void)hipEventSynchronize(done[k]);
            c->normals.fill(slot[k], len);
            {
                std::lock_guard<std::mutex> lk(mu);
                filled = s + 1;
            }
            cv.notify_all();
        }
    }
    double* take(int s) {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return filled > s; });
        return slot[s & 1];
    }
    ps_err release(int s) {
        PS_HIP(c, hipEventRecord(done[s & 1], c->stream));
        {
            std::lock_guard<std::mutex> lk(mu);
            consumed = s + 1;
        }
        cv.notify_all();
        return PS_OK;
    }
    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu);
            abort = true;
        }
        cv.notify_all();
        if (th.joinable()) th.join();
    }
};

// The step's saddle solve of rhs into X; `uout` (optional) receives its velocity rows.  No solve
// information is asked for, so a block-eliminated solve makes no host round trip.
ps_err solve(Ctx* c, double* X, int warm, double* uout = nullptr) {
    ps_err r = saddle_solve_dev(c, c->rhs, X, warm, nullptr, uout);
    if (r == PS_ERR_NOCONV) r = PS_OK;   // the reference prints and continues (src/vel.cpp:45-51)
    return r;
}

__global__ void k_zero(double* __restrict__ a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// One iteration of the run() loop; `pipe` supplies this step's noise (step index s), or is null
// when kT <= 0.  With `noise_fixed` (graph capture) the noise is read from that pinned slot and the
// pipe is not touched.
ps_err one_step_joint(Ctx* c, NoisePipe* pipe, int s, double* noise_fixed);
ps_err one_step(Ctx* c, NoisePipe* pipe, int s, double* noise_fixed = nullptr) {
    if (c->nmslot > 0) return one_step_joint(c, pipe, s, noise_fixed);
    const int64_t n = c->n, nu = c->nu, nc5 = 5 * (int64_t)c->Nc;
    hipStream_t st = c->stream;
    const double dt = c->p.dt, kT = c->p.kT, eps = 1.0e-3;
    c->t += dt;
    // ---- predictor (x_n = x is saved by the same pass that bins the positions)
    PS_TRY(check_dist_dev(c, c->x_n));
    PS_TRY(rhs_dev(c, 0));
    c->ar_gram_next = c->arrow;
    PS_TRY(mob_prepare(c));
    if (c->arrow) PS_TRY(arrow_schur(c, 0));   // sync_mcc_schur_correction (src/run.cpp:92)
    PS_TRY(solve(c, c->Xdet, 0, c->v_det_n));
    if (kT > 0) {
        // solve_slip_vel: W = sqrt(M) xi, rhs[0:n] += sqrt(2kT/dt) W
        double* noise = noise_fixed ? noise_fixed : pipe->take(s);
        PS_HIP(c, hipMemcpyAsync(c->xi, noise, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, st));
        PS_TRY(sqrt_apply_dev(c, c->xi, c->W, nullptr));
        k_add_slip<<<eb(n), TB, 0, st>>>(c->rhs, c->W, std::sqrt(2 * kT / dt), n);
        PS_LAUNCH(c);
        // new_vel
        PS_TRY(solve(c, c->X, 0, c->up));
        if (c->p.record_forces) {
            k_extract<<<eb(nu), TB, 0, st>>>(c->fext, c->X, 0, nu);
            PS_LAUNCH(c);
        }
        k_sub<<<eb(nu), TB, 0, st>>>(c->v_brown, c->up, c->v_det_n, nu);
        PS_LAUNCH(c);
        // drift(): RFD probes at x +/- eps W
        PS_HIP(c, hipMemcpyAsync(c->fext, noise + n, sizeof(double) * (size_t)nu, hipMemcpyHostToDevice, st));
        if (!noise_fixed) PS_TRY(pipe->release(s));
        PS_TRY(probe_move(c, eps));
        PS_TRY(mob_prepare(c));
        PS_TRY(rhs_dev(c, 1));
        PS_TRY(solve(c, c->Xd, 0, c->drift_v));
        if (nc5 > 0) {
            k_sb<<<eb(nc5), TB, 0, st>>>(c->sb, c->Xd, nu, nc5, 0);
            PS_LAUNCH(c);
        }
        PS_TRY(probe_move(c, -2.0 * eps));
        PS_TRY(mob_prepare(c));
        PS_TRY(rhs_dev(c, 1));
        PS_TRY(solve(c, c->Xd, 1, c->udiff));
        k_sub<<<eb(nu), TB, 0, st>>>(c->udiff, c->udiff, c->drift_v, nu);
        PS_LAUNCH(c);
        if (nc5 > 0) {
            k_sb<<<eb(nc5), TB, 0, st>>>(c->sb, c->Xd, nu, nc5, 1);
            PS_LAUNCH(c);
        }
    } else {
        PS_TRY(solve(c, c->X, 0, c->up));
        if (c->p.record_forces) {
            k_extract<<<eb(nu), TB, 0, st>>>(c->fext, c->X, 0, nu);
            PS_LAUNCH(c);
        }
        k_zero<<<eb(nu), TB, 0, st>>>(c->v_brown, nu);
        PS_LAUNCH(c);
    }
    k_step<<<eb(c->Np), TB, 0, st>>>(c->sc, c->x, c->fext, c->udiff, c->up, dt, kT);
    PS_LAUNCH(c);
    // ---- corrector
    PS_TRY(check_dist_dev(c));
    PS_TRY(rhs_dev(c, 0));
    c->ar_gram_next = c->arrow;
    PS_TRY(mob_prepare(c));
    // src/run.cpp:129 (the drift's mob() calls take none); lazy: only the corrector's solve reads it
    if (c->arrow) PS_TRY(arrow_schur(c, 1));
    PS_TRY(solve(c, c->Xdet, 0, c->v_det_c));
    k_trapz<<<eb(nu), TB, 0, st>>>(c->up, c->v_det_n, c->v_det_c, c->v_brown, nu);
    PS_LAUNCH(c);
    k_step_from<<<eb(c->Np), TB, 0, st>>>(c->sc, c->x, c->x_n, c->udiff, c->up, dt, kT);
    PS_LAUNCH(c);
    if (c->Nc > 0) {
        k_quat<<<eb(c->Nc), TB, 0, st>>>(c->Nc, 3 * (int64_t)c->Np, c->q, c->up, dt);
        PS_LAUNCH(c);
    }
    c->have_mob = true;
    return PS_OK;
}

// rhs_s = rhs with a * w added to its first n entries (the slip solve's right-hand side)
__global__ void k_slip_rhs(double* __restrict__ rhs_s, const double* __restrict__ rhs, const double* __restrict__ w,
                           double a, int64_t n, int64_t ns) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i < ns) rhs_s[i] = i < n ? rhs[i] + a * w[i] : rhs[i];
}

// Whether the step runs as one_step_joint: full monomer HI, block-eliminated solves with no
// refinement sweep, thermal noise, colloids whose four stresslet systems fit one launch, and no
// lubrication.  (record_forces needs nothing here: with thermal noise the forces it copies into fext
// are overwritten by the RFD normals in the same step, src/vel.cpp:83-85, src/drift.cpp:9-15.)
bool joint_step_ok(const Ctx* c) {
    if (c->arrow || c->p.lub || c->p.solver != PS_SOLVER_BLOCK || !(c->p.kT > 0)) return false;
    if (c->p.precision == PS_PREC_MIXED && c->p.ir_maxit > 0) return false;
    if (c->Nm <= 0 || c->Nc <= 0) return false;
    // PS_JOINT_STEP=0 keeps the sequential order (one_step), e.g. to compare the two
    const char* env = std::getenv("PS_JOINT_STEP");
    if (env && std::strcmp(env, "0") == 0) return false;
    return mss_joint_fits(c, 4);
}

// One iteration of the run() loop with the same operations as one_step, reordered so that the
// stresslet solves of the predictor, the slip solve and both RFD probes share one launch:
//   - stage 1 of the predictor's solve (x = [-b_lambda; 0], the stresslet right-hand side) at x;
//   - the Lanczos recurrence at x;
//   - on the side stream, while the main one solves the tridiagonal square roots: both probes'
//     displacement, geometry, mob() (slots 1 and 2) and stage 1 of their solves;
//   - stage 1 of the slip solve at x;
//   - the four stresslet MINRES solves in one launch: its cost is the slowest one's iterations, not
//     their sum;
//   - stage 3 (the full apply and the velocities) of the four solves;
//   - the integrator and the corrector, as in one_step.
// The mob() calls run in the reference's order (x, x + eps W, x - eps W, corrector), so the carried
// tile passes from each to the next exactly as there, and each solve is the same block elimination
// as one_step's; only their order differs.
ps_err one_step_joint(Ctx* c, NoisePipe* pipe, int s, double* noise_fixed) {
    const int64_t n = c->n, nu = c->nu, ns = c->ns, nc5 = 5 * (int64_t)c->Nc;
    hipStream_t A = c->stream, B = c->side;
    const double dt = c->p.dt, kT = c->p.kT, eps = 1.0e-3;
    const int prec = c->p.precision;
    // every exit leaves slot 0 current and the main stream in c->stream
    struct Restore {
        Ctx* c;
        hipStream_t a;
        ~Restore() {
            c->stream = a;
            mob_use_slot(c, 0);
        }
    } restore{c, A};
    c->t += dt;
    // ---- predictor point x: forces (x_n = x saved by the binning pass), mob() into slot 0, stage 1
    // of the deterministic solve
    PS_TRY(check_dist_dev(c, c->x_n));
    PS_TRY(rhs_dev(c, 0));
    mob_use_slot(c, 0);
    PS_TRY(mob_prepare(c));
    double* bs0 = c->bsr4;
    double* bs1 = c->bsr4 + nc5;
    double* bs2 = c->bsr4 + 2 * nc5;
    double* bs3 = c->bsr4 + 3 * nc5;
    PS_TRY(block_stage1(c, c->rhs, c->Xdet, bs0));
    // ---- noise: the slip normals, then the RFD normals (the reference's draw order)
    double* noise = noise_fixed ? noise_fixed : pipe->take(s);
    PS_HIP(c, hipMemcpyAsync(c->xi, noise, sizeof(double) * (size_t)(n + nu), hipMemcpyHostToDevice, A));
    if (!noise_fixed) PS_TRY(pipe->release(s));
    PS_TRY(sqrt_recurrence_dev(c, c->xi));
    // ---- fork: the probes on the side stream
    PS_HIP(c, hipEventRecord(c->jev[0], A));
    PS_HIP(c, hipStreamWaitEvent(B, c->jev[0], 0));
    c->stream = B;
    PS_TRY(probe_move(c, eps));
    mob_use_slot(c, 1);
    PS_TRY(mob_prepare(c));
    PS_TRY(rhs_drift_dev(c, c->rhs_d));
    PS_TRY(block_stage1(c, c->rhs_d, c->Xd, bs2));
    PS_TRY(probe_move(c, -2.0 * eps));
    mob_use_slot(c, 2);
    PS_TRY(mob_prepare(c));
    PS_TRY(block_stage1(c, c->rhs_d, c->Xd2, bs3));
    PS_HIP(c, hipEventRecord(c->jev[1], B));
    // ---- main stream: square-root weights, W, the slip right-hand side; join; slip stage 1
    c->stream = A;
    mob_use_slot(c, 0);
    PS_TRY(sqrt_finish_dev(c, c->W, nullptr));
    k_slip_rhs<<<eb(ns), TB, 0, A>>>(c->rhs_s, c->rhs, c->W, std::sqrt(2 * kT / dt), n, ns);
    PS_LAUNCH(c);
    PS_HIP(c, hipStreamWaitEvent(A, c->jev[1], 0));
    PS_TRY(block_stage1(c, c->rhs_s, c->X, bs1));
    // ---- the four stresslet solves in one launch
    {
        const double* Ms[4] = {c->mslot[0].Mss, c->mslot[0].Mss, c->mslot[1].Mss, c->mslot[2].Mss};
        const double* rs[4] = {bs0, bs1, bs2, bs3};
        double* xs[4] = {c->Xdet + nu, c->X + nu, c->Xd + nu, c->Xd2 + nu};
        PS_TRY(mss_joint(c, 4, Ms, rs, xs));
    }
    // ---- stage 3: velocities of the predictor, the slip solve and both probes
    PS_TRY(block_stage3(c, c->rhs, c->Xdet, prec, c->v_det_n));
    PS_TRY(block_stage3(c, c->rhs_s, c->X, prec, c->up));
    k_sub<<<eb(nu), TB, 0, A>>>(c->v_brown, c->up, c->v_det_n, nu);
    PS_LAUNCH(c);
    mob_use_slot(c, 1);
    PS_TRY(block_stage3(c, c->rhs_d, c->Xd, prec, c->drift_v));
    k_sb<<<eb(nc5), TB, 0, A>>>(c->sb, c->Xd, nu, nc5, 0);
    PS_LAUNCH(c);
    mob_use_slot(c, 2);
    PS_TRY(block_stage3(c, c->rhs_d, c->Xd2, prec, c->udiff));
    k_sub<<<eb(nu), TB, 0, A>>>(c->udiff, c->udiff, c->drift_v, nu);
    PS_LAUNCH(c);
    k_sb<<<eb(nc5), TB, 0, A>>>(c->sb, c->Xd2, nu, nc5, 1);
    PS_LAUNCH(c);
    mob_use_slot(c, 0);
    k_step<<<eb(c->Np), TB, 0, A>>>(c->sc, c->x, c->fext, c->udiff, c->up, dt, kT);
    PS_LAUNCH(c);
    // ---- corrector (slot 0)
    PS_TRY(check_dist_dev(c));
    PS_TRY(rhs_dev(c, 0));
    PS_TRY(mob_prepare(c));
    PS_TRY(solve(c, c->Xdet, 0, c->v_det_c));
    k_trapz<<<eb(nu), TB, 0, A>>>(c->up, c->v_det_n, c->v_det_c, c->v_brown, nu);
    PS_LAUNCH(c);
    k_step_from<<<eb(c->Np), TB, 0, A>>>(c->sc, c->x, c->x_n, c->udiff, c->up, dt, kT);
    PS_LAUNCH(c);
    k_quat<<<eb(c->Nc), TB, 0, A>>>(c->Nc, 3 * (int64_t)c->Np, c->q, c->up, dt);
    PS_LAUNCH(c);
    c->have_mob = true;
    return PS_OK;
}

// Allocate what the joint step needs (once, at create, when joint_step_ok holds).
ps_err joint_alloc(Ctx* c) {
    const size_t ns = (size_t)c->ns, nc5 = 5 * (size_t)c->Nc;
    PS_TRY(mob_slots_alloc(c));
    PS_HIP(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    for (hipEvent_t& e : c->jev) PS_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    PS_HIP(c, hipMalloc(&c->rhs_s, sizeof(double) * ns));
    PS_HIP(c, hipMalloc(&c->rhs_d, sizeof(double) * ns));
    PS_HIP(c, hipMalloc(&c->Xd2, sizeof(double) * ns));
    PS_HIP(c, hipMalloc(&c->bsr4, sizeof(double) * 4 * nc5));
    PS_HIP(c, hipMemsetAsync(c->Xd2, 0, sizeof(double) * ns, c->stream));
    return PS_OK;
}

void joint_free(Ctx* c) {
    if (c->side) (void)hipStreamSynchronize(c->side);
    mob_slots_free(c);
    for (void* p : {(void*)c->rhs_s, (void*)c->rhs_d, (void*)c->Xd2, (void*)c->bsr4})
        if (p) (void)hipFree(p);
    c->rhs_s = c->rhs_d = c->Xd2 = c->bsr4 = nullptr;
    for (hipEvent_t& e : c->jev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    if (c->side) (void)hipStreamDestroy(c->side);
    c->side = nullptr;
}

Ctx* as_ctx(void* h) { return static_cast<Ctx*>(h); }

constexpr int EV_POOL = 4096;   // event pairs held before a harvest

ps_err fail(Ctx* c, ps_err code, const std::string& msg) {
    if (c) c->set_error(msg);
    return code;
}

// A step with no host round trip can be replayed, since every launch's arguments are then fixed for
// the run: block-eliminated solves (the stresslet MINRES runs to convergence inside one launch) and a
// Lanczos square root whose stopping test runs on the device.  In the arrowhead regime: one colloid,
// no WCA grid from the host.
bool step_graphable(const Ctx* c) {
    if (!c->p.graph || c->p.solver != PS_SOLVER_BLOCK || !(c->p.kT > 0) || c->prof_on) return false;
    // the lubricated block solve polls its colloid MINRES on the host
    if (c->p.lub) return false;
    if (c->arrow) return c->Nc == 1 && c->Nm > 0 && !c->p.mono_ev && !c->p.record_forces;
    // FP64 refinement of the mixed form reads its residual norms on the host
    return !(c->p.precision == PS_PREC_MIXED && c->p.ir_maxit > 0);
}

// Sticky device flags of the last run: a Verlet list over its fixed capacity, a grid-wide wait that
// ran out of time.  Reads the device counters (a host round trip).
ps_err check_device_flags(Ctx* c) {
    PS_HIP(c, hipMemcpyAsync(c->dcnt_h, c->dcnt, sizeof(int64_t) * DC_N, hipMemcpyDeviceToHost, c->stream));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    if (c->dcnt_h[DC_VLIST_OVERFLOW]) {
        c->set_error("a Verlet list exceeded its fixed capacity (511 partners per particle, 64 per particle "
                     "on average); the step results are not valid");
        return PS_ERR_STATE;
    }
    if (c->dcnt_h[DC_BARRIER_TIMEOUT]) {
        c->set_error("the stresslet MINRES grid barrier ran out of time; the step results are not valid");
        return PS_ERR_HIP;
    }
    return PS_OK;
}

void step_graphs_reset(Ctx* c) {
    for (auto& g : c->step_graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// The step as a HIP graph reading its noise from pinned slot `slot` (captured, not run): the host's
// t and counters are restored, and the counter increments of one step kept for the replays.  The
// launch sequence of a step does not depend on its index, only on the slot's address.  Captured
// before the noise producer starts, so no other thread makes HIP calls during a capture.
ps_err capture_step(Ctx* c, int k, double* slot) {
    const double t0 = c->t;
    const ps_counters cnt0 = c->cnt;
    PS_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const ps_err r = one_step(c, nullptr, k, slot);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->t = t0;
    if (r != PS_OK) {
        if (g) (void)hipGraphDestroy(g);
        c->cnt = cnt0;
        return r;
    }
    if (e != hipSuccess) return fail(c, PS_ERR_HIP, std::string("step capture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&c->step_graph[k], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        c->step_graph[k] = nullptr;
        c->cnt = cnt0;
        return fail(c, PS_ERR_HIP, std::string("step graph: ") + hipGetErrorString(ei));
    }
    const int64_t* a = &c->cnt.mob_calls;
    const int64_t* b = &cnt0.mob_calls;
    int64_t* d = &c->step_graph_delta.mob_calls;
    for (size_t q = 0; q < sizeof(ps_counters) / sizeof(int64_t); q++) d[q] = a[q] - b[q];
    c->cnt = cnt0;
    return PS_OK;
}

// One step by replaying the graph of its noise slot; false in *eager when the step had to run
// eagerly: a step that weakens the trap (ktrap is a kernel argument), after which the run continues
// without graphs.
ps_err graph_step(Ctx* c, NoisePipe* pipe, int s, bool* eager) {
    const int k = s & 1;
    const bool trap_on = !(std::fabs(c->ktrap_cur) < 1e-12);
    if (trap_on && c->p.weaken_trap >= 0 && c->ktrap_cur != c->p.weaken_trap &&
        c->t + c->p.dt > c->p.tstart + c->p.trun) {
        *eager = true;
        return one_step(c, pipe, s);
    }
    (void)pipe->take(s);   // the slot is filled (the graph reads it)
    PS_HIP(c, hipGraphLaunch(c->step_graph[k], c->stream));
    PS_TRY(pipe->release(s));
    c->t += c->p.dt;
    int64_t* a = &c->cnt.mob_calls;
    const int64_t* d = &c->step_graph_delta.mob_calls;
    for (size_t q = 0; q < sizeof(ps_counters) / sizeof(int64_t); q++)
        if (&a[q] != &c->cnt.nvlist) a[q] += d[q];
    return PS_OK;
}


template <class F>
ps_err with_io(Ctx* c, int64_t na, const double* in, int64_t nb, double* out, F&& fn) {
    // host in -> device a; fn(a, b); device b -> host out
    PS_HIP(c, hipMemcpyAsync(c->io_a, in, sizeof(double) * (size_t)na, hipMemcpyHostToDevice, c->stream));
    PS_TRY(fn(c->io_a, c->io_b));
    PS_HIP(c, hipMemcpyAsync(out, c->io_b, sizeof(double) * (size_t)nb, hipMemcpyDeviceToHost, c->stream));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    return PS_OK;
}

void destroy_ctx(Ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    c->ev_pool.clear();
    joint_free(c);
    apply_free(c);
    krylov_free(c);
    pairs_free(c);
    arrow_free(c);
    dense_free(c);
    lub_free(c);
    if (c->dense_io) (void)hipFree(c->dense_io);
    for (void* p : {(void*)c->x, (void*)c->xpd, (void*)c->xm, (void*)c->xpos0, (void*)c->q, (void*)c->x_n,
                    (void*)c->chain, (void*)c->up, (void*)c->udiff, (void*)c->v_det_n, (void*)c->v_det_c,
                    (void*)c->v_brown, (void*)c->drift_v, (void*)c->sb, (void*)c->rhs,
                    (void*)c->X, (void*)c->Xdet, (void*)c->Xd, (void*)c->W, (void*)c->xi, (void*)c->stale,
                    (void*)c->io_a, (void*)c->io_b})
        if (p) (void)hipFree(p);
    if (c->noise_h) (void)hipHostFree(c->noise_h);
    for (hipEvent_t e : c->noise_ev)
        if (e) (void)hipEventDestroy(e);
    step_graphs_reset(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

thread_local std::string g_create_error;

}  // namespace

// Errors of the entry points that have no context (ps_last_error(NULL) reads them).
void ps_set_context_free_error(const std::string& msg) { g_create_error = msg; }

ps_err Ctx::prof_harvest() {
    Ctx* c = this;
    if (ev_used == 0) return PS_OK;
    PS_HIP(c, hipStreamSynchronize(stream));
    for (int k = 0; k < ev_used; k++) {
        float ms = 0.f;
        PS_HIP(c, hipEventElapsedTime(&ms, ev_pool[2 * k], ev_pool[2 * k + 1]));
        if (ev_kind[k] == 0) {
            prof.aa_launches++;
            prof.aa_ms += ms;
        } else if (ev_kind[k] == 1) {
            prof.apply_launches++;
            prof.apply_ms += ms;
        } else if (ev_kind[k] == 2) {
            prof.dense_launches++;
            prof.dense_ms += ms;
        } else if (ev_kind[k] == 3) {
            prof.arrow_launches++;
            prof.arrow_ms += ms;
        } else {
            prof.arsolve_launches++;
            prof.arsolve_ms += ms;
        }
    }
    ev_used = 0;
    return PS_OK;
}

ps_err Ctx::prof_record(int kind, hipEvent_t* a, hipEvent_t* b) {
    if (ev_used >= EV_POOL) PS_TRY(prof_harvest());
    ev_kind[ev_used] = kind;
    *a = ev_pool[2 * ev_used];
    *b = ev_pool[2 * ev_used + 1];
    ev_used++;
    return PS_OK;
}

extern "C" {

ps_err ps_profile_begin(void* h) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    PS_TRY(c->prof_harvest());
    if (c->ev_pool.empty()) {   // created here, outside any timed region
        c->ev_pool.resize(2 * EV_POOL);
        c->ev_kind.resize(EV_POOL);
        for (auto& e : c->ev_pool) PS_HIP(c, hipEventCreate(&e));
    }
    c->prof = ps_profile{};
    c->prof.aa_pairs = (int64_t)c->Nm * c->Nm;
    c->prof_dispatch_events = std::getenv("ROCPROF_OUTPUT_PATH") == nullptr;
    c->prof_on = true;
    return PS_OK;
}

ps_err ps_profile_end(void* h, ps_profile* out) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    PS_TRY(c->prof_harvest());
    c->prof_on = false;
    if (out) *out = c->prof;
    return PS_OK;
}

void ps_params_default(ps_params* p) {
    std::memset(p, 0, sizeof(*p));
    // PolyStokes ctor defaults as bound (src/bindings.cpp:56-69)
    p->samplerate = 1;
    p->mm_HI = 1;
    p->chain_HI = 0;
    p->fene = 1;
    p->record_forces = 1;
    p->tether = 0;
    p->mono_ev = 0;
    p->box_active = 0;
    p->t = 0.0;
    p->seed = 12345;
    p->restart = 0.0;
    p->weaken_trap = -1;
    // KSP: rtol 1e-6 (src/init.cpp:64), PETSc default abstol / maxit
    p->ksp_rtol = 1e-6;
    p->ksp_abstol = 1e-50;
    p->ksp_maxit = 10000;
    p->sqrt_rtol = 1e-8;
    p->sqrt_kmax = 64;
    p->force_sel = PS_FORCE_HS;
    p->hs_kpref = 1.0;
    p->hs_kmode = 0;
    p->hs_exp_k = 100.0;
    p->precision = PS_PREC_FP64;
    p->ir_rtol = 1e-6;
    p->ir_maxit = 20;
    p->solver = PS_SOLVER_MINRES;
    p->aa_kernel = PS_AA_AUTO;
    p->graph = 1;
    p->lub = 0;
    p->lub_rcut = 4.0;
}

const char* ps_last_error(const void* ctx) {
    if (!ctx) return g_create_error.c_str();
    return static_cast<const Ctx*>(ctx)->err.c_str();
}

ps_err ps_create(const ps_params* p, void** out) {
    if (!p || !out) return PS_ERR_ARG;
    *out = nullptr;
    g_create_error.clear();
    if (p->Np != p->Nm + p->Nc || p->Np < 2 || p->Nm < 0 || p->Nc < 0) {
        g_create_error = "Np must equal Nm + Nc and be at least 2";
        return PS_ERR_ARG;
    }
    if (!p->mm_HI && p->Nc > 1) {
        // the arrowhead operator has no colloid-colloid block (src/matfree_A.cpp:13-16, :124-127)
        g_create_error = "mm_HI=False supports at most one colloid (the reference's arrowhead operator)";
        return PS_ERR_ARG;
    }
    if (p->dt <= 0 || p->beta <= 0) {
        g_create_error = "dt and beta must be positive";
        return PS_ERR_ARG;
    }
    if (p->box_active && (p->box[0] <= 0 || p->box[1] <= 0 || p->box[2] <= 0)) {
        g_create_error = "Box lengths must be positive";   // Box::configure (src/box.cpp:10-14)
        return PS_ERR_ARG;
    }
    if (p->precision < PS_PREC_FP64 || p->precision > PS_PREC_MIXED || p->solver < PS_SOLVER_MINRES ||
        p->solver > PS_SOLVER_BLOCK || p->aa_kernel < PS_AA_AUTO || p->aa_kernel > PS_AA_SYMMETRIC) {
        g_create_error = "precision, solver or aa_kernel out of range";
        return PS_ERR_ARG;
    }
    if (p->lub != 0 && (p->lub != 1 || !p->mm_HI || !(p->lub_rcut > 0.0))) {
        g_create_error = "lub must be 0 or 1; lubrication needs mm_HI and lub_rcut > 0";
        return PS_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_error = "no HIP device visible";
        return PS_ERR_NODEV;
    }
    Ctx* c = new Ctx();
    c->p = *p;
    c->Np = p->Np;
    c->Nm = p->Nm;
    c->Nc = p->Nc;
    c->arrow = !p->mm_HI;
    c->n = 3 * (int64_t)c->Nm + 11 * (int64_t)c->Nc;
    c->nu = 3 * (int64_t)c->Nm + 6 * (int64_t)c->Nc;
    c->ns = c->n + c->nu;
    SysConst& s = c->sc;
    s.Np = c->Np;
    s.Nm = c->Nm;
    s.Nc = c->Nc;
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
    s.n = c->n;
    s.nu = c->nu;
    s.ns = c->ns;
    s.nm3 = 3 * (int64_t)c->Nm;
    s.nc3 = 3 * (int64_t)c->Nc;
    auto bail = [&](ps_err code) {
        g_create_error = c->err;
        destroy_ctx(c);
        return code;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(PS_ERR_HIP);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) c->n_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        c->err = "hipStreamCreate failed";
        return bail(PS_ERR_HIP);
    }
    const size_t n3 = 3 * (size_t)c->Np;
    auto dalloc = [&](double** ptr, size_t cnt) -> bool {
        return hipMalloc(ptr, sizeof(double) * std::max<size_t>(cnt, 1)) == hipSuccess;
    };
    bool ok = dalloc(&c->x, n3) && dalloc(&c->xpd, n3) && dalloc(&c->xm, n3) && dalloc(&c->xpos0, 3 * (size_t)c->Nc) &&
              dalloc(&c->q, 4 * (size_t)c->Nc) && dalloc(&c->x_n, n3) && dalloc(&c->up, c->nu) &&
              dalloc(&c->udiff, c->nu) && dalloc(&c->v_det_n, c->nu) && dalloc(&c->v_det_c, c->nu) &&
              dalloc(&c->v_brown, c->nu) && dalloc(&c->drift_v, c->nu) &&
              dalloc(&c->sb, 5 * (size_t)c->Nc) && dalloc(&c->rhs, c->ns) && dalloc(&c->X, c->ns) &&
              dalloc(&c->Xdet, c->ns) && dalloc(&c->Xd, c->ns) && dalloc(&c->W, c->n) && dalloc(&c->xi, c->n + c->nu) &&
              dalloc(&c->stale, 18) && dalloc(&c->io_a, c->ns) && dalloc(&c->io_b, c->ns);
    if (!ok) {
        c->err = "device allocation failed";
        return bail(PS_ERR_ALLOC);
    }
    // one step's noise as drawn: the slip normals (n), then the RFD normals (nu), so that one copy
    // moves both
    c->fext = c->xi + c->n;
    if (hipEventCreateWithFlags(&c->noise_ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->noise_ev[1], hipEventDisableTiming) != hipSuccess) {
        c->err = "event creation failed";
        return bail(PS_ERR_HIP);
    }
    if (hipHostMalloc(&c->noise_h, sizeof(double) * 2 * (size_t)(c->n + c->nu), hipHostMallocDefault) != hipSuccess) {
        c->err = "pinned allocation failed";
        return bail(PS_ERR_ALLOC);
    }
    if (hipMalloc(&c->chain, sizeof(int) * std::max(c->Nm, 1)) != hipSuccess) return bail(PS_ERR_ALLOC);
    // chain ids (src/init.cpp:205-212); same-chain pairs skipped under chain_HI
    std::vector<int> chain(std::max(c->Nm, 1), 0);
    for (int ii = 0; ii < p->Npoly; ii++)
        for (int jj = 0; jj < p->Nmono_per_chain; jj++) {
            const int kk = ii * p->Nmono_per_chain + jj;
            if (kk < c->Nm) chain[kk] = ii;
        }
    if (hipMemcpy(c->chain, chain.data(), sizeof(int) * chain.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(PS_ERR_HIP);
    // last processed AA pair of mob()'s monomer loop (src/mob.cpp:845-897)
    c->stale_i = c->stale_j = -1;
    for (int i = c->Nm - 2; i >= 0 && c->stale_i < 0; i--)
        for (int j = c->Nm - 1; j > i; j--) {
            if (p->chain_HI && chain[i] == chain[j]) continue;
            c->stale_i = i;
            c->stale_j = j;
            break;
        }
    // mob_a after fill_self is the identity (src/mob.cpp:65-86, :568)
    double st0[18] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (hipMemcpy(c->stale, st0, sizeof(st0), hipMemcpyHostToDevice) != hipSuccess) return bail(PS_ERR_HIP);
    c->ktrap_cur = p->ktrap;
    c->t = p->restart;   // src/Stokes.cpp:21
    c->normals.seed(p->seed + (uint64_t)std::llround(p->restart / p->dt));
    if (apply_alloc(c) != PS_OK) return bail(PS_ERR_ALLOC);
    if (krylov_alloc(c) != PS_OK) return bail(PS_ERR_ALLOC);
    if (pairs_alloc(c) != PS_OK) return bail(PS_ERR_ALLOC);
    if (c->arrow && arrow_alloc(c) != PS_OK) return bail(PS_ERR_ALLOC);
    if (c->p.lub && lub_alloc(c) != PS_OK) return bail(PS_ERR_ALLOC);
    if (joint_step_ok(c) && joint_alloc(c) != PS_OK) return bail(PS_ERR_ALLOC);
    for (double* v : {c->up, c->udiff, c->v_det_n, c->v_det_c, c->v_brown, c->fext, c->drift_v})
        (void)hipMemsetAsync(v, 0, sizeof(double) * (size_t)c->nu, c->stream);
    for (double* v : {c->X, c->Xdet, c->Xd})
        (void)hipMemsetAsync(v, 0, sizeof(double) * (size_t)c->ns, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(PS_ERR_HIP);
    *out = c;
    return PS_OK;
}

ps_err ps_destroy(void* h) {
    destroy_ctx(as_ctx(h));
    return PS_OK;
}

ps_err ps_device_info(void* h, char* name, int name_len, int* n_cu) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    hipDeviceProp_t prop;
    PS_HIP(c, hipGetDeviceProperties(&prop, c->device));
    if (name && name_len > 0) std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    if (n_cu) *n_cu = prop.multiProcessorCount;
    return PS_OK;
}

ps_err ps_initial_configuration(void* h, const double* x0, const double* q0) {
    Ctx* c = as_ctx(h);
    if (!c || !x0) return PS_ERR_ARG;
    const size_t n3 = 3 * (size_t)c->Np;
    PS_HIP(c, hipMemcpy(c->x, x0, sizeof(double) * n3, hipMemcpyHostToDevice));
    // trap centres = initial colloid positions on a fresh start (src/Stokes.cpp:279-284)
    if (c->p.restart == 0.0 && c->Nc > 0)
        PS_HIP(c, hipMemcpy(c->xpos0, x0 + 3 * (size_t)c->Nm, sizeof(double) * 3 * c->Nc, hipMemcpyHostToDevice));
    else if (c->Nc > 0)
        PS_HIP(c, hipMemset(c->xpos0, 0, sizeof(double) * 3 * c->Nc));
    std::vector<double> q(4 * (size_t)std::max(c->Nc, 1), 0.0);
    for (int i = 0; i < c->Nc; i++)
        for (int k = 0; k < 4; k++) q[4 * i + k] = q0 ? q0[4 * i + k] : (k == 0 ? 1.0 : 0.0);
    if (c->Nc > 0) PS_HIP(c, hipMemcpy(c->q, q.data(), sizeof(double) * 4 * c->Nc, hipMemcpyHostToDevice));
    c->have_x = true;
    return PS_OK;
}

ps_err ps_set_positions(void* h, const double* x) {
    Ctx* c = as_ctx(h);
    if (!c || !x) return PS_ERR_ARG;
    PS_HIP(c, hipMemcpy(c->x, x, sizeof(double) * 3 * (size_t)c->Np, hipMemcpyHostToDevice));
    c->have_x = true;
    return PS_OK;
}

ps_err ps_get_state(void* h, double* x, double* q, double* t) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    PS_HIP(c, hipStreamSynchronize(c->stream));
    if (x) PS_HIP(c, hipMemcpy(x, c->x, sizeof(double) * 3 * (size_t)c->Np, hipMemcpyDeviceToHost));
    if (q && c->Nc > 0) PS_HIP(c, hipMemcpy(q, c->q, sizeof(double) * 4 * c->Nc, hipMemcpyDeviceToHost));
    if (t) *t = c->t;
    return PS_OK;
}

ps_err ps_get_vectors(void* h, double* up, double* vdn, double* vdc, double* vb, double* ud, double* sb, double* fext) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    PS_HIP(c, hipStreamSynchronize(c->stream));
    const size_t b = sizeof(double) * (size_t)c->nu;
    if (up) PS_HIP(c, hipMemcpy(up, c->up, b, hipMemcpyDeviceToHost));
    if (vdn) PS_HIP(c, hipMemcpy(vdn, c->v_det_n, b, hipMemcpyDeviceToHost));
    if (vdc) PS_HIP(c, hipMemcpy(vdc, c->v_det_c, b, hipMemcpyDeviceToHost));
    if (vb) PS_HIP(c, hipMemcpy(vb, c->v_brown, b, hipMemcpyDeviceToHost));
    if (ud) PS_HIP(c, hipMemcpy(ud, c->udiff, b, hipMemcpyDeviceToHost));
    if (sb && c->Nc > 0) PS_HIP(c, hipMemcpy(sb, c->sb, sizeof(double) * 5 * c->Nc, hipMemcpyDeviceToHost));
    if (fext) PS_HIP(c, hipMemcpy(fext, c->fext, b, hipMemcpyDeviceToHost));
    return PS_OK;
}

ps_err ps_kernel_choice(void* h, int* kernel64, int* kernel32, float* ms64, float* ms32) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    const bool has32 = c->p.precision == PS_PREC_MIXED;
    if (kernel64) *kernel64 = c->aw64.sym ? PS_AA_SYMMETRIC : PS_AA_ORDERED;
    if (kernel32) *kernel32 = has32 ? (c->aw32.sym ? PS_AA_SYMMETRIC : PS_AA_ORDERED) : 0;
    if (ms64) *ms64 = c->aa_tuned_ms[0];
    if (ms32) *ms32 = has32 ? c->aa_tuned_ms[1] : 0.f;
    return PS_OK;
}

ps_err ps_kernel_shape(void* h, int* split64, int* split32) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    const bool has32 = c->p.precision == PS_PREC_MIXED;
    if (split64) *split64 = c->aw64.sym ? 0 : c->aw64.per_split_aa;
    if (split32) *split32 = has32 && !c->aw32.sym ? c->aw32.per_split_aa : 0;
    return PS_OK;
}

ps_err ps_get_counters(void* h, ps_counters* out) {
    Ctx* c = as_ctx(h);
    if (!c || !out) return PS_ERR_ARG;
    // the counts the kernels decide (Lanczos iterations, negative Ritz values, stresslet iterations)
    // accumulate on the device
    PS_HIP(c, hipMemcpyAsync(c->dcnt_h, c->dcnt, sizeof(int64_t) * DC_N, hipMemcpyDeviceToHost, c->stream));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    if (!c->arrow && c->have_pd) PS_TRY(vlist_count(c, nullptr));
    *out = c->cnt;
    out->lanczos_iterations += c->dcnt_h[DC_LANCZOS_ITERS];
    out->lanczos_nneg_events += c->dcnt_h[DC_NNEG_EVENTS];
    out->lanczos_nneg_count += c->dcnt_h[DC_NNEG_COUNT];
    out->stresslet_iterations += c->dcnt_h[DC_MSS_ITERS];
    out->stresslet_capped += c->dcnt_h[DC_MSS_CAPPED];
    return PS_OK;
}

ps_err ps_check_dist(void* h, int64_t* n_pairs) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    if (!c->have_x) return fail(c, PS_ERR_STATE, "check_dist before initial_configuration");
    PS_TRY(check_dist_dev(c));
    if (c->arrow) {
        PS_TRY(arrow_count_pairs(c));
        PS_HIP(c, hipStreamSynchronize(c->stream));
        if (n_pairs) *n_pairs = c->nvlist;
        return PS_OK;
    }
    return vlist_count(c, n_pairs);
}

ps_err ps_get_vlist(void* h, int32_t* rows, int32_t* partners, int64_t* pidx, int64_t cap, int64_t* n_pairs) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    if (!c->have_pd) return fail(c, PS_ERR_STATE, "get_vlist before check_dist");
    return vlist_copy(c, rows, partners, pidx, cap, n_pairs);
}

ps_err ps_mob(void* h) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    if (!c->have_pd) return fail(c, PS_ERR_STATE, "mob before check_dist (mob reads the pair geometry)");
    PS_TRY(mob_prepare(c));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    c->have_mob = true;
    return PS_OK;
}

ps_err ps_stale_tile(void* h, double* t9) {
    Ctx* c = as_ctx(h);
    if (!c || !t9) return PS_ERR_ARG;
    PS_HIP(c, hipStreamSynchronize(c->stream));
    PS_HIP(c, hipMemcpy(t9, c->stale, sizeof(double) * 9, hipMemcpyDeviceToHost));
    return PS_OK;
}

ps_err ps_mob_mult_dev(void* h, const double* f, double* u) {
    Ctx* c = as_ctx(h);
    if (!c || !f || !u) return PS_ERR_ARG;
    if (!c->have_mob) return fail(c, PS_ERR_STATE, "apply before mob");
    return mob_apply(c, f, u, c->p.precision);
}
ps_err ps_mob_mult(void* h, const double* f, double* u) {
    Ctx* c = as_ctx(h);
    if (!c || !f || !u) return PS_ERR_ARG;
    if (!c->have_mob) return fail(c, PS_ERR_STATE, "apply before mob");
    return with_io(c, c->n, f, c->n, u, [&](double* a, double* b) { return mob_apply(c, a, b, c->p.precision); });
}
ps_err ps_sync_mcc(void* h) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    if (!c->arrow) return fail(c, PS_ERR_STATE, "sync_mcc: the mm_HI=False regime only");
    if (!c->have_mob) return fail(c, PS_ERR_STATE, "sync_mcc before mob");
    return arrow_schur(c, 0);
}
ps_err ps_get_schur(void* h, double* mcc, double* q, double* lam) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    if (!c->arrow || c->Nc != 1) return fail(c, PS_ERR_STATE, "get_schur: mm_HI=False with one colloid only");
    double buf[301];
    PS_HIP(c, hipMemcpyAsync(buf, c->ar_mcc, sizeof(buf), hipMemcpyDeviceToHost, c->stream));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    if (buf[300] != 0.0) {   // a lazy sync left Q and the spectrum of an earlier one: redo it in full
        PS_TRY(arrow_schur_redo(c));
        PS_HIP(c, hipMemcpyAsync(buf, c->ar_mcc, sizeof(buf), hipMemcpyDeviceToHost, c->stream));
        PS_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (mcc) std::copy(buf, buf + 121, mcc);
    if (q) std::copy(buf + 121, buf + 242, q);
    if (lam) std::copy(buf + 242, buf + 253, lam);
    return PS_OK;
}
ps_err ps_get_cell_list(void* h, int32_t* cell_of, int32_t* dims, double* geom) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    if (!c->arrow || !c->wc_built) return fail(c, PS_ERR_STATE, "get_cell_list: no monomer cell list built yet");
    if (cell_of && c->Nm > 0) {
        PS_HIP(c, hipMemcpyAsync(cell_of, c->wc_cell, sizeof(int) * (size_t)c->Nm, hipMemcpyDeviceToHost, c->stream));
        PS_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (dims) std::copy(c->wc_dims, c->wc_dims + 3, dims);
    if (geom) std::copy(c->wc_geom, c->wc_geom + 7, geom);
    return PS_OK;
}
ps_err ps_dense_build(void* h, int storage) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    if (c->arrow) return fail(c, PS_ERR_STATE, "dense_build: the arrowhead mobility is O(N); nothing to store");
    return dense_build(c, storage);
}
ps_err ps_dense_mult_dev(void* h, const double* X, double* Y, int nrhs) {
    Ctx* c = as_ctx(h);
    if (!c || !X || !Y) return PS_ERR_ARG;
    return dense_mult_dev(c, X, Y, nrhs);
}
ps_err ps_dense_mult(void* h, const double* X, double* Y, int nrhs) {
    Ctx* c = as_ctx(h);
    if (!c || !X || !Y) return PS_ERR_ARG;
    if (nrhs < 1 || nrhs > 16) return fail(c, PS_ERR_ARG, "dense_mult: 1 <= nrhs <= 16");
    const size_t len = (size_t)nrhs * (size_t)c->n;
    if (!c->dense_io) PS_HIP(c, hipMalloc(&c->dense_io, sizeof(double) * 2 * 16 * (size_t)c->n));
    double* dx = c->dense_io;
    double* dy = c->dense_io + 16 * (size_t)c->n;
    PS_HIP(c, hipMemcpyAsync(dx, X, sizeof(double) * len, hipMemcpyHostToDevice, c->stream));
    PS_TRY(dense_mult_dev(c, dx, dy, nrhs));
    PS_HIP(c, hipMemcpyAsync(Y, dy, sizeof(double) * len, hipMemcpyDeviceToHost, c->stream));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    return PS_OK;
}
ps_err ps_lub_set_table(void* h, int which, const double* rows, int nrows, int ncols) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    return lub_set_table(c, which, rows, nrows, ncols);
}
ps_err ps_lub_build(void* h, double rcut, int64_t* npairs) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    return lub_build(c, rcut, npairs);
}
ps_err ps_lub_apply_dev(void* h, const double* uw, const double* einf, double* fts) {
    Ctx* c = as_ctx(h);
    if (!c || !uw || !fts) return PS_ERR_ARG;
    return lub_apply_dev(c, uw, einf, fts);
}
ps_err ps_lub_apply(void* h, const double* uw, const double* einf, double* fts) {
    Ctx* c = as_ctx(h);
    if (!c || !uw || !fts) return PS_ERR_ARG;
    const int64_t Nc = c->Nc;
    if (Nc == 0) return PS_OK;
    // io_a holds [uw (6Nc); E (5)], io_b the 11Nc result (both are ns >= 17 Nc long)
    if (einf)
        PS_HIP(c, hipMemcpyAsync(c->io_a + 6 * Nc, einf, sizeof(double) * 5, hipMemcpyHostToDevice, c->stream));
    return with_io(c, 6 * Nc, uw, 11 * Nc, fts, [&](const double* a, double* b) {
        return lub_apply_dev(c, a, einf ? a + 6 * Nc : nullptr, b);
    });
}
ps_err ps_lub_interp(const double* rows, int nrows, int ncols, double r, double* out) {
    return lub_interp_host(rows, nrows, ncols, r, out);
}
ps_err ps_lub_slopes(const double* rows, int nrows, int ncols, double* slopes) {
    return lub_slopes_host(rows, nrows, ncols, slopes);
}
ps_err ps_saddle_mult_dev(void* h, const double* x, double* y) {
    Ctx* c = as_ctx(h);
    if (!c || !x || !y) return PS_ERR_ARG;
    if (!c->have_mob) return fail(c, PS_ERR_STATE, "apply before mob");
    return saddle_apply(c, x, y, c->p.precision);
}
ps_err ps_saddle_mult(void* h, const double* x, double* y) {
    Ctx* c = as_ctx(h);
    if (!c || !x || !y) return PS_ERR_ARG;
    if (!c->have_mob) return fail(c, PS_ERR_STATE, "apply before mob");
    return with_io(c, c->ns, x, c->ns, y, [&](double* a, double* b) { return saddle_apply(c, a, b, c->p.precision); });
}
ps_err ps_saddle_diag(void* h, double* d) {
    Ctx* c = as_ctx(h);
    if (!c || !d) return PS_ERR_ARG;
    PS_TRY(saddle_diag_dev(c, c->io_b));
    PS_HIP(c, hipMemcpyAsync(d, c->io_b, sizeof(double) * (size_t)c->ns, hipMemcpyDeviceToHost, c->stream));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    return PS_OK;
}
ps_err ps_saddle_solve_dev(void* h, const double* b, double* x, int warm, ps_solve_info* info) {
    Ctx* c = as_ctx(h);
    if (!c || !b || !x) return PS_ERR_ARG;
    if (!c->have_mob) return fail(c, PS_ERR_STATE, "solve before mob");
    ps_err r = saddle_solve_dev(c, b, x, warm, info);
    if (r == PS_ERR_NOCONV) c->set_error("MINRES reached its iteration cap");
    return r;
}
ps_err ps_saddle_solve(void* h, const double* b, double* x, int warm, ps_solve_info* info) {
    Ctx* c = as_ctx(h);
    if (!c || !b || !x) return PS_ERR_ARG;
    if (!c->have_mob) return fail(c, PS_ERR_STATE, "solve before mob");
    PS_HIP(c, hipMemcpyAsync(c->io_a, b, sizeof(double) * (size_t)c->ns, hipMemcpyHostToDevice, c->stream));
    if (warm) PS_HIP(c, hipMemcpyAsync(c->io_b, x, sizeof(double) * (size_t)c->ns, hipMemcpyHostToDevice, c->stream));
    ps_err r = saddle_solve_dev(c, c->io_a, c->io_b, warm, info);
    if (r != PS_OK && r != PS_ERR_NOCONV) return r;
    PS_HIP(c, hipMemcpyAsync(x, c->io_b, sizeof(double) * (size_t)c->ns, hipMemcpyDeviceToHost, c->stream));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    if (r == PS_ERR_NOCONV) c->set_error("MINRES reached its iteration cap");
    return r;
}
ps_err ps_sqrt_apply_dev(void* h, const double* xi, double* w, ps_sqrt_info* info) {
    Ctx* c = as_ctx(h);
    if (!c || !xi || !w) return PS_ERR_ARG;
    if (!c->have_mob) return fail(c, PS_ERR_STATE, "sqrt before mob");
    return sqrt_apply_dev(c, xi, w, info);
}
ps_err ps_sqrt_apply(void* h, const double* xi, double* w, ps_sqrt_info* info) {
    Ctx* c = as_ctx(h);
    if (!c || !xi || !w) return PS_ERR_ARG;
    if (!c->have_mob) return fail(c, PS_ERR_STATE, "sqrt before mob");
    return with_io(c, c->n, xi, c->n, w, [&](double* a, double* b) { return sqrt_apply_dev(c, a, b, info); });
}
ps_err ps_rhs(void* h, int drift, double* rhs) {
    Ctx* c = as_ctx(h);
    if (!c) return PS_ERR_ARG;
    if (!c->have_pd) return fail(c, PS_ERR_STATE, "RHS before check_dist");
    PS_TRY(rhs_dev(c, drift));
    if (rhs) {
        PS_HIP(c, hipMemcpyAsync(rhs, c->rhs, sizeof(double) * (size_t)c->ns, hipMemcpyDeviceToHost, c->stream));
    }
    PS_HIP(c, hipStreamSynchronize(c->stream));
    return PS_OK;
}
ps_err ps_run(void* h, int nsteps) {
    Ctx* c = as_ctx(h);
    if (!c || nsteps < 0) return PS_ERR_ARG;
    if (!c->have_x) return fail(c, PS_ERR_STATE, "run before initial_configuration");
    if (nsteps == 0) return PS_OK;
    std::unique_ptr<NoisePipe> pipe;
    if (c->p.kT > 0) {
        pipe.reset(new NoisePipe());
        pipe->c = c;
        pipe->len = c->n + c->nu;
        pipe->nsteps = nsteps;
        pipe->slot[0] = c->noise_h;
        pipe->slot[1] = c->noise_h + pipe->len;
        pipe->done[0] = c->noise_ev[0];
        pipe->done[1] = c->noise_ev[1];
    }
    ps_err rc = PS_OK;
    bool graph = pipe && step_graphable(c) && nsteps >= 3;   // two captures pay off from a few steps
    if (graph) {   // both slots' graphs, before the producer thread makes any HIP call
        rc = capture_step(c, 0, pipe->slot[0]);
        if (rc == PS_OK) rc = capture_step(c, 1, pipe->slot[1]);
        if (rc != PS_OK) {
            step_graphs_reset(c);
            return rc;
        }
    }
    if (pipe) pipe->th = std::thread([p = pipe.get()] { p->produce(); });
    for (int i = 0; i < nsteps && rc == PS_OK; i++) {
        bool eager = false;
        rc = graph ? graph_step(c, pipe.get(), i, &eager) : one_step(c, pipe.get(), i);
        if (eager) graph = false;
    }
    if (pipe) pipe->stop();
    step_graphs_reset(c);   // the graphs read this run's noise slots
    if (rc != PS_OK) return rc;
    return check_device_flags(c);
}
// File name of a frame: output_dir + "/" + prefix + t + "_.txt" with t printed by a default
// std::ostream (6 significant digits, %g style), as src/write_configuration.cpp:18-19 does.
static std::string frame_name(const char* dir, const char* prefix, double t) {
    std::ostringstream oss;
    oss << dir << "/" << prefix << t << "_.txt";
    return oss.str();
}

// rows x cols values, "%.10f" separated by tabs, one row per line (src/write_configuration.cpp).
static bool write_table(const std::string& name, const double* v, int64_t rows, int cols) {
    FILE* f = std::fopen(name.c_str(), "w");
    if (!f) return false;
    for (int64_t i = 0; i < rows; i++)
        for (int k = 0; k < cols; k++) std::fprintf(f, k + 1 < cols ? "%.10f\t" : "%.10f\n", v[i * cols + k]);
    return std::fclose(f) == 0;
}

ps_err ps_normal_stream(uint64_t seed, double* out, int64_t n, int64_t chunk, double margin) {
    if (!out || n < 0 || !(margin >= 0.0)) return PS_ERR_ARG;
    NormalStream g;
    g.seed(seed);
    if (margin > 0) g.set_margin(margin);
    if (chunk <= 0) chunk = n;
    for (int64_t i = 0; i < n; i += chunk) g.fill(out + i, std::min(chunk, n - i));
    return PS_OK;
}

ps_err ps_format_time(double t, char* buf, int len) {
    if (!buf || len <= 0) return PS_ERR_ARG;
    std::ostringstream oss;
    oss << t;
    const std::string s = oss.str();
    if ((int)s.size() + 1 > len) return PS_ERR_ARG;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return PS_OK;
}

ps_err ps_write_outputs(void* h, const char* output_dir) {
    Ctx* c = as_ctx(h);
    if (!c || !output_dir) return PS_ERR_ARG;
    if (!c->have_x) return fail(c, PS_ERR_STATE, "write_outputs before initial_configuration");
    const int64_t Np = c->Np, Nc = c->Nc;
    std::vector<double> x(3 * Np), q(4 * std::max<int64_t>(Nc, 1)), fx(c->nu), sb(5 * std::max<int64_t>(Nc, 1));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    PS_HIP(c, hipMemcpy(x.data(), c->x, sizeof(double) * 3 * Np, hipMemcpyDeviceToHost));
    if (Nc > 0) {
        PS_HIP(c, hipMemcpy(q.data(), c->q, sizeof(double) * 4 * Nc, hipMemcpyDeviceToHost));
        PS_HIP(c, hipMemcpy(sb.data(), c->sb, sizeof(double) * 5 * Nc, hipMemcpyDeviceToHost));
    }
    PS_HIP(c, hipMemcpy(fx.data(), c->fext, sizeof(double) * c->nu, hipMemcpyDeviceToHost));
    const double t = c->t;
    bool ok = write_table(frame_name(output_dir, "config_", t), x.data(), Np, 3);                 // :10-40
    ok = ok && write_table(frame_name(output_dir, "quats_", t), q.data(), Nc, 4);                  // :42-73
    if (c->p.record_forces) {                                                                      // :75-120
        // monomer + colloid forces (3 Np), then the colloid torques (3 Nc, from offset 3 Np)
        ok = ok && write_table(frame_name(output_dir, "force_", t), fx.data(), Np + Nc, 3);
    }
    ok = ok && write_table(frame_name(output_dir, "brownian_stresslet_", t), sb.data(), Nc, 5);    // :123-165
    if (!ok) return fail(c, PS_ERR_ARG, std::string("cannot write the frame files in ") + output_dir);
    return PS_OK;
}

// run() with the reference's output schedule (src/run.cpp:148-160): the frame files are written after
// loop index i when i == 0 or i % samplerate == 0.  Runs the steps between writes in one ps_run each.
ps_err ps_run_output(void* h, int nsteps, const char* output_dir) {
    Ctx* c = as_ctx(h);
    if (!c || nsteps < 0 || !output_dir) return PS_ERR_ARG;
    const int sr = std::max(1, c->p.samplerate);
    int i = 0;   // next loop index
    while (i < nsteps) {
        // steps up to and including the next index that writes
        const int next = (i == 0) ? 0 : ((i + sr - 1) / sr) * sr;
        const int last = std::min(next, nsteps - 1);
        PS_TRY(ps_run(h, last - i + 1));
        if (last == next) PS_TRY(ps_write_outputs(h, output_dir));
        i = last + 1;
    }
    return PS_OK;
}

ps_err ps_draw_normals(void* h, double* out, int64_t n) {
    Ctx* c = as_ctx(h);
    if (!c || (!out && n > 0) || n < 0) return PS_ERR_ARG;
    c->normals.fill(out, n);
    return PS_OK;
}

}  // extern "C"
