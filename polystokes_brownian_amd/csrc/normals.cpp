// normals.cpp -- host code of NormalStream (normals.h).  Host-only C++, compiled by g++ without
// floating-point contraction.
#include "normals.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {
constexpr int MT_N = 312, MT_M = 156;
constexpr uint64_t MT_A = 0xB5026F5AA96619E9ULL;
constexpr uint64_t MT_UPPER = 0xFFFFFFFF80000000ULL;   // the upper 33 bits
constexpr uint64_t MT_LOWER = 0x000000007FFFFFFFULL;   // the lower 31 bits
constexpr int64_t ATTEMPTS_PER_PART_MIN = 4096;
constexpr int SPIN_LIMIT = 1 << 16;   // pause-loop iterations, a few milliseconds

inline void cpu_relax() { __builtin_ia32_pause(); }

inline uint64_t mt_next(uint64_t cur, uint64_t nxt, uint64_t far) {
    const uint64_t y = (cur & MT_UPPER) | (nxt & MT_LOWER);
    return far ^ (y >> 1) ^ (((uint64_t)0 - (y & 1)) & MT_A);
}

// `nblocks` whole blocks of the MT19937-64 output: the twist of all 312 words, then the tempering
// of each.  Integer-only, so every clone gives the same words; the loops are branch-free so the
// compiler vectorises them (the first 156 words read only old state, the next 155 the new words
// 156 behind them).
__attribute__((target_clones("avx512f", "avx2", "default")))
void mt_blocks(uint64_t* mt, uint64_t* out, int64_t nblocks) {
    for (int64_t b = 0; b < nblocks; b++) {
        for (int i = 0; i < MT_N - MT_M; i++) mt[i] = mt_next(mt[i], mt[i + 1], mt[i + MT_M]);
        for (int i = MT_N - MT_M; i < MT_N - 1; i++) mt[i] = mt_next(mt[i], mt[i + 1], mt[i + MT_M - MT_N]);
        mt[MT_N - 1] = mt_next(mt[MT_N - 1], mt[0], mt[MT_M - 1]);
        uint64_t* o = out + (size_t)b * MT_N;
        for (int i = 0; i < MT_N; i++) {
            uint64_t y = mt[i];
            y ^= (y >> 29) & 0x5555555555555555ULL;
            y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
            y ^= (y << 37) & 0xFFF7EEE000000000ULL;
            y ^= y >> 43;
            o[i] = y;
        }
    }
}

// generate_canonical<double, 53> on one 64-bit draw, then the polar method's 2u - 1.  The draw is
// converted as hi * 2^32 + lo: both terms are exact doubles, so the one rounding of their sum is
// the correctly rounded double(draw), without the branch on the top bit a direct unsigned
// conversion compiles to.  u >= 1 (a draw that rounds to 2^64) is clamped to nextafter(1, 0).
inline double polar_coord(uint64_t draw) {
    const double v = (double)(uint32_t)(draw >> 32) * 4294967296.0 + (double)(uint32_t)draw;
    const double u = std::min(v * 0x1p-64, 0x1.fffffffffffffp-1);
    return 2.0 * u - 1.0;
}

// Attempts [a, b) of the polar method on the draws d (attempt j uses d[2j], d[2j+1]): the accepted
// ones' x, y, r2 and attempt index are written, in order, from index a on; returns how many.
// Branch-free (the store index advances by the acceptance), so rejections cost no mispredictions.
int64_t polar_compact(const uint64_t* d, int64_t a, int64_t b, double* xs, double* ys, double* r2s, int64_t* js) {
    int64_t c = a;
    for (int64_t j = a; j < b; j++) {
        const double x = polar_coord(d[2 * j]), y = polar_coord(d[2 * j + 1]);
        const double r2 = x * x + y * y;
        xs[c] = x;
        ys[c] = y;
        r2s[c] = r2;
        js[c] = j;
        c += !(r2 > 1.0 || r2 == 0.0);
    }
    return c - a;
}

// The polar method's outputs for cnt accepted pairs, given lg = log(r2): m = sqrt(-2 lg / r2),
// then y m and x m (plus zero, as the distribution adds its mean).  Division and square root are
// correctly rounded in every SIMD width, so the clones give identical values.
__attribute__((target_clones("avx512f", "avx2", "default")))
void polar_finish(const double* xs, const double* ys, const double* r2s, const double* lg, double* o, int64_t cnt) {
    for (int64_t q = 0; q < cnt; q++) {
        const double mult = std::sqrt(-2 * lg[q] / r2s[q]);
        o[2 * q] = (ys[q] * mult) * 1.0 + 0.0;
        o[2 * q + 1] = (xs[q] * mult) * 1.0 + 0.0;
    }
}
}  // namespace

NormalStream::~NormalStream() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_.store(true);
    }
    cv_.notify_all();
    for (auto& t : th_)
        if (t.joinable()) t.join();
}

void NormalStream::seed(uint64_t s) {
    head_ = tail_ = 0;
    mt_[0] = s;
    for (int i = 1; i < MT_N; i++) mt_[i] = 6364136223846793005ULL * (mt_[i - 1] ^ (mt_[i - 1] >> 62)) + (uint64_t)i;
    saved_ok_ = false;
    saved_ = 0.0;
}

void NormalStream::draw_at_least(size_t n) {
    const size_t have = tail_ - head_;
    if (have >= n) return;
    const size_t blocks = (n - have + MT_N - 1) / MT_N;
    if (head_ > 0) {   // move the unused draws to the front
        std::memmove(buf_.data(), buf_.data() + head_, have * sizeof(uint64_t));
        head_ = 0;
        tail_ = have;
    }
    const size_t cap = tail_ + blocks * MT_N;
    if (buf_.size() < cap) buf_.resize(std::max(cap, buf_.size() + buf_.size() / 2));
    mt_blocks(mt_, buf_.data() + tail_, (int64_t)blocks);
    tail_ += blocks * MT_N;
}

void NormalStream::worker(int id, int64_t seen) {
    for (;;) {
        int64_t g = gen_.load(std::memory_order_acquire);
        for (int spin = 0; g == seen && spin < SPIN_LIMIT && !stop_.load(std::memory_order_relaxed); spin++) {
            cpu_relax();
            g = gen_.load(std::memory_order_acquire);
        }
        if (g == seen) {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_.load() || gen_.load() != seen; });
            g = gen_.load(std::memory_order_acquire);
        }
        if (stop_.load()) return;
        seen = g;
        if (id < nparts_) (*job_)(id);
        // every worker acknowledges every job, so none can lag into the next one's data
        pending_.fetch_sub(1, std::memory_order_acq_rel);
    }
}

void NormalStream::run_parallel(int parts, const std::function<void(int)>& job) {
    if (parts <= 1) {
        job(0);
        return;
    }
    for (int id = (int)th_.size() + 1; id < parts; id++) {
        const int64_t now = gen_.load();   // a new worker waits for the next job, never a finished one
        th_.emplace_back([this, id, now] { worker(id, now); });
    }
    job_ = &job;
    nparts_ = parts;
    pending_.store((int)th_.size(), std::memory_order_relaxed);
    {
        std::lock_guard<std::mutex> lk(mu_);   // a worker about to sleep sees the new generation
        gen_.fetch_add(1, std::memory_order_release);
    }
    cv_.notify_all();
    job(0);
    for (int spin = 0; pending_.load(std::memory_order_acquire) != 0; spin++) {
        if (spin < SPIN_LIMIT) cpu_relax();
        else std::this_thread::yield();
    }
}

void NormalStream::fill(double* out, int64_t n) {
    if (n <= 0) return;
    int64_t k = 0;
    if (saved_ok_) {
        out[0] = saved_ * 1.0 + 0.0;
        saved_ok_ = false;
        k = 1;
    }
    const int64_t need = n - k;
    if (need == 0) return;
    const int64_t np = (need + 1) / 2;   // pairs; with an odd count the last keeps its x value
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int maxparts = (int)std::min<unsigned>((unsigned)max_parts_, std::max(1u, hw / 2));
    auto parts_for = [&](int64_t attempts) {
        return (int)std::max<int64_t>(1, std::min<int64_t>(maxparts, attempts / ATTEMPTS_PER_PART_MIN));
    };
    // pass 1: each segment of attempts compacts its accepted (x, y, r2) to its own front.
    // Acceptance is pi/4 per attempt, so margin_ (1.35) attempts per pair still missing are taken
    // at once (plus 64); a shortfall takes a further range.
    seg_.clear();
    int64_t scanned = 0, total = 0;
    int parts2 = 1;
    while (total < np) {
        const int64_t want = scanned + (int64_t)((double)(np - total) * margin_) + 64;
        draw_at_least(2 * (size_t)want);
        if ((int64_t)ax_.size() < want) {
            const size_t cap = (size_t)want + (size_t)want / 4;
            ax_.resize(cap);
            ay_.resize(cap);
            ar_.resize(cap);
            al_.resize(cap);
            aj_.resize(cap);
        }
        const uint64_t* d = buf_.data() + head_;
        const int parts = parts_for(want - scanned);
        parts2 = std::max(parts2, parts);
        const size_t s0 = seg_.size();
        for (int i = 0; i < parts; i++)
            seg_.push_back({scanned + (want - scanned) * i / parts, scanned + (want - scanned) * (i + 1) / parts, 0, 0});
        Segment* sg = seg_.data() + s0;
        run_parallel(parts, [&](int id) {
            sg[id].accepted = polar_compact(d, sg[id].a, sg[id].b, ax_.data(), ay_.data(), ar_.data(), aj_.data());
        });
        for (int i = 0; i < parts; i++) {
            sg[i].base = total;
            total += sg[i].accepted;
        }
        scanned = want;
    }
    // pass 2: each segment transforms its pairs into the output from its first pair index on
    const bool odd = (need & 1) != 0;
    double* o = out + k;
    const int nseg = (int)seg_.size();
    const Segment* sg = seg_.data();
    run_parallel(std::min(parts2, nseg), [&, parts = std::min(parts2, nseg)](int id) {
        for (int s = id; s < nseg; s += parts) {
            const int64_t p0 = sg[s].base, a = sg[s].a;
            if (p0 >= np) continue;
            const int64_t cnt = std::min(sg[s].accepted, np - p0);
            const bool last = p0 + cnt == np;   // this segment holds the fill's last pair
            const double *x = ax_.data() + a, *y = ay_.data() + a, *r2 = ar_.data() + a;
            double* lg = al_.data() + a;
            for (int64_t q = 0; q < cnt; q++) lg[q] = std::log(r2[q]);
            polar_finish(x, y, r2, lg, o + 2 * p0, last ? cnt - 1 : cnt);
            if (last) {
                const int64_t q = cnt - 1;
                const double mult = std::sqrt(-2 * lg[q] / r2[q]);
                o[2 * (p0 + q)] = (y[q] * mult) * 1.0 + 0.0;
                if (odd) {
                    saved_ = x[q] * mult;
                    saved_ok_ = true;
                } else {
                    o[2 * (p0 + q) + 1] = (x[q] * mult) * 1.0 + 0.0;
                }
                last_attempt_ = aj_[a + q];
            }
        }
    });
    head_ += 2 * (size_t)(last_attempt_ + 1);   // the draws after the last attempt used stay buffered
}
