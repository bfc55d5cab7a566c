// normals.h -- the reference's noise stream, generated in bulk.
//
// The reference draws its noise with std::mt19937_64 and std::normal_distribution<double>(0, 1)
// (src/init.cpp:105-110).  This class produces the identical sequence of doubles, bit for bit, as
// libstdc++ computes it:
//   - MT19937-64 (Matsumoto & Nishimura 2000; the parameters the C++ standard fixes for
//     mt19937_64), seeded by the standard's initialisation recurrence;
//   - generate_canonical<double, 53>: u = double(draw) / 2^64, clamped to nextafter(1, 0);
//   - Marsaglia's polar method: x = 2u - 1, y = 2u' - 1 until 0 < x^2 + y^2 <= 1, then
//     m = sqrt(-2 log(r2) / r2), return y m and keep x m for the next call.
// Every attempt of the polar method consumes exactly two consecutive draws, so which attempts are
// accepted depends on each pair of draws alone: only the MT recurrence is serial.  The recurrence
// runs 312-word blocks straight into a draw buffer (vectorised for the host's widest SIMD); the
// attempts are then split into segments on a small pool of host threads, one pass counting the
// accepted attempts of each segment and a second writing each segment's normals at its prefix
// offset.  Draws past the last attempt a fill uses stay buffered for the next fill.  The file is
// compiled without floating-point contraction so r2 rounds as the reference's does.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

class NormalStream {
public:
    NormalStream() = default;
    ~NormalStream();
    NormalStream(const NormalStream&) = delete;
    NormalStream& operator=(const NormalStream&) = delete;
    void seed(uint64_t s);
    // the next n values of the stream, exactly as n successive normal_distribution calls
    void fill(double* out, int64_t n);
    // attempts a fill first takes per pair still missing (any value > 0 gives the same stream)
    void set_margin(double m) { margin_ = m; }
    // cap on the host threads a fill uses (1: the caller's thread only)
    void set_max_parts(int m) { max_parts_ = m < 1 ? 1 : m; }

private:
    uint64_t mt_[312] = {};        // MT state; always at the end of a block (all 312 words used)
    double margin_ = 1.35;
    int max_parts_ = 8;
    bool saved_ok_ = false;
    double saved_ = 0.0;
    std::vector<uint64_t> buf_;    // buffered draws [head_, tail_)
    size_t head_ = 0, tail_ = 0;
    void draw_at_least(size_t n);  // whole blocks until n draws are buffered
    struct Segment {
        int64_t a, b;              // attempts [a, b) counted from head_
        int64_t accepted, base;    // accepted attempts; pairs before the segment
    };
    std::vector<Segment> seg_;
    std::vector<double> ax_, ay_, ar_, al_;   // accepted attempts' x, y, r2, log r2 (per segment)
    std::vector<int64_t> aj_;                 // and their attempt indices
    int64_t last_attempt_ = 0;
    // worker pool: a job is published by bumping gen_; workers spin on it for a while, then sleep
    // on cv_ (fills come once per time step, a fraction of a millisecond apart, and a sleeping
    // thread takes tens of microseconds to wake)
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_;
    const std::function<void(int)>* job_ = nullptr;
    std::atomic<int64_t> gen_{0};
    std::atomic<int> pending_{0};
    std::atomic<bool> stop_{false};
    int nparts_ = 1;
    void worker(int id, int64_t seen);
    void run_parallel(int parts, const std::function<void(int)>& job);
};
