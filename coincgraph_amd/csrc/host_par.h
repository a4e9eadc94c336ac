// Host-side thread helpers of libcig (graph synthesis and graph loading).
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <thread>
#include <vector>

namespace cig {

// Worker threads: the requested count, else the hardware's, at most 16.
inline unsigned pick_threads(uint32_t requested) {
    if (requested) return requested;
    unsigned hw = std::thread::hardware_concurrency();
    return std::max(1u, std::min(hw ? hw : 1u, 16u));
}

// f(lo, hi, t) over `threads` contiguous chunks of [0, n).
template <class F>
void parallel_for(uint64_t n, unsigned threads, F f) {
    if (threads <= 1 || n < 4096) { f(0, n, 0u); return; }
    std::vector<std::thread> pool;
    uint64_t chunk = (n + threads - 1) / threads;
    pool.reserve(threads);
    try {
        for (unsigned t = 0; t < threads; ++t) {
            uint64_t b = std::min<uint64_t>(n, t * chunk), e = std::min<uint64_t>(n, b + chunk);
            pool.emplace_back([=] { f(b, e, t); });
        }
    } catch (...) {
        // thread creation failed: join what started before rethrowing, so no
        // joinable std::thread is destroyed (that would call std::terminate)
        for (auto &th : pool) th.join();
        throw;
    }
    for (auto &th : pool) th.join();
}

// Barrier for a fixed team of threads (spins, then yields).
class SpinBarrier {
  public:
    explicit SpinBarrier(unsigned n) : n_(n) {}
    void wait() {
        const unsigned g = gen_.load(std::memory_order_acquire);
        if (count_.fetch_add(1, std::memory_order_acq_rel) + 1 == n_) {
            count_.store(0, std::memory_order_relaxed);
            gen_.fetch_add(1, std::memory_order_release);
            return;
        }
        for (unsigned spin = 0; gen_.load(std::memory_order_acquire) == g; ++spin)
            if (spin > 1024) std::this_thread::yield();
    }

  private:
    std::atomic<unsigned> count_{0}, gen_{0};
    const unsigned n_;
};

// Runs f(t, team) on a team of `threads` threads (t = 0 on the caller) that
// start together; if a thread cannot be created, the started ones return at
// once and f runs alone as a team of one.
template <class F>
void run_team(unsigned threads, F f) {
    std::atomic<int> go{0};  // 1: run, -1: give up
    std::vector<std::thread> pool;
    pool.reserve(threads);
    auto member = [&](unsigned t) {
        int g;
        while ((g = go.load(std::memory_order_acquire)) == 0) std::this_thread::yield();
        if (g > 0) f(t, threads);
    };
    bool ok = true;
    try {
        for (unsigned t = 1; t < threads; ++t) pool.emplace_back(member, t);
    } catch (...) {
        ok = false;
    }
    go.store(ok ? 1 : -1, std::memory_order_release);
    if (ok) f(0u, threads);
    for (auto &th : pool) th.join();
    if (!ok) f(0u, 1u);
}

}  // namespace cig
