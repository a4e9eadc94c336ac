// Host-side thread helpers of libcig (graph synthesis and graph loading).
#pragma once
#include <algorithm>
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

}  // namespace cig
