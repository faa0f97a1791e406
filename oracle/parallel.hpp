// Test infrastructure only: the CPU oracle.  Nothing outside tests/, the smoke
// check and bench.py's cpu_baseline leg may link or call this code.
//
// A std::thread work pool standing in for rayon's global pool: `par_for`
// hands out indices in small dynamic chunks, as rayon's work stealing does.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstddef>
#include <thread>
#include <vector>

namespace oracle {

inline int default_threads() {
    unsigned n = std::thread::hardware_concurrency();
    return n == 0 ? 1 : (int)n;
}

template <class F>
void par_for(size_t n, int threads, F&& f, size_t chunk = 64) {
    if (threads <= 1 || n <= chunk) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (;;) {
            size_t s = next.fetch_add(chunk);
            if (s >= n) return;
            size_t e = std::min(n, s + chunk);
            for (size_t i = s; i < e; ++i) f(i);
        }
    };
    std::vector<std::thread> pool;
    int t = std::min<size_t>(threads, (n + chunk - 1) / chunk);
    for (int k = 1; k < t; ++k) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

}  // namespace oracle
