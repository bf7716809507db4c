// Static partition of a host loop over up to 16 threads (set-up paths only).
#pragma once

#include <algorithm>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace astg {

inline int hostThreads() { return std::max(1, std::min(16, (int)std::thread::hardware_concurrency())); }

// Calls f(begin, end) on contiguous slices of [0, n); runs inline when n is
// small. The first exception thrown by a slice is rethrown after all join.
template <class F>
void parallelFor(long n, F&& f, long minSlice = 16384) {
    if (n <= 0) return;
    const int nt = (int)std::min<long>(hostThreads(), (n + minSlice - 1) / minSlice);
    if (nt <= 1) {
        f(0L, n);
        return;
    }
    std::exception_ptr first;
    std::mutex mu;
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int w = 0; w < nt; ++w)
        th.emplace_back([&, w] {
            try {
                f(n * w / nt, n * (w + 1) / nt);
            } catch (...) {
                std::lock_guard<std::mutex> g(mu);
                if (!first) first = std::current_exception();
            }
        });
    for (auto& x : th) x.join();
    if (first) std::rethrow_exception(first);
}

}  // namespace astg
