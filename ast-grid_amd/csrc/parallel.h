// Static partition of a host loop over up to 16 threads (set-up paths only).
//
// The worker threads persist for the life of the process, so a loop of a
// few thousand rows can be split without paying for thread creation each
// time. One loop runs on the pool at a time; a second caller that finds it
// busy (another thread of the host program) starts threads of its own.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include <unistd.h>

namespace astg {

inline int hostThreads() { return std::max(1, std::min(16, (int)std::thread::hardware_concurrency())); }

class HostPool {
public:
    // Never destroyed: the workers wait until the process ends (no join at
    // exit, nothing to tear down in a forked child).
    static HostPool& get() {
        static HostPool* pool = new HostPool(hostThreads());
        return *pool;
    }
    // Runs job(w) for w in [0, nt): w = 0 on the caller, the rest on workers
    // 1 .. nt-1, each woken on its own. False (nothing run) if the pool is
    // busy, has fewer than nt threads or belongs to the parent of a forked
    // process (whose workers are gone).
    bool run(int nt, const std::function<void(int)>& job) {
        if (nt > (int)workers_.size() + 1 || getpid() != pid_) return false;
        std::unique_lock<std::mutex> busy(busy_, std::try_to_lock);
        if (!busy.owns_lock()) return false;
        job_ = &job;
        pending_.store(nt - 1);
        for (int w = 1; w < nt; ++w) {
            Worker& k = *workers_[w - 1];
            {
                std::lock_guard<std::mutex> g(k.m);
                ++k.gen;
            }
            k.cv.notify_one();
        }
        job(0);
        std::unique_lock<std::mutex> l(doneM_);
        done_.wait(l, [&] { return pending_.load() == 0; });
        job_ = nullptr;
        return true;
    }

private:
    struct Worker {
        std::mutex m;
        std::condition_variable cv;
        unsigned gen = 0;
    };
    explicit HostPool(int n) : pid_(getpid()) {
        for (int w = 1; w < n; ++w) workers_.emplace_back(new Worker);
        for (int w = 1; w < n; ++w) std::thread([this, w] { loop(w); }).detach();
    }
    void loop(int w) {
        Worker& me = *workers_[w - 1];
        unsigned seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> l(me.m);
                me.cv.wait(l, [&] { return me.gen != seen; });
                seen = me.gen;
            }
            (*job_)(w);
            if (pending_.fetch_sub(1) == 1) {
                std::lock_guard<std::mutex> g(doneM_);
                done_.notify_one();
            }
        }
    }
    const pid_t pid_;
    std::vector<std::unique_ptr<Worker>> workers_;
    std::mutex busy_, doneM_;
    std::condition_variable done_;
    const std::function<void(int)>* job_ = nullptr;
    std::atomic<int> pending_{0};
};

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
    const std::function<void(int)> slice = [&](int w) {
        try {
            f(n * w / nt, n * (w + 1) / nt);
        } catch (...) {
            std::lock_guard<std::mutex> g(mu);
            if (!first) first = std::current_exception();
        }
    };
    if (!HostPool::get().run(nt, slice)) {
        std::vector<std::thread> th;
        th.reserve(nt);
        for (int w = 0; w < nt; ++w) th.emplace_back(slice, w);
        for (auto& x : th) x.join();
    }
    if (first) std::rethrow_exception(first);
}

}  // namespace astg
