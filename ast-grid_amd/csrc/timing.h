// Host phase timer for set-up paths, printed to stderr when ASTG_TIMING is
// set in the environment (diagnostics only; no effect otherwise).
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>

namespace astg {

class PhaseTimer {
public:
    explicit PhaseTimer(const char* scope) : scope_(scope), on_(std::getenv("ASTG_TIMING") != nullptr) {
        t_ = std::chrono::steady_clock::now();
    }
    // a count and the milliseconds it took, outside the phase clock
    void note(const char* what, double count, double ms) const {
        if (on_) std::fprintf(stderr, "[astg timing] %s/%s %.0f in %.2f ms\n", scope_, what, count, ms);
    }
    void mark(const char* phase) {
        if (!on_) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[astg timing] %s/%s %.2f ms\n", scope_, phase,
                     std::chrono::duration<double, std::milli>(now - t_).count());
        t_ = now;
    }

private:
    const char* scope_;
    bool on_;
    std::chrono::steady_clock::time_point t_;
};

}  // namespace astg
