// Ownership of the per-topology kernel constants (see consts.h).
#include "consts.h"

#include <hip/hip_runtime.h>

#include <atomic>

namespace astg {

namespace {
constexpr int kMaxGpus = 64;
uint64_t g_owner[kMaxGpus][CT_COUNT] = {};
std::atomic<uint64_t> g_next{1};
}  // namespace

std::recursive_mutex& constMutex() {
    static std::recursive_mutex m;
    return m;
}

uint64_t newConstOwner() { return g_next.fetch_add(1); }

bool takeConstants(int table, uint64_t owner, bool force) {
    std::lock_guard<std::recursive_mutex> lk(constMutex());
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxGpus) return true;
    uint64_t& cur = g_owner[dev][table];
    if (cur == owner && !force) return false;
    if (cur != 0 && cur != owner) (void)hipDeviceSynchronize();
    cur = owner;
    return true;
}

}  // namespace astg
