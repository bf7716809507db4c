// Device helpers shared by the projection kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sweep_merge.h"

namespace astg {

// Unpublished-value sentinels: signalling-NaN bit patterns. GPU arithmetic
// only ever produces quiet NaNs, so a computed value can never equal them.
constexpr unsigned long long kSent64 = 0x7FF4DEADBEEF0001ull;
constexpr unsigned int kSent32 = 0x7FA0BEEFu;
constexpr int kSentLevel = -1;

// Polls allowed per lane before a sweep gives up (error flag + NaN output).
// A healthy sweep needs at most a few thousand polls per lane.
constexpr int kSpinLimit = 1 << 22;

// scalars[] slots in the per-solver scalar block (doubles).
enum ScalarSlot {
    S_RHO = 0,      // rho = z.r
    S_RHO_NEW,      // new rho
    S_SQ,           // s.q
    S_RES,          // max |r| (or sum r^2 for L2)
    S_TOL,          // absolute tolerance
    S_DONE,         // 0 running; 1 converged, 2 breakdown, 3 iteration limit, 4 NaN or infinity
    S_ITER,         // iterations performed
    S_BNORM,        // ||b||
    S_ERR,          // nonzero: sweep spin limit hit / breakdown
    S_ALPHA,
    S_BETA,
    S_OUTER_RES,    // mixed form: FP64 residual norm of the outer iteration
    S_COUNT = 16
};
// int counters in the counter block
enum CounterSlot {
    C_TICKET_F = 0, C_TICKET_B, C_TICKET_L, C_TICKET_M, C_DONE_SQ, C_DONE_RES, C_DONE_RHO, C_DONE_B, C_ERR,
    C_PATTERN,  // set up: the pattern (or a symbolic plan) breaks a structural rule
    C_COUNT = 16
};
// Sticky error bits in cnt[C_ERR] (atomicOr): a spin limit was hit.
enum ErrBit { E_LEVELS = 1, E_MIC = 2, E_FWD = 4, E_BWD = 8 };

__device__ inline unsigned long long ld_agent_u64(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_agent_u64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned int ld_agent_u32(const unsigned int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_agent_u32(unsigned int* p, unsigned int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_agent_i32(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void st_agent_i32(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

template <typename T>
struct Sentinel;
template <>
struct Sentinel<double> {
    typedef unsigned long long bits_t;
    static constexpr bits_t value = kSent64;
    __device__ static bits_t load(const double* p) { return ld_agent_u64(reinterpret_cast<const bits_t*>(p)); }
    __device__ static void store(double* p, double v) { st_agent_u64(reinterpret_cast<bits_t*>(p), (bits_t)__double_as_longlong(v)); }
    __device__ static void storeSent(double* p) { *reinterpret_cast<bits_t*>(p) = value; }
    __device__ static double val(bits_t b) { return __longlong_as_double((long long)b); }
};
template <>
struct Sentinel<float> {
    typedef unsigned int bits_t;
    static constexpr bits_t value = kSent32;
    __device__ static bits_t load(const float* p) { return ld_agent_u32(reinterpret_cast<const bits_t*>(p)); }
    __device__ static void store(float* p, float v) { st_agent_u32(reinterpret_cast<bits_t*>(p), (bits_t)__float_as_uint(v)); }
    __device__ static void storeSent(float* p) { *reinterpret_cast<bits_t*>(p) = value; }
    __device__ static float val(bits_t b) { return __uint_as_float(b); }
};

// Wave64 reductions via DPP/shuffle.
__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// max that keeps a NaN from either side (fmax drops it), so that a norm over
// vectors holding a NaN is NaN
__device__ inline double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }
__device__ inline double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = nan_max(v, __shfl_xor(v, o, 64));
    return v;
}

// Polls of the advisory gate before a wavefront proceeds anyway.
constexpr int kGateLimit = 1 << 16;

// A row's dependencies held in registers: column indices and factor values
// are loaded before the row waits (they do not depend on other rows), then
// all published values are polled together, one round trip per poll. N is
// the register capacity (kMaxDeps, sweep_merge.h, by default).
template <typename T, int N = kMaxDeps>
struct DepRow {
    int nd;
    int c[N];
    T f[N];
    __device__ void preload(const int* col, const T* F, int p0, int pe) {
        nd = pe - p0;
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (j < nd) {
                c[j] = col[p0 + j];
                f[j] = F[p0 + j];
            }
    }
    // One polling round over the dependencies not seen yet. All loads are
    // issued before any result is examined, so a round costs one memory
    // round trip, not one per dependency.
    __device__ void poll_round(const T* vec, T* v, unsigned& ready) {
        typename Sentinel<T>::bits_t b[N];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            b[j] = Sentinel<T>::value;
            if (j < nd && !((ready >> j) & 1u)) b[j] = Sentinel<T>::load(&vec[c[j]]);
        }
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (b[j] != Sentinel<T>::value) {
                v[j] = Sentinel<T>::val(b[j]);
                ready |= 1u << j;
            }
    }
    // acc -= sum_j f[j] * vec[c[j]] in ascending j once every vec[c[j]] is
    // published, then publish(acc) -- inside the polling loop, in the round
    // the lane's own dependencies became ready, so a lane never waits for the
    // other lanes of its wavefront before publishing. Between polling rounds
    // the lane sleeps 1, 2, 4, ... up to `cap` units of 64 cycles. Returns
    // false (after publishing NaN) if the spin limit was hit, or if another
    // row of the launch has raised an error bit in *errBits (checked every
    // 256 rounds): once one row has given up, the others stop polling too, so
    // one non-resident block costs one spin budget, not one per waiting row.
    template <class Pub>
    __device__ bool run(const T* vec, T acc, int cap, const int* errBits, Pub publish) {
        T v[N];
        unsigned ready = 0;
        const unsigned full = nd >= 32 ? ~0u : (1u << nd) - 1u;
        int backoff = 1;
        bool pending = true, ok = true;
        // The loop exit is wave-uniform (vote), so the publishing branch is
        // part of the loop body: a lane publishes in the round it became ready.
        for (int spins = 0;; ++spins) {
            if (pending) {
                poll_round(vec, v, ready);
                if (ready == full) {
                    T a = acc;
#pragma unroll
                    for (int j = 0; j < N; ++j)
                        if (j < nd) a -= f[j] * v[j];
                    publish(a);
                    pending = false;
                } else if (spins > kSpinLimit ||
                           ((spins & 255) == 255 && ld_agent_i32(errBits) != 0)) {
                    publish((T)NAN);
                    pending = false;
                    ok = false;
                }
            }
            if (!__any(pending)) break;
            for (int k = 0; k < backoff; ++k) __builtin_amdgcn_s_sleep(1);
            backoff = backoff < cap ? 2 * backoff : cap;
        }
        return ok;
    }
};

}  // namespace astg
