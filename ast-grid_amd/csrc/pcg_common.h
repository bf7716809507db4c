// Device helpers shared by the single-system and batched PCG translation
// units: launch constants, deterministic block reductions, the SpMV row dot.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "device_common.h"

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);                 \
            return ASTG_ERR_HIP;                                                     \
        }                                                                            \
    } while (0)

namespace astg {

constexpr int BS = 256;
constexpr int kLpr = 8;  // lanes per row in the SpMV kernels

// ---------------------------------------------------------------------------
// deterministic block reductions
// ---------------------------------------------------------------------------
template <bool MAX>
__device__ inline double block_reduce(double v) {
    __shared__ double sm[BS / 64];
    v = MAX ? wave_max(v) : wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();  // sm may still be read by a previous call
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    double t = sm[0];
    for (int w = 1; w < BS / 64; ++w) t = MAX ? fmax(t, sm[w]) : t + sm[w];
    return t;
}
// Sum (or max) of np partials written by the previous kernel, same order in
// every block.
template <bool MAX>
__device__ inline double reduce_prev(const double* part, int np) {
    double acc = 0.0;
    for (int s = threadIdx.x; s < np; s += BS) acc = MAX ? fmax(acc, part[s]) : acc + part[s];
    return block_reduce<MAX>(acc);
}

// (A s)_i with LPR lanes per row (rows are short: 8 entries on average), so
// a wavefront reads the entries of 64/LPR consecutive rows contiguously; the
// row sum is a shuffle reduction over the row's lanes. i < 0: no row.
template <typename T, int LPR = kLpr>
__device__ inline T row_dot(const int* rs, const int* col, const T* A, const T* s, int i, int sub) {
    T acc = (T)0;
    if (i >= 0) {
        for (int p = rs[i] + sub, pe = rs[i + 1]; p < pe; p += LPR) acc += A[p] * s[col[p]];
    }
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
    return acc;
}

}  // namespace astg
