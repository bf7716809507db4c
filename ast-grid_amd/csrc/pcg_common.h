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
    for (int w = 1; w < BS / 64; ++w) t = MAX ? nan_max(t, sm[w]) : t + sm[w];
    return t;
}
// Sum (or max) of np partials written by the previous kernel, same order in
// every block.
template <bool MAX>
__device__ inline double reduce_prev(const double* part, int np) {
    double acc = 0.0;
    for (int s = threadIdx.x; s < np; s += BS) acc = MAX ? nan_max(acc, part[s]) : acc + part[s];
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

// The same product with more loads in flight: a lane group holds U rows at
// once and each lane's first four entries (rows of up to 4 * LPR entries; the
// octree's hold at most 15) are loaded without a loop, so the index, value and
// gather loads of all U rows issue before any is used. Each lane sums its
// entries in the same order as row_dot, so the results are bit-identical.
template <typename T, int LPR, int U>
__device__ inline void row_dot_u(const int* rs, const int* col, const T* A, const T* s, const int* i, int sub,
                                 T* out) {
    int p[U], e[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        p[u] = i[u] >= 0 ? rs[i[u]] + sub : 0;
        e[u] = i[u] >= 0 ? rs[i[u] + 1] : 0;
    }
    T a[U][4];
    int c[U][4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = p[u] + k * LPR;
            c[u][k] = q < e[u] ? col[q] : -1;
            a[u][k] = q < e[u] ? A[q] : (T)0;
        }
    T x[U][4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int u = 0; u < U; ++u) x[u][k] = c[u][k] >= 0 ? s[c[u][k]] : (T)0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        T acc = (T)0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (c[u][k] >= 0) acc += a[u][k] * x[u][k];
        for (int q = p[u] + 4 * LPR; q < e[u]; q += LPR) acc += A[q] * s[col[q]];
#pragma unroll
        for (int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
        out[u] = acc;
    }
}

}  // namespace astg
