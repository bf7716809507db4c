// Kernels of the device-built symbolic plans (plan_device.hip): one lane per
// row (factorisation plan, mic_plan.h) or per sweep position (merged sweep
// stencils, sweep_merge.h), each a count pass and a fill pass.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sweep_merge.h"

namespace astg {
namespace plan {

__global__ void k_set_i64(long long* p, long long v) {
    if (threadIdx.x == 0) p[0] = v;
}
__global__ void k_set_i32(int* p, int v) {
    if (threadIdx.x == 0) p[0] = v;
}

// Factorisation plan of row k (mic_plan.h): a count pass stores each entry's
// segment length at len[a]; the fill pass writes the segment at start[a].
template <bool FILL>
__global__ __launch_bounds__(256) void k_plan_mic(int n, const int* __restrict__ rs, const int* __restrict__ lend,
                                                  const int* __restrict__ ubeg, const int* __restrict__ col,
                                                  long long* __restrict__ len, const long long* __restrict__ start,
                                                  int* __restrict__ prog) {
    const int k = (int)(blockIdx.x * (long)blockDim.x + threadIdx.x);
    if (k >= n) return;
    const int r0 = rs[k], le = lend[k], ub = ubeg[k], r1 = rs[k + 1];
    if (!FILL)
        for (int a = le; a < ub; ++a) len[a] = 0;  // the diagonal
    // lower entry (k,m): upper entries of row m missing from row k's pattern
    for (int a = r0; a < le; ++a) {
        const int m = col[a];
        long long w = FILL ? start[a] : 0;
        int b = r0, cnt = 0;
        for (int p = ubeg[m], pe = rs[m + 1]; p < pe; ++p) {
            const int c = col[p];
            if (c == k) continue;
            while (b < r1 && col[b] < c) ++b;
            if (b < r1 && col[b] == c) continue;
            if (FILL)
                prog[w++] = p;
            else
                ++cnt;
        }
        if (!FILL) len[a] = cnt;
    }
    // upper entry (k,i): columns common to lower(k) and lower(i), as pairs
    for (int p = ub; p < r1; ++p) {
        const int i = col[p];
        long long w = FILL ? start[p] : 0;
        int a = r0, c = rs[i], cnt = 0;
        const int ce = lend[i];
        while (a < le && c < ce) {
            const int ca = col[a], cc = col[c];
            if (ca < cc) {
                ++a;
            } else if (cc < ca) {
                ++c;
            } else {
                if (FILL) {
                    prog[w++] = a - r0;
                    prog[w++] = c;
                } else {
                    cnt += 2;
                }
                ++a;
                ++c;
            }
        }
        if (!FILL) len[p] = cnt;
    }
}

__global__ void k_sweep_pos(int P, const int* __restrict__ order, int* __restrict__ pos) {
    const int t = (int)(blockIdx.x * (long)blockDim.x + threadIdx.x);
    if (t >= P) return;
    const int i = order[t];
    if (i >= 0) pos[i] = t;
}

constexpr int kKeyCap = kMaxDeps + kMaxRTerms;  // live keys <= maxY <= kMaxDeps, plus the substituted ones

template <int DIR>
__device__ inline int depBegin(const int* rs, const int* ubeg, int x) {
    return DIR == 0 ? rs[x] : ubeg[x];
}
template <int DIR>
__device__ inline int depEnd(const int* rs, const int* lend, int x) {
    return DIR == 0 ? lend[x] : rs[x + 1];
}

__device__ inline int findKey(const int* keys, int nk, int v) {
    for (int a = 0; a < nk; ++a)
        if (keys[a] == v) return a;
    return -1;
}

// Merged stencil of the row at sweep position t (sweep_merge.h): the greedy
// substitution, then the entry list (r-terms in substitution order, y-terms
// ascending) and the slot program. The count pass writes the sizes at
// cntIdx[t] / cntUpd[t] and the r-term count at nr[t]; the fill pass writes
// the entries at start[t] and the slots at ustart[t]. err receives 1 if a
// row breaks the key bound (cannot happen for patterns setPattern accepts).
template <int DIR, bool FILL>
__global__ __launch_bounds__(64) void k_plan_stencil(int P, const int* __restrict__ rs, const int* __restrict__ lend,
                                                     const int* __restrict__ ubeg, const int* __restrict__ col,
                                                     const int* __restrict__ lev, int q, int maxY,
                                                     const int* __restrict__ order, const int* __restrict__ pos,
                                                     int* __restrict__ cntIdx, int* __restrict__ cntUpd,
                                                     int* __restrict__ nr, const int* __restrict__ start,
                                                     const int* __restrict__ ustart, int* __restrict__ idx,
                                                     uint8_t* __restrict__ upd, int* __restrict__ err) {
    const int t = (int)(blockIdx.x * (long)blockDim.x + threadIdx.x);
    if (t >= P) return;
    const int i = order[t];
    if (i < 0) {
        if (!FILL) {
            cntIdx[t] = 0;
            cntUpd[t] = 0;
            nr[t] = 0;
        }
        return;
    }
    int keys[kKeyCap];
    uint8_t state[kKeyCap];  // 0 open, 1 substituted, 2 kept
    uint8_t sub[kMaxRTerms];  // key slots of the substituted keys, in order
    int nk = 0, ns = 0;
    {
        const int b = depBegin<DIR>(rs, ubeg, i), e = depEnd<DIR>(rs, lend, i);
        if (e - b > kMaxDeps) {
            atomicOr(err, 1);
            return;
        }
        for (int p = b; p < e; ++p) {
            keys[nk] = col[p];
            state[nk++] = 0;
        }
    }
    int live = nk;
    const int g = lev[i] / q;
    for (;;) {
        int best = -1, bl = 0;
        for (int a = 0; a < nk; ++a) {
            if (state[a] != 0) continue;
            const int la = lev[keys[a]];
            if (la / q != g) continue;
            if (best < 0 || (DIR == 0 ? la > bl : la < bl)) {
                best = a;
                bl = la;
            }
        }
        if (best < 0) break;
        const int j = keys[best];
        const int jb = depBegin<DIR>(rs, ubeg, j), je = depEnd<DIR>(rs, lend, j);
        int add = 0;
        for (int p = jb; p < je; ++p)
            if (findKey(keys, nk, col[p]) < 0) ++add;
        if (live - 1 + add > maxY || ns >= kMaxRTerms) {
            state[best] = 2;
            continue;
        }
        if (nk + add > kKeyCap) {  // excluded by live <= maxY <= kMaxDeps
            atomicOr(err, 1);
            return;
        }
        state[best] = 1;
        sub[ns++] = (uint8_t)best;
        live += add - 1;
        const int nk0 = nk;
        for (int p = jb; p < je; ++p) {
            const int c = col[p];
            if (findKey(keys, nk0, c) < 0) {  // the row's columns are distinct
                keys[nk] = c;
                state[nk++] = 0;
            }
        }
    }
    // y-terms: the keys not substituted, ascending
    int ys[kMaxDeps];
    int ny = 0;
    for (int a = 0; a < nk; ++a) {
        if (state[a] == 1) continue;
        if (ny >= kMaxDeps) {  // excluded by live <= maxY
            atomicOr(err, 1);
            return;
        }
        int v = keys[a], s = ny++;
        while (s > 0 && ys[s - 1] > v) {
            ys[s] = ys[s - 1];
            --s;
        }
        ys[s] = v;
    }
    if (!FILL) {
        int u = depEnd<DIR>(rs, lend, i) - depBegin<DIR>(rs, ubeg, i);
        for (int s = 0; s < ns; ++s) {
            const int j = keys[sub[s]];
            u += depEnd<DIR>(rs, lend, j) - depBegin<DIR>(rs, ubeg, j);
        }
        cntIdx[t] = ns + ny;
        cntUpd[t] = u;
        nr[t] = ns;
        return;
    }
    // slot of every key: substituted keys first (in order), then the y-terms
    uint8_t slot[kKeyCap];
    for (int s = 0; s < ns; ++s) slot[sub[s]] = (uint8_t)s;
    for (int a = 0; a < nk; ++a) {
        if (state[a] == 1) continue;
        int r = 0;
        while (ys[r] != keys[a]) ++r;
        slot[a] = (uint8_t)(ns + r);
    }
    // every dependency of i and of a substituted key is a key
    auto slotOf = [&](int c) {
        const int a = findKey(keys, nk, c);
        return a >= 0 ? slot[a] : (uint8_t)255;
    };
    int w = ustart[t];
    for (int p = depBegin<DIR>(rs, ubeg, i), e = depEnd<DIR>(rs, lend, i); p < e; ++p) upd[w++] = slotOf(col[p]);
    for (int s = 0; s < ns; ++s) {
        const int j = keys[sub[s]];
        for (int p = depBegin<DIR>(rs, ubeg, j), e = depEnd<DIR>(rs, lend, j); p < e; ++p) upd[w++] = slotOf(col[p]);
    }
    int x = start[t];
    for (int s = 0; s < ns; ++s) {
        const int j = keys[sub[s]];
        idx[x++] = DIR == 0 ? j : pos[j];
    }
    for (int s = 0; s < ny; ++s) idx[x++] = pos[ys[s]];
}

}  // namespace plan
}  // namespace astg
