// Device MIC(0)-PCG (see pcg.h). Kernels are written for gfx950 wave64.
//
// Synchronisation design
//   * The sweeps process rows in order of groups of q consecutive dependency
//     levels, every group padded so a 64-row wavefront segment holds rows of
//     one group only. Within a group, rows depend only on earlier groups
//     (dependencies inside the group are substituted away, sweep_merge.h).
//   * The sweeps are persistent kernels with a grid at the occupancy limit:
//     all blocks are co-resident and block b takes chunks b, b+G, b+2G, ...
//     in order. The smallest unfinished chunk therefore always belongs to a
//     running block whose dependencies (earlier chunks) are complete, so the
//     spin-waits end. No ticket atomics are needed. They are launched as
//     ordinary kernels (an idle device makes such a grid resident at once);
//     a spin-limit error makes the solve retry with cooperative launches,
//     which guarantee co-residency.
//   * A row publishes its value with one agent-scope atomic store over a
//     signalling-NaN sentinel; dependants poll the values they need, all at
//     once. A wavefront first waits (one lane, with backoff) on a single
//     representative row of the group `dist` groups back, which keeps the
//     number of polling lanes to a few groups' worth.
//   * The factorisation runs one plain launch per dependency level.
//   * Reductions are deterministic and atomic-free: every block writes one
//     partial, and the next kernel in the stream reduces the partials it needs
//     in a fixed order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "device_common.h"
#include "pcg.h"
#include "pcg_common.h"
#include "mic_plan.h"
#include "sweep_merge.h"
#include "timing.h"
#include "parallel.h"

namespace astg {

namespace {
std::atomic<long long> g_allocCount{0}, g_allocNs{0};
}
void allocStats(long long& count, double& ms) {
    count = g_allocCount.load();
    ms = g_allocNs.load() * 1e-6;
}

DevBuf::~DevBuf() { release(); }
void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}
hipError_t DevBuf::alloc(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    const auto t0 = std::chrono::steady_clock::now();
    // A buffer that has to grow gets 1/8 more than asked: tilt updates move
    // the unknown count up and down by a little every step, and each
    // reallocation costs a device synchronisation (hipFree) and a hipMalloc.
    const size_t cap = p ? b + b / 8 : b;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, cap ? cap : 16);
    if (e == hipSuccess) bytes = cap;
    g_allocCount.fetch_add(1);
    g_allocNs.fetch_add(
        std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count());
    return e;
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
__global__ void k_fill_i32(int* p, int v, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_iota(int* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}
template <typename T>
__global__ void k_fill_sent(T* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Sentinel<T>::storeSent(&p[i]);
}
__global__ void k_d2f(const double* a, float* b, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void k_set_err(int* c, int bit) {
    if (threadIdx.x == 0) atomicOr(&c[C_ERR], bit);
}
__global__ void k_zero_counters(int* c) {
    if (threadIdx.x < C_COUNT) c[threadIdx.x] = 0;
}
__global__ void k_zero64(int n, double* x) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x[i] = 0.0;
}
__global__ void k_fill64(int n, double* x, double v) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x[i] = v;
}
// Plain streaming copy (calibration of the HBM counters: 16 bytes per element).
__global__ void __launch_bounds__(BS) k_copy64(int n, const double* __restrict__ a, double* __restrict__ b) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// ---------------------------------------------------------------------------
// dependency levels of tril(A): level(i) = 1 + max level(k), k in lower(i).
// Setup only. One lane per row, blocks of BS consecutive rows handed out in
// index order through a ticket: every row a lane waits on has a smaller index,
// so it belongs to this wave, an earlier wave of this block or an earlier,
// already running block. Each lane keeps its row's lower columns in registers
// (at most kMaxDeps, checked by k_split_rows) and polls the ones not seen yet
// in one round (all loads issued before any is examined); the loop is
// wave-uniform and a lane publishes in the round its last dependency appears,
// so lanes of one wave that depend on each other finish in turn. A row that
// waits past the spin limit raises E_LEVELS and publishes what it has.
// (One wavefront per row kept only ~8 k rows in flight; rows depend on their
// predecessors along a grid line, so that form crawled: 276 ms at 16 M rows.)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BS) k_levels(int n, const int* rs, const int* lend, const int* col, int* lev,
                                               int* cnt) {
    if (cnt[C_PATTERN]) return;  // invalid pattern (k_split_rows ran earlier in the stream)
    __shared__ int chunk;
    if (threadIdx.x == 0) chunk = atomicAdd(&cnt[C_TICKET_L], 1);
    __syncthreads();
    const int i = chunk * BS + (int)threadIdx.x;
    bool pending = i < n;
    int nd = 0;
    int c[kMaxDeps];
    if (pending) {
        const int p0 = rs[i];
        nd = lend[i] - p0;
#pragma unroll
        for (int j = 0; j < kMaxDeps; ++j) c[j] = j < nd ? col[p0 + j] : 0;
    }
    unsigned seen = 0;
    const unsigned full = nd >= 32 ? ~0u : (1u << nd) - 1u;
    int L = 0, spins = 0;
    for (;;) {
        if (pending) {
            int v[kMaxDeps];
#pragma unroll
            for (int j = 0; j < kMaxDeps; ++j)
                v[j] = (j < nd && !((seen >> j) & 1u)) ? ld_agent_i32(&lev[c[j]]) : kSentLevel;
#pragma unroll
            for (int j = 0; j < kMaxDeps; ++j)
                if (v[j] >= 0) {
                    L = max(L, v[j] + 1);
                    seen |= 1u << j;
                }
            if (seen == full) {
                st_agent_i32(&lev[i], L);
                pending = false;
            } else if (++spins > kSpinLimit) {
                atomicOr(&cnt[C_ERR], E_LEVELS);
                st_agent_i32(&lev[i], L);
                pending = false;
            }
        }
        if (!__any(pending)) break;
        __builtin_amdgcn_s_sleep(1);
    }
}

// Split of row i at the diagonal (lend, ubeg), transpose map (tpos: position
// of (j,i) for every off-diagonal (i,j)) and structural checks: sorted rows,
// columns in range, symmetry, at most kMaxDeps entries on each side of the
// diagonal. A failing row sets C_PATTERN. One lane per row.
__global__ void __launch_bounds__(BS) k_split_rows(int n, const int* __restrict__ rs, const int* __restrict__ col,
                                                   int* __restrict__ lend, int* __restrict__ ubeg,
                                                   int* __restrict__ tpos, int* cnt) {
    const int i = (int)(blockIdx.x * (long)blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int p0 = rs[i], e = rs[i + 1];
    bool bad = false;
    for (int q = p0 + 1; q < e; ++q)
        if (col[q] <= col[q - 1]) bad = true;
    int p = p0;
    while (p < e && col[p] < i) ++p;
    const int ub = (p < e && col[p] == i) ? p + 1 : p;
    lend[i] = p;
    ubeg[i] = ub;
    for (int q = p0; q < e && !bad; ++q) {
        const int j = col[q];
        if (j < 0 || j >= n) {
            bad = true;
            break;
        }
        if (j == i) continue;
        int lo = rs[j], hi = rs[j + 1];
        const int end = hi;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (col[mid] < i)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo == end || col[lo] != i) {
            bad = true;
            break;
        }
        tpos[q] = lo;
    }
    if (p - p0 > kMaxDeps || e - ub > kMaxDeps) bad = true;
    if (bad) atomicOr(&cnt[C_PATTERN], 1);
}

// One lane of a wavefront polls the published value of a gate row (advisory
// throttle; bounded).
template <typename T>
__device__ inline void gate_row(const T* vec, int target, int cap) {
    if (target < 0 || (threadIdx.x & 63) != 0) return;
    int backoff = 1;
    for (int s = 0; s < kGateLimit; ++s) {
        if (Sentinel<T>::load(&vec[target]) != Sentinel<T>::value) return;
        for (int k = 0; k < backoff; ++k) __builtin_amdgcn_s_sleep(2);
        backoff = backoff < cap ? backoff * 2 : cap;
    }
}

// ---------------------------------------------------------------------------
// MIC(0) factorisation, level-synchronous: one launch per dependency level.
// Row k computes its pivot and column k of L (stored at the upper positions of
// row k and, through tpos, at the matching lower positions of the rows below);
// everything it reads was written by rows of lower levels, i.e. by earlier
// launches, so plain loads suffice. A row takes 32 lanes: lane j < 16 gathers
// the pivot term of lower entry j, lane 16 + j computes upper entry j, each
// from its own operand list (mic_plan.h), so all gathers are independent.
// Accumulation order follows the reference's right-looking loop
// (pcg_solver.h:129-189): contributions to row/column k arrive in ascending m.
// ---------------------------------------------------------------------------
constexpr int kMicLanes = 32;

// Rows whose column is eliminated: a stored, nonzero diagonal.
__global__ void k_mic_live(int n, const int* rs, const int* lend, const int* col, const double* A, uint8_t* live) {
    const int m = blockIdx.x * BS + threadIdx.x;
    if (m >= n) return;
    const int le = lend[m];
    live[m] = (le < rs[m + 1] && col[le] == m && A[le] != 0.0) ? 1 : 0;
}

// Rows at level-order positions [t0, t1): one dependency level.
__global__ void __launch_bounds__(BS) k_mic_level(int t0, int t1, const int* order, const long long* pst,
                                                  const int* prog, const int* rs, const int* lend, const int* ubeg,
                                                  const int* col, const double* A, const uint8_t* live,
                                                  const int* tpos, double* F, double* invd, double tau, double ratio) {
    const int sub = threadIdx.x & (kMicLanes - 1);
    const int t = t0 + (blockIdx.x * BS + threadIdx.x) / kMicLanes;
    const int k = t < t1 ? order[t] : -1;  // lanes of absent rows still join the shuffles
    int r0 = 0, le = 0, ub = 0, r1 = 0;
    if (k >= 0) {
        r0 = rs[k];
        le = lend[k];
        ub = ubeg[k];
        r1 = rs[k + 1];
    }
    const int nl = le - r0;
    // pivot term of lower entry sub: L(k,m) and column m's entries missing from row k
    double lkm = 0.0, missing = 0.0;
    int lv = 0;
    if (sub < nl) {
        const int a = r0 + sub;
        lv = live[col[a]];
        lkm = F[a];
        for (long long u = pst[a], ue = pst[a + 1]; u < ue; ++u) missing += F[prog[u]];
    }
    const double akk = (k >= 0 && le < r1 && col[le] == k) ? A[le] : 0.0;
    double d = akk;
    unsigned liveMask = 0;  // lower entries whose column is eliminated (null columns never are)
#pragma unroll
    for (int j = 0; j < kMaxDeps; ++j) {
        const double lj = __shfl(lkm, j, kMicLanes), mj = __shfl(missing, j, kMicLanes);
        const int vj = __shfl(lv, j, kMicLanes);
        if (j < nl && vj) {
            liveMask |= 1u << j;
            d -= lj * lj;
            d -= tau * lj * mj;
        }
    }
    double inv;
    if (akk == 0.0) {
        inv = d;  // null row: left as accumulated, column not scaled
    } else if (d < ratio * akk) {
        inv = 1.0 / sqrt(akk);
    } else {
        inv = 1.0 / sqrt(d);
    }
    // entry (i,k) of column k, i = upper entry sub - 16 of row k
    const int p = ub + sub - kMaxDeps;
    if (k >= 0 && sub >= kMaxDeps && p < r1) {
        double v = A[p];
        for (long long u = pst[p], ue = pst[p + 1]; u < ue; u += 2) {
            const int j = prog[u], c = prog[u + 1];
            if ((liveMask >> j) & 1u) v -= F[r0 + j] * F[c];
        }
        if (akk != 0.0) v *= inv;
        F[p] = v;
        F[tpos[p]] = v;
    }
    if (k >= 0 && sub == 0) invd[k] = inv;
}

// ---------------------------------------------------------------------------
// PCG kernels (T = double or float; scalars are always double). The loop is
//   q = A s, sq = s.q                                   k_spmv_dot
//   alpha = rho/sq; x += alpha s; r -= alpha q; |r|      k_axpy_res
//   convergence test; y = L^-1 r                         k_fwd
//   z = L^-T y, rho' = z.r                               k_bwd_dot
//   s = z + (rho'/rho) s                                 k_xpby
// with the reference's order of operations (pcg_solver.h:396-421). rho of
// iteration it lives in the partials buffer rho[it & 1].
// ---------------------------------------------------------------------------
__device__ inline const double* rho_buf(const Parts& P, int parity) { return parity ? P.rho1 : P.rho0; }

// X: type of the solution vector (FP64 also when the iteration runs in FP32).
template <typename T, typename X>
__global__ void __launch_bounds__(BS) k_init(int n, const T* b, X* x, T* r, Parts Pt, int norm) {
    double v = 0.0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const T bi = b[i];
        x[i] = (X)0;
        r[i] = bi;
        const double a = (double)bi;
        v = norm == ASTG_NORM_INF ? nan_max(v, fabs(a)) : v + a * a;
    }
    const double t = norm == ASTG_NORM_INF ? block_reduce<true>(v) : block_reduce<false>(v);
    if (threadIdx.x == 0) Pt.res[blockIdx.x] = t;
}

// Forward sweep y = L^-1 r. mode 0 (start of the solve): ||b|| from the
// partials, tolerance, zero-rhs exit. mode 1 (iteration `it`): the reference's
// convergence test on the updated residual, before preconditioning. mode 2:
// no test here (the batched solve tests each system in its own kernel).
// Returns true when the solve stops here (every block takes the same decision).
__device__ inline bool fwd_prologue(const Parts& Pt, double* scal, int mode, int it, int maxIt, int norm,
                                    double tolRel) {
    double red = norm == ASTG_NORM_INF ? reduce_prev<true>(Pt.res, Pt.gs) : reduce_prev<false>(Pt.res, Pt.gs);
    if (norm != ASTG_NORM_INF) red = sqrt(red);
    if (!(red <= DBL_MAX)) {  // a NaN or infinity in b or in the residual: the solve stops (state 4)
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scal[S_RES] = red;
            scal[S_ITER] = mode == 0 ? 0.0 : (double)(it + 1);
            scal[S_DONE] = 4.0;
        }
        return true;
    }
    if (mode == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scal[S_BNORM] = red;
            scal[S_RES] = red;
            scal[S_TOL] = tolRel * red;
            scal[S_ITER] = 0;
            if (red == 0.0) scal[S_DONE] = 1.0;
        }
        return red == 0.0;
    }
    const bool conv = red <= scal[S_TOL];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[S_RES] = red;
        scal[S_ITER] = (double)(it + 1);
        if (conv) scal[S_DONE] = 1.0;
        else if (it + 1 >= maxIt) scal[S_DONE] = 3.0;
    }
    return conv || it + 1 >= maxIt;
}

// Batched form: a row of a system that has finished takes no part in a sweep.
__device__ inline bool row_finished(const SweepPlan& S, int i) {
    return S.chunkSys && S.sysState[S.chunkSys[i / BS] * B_COUNT + B_DONE] != 0.0;
}

// Forward sweep y = L^-1 r over the merged stencil:
//   y_i = d_i (r_i - sum_j R_ij r_j - sum_m W_im y_m)
// (with no merging R is empty and W is row i of L, in the reference's order).
// The r-terms need no other row, so they are summed before the row waits.
template <typename T>
__global__ void __launch_bounds__(BS) k_fwd(SweepPlan S, StencilView<T> St, const T* invd, const T* r, T* y, T* zp,
                                            Parts Pt, int* cnt, double* scal, int mode, int it, int maxIt, int norm,
                                            double tolRel) {
    // a sweep of this solve already failed: the host retries the whole solve
    if (scal[S_DONE] != 0.0 || ld_agent_i32(&cnt[C_ERR]) != 0) return;
    if (mode != 2 && fwd_prologue(Pt, scal, mode, it, maxIt, norm, tolRel)) return;
    for (int c = blockIdx.x; c < S.nchunks; c += gridDim.x) {
        const int t = c * BS + threadIdx.x;
        const int i0 = S.order[t];
        const bool fin = i0 >= 0 && row_finished(S, i0);
        const int i = fin ? -1 : i0;  // live row, or -1
        // row data that does not depend on other rows is loaded before the gate
        DepRow<T> d;
        T acc = 0, inv = 0;
        if (i >= 0) {
            const int b = St.start[t], e = St.start[t + 1], br = b + St.nr[t];
            const int vi = S.vecPos ? t : i;  // vectors in sweep or row order
            acc = r[vi];
            inv = invd[vi];
            for (int p = b; p < br; ++p) acc -= St.coef[p] * r[St.idx[p]];
            d.preload(St.idx, St.coef, br, e);
        }
        gate_row(y, S.gateF[t >> 6], S.gateCap);
        const long long tg = S.trace ? wall_clock64() : 0;
        if (i >= 0) {
            const bool ok = d.run(y, acc, S.pollCap, &cnt[C_ERR], [&](T a) {
                const long long td = S.trace ? wall_clock64() : 0;
                Sentinel<T>::store(&y[t], a * inv);
                if (S.trace) {
                    S.trace[3 * (long)i] = tg;
                    S.trace[3 * (long)i + 1] = td;
                    S.trace[3 * (long)i + 2] = wall_clock64();
                }
            });
            if (!ok) atomicOr(&cnt[C_ERR], E_FWD);
            // armed for the backward sweep; after publishing, off the dependency path
            Sentinel<T>::storeSent(&zp[t]);
        } else if (fin) {
            // nobody reads a finished system's rows; published so that gates
            // placed on them pass
            Sentinel<T>::store(&y[t], (T)0);
            Sentinel<T>::storeSent(&zp[t]);
        }
    }
}

// Backward sweep z = L^-T y in reverse sweep order over the merged stencil
//   z_k = d_k (y_k - sum_j R_kj y_j - sum_m W_km z_m),
// fused with the partial dot z.r (one partial per block, fixed chunk
// assignment: deterministic). y is complete (previous kernel); the other y
// buffer is re-armed here for the next forward sweep.
template <typename T>
__global__ void __launch_bounds__(BS) k_bwd_dot(SweepPlan S, StencilView<T> St, const T* invd, const T* y, T* ynext,
                                                const T* r, T* z, T* zp, Parts Pt, int* cnt, const double* scal,
                                                int parityOut) {
    if (scal[S_DONE] != 0.0 || ld_agent_i32(&cnt[C_ERR]) != 0) return;
    double v = 0.0;
    for (int c = blockIdx.x; c < S.nchunks; c += gridDim.x) {
        const int pos = S.P - 1 - (c * BS + threadIdx.x);
        const int k0 = S.order[pos];
        const bool fin = k0 >= 0 && row_finished(S, k0);
        const int k = fin ? -1 : k0;  // live row, or -1
        DepRow<T> d;
        T acc = 0, inv = 0, rk = 0;
        if (k >= 0) {
            const int b = St.start[pos], e = St.start[pos + 1], br = b + St.nr[pos];
            acc = y[pos];
            for (int p = b; p < br; ++p) acc -= St.coef[p] * y[St.idx[p]];
            inv = invd[S.vecPos ? pos : k];
            d.preload(St.idx, St.coef, br, e);
        }
        gate_row(zp, S.gateB[pos >> 6], S.gateCap);
        const long long tg = S.trace ? wall_clock64() : 0;
        if (k >= 0) {
            T zk = 0;
            const bool ok = d.run(zp, acc, S.pollCap, &cnt[C_ERR], [&](T a) {
                const long long td = S.trace ? wall_clock64() : 0;
                zk = a * inv;
                Sentinel<T>::store(&zp[pos], zk);
                // in sweep order the PCG reads z from zp itself (same position)
                if (!S.vecPos) z[k] = zk;
                if (S.trace) {
                    S.trace[3 * (long)k] = tg;
                    S.trace[3 * (long)k + 1] = td;
                    S.trace[3 * (long)k + 2] = wall_clock64();
                }
            });
            if (!ok) atomicOr(&cnt[C_ERR], E_BWD);
            // after publishing, off the dependency path: the dot with r and the
            // re-arming of the next forward sweep's vector
            rk = r[S.vecPos ? pos : k];
            Sentinel<T>::storeSent(&ynext[pos]);
            v += (double)zk * (double)rk;
        } else if (fin) {
            Sentinel<T>::store(&zp[pos], (T)0);
            Sentinel<T>::storeSent(&ynext[pos]);
        } else if (S.vecPos) {
            Sentinel<T>::store(&zp[pos], (T)0);  // padding: z (read from zp) stays zero there
        }
    }
    const double t = block_reduce<false>(v);
    if (threadIdx.x == 0) (parityOut ? Pt.rho1 : Pt.rho0)[blockIdx.x] = t;
}

// Coefficients of the merged stencils from the current factor: replays the
// row's slot program (sweep_merge.h). Slots start at zero, receive the row's
// own factor entries, then each substituted key j in order turns its slot into
// the r-coefficient a = W_j d_j and subtracts a * (row j of the factor) from
// the slots of j's dependencies. DIR 0: lower rows (forward), 1: upper.
template <int DIR>
__global__ void __launch_bounds__(BS) k_merge(int P, const int* order, const int* rs, const int* lend,
                                              const int* ubeg, const double* F, const double* invd, const int* sb,
                                              const int* nr, const int* idx, const int* us, const uint8_t* upd,
                                              double* C, int fwdPos) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= P) return;
    const int i = order[t];
    if (i < 0) return;
    const int b = sb[t], e = sb[t + 1], nR = nr[t];
    for (int s = b; s < e; ++s) C[s] = 0.0;
    int u = us[t];
    const int p0 = DIR == 0 ? rs[i] : ubeg[i], p1 = DIR == 0 ? lend[i] : rs[i + 1];
    for (int p = p0; p < p1; ++p) C[b + upd[u++]] += F[p];
    for (int x = 0; x < nR; ++x) {
        const int j = (DIR == 1 || fwdPos) ? order[idx[b + x]] : idx[b + x];
        const double a = C[b + x] * invd[j];
        C[b + x] = a;
        const int q0 = DIR == 0 ? rs[j] : ubeg[j], q1 = DIR == 0 ? lend[j] : rs[j + 1];
        for (int p = q0; p < q1; ++p) C[b + upd[u++]] -= a * F[p];
    }
}

// ---------------------------------------------------------------------------
// Sweep-ordered solves: the PCG vectors are stored by sweep position, so the
// sweeps read r and the pivots and write z contiguously. Padding positions
// are empty rows whose vector entries stay zero.
// ---------------------------------------------------------------------------
__global__ void k_row_pos(int P, const int* __restrict__ order, int* __restrict__ pos) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= P) return;
    const int i = order[t];
    if (i >= 0) pos[i] = t;
}

// Forward r-term keys: rows -> sweep positions.
__global__ void k_rterm_pos(int P, const int* __restrict__ start, const int* __restrict__ nr, int* __restrict__ idx,
                            const int* __restrict__ pos) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= P) return;
    for (int x = start[t], e = start[t] + nr[t]; x < e; ++x) idx[x] = pos[idx[x]];
}

__global__ void k_perm_len(int P, const int* __restrict__ order, const int* __restrict__ rs, int* __restrict__ len) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= P) return;
    const int k = order[t];
    len[t] = k >= 0 ? rs[k + 1] - rs[k] : 0;
}

// Row t of the sweep-ordered CSR: row order[t]'s entries in their stored
// order (so every row sum keeps its order), columns as sweep positions, and
// the source entry of each (for the values of every new assembly).
__global__ void k_perm_fill(int P, const int* __restrict__ order, const int* __restrict__ rs,
                            const int* __restrict__ col, const int* __restrict__ pos, const int* __restrict__ prs,
                            int* __restrict__ pcol, int* __restrict__ pmap) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= P) return;
    const int k = order[t];
    if (k < 0) return;
    int w = prs[t];
    for (int p = rs[k], e = rs[k + 1]; p < e; ++p, ++w) {
        pcol[w] = pos[col[p]];
        pmap[w] = p;
    }
}

__global__ void k_gather_vals(long nnz, const int* __restrict__ pmap, const double* __restrict__ val,
                              double* __restrict__ out) {
    for (long e = (long)blockIdx.x * BS + threadIdx.x; e < nnz; e += (long)gridDim.x * BS) out[e] = val[pmap[e]];
}

// out[t] = in[order[t]], zero at padding positions.
__global__ void k_to_sweep(int P, const int* __restrict__ order, const double* __restrict__ in,
                           double* __restrict__ out) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= P) return;
    const int k = order[t];
    out[t] = k >= 0 ? in[k] : 0.0;
}

// out[order[t]] = in[t].
__global__ void k_from_sweep(int P, const int* __restrict__ order, const double* __restrict__ in,
                             double* __restrict__ out) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= P) return;
    const int k = order[t];
    if (k >= 0) out[k] = in[t];
}

// s = z + beta s with beta = rho(it+1) / rho(it); mode 0: s = z after the first
// preconditioning, with the reference's rho == 0 / NaN breakdown test.
template <typename T>
__global__ void __launch_bounds__(BS) k_xpby(int n, const T* z, T* s, Parts Pt, double* scal, int mode, int it) {
    if (scal[S_DONE] != 0.0) return;
    if (mode == 0) {
        const double rho = reduce_prev<false>(Pt.rho0, Pt.gb);
        if (rho == 0.0 || rho != rho) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                scal[S_DONE] = 2.0;
                scal[S_ITER] = 0;
            }
            return;
        }
        for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) s[i] = z[i];
        return;
    }
    const double rnew = reduce_prev<false>(rho_buf(Pt, (it + 1) & 1), Pt.gb);
    const double rold = reduce_prev<false>(rho_buf(Pt, it & 1), Pt.gb);
    const T beta = (T)(rnew / rold);
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) s[i] = z[i] + beta * s[i];
}

// q = A s (row_dot, LPR lanes per row), fused with s.q.
template <typename T, int LPR = kLpr>
__global__ void __launch_bounds__(BS) k_spmv_dot(int n, const int* rs, const int* col, const T* A, const T* s, T* q,
                                                 Parts Pt, const double* scal) {
    if (scal[S_DONE] != 0.0) return;
    const int sub = threadIdx.x & (LPR - 1);
    double v = 0.0;
    const int rowsPerIter = gridDim.x * (BS / LPR);
    for (int base = blockIdx.x * (BS / LPR); base < n; base += rowsPerIter) {
        const int ii = base + threadIdx.x / LPR;
        const int i = ii < n ? ii : -1;
        const T acc = row_dot<T, LPR>(rs, col, A, s, i, sub);
        if (i >= 0 && sub == 0) {
            q[i] = acc;
            v += (double)s[i] * (double)acc;
        }
    }
    const double t = block_reduce<false>(v);
    if (threadIdx.x == 0) Pt.sq[blockIdx.x] = t;
}

template <typename T, int LPR, int U>
__global__ void __launch_bounds__(BS) k_spmv_u_dot(int n, const int* rs, const int* col, const T* A, const T* s,
                                                   T* q, Parts Pt, const double* scal) {
    if (scal[S_DONE] != 0.0) return;
    const int sub = threadIdx.x & (LPR - 1);
    double v = 0.0;
    const int rowsPerIter = gridDim.x * (BS / LPR);
    for (int base = blockIdx.x * (BS / LPR); base < n; base += U * rowsPerIter) {
        int i[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int ii = base + threadIdx.x / LPR + u * rowsPerIter;
            i[u] = ii < n ? ii : -1;
        }
        T acc[U];
        row_dot_u<T, LPR, U>(rs, col, A, s, i, sub, acc);
        if (sub == 0)
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (i[u] >= 0) {
                    q[i[u]] = acc[u];
                    v += (double)s[i[u]] * (double)acc[u];
                }
    }
    const double t = block_reduce<false>(v);
    if (threadIdx.x == 0) Pt.sq[blockIdx.x] = t;
}

// x accumulates in X (FP64 in the mixed form: the FP32 iteration's steps are
// summed without FP32 rounding).
template <typename T, typename X>
__global__ void __launch_bounds__(BS) k_axpy_res(int n, const T* s, const T* q, X* x, T* r, Parts Pt,
                                                 const double* scal, int norm, int it) {
    if (scal[S_DONE] != 0.0) return;
    const double sq = reduce_prev<false>(Pt.sq, Pt.gs);
    const double rho = reduce_prev<false>(rho_buf(Pt, it & 1), Pt.gb);
    const double ad = rho / sq;
    const T alpha = (T)ad;
    const X ax = sizeof(X) > sizeof(T) ? (X)ad : (X)alpha;
    double v = 0.0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        x[i] += ax * (X)s[i];
        const T ri = r[i] - alpha * q[i];
        r[i] = ri;
        const double a = (double)ri;
        v = norm == ASTG_NORM_INF ? nan_max(v, fabs(a)) : v + a * a;
    }
    const double t = norm == ASTG_NORM_INF ? block_reduce<true>(v) : block_reduce<false>(v);
    if (threadIdx.x == 0) Pt.res[blockIdx.x] = t;
}

// Mixed form: r64 = b - A x (FP64), |r| partials.
__global__ void __launch_bounds__(BS) k_residual64(int n, const int* rs, const int* col, const double* A,
                                                   const double* x, const double* b, double* r, Parts Pt, int norm) {
    const int sub = threadIdx.x & (kLpr - 1);
    double v = 0.0;
    const int rowsPerIter = gridDim.x * (BS / kLpr);
    for (int base = blockIdx.x * (BS / kLpr); base < n; base += rowsPerIter) {
        const int ii = base + threadIdx.x / kLpr;
        const int i = ii < n ? ii : -1;
        const double acc = row_dot(rs, col, A, x, i, sub);
        if (i >= 0 && sub == 0) {
            const double ri = b[i] - acc;
            r[i] = ri;
            v = norm == ASTG_NORM_INF ? nan_max(v, fabs(ri)) : v + ri * ri;
        }
    }
    const double t = norm == ASTG_NORM_INF ? block_reduce<true>(v) : block_reduce<false>(v);
    if (threadIdx.x == 0) Pt.res[blockIdx.x] = t;
}
__global__ void k_res_final(Parts Pt, double* scal, int norm) {
    const double t = norm == ASTG_NORM_INF ? reduce_prev<true>(Pt.res, Pt.gs) : reduce_prev<false>(Pt.res, Pt.gs);
    if (threadIdx.x == 0) scal[S_OUTER_RES] = norm == ASTG_NORM_INF ? t : sqrt(t);
}
// Mixed form, inner right-hand side: r64 / |r64| in FP64 and FP32.
__global__ void k_scale_rhs(int n, const double* r, double inv, double* o64, float* o32) {
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const double v = r[i] * inv;
        o64[i] = v;
        o32[i] = (float)v;
    }
}
__global__ void k_add_corr(int n, double* x, const double* d, double scale) {
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) x[i] += scale * d[i];
}
// Mixed form, reliable update: the FP32 recurrence residual is replaced by the
// true residual b - A x of the inner system, computed in FP64 from the FP64
// iterate, with its norm partials for the convergence test that follows.
__global__ void __launch_bounds__(BS) k_replace_res(int n, const int* rs, const int* col, const double* A,
                                                    const double* x, const double* b, float* r, Parts Pt,
                                                    const double* scal, int norm) {
    if (scal[S_DONE] != 0.0) return;
    const int sub = threadIdx.x & (kLpr - 1);
    double v = 0.0;
    const int rowsPerIter = gridDim.x * (BS / kLpr);
    for (int base = blockIdx.x * (BS / kLpr); base < n; base += rowsPerIter) {
        const int ii = base + threadIdx.x / kLpr;
        const int i = ii < n ? ii : -1;
        const double acc = row_dot(rs, col, A, x, i, sub);
        if (i >= 0 && sub == 0) {
            const float ri = (float)(b[i] - acc);
            r[i] = ri;
            v = norm == ASTG_NORM_INF ? nan_max(v, fabs((double)ri)) : v + (double)ri * (double)ri;
        }
    }
    const double t = norm == ASTG_NORM_INF ? block_reduce<true>(v) : block_reduce<false>(v);
    if (threadIdx.x == 0) Pt.res[blockIdx.x] = t;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
PcgDevice::~PcgDevice() {
    if (hscal_) (void)hipHostFree(hscal_);
    if (hplan_) (void)hipHostFree(hplan_);
}

static inline int blocks(long n) { return (int)((n + BS - 1) / BS); }

template <typename K>
static int coopBlocks(K kernel, int cus, int want) {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, BS, 0) != hipSuccess || per < 1) per = 1;
    return std::max(1, std::min(want, per * cus));
}

// Persistent sweep launch. coop: the cooperative API, which guarantees that
// all blocks are co-resident; otherwise an ordinary launch of the same grid
// (at most the occupancy limit, so on an otherwise idle device every block is
// resident at once; if one were not, the sweeps' bounded spin ends the kernel
// with an error flag instead of waiting for it).
template <typename K, typename... Args>
static hipError_t launchCoop(bool coop, K kernel, int grid, hipStream_t s, Args... args) {
    if (!coop) {
        kernel<<<grid, BS, 0, s>>>(args...);
        return hipGetLastError();
    }
    void* ptrs[] = {(void*)&args...};
    return hipLaunchCooperativeKernel((const void*)kernel, dim3(grid), dim3(BS), ptrs, 0, s);
}

// Split at the diagonal and structural checks on the host threads: sorted
// rows, columns in range, structural symmetry, at most kMaxDeps entries on
// each side of the diagonal. The first failing row is reported. Serves the
// host plans (which need lend / ubeg) and words the error when the device
// check (k_split_rows) fails.
static int splitPatternHost(int n, const int* rowstart, const int* col, std::vector<int>& lend,
                            std::vector<int>& ubeg, std::string& err) {
    lend.assign(n, 0);
    ubeg.assign(n, 0);
    int badRow = n, badKind = 0, badCol = 0;
    std::mutex mu;
    parallelFor(n, [&](long i0, long i1) {
        int row = n, kind = 0, cb = 0;
        for (int i = (int)i0; i < (int)i1 && row == n; ++i) {
            int p = rowstart[i];
            const int e = rowstart[i + 1];
            for (int q = p + 1; q < e; ++q)
                if (col[q] <= col[q - 1]) {
                    row = i;
                    kind = 1;
                    break;
                }
            if (row != n) break;
            while (p < e && col[p] < i) ++p;
            lend[i] = p;
            ubeg[i] = (p < e && col[p] == i) ? p + 1 : p;
            for (int q = rowstart[i]; q < e; ++q) {
                const int j = col[q];
                if (j < 0 || j >= n) {
                    row = i;
                    kind = 2;
                    break;
                }
                if (j == i) continue;
                const int* b = col + rowstart[j];
                const int* f = std::lower_bound(b, col + rowstart[j + 1], i);
                if (f == col + rowstart[j + 1] || *f != i) {
                    row = i;
                    kind = 3;
                    cb = j;
                    break;
                }
            }
            if (row == n && (lend[i] - rowstart[i] > kMaxDeps || rowstart[i + 1] - ubeg[i] > kMaxDeps)) {
                row = i;
                kind = 4;
            }
        }
        if (row != n) {
            std::lock_guard<std::mutex> g(mu);
            if (row < badRow) {
                badRow = row;
                badKind = kind;
                badCol = cb;
            }
        }
    });
    if (badRow < n) {
        const int i = badRow;
        switch (badKind) {
            case 1: err = "row " + std::to_string(i) + " is not sorted by column"; break;
            case 2: err = "column index out of range"; break;
            case 3:
                err = "matrix is not structurally symmetric at (" + std::to_string(i) + "," +
                      std::to_string(badCol) + ")";
                break;
            default:
                err = "row " + std::to_string(i) + " has more than " + std::to_string(kMaxDeps) +
                      " entries on one side of the diagonal (unsupported)";
        }
        return ASTG_ERR_ARG;
    }
    return ASTG_OK;
}

// Host construction of a padded level order (the device builds the same with
// k_level_starts / k_pad_order; the host planners check one against the
// other): groups of q consecutive levels, each padded to a multiple of 64
// rows, the whole to a multiple of BS; gate rows per 64-row segment from
// groups G-dist and G+dist.
static void hostOrder(int q, int dist, int n, int depth, const std::vector<int>& sortedLev,
                      const std::vector<int>& sortedRow, std::vector<int>& padded, std::vector<int>& gF,
                      std::vector<int>& gB) {
    const int ng = (depth + q - 1) / q;
    std::vector<int> gStart(ng, 0), gRows(ng, 0);
    padded.clear();
    padded.reserve((size_t)n + (size_t)ng * 64 + BS);
    for (int i = 0; i < n;) {
        int j = i;
        const int G = sortedLev[i] / q;
        gStart[G] = (int)padded.size();
        while (j < n && sortedLev[j] / q == G) padded.push_back(sortedRow[j++]);
        gRows[G] = j - i;
        while (padded.size() % 64) padded.push_back(-1);
        i = j;
    }
    while (padded.size() % BS) padded.push_back(-1);
    const int nseg = (int)padded.size() / 64;
    gF.assign(nseg, -1);
    gB.assign(nseg, -1);
    std::vector<int> segG(nseg, -1);
    for (int G = 0; G < ng; ++G) {
        const int s0 = gStart[G] / 64, s1 = (gStart[G] + gRows[G] + 63) / 64;
        for (int sg = s0; sg < s1; ++sg) segG[sg] = G;
    }
    for (int sg = 0; sg < nseg; ++sg) {
        const int G = segG[sg];
        if (G < 0) continue;
        const unsigned h = (unsigned)sg * 2654435761u;
        const int gf = G - dist, gb = G + dist;
        if (gf >= 0) gF[sg] = gStart[gf] + (int)(h % (unsigned)gRows[gf]);
        if (gb < ng) gB[sg] = gStart[gb] + (int)(h % (unsigned)gRows[gb]);
    }
}

// First sorted position of every level (rows sorted by level; every level
// from 0 to the last holds a row), and n after the last.
__global__ void k_level_starts(int n, const int* __restrict__ key, int* __restrict__ lvlStart) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    if (t == 0 || key[t] != key[t - 1]) lvlStart[key[t]] = t;
    if (t == n - 1) lvlStart[key[t] + 1] = n;
}
// Rows in level order into their group's padded slots: the row at sorted
// position t goes to gStart[G] + (t - gFirst[G]), G = level / q.
__global__ void k_pad_order(int n, const int* __restrict__ key, const int* __restrict__ row, int q,
                            const int* __restrict__ gStart, const int* __restrict__ gFirst, int* __restrict__ out) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const int G = key[t] / q;
    out[gStart[G] + (t - gFirst[G])] = row[t];
}

int PcgDevice::setPattern(int n, const int* rowstart, const int* col, std::string& err) {
    if (n <= 0) {
        err = "empty system";
        return ASTG_ERR_ARG;
    }
    PhaseTimer tm("setPattern");
    // host copies of the pattern (host plans, factor download, checks), by
    // the worker pool: at 16 M rows the column array alone is 0.5 GB
    const long long nnz = rowstart[n];
    hrs_.resize((size_t)n + 1);
    hcol_.resize((size_t)nnz);
    parallelFor((long)n + 1, [&](long a, long b) { std::memcpy(hrs_.data() + a, rowstart + a, (b - a) * sizeof(int)); },
                1 << 20);
    parallelFor((long)nnz, [&](long a, long b) { std::memcpy(hcol_.data() + a, col + a, (b - a) * sizeof(int)); },
                1 << 20);
    hostPatternValid_ = true;
    tm.mark("host copy of the pattern");
    return setPatternImpl(n, nnz, nullptr, nullptr, tm, err);
}

int PcgDevice::setPatternDevice(int n, long long nnz, const int* dRowstart, const int* dCol, std::string& err) {
    if (n <= 0) {
        err = "empty system";
        return ASTG_ERR_ARG;
    }
    PhaseTimer tm("setPattern");
    hostPatternValid_ = false;
    return setPatternImpl(n, nnz, dRowstart, dCol, tm, err);
}

int PcgDevice::ensureHostPattern(std::string& err) {
    if (hostPatternValid_) return ASTG_OK;
    hrs_.resize((size_t)n_ + 1);
    hcol_.resize((size_t)nnz_);
    HIPCHK(hipMemcpyAsync(hrs_.data(), rs_.p, ((size_t)n_ + 1) * 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(hcol_.data(), col_.p, (size_t)nnz_ * 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    hostPatternValid_ = true;
    return ASTG_OK;
}

// The pattern is in hrs_ / hcol_ (dRowstart null) or in device memory.
int PcgDevice::setPatternImpl(int n, long long nnz, const int* dRowstart, const int* dCol, PhaseTimer& tm,
                              std::string& err) {
    f32Valid_ = false;
    n_ = n;
    nnz_ = nnz;
    if (const char* g = getenv("ASTG_GATE_DIST")) gateDist_ = std::max(1, atoi(g));
    if (const char* g = getenv("ASTG_POLL_CAP")) pollCap_ = std::max(1, atoi(g));
    if (const char* g = getenv("ASTG_GATE_CAP")) gateCap_ = std::max(1, atoi(g));
    // sweep merging: levels per group (1 = plain level order)
    merge_ = 2;
    if (const char* g = getenv("ASTG_MERGE")) merge_ = std::max(1, std::min(8, atoi(g)));
    if (mergeLevels_ > 0) merge_ = std::min(8, mergeLevels_);
    if (const char* g = getenv("ASTG_BATCH_GRAPH")) batchGraph_ = atoi(g) != 0;
    if (const char* g = getenv("ASTG_COOP")) coop_ = atoi(g) != 0;
    if (const char* g = getenv("ASTG_TEST_SPIN_ERROR_ONCE")) forceSpinError_ = atoi(g) != 0;
    hostPlan_ = false;
    if (const char* g = getenv("ASTG_HOST_PLAN")) hostPlan_ = atoi(g) != 0;
    // 4 lanes per row: 16 rows per wavefront, 2 entries per lane on average
    spmvLpr_ = 4;
    if (const char* g = getenv("ASTG_SPMV_LPR")) spmvLpr_ = (atoi(g) == 8 || atoi(g) == 16) ? atoi(g) : 4;
    spmvRows_ = 2;
    if (const char* g = getenv("ASTG_SPMV_ROWS")) spmvRows_ = atoi(g) == 4 ? 4 : atoi(g) == 1 ? 1 : 2;
    // mixed form: residual replacement period and inner tolerance factor
    if (const char* g = getenv("ASTG_MIXED_REPLACE")) mixedReplace_ = std::max(0, atoi(g));
    if (const char* g = getenv("ASTG_MIXED_ETA")) mixedEta_ = std::max(0.01, atof(g));
    int dev = 0, cus = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    cus_ = cus;
    // streaming kernels: 8 blocks (32 wavefronts) per CU in a grid-stride loop;
    // this many partials is still cheap to reduce in the consumer kernel
    gs_ = std::max(1, std::min(8 * cus, blocks(n)));
    const size_t n8 = (size_t)n * 8, n4 = (size_t)n * 4;
    HIPCHK(rs_.alloc(((size_t)n + 1) * 4));
    HIPCHK(col_.alloc((size_t)nnz_ * 4));
    HIPCHK(lend_.alloc(n4));
    HIPCHK(ubeg_.alloc(n4));
    HIPCHK(tpos_.alloc((size_t)nnz_ * 4));
    HIPCHK(order_.alloc(n4));
    HIPCHK(lev_.alloc(n4));
    HIPCHK(val_.alloc((size_t)nnz_ * 8));
    HIPCHK(F_.alloc((size_t)nnz_ * 8));
    HIPCHK(invd_.alloc(n8));
    HIPCHK(b_.alloc(n8));
    HIPCHK(x_.alloc(n8));
    HIPCHK(scal_.alloc(32 * 8));
    HIPCHK(cnt_.alloc(C_COUNT * 4));
    if (!hscal_) HIPCHK(hipHostMalloc((void**)&hscal_, 32 * 8, hipHostMallocDefault));
    if (!hplan_) HIPCHK(hipHostMalloc((void**)&hplan_, 16 * 8, hipHostMallocDefault));
    if (dRowstart) {
        HIPCHK(hipMemcpyAsync(rs_.p, dRowstart, ((size_t)n + 1) * 4, hipMemcpyDeviceToDevice, stream_));
        HIPCHK(hipMemcpyAsync(col_.p, dCol, (size_t)nnz_ * 4, hipMemcpyDeviceToDevice, stream_));
    } else {
        HIPCHK(hipMemcpyAsync(rs_.p, hrs_.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice, stream_));
        HIPCHK(hipMemcpyAsync(col_.p, hcol_.data(), (size_t)nnz_ * 4, hipMemcpyHostToDevice, stream_));
    }
    k_zero_counters<<<1, 64, 0, stream_>>>(cnt_.as<int>());
    k_zero64<<<1, BS, 0, stream_>>>(32, scal_.as<double>());
    // split at the diagonal, transpose map and structural checks
    k_split_rows<<<blocks(n), BS, 0, stream_>>>(n, rs_.as<int>(), col_.as<int>(), lend_.as<int>(), ubeg_.as<int>(),
                                                tpos_.as<int>(), cnt_.as<int>());
    // dependency levels on the device (skipped if the check failed)
    k_fill_i32<<<blocks(n), BS, 0, stream_>>>(lev_.as<int>(), kSentLevel, n);
    k_levels<<<blocks(n), BS, 0, stream_>>>(n, rs_.as<int>(), lend_.as<int>(), col_.as<int>(), lev_.as<int>(),
                                            cnt_.as<int>());
    HIPCHK(hipGetLastError());
    // stable sort of (level, row): rows in level order, ties by index
    HIPCHK(sortKeys_.alloc(n4));
    HIPCHK(sortIdx_.alloc(n4));
    HIPCHK(sortRows_.alloc(n4));
    k_iota<<<blocks(n), BS, 0, stream_>>>(sortIdx_.as<int>(), n);
    size_t tmpBytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmpBytes, lev_.as<int>(), sortKeys_.as<int>(),
                                              sortIdx_.as<int>(), sortRows_.as<int>(), n, 0, 31, stream_));
    HIPCHK(planTmp_.alloc(tmpBytes));
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(planTmp_.p, tmpBytes, lev_.as<int>(), sortKeys_.as<int>(),
                                              sortIdx_.as<int>(), sortRows_.as<int>(), n, 0, 31, stream_));
    int* hcnt = reinterpret_cast<int*>(hplan_);
    HIPCHK(hipMemcpyAsync(hcnt, cnt_.p, C_COUNT * 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(hcnt + C_COUNT, sortKeys_.as<int>() + (n - 1), 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    if (hcnt[C_PATTERN]) {
        std::vector<int> lend, ubeg;
        if (int rc = ensureHostPattern(err)) return rc;
        if (splitPatternHost(n, hrs_.data(), hcol_.data(), lend, ubeg, err)) return ASTG_ERR_ARG;
        err = "the device pattern check failed where the host check passes";
        return ASTG_ERR_HIP;
    }
    if (hcnt[C_ERR]) {
        err = "level computation did not complete";
        return ASTG_ERR_HIP;
    }
    depth_ = hcnt[C_COUNT] + 1;
    tm.mark("split, transpose map, levels, sort");

    // Orders: rows in order of groups of q consecutive levels, every group
    // padded to a multiple of 64 rows so that a wavefront never holds rows of
    // two groups, the whole padded to a multiple of BS; gate rows: for each
    // 64-row segment of group G, one row of group G-dist (forward) and one of
    // group G+dist (backward), spread over that group. Every level 0..depth-1
    // holds at least one row, so every group is non-empty. The group layout
    // follows from the first sorted row of each level (lvlStart, depth+1
    // entries): computed on the host from that, rows scattered on the device.
    std::vector<int> lvlStart(depth_ + 1);
    struct Layout {
        int ng = 0, P = 0;
        std::vector<int> gStart, gFirst, gRows, gF, gB;
    };
    auto layout = [&](int q, int dist, bool gates, Layout& o) {
        o.ng = (depth_ + q - 1) / q;
        o.gStart.assign(o.ng, 0);
        o.gFirst.assign(o.ng, 0);
        o.gRows.assign(o.ng, 0);
        long long pos = 0;
        for (int G = 0; G < o.ng; ++G) {
            const int first = lvlStart[G * q], end = lvlStart[std::min((G + 1) * q, depth_)];
            o.gFirst[G] = first;
            o.gRows[G] = end - first;
            o.gStart[G] = (int)pos;
            pos = (pos + o.gRows[G] + 63) / 64 * 64;
        }
        pos = (pos + BS - 1) / BS * BS;
        o.P = (int)pos;
        if (!gates) return;
        const int nseg = o.P / 64;
        o.gF.assign(nseg, -1);
        o.gB.assign(nseg, -1);
        std::vector<int> segG(nseg, -1);
        for (int G = 0; G < o.ng; ++G) {
            const int s0 = o.gStart[G] / 64, s1 = (o.gStart[G] + o.gRows[G] + 63) / 64;
            for (int sg = s0; sg < s1; ++sg) segG[sg] = G;
        }
        for (int sg = 0; sg < nseg; ++sg) {
            const int G = segG[sg];
            if (G < 0) continue;
            const unsigned h = (unsigned)sg * 2654435761u;
            const int gf = G - dist, gb = G + dist;
            if (gf >= 0) o.gF[sg] = o.gStart[gf] + (int)(h % (unsigned)o.gRows[gf]);
            if (gb < o.ng) o.gB[sg] = o.gStart[gb] + (int)(h % (unsigned)o.gRows[gb]);
        }
    };
    auto upload = [&](DevBuf& d, const void* src, size_t bytes) -> hipError_t {
        hipError_t e = d.alloc(bytes);
        if (e != hipSuccess) return e;
        return hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, stream_);
    };
    HIPCHK(lvlStart_.alloc(((size_t)depth_ + 1) * 4));
    k_fill_i32<<<blocks(depth_ + 1), BS, 0, stream_>>>(lvlStart_.as<int>(), -1, depth_ + 1);
    k_level_starts<<<blocks(n), BS, 0, stream_>>>(n, sortKeys_.as<int>(), lvlStart_.as<int>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(lvlStart.data(), lvlStart_.p, ((size_t)depth_ + 1) * 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    // the layout below indexes by these: every level must start after the last
    bool starts = lvlStart[0] == 0 && lvlStart[depth_] == n;
    for (int L = 0; starts && L < depth_; ++L) starts = lvlStart[L] < lvlStart[L + 1];
    if (!starts) {
        err = "dependency levels are not contiguous";
        return ASTG_ERR_HIP;
    }
    // the host-planner path (ASTG_HOST_PLAN) also needs the sorted rows and
    // both orders on the host
    std::vector<int> sortedLev, sortedRow, spadded;
    if (hostPlan_) {
        sortedLev.resize(n);
        sortedRow.resize(n);
        HIPCHK(hipMemcpyAsync(sortedLev.data(), sortKeys_.p, n4, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipMemcpyAsync(sortedRow.data(), sortRows_.p, n4, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
    }
    // The host vectors uploaded below live until the synchronisation at the
    // end of this function.
    auto scatter = [&](const Layout& o, int q, DevBuf& ord) -> int {
        HIPCHK(upload(grpStart_, o.gStart.data(), (size_t)o.ng * 4));
        HIPCHK(upload(grpFirst_, o.gFirst.data(), (size_t)o.ng * 4));
        HIPCHK(ord.alloc((size_t)o.P * 4));
        k_fill_i32<<<blocks(o.P), BS, 0, stream_>>>(ord.as<int>(), -1, o.P);
        k_pad_order<<<blocks(n), BS, 0, stream_>>>(n, sortKeys_.as<int>(), sortRows_.as<int>(), q,
                                                  grpStart_.as<int>(), grpFirst_.as<int>(), ord.as<int>());
        HIPCHK(hipGetLastError());
        // the small group arrays are reused by the next call: finish this one
        HIPCHK(hipStreamSynchronize(stream_));
        return ASTG_OK;
    };
    // factorisation: rows in level order, one launch per level
    Layout lf, ls;
    layout(1, 1, false, lf);
    P_ = lf.P;
    if (int rc = scatter(lf, 1, order_)) return rc;
    levRange_.assign(2 * (size_t)depth_, 0);
    for (int L = 0; L < depth_; ++L) {
        levRange_[2 * L] = lf.gStart[L];
        levRange_[2 * L + 1] = lf.gStart[L] + lf.gRows[L];
    }
    // sweeps: q levels per group
    layout(merge_, gateDist_, true, ls);
    PS_ = ls.P;
    if (int rc = scatter(ls, merge_, sorder_)) return rc;
    HIPCHK(upload(sgateF_, ls.gF.data(), ls.gF.size() * 4));
    HIPCHK(upload(sgateB_, ls.gB.data(), ls.gB.size() * 4));
    if (hostPlan_) {
        // the host construction of both orders and the gates, from the sorted
        // rows, must give the device's entry for entry
        std::vector<int> padded, gF, gB, devOrd, devF, devB;
        hostOrder(1, 1, n, depth_, sortedLev, sortedRow, padded, gF, gB);
        devOrd.resize(P_);
        HIPCHK(hipMemcpyAsync(devOrd.data(), order_.p, (size_t)P_ * 4, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        if ((int)padded.size() != P_ || padded != devOrd) {
            err = "device-built factorisation order differs from the host construction";
            return ASTG_ERR_HIP;
        }
        hostOrder(merge_, gateDist_, n, depth_, sortedLev, sortedRow, spadded, gF, gB);
        devOrd.resize(PS_);
        devF.resize(ls.gF.size());
        devB.resize(ls.gB.size());
        HIPCHK(hipMemcpyAsync(devOrd.data(), sorder_.p, (size_t)PS_ * 4, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        if ((int)spadded.size() != PS_ || spadded != devOrd || gF != ls.gF || gB != ls.gB) {
            err = "device-built sweep order or gates differ from the host construction";
            return ASTG_ERR_HIP;
        }
    }
    tm.mark("level and sweep orders");
    // symbolic plans: factorisation programs and merged stencils
    const int maxY = kMaxDeps;
    if (!hostPlan_) {
        const int rc = planDevice(maxY, err);
        if (rc) return rc;
    } else {
        std::vector<int> lend, ubeg;
        if (int rc = ensureHostPattern(err)) return rc;
        const int* rowstart = hrs_.data();
        const int* col = hcol_.data();
        if (splitPatternHost(n, rowstart, col, lend, ubeg, err)) return ASTG_ERR_ARG;
        std::vector<long long> ms;
        std::vector<int> mp;
        planMic(n, rowstart, lend.data(), ubeg.data(), col, ms, mp);
        HIPCHK(upload(micStart_, ms.data(), ms.size() * 8));
        HIPCHK(upload(micProg_, mp.data(), mp.size() * 4));
        HIPCHK(hipStreamSynchronize(stream_));
        std::vector<int> lev(n), pos(n);
        for (int q = 0; q < n; ++q) lev[sortedRow[q]] = sortedLev[q];
        for (int q = 0; q < PS_; ++q)
            if (spadded[q] >= 0) pos[spadded[q]] = q;
        DevStencil* dst[2] = {&stF_, &stB_};
        for (int dir = 0; dir < 2; ++dir) {
            SweepStencil st;
            planStencil(dir, n, rowstart, lend.data(), ubeg.data(), col, lev.data(), merge_, maxY, spadded.data(),
                        PS_, pos.data(), st);
            DevStencil& D = *dst[dir];
            D.E = (long long)st.idx.size();
            HIPCHK(upload(D.start, st.start.data(), st.start.size() * 4));
            HIPCHK(upload(D.nr, st.nr.data(), st.nr.size() * 4));
            HIPCHK(upload(D.idx, st.idx.data(), st.idx.size() * 4));
            HIPCHK(upload(D.ustart, st.ustart.data(), st.ustart.size() * 4));
            HIPCHK(upload(D.upd, st.upd.data(), st.upd.size()));
            HIPCHK(D.coef.alloc((size_t)D.E * 8));
            // the host vectors must outlive the copies
            HIPCHK(hipStreamSynchronize(stream_));
        }
    }
    HIPCHK(live_.alloc((size_t)n));
    // PCG work vectors: sweep-ordered (PS entries) unless the row order is kept
    perm_ = !rowOrderForced_;
    if (const char* g = getenv("ASTG_ROW_ORDER")) perm_ = perm_ && atoi(g) == 0;
    if (perm_) {
        const int rc = permuteSetup(err);
        if (rc) return rc;
    }
    const size_t nv8 = (size_t)vlen() * 8;
    HIPCHK(r_.alloc(nv8));
    HIPCHK(z_.alloc(nv8));
    HIPCHK(s_.alloc(nv8));
    HIPCHK(q_.alloc(nv8));
    HIPCHK(y_.alloc((size_t)PS_ * 2 * 8));
    HIPCHK(zp_.alloc((size_t)PS_ * 8));
    tm.mark("symbolic plans");
    // persistent grids (all blocks co-resident)
    const int schunks = PS_ / BS;
    gF64_ = coopBlocks(k_fwd<double>, cus, schunks);
    gF32_ = coopBlocks(k_fwd<float>, cus, schunks);
    gB64_ = coopBlocks(k_bwd_dot<double>, cus, schunks);
    gB32_ = coopBlocks(k_bwd_dot<float>, cus, schunks);
    gb_ = std::max(std::max(gB64_, gB32_), 1);
    HIPCHK(part_.alloc((size_t)(2 * gs_ + 2 * gb_ + 8) * 8));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

template <typename T>
void PcgDevice::launchSpmv(int n, const int* rs, const int* col, const T* A, const T* s, T* q, const Parts& Pt,
                           const double* scal) {
    if (spmvLpr_ == 4 && spmvRows_ == 4)
        k_spmv_u_dot<T, 4, 4><<<gs_, BS, 0, stream_>>>(n, rs, col, A, s, q, Pt, scal);
    else if (spmvLpr_ == 4 && spmvRows_ == 2)
        k_spmv_u_dot<T, 4, 2><<<gs_, BS, 0, stream_>>>(n, rs, col, A, s, q, Pt, scal);
    else if (spmvLpr_ == 4)
        k_spmv_dot<T, 4><<<gs_, BS, 0, stream_>>>(n, rs, col, A, s, q, Pt, scal);
    else if (spmvLpr_ == 16)
        k_spmv_dot<T, 16><<<gs_, BS, 0, stream_>>>(n, rs, col, A, s, q, Pt, scal);
    else
        k_spmv_dot<T, 8><<<gs_, BS, 0, stream_>>>(n, rs, col, A, s, q, Pt, scal);
}

int PcgDevice::permuteSetup(std::string& err) {
    const int n = n_, P = PS_;
    const long long nnz = nnz_;
    HIPCHK(pos_.alloc((size_t)n * 4));
    HIPCHK(prs_.alloc(((size_t)P + 1) * 4));
    HIPCHK(pcol_.alloc((size_t)nnz * 4));
    HIPCHK(pmap_.alloc((size_t)nnz * 4));
    HIPCHK(valP_.alloc((size_t)nnz * 8));
    HIPCHK(invdP_.alloc((size_t)P * 8));
    HIPCHK(bP_.alloc((size_t)P * 8));
    HIPCHK(xP_.alloc((size_t)P * 8));
    HIPCHK(permLen_.alloc((size_t)P * 4));
    const int* order = sorder_.as<int>();
    k_row_pos<<<blocks(P), BS, 0, stream_>>>(P, order, pos_.as<int>());
    k_rterm_pos<<<blocks(P), BS, 0, stream_>>>(P, stF_.start.as<int>(), stF_.nr.as<int>(), stF_.idx.as<int>(),
                                               pos_.as<int>());
    k_perm_len<<<blocks(P), BS, 0, stream_>>>(P, order, rs_.as<int>(), permLen_.as<int>());
    size_t tb = 0;
    HIPCHK(hipcub::DeviceScan::InclusiveSum(nullptr, tb, permLen_.as<int>(), prs_.as<int>() + 1, P, stream_));
    HIPCHK(planTmp_.alloc(tb));
    HIPCHK(hipcub::DeviceScan::InclusiveSum(planTmp_.p, tb, permLen_.as<int>(), prs_.as<int>() + 1, P, stream_));
    k_fill_i32<<<1, BS, 0, stream_>>>(prs_.as<int>(), 0, 1);
    k_perm_fill<<<blocks(P), BS, 0, stream_>>>(P, order, rs_.as<int>(), col_.as<int>(), pos_.as<int>(),
                                               prs_.as<int>(), pcol_.as<int>(), pmap_.as<int>());
    HIPCHK(hipGetLastError());
    return ASTG_OK;
}

int PcgDevice::prepareSolve(std::string& err) {
    if (!perm_) return ASTG_OK;
    const int P = PS_;
    const int* order = sorder_.as<int>();
    k_gather_vals<<<std::max(1, std::min(8 * cus_, blocks(nnz_))), BS, 0, stream_>>>(
        (long)nnz_, pmap_.as<int>(), val_.as<double>(), valP_.as<double>());
    k_to_sweep<<<blocks(P), BS, 0, stream_>>>(P, order, invd_.as<double>(), invdP_.as<double>());
    k_to_sweep<<<blocks(P), BS, 0, stream_>>>(P, order, b_.as<double>(), bP_.as<double>());
    HIPCHK(hipGetLastError());
    return ASTG_OK;
}

bool PcgDevice::isSweepKernel(const void* f) const {
    return f == (const void*)k_fwd<double> || f == (const void*)k_fwd<float> || f == (const void*)k_bwd_dot<double> ||
           f == (const void*)k_bwd_dot<float>;
}

SweepPlan PcgDevice::splan() const {
    SweepPlan s;
    s.order = sorder_.as<int>();
    s.gateF = sgateF_.as<int>();
    s.gateB = sgateB_.as<int>();
    s.P = PS_;
    s.nchunks = PS_ / BS;
    s.pollCap = pollCap_;
    s.gateCap = gateCap_;
    s.trace = trace_;
    s.chunkSys = K_ ? chunkSeg_.as<int>() : nullptr;
    s.sysState = K_ ? bst_.as<double>() : nullptr;
    s.vecPos = perm_ ? 1 : 0;
    return s;
}

template <typename T>
StencilView<T> PcgDevice::view(const DevStencil& d) const {
    StencilView<T> v;
    v.start = d.start.as<int>();
    v.nr = d.nr.as<int>();
    v.idx = d.idx.as<int>();
    v.coef = sizeof(T) == 8 ? reinterpret_cast<const T*>(d.coef.p) : reinterpret_cast<const T*>(d.coef32.p);
    return v;
}

Parts PcgDevice::parts(int gb) const {
    Parts p;
    double* base = part_.as<double>();
    p.sq = base;
    p.res = base + gs_;
    p.rho0 = base + 2 * gs_;
    p.rho1 = base + 2 * gs_ + gb_;
    p.gs = gs_;
    p.gb = gb;
    return p;
}

int PcgDevice::checkErrors(const char* where, std::string& err) {
    int bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, cnt_.as<int>() + C_ERR, 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    if (bits) {
        spinFailed_ = true;
        err = std::string(where) + ": a sync-free kernel hit its spin limit (error bits " + std::to_string(bits) + ")";
        return ASTG_ERR_HIP;
    }
    return ASTG_OK;
}

bool PcgDevice::fallBackToCooperative(std::string& err) {
    if (!spinFailed_ || coop_) return false;
    std::fprintf(stderr, "astg: a sweep block was not resident under an ordinary launch (%s); "
                         "retrying with cooperative launches\n", err.c_str());
    coop_ = true;
    spinFailed_ = false;
    k_zero_counters<<<1, 64, 0, stream_>>>(cnt_.as<int>());
    if (hipStreamSynchronize(stream_) != hipSuccess) return false;
    err.clear();
    return true;
}

int PcgDevice::launchMic(double tau, double ratio, std::string& err) {
    k_mic_live<<<blocks(n_), BS, 0, stream_>>>(n_, rs_.as<int>(), lend_.as<int>(), col_.as<int>(), val_.as<double>(),
                                                 live_.as<uint8_t>());
    for (int L = 0; L < depth_; ++L) {
        const int t0 = levRange_[2 * L], t1 = levRange_[2 * L + 1];
        k_mic_level<<<blocks((long)(t1 - t0) * kMicLanes), BS, 0, stream_>>>(
            t0, t1, order_.as<int>(), micStart_.as<long long>(), micProg_.as<int>(), rs_.as<int>(), lend_.as<int>(),
            ubeg_.as<int>(), col_.as<int>(), val_.as<double>(), live_.as<uint8_t>(), tpos_.as<int>(), F_.as<double>(),
            invd_.as<double>(), tau, ratio);
    }
    HIPCHK(hipGetLastError());
    return ASTG_OK;
}

int PcgDevice::launchMerge(std::string& err) {
    const DevStencil* st[2] = {&stF_, &stB_};
    for (int dir = 0; dir < 2; ++dir) {
        const DevStencil& d = *st[dir];
        auto k = dir == 0 ? k_merge<0> : k_merge<1>;
        k<<<blocks(PS_), BS, 0, stream_>>>(PS_, sorder_.as<int>(), rs_.as<int>(), lend_.as<int>(), ubeg_.as<int>(),
                                           F_.as<double>(), invd_.as<double>(), d.start.as<int>(), d.nr.as<int>(),
                                           d.idx.as<int>(), d.ustart.as<int>(), d.upd.as<uint8_t>(),
                                           d.coef.as<double>(), perm_ ? 1 : 0);
    }
    HIPCHK(hipGetLastError());
    return ASTG_OK;
}

// FP32 copies of A, the pivots and the merged stencil coefficients.
int PcgDevice::convert32(std::string& err) {
    const long long nnz = nnz_;
    HIPCHK(val32_.alloc((size_t)nnz * 4));
    HIPCHK(invd32_.alloc((size_t)vlen() * 4));
    HIPCHK(stF_.coef32.alloc((size_t)stF_.E * 4));
    HIPCHK(stB_.coef32.alloc((size_t)stB_.E * 4));
    k_d2f<<<blocks(nnz), BS, 0, stream_>>>(sVal64(), val32_.as<float>(), nnz);
    k_d2f<<<blocks(vlen()), BS, 0, stream_>>>(sInvd64(), invd32_.as<float>(), vlen());
    k_d2f<<<blocks(stF_.E), BS, 0, stream_>>>(stF_.coef.as<double>(), stF_.coef32.as<float>(), stF_.E);
    k_d2f<<<blocks(stB_.E), BS, 0, stream_>>>(stB_.coef.as<double>(), stB_.coef32.as<float>(), stB_.E);
    HIPCHK(hipGetLastError());
    f32Valid_ = true;
    return ASTG_OK;
}

int PcgDevice::factor(double tau, double ratio, std::string& err) {
    f32Valid_ = false;  // new values and pivots: the FP32 copies are stale
    int rc = launchMic(tau, ratio, err);
    if (rc) return rc;
    return launchMerge(err);
}

int PcgDevice::syncScalars(std::string& err) {
    HIPCHK(hipMemcpyAsync(hscal_, scal_.p, 32 * 8, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

template <typename T>
int PcgDevice::launchFwd(const T* invd, const T* r, int k, int mode, int it, int maxIt, int norm, double tolRel,
                         std::string& err) {
    const int g = fwdGrid<T>(), gb = bwdGrid<T>();
    const StencilView<T> V = view<T>(stF_);
    T* y = ybuf<T>(k);
    HIPCHK(launchCoop(coop_, k_fwd<T>, g, stream_, splan(), V, invd, r, y, zp_.as<T>(), parts(gb), cnt_.as<int>(),
                      scal_.as<double>(), mode, it, maxIt, norm, tolRel));
    return ASTG_OK;
}

template <typename T>
int PcgDevice::launchBwd(const T* invd, const T* r, T* z, int k, int parityOut, std::string& err) {
    const int gb = bwdGrid<T>();
    const StencilView<T> V = view<T>(stB_);
    const T* y = ybuf<T>(k);
    T* ynext = ybuf<T>(k + 1);
    HIPCHK(launchCoop(coop_, k_bwd_dot<T>, gb, stream_, splan(), V, invd, y, ynext, r, z, zp_.as<T>(), parts(gb),
                      cnt_.as<int>(), (const double*)scal_.as<double>(), parityOut));
    return ASTG_OK;
}

template <typename T>
int PcgDevice::pcgInner(const T* A, const T* invd, const T* b, double* x, double tolRel, int maxIt, int norm,
                        int& iters, double& resOut, bool& conv, std::string& err, const double* b64,
                        int replaceEvery) {
    const int n = vlen(), gs = gs_;
    const int gb = bwdGrid<T>();
    T* r = r_.as<T>();
    T* z = z_.as<T>();
    T* s = s_.as<T>();
    T* q = q_.as<T>();
    const int* rs = sRs();
    const int* col = sCol();
    double* scal = scal_.as<double>();
    const Parts Pt = parts(gb);
    int rc;
    k_zero64<<<1, BS, 0, stream_>>>(32, scal);
    k_init<T, double><<<gs, BS, 0, stream_>>>(n, b, x, r, Pt, norm);
    k_fill_sent<T><<<blocks(2 * PS_), BS, 0, stream_>>>(y_.as<T>(), 2 * PS_);
    if ((rc = launchFwd<T>(invd, r, 0, 0, 0, maxIt, norm, tolRel, err))) return rc;
    if ((rc = launchBwd<T>(invd, r, z, 0, 0, err))) return rc;
    if (forceSpinError_ && !coop_) {  // test hook: as if a sweep had hit its spin limit
        forceSpinError_ = false;
        k_set_err<<<1, 64, 0, stream_>>>(cnt_.as<int>(), E_FWD);
    }
    // the first sweep pair shows whether every sweep block was resident: a
    // failure is caught here, before any iteration is spent on it
    rc = checkErrors("PCG", err);
    if (rc) return rc;
    // z as the PCG reads it: the backward sweep's published copy in sweep
    // order (valid until the next forward sweep re-arms it)
    const T* zv = perm_ ? zp_.as<T>() : z;
    k_xpby<T><<<gs, BS, 0, stream_>>>(n, zv, s, Pt, scal, 0, 0);
    HIPCHK(hipGetLastError());
    int it = 0;
    int batch = 4;
    for (;;) {
        for (int j = 0; j < batch && it < maxIt; ++j, ++it) {
            launchSpmv<T>(n, rs, col, A, s, q, Pt, scal);
            k_axpy_res<T, double><<<gs, BS, 0, stream_>>>(n, s, q, x, r, Pt, scal, norm, it);
            if (sizeof(T) == 4 && replaceEvery > 0 && (it + 1) % replaceEvery == 0)
                k_replace_res<<<gs, BS, 0, stream_>>>(n, rs, col, sVal64(), x, b64,
                                                      reinterpret_cast<float*>(r), Pt, scal, norm);
            if ((rc = launchFwd<T>(invd, r, it + 1, 1, it, maxIt, norm, tolRel, err))) return rc;
            if ((rc = launchBwd<T>(invd, r, z, it + 1, (it + 1) & 1, err))) return rc;
            k_xpby<T><<<gs, BS, 0, stream_>>>(n, zv, s, Pt, scal, 1, it);
        }
        HIPCHK(hipGetLastError());
        rc = syncScalars(err);
        if (rc) return rc;
        rc = checkErrors("PCG", err);
        if (rc) return rc;
        if (hscal_[S_DONE] != 0.0 || it >= maxIt) break;
        batch = std::min(batch * 2, 32);
    }
    if (hscal_[S_DONE] == 4.0) {
        const int at = (int)hscal_[S_ITER];
        err = at == 0 ? "PCG: the right-hand side holds a NaN or infinity"
                      : "PCG: the residual holds a NaN or infinity at iteration " + std::to_string(at);
        return ASTG_ERR_NOT_CONVERGED;
    }
    iters = (int)hscal_[S_ITER];
    resOut = hscal_[S_RES];
    conv = hscal_[S_DONE] == 1.0;
    if (hscal_[S_DONE] == 2.0) {
        iters = 0;
        conv = false;
    }
    return ASTG_OK;
}

int PcgDevice::solve(const astg_pcg_params& prm, astg_pcg_result& res, std::string& err) {
    spinFailed_ = false;
    int rc = solveOnce(prm, res, err);
    if (rc != ASTG_OK && fallBackToCooperative(err)) rc = solveOnce(prm, res, err);
    return rc;
}

int PcgDevice::solveOnce(const astg_pcg_params& prm, astg_pcg_result& res, std::string& err) {
    const int n = vlen(), gs = gs_;
    res.n = n_;
    res.nnz = nnz_;
    res.depth = depth_;
    res.outer = 0;
    if (int rc = prepareSolve(err)) return rc;
    // the solution in row order (x_) from the sweep-ordered one
    auto finish = [&]() -> int {
        if (perm_) k_from_sweep<<<blocks(PS_), BS, 0, stream_>>>(PS_, sorder_.as<int>(), sX(), x_.as<double>());
        HIPCHK(hipGetLastError());
        return ASTG_OK;
    };
    if (prm.precision == ASTG_PREC_FP64) {
        int it = 0;
        double r = 0;
        bool conv = false;
        int rc = pcgInner<double>(sVal64(), sInvd64(), sB(), sX(), prm.tol, prm.max_iterations, prm.norm, it, r,
                                  conv, err, nullptr, 0);
        if (rc) return rc;
        if ((rc = finish())) return rc;
        res.iterations = it;
        res.residual = r;
        res.rel_residual = hscal_[S_BNORM] > 0 ? r / hscal_[S_BNORM] : 0.0;
        res.converged = conv ? 1 : 0;
        if (hscal_[S_BNORM] == 0.0) res.converged = 1;
        return ASTG_OK;
    }
    // mixed: FP32 inner PCG whose correction accumulates in FP64 and whose
    // recurrence residual is replaced every `replace` iterations by the FP64
    // residual of the inner system (reliable updates); FP64 outer residual.
    HIPCHK(r64_.alloc((size_t)n * 8));
    HIPCHK(d64_.alloc((size_t)n * 8));
    HIPCHK(rhs64_.alloc((size_t)n * 12));
    if (int rc = convert32(err)) return rc;
    k_zero64<<<blocks(n), BS, 0, stream_>>>(n, sX());
    double* rhs64 = rhs64_.as<double>();
    float* rhs32 = reinterpret_cast<float*>(rhs64 + n);
    double* d64 = d64_.as<double>();
    double bnorm = -1, target = 0;
    int total = 0, outer = 0;
    const int maxOuter = prm.max_outer > 0 ? prm.max_outer : 20;
    double rnorm = 0;
    bool conv = false;
    for (;;) {
        const Parts Pt = parts(gB64_);
        k_residual64<<<gs, BS, 0, stream_>>>(n, sRs(), sCol(), sVal64(), sX(), sB(), r64_.as<double>(), Pt,
                                             prm.norm);
        k_res_final<<<1, BS, 0, stream_>>>(Pt, scal_.as<double>(), prm.norm);
        int rc = syncScalars(err);
        if (rc) return rc;
        rnorm = hscal_[S_OUTER_RES];
        if (!(rnorm <= DBL_MAX)) {
            err = outer == 0 ? "PCG: the right-hand side holds a NaN or infinity"
                             : "PCG: the FP64 residual holds a NaN or infinity after " + std::to_string(outer) +
                                   " inner solves";
            return ASTG_ERR_NOT_CONVERGED;
        }
        if (bnorm < 0) {
            bnorm = rnorm;
            target = prm.tol * bnorm;
        }
        if (rnorm <= target) {
            conv = true;
            break;
        }
        if (outer >= maxOuter || total >= prm.max_iterations) break;
        // Inner relative tolerance: a fraction of the outer target; with the
        // residual replaced by the true one, the inner recurrence residual
        // tracks the FP64 residual, so one inner solve normally suffices.
        const double eta = std::max(mixedEta_ * target / rnorm, 1e-7);
        k_scale_rhs<<<gs, BS, 0, stream_>>>(n, r64_.as<double>(), 1.0 / rnorm, rhs64, rhs32);
        int it = 0;
        double ir = 0;
        bool ic = false;
        rc = pcgInner<float>(val32_.as<float>(), invd32_.as<float>(), rhs32, d64, eta, prm.max_iterations - total,
                             prm.norm, it, ir, ic, err, rhs64, mixedReplace_);
        if (rc) return rc;
        total += it;
        ++outer;
        k_add_corr<<<gs, BS, 0, stream_>>>(n, sX(), d64, rnorm);
        if (it == 0) break;
    }
    if (int rc = finish()) return rc;
    res.iterations = total;
    res.outer = outer;
    res.residual = rnorm;
    res.rel_residual = bnorm > 0 ? rnorm / bnorm : 0.0;
    res.converged = conv ? 1 : 0;
    return ASTG_OK;
}

int PcgDevice::downloadFactor(std::vector<double>& invdiag, std::vector<unsigned>& colstart,
                              std::vector<unsigned>& rowindex, std::vector<double>& value) {
    const int n = n_;
    std::string perr;
    if (ensureHostPattern(perr)) return ASTG_ERR_HIP;
    std::vector<double> F(nnz_);
    std::vector<int> ubeg(n);
    invdiag.resize(n);
    if (hipMemcpyAsync(F.data(), F_.p, (size_t)nnz_ * 8, hipMemcpyDeviceToHost, stream_) != hipSuccess)
        return ASTG_ERR_HIP;
    if (hipMemcpyAsync(invdiag.data(), invd_.p, (size_t)n * 8, hipMemcpyDeviceToHost, stream_) != hipSuccess)
        return ASTG_ERR_HIP;
    if (hipMemcpyAsync(ubeg.data(), ubeg_.p, (size_t)n * 4, hipMemcpyDeviceToHost, stream_) != hipSuccess)
        return ASTG_ERR_HIP;
    if (hipStreamSynchronize(stream_) != hipSuccess) return ASTG_ERR_HIP;
    colstart.assign(n + 1, 0);
    rowindex.clear();
    value.clear();
    for (int k = 0; k < n; ++k) {
        colstart[k] = (unsigned)rowindex.size();
        for (int p = ubeg[k]; p < hrs_[k + 1]; ++p) {
            rowindex.push_back((unsigned)hcol_[p]);
            value.push_back(F[p]);
        }
    }
    colstart[n] = (unsigned)rowindex.size();
    return ASTG_OK;
}

int PcgDevice::benchKernel(int which, int reps, double& avgMs, double& bytes, std::string& err) {
    const int n = n_, gs = gs_, nv = vlen();
    const long long nnz = nnz_, nl = nnzL();
    // the solve's own layout: sweep-ordered vectors and CSR unless the row order is kept
    if (int rc = prepareSolve(err)) return rc;
    const int* rs = sRs();
    const int* col = sCol();
    double* scal = scal_.as<double>();
    const bool f32 = which == ASTG_K_SPMV32 || which == ASTG_K_FWD32 || which == ASTG_K_BWD32;
    if (f32 && !f32Valid_) {
        int rc = convert32(err);
        if (rc) return rc;
    }
    const double w = f32 ? 4.0 : 8.0;
    switch (which) {
        case ASTG_K_SPMV64: case ASTG_K_SPMV32:
            // values + columns, rowstart, s gathered once, s read for the dot, q written
            bytes = (w + 4) * nnz + 4.0 * (n + 1) + 3 * w * n;
            break;
        case ASTG_K_FWD64: case ASTG_K_FWD32:
            // stencil coefficients + indices; per row: order, start, r-term count, r, pivot,
            // y written, y gathered once, backward vector armed
            bytes = (w + 4) * stF_.E + (5.0 * w + 12) * n;
            break;
        case ASTG_K_BWD64: case ASTG_K_BWD32:
            // stencil coefficients + indices; per row: order, start, r-term count, y, next y
            // armed, pivot, r, z published (and, in row order, written to the PCG vector),
            // z gathered once
            bytes = (w + 4) * stB_.E + ((perm_ ? 6.0 : 7.0) * w + 12) * n;
            break;
        case ASTG_K_MIC:
            // A values + pattern read, factor written twice (both triangles), invd
            bytes = 12.0 * nnz + 8.0 * 2 * nl + 24.0 * n;
            break;
        case ASTG_K_COPY:
            bytes = 16.0 * n;
            break;
        default:
            err = "unknown kernel id";
            return ASTG_ERR_ARG;
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    k_zero_counters<<<1, 64, 0, stream_>>>(cnt_.as<int>());
    double total = 0;
    int rc = 0;
    const bool sweep = which == ASTG_K_FWD64 || which == ASTG_K_BWD64 || which == ASTG_K_FWD32 ||
                       which == ASTG_K_BWD32;
    for (int rep = 0; rep < reps; ++rep) {
        // re-arm the device state so no kernel early-exits
        k_zero64<<<1, BS, 0, stream_>>>(32, scal);
        const Parts P64 = parts(bwdGrid<double>()), P32 = parts(bwdGrid<float>());
        if (which == ASTG_K_FWD64 || which == ASTG_K_BWD64) {
            k_fill_sent<double><<<blocks(2 * PS_), BS, 0, stream_>>>(y_.as<double>(), 2 * PS_);
            k_fill_sent<double><<<blocks(PS_), BS, 0, stream_>>>(zp_.as<double>(), PS_);
        } else if (which == ASTG_K_FWD32 || which == ASTG_K_BWD32) {
            k_fill_sent<float><<<blocks(2 * PS_), BS, 0, stream_>>>(y_.as<float>(), 2 * PS_);
            k_fill_sent<float><<<blocks(PS_), BS, 0, stream_>>>(zp_.as<float>(), PS_);
        }
        // the forward sweep starts in mode 0, which needs nonzero ||b|| partials
        if (sweep) k_fill64<<<blocks(gs_), BS, 0, stream_>>>(gs_, P64.res, 1.0);
        if (which == ASTG_K_BWD64)
            rc = launchFwd<double>(sInvd64(), r_.as<double>(), 0, 0, 0, 1 << 30, ASTG_NORM_INF, 1e-30, err);
        else if (which == ASTG_K_BWD32)
            rc = launchFwd<float>(invd32_.as<float>(), r_.as<float>(), 0, 0, 0, 1 << 30, ASTG_NORM_INF, 1e-30, err);
        if (rc) break;
        HIPCHK(hipEventRecord(e0, stream_));
        switch (which) {
            case ASTG_K_SPMV64:
                launchSpmv<double>(nv, rs, col, sVal64(), s_.as<double>(),
                                                           q_.as<double>(), P64, scal);
                break;
            case ASTG_K_SPMV32:
                launchSpmv<float>(nv, rs, col, val32_.as<float>(), s_.as<float>(),
                                                          q_.as<float>(), P32, scal);
                break;
            case ASTG_K_FWD64:
                rc = launchFwd<double>(sInvd64(), r_.as<double>(), 0, 0, 0, 1 << 30, ASTG_NORM_INF, 1e-30, err);
                break;
            case ASTG_K_FWD32:
                rc = launchFwd<float>(invd32_.as<float>(), r_.as<float>(), 0, 0, 0, 1 << 30, ASTG_NORM_INF, 1e-30,
                                      err);
                break;
            case ASTG_K_BWD64:
                rc = launchBwd<double>(sInvd64(), r_.as<double>(), z_.as<double>(), 0, 0, err);
                break;
            case ASTG_K_BWD32:
                rc = launchBwd<float>(invd32_.as<float>(), r_.as<float>(), z_.as<float>(), 0, 0, err);
                break;
            case ASTG_K_MIC:
                rc = launchMic(0.97, 0.25, err);
                break;
            case ASTG_K_COPY:
                k_copy64<<<blocks(n), BS, 0, stream_>>>(n, b_.as<double>(), q_.as<double>());
                break;
        }
        if (rc) break;
        HIPCHK(hipEventRecord(e1, stream_));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        total += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    rc = checkErrors("kernel timing", err);
    if (rc) return rc;
    avgMs = total / reps;
    return ASTG_OK;
}

int PcgDevice::sweepTimeline(int which, std::vector<long long>& t, std::vector<int>& lev, std::string& err) {
    if (which != ASTG_K_FWD64 && which != ASTG_K_BWD64) {
        err = "timeline: only the FP64 sweeps are instrumented";
        return ASTG_ERR_ARG;
    }
    DevBuf tr;
    HIPCHK(tr.alloc((size_t)n_ * 3 * sizeof(long long)));
    HIPCHK(hipMemsetAsync(tr.p, 0, (size_t)n_ * 3 * sizeof(long long), stream_));
    trace_ = tr.as<long long>();
    double ms = 0, bytes = 0;
    int rc = benchKernel(which, 1, ms, bytes, err);
    trace_ = nullptr;
    if (rc) return rc;
    t.resize((size_t)n_ * 3);
    lev.resize(n_);
    HIPCHK(hipMemcpyAsync(t.data(), tr.p, t.size() * sizeof(long long), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(lev.data(), lev_.p, (size_t)n_ * 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

template int PcgDevice::launchFwd<double>(const double*, const double*, int, int, int, int, int, double,
                                          std::string&);
template int PcgDevice::launchFwd<float>(const float*, const float*, int, int, int, int, int, double, std::string&);
template int PcgDevice::launchBwd<double>(const double*, const double*, double*, int, int, std::string&);
template int PcgDevice::launchBwd<float>(const float*, const float*, float*, int, int, std::string&);
template int PcgDevice::pcgInner<double>(const double*, const double*, const double*, double*, double, int, int, int&,
                                         double&, bool&, std::string&, const double*, int);
template int PcgDevice::pcgInner<float>(const float*, const float*, const float*, double*, double, int, int, int&,
                                        double&, bool&, std::string&, const double*, int);

}  // namespace astg
