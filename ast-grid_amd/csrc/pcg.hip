// Device MIC(0)-PCG (see pcg.h). Kernels are written for gfx950 wave64.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_common.h"
#include "pcg.h"

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

DevBuf::~DevBuf() {
    if (p) (void)hipFree(p);
}
hipError_t DevBuf::alloc(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, b ? b : 16);
    if (e == hipSuccess) bytes = b;
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
__global__ void k_zero_counters(int* c) {
    if (threadIdx.x < C_COUNT) c[threadIdx.x] = 0;
}

// Block-level reduction into one partial per slot, and the final reduction by
// the last block to finish (fixed order over slots, so results are
// deterministic). Returns true in the block that performed the final step.
template <bool MAX>
__device__ inline bool reduce_partials(double v, double* part, int slot, int nslots, int* doneCtr, double* out) {
    __shared__ double sm[BS / 64];
    __shared__ bool last;
    v = MAX ? wave_max(v) : wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sm[0];
        for (int w = 1; w < BS / 64; ++w) t = MAX ? fmax(t, sm[w]) : t + sm[w];
        st_agent_u64(reinterpret_cast<unsigned long long*>(&part[slot]), (unsigned long long)__double_as_longlong(t));
        __threadfence();
        int prev = atomicAdd(doneCtr, 1);
        last = (prev == nslots - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double acc = MAX ? 0.0 : 0.0;
    for (int s = threadIdx.x; s < nslots; s += BS) {
        double t = __longlong_as_double((long long)ld_agent_u64(reinterpret_cast<const unsigned long long*>(&part[s])));
        acc = MAX ? fmax(acc, t) : acc + t;
    }
    acc = MAX ? wave_max(acc) : wave_sum(acc);
    if (lane == 0) sm[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sm[0];
        for (int w = 1; w < BS / 64; ++w) t = MAX ? fmax(t, sm[w]) : t + sm[w];
        *out = t;
        *doneCtr = 0;
    }
    return true;
}

// ---------------------------------------------------------------------------
// dependency levels of tril(A): level(i) = 1 + max level(k), k in lower(i).
// One row per wavefront (its lanes poll the row's dependencies in parallel),
// rows handed out in index order through a ticket: every row a wave waits on
// belongs to an earlier, already running wave, and no wave ever waits on one
// of its own lanes.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BS) k_levels(int n, const int* rs, const int* lend, const int* col, int* lev, int* cnt,
                                               double* scal) {
    __shared__ int chunk;
    if (threadIdx.x == 0) chunk = atomicAdd(&cnt[C_TICKET_L], 1);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int i = chunk * (BS / 64) + (threadIdx.x >> 6);
    if (i >= n) return;
    const int p0 = rs[i], pe = lend[i];
    int L = 0;
    for (int base = p0; base < pe; base += 64) {
        const int p = base + lane;
        if (p < pe) {
            const int k = col[p];
            int lk, spins = 0;
            while ((lk = ld_agent_i32(&lev[k])) < 0) {
                if (++spins > kSpinLimit) {
                    atomicOr(&cnt[C_ERR], E_LEVELS);
                    lk = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            L = max(L, lk + 1);
        }
    }
    for (int o = 32; o > 0; o >>= 1) L = max(L, __shfl_xor(L, o, 64));
    if (lane == 0) st_agent_i32(&lev[i], L);
}

// ---------------------------------------------------------------------------
// MIC(0) factorisation, one row per lane, rows in level order.
// Row k waits until every column m in lower(k) is final (invdiag[m] published),
// then computes its pivot and column k of L (stored at the upper positions of
// row k and, through tpos, at the matching lower positions), then publishes
// invdiag[k]. Accumulation order follows the reference's right-looking loop:
// contributions to row/column k arrive in ascending m.
// ---------------------------------------------------------------------------
__device__ inline void mic_row(const int k, const int r0, const int le, const int ub, const int r1, const int* rs,
                               const int* lend, const int* ubeg, const int* col, const double* A, const int* tpos,
                               double* F, double* invd, double tau, double ratio) {
    const double akk = (le < r1 && col[le] == k) ? A[le] : 0.0;
    double d = akk;
    for (int p = r0; p < le; ++p) {
        const int m = col[p];
        const int mr0 = rs[m], mle = lend[m];
        const double amm = (mle < rs[m + 1] && col[mle] == m) ? A[mle] : 0.0;
        if (amm == 0.0) continue;  // null column: never eliminated
        const double Lkm = F[p];
        // missing fill of column m with respect to row k's pattern
        double missing = 0.0;
        int a = ubeg[m];
        const int ae = rs[m + 1];
        int b = r0;
        for (; a < ae; ++a) {
            const int q = col[a];
            if (q == k) continue;
            while (b < r1 && col[b] < q) ++b;
            if (b < r1 && col[b] == q) continue;
            missing += F[a];
        }
        d -= Lkm * Lkm;
        d -= tau * Lkm * missing;
        (void)mr0;
    }
    double inv;
    if (akk == 0.0) {
        inv = d;  // null row: left as accumulated, column not scaled
    } else if (d < ratio * akk) {
        inv = 1.0 / sqrt(akk);
    } else {
        inv = 1.0 / sqrt(d);
    }
    // column k of L: entries (i,k), i in upper(k)
    for (int p = ub; p < r1; ++p) {
        const int i = col[p];
        double v = A[p];
        // subtract sum over m in lower(k) and lower(i), ascending
        int a = r0, c = rs[i];
        const int ce = lend[i];
        while (a < le && c < ce) {
            const int ma = col[a], mc = col[c];
            if (ma < mc) {
                ++a;
            } else if (mc < ma) {
                ++c;
            } else {
                const int mr = lend[ma];
                const double amm = (mr < rs[ma + 1] && col[mr] == ma) ? A[mr] : 0.0;
                if (amm != 0.0) v -= F[a] * F[c];
                ++a;
                ++c;
            }
        }
        if (akk != 0.0) v *= inv;
        F[p] = v;
        F[tpos[p]] = v;
    }
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(&invd[k]), (unsigned long long)__double_as_longlong(inv),
                       __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(BS) k_mic(int P, const int* order, LevelGate G, const int* rs, const int* lend,
                                            const int* ubeg, const int* col, const double* A, const int* tpos,
                                            double* F, double* invd, double tau, double ratio, int* cnt) {
    __shared__ int chunk;
    if (threadIdx.x == 0) chunk = atomicAdd(&cnt[C_TICKET_M], 1);
    __syncthreads();
    const int t = chunk * BS + threadIdx.x;
    const int lev = (t >> 6) < G.nseg ? G.wlev[t >> 6] : -1;
    gate_wait(G, lev - kGate);
    const int k = t < P ? order[t] : -1;
    if (k >= 0) {
        const int r0 = rs[k], le = lend[k], ub = ubeg[k], r1 = rs[k + 1];
        // Wait for all lower dependencies, then factor the row inside the same
        // loop iteration.
        int pw = r0, spins = 0;
        for (;;) {
            while (pw < le) {
                unsigned long long bb = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(&invd[col[pw]]),
                                                          __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                if (bb == kSent64) break;
                ++pw;
            }
            if (pw < le) {
                if (++spins <= kSpinLimit) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                atomicOr(&cnt[C_ERR], E_MIC);
            }
            mic_row(k, r0, le, ub, r1, rs, lend, ubeg, col, A, tpos, F, invd, tau, ratio);
            break;
        }
    }
    gate_arrive(G, lev);
}

// ---------------------------------------------------------------------------
// PCG kernels (T = double or float; scalars are always double)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(BS) k_init(int n, const T* b, T* x, T* r, T* y, T* z, double* part, int* cnt,
                                             double* scal, int norm, double tolRel, int nblk, int* gateF, int* gateB,
                                             int depth) {
    const int i = blockIdx.x * BS + threadIdx.x;
    for (int q = i; q < depth; q += gridDim.x * BS) {
        gateF[q] = 0;
        gateB[q] = 0;
    }
    double v = 0.0;
    if (i < n) {
        T bi = b[i];
        x[i] = (T)0;
        r[i] = bi;
        Sentinel<T>::storeSent(&y[i]);
        Sentinel<T>::storeSent(&z[i]);
        v = norm == ASTG_NORM_INF ? fabs((double)bi) : (double)bi * (double)bi;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cnt[C_TICKET_F] = 0;
        cnt[C_TICKET_B] = 0;
    }
    double tot;
    bool last = norm == ASTG_NORM_INF ? reduce_partials<true>(v, part, blockIdx.x, nblk, &cnt[C_DONE_RES], &tot)
                                      : reduce_partials<false>(v, part, blockIdx.x, nblk, &cnt[C_DONE_RES], &tot);
    if (last && threadIdx.x == 0) {
        double bn = norm == ASTG_NORM_INF ? tot : sqrt(tot);
        scal[S_BNORM] = bn;
        scal[S_RES] = bn;
        scal[S_TOL] = tolRel * bn;
        scal[S_ITER] = 0;
        scal[S_DONE] = (bn == 0.0) ? 1.0 : 0.0;
    }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_fwd(int P, const int* order, LevelGate G, int* gateOther, const int* rs,
                                            const int* lend, const int* col, const T* F, const T* invd, const T* r,
                                            T* y, T* z, int* cnt, double* scal) {
    if (scal[S_DONE] != 0.0) return;
    __shared__ int chunk;
    if (threadIdx.x == 0) chunk = atomicAdd(&cnt[C_TICKET_F], 1);
    __syncthreads();
    if (chunk == 0 && threadIdx.x == 0) cnt[C_TICKET_B] = 0;
    for (int q = blockIdx.x * BS + threadIdx.x; q < G.depth; q += gridDim.x * BS) gateOther[q] = 0;
    const int t = chunk * BS + threadIdx.x;
    const int lev = (t >> 6) < G.nseg ? G.wlev[t >> 6] : -1;
    gate_wait(G, lev - kGate);
    const int i = t < P ? order[t] : -1;
    if (i >= 0) {
        T acc = r[i];
        int p = rs[i];
        const int pe = lend[i];
        int spins = 0;
        Sentinel<T>::storeSent(&z[i]);
        for (;;) {
            while (p < pe) {
                auto bits = Sentinel<T>::load(&y[col[p]]);
                if (bits == Sentinel<T>::value) break;
                acc -= F[p] * Sentinel<T>::val(bits);
                ++p;
            }
            if (p == pe) {
                Sentinel<T>::store(&y[i], acc * invd[i]);
                break;
            }
            if (++spins > kSpinLimit) {
                atomicOr(&cnt[C_ERR], E_FWD);
                Sentinel<T>::store(&y[i], (T)NAN);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    gate_arrive(G, lev);
}

// Backward sweep z = L^-T y fused with the partial dot z.r. Partials are
// indexed by chunk so the reduction order is fixed.
template <typename T>
__global__ void __launch_bounds__(BS) k_bwd_dot(int P, const int* order, LevelGate G, const int* rs, const int* ubeg,
                                                const int* col, const T* F, const T* invd, const T* y, const T* r, T* z,
                                                double* part, int* cnt, double* scal, int init, int nblk) {
    if (scal[S_DONE] != 0.0) return;
    __shared__ int chunk;
    if (threadIdx.x == 0) chunk = atomicAdd(&cnt[C_TICKET_B], 1);
    __syncthreads();
    const int t = chunk * BS + threadIdx.x;
    const int pos = P - 1 - t;  // reverse level order
    const int lev = (pos >= 0 && (pos >> 6) < G.nseg) ? G.wlev[pos >> 6] : -1;
    if (lev >= 0) gate_wait(G, lev + kGate);
    double v = 0.0;
    const int k = pos >= 0 ? order[pos] : -1;
    if (k >= 0) {
        T acc = y[k];
        int p = ubeg[k];
        const int pe = rs[k + 1];
        int spins = 0;
        T zk;
        for (;;) {
            while (p < pe) {
                auto bits = Sentinel<T>::load(&z[col[p]]);
                if (bits == Sentinel<T>::value) break;
                acc -= F[p] * Sentinel<T>::val(bits);
                ++p;
            }
            if (p == pe) {
                zk = acc * invd[k];
                Sentinel<T>::store(&z[k], zk);
                break;
            }
            if (++spins > kSpinLimit) {
                atomicOr(&cnt[C_ERR], E_BWD);
                zk = (T)NAN;
                Sentinel<T>::store(&z[k], zk);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        v = (double)zk * (double)r[k];
    }
    gate_arrive(G, lev);
    double tot;
    if (reduce_partials<false>(v, part, chunk, nblk, &cnt[C_DONE_RHO], &tot) && threadIdx.x == 0) {
        if (init) {
            scal[S_RHO] = tot;
            if (tot == 0.0 || tot != tot) {
                scal[S_DONE] = 2.0;  // breakdown before the first iteration (reference returns false)
                scal[S_ITER] = 0;
            }
        } else {
            scal[S_BETA] = tot / scal[S_RHO];
            scal[S_RHO] = tot;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_xpby(int n, const T* z, T* s, double* scal, int init) {
    if (scal[S_DONE] != 0.0) return;
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    if (init) {
        s[i] = z[i];
    } else {
        const T beta = (T)scal[S_BETA];
        s[i] = z[i] + beta * s[i];
    }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_spmv_dot(int n, const int* rs, const int* col, const T* A, const T* s, T* q,
                                                 double* part, int* cnt, double* scal, int nblk) {
    if (scal[S_DONE] != 0.0) return;
    const int i = blockIdx.x * BS + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        T acc = (T)0;
        for (int p = rs[i]; p < rs[i + 1]; ++p) acc += A[p] * s[col[p]];
        q[i] = acc;
        v = (double)s[i] * (double)acc;
    }
    double tot;
    if (reduce_partials<false>(v, part, blockIdx.x, nblk, &cnt[C_DONE_SQ], &tot) && threadIdx.x == 0) {
        scal[S_SQ] = tot;
        scal[S_ALPHA] = scal[S_RHO] / tot;
    }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_axpy_res(int n, const T* s, const T* q, T* x, T* r, T* y, double* part, int* cnt,
                                                 double* scal, int norm, int maxIt, int nblk, int* gateF, int depth) {
    if (scal[S_DONE] != 0.0) return;
    const int i = blockIdx.x * BS + threadIdx.x;
    for (int g = i; g < depth; g += gridDim.x * BS) gateF[g] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[C_TICKET_F] = 0;
    double v = 0.0;
    if (i < n) {
        const T alpha = (T)scal[S_ALPHA];
        x[i] += alpha * s[i];
        T ri = r[i] - alpha * q[i];
        r[i] = ri;
        Sentinel<T>::storeSent(&y[i]);
        v = norm == ASTG_NORM_INF ? fabs((double)ri) : (double)ri * (double)ri;
    }
    double tot;
    bool last = norm == ASTG_NORM_INF ? reduce_partials<true>(v, part, blockIdx.x, nblk, &cnt[C_DONE_RES], &tot)
                                      : reduce_partials<false>(v, part, blockIdx.x, nblk, &cnt[C_DONE_RES], &tot);
    if (last && threadIdx.x == 0) {
        double res = norm == ASTG_NORM_INF ? tot : sqrt(tot);
        scal[S_RES] = res;
        double it = scal[S_ITER] + 1.0;
        scal[S_ITER] = it;
        if (res <= scal[S_TOL]) scal[S_DONE] = 1.0;
        else if (it >= (double)maxIt) scal[S_DONE] = 3.0;
    }
}

// Mixed form: r64 = b - A x (FP64) with max|r| / sum r^2.
__global__ void __launch_bounds__(BS) k_residual64(int n, const int* rs, const int* col, const double* A,
                                                   const double* x, const double* b, double* r, double* part,
                                                   int* cnt, double* scal, int norm, int nblk) {
    const int i = blockIdx.x * BS + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        double acc = 0.0;
        for (int p = rs[i]; p < rs[i + 1]; ++p) acc += A[p] * x[col[p]];
        double ri = b[i] - acc;
        r[i] = ri;
        v = norm == ASTG_NORM_INF ? fabs(ri) : ri * ri;
    }
    double tot;
    bool last = norm == ASTG_NORM_INF ? reduce_partials<true>(v, part, blockIdx.x, nblk, &cnt[C_DONE_RES], &tot)
                                      : reduce_partials<false>(v, part, blockIdx.x, nblk, &cnt[C_DONE_RES], &tot);
    if (last && threadIdx.x == 0) scal[S_RES + 8] = norm == ASTG_NORM_INF ? tot : sqrt(tot);
}
__global__ void k_scale_f32(int n, const double* r, float* o, double inv) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) o[i] = (float)(r[i] * inv);
}
__global__ void k_add_corr(int n, double* x, const float* d, double scale) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x[i] += scale * (double)d[i];
}
__global__ void k_zero64(int n, double* x) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
PcgDevice::~PcgDevice() {
    if (hscal_) (void)hipHostFree(hscal_);
}

static inline int blocks(long n) { return (int)((n + BS - 1) / BS); }

int PcgDevice::setPattern(int n, const int* rowstart, const int* col, std::string& err) {
    if (n <= 0) {
        err = "empty system";
        return ASTG_ERR_ARG;
    }
    n_ = n;
    nnz_ = rowstart[n];
    hrs_.assign(rowstart, rowstart + n + 1);
    hcol_.assign(col, col + nnz_);
    // split each row at the diagonal; build the transpose map and check symmetry
    std::vector<int> lend(n), ubeg(n), tpos(nnz_, 0);
    for (int i = 0; i < n; ++i) {
        int p = rowstart[i];
        const int e = rowstart[i + 1];
        for (int q = p + 1; q < e; ++q)
            if (col[q] <= col[q - 1]) {
                err = "row " + std::to_string(i) + " is not sorted by column";
                return ASTG_ERR_ARG;
            }
        while (p < e && col[p] < i) ++p;
        lend[i] = p;
        ubeg[i] = (p < e && col[p] == i) ? p + 1 : p;
    }
    for (int i = 0; i < n; ++i)
        for (int p = rowstart[i]; p < rowstart[i + 1]; ++p) {
            const int j = col[p];
            if (j < 0 || j >= n) {
                err = "column index out of range";
                return ASTG_ERR_ARG;
            }
            if (j == i) continue;
            const int* b = col + rowstart[j];
            const int* e = col + rowstart[j + 1];
            const int* f = std::lower_bound(b, e, i);
            if (f == e || *f != i) {
                err = "matrix is not structurally symmetric at (" + std::to_string(i) + "," + std::to_string(j) + ")";
                return ASTG_ERR_ARG;
            }
            tpos[p] = (int)(f - col);
        }
    nblk_ = blocks(n);
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
    HIPCHK(r_.alloc(n8));
    HIPCHK(z_.alloc(n8));
    HIPCHK(s_.alloc(n8));
    HIPCHK(q_.alloc(n8));
    HIPCHK(y_.alloc(n8));
    HIPCHK(scal_.alloc(32 * 8));
    HIPCHK(cnt_.alloc(C_COUNT * 4));
    HIPCHK(part_.alloc((size_t)(nblk_ + 1) * 8));
    if (!hscal_) HIPCHK(hipHostMalloc((void**)&hscal_, 32 * 8, hipHostMallocDefault));
    HIPCHK(hipMemcpyAsync(rs_.p, rowstart, ((size_t)n + 1) * 4, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyAsync(col_.p, col, (size_t)nnz_ * 4, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyAsync(lend_.p, lend.data(), n4, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyAsync(ubeg_.p, ubeg.data(), n4, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyAsync(tpos_.p, tpos.data(), (size_t)nnz_ * 4, hipMemcpyHostToDevice, stream_));
    // zero scalars/counters with kernels (no memset nodes)
    k_zero_counters<<<1, 64, 0, stream_>>>(cnt_.as<int>());
    k_zero64<<<1, BS, 0, stream_>>>(32, scal_.as<double>());
    // dependency levels on the device
    k_fill_i32<<<blocks(n), BS, 0, stream_>>>(lev_.as<int>(), kSentLevel, n);
    k_levels<<<(n + BS / 64 - 1) / (BS / 64), BS, 0, stream_>>>(n, rs_.as<int>(), lend_.as<int>(), col_.as<int>(), lev_.as<int>(),
                                            cnt_.as<int>(), scal_.as<double>());
    HIPCHK(hipGetLastError());
    // stable sort of (level, row): rows in level order, ties by index
    DevBuf keysOut, idxIn, tmp;
    HIPCHK(keysOut.alloc(n4));
    HIPCHK(idxIn.alloc(n4));
    k_iota<<<blocks(n), BS, 0, stream_>>>(idxIn.as<int>(), n);
    size_t tmpBytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmpBytes, lev_.as<int>(), keysOut.as<int>(), idxIn.as<int>(),
                                              order_.as<int>(), n, 0, 31, stream_));
    HIPCHK(tmp.alloc(tmpBytes));
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmpBytes, lev_.as<int>(), keysOut.as<int>(), idxIn.as<int>(),
                                              order_.as<int>(), n, 0, 31, stream_));
    std::vector<int> sortedLev(n), sortedRow(n);
    HIPCHK(hipMemcpyAsync(sortedLev.data(), keysOut.p, n4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(sortedRow.data(), order_.p, n4, hipMemcpyDeviceToHost, stream_));
    int errBits = 0;
    HIPCHK(hipMemcpyAsync(&errBits, cnt_.as<int>() + C_ERR, 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    if (errBits) {
        err = "level computation did not complete";
        return ASTG_ERR_HIP;
    }
    depth_ = sortedLev[n - 1] + 1;
    // Pad every level to a multiple of 64 rows so that a wavefront never holds
    // two rows of different levels (rows of one level are independent).
    std::vector<int> padded;
    padded.reserve((size_t)n + (size_t)depth_ * 64 + BS);
    for (int i = 0; i < n;) {
        int j = i;
        while (j < n && sortedLev[j] == sortedLev[i]) padded.push_back(sortedRow[j++]);
        while (padded.size() % 64) padded.push_back(-1);
        i = j;
    }
    while (padded.size() % BS) padded.push_back(-1);
    P_ = (int)padded.size();
    nblkS_ = P_ / BS;
    {
        std::vector<int> rowLev(n);
        for (int q = 0; q < n; ++q) rowLev[sortedRow[q]] = sortedLev[q];
        const int nseg = P_ / 64;
        std::vector<int> wlev(nseg, -1), wcnt(depth_, 0);
        for (int sgi = 0; sgi < nseg; ++sgi) {
            int r = padded[(size_t)sgi * 64];
            if (r >= 0) {
                wlev[sgi] = rowLev[r];
                wcnt[rowLev[r]] += 1;
            }
        }
        HIPCHK(wlev_.alloc((size_t)nseg * 4));
        HIPCHK(wcnt_.alloc((size_t)depth_ * 4));
        HIPCHK(gateF_.alloc((size_t)depth_ * 4));
        HIPCHK(gateB_.alloc((size_t)depth_ * 4));
        HIPCHK(gateM_.alloc((size_t)depth_ * 4));
        HIPCHK(hipMemcpyAsync(wlev_.p, wlev.data(), (size_t)nseg * 4, hipMemcpyHostToDevice, stream_));
        HIPCHK(hipMemcpyAsync(wcnt_.p, wcnt.data(), (size_t)depth_ * 4, hipMemcpyHostToDevice, stream_));
    }
    HIPCHK(order_.alloc((size_t)P_ * 4));
    HIPCHK(part_.alloc((size_t)(std::max(nblk_, nblkS_) + 1) * 8));
    HIPCHK(hipMemcpyAsync(order_.p, padded.data(), (size_t)P_ * 4, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

LevelGate PcgDevice::gate(DevBuf& counters) const {
    LevelGate g;
    g.wlev = wlev_.as<int>();
    g.wcnt = wcnt_.as<int>();
    g.done = counters.as<int>();
    g.nseg = P_ / 64;
    g.depth = depth_;
    return g;
}

int PcgDevice::checkErrors(const char* where, std::string& err) {
    int bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, cnt_.as<int>() + C_ERR, 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    if (bits) {
        err = std::string(where) + ": a sync-free kernel hit its spin limit (error bits " + std::to_string(bits) + ")";
        return ASTG_ERR_HIP;
    }
    return ASTG_OK;
}

int PcgDevice::factor(double tau, double ratio, std::string& err) {
    const int n = n_;
    k_zero_counters<<<1, 64, 0, stream_>>>(cnt_.as<int>());
    k_zero64<<<1, BS, 0, stream_>>>(32, scal_.as<double>());
    k_fill_sent<double><<<blocks(n), BS, 0, stream_>>>(invd_.as<double>(), n);
    k_fill_i32<<<blocks(depth_), BS, 0, stream_>>>(gateM_.as<int>(), 0, depth_);
    k_mic<<<nblkS_, BS, 0, stream_>>>(P_, order_.as<int>(), gate(gateM_), rs_.as<int>(), lend_.as<int>(),
                                      ubeg_.as<int>(), col_.as<int>(), val_.as<double>(), tpos_.as<int>(),
                                      F_.as<double>(), invd_.as<double>(), tau, ratio, cnt_.as<int>());
    HIPCHK(hipGetLastError());
    return checkErrors("MIC(0) factorisation", err);
}

int PcgDevice::syncScalars(std::string& err) {
    HIPCHK(hipMemcpyAsync(hscal_, scal_.p, 32 * 8, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

template <typename T>
int PcgDevice::pcgInner(const T* A, const T* F, const T* invd, const T* b, T* x, double tolRel, int maxIt, int norm,
                        int& iters, double& resOut, bool& conv, std::string& err) {
    const int n = n_, nb = nblk_, P = P_, nbs = nblkS_;
    T* r = r_.as<T>();
    T* z = z_.as<T>();
    T* s = s_.as<T>();
    T* q = q_.as<T>();
    T* y = y_.as<T>();
    const int* rs = rs_.as<int>();
    const int* col = col_.as<int>();
    const int* lend = lend_.as<int>();
    const int* ubeg = ubeg_.as<int>();
    const int* order = order_.as<int>();
    double* part = part_.as<double>();
    int* cnt = cnt_.as<int>();
    double* scal = scal_.as<double>();
    const LevelGate gF = gate(gateF_), gB = gate(gateB_);
    const int D = depth_;
    k_init<T><<<nb, BS, 0, stream_>>>(n, b, x, r, y, z, part, cnt, scal, norm, tolRel, nb, gateF_.as<int>(),
                                      gateB_.as<int>(), D);
    k_fwd<T><<<nbs, BS, 0, stream_>>>(P, order, gF, gateB_.as<int>(), rs, lend, col, F, invd, r, y, z, cnt, scal);
    k_bwd_dot<T><<<nbs, BS, 0, stream_>>>(P, order, gB, rs, ubeg, col, F, invd, y, r, z, part, cnt, scal, 1, nbs);
    k_xpby<T><<<nb, BS, 0, stream_>>>(n, z, s, scal, 1);
    HIPCHK(hipGetLastError());
    int enq = 0;
    int batch = 4;
    for (;;) {
        for (int j = 0; j < batch && enq < maxIt; ++j, ++enq) {
            k_spmv_dot<T><<<nb, BS, 0, stream_>>>(n, rs, col, A, s, q, part, cnt, scal, nb);
            k_axpy_res<T><<<nb, BS, 0, stream_>>>(n, s, q, x, r, y, part, cnt, scal, norm, maxIt, nb,
                                                  gateF_.as<int>(), D);
            k_fwd<T><<<nbs, BS, 0, stream_>>>(P, order, gF, gateB_.as<int>(), rs, lend, col, F, invd, r, y, z, cnt,
                                              scal);
            k_bwd_dot<T><<<nbs, BS, 0, stream_>>>(P, order, gB, rs, ubeg, col, F, invd, y, r, z, part, cnt, scal, 0,
                                                  nbs);
            k_xpby<T><<<nb, BS, 0, stream_>>>(n, z, s, scal, 0);
        }
        HIPCHK(hipGetLastError());
        int rc = syncScalars(err);
        if (rc) return rc;
        rc = checkErrors("PCG", err);
        if (rc) return rc;
        if (hscal_[S_DONE] != 0.0 || enq >= maxIt) break;
        batch = std::min(batch * 2, 32);
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
    const int n = n_, nb = nblk_;
    res.n = n;
    res.nnz = nnz_;
    res.depth = depth_;
    res.outer = 0;
    if (prm.precision == ASTG_PREC_FP64) {
        int it = 0;
        double r = 0;
        bool conv = false;
        int rc = pcgInner<double>(val_.as<double>(), F_.as<double>(), invd_.as<double>(), b_.as<double>(),
                                  x_.as<double>(), prm.tol, prm.max_iterations, prm.norm, it, r, conv, err);
        if (rc) return rc;
        res.iterations = it;
        res.residual = r;
        res.rel_residual = hscal_[S_BNORM] > 0 ? r / hscal_[S_BNORM] : 0.0;
        res.converged = conv ? 1 : 0;
        if (hscal_[S_BNORM] == 0.0) res.converged = 1;
        return ASTG_OK;
    }
    // mixed: FP32 inner PCG, FP64 outer residual
    const long nnz = nnz_;
    HIPCHK(val32_.alloc((size_t)nnz * 4));
    HIPCHK(F32_.alloc((size_t)nnz * 4));
    HIPCHK(invd32_.alloc((size_t)n * 4));
    HIPCHK(r64_.alloc((size_t)n * 8));
    HIPCHK(d64_.alloc((size_t)n * 8));
    k_d2f<<<blocks(nnz), BS, 0, stream_>>>(val_.as<double>(), val32_.as<float>(), nnz);
    k_d2f<<<blocks(nnz), BS, 0, stream_>>>(F_.as<double>(), F32_.as<float>(), nnz);
    k_d2f<<<blocks(n), BS, 0, stream_>>>(invd_.as<double>(), invd32_.as<float>(), n);
    k_zero64<<<nb, BS, 0, stream_>>>(n, x_.as<double>());
    float* rhs32 = reinterpret_cast<float*>(d64_.p);
    float* d32 = reinterpret_cast<float*>(d64_.as<char>() + (size_t)n * 4);
    double bnorm = -1, target = 0;
    int total = 0, outer = 0;
    const int maxOuter = prm.max_outer > 0 ? prm.max_outer : 20;
    double rnorm = 0;
    bool conv = false;
    for (;;) {
        k_residual64<<<nb, BS, 0, stream_>>>(n, rs_.as<int>(), col_.as<int>(), val_.as<double>(), x_.as<double>(),
                                             b_.as<double>(), r64_.as<double>(), part_.as<double>(), cnt_.as<int>(),
                                             scal_.as<double>(), prm.norm, nb);
        int rc = syncScalars(err);
        if (rc) return rc;
        rnorm = hscal_[S_RES + 8];
        if (bnorm < 0) {
            bnorm = rnorm;
            target = prm.tol * bnorm;
        }
        if (rnorm <= target) {
            conv = true;
            break;
        }
        if (outer >= maxOuter || total >= prm.max_iterations) break;
        // inner relative tolerance: aim just below the outer target, not below FP32 reach
        double eta = std::max(0.5 * target / rnorm, 2e-5);
        k_scale_f32<<<nb, BS, 0, stream_>>>(n, r64_.as<double>(), rhs32, 1.0 / rnorm);
        int it = 0;
        double ir = 0;
        bool ic = false;
        rc = pcgInner<float>(val32_.as<float>(), F32_.as<float>(), invd32_.as<float>(), rhs32, d32, eta,
                             prm.max_iterations - total, prm.norm, it, ir, ic, err);
        if (rc) return rc;
        total += it;
        ++outer;
        k_add_corr<<<nb, BS, 0, stream_>>>(n, x_.as<double>(), d32, rnorm);
        if (it == 0) break;
    }
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
    std::vector<double> F(nnz_);
    std::vector<int> ubeg(n);
    invdiag.resize(n);
    if (hipMemcpyAsync(F.data(), F_.p, (size_t)nnz_ * 8, hipMemcpyDeviceToHost, stream_) != hipSuccess) return ASTG_ERR_HIP;
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
    const int n = n_, nb = nblk_, P = P_, nbs = nblkS_;
    const long long nnz = nnz_, nl = nnzL();
    const int* rs = rs_.as<int>();
    const int* col = col_.as<int>();
    const int* lend = lend_.as<int>();
    const int* ubeg = ubeg_.as<int>();
    const int* order = order_.as<int>();
    double* part = part_.as<double>();
    int* cnt = cnt_.as<int>();
    double* scal = scal_.as<double>();
    const bool f32 = which == ASTG_K_SPMV32 || which == ASTG_K_FWD32 || which == ASTG_K_BWD32;
    if (f32 && !val32_.p) {
        HIPCHK(val32_.alloc((size_t)nnz * 4));
        HIPCHK(F32_.alloc((size_t)nnz * 4));
        HIPCHK(invd32_.alloc((size_t)n * 4));
        k_d2f<<<blocks(nnz), BS, 0, stream_>>>(val_.as<double>(), val32_.as<float>(), nnz);
        k_d2f<<<blocks(nnz), BS, 0, stream_>>>(F_.as<double>(), F32_.as<float>(), nnz);
        k_d2f<<<blocks(n), BS, 0, stream_>>>(invd_.as<double>(), invd32_.as<float>(), n);
    }
    const double w = f32 ? 4.0 : 8.0;
    switch (which) {
        case ASTG_K_SPMV64: case ASTG_K_SPMV32:
            bytes = (w + 4) * nnz + (2 * w + 4) * (double)n + w * n;  // values+cols, rowstart, s gather, q write, s read
            break;
        case ASTG_K_FWD64: case ASTG_K_FWD32:
            bytes = (w + 4) * nl + (5 * w + 12) * (double)n;  // L values+cols, order+rs+lend, r, invd, y write, y gather
            break;
        case ASTG_K_BWD64: case ASTG_K_BWD32:
            bytes = (w + 4) * nl + (5 * w + 12) * (double)n;  // L^T values+cols, order+ubeg+rs, y, r, invd, z write/gather
            break;
        case ASTG_K_MIC:
            bytes = 12.0 * nnz + 8.0 * nnz + 36.0 * n;  // A + pattern read, L written, invd/order/splits
            break;
        default:
            err = "unknown kernel id";
            return ASTG_ERR_ARG;
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double total = 0;
    for (int rep = 0; rep < reps; ++rep) {
        // reset device flags so the kernels do not early-exit, and re-arm the sweep inputs
        k_zero_counters<<<1, 64, 0, stream_>>>(cnt);
        k_zero64<<<1, BS, 0, stream_>>>(32, scal);
        k_fill_i32<<<blocks(depth_), BS, 0, stream_>>>(gateF_.as<int>(), 0, depth_);
        k_fill_i32<<<blocks(depth_), BS, 0, stream_>>>(gateB_.as<int>(), 0, depth_);
        k_fill_i32<<<blocks(depth_), BS, 0, stream_>>>(gateM_.as<int>(), 0, depth_);
        if (which == ASTG_K_FWD64 || which == ASTG_K_BWD64) {
            k_fill_sent<double><<<nb, BS, 0, stream_>>>(y_.as<double>(), n);
            k_fill_sent<double><<<nb, BS, 0, stream_>>>(z_.as<double>(), n);
        } else if (which == ASTG_K_FWD32 || which == ASTG_K_BWD32) {
            k_fill_sent<float><<<nb, BS, 0, stream_>>>(y_.as<float>(), n);
            k_fill_sent<float><<<nb, BS, 0, stream_>>>(z_.as<float>(), n);
        } else if (which == ASTG_K_MIC) {
            k_fill_sent<double><<<nb, BS, 0, stream_>>>(invd_.as<double>(), n);
        }
        if (which == ASTG_K_BWD64) {  // the backward sweep needs a complete y
            k_fwd<double><<<nbs, BS, 0, stream_>>>(P, order, gate(gateF_), gateB_.as<int>(), rs, lend, col, F_.as<double>(), invd_.as<double>(),
                                                  r_.as<double>(), y_.as<double>(), z_.as<double>(), cnt, scal);
        } else if (which == ASTG_K_BWD32) {
            k_fwd<float><<<nbs, BS, 0, stream_>>>(P, order, gate(gateF_), gateB_.as<int>(), rs, lend, col, F32_.as<float>(), invd32_.as<float>(),
                                                 r_.as<float>(), y_.as<float>(), z_.as<float>(), cnt, scal);
        }
        HIPCHK(hipEventRecord(e0, stream_));
        switch (which) {
            case ASTG_K_SPMV64:
                k_spmv_dot<double><<<nb, BS, 0, stream_>>>(n, rs, col, val_.as<double>(), s_.as<double>(),
                                                           q_.as<double>(), part, cnt, scal, nb);
                break;
            case ASTG_K_SPMV32:
                k_spmv_dot<float><<<nb, BS, 0, stream_>>>(n, rs, col, val32_.as<float>(), s_.as<float>(),
                                                          q_.as<float>(), part, cnt, scal, nb);
                break;
            case ASTG_K_FWD64:
                k_fwd<double><<<nbs, BS, 0, stream_>>>(P, order, gate(gateF_), gateB_.as<int>(), rs, lend, col, F_.as<double>(), invd_.as<double>(),
                                                      r_.as<double>(), y_.as<double>(), z_.as<double>(), cnt, scal);
                break;
            case ASTG_K_FWD32:
                k_fwd<float><<<nbs, BS, 0, stream_>>>(P, order, gate(gateF_), gateB_.as<int>(), rs, lend, col, F32_.as<float>(), invd32_.as<float>(),
                                                     r_.as<float>(), y_.as<float>(), z_.as<float>(), cnt, scal);
                break;
            case ASTG_K_BWD64:
                k_bwd_dot<double><<<nbs, BS, 0, stream_>>>(P, order, gate(gateB_), rs, ubeg, col, F_.as<double>(),
                                                          invd_.as<double>(), y_.as<double>(), r_.as<double>(),
                                                          z_.as<double>(), part, cnt, scal, 1, nbs);
                break;
            case ASTG_K_BWD32:
                k_bwd_dot<float><<<nbs, BS, 0, stream_>>>(P, order, gate(gateB_), rs, ubeg, col, F32_.as<float>(),
                                                         invd32_.as<float>(), y_.as<float>(), r_.as<float>(),
                                                         z_.as<float>(), part, cnt, scal, 1, nbs);
                break;
            case ASTG_K_MIC:
                k_mic<<<nbs, BS, 0, stream_>>>(P, order, gate(gateM_), rs, lend, ubeg, col, val_.as<double>(),
                                               tpos_.as<int>(), F_.as<double>(), invd_.as<double>(), 0.97, 0.25, cnt);
                break;
        }
        HIPCHK(hipEventRecord(e1, stream_));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        total += ms;
    }
    HIPCHK(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int rc = checkErrors("kernel timing", err);
    if (rc) return rc;
    avgMs = total / reps;
    return ASTG_OK;
}

template int PcgDevice::pcgInner<double>(const double*, const double*, const double*, const double*, double*, double,
                                         int, int, int&, double&, bool&, std::string&);
template int PcgDevice::pcgInner<float>(const float*, const float*, const float*, const float*, float*, double, int,
                                        int, int&, double&, bool&, std::string&);

}  // namespace astg
