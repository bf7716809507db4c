// Batched MIC(0)-PCG: many independent systems solved by one sequence of
// launches.
//
// The systems are stacked into one block-diagonal matrix, every system's rows
// padded with identity rows (zero right-hand side) to a multiple of 256, so
// that every 256-row chunk belongs to one system. The pattern-only work
// (dependency levels, factorisation plan, merged sweep stencils) and the
// factorisation are PcgDevice's own on the stacked matrix. With no coupling
// between the blocks, the MIC(0) factor of the stack is the stack of the
// systems' factors and a row's dependency level is its level inside its own
// system, so the sweeps compute what the per-system sweeps compute, and a
// batch is as deep as its deepest system rather than the sum of the depths.
//
// Per system: the PCG scalars (rho, alpha, beta, the norms), the reference's
// convergence test (pcg_solver.h:396-421) and the iteration count. Streaming
// kernels write one partial per chunk; kb_reduce, one block per system, sums
// the system's partials in chunk order (deterministic) and advances the
// system's scalars. A finished system's chunks are skipped by the streaming
// kernels. The sweeps run over all rows: a finished system's residual no
// longer changes, so its rows recompute values nobody reads.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "pcg.h"
#include "pcg_common.h"

namespace astg {

namespace {

enum BatchPhase { P_INIT = 0, P_RHO0, P_SQ, P_RES, P_RHO, P_OUTER };

__device__ inline bool chunk_live(const BatchView& V, int c) {
    return V.st[V.chunkSeg[c] * B_COUNT + B_DONE] == 0.0;
}
__device__ inline double norm_part(double a, int norm) { return norm == ASTG_NORM_INF ? fabs(a) : a * a; }
__device__ inline double norm_reduce(double v, int norm) {
    return norm == ASTG_NORM_INF ? block_reduce<true>(v) : block_reduce<false>(v);
}

template <typename T, typename X>
__global__ void __launch_bounds__(BS) kb_init(BatchView V, const T* b, X* x, T* r, int norm) {
    for (int c = blockIdx.x; c < V.nchunks; c += gridDim.x) {
        const int i = c * BS + threadIdx.x;
        const T bi = b[i];
        x[i] = (X)0;
        r[i] = bi;
        const double t = norm_reduce(norm_part((double)bi, norm), norm);
        if (threadIdx.x == 0) V.part[c] = t;
    }
}

// One block per system: the system's partials in chunk order, then the
// phase's scalar update (thread 0). A system that finishes here is counted;
// the last one to finish raises the all-done flag that ends the sweeps.
__global__ void __launch_bounds__(BS) kb_reduce(BatchView V, int phase, int norm, double* scal) {
    const int k = blockIdx.x;
    double* s = V.st + (size_t)k * B_COUNT;
    if (phase != P_OUTER && s[B_DONE] != 0.0) return;
    const bool isNorm = phase == P_INIT || phase == P_RES || phase == P_OUTER;
    const bool mx = isNorm && norm == ASTG_NORM_INF;
    const int c0 = V.segChunk[k], c1 = V.segChunk[k + 1];
    double acc = 0.0;
    for (int c = c0 + threadIdx.x; c < c1; c += BS) acc = mx ? nan_max(acc, V.part[c]) : acc + V.part[c];
    double red = mx ? block_reduce<true>(acc) : block_reduce<false>(acc);
    if (threadIdx.x != 0) return;
    if (isNorm && !mx) red = sqrt(red);
    bool fin = false;
    if ((phase == P_INIT || phase == P_RES) && s[B_SKIP] == 0.0 && !(red <= DBL_MAX)) {
        s[B_RES] = red;  // a NaN or infinity in this system's b or residual: it stops (state 5)
        if (phase == P_RES) s[B_ITER] += 1.0;
        s[B_DONE] = 5.0;
        if (atomicAdd(V.ndone, 1) + 1 == V.K) scal[S_DONE] = 1.0;
        return;
    }
    switch (phase) {
        case P_INIT:
            s[B_ITER] = 0.0;
            if (s[B_SKIP] != 0.0) {
                s[B_DONE] = 4.0;
                fin = true;
                break;
            }
            s[B_BNORM] = red;
            s[B_RES] = red;
            s[B_TOL] = s[B_TOLREL] * red;
            if (red == 0.0) {
                s[B_DONE] = 1.0;
                fin = true;
            } else if (s[B_MAXIT] <= 0.0) {
                s[B_DONE] = 3.0;
                fin = true;
            }
            break;
        case P_RHO0:  // the reference's breakdown test after the first preconditioning
            if (red == 0.0 || red != red) {
                s[B_DONE] = 2.0;
                s[B_ITER] = 0.0;
                fin = true;
            } else {
                s[B_RHO] = red;
            }
            break;
        case P_SQ:
            s[B_ALPHA] = s[B_RHO] / red;
            break;
        case P_RES: {
            const double it1 = s[B_ITER] + 1.0;
            s[B_RES] = red;
            s[B_ITER] = it1;
            if (red <= s[B_TOL]) {
                s[B_DONE] = 1.0;
                fin = true;
            } else if (it1 >= s[B_MAXIT]) {
                s[B_DONE] = 3.0;
                fin = true;
            }
            break;
        }
        case P_RHO:
            s[B_BETA] = red / s[B_RHO];
            s[B_RHO] = red;
            break;
        case P_OUTER:
            s[B_OUTER] = red;
            break;
    }
    if (fin && atomicAdd(V.ndone, 1) + 1 == V.K) scal[S_DONE] = 1.0;
}

// Partials of z.r per chunk (the single-system form fuses this dot into the
// backward sweep, whose rows are in sweep order and mix the systems).
template <typename T>
__global__ void __launch_bounds__(BS) kb_dot(BatchView V, const T* z, const T* r) {
    for (int c = blockIdx.x; c < V.nchunks; c += gridDim.x) {
        if (!chunk_live(V, c)) continue;
        const int i = c * BS + threadIdx.x;
        const double t = block_reduce<false>((double)z[i] * (double)r[i]);
        if (threadIdx.x == 0) V.part[c] = t;
    }
}

// q = A s and the partials of s.q, one chunk (256 rows) per block pass, LPR
// lanes per row.
template <typename T, int LPR>
__global__ void __launch_bounds__(BS) kb_spmv_dot(BatchView V, const int* rs, const int* col, const T* A, const T* s,
                                                  T* q) {
    const int sub = threadIdx.x & (LPR - 1);
    for (int c = blockIdx.x; c < V.nchunks; c += gridDim.x) {
        if (!chunk_live(V, c)) continue;
        double v = 0.0;
        // two of the lane group's rows at a time (row_dot_u), in row order
        for (int j = 0; j < LPR; j += 2) {
            const int lane = (int)threadIdx.x / LPR;
            const int i[2] = {c * BS + j * (BS / LPR) + lane, c * BS + (j + 1) * (BS / LPR) + lane};
            T acc[2];
            row_dot_u<T, LPR, 2>(rs, col, A, s, i, sub, acc);
            if (sub == 0)
                for (int u = 0; u < 2; ++u) {
                    q[i[u]] = acc[u];
                    v += (double)s[i[u]] * (double)acc[u];
                }
        }
        const double t = block_reduce<false>(v);
        if (threadIdx.x == 0) V.part[c] = t;
    }
}

template <typename T, typename X>
__global__ void __launch_bounds__(BS) kb_axpy_res(BatchView V, const T* s, const T* q, X* x, T* r, int norm) {
    for (int c = blockIdx.x; c < V.nchunks; c += gridDim.x) {
        if (!chunk_live(V, c)) continue;
        const double ad = V.st[V.chunkSeg[c] * B_COUNT + B_ALPHA];
        const T alpha = (T)ad;
        const X ax = sizeof(X) > sizeof(T) ? (X)ad : (X)alpha;
        const int i = c * BS + threadIdx.x;
        x[i] += ax * (X)s[i];
        const T ri = r[i] - alpha * q[i];
        r[i] = ri;
        const double t = norm_reduce(norm_part((double)ri, norm), norm);
        if (threadIdx.x == 0) V.part[c] = t;
    }
}

// s = z (mode 0) or s = z + beta s (mode 1).
template <typename T>
__global__ void __launch_bounds__(BS) kb_xpby(BatchView V, const T* z, T* s, int mode) {
    for (int c = blockIdx.x; c < V.nchunks; c += gridDim.x) {
        if (!chunk_live(V, c)) continue;
        const int i = c * BS + threadIdx.x;
        if (mode == 0) {
            s[i] = z[i];
        } else {
            const T beta = (T)V.st[V.chunkSeg[c] * B_COUNT + B_BETA];
            s[i] = z[i] + beta * s[i];
        }
    }
}

// Mixed form, reliable update: r = b64 - A x in FP64, stored in FP32.
template <int LPR>
__global__ void __launch_bounds__(BS) kb_replace_res(BatchView V, const int* rs, const int* col, const double* A,
                                                     const double* x, const double* b, float* r, int norm) {
    const int sub = threadIdx.x & (LPR - 1);
    for (int c = blockIdx.x; c < V.nchunks; c += gridDim.x) {
        if (!chunk_live(V, c)) continue;
        double v = 0.0;
        for (int j = 0; j < LPR; ++j) {
            const int i = c * BS + j * (BS / LPR) + threadIdx.x / LPR;
            const double acc = row_dot<double, LPR>(rs, col, A, x, i, sub);
            if (sub == 0) {
                const float ri = (float)(b[i] - acc);
                r[i] = ri;
                v = norm == ASTG_NORM_INF ? nan_max(v, fabs((double)ri)) : v + (double)ri * (double)ri;
            }
        }
        const double t = norm_reduce(v, norm);
        if (threadIdx.x == 0) V.part[c] = t;
    }
}

// Mixed form, outer residual r = b - A x (FP64) of every system.
template <int LPR>
__global__ void __launch_bounds__(BS) kb_residual64(BatchView V, const int* rs, const int* col, const double* A,
                                                    const double* x, const double* b, double* r, int norm) {
    const int sub = threadIdx.x & (LPR - 1);
    for (int c = blockIdx.x; c < V.nchunks; c += gridDim.x) {
        double v = 0.0;
        for (int j = 0; j < LPR; ++j) {
            const int i = c * BS + j * (BS / LPR) + threadIdx.x / LPR;
            const double acc = row_dot<double, LPR>(rs, col, A, x, i, sub);
            if (sub == 0) {
                const double ri = b[i] - acc;
                r[i] = ri;
                v = norm == ASTG_NORM_INF ? nan_max(v, fabs(ri)) : v + ri * ri;
            }
        }
        const double t = norm_reduce(v, norm);
        if (threadIdx.x == 0) V.part[c] = t;
    }
}

// Mixed form: inner right-hand side r64 / ||r64|| (zero for systems not
// solved this time), and the correction x += ||r64|| d of the solved ones.
__global__ void __launch_bounds__(BS) kb_scale_rhs(BatchView V, const double* r, double* o64, float* o32) {
    for (int c = blockIdx.x; c < V.nchunks; c += gridDim.x) {
        const double* s = V.st + V.chunkSeg[c] * B_COUNT;
        const int i = c * BS + threadIdx.x;
        const double v = s[B_SKIP] != 0.0 ? 0.0 : r[i] * s[B_SCALE];
        o64[i] = v;
        o32[i] = (float)v;
    }
}
__global__ void __launch_bounds__(BS) kb_add_corr(BatchView V, double* x, const double* d) {
    for (int c = blockIdx.x; c < V.nchunks; c += gridDim.x) {
        const double* s = V.st + V.chunkSeg[c] * B_COUNT;
        if (s[B_SKIP] != 0.0) continue;
        const int i = c * BS + threadIdx.x;
        x[i] += s[B_OUTER] * d[i];
    }
}

__global__ void kb_zero_count(int* c) {
    if (threadIdx.x == 0) c[0] = 0;
}
__global__ void kb_zero64(int n, double* x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = 0.0;
}
template <typename T>
__global__ void kb_fill_sent(T* p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Sentinel<T>::storeSent(&p[i]);
}
inline int nblocks(long n) { return (int)((n + BS - 1) / BS); }

}  // namespace

int PcgDevice::setBatch(const std::vector<int>& segRow, std::string& err) {
    const int K = (int)segRow.size() - 1;
    if (K < 1 || segRow[0] != 0 || segRow[K] != n_) {
        err = "batch layout does not cover the stacked system";
        return ASTG_ERR_ARG;
    }
    const int nch = n_ / BS;
    std::vector<int> segChunk(K + 1), chunkSeg(nch);
    for (int k = 0; k <= K; ++k) {
        if (segRow[k] % BS || (k && segRow[k] <= segRow[k - 1])) {
            err = "batch layout: system starts must be increasing multiples of 256";
            return ASTG_ERR_ARG;
        }
        segChunk[k] = segRow[k] / BS;
    }
    for (int k = 0; k < K; ++k)
        for (int c = segChunk[k]; c < segChunk[k + 1]; ++c) chunkSeg[c] = k;
    HIPCHK(chunkSeg_.alloc((size_t)nch * 4));
    HIPCHK(segChunk_.alloc((size_t)(K + 1) * 4));
    HIPCHK(bst_.alloc((size_t)K * B_COUNT * 8));
    HIPCHK(bpart_.alloc((size_t)nch * 8));
    HIPCHK(bcnt_.alloc(64));
    HIPCHK(hipMemcpyAsync(chunkSeg_.p, chunkSeg.data(), (size_t)nch * 4, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyAsync(segChunk_.p, segChunk.data(), (size_t)(K + 1) * 4, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    K_ = K;
    return ASTG_OK;
}

int PcgDevice::batchGrid() const {
    const int want = blocksPerCu_ > 0 ? blocksPerCu_ * cus_ : gs_;
    return std::max(1, std::min(want, n_ / BS));
}

void PcgDevice::setBatchTuning(int blocksPerCu, int lanesPerRow) {
    blocksPerCu_ = std::max(0, blocksPerCu);
    lpr_ = (lanesPerRow == 4 || lanesPerRow == 16) ? lanesPerRow : 8;
}

BatchView PcgDevice::bview() const {
    BatchView v;
    v.chunkSeg = chunkSeg_.as<int>();
    v.segChunk = segChunk_.as<int>();
    v.st = bst_.as<double>();
    v.part = bpart_.as<double>();
    v.ndone = bcnt_.as<int>();
    v.K = K_;
    v.nchunks = n_ / BS;
    return v;
}

// A system stopped on a NaN or infinity (state 5) fails the whole call.
static int nonFinite(const std::vector<double>& st, std::string& err) {
    for (size_t k = 0; k * B_COUNT < st.size(); ++k)
        if (st[k * B_COUNT + B_DONE] == 5.0) {
            const int at = (int)st[k * B_COUNT + B_ITER];
            err = "PCG (batched): system " + std::to_string(k) +
                  (at == 0 ? " has a NaN or infinity in its right-hand side"
                           : " has a NaN or infinity in its residual at iteration " + std::to_string(at));
            return ASTG_ERR_NOT_CONVERGED;
        }
    return ASTG_OK;
}

int PcgDevice::batchState(std::vector<double>& st, bool toDevice, std::string& err) {
    const size_t bytes = (size_t)K_ * B_COUNT * 8;
    if (toDevice) {
        HIPCHK(hipMemcpyAsync(bst_.p, st.data(), bytes, hipMemcpyHostToDevice, stream_));
    } else {
        st.resize((size_t)K_ * B_COUNT);
        HIPCHK(hipMemcpyAsync(st.data(), bst_.p, bytes, hipMemcpyDeviceToHost, stream_));
    }
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

template <typename T>
int PcgDevice::pcgBatchInner(const T* A, const T* invd, const T* b, double* x, int maxIt, int norm, std::string& err,
                             const double* b64, int replaceEvery) {
    const BatchView V = bview();
    const int g = batchGrid();
    T* r = r_.as<T>();
    T* z = z_.as<T>();
    T* s = s_.as<T>();
    T* q = q_.as<T>();
    const int* rs = rs_.as<int>();
    const int* col = col_.as<int>();
    double* scal = scal_.as<double>();
    int rc;
    kb_zero64<<<1, BS, 0, stream_>>>(32, scal);
    kb_zero_count<<<1, 64, 0, stream_>>>(V.ndone);
    kb_init<T, double><<<g, BS, 0, stream_>>>(V, b, x, r, norm);
    kb_reduce<<<V.K, BS, 0, stream_>>>(V, P_INIT, norm, scal);
    kb_fill_sent<T><<<nblocks(2 * (long)PS_), BS, 0, stream_>>>(y_.as<T>(), 2 * (long)PS_);
    if ((rc = launchFwd<T>(invd, r, 0, 2, 0, maxIt, norm, 0.0, err))) return rc;
    if ((rc = launchBwd<T>(invd, r, z, 0, 0, err))) return rc;
    kb_dot<T><<<g, BS, 0, stream_>>>(V, z, r);
    kb_reduce<<<V.K, BS, 0, stream_>>>(V, P_RHO0, norm, scal);
    kb_xpby<T><<<g, BS, 0, stream_>>>(V, z, s, 0);
    HIPCHK(hipGetLastError());
    // one PCG iteration of every running system
    auto iteration = [&](int itx, bool replace) -> int {
        if (lpr_ == 4)
            kb_spmv_dot<T, 4><<<g, BS, 0, stream_>>>(V, rs, col, A, s, q);
        else if (lpr_ == 16)
            kb_spmv_dot<T, 16><<<g, BS, 0, stream_>>>(V, rs, col, A, s, q);
        else
            kb_spmv_dot<T, 8><<<g, BS, 0, stream_>>>(V, rs, col, A, s, q);
        kb_reduce<<<V.K, BS, 0, stream_>>>(V, P_SQ, norm, scal);
        kb_axpy_res<T, double><<<g, BS, 0, stream_>>>(V, s, q, x, r, norm);
        if (replace) {
            float* r32 = reinterpret_cast<float*>(r);
            if (lpr_ == 4)
                kb_replace_res<4><<<g, BS, 0, stream_>>>(V, rs, col, val_.as<double>(), x, b64, r32, norm);
            else if (lpr_ == 16)
                kb_replace_res<16><<<g, BS, 0, stream_>>>(V, rs, col, val_.as<double>(), x, b64, r32, norm);
            else
                kb_replace_res<8><<<g, BS, 0, stream_>>>(V, rs, col, val_.as<double>(), x, b64, r32, norm);
        }
        kb_reduce<<<V.K, BS, 0, stream_>>>(V, P_RES, norm, scal);
        int e = launchFwd<T>(invd, r, itx + 1, 2, itx, maxIt, norm, 0.0, err);
        if (!e) e = launchBwd<T>(invd, r, z, itx + 1, (itx + 1) & 1, err);
        if (e) return e;
        kb_dot<T><<<g, BS, 0, stream_>>>(V, z, r);
        kb_reduce<<<V.K, BS, 0, stream_>>>(V, P_RHO, norm, scal);
        kb_xpby<T><<<g, BS, 0, stream_>>>(V, z, s, 1);
        return ASTG_OK;
    };
    auto replaceAt = [&](int itx) {
        return sizeof(T) == 4 && replaceEvery > 0 && (itx + 1) % replaceEvery == 0;
    };
    // Two consecutive iterations captured as one hipGraph: the sweep vectors
    // alternate every iteration and nothing else depends on the iteration
    // number. Variant 1 replaces the residual in its second iteration. Under
    // cooperative launches the sweep nodes carry the cooperative attribute.
    hipGraphExec_t pair[2] = {nullptr, nullptr};
    bool useGraph = batchGraph_;
    auto buildPair = [&](int variant) -> bool {
        hipGraph_t gr = nullptr;
        if (hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        const int c1 = iteration(0, false), c2 = iteration(1, variant == 1);
        const hipError_t ce = hipStreamEndCapture(stream_, &gr);
        bool ok = c1 == 0 && c2 == 0 && ce == hipSuccess && gr != nullptr;
        if (ok) {
            size_t nn = 0;
            ok = hipGraphGetNodes(gr, nullptr, &nn) == hipSuccess;
            std::vector<hipGraphNode_t> nodes(nn);
            if (ok) ok = hipGraphGetNodes(gr, nodes.data(), &nn) == hipSuccess;
            for (size_t a = 0; ok && a < nn; ++a) {
                hipGraphNodeType ty;
                if (hipGraphNodeGetType(nodes[a], &ty) != hipSuccess || ty != hipGraphNodeTypeKernel) continue;
                hipKernelNodeParams kp;
                if (hipGraphKernelNodeGetParams(nodes[a], &kp) != hipSuccess) {
                    ok = false;
                    break;
                }
                if (!coop_ || !isSweepKernel(kp.func)) continue;
                hipKernelNodeAttrValue v;
                std::memset(&v, 0, sizeof v);
                v.cooperative = 1;
                ok = hipGraphKernelNodeSetAttribute(nodes[a], hipKernelNodeAttributeCooperative, &v) == hipSuccess;
            }
        }
        if (ok) ok = hipGraphInstantiate(&pair[variant], gr, nullptr, nullptr, 0) == hipSuccess;
        if (gr) (void)hipGraphDestroy(gr);
        if (!ok) {
            pair[variant] = nullptr;
            (void)hipGetLastError();
        }
        return ok;
    };
    auto destroyPairs = [&] {
        for (auto& ex : pair)
            if (ex) (void)hipGraphExecDestroy(ex);
    };
    int it = 0, batch = 4, replays = 0;
    for (;;) {
        for (int j = 0; j < batch && it < maxIt;) {
            const int variant = replaceAt(it + 1) ? 1 : 0;
            if (useGraph && it % 2 == 0 && it + 2 <= maxIt && j + 2 <= batch && !replaceAt(it)) {
                if (!pair[variant] && !buildPair(variant)) {
                    useGraph = false;
                    continue;
                }
                if (hipGraphLaunch(pair[variant], stream_) != hipSuccess) {
                    destroyPairs();
                    err = "batched PCG: graph launch failed";
                    return ASTG_ERR_HIP;
                }
                it += 2;
                j += 2;
                ++replays;
                continue;
            }
            if ((rc = iteration(it, replaceAt(it)))) {
                destroyPairs();
                return rc;
            }
            ++it;
            ++j;
        }
        if (hipGetLastError() != hipSuccess || (rc = syncScalars(err)) || (rc = checkErrors("batched PCG", err))) {
            destroyPairs();
            return rc ? rc : ASTG_ERR_HIP;
        }
        if (hscal_[S_DONE] != 0.0 || it >= maxIt) break;
        batch = std::min(batch * 2, 32);
    }
    destroyPairs();
    if (getenv("ASTG_TIMING"))
        fprintf(stderr, "[astg timing] batched PCG: %d iterations, %d two-iteration graph replays%s\n", it, replays,
                batchGraph_ && !useGraph ? " (capture failed: eager)" : "");
    return ASTG_OK;
}

int PcgDevice::solveBatch(const astg_pcg_params& prm, std::vector<astg_pcg_result>& res, std::string& err) {
    spinFailed_ = false;
    int rc = solveBatchOnce(prm, res, err);
    if (rc != ASTG_OK && fallBackToCooperative(err)) rc = solveBatchOnce(prm, res, err);
    return rc;
}

int PcgDevice::solveBatchOnce(const astg_pcg_params& prm, std::vector<astg_pcg_result>& res, std::string& err) {
    const int K = K_;
    if (K < 1) {
        err = "no batch layout";
        return ASTG_ERR_ARG;
    }
    const BatchView V = bview();
    const int g = batchGrid();
    const int maxIt = std::max(0, prm.max_iterations);
    res.assign(K, astg_pcg_result());
    for (auto& r : res) {
        std::memset(&r, 0, sizeof(r));
        r.depth = depth_;
    }
    std::vector<double> st((size_t)K * B_COUNT, 0.0);
    auto S = [&](int k, int slot) -> double& { return st[(size_t)k * B_COUNT + slot]; };
    if (prm.precision == ASTG_PREC_FP64) {
        for (int k = 0; k < K; ++k) {
            S(k, B_TOLREL) = prm.tol;
            S(k, B_MAXIT) = maxIt;
        }
        int rc = batchState(st, true, err);
        if (rc) return rc;
        rc = pcgBatchInner<double>(val_.as<double>(), invd_.as<double>(), b_.as<double>(), x_.as<double>(), maxIt,
                                   prm.norm, err, nullptr, 0);
        if (rc) return rc;
        if ((rc = batchState(st, false, err))) return rc;
        if ((rc = nonFinite(st, err))) return rc;
        for (int k = 0; k < K; ++k) {
            astg_pcg_result& R = res[k];
            const double done = S(k, B_DONE), bn = S(k, B_BNORM);
            R.iterations = done == 2.0 ? 0 : (int)S(k, B_ITER);
            R.residual = S(k, B_RES);
            R.rel_residual = bn > 0 ? R.residual / bn : 0.0;
            R.converged = (done == 1.0 || bn == 0.0) ? 1 : 0;
        }
        return ASTG_OK;
    }
    // Mixed form, per system as PcgDevice::solve: FP64 outer residual, FP32
    // inner PCG on r / ||r|| with FP64 accumulation and reliable updates,
    // correction x += ||r|| d; systems drop out as they converge.
    const int n = n_;
    HIPCHK(r64_.alloc((size_t)n * 8));
    HIPCHK(d64_.alloc((size_t)n * 8));
    HIPCHK(rhs64_.alloc((size_t)n * 12));
    if (int rc = convert32(err)) return rc;
    kb_zero64<<<nblocks(n), BS, 0, stream_>>>(n, x_.as<double>());
    double* rhs64 = rhs64_.as<double>();
    float* rhs32 = reinterpret_cast<float*>(rhs64 + n);
    double* d64 = d64_.as<double>();
    const int maxOuter = prm.max_outer > 0 ? prm.max_outer : 20;
    std::vector<double> bnorm(K, -1.0), target(K, 0.0), rnorm(K, 0.0);
    std::vector<int> total(K, 0), outer(K, 0);
    std::vector<char> conv(K, 0), stopped(K, 0);
    for (;;) {
        {
            const int* rs = rs_.as<int>();
            const int* col = col_.as<int>();
            const double* A = val_.as<double>();
            if (lpr_ == 4)
                kb_residual64<4><<<g, BS, 0, stream_>>>(V, rs, col, A, x_.as<double>(), b_.as<double>(),
                                                        r64_.as<double>(), prm.norm);
            else if (lpr_ == 16)
                kb_residual64<16><<<g, BS, 0, stream_>>>(V, rs, col, A, x_.as<double>(), b_.as<double>(),
                                                         r64_.as<double>(), prm.norm);
            else
                kb_residual64<8><<<g, BS, 0, stream_>>>(V, rs, col, A, x_.as<double>(), b_.as<double>(),
                                                        r64_.as<double>(), prm.norm);
        }
        kb_reduce<<<K, BS, 0, stream_>>>(V, P_OUTER, prm.norm, scal_.as<double>());
        HIPCHK(hipGetLastError());
        int rc = batchState(st, false, err);
        if (rc) return rc;
        int active = 0, innerMax = 0;
        for (int k = 0; k < K; ++k) {
            const double rn = S(k, B_OUTER);
            if (!stopped[k] && !(rn <= DBL_MAX)) {
                err = "PCG (batched): the FP64 residual of system " + std::to_string(k) +
                      " holds a NaN or infinity";
                return ASTG_ERR_NOT_CONVERGED;
            }
            if (!stopped[k]) rnorm[k] = rn;
            if (bnorm[k] < 0) {
                bnorm[k] = rn;
                target[k] = prm.tol * rn;
            }
            if (!stopped[k]) {
                if (rn <= target[k]) {
                    conv[k] = 1;
                    stopped[k] = 1;
                } else if (outer[k] >= maxOuter || total[k] >= maxIt) {
                    stopped[k] = 1;
                }
            }
            std::fill(st.begin() + (size_t)k * B_COUNT, st.begin() + (size_t)(k + 1) * B_COUNT, 0.0);
            S(k, B_OUTER) = rn;
            if (stopped[k]) {
                S(k, B_SKIP) = 1.0;
                continue;
            }
            ++active;
            S(k, B_TOLREL) = std::max(mixedEta_ * target[k] / rn, 1e-7);
            S(k, B_MAXIT) = maxIt - total[k];
            S(k, B_SCALE) = 1.0 / rn;
            innerMax = std::max(innerMax, maxIt - total[k]);
        }
        if (!active) break;
        if ((rc = batchState(st, true, err))) return rc;
        kb_scale_rhs<<<g, BS, 0, stream_>>>(V, r64_.as<double>(), rhs64, rhs32);
        rc = pcgBatchInner<float>(val32_.as<float>(), invd32_.as<float>(), rhs32, d64, innerMax, prm.norm, err, rhs64,
                                  mixedReplace_);
        if (rc) return rc;
        kb_add_corr<<<g, BS, 0, stream_>>>(V, x_.as<double>(), d64);
        HIPCHK(hipGetLastError());
        if ((rc = batchState(st, false, err))) return rc;
        if ((rc = nonFinite(st, err))) return rc;
        for (int k = 0; k < K; ++k) {
            if (S(k, B_SKIP) != 0.0) continue;
            const int it = S(k, B_DONE) == 2.0 ? 0 : (int)S(k, B_ITER);
            total[k] += it;
            ++outer[k];
            if (it == 0) stopped[k] = 1;
        }
    }
    for (int k = 0; k < K; ++k) {
        astg_pcg_result& R = res[k];
        R.iterations = total[k];
        R.outer = outer[k];
        R.residual = rnorm[k];
        R.rel_residual = bnorm[k] > 0 ? rnorm[k] / bnorm[k] : 0.0;
        R.converged = conv[k];
    }
    return ASTG_OK;
}

template int PcgDevice::pcgBatchInner<double>(const double*, const double*, const double*, double*, int, int,
                                              std::string&, const double*, int);
template int PcgDevice::pcgBatchInner<float>(const float*, const float*, const float*, double*, int, int,
                                             std::string&, const double*, int);

}  // namespace astg
