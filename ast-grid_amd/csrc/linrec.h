// Transposes of the step's fixed linear passes, recorded from the passes'
// own kernel bodies.
//
// A linear pass is written once, as a functor body(t, pol) run by one lane per
// work item t, that reads field values with pol.ld(base, i) and writes them
// with pol.st(base, i, v). With PlainPol it is the forward kernel. With
// ProbePol the same body runs once per load k of the lane with that load
// returning 1 and every other load 0: since the body is linear in the loaded
// values (its weights depend on the topology only), the value it then stores
// to output o is the weight of load k in o. Recording these (o, input, weight)
// triplets once per topology gives the pass's matrix M, and the adjoint of
// the pass applies M^T:
//     x_in = x_out with the outputs o cleared, plus sum over triplets
//            (o, j, w) of w * x_out[o] added into x_in[j].
// Loads of a location the lane itself stored earlier return the stored
// (probe) value, as memory would. A pass's loads see the values from before
// the pass (no lane reads what another lane of the same pass writes); where
// two lanes write the same location they write the same function of the
// inputs, and one of them is kept.
//
// State arrays are registered in a StateMap: each base pointer is one region
// of a flat state vector, so (base, i) names a global index.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "pcg.h"

namespace astg {

constexpr int kRegions = 10;
constexpr int kMaxStores = 24;

struct StateMap {
    const double* base[kRegions];
    long long len[kRegions];
    long long off[kRegions];
    int n;
    __host__ __device__ long long glob(const double* b, size_t i) const {
        for (int r = 0; r < n; ++r)
            if (b == base[r]) return off[r] + (long long)i;
        return -1;
    }
};

struct PlainPol {
    static constexpr bool kAux = true;  // write outputs that are not state (matrix values)
    __device__ double ld(const double* b, size_t i) { return b[i]; }
    __device__ void st(double* b, size_t i, double v) { b[i] = v; }
};

struct ProbePol {
    static constexpr bool kAux = false;
    const StateMap* M;
    int target;       // the load that returns 1 (-1: none, a counting run)
    int cnt = 0;      // loads so far
    long long tgt = -1;
    int nst = 0;
    bool overflow = false;
    long long so[kMaxStores];
    double sv[kMaxStores];
    __device__ ProbePol(const StateMap* m, int k) : M(m), target(k) {}
    __device__ double ld(const double* b, size_t i) {
        const long long g = M->glob(b, i);
        if (g < 0) return b[i];  // not state: a constant of the pass
        for (int s = 0; s < nst; ++s)
            if (so[s] == g) return sv[s];
        const int c = cnt++;
        if (c == target) {
            tgt = g;
            return 1.0;
        }
        return 0.0;
    }
    __device__ void st(double* b, size_t i, double v) {
        const long long g = M->glob(b, i);
        if (g < 0) return;
        for (int s = 0; s < nst; ++s)
            if (so[s] == g) {
                sv[s] = v;
                return;
            }
        if (nst < kMaxStores) {
            so[nst] = g;
            sv[nst] = v;
            ++nst;
        } else {
            overflow = true;
        }
    }
};

template <class Body>
__global__ void __launch_bounds__(256) k_run(Body b, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    PlainPol P;
    b(t, P);
}

// Claims the outputs of every lane (the first lane to store a location owns
// it) and counts the owned nonzero triplets.
template <class Body>
__global__ void __launch_bounds__(256) k_rec_count(Body b, int n, StateMap M, int* claim, unsigned* own, int* nStore, int* nTrip,
                            int* bad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    ProbePol P0(&M, -1);
    b(t, P0);
    if (P0.overflow) atomicOr(bad, 1);
    unsigned mask = 0;
    for (int s = 0; s < P0.nst; ++s)
        if (atomicCAS(&claim[P0.so[s]], 0, 1) == 0) mask |= 1u << s;
    own[t] = mask;
    int ns = 0;
    for (int s = 0; s < P0.nst; ++s) ns += (mask >> s) & 1u;
    nStore[t] = ns;
    int nz = 0;
    for (int k = 0; k < P0.cnt; ++k) {
        ProbePol P(&M, k);
        b(t, P);
        for (int s = 0; s < P.nst; ++s)
            if (((mask >> s) & 1u) && P.sv[s] != 0.0) ++nz;
    }
    nTrip[t] = nz;
}

// Writes the owned outputs (list O) and the triplets (output position in O,
// input index, weight) at the offsets from the scans of the counts.
template <class Body>
__global__ void __launch_bounds__(256) k_rec_fill(Body b, int n, StateMap M, const unsigned* own, const long long* oStore,
                           const long long* oTrip, int* O, int* trO, int* trIn, double* trW) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const unsigned mask = own[t];
    ProbePol P0(&M, -1);
    b(t, P0);
    int pos[kMaxStores];
    int q = 0;
    for (int s = 0; s < P0.nst; ++s)
        if ((mask >> s) & 1u) {
            pos[s] = (int)(oStore[t] + q);
            O[pos[s]] = (int)P0.so[s];
            ++q;
        }
    long long e = oTrip[t];
    for (int k = 0; k < P0.cnt; ++k) {
        ProbePol P(&M, k);
        b(t, P);
        for (int s = 0; s < P.nst; ++s)
            if (((mask >> s) & 1u) && P.sv[s] != 0.0) {
                trO[e] = pos[s];
                trIn[e] = (int)P.tgt;
                trW[e] = P.sv[s];
                ++e;
            }
    }
}

// One recorded pass: its outputs and triplets on the device.
struct PassRecord {
    DevBuf O, trO, trIn, trW;
    long long nO = 0, nT = 0;
};

// Adjoint of one pass on the flat adjoint vector x (in place): the outputs'
// adjoints move to tmp and are cleared, then every triplet adds w * tmp[o]
// into its input.
__global__ void k_tr_take(const int* O, long long nO, double* x, double* tmp);
__global__ void k_tr_scatter(const int* trO, const int* trIn, const double* trW, long long nT, const double* tmp,
                             double* x);

// Host side: exclusive scans and the two recording launches.
int recordScan(const int* counts, int n, long long* offsets, long long& total, hipStream_t s, std::string& err);

template <class Body>
int recordPass(const Body& b, int n, const StateMap& M, long long stateLen, DevBuf& claim, DevBuf& scratch,
               hipStream_t s, PassRecord& out, std::string& err) {
    out.nO = out.nT = 0;
    if (n <= 0) return 0;
    const int bs = 128, nb = (n + bs - 1) / bs;
    if (claim.alloc((size_t)stateLen * 4) != hipSuccess) {
        err = "record: allocation failed";
        return ASTG_ERR_HIP;
    }
    // counts (own, nStore, nTrip, bad) then offsets
    const size_t head = ((size_t)n * 12 + 16 + 7) & ~(size_t)7;  // 8-byte aligned offsets follow
    const size_t need = head + (size_t)(n + 1) * 8 * 2;
    if (scratch.alloc(need) != hipSuccess) {
        err = "record: allocation failed";
        return ASTG_ERR_HIP;
    }
    char* p = static_cast<char*>(scratch.p);
    unsigned* own = reinterpret_cast<unsigned*>(p);
    int* nStore = reinterpret_cast<int*>(p + (size_t)n * 4);
    int* nTrip = reinterpret_cast<int*>(p + (size_t)n * 8);
    int* bad = reinterpret_cast<int*>(p + (size_t)n * 12);
    long long* oStore = reinterpret_cast<long long*>(p + head);
    long long* oTrip = oStore + (n + 1);
    if (hipMemsetAsync(claim.p, 0, (size_t)stateLen * 4, s) != hipSuccess || hipMemsetAsync(bad, 0, 4, s) != hipSuccess) {
        err = "record: memset failed";
        return ASTG_ERR_HIP;
    }
    k_rec_count<<<nb, bs, 0, s>>>(b, n, M, claim.as<int>(), own, nStore, nTrip, bad);
    int rc = recordScan(nStore, n, oStore, out.nO, s, err);
    if (rc) return rc;
    rc = recordScan(nTrip, n, oTrip, out.nT, s, err);
    if (rc) return rc;
    int hb = 0;
    if (hipMemcpyAsync(&hb, bad, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        err = "record: count pass failed";
        return ASTG_ERR_HIP;
    }
    if (hb) {
        err = "record: a lane stores more values than the probe keeps";
        return ASTG_ERR_STATE;
    }
    if (out.O.alloc((size_t)std::max(out.nO, 1LL) * 4) != hipSuccess ||
        out.trO.alloc((size_t)std::max(out.nT, 1LL) * 4) != hipSuccess ||
        out.trIn.alloc((size_t)std::max(out.nT, 1LL) * 4) != hipSuccess ||
        out.trW.alloc((size_t)std::max(out.nT, 1LL) * 8) != hipSuccess) {
        err = "record: allocation failed";
        return ASTG_ERR_HIP;
    }
    k_rec_fill<<<nb, bs, 0, s>>>(b, n, M, own, oStore, oTrip, out.O.as<int>(), out.trO.as<int>(), out.trIn.as<int>(),
                                 out.trW.as<double>());
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        err = "record: fill pass failed";
        return ASTG_ERR_HIP;
    }
    return 0;
}

// Applies the adjoint of a recorded pass to x (tmp: at least nO doubles).
int applyTransposed(const PassRecord& r, double* x, double* tmp, hipStream_t s, std::string& err);

}  // namespace astg
