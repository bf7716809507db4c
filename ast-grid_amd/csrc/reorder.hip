// Block-colour ordering of the projection system (device part): moving the
// assembled values, the right-hand side and the solution between the
// reference numbering and the block-colour one.
#include <hipcub/hipcub.hpp>

#include "../../include/astg.h"
#include "pcg_common.h"
#include "reorder.h"

namespace astg {
namespace {

constexpr int kBS = 256;

__global__ void k_gather_vals_perm(long long nnz, const int* src, const double* in, double* out) {
    const long long e = (long long)blockIdx.x * kBS + threadIdx.x;
    if (e < nnz) out[e] = in[src[e]];
}

__global__ void k_gather_rows(int n, const int* perm, const double* in, double* out) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) out[i] = in[perm[i]];
}

__global__ void k_scatter_rows(int n, const int* perm, const double* in, double* out) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) out[perm[i]] = in[i];
}

__global__ void k_iota_i(int n, int* p) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) p[i] = i;
}

// inv[perm[i]] = i; len[i + 1] = length of row perm[i] of A (len[0] = 0)
__global__ void k_inv_len(int n, const int* perm, const int* rs, int* inv, int* len) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i == 0) len[0] = 0;
    if (i >= n) return;
    const int r = perm[i];
    inv[r] = i;
    len[i + 1] = rs[r + 1] - rs[r];
}

constexpr int kMaxRow = 64;

// Row i of P A P^T: row perm[i] of A with its columns renamed by inv and put
// in increasing order (a column's rank among the row's, columns distinct),
// src the entry of A each comes from.
__global__ void k_perm_rows(int n, const int* perm, const int* inv, const int* rs, const int* col, const int* nrs,
                            int* ncol, int* src, int* flag) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    const int r = perm[i], p0 = rs[r], len = rs[r + 1] - p0;
    if (len > kMaxRow) {
        atomicOr(flag, 1);
        return;
    }
    int m[kMaxRow];
    for (int k = 0; k < len; ++k) m[k] = inv[col[p0 + k]];
    const int base = nrs[i];
    for (int k = 0; k < len; ++k) {
        int rank = 0;
        for (int j = 0; j < len; ++j) rank += m[j] < m[k] ? 1 : 0;
        ncol[base + rank] = m[k];
        src[base + rank] = p0 + k;
    }
}

}  // namespace

int BlockOrderDevice::build(hipStream_t stream, int n_, long long nnz_, const int* hostKeys, const int* rs,
                            const int* colA, std::string& err, bool keysOnDevice) {
    n = n_;
    nnz = nnz_;
    const size_t n4 = (size_t)n * 4;
    const unsigned gb = (unsigned)((n + kBS - 1) / kBS);
    HIPCHK(keys.alloc(n4));
    HIPCHK(sortedKeys.alloc(n4));
    HIPCHK(iota.alloc(n4));
    HIPCHK(perm.alloc(n4));
    HIPCHK(inv.alloc(n4));
    HIPCHK(rowstart.alloc(n4 + 4));
    HIPCHK(col.alloc((size_t)nnz * 4));
    HIPCHK(src.alloc((size_t)nnz * 4));
    HIPCHK(flag.alloc(4));
    HIPCHK(hipMemcpyAsync(keys.p, hostKeys, n4, keysOnDevice ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                          stream));
    HIPCHK(hipMemsetAsync(flag.p, 0, 4, stream));
    k_iota_i<<<gb, kBS, 0, stream>>>(n, iota.as<int>());
    // stable sort by key: ties keep the reference index order
    size_t tb = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.as<int>(), sortedKeys.as<int>(), iota.as<int>(),
                                              perm.as<int>(), n, 0, 32, stream));
    HIPCHK(tmp.alloc(tb));
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, keys.as<int>(), sortedKeys.as<int>(), iota.as<int>(),
                                              perm.as<int>(), n, 0, 32, stream));
    k_inv_len<<<gb, kBS, 0, stream>>>(n, perm.as<int>(), rs, inv.as<int>(), rowstart.as<int>());
    size_t sb = 0;
    HIPCHK(hipcub::DeviceScan::InclusiveSum(nullptr, sb, rowstart.as<int>() + 1, rowstart.as<int>() + 1, n, stream));
    HIPCHK(tmp.alloc(sb));
    HIPCHK(hipcub::DeviceScan::InclusiveSum(tmp.p, sb, rowstart.as<int>() + 1, rowstart.as<int>() + 1, n, stream));
    k_perm_rows<<<gb, kBS, 0, stream>>>(n, perm.as<int>(), inv.as<int>(), rs, colA, rowstart.as<int>(), col.as<int>(),
                                        src.as<int>(), flag.as<int>());
    HIPCHK(hipGetLastError());
    int hf = 0;
    HIPCHK(hipMemcpyAsync(&hf, flag.p, 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (hf) {
        err = "block order: row wider than 64 entries";
        return ASTG_ERR_ARG;
    }
    return ASTG_OK;
}

int gatherPermuted(hipStream_t stream, long long nnz, const int* src, const double* valIn, double* valOut, int n,
                   const int* perm, const double* bIn, double* bOut, std::string& err) {
    if (valIn && nnz > 0)
        k_gather_vals_perm<<<(unsigned)((nnz + kBS - 1) / kBS), kBS, 0, stream>>>(nnz, src, valIn, valOut);
    if (bIn && n > 0) k_gather_rows<<<(n + kBS - 1) / kBS, kBS, 0, stream>>>(n, perm, bIn, bOut);
    HIPCHK(hipGetLastError());
    return ASTG_OK;
}

int scatterPermuted(hipStream_t stream, int n, const int* perm, const double* xIn, double* xOut, std::string& err) {
    if (n > 0) k_scatter_rows<<<(n + kBS - 1) / kBS, kBS, 0, stream>>>(n, perm, xIn, xOut);
    HIPCHK(hipGetLastError());
    return ASTG_OK;
}

}  // namespace astg
