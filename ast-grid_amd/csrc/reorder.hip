// Block-colour ordering of the projection system (device part): moving the
// assembled values, the right-hand side and the solution between the
// reference numbering and the block-colour one.
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

}  // namespace

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
