// Recorded linear passes: scans of the record counts and the transposed
// application (see linrec.h).
#include <hipcub/hipcub.hpp>

#include "linrec.h"

namespace astg {

__global__ void k_tr_take(const int* O, long long nO, double* x, double* tmp) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nO) return;
    const int o = O[k];
    tmp[k] = x[o];
    x[o] = 0.0;
}

__global__ void k_tr_scatter(const int* trO, const int* trIn, const double* trW, long long nT, const double* tmp,
                             double* x) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nT) return;
    const double v = trW[e] * tmp[trO[e]];
    if (v != 0.0) atomicAdd(&x[trIn[e]], v);
}

int recordScan(const int* counts, int n, long long* offsets, long long& total, hipStream_t s, std::string& err) {
    size_t bytes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, counts, offsets, n, s) != hipSuccess) {
        err = "record: scan sizing failed";
        return ASTG_ERR_HIP;
    }
    DevBuf tmp;
    if (tmp.alloc(bytes + 16) != hipSuccess ||
        hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, counts, offsets, n, s) != hipSuccess) {
        err = "record: scan failed";
        return ASTG_ERR_HIP;
    }
    long long last = 0;
    int lastCount = 0;
    if (hipMemcpyAsync(&last, offsets + (n - 1), 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&lastCount, counts + (n - 1), 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        err = "record: scan read-back failed";
        return ASTG_ERR_HIP;
    }
    total = last + lastCount;
    return 0;
}

int applyTransposed(const PassRecord& r, double* x, double* tmp, hipStream_t s, std::string& err) {
    const int bs = 256;
    if (r.nO > 0)
        k_tr_take<<<(int)((r.nO + bs - 1) / bs), bs, 0, s>>>(r.O.as<int>(), r.nO, x, tmp);
    if (r.nT > 0)
        k_tr_scatter<<<(int)((r.nT + bs - 1) / bs), bs, 0, s>>>(r.trO.as<int>(), r.trIn.as<int>(),
                                                                 r.trW.as<double>(), r.nT, tmp, x);
    if (hipGetLastError() != hipSuccess) {
        err = "transposed pass: launch failed";
        return ASTG_ERR_HIP;
    }
    return 0;
}

}  // namespace astg
