// Symbolic plans of the factorisation (mic_plan.h) and of the merged sweep
// stencils (sweep_merge.h) built on the device. Each is a count pass (one
// lane per row or sweep position), a prefix sum of the counts and a fill pass
// that repeats the count pass's walk and writes at the scanned offsets, so the
// arrays equal the host planners' entry for entry. Pattern arrays are read
// from the copies the solver already holds on the device.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>

#include "device_common.h"
#include "pcg.h"
#include "pcg_common.h"
#include "plan_kernels.h"
#include "sweep_merge.h"

namespace astg {

namespace {

inline int grid(long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

using namespace plan;

int PcgDevice::planMicDevice(std::string& err) {
    const int n = n_;
    const long long nnz = nnz_;
    DevBuf len, tmp;
    HIPCHK(len.alloc((size_t)nnz * 8));
    HIPCHK(micStart_.alloc(((size_t)nnz + 1) * 8));
    long long* start = micStart_.as<long long>();
    k_plan_mic<false><<<grid(n, 256), 256, 0, stream_>>>(n, rs_.as<int>(), lend_.as<int>(), ubeg_.as<int>(),
                                                         col_.as<int>(), len.as<long long>(), nullptr, nullptr);
    HIPCHK(hipGetLastError());
    size_t tmpBytes = 0;
    HIPCHK(hipcub::DeviceScan::InclusiveSum(nullptr, tmpBytes, len.as<long long>(), start + 1, (int)nnz, stream_));
    HIPCHK(tmp.alloc(tmpBytes));
    HIPCHK(hipcub::DeviceScan::InclusiveSum(tmp.p, tmpBytes, len.as<long long>(), start + 1, (int)nnz, stream_));
    k_set_i64<<<1, 64, 0, stream_>>>(start, 0);
    long long total = 0;
    HIPCHK(hipMemcpyAsync(&total, start + nnz, 8, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    HIPCHK(micProg_.alloc((size_t)total * 4));
    k_plan_mic<true><<<grid(n, 256), 256, 0, stream_>>>(n, rs_.as<int>(), lend_.as<int>(), ubeg_.as<int>(),
                                                        col_.as<int>(), nullptr, start, micProg_.as<int>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));  // the temporaries are freed on return
    return ASTG_OK;
}

int PcgDevice::planStencilDevice(int dir, int maxY, DevStencil& D, std::string& err) {
    if (maxY > kMaxDeps) {
        err = "sweep stencil: maxY above the dependency limit";
        return ASTG_ERR_ARG;
    }
    const int P = PS_, n = n_;
    DevBuf pos, cI, cU, tmp, ebuf;
    HIPCHK(pos.alloc((size_t)n * 4));
    HIPCHK(cI.alloc((size_t)P * 4));
    HIPCHK(cU.alloc((size_t)P * 4));
    HIPCHK(ebuf.alloc(4));
    HIPCHK(D.start.alloc(((size_t)P + 1) * 4));
    HIPCHK(D.ustart.alloc(((size_t)P + 1) * 4));
    HIPCHK(D.nr.alloc((size_t)P * 4));
    k_set_i32<<<1, 64, 0, stream_>>>(ebuf.as<int>(), 0);
    k_sweep_pos<<<grid(P, 256), 256, 0, stream_>>>(P, sorder_.as<int>(), pos.as<int>());
    auto kc = dir == 0 ? k_plan_stencil<0, false> : k_plan_stencil<1, false>;
    kc<<<grid(P, 64), 64, 0, stream_>>>(P, rs_.as<int>(), lend_.as<int>(), ubeg_.as<int>(), col_.as<int>(),
                                        lev_.as<int>(), merge_, maxY, sorder_.as<int>(), pos.as<int>(), cI.as<int>(),
                                        cU.as<int>(), D.nr.as<int>(), nullptr, nullptr, nullptr, nullptr,
                                        ebuf.as<int>());
    HIPCHK(hipGetLastError());
    size_t b1 = 0, b2 = 0;
    HIPCHK(hipcub::DeviceScan::InclusiveSum(nullptr, b1, cI.as<int>(), D.start.as<int>() + 1, P, stream_));
    HIPCHK(hipcub::DeviceScan::InclusiveSum(nullptr, b2, cU.as<int>(), D.ustart.as<int>() + 1, P, stream_));
    HIPCHK(tmp.alloc(std::max(b1, b2)));
    HIPCHK(hipcub::DeviceScan::InclusiveSum(tmp.p, b1, cI.as<int>(), D.start.as<int>() + 1, P, stream_));
    HIPCHK(hipcub::DeviceScan::InclusiveSum(tmp.p, b2, cU.as<int>(), D.ustart.as<int>() + 1, P, stream_));
    k_set_i32<<<1, 64, 0, stream_>>>(D.start.as<int>(), 0);
    k_set_i32<<<1, 64, 0, stream_>>>(D.ustart.as<int>(), 0);
    int tot[3] = {0, 0, 0};
    HIPCHK(hipMemcpyAsync(&tot[0], D.start.as<int>() + P, 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(&tot[1], D.ustart.as<int>() + P, 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(&tot[2], ebuf.p, 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    if (tot[2]) {
        err = "sweep stencil: a row exceeds the key bound";
        return ASTG_ERR_HIP;
    }
    D.E = tot[0];
    HIPCHK(D.idx.alloc((size_t)D.E * 4));
    HIPCHK(D.upd.alloc((size_t)tot[1]));
    HIPCHK(D.coef.alloc((size_t)D.E * 8));
    auto kf = dir == 0 ? k_plan_stencil<0, true> : k_plan_stencil<1, true>;
    kf<<<grid(P, 64), 64, 0, stream_>>>(P, rs_.as<int>(), lend_.as<int>(), ubeg_.as<int>(), col_.as<int>(),
                                        lev_.as<int>(), merge_, maxY, sorder_.as<int>(), pos.as<int>(), nullptr,
                                        nullptr, nullptr, D.start.as<int>(), D.ustart.as<int>(), D.idx.as<int>(),
                                        D.upd.as<uint8_t>(), ebuf.as<int>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));  // the temporaries are freed on return
    return ASTG_OK;
}

}  // namespace astg
