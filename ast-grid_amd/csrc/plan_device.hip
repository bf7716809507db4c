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

int PcgDevice::planDevice(int maxY, std::string& err) {
    if (maxY > kMaxDeps) {
        err = "sweep stencil: maxY above the dependency limit";
        return ASTG_ERR_ARG;
    }
    const int n = n_, P = PS_;
    const long long nnz = nnz_;
    const int* rs = rs_.as<int>();
    const int* lend = lend_.as<int>();
    const int* ubeg = ubeg_.as<int>();
    const int* col = col_.as<int>();
    const int* lev = lev_.as<int>();
    const int* order = sorder_.as<int>();
    int* cnt = cnt_.as<int>();
    HIPCHK(planLen_.alloc((size_t)nnz * 8));
    HIPCHK(planPos_.alloc((size_t)n * 4));
    HIPCHK(planCnt_.alloc((size_t)P * 16));  // per direction: entry and slot counts
    HIPCHK(micStart_.alloc(((size_t)nnz + 1) * 8));
    DevStencil* dst[2] = {&stF_, &stB_};
    int* cI[2] = {planCnt_.as<int>(), planCnt_.as<int>() + 2 * (size_t)P};
    int* cU[2] = {planCnt_.as<int>() + (size_t)P, planCnt_.as<int>() + 3 * (size_t)P};
    for (DevStencil* D : dst) {
        HIPCHK(D->start.alloc(((size_t)P + 1) * 4));
        HIPCHK(D->ustart.alloc(((size_t)P + 1) * 4));
        HIPCHK(D->nr.alloc((size_t)P * 4));
    }
    long long* mstart = micStart_.as<long long>();
    size_t b64 = 0, b32 = 0;
    HIPCHK(hipcub::DeviceScan::InclusiveSum(nullptr, b64, planLen_.as<long long>(), mstart + 1, (int)nnz, stream_));
    HIPCHK(hipcub::DeviceScan::InclusiveSum(nullptr, b32, cI[0], dst[0]->start.as<int>() + 1, P, stream_));
    size_t tb = std::max(b64, b32);
    HIPCHK(planTmp_.alloc(tb));
    // count passes and prefix sums
    k_plan_mic<false><<<grid(n, 256), 256, 0, stream_>>>(n, rs, lend, ubeg, col, planLen_.as<long long>(), nullptr,
                                                         nullptr);
    HIPCHK(hipcub::DeviceScan::InclusiveSum(planTmp_.p, tb, planLen_.as<long long>(), mstart + 1, (int)nnz,
                                            stream_));
    k_set_i64<<<1, 64, 0, stream_>>>(mstart, 0);
    k_sweep_pos<<<grid(P, 256), 256, 0, stream_>>>(P, order, planPos_.as<int>());
    for (int dir = 0; dir < 2; ++dir) {
        DevStencil& D = *dst[dir];
        auto kc = dir == 0 ? k_plan_stencil<0, false> : k_plan_stencil<1, false>;
        kc<<<grid(P, 64), 64, 0, stream_>>>(P, rs, lend, ubeg, col, lev, merge_, maxY, order, planPos_.as<int>(),
                                            cI[dir], cU[dir], D.nr.as<int>(), nullptr, nullptr, nullptr, nullptr,
                                            cnt + C_PATTERN);
        HIPCHK(hipcub::DeviceScan::InclusiveSum(planTmp_.p, tb, cI[dir], D.start.as<int>() + 1, P, stream_));
        HIPCHK(hipcub::DeviceScan::InclusiveSum(planTmp_.p, tb, cU[dir], D.ustart.as<int>() + 1, P, stream_));
        k_set_i32<<<1, 64, 0, stream_>>>(D.start.as<int>(), 0);
        k_set_i32<<<1, 64, 0, stream_>>>(D.ustart.as<int>(), 0);
    }
    HIPCHK(hipGetLastError());
    // sizes: program length, entries and slots per direction, error flag
    int* hi = reinterpret_cast<int*>(hplan_ + 1);
    HIPCHK(hipMemcpyAsync(hplan_, mstart + nnz, 8, hipMemcpyDeviceToHost, stream_));
    for (int dir = 0; dir < 2; ++dir) {
        HIPCHK(hipMemcpyAsync(hi + 2 * dir, dst[dir]->start.as<int>() + P, 4, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipMemcpyAsync(hi + 2 * dir + 1, dst[dir]->ustart.as<int>() + P, 4, hipMemcpyDeviceToHost, stream_));
    }
    HIPCHK(hipMemcpyAsync(hi + 4, cnt + C_PATTERN, 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    if (hi[4]) {
        err = "sweep stencil: a row exceeds the key bound";
        return ASTG_ERR_HIP;
    }
    // fill passes
    HIPCHK(micProg_.alloc((size_t)hplan_[0] * 4));
    k_plan_mic<true><<<grid(n, 256), 256, 0, stream_>>>(n, rs, lend, ubeg, col, nullptr, mstart, micProg_.as<int>());
    for (int dir = 0; dir < 2; ++dir) {
        DevStencil& D = *dst[dir];
        D.E = hi[2 * dir];
        HIPCHK(D.idx.alloc((size_t)D.E * 4));
        HIPCHK(D.upd.alloc((size_t)hi[2 * dir + 1]));
        HIPCHK(D.coef.alloc((size_t)D.E * 8));
        auto kf = dir == 0 ? k_plan_stencil<0, true> : k_plan_stencil<1, true>;
        kf<<<grid(P, 64), 64, 0, stream_>>>(P, rs, lend, ubeg, col, lev, merge_, maxY, order, planPos_.as<int>(),
                                            nullptr, nullptr, nullptr, D.start.as<int>(), D.ustart.as<int>(),
                                            D.idx.as<int>(), D.upd.as<uint8_t>(), cnt + C_PATTERN);
    }
    HIPCHK(hipGetLastError());
    return ASTG_OK;
}

}  // namespace astg
