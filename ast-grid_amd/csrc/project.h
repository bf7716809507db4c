// Device-side projection driver: owns the stencil tables and velocity fields on
// the device; assembly and velocity correction kernels live in project.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "consts.h"
#include "linrec.h"
#include "pcg.h"
#include "project_host.h"

namespace astg {

struct AsmPasses;
struct VelPasses;

class ProjectDevice {
public:
    explicit ProjectDevice(hipStream_t s) : stream_(s) {}
    int upload(const ProjectTables& t, const Octree& o, std::string& err);
    // Writes A's values (CSR of the tables' pattern) and b (scaled by rho/dt).
    int assemble(double* A, double* b, const int* rs, double rho, double dt, std::string& err);
    // u -= dt/rho * grad p on MAC, tilt and half-tilt slots.
    int correct(const double* p, double rho, double dt, std::string& err);
    // Device-side copy of the velocity fields (restore = false saves).
    int snapshot(bool restore, std::string& err);
    double assembleBytes() const;
    // Exchanges the live velocity with the advection's output pair (the
    // advected velocity becomes live, the old one the next output buffer).
    void swapVelocity() {
        std::swap(u_mac_, t_mac_);
        std::swap(u_tilt_, t_tilt_);
    }
    double* uMac() { return u_mac_.as<double>(); }    // 3*G, axis-major
    double* uTilt() { return u_tilt_.as<double>(); }  // 8*G, slot-major
    double* tMac() { return t_mac_.as<double>(); }    // advection output, as uMac
    double* tTilt() { return t_tilt_.as<double>(); }  // advection output, as uTilt
    // The assembly's b part and the velocity correction as linear passes
    // (project.hip): their bodies with this context's arrays.
    void asmPasses(double* A, double* b, const int* rs, double rho, double dt, AsmPasses& ps);
    void velPasses(const double* p, double rho, double dt, VelPasses& ps);
    // Records the transposes of the assembly's b part and of the velocity
    // correction (linrec.h) in the flat state M.
    int recordPasses(const StateMap& M, long long stateLen, double* A, double* b, const int* rs, const double* p,
                     double rho, double dt, DevBuf& claim, DevBuf& scratch, std::vector<PassRecord>& asmRec,
                     std::vector<PassRecord>& velRec, std::string& err);

private:
    // level spacings in the kernels' constant table: uploaded by upload(),
    // again whenever another context used the table in between
    int bindConstants(bool force, std::string& err);
    hipStream_t stream_;
    uint64_t constOwner_ = newConstOwner();
    double hConst_[16] = {};
    const ProjectTables* T_ = nullptr;
    DevBuf level_, tilt_e_, tilt_bits_, tilt_hd_, pid_tilt_, w_mac_, w_tilt_;
    DevBuf oct_row_, oct_corner_, oct_face_, oct_pos_;
    DevBuf tilt_row_, tilt_node_, tilt_nb_, tilt_nblev_, tilt_pos_;
    DevBuf half_node_, half_c_, half_t_, half_closed_, half_nb_;
    DevBuf hc_row_, hc_start_, hc_q_, hc_pos_diag_, hc_pos_off_;
    DevBuf hr_row_, hr_start_, hr_q_, hr_pos_;
    DevBuf vm_node_, vm_c_, vm_nb_, vt_node_, vt_c_, vt_nb_, vt_lev_;
    DevBuf u_mac_, u_tilt_, t_mac_, t_tilt_, snap_mac_, snap_tilt_;
    long long velG_ = -1;  // node count the four velocity buffers are sized for
    bool haveSnap_ = false;
};

}  // namespace astg
