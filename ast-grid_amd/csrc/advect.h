// Device-side step around the projection on the AST octree: wall constraint
// of the MAC velocity, ghost updates of the omni fields, the dual (cell-centre
// and vertex) velocity grid, point sampling and semi-Lagrangian advection of
// the scalar fields (density, heat) and of the velocity itself.
//
// Reference (file:line into the AST-Grid sources):
//   onUpdate chain after project       src/octree_omnigrid_solver3.cpp:234-246
//   constrainVelocity (no collider)    include/omni_boundary_solver3.h:26-162
//   omni face ghost update             include/octree_omni_face_centered_grid3.h:97-120
//   updateDualGrid                     src/octree_omni_face_centered_grid3.cpp:32-159
//   omni grid ghost update             include/octree_omni_grid3.h:136-143, 287-337
//   point location                     src/octree_grid_layout3.cpp:16-41
//   MAC / omni velocity sampling       src/octree_face_centered_grid3.cpp:36-113,
//                                      include/octree_omni_face_centered_grid3.h:37-80
//   omni scalar sampling / sampleSub   include/octree_omni_grid3.h:146-286
//   getNeighborCubeValue               include/octree_util.h:124-185
//   BackTrace                          include/advection.h:15-64 (midpoint branch)
//   advect                             src/octree_omnigrid_solver3.cpp:850-914,
//                                      include/advection.h:195-239
//   adjustE (decision)                 src/octree_omnigrid_solver3.cpp:1125-1149
//   addForce                           src/octree_omnigrid_solver3.cpp:917-972
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "consts.h"
#include "linrec.h"
#include "octree.h"
#include "pcg.h"
#include "project.h"
#include "project_host.h"

namespace astg {

// astg_sample_points_noalloc: largest point count of one call, and the device
// ordinals that can hold the staging array's address.
constexpr int kNoAllocPoints = 1 << 18;
constexpr int kMaxDevices = 64;
// Looks up this GPU's copy of the staging array (hipGetSymbolAddress) once per
// device, with `device` current; later calls return at once.
int noallocScratchInit(int device, std::string& err);
// The address recorded for `device`, or null before noallocScratchInit.
double* noallocScratch(int device);

// Device views passed to the kernels.
struct AdvTopo {
    const int* map;                 // concatenated dense index maps -> global node, -1
    const uint16_t *nx, *ny, *nz;   // node coordinates at their level
    const uint8_t* nf;              // NF_* node flags
    const double* te;               // tilt size e
    const uint8_t* tb;              // TB_* tilt bits
    const int8_t* thd;              // half-tilt direction
    const double* sdf;              // boundary SDF on the finest grid (R^3 cell centres), null: no collider
};
struct AdvVel {
    const double* mac;    // 3G axis-major
    const double* tilt;   // 8G slot-major
    const double* doct;   // dual cell vectors, 3 per node
    const double* dtilt;  // dual vertex vectors, 3 per node
};

class AdvectDevice {
public:
    explicit AdvectDevice(hipStream_t s) : stream_(s) {}
    // keepFields: the node sets are unchanged (a tilt update); the fields stay
    // and keep being sampled with the previous tilt state until advect.
    int upload(const Octree& o, const ProjectTables& t, bool keepFields, std::string& err);

    // Scalar fields: 0 density, 1 heat; per level, octagon and tilt value per node.
    int setScalar(int field, int level, const double* oct, const double* tilt, std::string& err);
    int getScalar(int field, int level, double* oct, double* tilt, std::string& err);
    int getDual(int level, double* oct3, double* tilt3, std::string& err);

    // MAC wall constraint and velocity ghosts (onUpdate before project).
    int constrainAndGhost(ProjectDevice& pd, std::string& err, bool constrain = true);
    // onUpdate between project and advect: the above, then the dual grid and
    // its ghosts.
    int postProject(ProjectDevice& pd, std::string& err, bool constrain = true);
    // The fields now carry the current tilt state (they were just written).
    void fieldsCurrent() { sampValid_ = false; }
    // Copies the current scalar fields (field f when scalars[f]) and, when
    // velocity is set, the current velocity into the advection destinations,
    // as restart loads tDensity / tHeat / tVelocity with the same values.
    int mirrorTemporaries(ProjectDevice& pd, const bool scalars[2], bool velocity, std::string& err);
    // Ghost update of one scalar field (omni grid).
    int scalarGhost(int field, std::string& err);
    // advect(dt) with tilt adaptivity off: scalars, their ghosts, velocity,
    // then the 0.999 decay of both scalars.
    int advect(ProjectDevice& pd, double dt, std::string& err);
    // Point queries on host arrays: what = 0 omni velocity, 1 MAC velocity,
    // 2 back-trace over dt, 3 density, 4 heat (out: 3 per point, or 1).
    int samplePoints(ProjectDevice& pd, int what, int n, const double* xyz, double dt, double* out,
                     std::string& err);
    // The same queries staged through `stage` (device memory of at least
    // 6 * kNoAllocPoints doubles: points, then results); allocates nothing.
    int samplePointsStaged(ProjectDevice& pd, int what, int n, const double* xyz, double dt, double* out,
                           double* stage, std::string& err);
    double advectBytes() const;
    // adjustE's decision (octree_omnigrid_solver3.cpp:1131-1149): the new
    // tilt_layout open flag of every leaf, levels ascending, leaf order.
    int decideTilts(ProjectDevice& pd, double dThr, double dt, double vThr, double maxThr, std::vector<uint8_t>& open,
                    std::string& err);
    int leaves(int level) const { return nleaf_[level]; }
    // Current buffer of a scalar field: field 0 density, 1 heat; kind 0 cell, 1 vertex (G values).
    double* scalarPtr(int field, int kind) { return scal_[field][cur_[field]][kind].as<double>(); }
    // Sphere smoke source of the octree smoke example (sphereSource,
    // example/octree_smoke_3d/main.cpp:23-81) with parameters; ghost: update
    // the scalar ghosts afterwards (the example does on its first call).
    int applySource(ProjectDevice& pd, double cx, double cy, double cz, double r, double dval, double hval,
                    double speed, bool ghost, std::string& err);
    // The same source with an inflow vector v (3 doubles) on every MAC face and
    // tilt slot inside the sphere; v = null is applySource with `speed`.
    int applySourceV(ProjectDevice& pd, double cx, double cy, double cz, double r, double dval, double hval,
                     const double* v, double speed, bool ghost, std::string& err);
    // addForce (octree_omnigrid_solver3.cpp:917-972): buoyancy along +z.
    int addForce(ProjectDevice& pd, double dt, double fd, double ft, std::string& err);
    // wipeBoundary (octree_omnigrid_solver3.cpp:1240-1292): density and heat
    // are zeroed where the boundary SDF is negative (no-op without a collider).
    int wipeBoundary(std::string& err);
    bool hasCollider() const { return hasSdf_; }

    // ---- differentiable step (frozen topology, no collider) ----
    // Flat state of a step: MAC 3G | tilt 8G | dual cell 3G | dual vertex 3G |
    // density cell G | density vertex G; the adjoint work vector appends the
    // projection's b and p (n each).
    long long diffStateLen() const { return 19LL * G_; }
    // Records the transposes of the step's fixed linear passes (once per
    // topology); pcg holds the projection system (n unknowns, b, p).
    int diffPrepare(ProjectDevice& pd, PcgDevice& pcg, double dt, double rho, std::string& err);
    bool diffReady() const { return diffReady_; }
    void diffInvalidate() { diffReady_ = false; }
    // Copies the current state to / from a flat device vector (stream order).
    int diffGetState(ProjectDevice& pd, double* dst, std::string& err);
    int diffSetState(ProjectDevice& pd, const double* src, std::string& err);
    // advect(dt) in which the entries the step does not write keep their
    // current value (the differentiable step's forward).
    int advectCarry(ProjectDevice& pd, double dt, std::string& err);
    // Adjoint of advect (with the decay and the density ghosts): x holds the
    // output adjoint on entry (first 19G entries) and the adjoint of the
    // post-projection state (velocity, dual grid) and of the input density
    // on exit. Needs the post-projection fields of the step in the context.
    int diffAdjointAdvect(ProjectDevice& pd, double dt, std::string& err);
    // Applies the transposes of the post-projection chain / the constraint
    // and ghosts before the projection to the work vector.
    int diffAdjointPost(std::string& err);
    int diffAdjointPre(std::string& err);
    // Transposes of the velocity correction (p's adjoint gathers in the p
    // region) and of the assembly of b (b's adjoint spreads to the velocity).
    int diffAdjointVel(std::string& err) { return applyList(recVel_, err); }
    int diffAdjointAsm(std::string& err) { return applyList(recAsm_, err); }
    int diffAdjointGhost1(std::string& err) { return applyList(recG1_, err); }
    // The work vector and the offsets of the regions of b and p in it.
    double* diffWork() { return work_.as<double>(); }
    long long diffOffB() const { return offB_; }
    long long diffOffP() const { return offP_; }
    double* diffTmp() { return wtmp_.as<double>(); }

private:
    // This context's values of the kernels' constant tables (level
    // resolutions, map offsets, spacings, slip centre, dual-grid matrices):
    // uploaded by upload(), again whenever another context used the tables.
    int bindConstants(bool force, std::string& err);
    uint64_t constOwner_ = newConstOwner();
    std::vector<int> cRes_, cMoff_;
    std::vector<double> cHl_, cDual_;
    double cSlip_[3] = {0.5, 0.5, 0.45};
    int omniGhost(double* cell, double* vert, int nc, std::string& err);
    int checkBad(std::string& err);
    AdvTopo topo() const;      // current tilt state
    AdvTopo sampTopo() const;  // tilt state the fields were written with
    AdvVel vel(ProjectDevice& pd) const;
    hipStream_t stream_;
    int L_ = 0, R_ = 0, G_ = 0;
    double h0_ = 0;
    int closed_ = 0;
    std::vector<int> off_, nleaf_, nnode_;
    int nLeaf_ = 0, nGhost_ = 0, nHalf_ = 0, nTJ_ = 0, nCBH_ = 0, nOpenTilt_ = 0, nVelTilt_ = 0, nHalfMac_ = 0;
    DevBuf map_, nx_, ny_, nz_, nflag_, te_, tb_, thd_, level_, inHalf_;
    std::vector<uint16_t> hnx_, hny_, hnz_;  // node coordinates on the host (kept across tilt updates)
    DevBuf leaf_, ghost_, half_, tj_, cbh_, openTilt_, velTilt_;
    DevBuf hmG_, hmT_, hmA_, hmS_;   // coarse MAC faces written from half-tilts
    DevBuf bad_;
    DevBuf scal_[2][2][2];           // [field][buffer][0 octagon, 1 tilt]
    int cur_[2] = {0, 0};            // current buffer of each scalar field
    DevBuf dualOct_, dualTilt_;      // 3 per node
    DevBuf pts_, out_;               // point queries
    DevBuf sampTe_, sampTb_, sampThd_, open_;  // tilt state before an update, decisions
    bool sampValid_ = false;
    // collider: SDF grid, MAC faces inside it in the reference's visiting order
    // (faceG node, faceA axis), their order index per (node, axis) (-1 if
    // outside), and the published new values of the serial slip constraint
    bool hasSdf_ = false;
    // differentiable step: recorded passes, masks of written entries, work vectors
    bool diffReady_ = false;
    std::vector<PassRecord> recCG_, recDual_, recG3_, recG1_, recAsm_, recVel_;
    DevBuf wMac_, wTilt_, wCell_, wVert_;  // 1 where advect writes the entry
    DevBuf work_, wtmp_, wadv_, recClaim_, recScratch_;
    long long offB_ = 0, offP_ = 0, workLen_ = 0;
    int applyList(std::vector<PassRecord>& recs, std::string& err);
    DevBuf sdf_, slipFaceG_, slipFaceA_, slipOrd_, slipPub_, slipErr_;
    int nSlip_ = 0, slipGrid_ = 0;
    int slipMac(ProjectDevice& pd, std::string& err);
};

}  // namespace astg
