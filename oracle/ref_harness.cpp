// Checker-side harness around the reference solver (TEST INFRASTRUCTURE ONLY).
//
// This file is compiled together with the reference's own sources, where they
// lie under the reference tree, by oracle/ref.mk into oracle/_ref/libastgref.so.
// Nothing in the product links or loads it: tests/golden/make_golden.py uses it
// to write fixtures, the CPU tests use it (when present) to pin the oracle, and
// bench.py may time it as the reference CPU baseline.
//
// It drives OctreeOmniGridSolver3 (include/octree_omnigrid_solver3.h:23-129)
// through a subclass that
//   * fills the split map from a caller-provided array (splitMapSetup callback,
//     octree_omnigrid_solver3.h:36, consumed at octree_omnigrid_solver3.cpp:67),
//   * replaces initObstacle (octree_omnigrid_solver3.cpp:152-219) by closed-wall
//     flags plus a constant +1e4 boundary SDF (no collider),
//   * exposes layout / tilt / numbering / matrix / velocity state for dumping.
// The C functions below take plain pointers and sizes only.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <exception>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include <spdlog/spdlog.h>

#include "advection.h"
#include "global_benchmark.h"
#include "geometry/rigid_body_collider3.h"
#include "geometry/sphere3.h"
#include "octree_omnigrid_solver3.h"
#include "octree_util.h"
#include "pcgsolver/pcg_solver.h"
#include "viewer/grid_dump.h"

using namespace Omni;

namespace {

std::string g_err;

class HarnessSolver : public OctreeOmniGridSolver3 {
public:
    HarnessSolver(int R, double h, int levels, const int* split, int closedFlags)
        : OctreeOmniGridSolver3(Size3(R, R, R), h, "astg_ref_harness"),
          split_(split, split + (size_t)R * R * R), R_(R), closed_(closedFlags) {
        numOctreeLayer = levels;
        isRestart = true;         // skips the layout dump in initData
        useFixedTimeStep = true;
        enableTiltAdaptivity = false;
        splitMapSetup = [this](OctreeOmniGridSolver3* s) {
            auto& sm = s->getSplitMap();
            for (int z = 0; z < R_; ++z)
                for (int y = 0; y < R_; ++y)
                    for (int x = 0; x < R_; ++x)
                        sm->set(Size3(x, y, z), split_[((size_t)z * R_ + y) * R_ + x]);
        };
    }

    void initObstacle() override {
        boundarySolver->setClosedDomainBoundaryFlag(closed_);
        if (collider_) {
            // the reference's own initObstacle collider (octree_omnigrid_solver3.cpp:152-157)
            // with the sphere's centre and radius as parameters
            Geometry::Sphere3Ptr sphere = std::make_shared<Geometry::Sphere3>(
                colR_, Geometry::Transform3(Vector3f(colC_[0], colC_[1], colC_[2]), Geometry::Quaternion()));
            boundarySolver->updateCollider(std::make_shared<Geometry::RigidBodyCollider3>(sphere));
        } else {
            auto sdf = std::make_shared<CellCenteredScalarGrid3>(resolution, delta_h, 10000);
            boundarySolver->updateSdf(sdf);
        }
        boundarySdf = boundarySolver->colliderSdf();
    }
    void setCollider(double cx, double cy, double cz, double r) {
        collider_ = true;
        colC_[0] = cx;
        colC_[1] = cy;
        colC_[2] = cz;
        colR_ = r;
    }
    CellCenteredScalarGrid3Ptr& sdf() { return boundarySdf; }

    // constrainVelocity(MAC) (omni_boundary_solver3.h:72-161) restated with the
    // reference's own grid calls, the faces visited one after another in the
    // iteration order (levels ascending, leaf order, axis x, y, z). The
    // reference's loop runs on its thread pool and updates faces in place
    // while sampling neighbours, so its result depends on where the pool cuts
    // the iteration into chunks; this serial form is the order-defined version
    // of the same loop (the slip part), followed by the closed-wall part.
    void constrainMacSerial() {
        auto& vel = velocity->getMACGrid();
        auto& sdfG = boundarySolver->colliderSdf();
        const Vector3f sCenter(0.5, 0.5, 0.45);  // omni_boundary_solver3.h:77
        for (auto it : vel->getParallelIteratorVec())
            for (; it.isValid(); it.next())
                for (int axis = 0; axis < 3; ++axis) {
                    if (!vel->isValid(axis, it.level, it.xIdx(), it.yIdx(), it.zIdx())) continue;
                    auto pos = vel->position(axis, it.level, it.xIdx(), it.yIdx(), it.zIdx());
                    if (!isInsideSdf(sdfG->sample(pos))) continue;
                    Vector3f n = pos - sCenter;
                    if (n.norm() < EPSILON) {
                        vel->set(axis, it.level, it.xIdx(), it.yIdx(), it.zIdx(), 0);
                    } else {
                        auto v = vel->sample(pos);
                        n.normalize();
                        vel->set(axis, it.levelCoor(), (v - v.dot(n) * n)[axis]);
                    }
                }
        auto& layout = vel->getLayout();
        const int f = boundarySolver->closedDomainBoundaryFlag();
        for (int i = 0; i < layout->level(); ++i) {
            const auto& res = layout->levelResolution(i);
            for (int a = 0; a < res.x(); ++a)
                for (int b = 0; b < res.x(); ++b) {
                    if (f & kDirectionXPositive) vel->trySetX(i, res.x() - 1, b, a, 0);
                    if (f & kDirectionXNegative) vel->trySetX(i, 0, b, a, 0);
                    if (f & kDirectionYPositive) vel->trySetY(i, b, res.y() - 1, a, 0);
                    if (f & kDirectionYNegative) vel->trySetY(i, b, 0, a, 0);
                    if (f & kDirectionZPositive) vel->trySetZ(i, b, a, res.z() - 1, 0);
                    if (f & kDirectionZNegative) vel->trySetZ(i, b, a, 0, 0);
                }
        }
    }
    // onUpdate's constraint and ghost update before project (222-232), with
    // the serial MAC constraint
    void constrainAndGhostSerial() {
        constrainMacSerial();
        boundarySolver->constrainVelocity(velocity->getTiltGrid(), velocity->getMACGrid());
        velocity->updateGhost();
    }
    // the chain between project and advect (234-241), serial MAC constraint
    void postProjectSerial() {
        constrainAndGhostSerial();
        velocity->updateDualGrid();
        boundarySolver->constrainVelocity(velocity->getDualGrid());
        velocity->getDualGrid()->updateGhost();
    }
    void doWipeBoundary() { wipeBoundary(); }

    // --- accessors -------------------------------------------------------
    OctreeGridLayout3Ptr& lay() { return layout; }
    OctreeTiltENodeGrid3Ptr& tilt() { return tiltE; }
    OctreeTiltENodeGrid3Ptr& ttilt() { return tTiltE; }
    OctreeOmniGridLayoutPtr<3>& tiltLayout() { return tilt_layout; }
    OctreeOmniFlagGrid3Ptr& pidx() { return pIndex; }
    OctreeOmniFaceCenteredGrid3Ptr& weights() { return weightsCC; }
    PCG::SparseMatrix<REAL>& mat() { return A; }
    VectorXd& rhs() { return b; }
    VectorXd& sol() { return vRes; }
    PCGSolver<REAL>& pcg() { return pcgSolver; }
    double h() const { return delta_h; }

    // onUpdate's chain between project and advect (octree_omnigrid_solver3.cpp:234-241)
    void postProject() {
        boundarySolver->constrainVelocity(velocity->getMACGrid());
        boundarySolver->constrainVelocity(velocity->getTiltGrid(), velocity->getMACGrid());
        velocity->updateGhost();
        velocity->updateDualGrid();
        boundarySolver->constrainVelocity(velocity->getDualGrid());
        velocity->getDualGrid()->updateGhost();
    }
    Vector3f backTrace(const Vector3f& p, double dt) { return BackTrace(boundarySdf, velocity, p, dt); }
    OctreeOmniScalarGrid3Ptr& scalar(int f) { return f == 0 ? density : heat; }

    void doProject(double dt) { project(dt); }
    void doUpdateVelocity(const VectorXd& p, double dt) { updateVelocityByPressure(p, dt); }
    int doNumber() { return updateIndexParallel(); }
    void doAdvect(double dt) { advect(dt); }
    void doAdjustE(double dThr, double dt, double vThr, double maxThr) { adjustE(dThr, dt, vThr, maxThr); }
    void setTiltAdaptivity(bool on) { enableTiltAdaptivity = on; }
    void setBuoyancy(double fd, double ft) {
        _buoyancySmokeDensityFactor = fd;
        _buoyancyTemperatureFactor = ft;
    }

    // sceneSetup of the reference's octree smoke example (sphereSource,
    // example/octree_smoke_3d/main.cpp:23-81) with its centre, radius, values
    // and inflow speed as parameters; onUpdate calls it first in every step.
    void setSource(double cx, double cy, double cz, double r, double dval, double hval, double speed) {
        srcC_ = Vector3f(cx, cy, cz);
        srcR_ = r;
        srcD_ = dval;
        srcH_ = hval;
        srcV_ = speed;
        srcFirst_ = true;
        sceneSetup = [this](OctreeOmniGridSolver3*) { applySource(); };
    }
    void applySource() {
        auto in = [&](const Vector3f& p) { return (p - srcC_).norm() < srcR_; };
        auto& dOct = density->getOctagonGrid();
        auto& dTilt = density->getTiltGrid();
        auto& hOct = heat->getOctagonGrid();
        auto& hTilt = heat->getTiltGrid();
        ThreadPool::parallelIterateGrid(dOct->getParallelIteratorVec(), [&](const OctreeGridIterator3& it) {
            if (in(dOct->position(it.levelCoor()))) {
                dOct->set(it.levelCoor(), srcD_);
                hOct->set(it.levelCoor(), srcH_);
            }
        });
        ThreadPool::parallelIterateGrid(dTilt->getParallelIteratorVec(), [&](const OctreeGridIterator3& it) {
            if (in(dTilt->position(it.levelCoor()))) {
                dTilt->set(it.levelCoor(), srcD_);
                hTilt->set(it.levelCoor(), srcH_);
            }
        });
        parallel_iterate_half_tilt(layout, [&](LCOOR_T c) {
            if (in(dTilt->position(c))) {
                dTilt->set(c, srcD_);
                hTilt->set(c, srcH_);
            }
        });
        auto& mac = velocity->getMACGrid();
        auto& tv = velocity->getTiltGrid();
        ThreadPool::parallelIterateGrid(mac->getParallelIteratorVec(), [&](const OctreeGridIterator3& it) {
            if (in(mac->positionZ(it.levelCoor()))) mac->setZ(it.levelCoor(), srcV_);
            if (in(tv->positionCenter(it.levelCoor()))) tv->set(it.levelCoor(), Vector8f::Ones() * srcV_ * INV_SQRT3);
        });
        parallel_iterate_half_tilt(layout, [&](LCOOR_T c) {
            if (!in(tv->positionCenter(c))) return;
            Vector8f v = Vector8f::Zero();
            const int dir = tiltE->get(c).half_direction;
            for (int i = 0; i < 4; ++i) v(HALF_TILT_SLOT_INDEX[dir][i]) = srcV_ * INV_SQRT3;
            if (dir % 3 == 2) v(HALF_TILT_SLOT_INDEX[dir][4]) = srcV_;
            tv->set(c, v);
        });
        if (srcFirst_) {
            density->updateGhost();
            heat->updateGhost();
            srcFirst_ = false;
        }
    }
    void doStep(double dt) { onUpdate(dt); }

private:
    std::vector<int> split_;
    int R_;
    int closed_;
    bool collider_ = false;
    double colC_[3] = {0, 0, 0}, colR_ = 0;
    Vector3f srcC_ = Vector3f::Zero();
    double srcR_ = 0, srcD_ = 0, srcH_ = 0, srcV_ = 0;
    bool srcFirst_ = true;
};

HarnessSolver* H(void* p) { return static_cast<HarnessSolver*>(p); }

LCOOR_T nodeCoor(OctreeGridLayout3Ptr& layout, int level, int i) {
    auto layer = layout->getLayerPtr(level);
    int nleaf = layer->nodeSize();
    return i < nleaf ? layer->getNodeIdx()[i] : layer->getGhostNodeIdx()[i - nleaf];
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    } catch (...) {
        g_err = "unknown exception";
        return -1;
    }
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void ref_quiet() { spdlog::set_level(spdlog::level::off); }

int ref_num_workers() { return TaskManager::tf.num_workers(); }

// split: R^3 ints, x fastest, value = level of the leaf covering that finest cell.
void* ref_create(int R, double h, int levels, const int* split, int closedFlags) {
    HarnessSolver* s = nullptr;
    int rc = guarded([&] {
        s = new HarnessSolver(R, h, levels, split, closedFlags);
        s->initialize();
    });
    if (rc != 0) {
        delete s;
        return nullptr;
    }
    return s;
}

// The same with the reference's sphere collider (centre, radius) in place of
// the constant boundary SDF.
void* ref_create_collider(int R, double h, int levels, const int* split, int closedFlags, double cx, double cy,
                          double cz, double r) {
    HarnessSolver* s = nullptr;
    int rc = guarded([&] {
        s = new HarnessSolver(R, h, levels, split, closedFlags);
        s->setCollider(cx, cy, cz, r);
        s->initialize();
    });
    if (rc != 0) {
        delete s;
        return nullptr;
    }
    return s;
}

// The boundary SDF grid (R^3 cell-centred values, x fastest).
void ref_get_sdf(void* p, double* out) {
    auto& g = H(p)->sdf();
    const auto res = g->resolution();
    for (int z = 0; z < res.z(); ++z)
        for (int y = 0; y < res.y(); ++y)
            for (int x = 0; x < res.x(); ++x)
                out[((size_t)z * res.y() + y) * res.x() + x] = g->get(Size3(x, y, z));
}

// Boundary SDF sampled at points (CellCenteredGrid::sample, cell_centered_grid.h:40-54).
int ref_sample_sdf(void* p, int n, const double* xyz, double* out) {
    return guarded([&] {
        for (int i = 0; i < n; ++i) out[i] = H(p)->sdf()->sample(Vector3f(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
    });
}

int ref_constrain_and_ghost_serial(void* p) {
    return guarded([&] { H(p)->constrainAndGhostSerial(); });
}
int ref_post_project_serial(void* p) {
    return guarded([&] { H(p)->postProjectSerial(); });
}
int ref_wipe_boundary(void* p) {
    return guarded([&] { H(p)->doWipeBoundary(); });
}

void ref_destroy(void* p) { delete H(p); }

int ref_level_size(void* p, int level, int includeGhost) {
    return H(p)->lay()->levelSize(level, includeGhost != 0);
}

// xyz: 3*n ints, flags: n ints (LayerNodeFlagMask bits), in layer-node index order.
void ref_nodes(void* p, int level, int* xyz, int* flags) {
    auto& layout = H(p)->lay();
    auto layer = layout->getLayerPtr(level);
    int n = layer->size();
    for (int i = 0; i < n; ++i) {
        LCOOR_T c = nodeCoor(layout, level, i);
        const auto& nd = layer->getMapTable()[c];
        xyz[3 * i + 0] = nd.xIdx;
        xyz[3 * i + 1] = nd.yIdx;
        xyz[3 * i + 2] = nd.zIdx;
        flags[i] = (int)nd.flag;
        if ((int)nd.index != i) flags[i] |= 1 << 30;  // index mismatch marker
    }
}

// Tilt state per layer node. bits: 0 is_largest, 1 is_closed, 2 is_fixed,
// 3 is_T_joint, 4 is_closed_by_T_joint, 5 is_half_tilt, 6 layout.is_open,
// 7 layout.is_fixed.
// which = 0: tiltE (the state the solver's fields use), 1: tTiltE (the state
// adjustE writes, swapped in at the end of advect).
void ref_tilt_nodes_of(void* p, int which, int level, double* e, int* bits, int* halfDir) {
    auto* s = H(p);
    auto& layout = s->lay();
    auto& data = (which ? s->ttilt() : s->tilt())->getData()[level];
    int n = (int)data.size();
    for (int i = 0; i < n; ++i) {
        const TiltENode& t = data[i];
        LCOOR_T c = nodeCoor(layout, level, i) | ((LCOOR_T)level << 60);
        const TiltENodeS& ts = s->tiltLayout()->getTiltNode(c);
        e[i] = t.e;
        bits[i] = (t.is_largest ? 1 : 0) | (t.is_closed ? 2 : 0) | (t.is_fixed ? 4 : 0) |
                  (t.is_T_joint ? 8 : 0) | (t.is_closed_by_T_joint ? 16 : 0) |
                  (t.is_half_tilt ? 32 : 0) | (ts.is_open ? 64 : 0) | (ts.is_fixed ? 128 : 0);
        halfDir[i] = t.half_direction;
    }
}

void ref_tilt_nodes(void* p, int level, double* e, int* bits, int* halfDir) {
    ref_tilt_nodes_of(p, 0, level, e, bits, halfDir);
}

// Returns the number of half-tilts; out (may be null) receives (level,x,y,z) quads.
int ref_half_tilts(void* p, int* out) {
    auto& v = H(p)->lay()->getHalfGhostIdx();
    if (out) {
        for (size_t i = 0; i < v.size(); ++i) {
            out[4 * i + 0] = UNPACK_LEVEL3(v[i]);
            out[4 * i + 1] = UNPACK_X3(v[i]);
            out[4 * i + 2] = UNPACK_Y3(v[i]);
            out[4 * i + 3] = UNPACK_Z3(v[i]);
        }
    }
    return (int)v.size();
}

// Opens/closes non-fixed leaf tilts of one level from a per-leaf mask, the way
// adjustE (octree_omnigrid_solver3.cpp:1125-1166) commits its decision:
// layout flag, then setE(h/2,h/2) or setE(0), then ghost propagation.
int ref_set_tilt_open(void* p, int level, const unsigned char* openMask) {
    auto* s = H(p);
    return guarded([&] {
        auto& layout = s->lay();
        auto layer = layout->getLayerPtr(level);
        int nleaf = layer->nodeSize();
        double maxE = layout->levelGridSpacing(level) * 0.5;
        for (int i = 0; i < nleaf; ++i) {
            LCOOR_T c = layer->getNodeIdx()[i] | ((LCOOR_T)level << 60);
            auto& ts = s->tiltLayout()->getTiltNode(c);
            if (ts.is_fixed) continue;
            ts.is_open = openMask[i] != 0;
            for (auto* g : {&s->tilt(), &s->ttilt()}) {
                auto& node = (*g)->get(c);
                if (ts.is_open) node.setE(maxE, maxE);
                else node.setE(0);
            }
        }
        s->tiltLayout()->updateGhost();
        s->tilt()->updateGhost();
        s->ttilt()->updateGhost();
    });
}

// mac: n*3 (x,y,z per node), tilt: n*8, in layer-node index order.
void ref_set_velocity(void* p, int level, const double* mac, const double* tilt8) {
    auto* s = H(p);
    auto& v = s->getVelocity();
    auto& mg = v->getMACGrid();
    auto& td = v->getTiltGrid()->getGridPtr()->getData()[level];
    int n = std::min((int)td.size(), s->lay()->levelSize(level));
    for (int i = 0; i < n; ++i) {
        mg->getUx()[level][i] = mac[3 * i + 0];
        mg->getUy()[level][i] = mac[3 * i + 1];
        mg->getUz()[level][i] = mac[3 * i + 2];
        for (int k = 0; k < 8; ++k) td[i](k) = tilt8[8 * i + k];
    }
}

void ref_get_velocity(void* p, int level, double* mac, double* tilt8) {
    auto* s = H(p);
    auto& v = s->getVelocity();
    auto& mg = v->getMACGrid();
    auto& td = v->getTiltGrid()->getGridPtr()->getData()[level];
    int n = std::min((int)td.size(), s->lay()->levelSize(level));
    for (int i = 0; i < n; ++i) {
        mac[3 * i + 0] = mg->getUx()[level][i];
        mac[3 * i + 1] = mg->getUy()[level][i];
        mac[3 * i + 2] = mg->getUz()[level][i];
        for (int k = 0; k < 8; ++k) tilt8[8 * i + k] = td[i](k);
    }
}

// ---- post-projection chain, sampling and advection (a19-a26) -------------
// Scalar fields: 0 density, 1 heat; per level, per layer node: octagon and tilt.
// The advection temporaries can hold more entries than the layer has nodes;
// only the first levelSize entries are node values.
void ref_set_scalar(void* p, int field, int level, const double* oct, const double* tilt) {
    auto& f = H(p)->scalar(field);
    auto& od = f->getOctagonGrid()->getData()[level];
    auto& td = f->getTiltGrid()->getData()[level];
    const size_t n = (size_t)H(p)->lay()->levelSize(level);
    for (size_t i = 0; i < n && i < od.size(); ++i) od[i] = oct[i];
    for (size_t i = 0; i < n && i < td.size(); ++i) td[i] = tilt[i];
}

void ref_get_scalar(void* p, int field, int level, double* oct, double* tilt) {
    auto& f = H(p)->scalar(field);
    auto& od = f->getOctagonGrid()->getData()[level];
    auto& td = f->getTiltGrid()->getData()[level];
    const size_t n = (size_t)H(p)->lay()->levelSize(level);
    for (size_t i = 0; i < n; ++i) oct[i] = i < od.size() ? od[i] : 0.0;
    for (size_t i = 0; i < n; ++i) tilt[i] = i < td.size() ? td[i] : 0.0;
}

int ref_post_project(void* p) {
    return guarded([&] { H(p)->postProject(); });
}

int ref_scalar_update_ghost(void* p, int field) {
    return guarded([&] { H(p)->scalar(field)->updateGhost(); });
}

// Dual grid (cell-centre and vertex velocity vectors): 3 per node.
void ref_get_dual(void* p, int level, double* oct3, double* tilt3) {
    auto& d = H(p)->getVelocity()->getDualGrid();
    auto& od = d->getOctagonGrid()->getData()[level];
    auto& td = d->getTiltGrid()->getData()[level];
    const size_t n = (size_t)H(p)->lay()->levelSize(level);
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) oct3[3 * i + k] = i < od.size() ? od[i](k) : 0.0;
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) tilt3[3 * i + k] = i < td.size() ? td[i](k) : 0.0;
}

// what: 0 omni velocity sample, 1 MAC-only sample, 2 back-trace (dt)
int ref_sample_points(void* p, int what, int n, const double* xyz, double dt, double* out3) {
    auto* s = H(p);
    return guarded([&] {
        for (int i = 0; i < n; ++i) {
            Vector3f q(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]), r;
            if (what == 0) r = s->getVelocity()->sample(q);
            else if (what == 1) r = s->getVelocity()->getMACGrid()->sample(q);
            else r = s->backTrace(q, dt);
            for (int k = 0; k < 3; ++k) out3[3 * i + k] = r(k);
        }
    });
}

int ref_sample_scalar(void* p, int field, int n, const double* xyz, double* out) {
    auto* s = H(p);
    return guarded([&] {
        for (int i = 0; i < n; ++i) out[i] = s->scalar(field)->sample(Vector3f(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
    });
}

// The reference's frame dumps (src/viewer/grid_dump.cpp) of the current
// state; a null path skips that file.
int ref_write_frame(void* p, const char* layout, const char* tilt, const char* density, const char* heat,
                    const char* velocity) {
    return guarded([&] {
        auto* s = H(p);
        if (layout) dumpOctreeLayout3dBinary(layout, *s->lay());
        if (tilt) dumpOctreeOmniTiltE3dBinary(tilt, *s->tilt());
        if (density) dumpOctreeOmniScalar3dBinary(density, *s->scalar(0));
        if (heat) dumpOctreeOmniScalar3dBinary(heat, *s->scalar(1));
        if (velocity) dumpOctreeOmniFaceCentered3dBinary(velocity, *s->getVelocity());
    });
}

// The reference's restart (octree_omnigrid_solver3.cpp:1015-1057) from frame
// files; it also writes a preview image under ./astg_ref_harness/.
int ref_restart(void* p, const char* layout, const char* tilt, const char* density, const char* heat,
                const char* velocity) {
    int ok = 0;
    int rc = guarded([&] { ok = H(p)->restart(0, layout, "", tilt, density, heat, velocity) ? 1 : 0; });
    if (rc) return rc;
    if (!ok) {
        g_err = "restart failed to read a file";
        return -1;
    }
    return 0;
}

// adjustE (octree_omnigrid_solver3.cpp:1125-1166) on the current fields.
int ref_adjust_e(void* p, double dThr, double dt, double vThr, double maxThr) {
    return guarded([&] { H(p)->doAdjustE(dThr, dt, vThr, maxThr); });
}

// enableTiltAdaptivity (octree_omnigrid_solver3.h:56): advect(dt) then starts
// with adjustE(0.02, dt, EPSILON, 1000) (octree_omnigrid_solver3.cpp:852-853).
void ref_set_tilt_adaptivity(void* p, int on) { H(p)->setTiltAdaptivity(on != 0); }

// Scene source for onUpdate's sceneSetup (see HarnessSolver::setSource) and
// addForce's buoyancy factors.
void ref_set_source(void* p, double cx, double cy, double cz, double r, double dval, double hval, double speed) {
    H(p)->setSource(cx, cy, cz, r, dval, hval, speed);
}
void ref_set_buoyancy(void* p, double fd, double ft) { H(p)->setBuoyancy(fd, ft); }

// One onUpdate(dt) (octree_omnigrid_solver3.cpp:222-249) with the PCG
// tolerance factor and iteration cap given here.
int ref_step(void* p, double dt, double tol, int maxit) {
    auto* s = H(p);
    return guarded([&] {
        s->pcg().set_solver_parameters(tol, maxit);
        s->doStep(dt);
    });
}

int ref_advect(void* p, double dt) {
    return guarded([&] { H(p)->doAdvect(dt); });
}

void ref_get_weights(void* p, int level, double* mac, double* tilt8) {
    auto* s = H(p);
    auto& w = s->weights();
    auto& mg = w->getMACGrid();
    auto& td = w->getTiltGrid()->getGridPtr()->getData()[level];
    int n = (int)td.size();
    for (int i = 0; i < n; ++i) {
        mac[3 * i + 0] = mg->getUx()[level][i];
        mac[3 * i + 1] = mg->getUy()[level][i];
        mac[3 * i + 2] = mg->getUz()[level][i];
        for (int k = 0; k < 8; ++k) tilt8[8 * i + k] = td[i](k);
    }
}

// Runs updateIndexParallel (octree_omnigrid_solver3.cpp:292-347); returns n.
int ref_number(void* p) {
    int n = -1;
    int rc = guarded([&] { n = H(p)->doNumber(); });
    return rc == 0 ? n : -1;
}

void ref_get_pindex(void* p, int level, int* oct, int* tiltIdx) {
    auto* s = H(p);
    auto& od = s->pidx()->getOctagonGrid()->getData()[level];
    auto& td = s->pidx()->getTiltGrid()->getData()[level];
    for (size_t i = 0; i < od.size(); ++i) {
        oct[i] = od[i];
        tiltIdx[i] = td[i];
    }
}

// Full reference project(dt) (octree_omnigrid_solver3.cpp:350-679) with the
// PCG tolerance factor and iteration cap given here.
int ref_project(void* p, double dt, double tol, int maxit) {
    auto* s = H(p);
    return guarded([&] {
        s->pcg().set_solver_parameters(tol, maxit);
        s->doProject(dt);
    });
}

// Re-solves the last assembled system with the same PCGSolver settings to
// recover the residual and iteration count that project() only logs.
int ref_resolve(void* p, double* x, double* residual, int* iters) {
    auto* s = H(p);
    return guarded([&] {
        VectorXd xs(s->rhs().rows());
        double r = 0;
        int it = 0;
        s->pcg().solve(s->mat(), s->rhs(), xs, r, it);
        for (int i = 0; i < xs.rows(); ++i) x[i] = xs[i];
        *residual = r;
        *iters = it;
    });
}

int ref_matrix_n(void* p) { return (int)H(p)->mat().n; }

int ref_matrix_nnz(void* p) {
    auto& A = H(p)->mat();
    long nnz = 0;
    for (unsigned i = 0; i < A.n; ++i) nnz += (long)A.index[i].size();
    return (int)nnz;
}

// CSR of the last assembled matrix, the scaled right-hand side b and solution.
void ref_get_system(void* p, int* rowstart, int* col, double* val, double* b, double* x) {
    auto* s = H(p);
    auto& A = s->mat();
    int k = 0;
    rowstart[0] = 0;
    for (unsigned i = 0; i < A.n; ++i) {
        for (size_t j = 0; j < A.index[i].size(); ++j) {
            col[k] = (int)A.index[i][j];
            val[k] = A.value[i][j];
            ++k;
        }
        rowstart[i + 1] = k;
    }
    for (unsigned i = 0; i < A.n; ++i) {
        b[i] = s->rhs()[i];
        x[i] = s->sol()[i];
    }
}

// Velocity correction alone (octree_omnigrid_solver3.cpp:681-847) with a given p.
int ref_update_velocity(void* p, const double* pres, int n, double dt) {
    auto* s = H(p);
    return guarded([&] {
        VectorXd pv(n);
        for (int i = 0; i < n; ++i) pv[i] = pres[i];
        s->doUpdateVelocity(pv, dt);
    });
}

// ---- solver-level entry points on a caller CSR (PCGSolver, pcg_solver.h) ----

static void csrToRows(int n, const int* rs, const int* ci, const double* v, PCG::SparseMatrix<double>& A) {
    A.resize(n);
    A.zero();
    for (int i = 0; i < n; ++i)
        for (int k = rs[i]; k < rs[i + 1]; ++k) {
            A.index[i].push_back((unsigned)ci[k]);
            A.value[i].push_back(v[k]);
        }
}

// MIC(0) factor (pcg_solver.h:89-190). Returns nnz of the strict lower part;
// arrays may be null on a size query.
int ref_mic_factor(int n, const int* rs, const int* ci, const double* v, double tau, double ratio,
                   double* invdiag, unsigned* colstart, unsigned* rowindex, double* value) {
    PCG::SparseMatrix<double> A;
    csrToRows(n, rs, ci, v, A);
    SparseColumnLowerFactor<double> f;
    factor_modified_incomplete_cholesky0(A, f, tau, ratio);
    if (invdiag) {
        for (int i = 0; i < n; ++i) invdiag[i] = f.invdiag[i];
        for (int i = 0; i <= n; ++i) colstart[i] = f.colstart[i];
        for (size_t k = 0; k < f.value.size(); ++k) {
            rowindex[k] = f.rowindex[k];
            value[k] = f.value[k];
        }
    }
    return (int)f.value.size();
}

// L L^T z = r with the factor above (pcg_solver.h:211-253).
void ref_mic_apply(int n, const int* rs, const int* ci, const double* v, double tau, double ratio,
                   const double* r, double* z) {
    PCG::SparseMatrix<double> A;
    csrToRows(n, rs, ci, v, A);
    SparseColumnLowerFactor<double> f;
    factor_modified_incomplete_cholesky0(A, f, tau, ratio);
    Eigen::VectorXd rr(n), zz(n);
    for (int i = 0; i < n; ++i) rr[i] = r[i];
    solve_lower(f, rr, zz);
    solve_lower_transpose_in_place(f, zz);
    for (int i = 0; i < n; ++i) z[i] = zz[i];
}

// MIC(0)-PCG (Eigen overload, pcg_solver.h:366-426) on a caller CSR.
int ref_pcg_csr(int n, const int* rs, const int* ci, const double* v, const double* b, double* x,
                double tol, int maxit, double tau, double ratio, double* residual, int* iters) {
    int ok = 0;
    int rc = guarded([&] {
        PCG::SparseMatrix<double> A;
        csrToRows(n, rs, ci, v, A);
        PCGSolver<double> solver;
        solver.set_solver_parameters(tol, maxit, tau, ratio);
        VectorXd bb(n), xx(n);
        for (int i = 0; i < n; ++i) bb[i] = b[i];
        double r = 0;
        int it = 0;
        ok = solver.solve(A, bb, xx, r, it) ? 1 : 0;
        for (int i = 0; i < n; ++i) x[i] = xx[i];
        *residual = r;
        *iters = it;
    });
    return rc == 0 ? ok : -1;
}

// The reference's own section timers (global_benchmark.cpp, built with
// GLOBALBENCHMARK): cleared here, read back as the text BenchmarkTimer::report
// writes (timer.h:214-225: "name<TAB>seconds<TAB>invocations" per line, a
// nested section indented 4 spaces per level, then "Full time").
void ref_benchmark_reset() { BENCHMARK_RESET(); }

// Copies the report into buf (at most cap-1 bytes and a NUL); returns its length.
int ref_benchmark_report(char* buf, int cap) {
    std::ostringstream os;
    std::streambuf* old = std::cout.rdbuf(os.rdbuf());
    BENCHMARK_REPORT_NO_MESSAGES();
    std::cout.rdbuf(old);
    const std::string s = os.str();
    if (buf && cap > 0) {
        const size_t n = std::min(s.size(), (size_t)cap - 1);
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return (int)s.size();
}

}  // extern "C"

