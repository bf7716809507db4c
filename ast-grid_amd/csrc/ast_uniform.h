// Uniform-layout AST projection on the device (the reference's
// ast::ASTSolver<d>::project, include/ast_solver.h:322-498, with its build
// configuration: CORRECTION_MAC, no edge-node unknowns, ast_general.h:11-15).
//
// Layout in HBM, every grid x fastest as the reference's own:
//   open    uint8 per tilt node, prod(res + 1) (boundary nodes held closed)
//   tid     int per tilt node: its unknown, or -1
//   tnode   int per open tilt: its tilt node (tilt rows nc .. n-1)
//   mac[a]  double per MAC face of axis a, prod(res + e_a)
//   tilt    double[kT] per tilt node (kT = 4 in 2D, 8 in 3D), node-major
// Unknowns: every cell in box order, then the open tilt nodes in box order
// (updateIndex 599-621). The system is built on the device (row lengths, a
// scan, sorted columns, then values and right-hand side, one lane per row) and
// solved by the same MIC(0)-PCG as the octree projection; the velocity
// correction and the wall constraint run one lane per face or tilt.
//
// Solver orderings: the reference numbering (the reference's MIC(0) and
// iterates), or block colours (an added form, as for the octree: squares or
// cubes of B = 4 cells coloured by the parity of their coordinates, 4 colours in
// 2D and 8 in 3D, rows ordered by colour, block and reference index; a tilt
// row belongs to the block of its lower corner cell). Every coupling joins
// unknowns at most one cell apart, so same-coloured blocks are uncoupled and
// the sweeps' dependency chains are one block deep per colour instead of
// nx + ny (+ nz) deep. ASTG_ORDER_AUTO picks block colours for the mixed form.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>

#include "pcg.h"
#include "reorder.h"

namespace astg {

struct AstGeom {
    int d = 0;
    int res[3] = {1, 1, 1};  // cells per axis (res[2] = 1 in 2D)
    int nc = 0;              // cells
    long long nt = 0;        // tilt nodes, prod(res + 1)
    long long nm[3] = {0, 0, 0};  // MAC faces per axis
    int kT = 0;              // tilt slots per node
    double dx = 0, half = 0;
    double tod = 0;    // distance between a tilt and a cell centre: half dx sqrt(d)
    double tarea = 0;  // one tilt face: half dx sqrt2 (2D), sqrt3 / 2 half dx^2 (3D)
    double cut = 0;    // MAC face area a tilt face takes away: half dx (2D), half dx^2 / 2 (3D)
    double full = 0;   // MAC face area: dx (2D), dx^2 (3D)
};

class AstUniform {
public:
    explicit AstUniform(hipStream_t s) : stream_(s), pcg_(s) {}
    ~AstUniform();
    AstUniform(const AstUniform&) = delete;
    AstUniform& operator=(const AstUniform&) = delete;
    // Sizes and spacing; fields zero, every tilt closed.
    int init(int dim, const int* res, double dx, std::string& err);
    bool ready() const { return ready_; }  // init succeeded
    const AstGeom& geom() const { return g_; }
    // Tilt-node flags (host, prod(res + 1)); boundary nodes are held closed
    // (initData 79-87, constrainFixedTiltNode). Renumbers and rebuilds the
    // pattern on the next projection.
    int setLayout(const uint8_t* open, std::string& err);
    // Host arrays in the reference's grid layouts; mac[2] unused in 2D.
    int setVelocity(const double* const* mac, const double* tilt, std::string& err);
    int getVelocity(double* const* mac, double* tilt, std::string& err);
    // Assembly, MIC(0)-PCG with tol an absolute infinity-norm bound (the
    // reference's std::vector PCGSolver::solve), correction, wall constraint.
    int project(double dt, double rho, const astg_pcg_params& prm, astg_pcg_result& res, std::string& err);
    int n() const { return n_; }
    long long nnz() const { return nnz_; }
    // Downloads of the last projection's system and pressure, and the numbering.
    int system(int* rowstart, int* col, double* val, double* b, double* p, std::string& err);
    int tiltIndex(int* tid, std::string& err);

private:
    int ensurePattern(std::string& err);
    int ensureBlockOrder(std::string& err);

    hipStream_t stream_;
    AstGeom g_;
    PcgDevice pcg_;
    DevBuf open_, tid_, tnode_, flag_, scan_, mac_[3], tilt_, len_, rs_, col_, tmp_, bmax_, keys_;
    BlockOrderDevice blk_;
    std::unique_ptr<PcgDevice> pcgBlk_;  // the block-colour copy of the system
    int n_ = 0;
    long long nnz_ = 0;
    bool patternReady_ = false;
    bool blkReady_ = false;
    bool solved_ = false;
    bool ready_ = false;
    hipEvent_t ev_[5] = {};
};

}  // namespace astg
