// Host-side preparation of the projection: cutting-cell weights, flattened
// node arrays, per-row stencil tables and the CSR pattern of A.
//
// Everything here depends on topology only (node sets, tilt state, numbering)
// and is rebuilt when the topology changes; the per-projection numeric work
// (assembly, MIC(0), PCG, velocity correction) runs on the device from these
// tables.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "octree.h"

namespace astg {

extern const int kHalfSlot[6][5];
extern const int kTiltNbOct[8][3];
extern const int kOctNbTilt[8][3];
extern const int kOctNb[6][3];
extern const int kOctNbFace[6][3];

struct Weights {
    std::vector<std::vector<double>> mac;   // per level: 3*nnode
    std::vector<std::vector<double>> tilt;  // per level: 8*nnode
};
void computeWeights(const Octree& o, Weights& w);

// Face-neighbour codes of an octagon row (values >= 0 are unknown indices).
enum : int32_t { FACE_EXCLUDED = -1, FACE_SKIP = -10, FACE_CLOSED = -11, FACE_OPEN = -12 };
// Tilt-row neighbour codes.
enum : int32_t { TNB_ZERO = -1, TNB_DIRICHLET = -13, TNB_NONE = -14 };
// Velocity-update neighbour codes.
enum : int32_t { VNB_P0 = -1, VNB_ZERO = -3, VNB_SKIP = -2 };

struct ProjectTables {
    // flattened nodes
    int G = 0;                      // total nodes (all levels, leaves + ghosts)
    std::vector<int> off;           // level offsets, size L+1
    std::vector<uint8_t> level;     // G
    std::vector<double> tilt_e;     // G
    std::vector<uint8_t> tilt_bits; // G
    std::vector<int8_t> tilt_hd;    // G
    std::vector<int32_t> pid_oct, pid_tilt;  // G
    std::vector<double> w_mac;      // 3*G, axis-major
    std::vector<double> w_tilt;     // 8*G, slot-major

    int n = 0;
    int n_oct = 0, n_tilt = 0, n_half_rows = 0;

    // octagon rows (j = 0..n_oct-1), SoA with stride n_oct
    std::vector<int32_t> oct_row;     // unknown index
    std::vector<int32_t> oct_corner;  // 8 * n_oct, global node of tilt corner k
    std::vector<int32_t> oct_face;    // 6 * n_oct, face code
    std::vector<uint8_t> oct_pos;     // 15 * n_oct, CSR offset in row (0xff = none)

    // tilt rows
    std::vector<int32_t> tilt_row, tilt_node;
    std::vector<int32_t> tilt_nb;     // 8 * n_tilt
    std::vector<uint8_t> tilt_nblev;  // 8 * n_tilt
    std::vector<uint8_t> tilt_pos;    // 9 * n_tilt

    // half-tilt entries q (halfGhostIdx order)
    int nh = 0;
    std::vector<int32_t> half_node, half_c, half_t;   // half_t: coarse octagon unknown, -1, or -20 (coarse invalid)
    std::vector<uint8_t> half_closed;                 // outer wall closed (coarse-invalid case)
    std::vector<int32_t> half_nb;                     // 4 * nh, same-level octagon unknown or -1
    // coarse-target lists: rows receiving half-tilt face terms, in q order
    std::vector<int32_t> hc_row, hc_start, hc_q;      // hc_start size = rows+1
    std::vector<uint8_t> hc_pos_diag, hc_pos_off;     // per list entry
    // half rows: occurrences per row, in q order
    std::vector<int32_t> hr_row, hr_start, hr_q;
    std::vector<uint8_t> hr_pos;                      // 6 per occurrence: coarse, 4 same-level, diag

    // CSR pattern of A
    std::vector<int32_t> rowstart, col;

    // velocity correction
    std::vector<int32_t> vm_node, vm_c;      // MAC: leaves
    std::vector<int32_t> vm_nb;              // 3 * n_vm, code
    std::vector<int32_t> vt_node, vt_c;      // open leaf tilts
    std::vector<int32_t> vt_nb;              // 8 * n_vt
    std::vector<uint8_t> vt_lev;             // 8 * n_vt
    bool half_duplicates = false;

    // builder scratch: pattern rows under construction, kept with the tables
    // so that a rebuild (every tilt update) reuses their memory
    std::vector<int32_t> scratch_rows;               // n * 16
    std::vector<uint8_t> scratch_len;                // n
    std::vector<std::pair<int, int>> scratch_extra;  // shared (row, column) pairs
};

void buildProjectTables(const Octree& o, const Numbering& num, const Weights& w, ProjectTables& t);

// CPU-side assembly from the tables (same arithmetic as the device kernels).
// Used only to cross-check the device kernels in tests through the C ABI.
}  // namespace astg
