// Cutting-cell weights without a collider.
//
// initCuttingCellWeights (reference src/octree_omnigrid_solver3.cpp:1169-1238)
// sets weight = 1 - fractionInside(phi_a, phi_b) (util.h:340-349). With the
// boundary SDF positive everywhere every computed weight is 1, and the grid
// starts at 1 (initData 77-78), so only the half-tilt records differ: they are
// reset to zero and get 1 in the face slot (if the coarse octagon is valid) and
// in the 4 same-level slots whose octagon is valid.
#include "project_host.h"

namespace astg {

// HALF_TILT_SLOT_INDEX (reference constants.h:111-118): per half direction,
// the 4 same-level slots then the face slot.
const int kHalfSlot[6][5] = {{0, 3, 4, 7, 1}, {0, 1, 4, 5, 2}, {0, 1, 2, 3, 5},
                             {1, 2, 5, 6, 0}, {2, 3, 6, 7, 4}, {4, 5, 6, 7, 3}};
// TILT_NB_OCT_DXDYDZ (constants.h:102-105): octagon of tilt slot k.
const int kTiltNbOct[8][3] = {{0, 0, 0}, {-1, 0, 0}, {-1, -1, 0}, {0, -1, 0},
                              {0, 0, -1}, {-1, 0, -1}, {-1, -1, -1}, {0, -1, -1}};
// OCT_NB_TILT_DXDYDZ (constants.h:90-93): tilt corner k of an octagon.
const int kOctNbTilt[8][3] = {{1, 1, 1}, {0, 1, 1}, {0, 0, 1}, {1, 0, 1},
                              {1, 1, 0}, {0, 1, 0}, {0, 0, 0}, {1, 0, 0}};
// OCT_NB_DXDYDZ (constants.h:85-88) and OCT_NB_FACE_DXDYDZ (96-99).
const int kOctNb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
const int kOctNbFace[6][3] = {{1, 0, 0}, {0, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 1}, {0, 0, 0}};

void computeWeights(const Octree& o, Weights& w) {
    w.mac.assign(o.L, {});
    w.tilt.assign(o.L, {});
    for (int l = 0; l < o.L; ++l) {
        w.mac[l].assign((size_t)o.lv[l].nnode * 3, 1.0);
        w.tilt[l].assign((size_t)o.lv[l].nnode * 8, 1.0);
    }
    for (lcoor_t h : o.half) {
        const int l = lc_level(h), x = lc_x(h), y = lc_y(h), z = lc_z(h);
        const int i = o.node(l, x, y, z);
        const int hd = o.tilt[l][i].half_dir;
        double* t = &w.tilt[l][(size_t)i * 8];
        for (int k = 0; k < 8; ++k) t[k] = 0.0;
        if (hd < 0 || hd > 5) continue;
        int c[3] = {x >> 1, y >> 1, z >> 1};
        c[hd % 3] -= hd < 3 ? 1 : 0;
        if (l + 1 < o.L && octValid(o.lv[l + 1], c[0], c[1], c[2])) t[kHalfSlot[hd][4]] = 1.0;
        for (int m = 0; m < 4; ++m) {
            const int k = kHalfSlot[hd][m];
            if (octValid(o.lv[l], x + kTiltNbOct[k][0], y + kTiltNbOct[k][1], z + kTiltNbOct[k][2])) t[k] = 1.0;
        }
    }
}

}  // namespace astg
