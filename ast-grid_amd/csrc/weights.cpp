// Cutting-cell weights (initCuttingCellWeights, reference
// src/octree_omnigrid_solver3.cpp:1169-1238): weight = 1 - fractionInside(phi_a,
// phi_b) (util.h:340-349) of the boundary SDF at the two cell centres a face
// or tilt slot joins. The grid starts at 1 (initData 77-78); leaf MAC faces
// and leaf tilt slots whose neighbouring cell is valid are overwritten, and
// half-tilt records are reset to zero and get their face slot (if the coarse
// octagon is valid) and the 4 same-level slots whose octagon is valid. With
// no collider every computed weight is 1.
//
// Two quirks of the reference are kept: the coarse octagon of a half-tilt is
// positioned with the half-tilt's level instead of level+1 (1223), and the 4
// same-level slots of a half-tilt sample the SDF at the stale position of the
// previous step (1232): the coarse octagon's if it was valid, else the
// half-tilt's own vertex.
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

namespace {
double fractionInside(double a, double b) {
    if (a < 0 && b < 0) return 1;
    if (a < 0 && b >= 0) return a / (a - b);
    if (a >= 0 && b < 0) return b / (b - a);
    return 0;
}
}  // namespace

void computeWeights(const Octree& o, Weights& w) {
    const BoundarySdf& B = o.sdf;
    w.mac.assign(o.L, {});
    w.tilt.assign(o.L, {});
    for (int l = 0; l < o.L; ++l) {
        w.mac[l].assign((size_t)o.lv[l].nnode * 3, 1.0);
        w.tilt[l].assign((size_t)o.lv[l].nnode * 8, 1.0);
    }
    if (B.present()) {
        for (int l = 0; l < o.L; ++l) {
            const Level& Lv = o.lv[l];
            const double h = Lv.h;
            // SDF at the centre of cell (x, y, z) of level l
            auto cell = [&](int x, int y, int z) { return B.sample((x + 0.5) * h, (y + 0.5) * h, (z + 0.5) * h); };
            for (int i = 0; i < Lv.nleaf; ++i) {
                const int x = Lv.x[i], y = Lv.y[i], z = Lv.z[i];
                const double cur = cell(x, y, z);
                double* m = &w.mac[l][(size_t)i * 3];
                // MAC faces (isValidX/Y/Z, octree_face_centered_grid3.h:29-31)
                const int r = Lv.res;
                if (x <= r - 1 && y <= r - 2 && z <= r - 2 && octValid(Lv, x - 1, y, z))
                    m[0] = 1 - fractionInside(cell(x - 1, y, z), cur);
                if (x <= r - 2 && y <= r - 1 && z <= r - 2 && octValid(Lv, x, y - 1, z))
                    m[1] = 1 - fractionInside(cell(x, y - 1, z), cur);
                if (x <= r - 2 && y <= r - 2 && z <= r - 1 && octValid(Lv, x, y, z - 1))
                    m[2] = 1 - fractionInside(cell(x, y, z - 1), cur);
                // tilt slots: the vertex against each of its 8 cells
                const double vtx = B.sample(x * h, y * h, z * h);
                double* t = &w.tilt[l][(size_t)i * 8];
                for (int k = 0; k < 8; ++k) {
                    const int nx = x + kTiltNbOct[k][0], ny = y + kTiltNbOct[k][1], nz = z + kTiltNbOct[k][2];
                    if (octValid(Lv, nx, ny, nz)) t[k] = 1 - fractionInside(vtx, cell(nx, ny, nz));
                }
            }
        }
    }
    for (lcoor_t hc : o.half) {
        const int l = lc_level(hc), x = lc_x(hc), y = lc_y(hc), z = lc_z(hc);
        const int i = o.node(l, x, y, z);
        const int hd = o.tilt[l][i].half_dir;
        double* t = &w.tilt[l][(size_t)i * 8];
        for (int k = 0; k < 8; ++k) t[k] = 0.0;
        if (hd < 0 || hd > 5) continue;
        const double h = o.lv[l].h;
        double pos[3] = {x * h, y * h, z * h};
        const double cur = B.present() ? B.sample(pos[0], pos[1], pos[2]) : 1.0;
        int c[3] = {x >> 1, y >> 1, z >> 1};
        c[hd % 3] -= hd < 3 ? 1 : 0;
        if (l + 1 < o.L && octValid(o.lv[l + 1], c[0], c[1], c[2])) {
            // positioned at level l, not l + 1 (the reference's quirk)
            pos[0] = (c[0] + 0.5) * h;
            pos[1] = (c[1] + 0.5) * h;
            pos[2] = (c[2] + 0.5) * h;
            t[kHalfSlot[hd][4]] = B.present() ? 1 - fractionInside(cur, B.sample(pos[0], pos[1], pos[2])) : 1.0;
        }
        for (int m = 0; m < 4; ++m) {
            const int k = kHalfSlot[hd][m];
            if (octValid(o.lv[l], x + kTiltNbOct[k][0], y + kTiltNbOct[k][1], z + kTiltNbOct[k][2]))
                // the stale position (the reference's quirk)
                t[k] = B.present() ? 1 - fractionInside(cur, B.sample(pos[0], pos[1], pos[2])) : 1.0;
        }
    }
}

}  // namespace astg
