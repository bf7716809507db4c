// Adaptive staggered-tilted octree: host-side topology in the reference's node
// order, with dense per-level index maps in place of the reference's hash maps.
//
// Node model (one record per layer node, leaves first then ghosts), as in
// OctreeLayer3 (reference include/octree_layer3.h:11-80, src/octree_layer3.cpp:7-88):
//   * level l has resolution R>>l and spacing h0*2^l;
//   * every node carries an octagon (cell, min corner at the node) and a tilt
//     (vertex at the node);
//   * a dense map level_map[l][(z*res+y)*res+x] -> node index or -1 replaces
//     the flat_hash_map lookup; node(l,x,y,z) out of range is -1.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace astg {

enum NodeFlag : uint8_t { NF_VALID = 1, NF_GHOST = 2, NF_FROM_COARSE = 4 };

// Tilt state bits, mirroring TiltENode / TiltENodeS (reference include/tilt_e_node.h:7-67).
enum TiltBit : uint8_t {
    TB_LARGEST = 1,
    TB_CLOSED = 2,
    TB_FIXED = 4,
    TB_TJOINT = 8,
    TB_CLOSED_BY_TJ = 16,
    TB_HALF = 32,
    TB_LAYOUT_OPEN = 64,   // tilt_layout is_open (TiltENodeS)
    TB_LAYOUT_FIXED = 128  // tilt_layout is_fixed (TiltENodeS)
};

struct Level {
    int res = 0;
    double h = 0;
    int nleaf = 0;
    int nnode = 0;
    std::vector<uint16_t> x, y, z;
    std::vector<uint8_t> flag;
    std::vector<int32_t> map;  // res^3 dense index map

    int get(int xi, int yi, int zi) const {
        if (xi < 0 || yi < 0 || zi < 0 || xi >= res || yi >= res || zi >= res) return -1;
        return map[((size_t)zi * res + yi) * res + xi];
    }
    bool ghost(int i) const { return (flag[i] & NF_GHOST) != 0; }
};

// Packed (level, x, y, z); -1 means "none". Same field widths as the
// reference's LCOOR (macro.h:37-45) so that c>0 / c>=0 tests keep their meaning.
typedef int64_t lcoor_t;
inline lcoor_t pack(int l, int x, int y, int z) {
    return ((int64_t)l << 60) | ((int64_t)x << 40) | ((int64_t)y << 20) | (int64_t)z;
}
inline int lc_level(lcoor_t c) { return (int)((c >> 60) & 0xf); }
inline int lc_x(lcoor_t c) { return (int)((c >> 40) & 0xfffff); }
inline int lc_y(lcoor_t c) { return (int)((c >> 20) & 0xfffff); }
inline int lc_z(lcoor_t c) { return (int)(c & 0xfffff); }

// Boundary (collider) signed distance on the finest grid, as the reference's
// colliderSdf (CellCenteredScalarGrid3 of resolution R and spacing h0,
// omni_boundary_solver3.cpp:11-31): R^3 values at the cell centres
// ((i+0.5) h0), x fastest; negative inside the collider. Absent (empty) means
// no collider: the reference then samples a constant positive SDF.
struct BoundarySdf {
    int R = 0;
    double h0 = 0;
    std::vector<double> phi;
    // centre the slip constraint takes its normals from (the reference
    // hard-codes (0.5, 0.5, 0.45), omni_boundary_solver3.h:27,57,77)
    double slip[3] = {0.5, 0.5, 0.45};
    bool present() const { return !phi.empty(); }
    // CellCenteredGrid::sample (cell_centered_grid.h:40-54): clamped
    // trilinear interpolation of the cell-centred values.
    double sample(double x, double y, double z) const;
};

struct TiltNode {
    double e = 0;
    uint8_t bits = TB_CLOSED;  // default TiltENode(0,false,true,false): closed, e=0
    int8_t half_dir = 0;
};

class Octree {
public:
    int R = 0, L = 0;
    double h0 = 0;
    int closed_flags = 0;  // reference kDirection* bits (constants.h:11-26)
    std::vector<Level> lv;
    std::vector<std::vector<TiltNode>> tilt;  // tiltE + tilt_layout bits per node
    std::vector<lcoor_t> half;                // halfGhostIdx order
    std::vector<lcoor_t> closed_by_half;      // closedByHalfTiltArr order
    BoundarySdf sdf;                          // collider SDF (empty: none)

    // Builds layers, ghosts, T-joint / half-tilt state and the initData
    // boundary fix-ups from a split map (R^3 ints, x fastest).
    void build(const int* split, int R, double h0, int levels, int closedFlags);

    // Opens (e = h/2, largest) or closes non-fixed leaf tilts of one level
    // following adjustE's commit step (octree_omnigrid_solver3.cpp:1150-1165).
    void setTiltOpen(int level, const uint8_t* openMask);
    // The same for every level at once (masks: leaves of each level, levels
    // ascending), ghosts updated once at the end; returns whether any leaf
    // tilt changed.
    bool setTiltOpenAll(const uint8_t* openMasks);

    int node(int l, int x, int y, int z) const { return (l < 0 || l >= L) ? -1 : lv[l].get(x, y, z); }
    lcoor_t getUntil(int l, int x, int y, int z, bool toFinest) const;
    long totalNodes() const;

    std::string err;

private:
    void buildLayer(int l, const int* split);
    void constrainTilt(std::vector<std::vector<TiltNode>>& t, bool record);
    void updateTiltGhost(std::vector<std::vector<TiltNode>>& t, bool layoutBits);
    void initFixups();
};

// Unknown numbering (reference updateIndexParallel, octree_omnigrid_solver3.cpp:292-347):
// a cell (octagon), open tilt or half-tilt is an unknown only where the
// boundary SDF at its position exceeds -3 h_l (always, with no collider).
struct Numbering {
    std::vector<std::vector<int32_t>> oct, tilt;  // per level, per node
    int n_oct_tilt = 0;  // rows before the half-tilt block
    int n = 0;
};
void numberUnknowns(const Octree& o, Numbering& num);

// Reference octagon-grid validity: 0 <= v <= res-2 (octree_cell_centered_grid3.h:35).
inline bool octValid(const Level& L, int x, int y, int z) {
    return x >= 0 && y >= 0 && z >= 0 && x < L.res - 1 && y < L.res - 1 && z < L.res - 1;
}

}  // namespace astg
