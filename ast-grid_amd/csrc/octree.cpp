// Host-side octree topology (see octree.h). Each routine cites the reference
// routine whose observable result it reproduces; the data structures differ
// (dense per-level maps instead of hash maps, one flat record per tilt).
#include "octree.h"

#include <algorithm>
#include <cstring>

namespace astg {

namespace {
// +x, -x, +y, -y, +z, -z (reference OCT_NB_DXDYDZ, constants.h:85-88)
const int kFace[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

const uint8_t kTiltMask = 0x3f;     // TiltENode fields
const uint8_t kLayoutMask = 0xc0;   // TiltENodeS fields

// TiltENode::setE with maxE (tilt_e_node.h:41-56).
void setE(TiltNode& t, double e, double maxE = -1) {
    if (t.bits & TB_FIXED) return;
    if (e <= 0) {
        t.bits |= TB_CLOSED;
        t.bits &= ~TB_LARGEST;
        e = 0;
    } else if (maxE > 0 && e >= maxE) {
        t.bits &= ~TB_CLOSED;
        t.bits |= TB_LARGEST;
        e = maxE;
    } else {
        t.bits &= ~(TB_CLOSED | TB_LARGEST);
    }
    t.e = e;
}
}  // namespace

long Octree::totalNodes() const {
    long s = 0;
    for (auto& l : lv) s += l.nnode;
    return s;
}

// OctreeLayer3 constructor (octree_layer3.cpp:7-88): leaves in (z, y, x) order,
// then ghosts in discovery order over leaves x 26 neighbours (x outer, z inner).
void Octree::buildLayer(int l, const int* split) {
    Level& L = lv[l];
    L.res = R >> l;
    L.h = h0 * (double)(1 << l);
    const int res = L.res;
    L.map.assign((size_t)res * res * res, -1);
    auto sp = [&](int x, int y, int z) { return split[(((size_t)(z << l)) * R + (y << l)) * R + (x << l)]; };
    for (int z = 0; z < res; ++z)
        for (int y = 0; y < res; ++y)
            for (int x = 0; x < res; ++x)
                if (sp(x, y, z) == l) {
                    L.map[((size_t)z * res + y) * res + x] = (int32_t)L.x.size();
                    L.x.push_back((uint16_t)x);
                    L.y.push_back((uint16_t)y);
                    L.z.push_back((uint16_t)z);
                    L.flag.push_back(NF_VALID);
                }
    L.nleaf = (int)L.x.size();
    for (int i = 0; i < L.nleaf; ++i) {
        const int x0 = L.x[i], y0 = L.y[i], z0 = L.z[i];
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    if (!dx && !dy && !dz) continue;
                    int nx = x0 + dx, ny = y0 + dy, nz = z0 + dz;
                    if (nx < 0 || ny < 0 || nz < 0 || nx >= res || ny >= res || nz >= res) continue;
                    int32_t& slot = L.map[((size_t)nz * res + ny) * res + nx];
                    if (slot != -1) continue;
                    int tv = sp(nx, ny, nz);
                    uint8_t f;
                    if (tv == l - 1) f = NF_VALID | NF_GHOST;
                    else if (tv == l + 1) f = NF_VALID | NF_GHOST | NF_FROM_COARSE;
                    else continue;
                    slot = (int32_t)L.x.size();
                    L.x.push_back((uint16_t)nx);
                    L.y.push_back((uint16_t)ny);
                    L.z.push_back((uint16_t)nz);
                    L.flag.push_back(f);
                }
    }
    L.nnode = (int)L.x.size();
}

// OctreeGridLayout3::getUntil (octree_grid_layout3.cpp:51-69).
lcoor_t Octree::getUntil(int l, int x, int y, int z, bool toFinest) const {
    const int res = lv[l].res;
    if (x < 0 || y < 0 || z < 0 || x >= res || y >= res || z >= res) return -1;
    for (;;) {
        int i = lv[l].get(x, y, z);
        if (i >= 0 && !lv[l].ghost(i)) break;
        if (toFinest) {
            x <<= 1; y <<= 1; z <<= 1;
            l -= 1;
        } else {
            x >>= 1; y >>= 1; z >>= 1;
            l += 1;
        }
        if (l < 0 || l >= L) return -1;
    }
    return pack(l, x, y, z);
}

// OctreeVertexCenteredGrid3<TiltENode / TiltENodeS>::updateGhost
// (octree_vertex_centered_grid3.cpp:5-44): from-coarse ghosts at even
// coordinates copy the coarse node, odd ones are left alone, from-fine ghosts
// copy the fine node at (2x,2y,2z). Sources are always leaves, so the order of
// the reference's parallel loop does not matter.
void Octree::updateTiltGhost(std::vector<std::vector<TiltNode>>& t, bool layoutBits) {
    for (int l = 0; l < L; ++l) {
        const Level& Lv = lv[l];
        for (int i = Lv.nleaf; i < Lv.nnode; ++i) {
            int x = Lv.x[i], y = Lv.y[i], z = Lv.z[i];
            int src;
            int sl;
            if (Lv.flag[i] & NF_FROM_COARSE) {
                if ((x & 1) || (y & 1) || (z & 1)) continue;
                sl = l + 1;
                src = node(sl, x >> 1, y >> 1, z >> 1);
            } else {
                sl = l - 1;
                src = node(sl, x << 1, y << 1, z << 1);
            }
            if (src < 0) continue;  // the reference would read out of bounds here
            const TiltNode& s = t[sl][src];
            TiltNode& d = t[l][i];
            if (layoutBits) {
                d.bits = (uint8_t)((d.bits & kTiltMask) | (s.bits & kLayoutMask));
            } else {
                d.e = s.e;
                d.half_dir = s.half_dir;
                d.bits = (uint8_t)((d.bits & kLayoutMask) | (s.bits & kTiltMask));
            }
        }
    }
}

// constrainOctreeTiltE3 (octree_util.cpp:38-169).
void Octree::constrainTilt(std::vector<std::vector<TiltNode>>& t, bool record) {
    // (1) T-joint detection over the 8 cells incident to each leaf vertex.
    for (int l = 0; l < L; ++l) {
        const Level& Lv = lv[l];
        for (int i = 0; i < Lv.nleaf; ++i) {
            int x = Lv.x[i], y = Lv.y[i], z = Lv.z[i];
            int minL = l, maxL = l;
            for (int a = -1; a <= 0; ++a)
                for (int b = -1; b <= 0; ++b)
                    for (int c = -1; c <= 0; ++c) {
                        lcoor_t q = getUntil(l, x + a, y + b, z + c, true);
                        if (q < 0) q = getUntil(l, x + a, y + b, z + c, false);
                        if (q > 0) {  // strict: the node (0,0,0,0) never counts
                            minL = std::min(minL, lc_level(q));
                            maxL = std::max(maxL, lc_level(q));
                        }
                    }
            if (minL < l || (maxL > l && !(x & 1) && !(y & 1) && !(z & 1))) {
                TiltNode& n = t[l][i];
                n.bits |= TB_TJOINT | TB_FIXED;
                n.bits &= ~TB_CLOSED;
                n.e = lv[minL].h;
            }
        }
    }
    // (2) close the 6 axis neighbours of every T-joint tilt.
    for (int l = 0; l < L; ++l) {
        const Level& Lv = lv[l];
        for (int i = 0; i < Lv.nleaf; ++i) {
            const TiltNode n = t[l][i];
            if (!(n.bits & TB_TJOINT)) continue;
            int tl = l, ix = Lv.x[i], iy = Lv.y[i], iz = Lv.z[i];
            if (n.e < lv[l].h) {
                tl -= 1;
                ix *= 2; iy *= 2; iz *= 2;
            }
            for (int f = 0; f < 6; ++f) {
                int j = node(tl, ix + kFace[f][0], iy + kFace[f][1], iz + kFace[f][2]);
                if (j < 0) continue;
                TiltNode& m = t[tl][j];
                setE(m, 0);
                m.bits |= TB_FIXED | TB_CLOSED_BY_TJ;
            }
        }
    }
    updateTiltGhost(t, false);
    // (3) half-tilts at the centre of coarse faces whose 4 corners are T-joints.
    for (int l = 0; l < L; ++l) {
        const Level& Lv = lv[l];
        for (int i = 0; i < Lv.nleaf; ++i) {
            const TiltNode n = t[l][i];
            if (!(n.bits & TB_TJOINT)) continue;
            const int isCoarse = n.e < lv[l].h ? 1 : 0;
            int ci[3];
            if (isCoarse) {
                ci[0] = Lv.x[i]; ci[1] = Lv.y[i]; ci[2] = Lv.z[i];
            } else {
                ci[0] = Lv.x[i] >> 1; ci[1] = Lv.y[i] >> 1; ci[2] = Lv.z[i] >> 1;
            }
            const int ql = l + (1 - isCoarse);   // level the coarse index lives on
            const int rl = l - isCoarse;         // level compared for the border test (as in the reference)
            for (int axis = 0; axis < 3; ++axis) {
                if (ci[axis] == 0 || ci[axis] == lv[rl].res - 1) continue;
                bool flag = true;
                for (int a = 0; a <= 1 && flag; ++a)
                    for (int b = 0; b <= 1 && flag; ++b) {
                        if (!a && !b) continue;
                        int id[3] = {ci[0], ci[1], ci[2]};
                        id[(axis + 1) % 3] += a;
                        id[(axis + 2) % 3] += b;
                        lcoor_t q = getUntil(ql, id[0], id[1], id[2], true);
                        if (q < 0) q = getUntil(ql, id[0], id[1], id[2], false);
                        if (q >= 0) {
                            int qi = node(lc_level(q), lc_x(q), lc_y(q), lc_z(q));
                            if (!(t[lc_level(q)][qi].bits & TB_TJOINT)) flag = false;
                        } else {
                            flag = false;
                        }
                    }
                if (flag) {
                    int nb[3] = {ci[0], ci[1], ci[2]};
                    nb[axis] -= 1;
                    lcoor_t c0 = getUntil(ql, ci[0], ci[1], ci[2], true);
                    lcoor_t c1 = getUntil(ql, nb[0], nb[1], nb[2], true);
                    if (c1 < 0 || lc_level(c0) == lc_level(c1)) flag = false;
                }
                if (!flag) continue;
                int id[3] = {ci[0] * 2, ci[1] * 2, ci[2] * 2};
                id[(axis + 1) % 3] += 1;
                id[(axis + 2) % 3] += 1;
                int hj = node(rl, id[0], id[1], id[2]);
                if (hj < 0) continue;
                TiltNode& hn = t[rl][hj];
                hn.bits |= TB_CLOSED | TB_HALF | TB_FIXED;
                hn.half_dir = (int8_t)(axis + (isCoarse ? 3 : 0));
                hn.e = lv[rl].h;
                if (record) half.push_back(pack(rl, id[0], id[1], id[2]));
                int td[3] = {id[0], id[1], id[2]};
                td[axis] += hn.half_dir < 3 ? 1 : -1;
                lcoor_t tc = pack(rl, td[0], td[1], td[2]);
                if (record && std::find(closed_by_half.begin(), closed_by_half.end(), tc) == closed_by_half.end())
                    closed_by_half.push_back(tc);
                int tj = node(rl, td[0], td[1], td[2]);
                if (tj < 0) {
                    err = "closed-by-half-tilt node missing";
                    continue;
                }
                TiltNode& cn = t[rl][tj];
                setE(cn, 0);
                cn.bits |= TB_CLOSED_BY_TJ | TB_FIXED;
                cn.half_dir = -1;
            }
        }
    }
}

// initData's tilt fix-ups (octree_omnigrid_solver3.cpp:128-145): T-joint and
// fixed tilts copy their open state into tilt_layout; tilts on a level's
// outer vertex layer are closed and fixed.
void Octree::initFixups() {
    for (int l = 0; l < L; ++l) {
        const Level& Lv = lv[l];
        for (int i = 0; i < Lv.nleaf; ++i) {
            TiltNode& n = tilt[l][i];
            if ((n.bits & TB_TJOINT) || (n.bits & TB_FIXED)) {
                n.bits = (uint8_t)(n.bits & ~(TB_LAYOUT_OPEN | TB_LAYOUT_FIXED));
                if (!(n.bits & TB_CLOSED)) n.bits |= TB_LAYOUT_OPEN;
                n.bits |= TB_LAYOUT_FIXED;
                continue;
            }
            int x = Lv.x[i], y = Lv.y[i], z = Lv.z[i], r = Lv.res;
            if (x == 0 || y == 0 || z == 0 || x == r - 1 || y == r - 1 || z == r - 1) {
                setE(n, 0);
                n.bits |= TB_FIXED;
                n.bits = (uint8_t)(n.bits & ~TB_LAYOUT_OPEN);
                if (!(n.bits & TB_CLOSED)) n.bits |= TB_LAYOUT_OPEN;
                n.bits |= TB_LAYOUT_FIXED;
            }
        }
    }
}

void Octree::build(const int* split, int R_, double h0_, int levels, int closedFlags) {
    R = R_;
    h0 = h0_;
    L = levels;
    closed_flags = closedFlags;
    lv.assign(L, Level());
    for (int l = 0; l < L; ++l) buildLayer(l, split);
    tilt.assign(L, {});
    for (int l = 0; l < L; ++l) tilt[l].assign(lv[l].nnode, TiltNode());
    half.clear();
    closed_by_half.clear();
    // initData 92-93: constrain with recording, then one more ghost pass.
    constrainTilt(tilt, true);
    updateTiltGhost(tilt, false);
    initFixups();
}

void Octree::setTiltOpen(int l, const uint8_t* openMask) {
    const Level& Lv = lv[l];
    const double maxE = Lv.h * 0.5;
    for (int i = 0; i < Lv.nleaf; ++i) {
        TiltNode& n = tilt[l][i];
        if (n.bits & TB_LAYOUT_FIXED) continue;
        if (openMask[i]) n.bits |= TB_LAYOUT_OPEN;
        else n.bits &= ~TB_LAYOUT_OPEN;
        if (openMask[i]) setE(n, maxE, maxE);
        else setE(n, 0);
    }
    updateTiltGhost(tilt, true);
    updateTiltGhost(tilt, false);
}

// adjustE's commit (octree_omnigrid_solver3.cpp:1150-1165) after its decision
// loop (1131-1149) has set the layout flags of every level.
bool Octree::setTiltOpenAll(const uint8_t* openMasks) {
    bool changed = false;
    for (int l = 0, base = 0; l < L; base += lv[l].nleaf, ++l) {
        const Level& Lv = lv[l];
        const double maxE = Lv.h * 0.5;
        for (int i = 0; i < Lv.nleaf; ++i) {
            TiltNode& n = tilt[l][i];
            if (n.bits & TB_LAYOUT_FIXED) continue;
            const bool open = openMasks[base + i] != 0;
            if (open != ((n.bits & TB_LAYOUT_OPEN) != 0)) changed = true;
            if (open) n.bits |= TB_LAYOUT_OPEN;
            else n.bits &= ~TB_LAYOUT_OPEN;
            if (open) setE(n, maxE, maxE);
            else setE(n, 0);
        }
    }
    updateTiltGhost(tilt, true);
    updateTiltGhost(tilt, false);
    return changed;
}

// updateIndexParallel (octree_omnigrid_solver3.cpp:292-347). The per-core
// counts plus prefix sum give the sequential order of the leaf iterator
// (levels ascending, leaf order), octagon before tilt at each node; then the
// half-tilts in halfGhostIdx order; then the int ghost rule
// (octree_vertex_centered_grid3.cpp:47-64) on the tilt indices.
// getBaryCentric, vector form (util.h:244-257): truncation, index and
// fraction clamped.
static inline void baryClamp(double x, int high, int& i, double& f) {
    i = (int)x;
    i = std::max(i, 0);
    i = std::min(i, high - 1);
    f = x - (double)i;
    f = std::max(f, 0.0);
    f = std::min(f, 1.0);
}

double BoundarySdf::sample(double x, double y, double z) const {
    if (phi.empty()) return 1e4;
    int i[3];
    double f[3];
    const double p[3] = {x / h0 - 0.5, y / h0 - 0.5, z / h0 - 0.5};
    for (int d = 0; d < 3; ++d) baryClamp(p[d], R - 1, i[d], f[d]);
    const int j[3] = {std::min(i[0] + 1, R - 1), std::min(i[1] + 1, R - 1), std::min(i[2] + 1, R - 1)};
    auto g = [&](int a, int b, int c) { return phi[((size_t)c * R + b) * R + a]; };
    auto lerp = [](double v0, double v1, double t) { return v0 * (1 - t) + v1 * t; };
    // trilerp (util.h:165-198): x within each (y, z) pair, then y, then z
    const double c00 = lerp(g(i[0], i[1], i[2]), g(j[0], i[1], i[2]), f[0]);
    const double c10 = lerp(g(i[0], j[1], i[2]), g(j[0], j[1], i[2]), f[0]);
    const double c01 = lerp(g(i[0], i[1], j[2]), g(j[0], i[1], j[2]), f[0]);
    const double c11 = lerp(g(i[0], j[1], j[2]), g(j[0], j[1], j[2]), f[0]);
    return lerp(lerp(c00, c10, f[1]), lerp(c01, c11, f[1]), f[2]);
}

void numberUnknowns(const Octree& o, Numbering& num) {
    num.oct.assign(o.L, {});
    num.tilt.assign(o.L, {});
    const BoundarySdf& B = o.sdf;
    // the collider cut: sdf(position) > -3 h_l (octree_omnigrid_solver3.cpp:294-315)
    auto keep = [&](int l, double x, double y, double z) {
        return !B.present() || B.sample(x, y, z) > -o.lv[l].h * 3;
    };
    int c = 0;
    for (int l = 0; l < o.L; ++l) {
        const Level& Lv = o.lv[l];
        const double h = Lv.h;
        num.oct[l].assign(Lv.nnode, -1);
        num.tilt[l].assign(Lv.nnode, -1);
        for (int i = 0; i < Lv.nleaf; ++i) {
            const int x = Lv.x[i], y = Lv.y[i], z = Lv.z[i];
            if (octValid(Lv, x, y, z) && keep(l, (x + 0.5) * h, (y + 0.5) * h, (z + 0.5) * h)) num.oct[l][i] = c++;
            if ((o.tilt[l][i].bits & TB_LAYOUT_OPEN) && keep(l, x * h, y * h, z * h)) num.tilt[l][i] = c++;
        }
    }
    num.n_oct_tilt = c;
    for (lcoor_t hc : o.half) {
        int l = lc_level(hc);
        int i = o.node(l, lc_x(hc), lc_y(hc), lc_z(hc));
        const double h = o.lv[l].h;
        if (keep(l, lc_x(hc) * h, lc_y(hc) * h, lc_z(hc) * h)) num.tilt[l][i] = c++;
        else num.tilt[l][i] = -1;
    }
    num.n = c;
    for (int l = 0; l < o.L; ++l) {
        const Level& Lv = o.lv[l];
        for (int i = Lv.nleaf; i < Lv.nnode; ++i) {
            int x = Lv.x[i], y = Lv.y[i], z = Lv.z[i];
            if (Lv.flag[i] & NF_FROM_COARSE) {
                if ((x & 1) || (y & 1) || (z & 1)) continue;
                int s = o.node(l + 1, x >> 1, y >> 1, z >> 1);
                if (s >= 0) num.tilt[l][i] = num.tilt[l + 1][s];
            } else {
                int s = o.node(l - 1, x << 1, y << 1, z << 1);
                if (s >= 0) num.tilt[l][i] = num.tilt[l - 1][s];
            }
        }
    }
}

}  // namespace astg
