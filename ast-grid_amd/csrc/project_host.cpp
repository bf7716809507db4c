// Stencil tables and the CSR pattern of the projection matrix.
//
// The pattern follows the reference's row-list matrix semantics
// (PCG::SparseMatrix::set_element / add_to_element keep each row sorted and
// create an entry on first touch, include/pcgsolver/sparse_matrix.h:69-103):
// an entry exists for every (row, column) the assembly touches, whatever its
// value, so explicit zeros are kept exactly where the reference keeps them.
#include "project_host.h"
#include "parallel.h"
#include "timing.h"

#include <algorithm>
#include <climits>
#include <set>
#include <stdexcept>

namespace astg {

namespace {

// tiltNeighborOctagon3 (reference src/octree_util.cpp:172-194): the octagon
// seen by tilt slot k, found finest-first then coarsest-first, with the
// coordinate shift onto the tilt when the octagon is finer than the tilt.
lcoor_t tiltNeighbor(const Octree& o, int l, int x, int y, int z, int k) {
    int nx = x + kTiltNbOct[k][0], ny = y + kTiltNbOct[k][1], nz = z + kTiltNbOct[k][2];
    const int res = o.lv[l].res;
    if (nx < 0 || ny < 0 || nz < 0 || nx >= res || ny >= res || nz >= res) return -1;
    lcoor_t c = o.getUntil(l, nx, ny, nz, true);
    if (c < 0) c = o.getUntil(l, nx, ny, nz, false);
    if (c >= 0 && lc_level(c) < l) {
        int tl = lc_level(c), tx = lc_x(c), ty = lc_y(c), tz = lc_z(c);
        int m4 = k % 4;
        tz += k >= 4 ? 1 : 0;
        ty += m4 >= 2 ? 1 : 0;
        tx += (m4 == 1 || m4 == 2) ? 1 : 0;
        c = pack(tl, tx, ty, tz);
    }
    return c;
}

// Appends off + i for every i in [0, n) with keep(i), in increasing order, on
// the worker pool: fixed chunks counted, offsets by a prefix sum, then filled.
template <class P>
void compactAppend(int n, int off, P keep, std::vector<int>& out) {
    constexpr int S = 256;
    auto lo = [&](long c) { return (int)((long long)n * c / S); };
    std::vector<long> cnt(S + 1, 0);
    parallelFor(S, [&](long c0, long c1) {
        for (long c = c0; c < c1; ++c) {
            long k = 0;
            for (int i = lo(c); i < lo(c + 1); ++i) k += keep(i) ? 1 : 0;
            cnt[c + 1] = k;
        }
    }, 1);
    for (int c = 0; c < S; ++c) cnt[c + 1] += cnt[c];
    const size_t base = out.size();
    out.resize(base + (size_t)cnt[S]);
    parallelFor(S, [&](long c0, long c1) {
        for (long c = c0; c < c1; ++c) {
            size_t w = base + (size_t)cnt[c];
            for (int i = lo(c); i < lo(c + 1); ++i)
                if (keep(i)) out[w++] = off + i;
        }
    }, 1);
}

// Pattern rows under construction. Octagon and tilt rows (each its own
// unknown, at most 15 entries: 6 faces, 8 tilt corners, the diagonal) are
// sorted lists in a flat array with stride kRowCap, each written by one
// thread; the half-tilt couplings, which several half-tilts add to one
// coarse row, are (row, column) pairs merged in when the CSR is formed.
constexpr int kRowCap = 16;
// rows per thread at least, for the loops that look neighbours up per row
constexpr long kRowSlice = 2048;
struct RowSet {
    int32_t* c;                                // n * kRowCap, only the first k[r] of a row are defined
    std::vector<uint8_t>& k;                   // n
    std::vector<std::pair<int, int>>& extra;
    RowSet(int n, ProjectTables& t) : k(t.scratch_len), extra(t.scratch_extra) {
        if (t.scratch_rows.size() < (size_t)n * kRowCap) t.scratch_rows.resize((size_t)n * kRowCap);
        c = t.scratch_rows.data();
        k.assign((size_t)n, 0);
        extra.clear();
    }
    void touch(int r, int col) {
        int32_t* v = c + (size_t)r * kRowCap;
        const int m = k[r];
        const int p = (int)(std::lower_bound(v, v + m, col) - v);
        if (p < m && v[p] == col) return;
        if (m == kRowCap) throw std::runtime_error("pattern row longer than expected");
        for (int q = m; q > p; --q) v[q] = v[q - 1];
        v[p] = col;
        k[r] = (uint8_t)(m + 1);
    }
    void touchShared(int r, int col) { extra.push_back({r, col}); }
};

}  // namespace

void buildProjectTables(const Octree& o, const Numbering& num, const Weights& w, ProjectTables& t) {
    PhaseTimer tm("tables");
    const int L = o.L;
    // A rebuild reuses the previous tables' memory: every array below is
    // either filled (assign), cleared before it is appended to, or resized
    // and then written at every element.
    for (auto* v : {&t.hc_row, &t.hc_start, &t.hc_q, &t.hr_row, &t.hr_start, &t.hr_q, &t.vt_node}) v->clear();
    for (auto* v : {&t.hc_pos_diag, &t.hc_pos_off, &t.hr_pos}) v->clear();
    t.half_duplicates = false;
    t.off.assign(L + 1, 0);
    for (int l = 0; l < L; ++l) t.off[l + 1] = t.off[l] + o.lv[l].nnode;
    const int G = t.G = t.off[L];
    t.level.resize(G);
    t.tilt_e.resize(G);
    t.tilt_bits.resize(G);
    t.tilt_hd.resize(G);
    t.pid_oct.resize(G);
    t.pid_tilt.resize(G);
    t.w_mac.resize((size_t)3 * G);
    t.w_tilt.resize((size_t)8 * G);
    for (int l = 0; l < L; ++l)
        parallelFor(o.lv[l].nnode, [&](long i0, long i1) {
            for (int i = (int)i0; i < (int)i1; ++i) {
                const int g = t.off[l] + i;
                t.level[g] = (uint8_t)l;
                t.tilt_e[g] = o.tilt[l][i].e;
                t.tilt_bits[g] = o.tilt[l][i].bits;
                t.tilt_hd[g] = o.tilt[l][i].half_dir;
                t.pid_oct[g] = num.oct[l][i];
                t.pid_tilt[g] = num.tilt[l][i];
                for (int a = 0; a < 3; ++a) t.w_mac[(size_t)a * G + g] = w.mac[l][(size_t)i * 3 + a];
                for (int k = 0; k < 8; ++k) t.w_tilt[(size_t)k * G + g] = w.tilt[l][(size_t)i * 8 + k];
            }
        }, 16 * kRowSlice);
    tm.mark("nodes");
    t.n = num.n;
    RowSet rs(t.n, t);

    // ---- octagon and tilt rows (project 367-552) ----
    std::vector<int> octJ, tiltJ;
    for (int l = 0; l < L; ++l) {
        const int* no_ = num.oct[l].data();
        const int* nt_ = num.tilt[l].data();
        const auto* ts = o.tilt[l].data();
        compactAppend(o.lv[l].nleaf, t.off[l], [&](int i) { return no_[i] >= 0; }, octJ);
        compactAppend(o.lv[l].nleaf, t.off[l], [&](int i) { return nt_[i] >= 0 && (ts[i].bits & TB_LAYOUT_OPEN); },
                      tiltJ);
    }
    tm.mark("row lists");
    const int no = t.n_oct = (int)octJ.size();
    t.oct_row.resize(no);
    t.oct_corner.resize((size_t)8 * no);
    t.oct_face.resize((size_t)6 * no);
    t.oct_pos.assign((size_t)15 * no, 0xff);
    // every octagon row and every tilt row is its own unknown: the rows of the
    // pattern are written by one thread each
    parallelFor(no, [&](long j0, long j1) {
    for (int j = (int)j0; j < (int)j1; ++j) {
        const int g = octJ[j], l = t.level[g], i = g - t.off[l];
        const Level& Lv = o.lv[l];
        const int x = Lv.x[i], y = Lv.y[i], z = Lv.z[i];
        const int r = num.oct[l][i];
        t.oct_row[j] = r;
        for (int k = 0; k < 8; ++k) {
            int c = o.node(l, x + kOctNbTilt[k][0], y + kOctNbTilt[k][1], z + kOctNbTilt[k][2]);
            if (c < 0) throw std::runtime_error("octagon corner node missing (split map not balanced)");
            t.oct_corner[(size_t)k * no + j] = t.off[l] + c;
        }
        for (int f = 0; f < 6; ++f) {
            int nx = x + kOctNb[f][0], ny = y + kOctNb[f][1], nz = z + kOctNb[f][2];
            int32_t code;
            if (octValid(Lv, nx, ny, nz)) {
                int nb = Lv.get(nx, ny, nz);
                if (nb < 0) throw std::runtime_error("octagon face neighbour missing (split map not balanced)");
                if (Lv.ghost(nb)) code = FACE_SKIP;
                else code = num.oct[l][nb] >= 0 ? num.oct[l][nb] : FACE_EXCLUDED;
            } else {
                // isClosedDomain(i) maps face i to the kDirection bit (omni_boundary_solver3.h:164-182)
                static const int bit[6] = {1 << 1, 1 << 0, 1 << 5, 1 << 4, 1 << 3, 1 << 2};
                code = (o.closed_flags & bit[f]) ? FACE_CLOSED : FACE_OPEN;
            }
            t.oct_face[(size_t)f * no + j] = code;
            if (code >= 0) rs.touch(r, code);
        }
        for (int k = 0; k < 8; ++k) {
            const int g2 = t.oct_corner[(size_t)k * no + j];
            const uint8_t b = t.tilt_bits[g2];
            if ((!(b & TB_CLOSED) || (b & TB_HALF)) && t.pid_tilt[g2] >= 0) rs.touch(r, t.pid_tilt[g2]);
        }
        rs.touch(r, r);
    }
    }, kRowSlice);
    const int nt = t.n_tilt = (int)tiltJ.size();
    t.tilt_row.resize(nt);
    t.tilt_node.resize(nt);
    t.tilt_nb.resize((size_t)8 * nt);
    t.tilt_nblev.resize((size_t)8 * nt);
    t.tilt_pos.assign((size_t)9 * nt, 0xff);
    parallelFor(nt, [&](long j0, long j1) {
    for (int j = (int)j0; j < (int)j1; ++j) {
        const int g = tiltJ[j], l = t.level[g], i = g - t.off[l];
        const Level& Lv = o.lv[l];
        const int x = Lv.x[i], y = Lv.y[i], z = Lv.z[i];
        const int r = num.tilt[l][i];
        t.tilt_row[j] = r;
        t.tilt_node[j] = g;
        for (int k = 0; k < 8; ++k) {
            lcoor_t c = tiltNeighbor(o, l, x, y, z, k);
            int32_t code;
            uint8_t lev = (uint8_t)l;
            if (c == -1) {
                bool valid = octValid(Lv, x + kTiltNbOct[k][0], y + kTiltNbOct[k][1], z + kTiltNbOct[k][2]);
                code = valid ? TNB_NONE : TNB_DIRICHLET;
            } else {
                const int cl = lc_level(c);
                const Level& Cv = o.lv[cl];
                if (octValid(Cv, lc_x(c), lc_y(c), lc_z(c))) {
                    int ci = Cv.get(lc_x(c), lc_y(c), lc_z(c));
                    if (ci < 0) throw std::runtime_error("tilt neighbour octagon missing");
                    code = num.oct[cl][ci] >= 0 ? num.oct[cl][ci] : TNB_ZERO;
                    lev = (uint8_t)cl;
                } else {
                    code = TNB_DIRICHLET;
                }
            }
            t.tilt_nb[(size_t)k * nt + j] = code;
            t.tilt_nblev[(size_t)k * nt + j] = lev;
            if (code >= 0) rs.touch(r, code);
        }
        rs.touch(r, r);
    }
    }, kRowSlice);

    tm.mark("octagon and tilt rows");
    // ---- half-tilt rows (project 554-649), serial order ----
    const int nh = t.nh = (int)o.half.size();
    t.half_node.resize(nh);
    t.half_c.resize(nh);
    t.half_t.resize(nh);
    t.half_closed.resize(nh);
    t.half_nb.resize((size_t)4 * nh);
    {
        std::set<lcoor_t> seen;
        for (lcoor_t h : o.half)
            if (!seen.insert(h).second) t.half_duplicates = true;
    }
    for (int q = 0; q < nh; ++q) {
        const lcoor_t h = o.half[q];
        const int l = lc_level(h), x = lc_x(h), y = lc_y(h), z = lc_z(h);
        const int i = o.node(l, x, y, z);
        const int g = t.off[l] + i;
        const int hd = o.tilt[l][i].half_dir;
        if (hd < 0 || hd > 5) throw std::runtime_error("half-tilt without a direction");
        t.half_node[q] = g;
        const int c = num.tilt[l][i];
        t.half_c[q] = c;
        int ci[3] = {x >> 1, y >> 1, z >> 1};
        ci[hd % 3] -= hd < 3 ? 1 : 0;
        int32_t tc = -20;
        if (l + 1 < L && octValid(o.lv[l + 1], ci[0], ci[1], ci[2])) {
            int cn = o.node(l + 1, ci[0], ci[1], ci[2]);
            if (cn < 0) throw std::runtime_error("half-tilt coarse octagon missing");
            tc = num.oct[l + 1][cn];
        }
        t.half_t[q] = tc;
        const int axisBit[6] = {1 << 1, 1 << 0, 1 << 5, 1 << 4, 1 << 3, 1 << 2};
        const int axis = (hd % 3) * 2 + (hd >= 3 ? 1 : 0);
        t.half_closed[q] = (o.closed_flags & axisBit[axis]) ? 1 : 0;
        for (int m = 0; m < 4; ++m) {
            const int k = kHalfSlot[hd][m];
            int nx = x + kTiltNbOct[k][0], ny = y + kTiltNbOct[k][1], nz = z + kTiltNbOct[k][2];
            int32_t code = -1;
            if (octValid(o.lv[l], nx, ny, nz)) {
                int ni = o.node(l, nx, ny, nz);
                if (ni < 0) throw std::runtime_error("half-tilt neighbour node missing");
                code = num.oct[l][ni];
            } else if (c >= 0) {
                throw std::runtime_error("half-tilt neighbour outside the domain");
            }
            t.half_nb[(size_t)4 * q + m] = code;
        }
        if (tc >= 0 && c >= 0) {
            rs.touchShared(tc, tc);
            rs.touchShared(tc, c);
            rs.touchShared(c, tc);
        }
        if (c >= 0) {
            for (int m = 0; m < 4; ++m)
                if (t.half_nb[(size_t)4 * q + m] >= 0) rs.touchShared(c, t.half_nb[(size_t)4 * q + m]);
            rs.touchShared(c, c);
        }
    }

    tm.mark("half-tilt rows");
    // ---- CSR: each row's own list merged with its shared couplings ----
    std::sort(rs.extra.begin(), rs.extra.end());
    rs.extra.erase(std::unique(rs.extra.begin(), rs.extra.end()), rs.extra.end());
    tm.mark("shared entries sorted");
    const auto& ex = rs.extra;
    // merge of row r's list and its shared entries [e0, e1); emit(col) per entry
    auto mergeRow = [&](int r, size_t e0, size_t e1, auto&& emit) {
        const int32_t* v = rs.c + (size_t)r * kRowCap;
        const int m = rs.k[r];
        int a = 0;
        size_t b = e0;
        while (a < m || b < e1) {
            if (b == e1 || (a < m && v[a] < ex[b].second)) emit(v[a++]);
            else if (a == m || ex[b].second < v[a]) emit(ex[b++].second);
            else {
                emit(v[a++]);
                ++b;
            }
        }
    };
    // rows in slices; each slice finds its first shared entry once
    auto sliceRows = [&](auto&& body) {
        parallelFor(t.n, [&](long r0, long r1) {
            size_t e = (size_t)(std::lower_bound(ex.begin(), ex.end(), std::make_pair((int)r0, INT32_MIN)) - ex.begin());
            for (int r = (int)r0; r < (int)r1; ++r) {
                const size_t e0 = e;
                while (e < ex.size() && ex[e].first == r) ++e;
                body(r, e0, e);
            }
        }, kRowSlice);
    };
    t.rowstart.assign(t.n + 1, 0);
    sliceRows([&](int r, size_t e0, size_t e1) {
        int cnt = 0;
        mergeRow(r, e0, e1, [&](int) { ++cnt; });
        t.rowstart[r + 1] = cnt;
    });
    for (int r = 0; r < t.n; ++r) t.rowstart[r + 1] += t.rowstart[r];
    t.col.resize(t.rowstart[t.n]);
    sliceRows([&](int r, size_t e0, size_t e1) {
        int32_t* o = t.col.data() + t.rowstart[r];
        mergeRow(r, e0, e1, [&](int c) { *o++ = c; });
    });
    tm.mark("CSR");
    auto posOf = [&](int r, int c) -> uint8_t {
        const int32_t* b = t.col.data() + t.rowstart[r];
        const int32_t* e = t.col.data() + t.rowstart[r + 1];
        const int32_t* it = std::lower_bound(b, e, c);
        if (it == e || *it != c) throw std::runtime_error("pattern lookup failed");
        size_t p = (size_t)(it - b);
        if (p > 254) throw std::runtime_error("row too long");
        return (uint8_t)p;
    };
    parallelFor(no, [&](long j0, long j1) {
    for (int j = (int)j0; j < (int)j1; ++j) {
        const int r = t.oct_row[j];
        for (int f = 0; f < 6; ++f) {
            int code = t.oct_face[(size_t)f * no + j];
            if (code >= 0) t.oct_pos[(size_t)f * no + j] = posOf(r, code);
        }
        for (int k = 0; k < 8; ++k) {
            const int g2 = t.oct_corner[(size_t)k * no + j];
            const uint8_t b = t.tilt_bits[g2];
            if ((!(b & TB_CLOSED) || (b & TB_HALF)) && t.pid_tilt[g2] >= 0)
                t.oct_pos[(size_t)(6 + k) * no + j] = posOf(r, t.pid_tilt[g2]);
        }
        t.oct_pos[(size_t)14 * no + j] = posOf(r, r);
    }
    }, kRowSlice);
    parallelFor(nt, [&](long j0, long j1) {
    for (int j = (int)j0; j < (int)j1; ++j) {
        const int r = t.tilt_row[j];
        for (int k = 0; k < 8; ++k) {
            int code = t.tilt_nb[(size_t)k * nt + j];
            if (code >= 0) t.tilt_pos[(size_t)k * nt + j] = posOf(r, code);
        }
        t.tilt_pos[(size_t)8 * nt + j] = posOf(r, r);
    }
    }, kRowSlice);
    tm.mark("CSR and positions");
    // coarse-target lists and half-row occurrence lists, both in q order
    {
        // (row, q) pairs sorted by row then q: rows ascending, q order within a row
        std::vector<std::pair<int, int>> byT, byC;
        for (int q = 0; q < nh; ++q) {
            if (t.half_t[q] >= 0 && t.half_c[q] >= 0) byT.push_back({t.half_t[q], q});
            if (t.half_c[q] >= 0) byC.push_back({t.half_c[q], q});
        }
        std::sort(byT.begin(), byT.end());
        std::sort(byC.begin(), byC.end());
        t.hc_start.push_back(0);
        for (size_t a = 0; a < byT.size();) {
            const int r = byT[a].first;
            t.hc_row.push_back(r);
            for (; a < byT.size() && byT[a].first == r; ++a) {
                const int q = byT[a].second;
                t.hc_q.push_back(q);
                t.hc_pos_diag.push_back(posOf(r, r));
                t.hc_pos_off.push_back(posOf(r, t.half_c[q]));
            }
            t.hc_start.push_back((int)t.hc_q.size());
        }
        t.hr_start.push_back(0);
        for (size_t a = 0; a < byC.size();) {
            const int r = byC[a].first;
            t.hr_row.push_back(r);
            for (; a < byC.size() && byC[a].first == r; ++a) {
                const int q = byC[a].second;
                t.hr_q.push_back(q);
                t.hr_pos.push_back(t.half_t[q] >= 0 ? posOf(r, t.half_t[q]) : 0xff);
                for (int m = 0; m < 4; ++m) {
                    int code = t.half_nb[(size_t)4 * q + m];
                    t.hr_pos.push_back(code >= 0 ? posOf(r, code) : 0xff);
                }
                t.hr_pos.push_back(posOf(r, r));
            }
            t.hr_start.push_back((int)t.hr_q.size());
        }
        t.n_half_rows = (int)t.hr_row.size();
    }

    tm.mark("half lists");
    // ---- velocity correction tables (updateVelocityByPressure 681-847) ----
    // MAC: every leaf of every level, in level then leaf order; neighbour
    // codes stored axis-major
    {
        std::vector<int> base(L + 1, 0);
        for (int l = 0; l < L; ++l) base[l + 1] = base[l] + o.lv[l].nleaf;
        const int nvm = base[L];
        t.vm_node.resize(nvm);
        t.vm_c.resize(nvm);
        t.vm_nb.resize((size_t)3 * nvm);
        for (int l = 0; l < L; ++l) {
            const Level& Lv = o.lv[l];
            const int res = Lv.res;
            parallelFor(Lv.nleaf, [&](long i0, long i1) {
                for (int i = (int)i0; i < (int)i1; ++i) {
                    const int j = base[l] + i;
                    const int g = t.off[l] + i, x = Lv.x[i], y = Lv.y[i], z = Lv.z[i];
                    t.vm_node[j] = g;
                    t.vm_c[j] = octValid(Lv, x, y, z) ? num.oct[l][i] : -1;
                    for (int a = 0; a < 3; ++a) {
                        // MAC validity: axis a may reach res-1, the other two res-2
                        int lim[3] = {res - 1, res - 1, res - 1};
                        lim[a] = res;
                        const int v[3] = {x, y, z};
                        bool macValid = v[0] < lim[0] && v[1] < lim[1] && v[2] < lim[2];
                        int32_t code = VNB_SKIP;
                        if (macValid) {
                            int nb[3] = {x, y, z};
                            nb[a] -= 1;
                            code = VNB_P0;
                            if (octValid(Lv, nb[0], nb[1], nb[2])) {
                                int ni = Lv.get(nb[0], nb[1], nb[2]);
                                if (ni < 0) throw std::runtime_error("MAC neighbour node missing");
                                if (num.oct[l][ni] >= 0) code = num.oct[l][ni];
                            }
                        }
                        t.vm_nb[(size_t)a * nvm + j] = code;
                    }
                }
            }, 4 * kRowSlice);
        }
    }
    // open tilts, in level then leaf order; codes and levels slot-major
    {
        for (int l = 0; l < L; ++l)
            for (int i = 0; i < o.lv[l].nleaf; ++i)
                if (o.tilt[l][i].bits & TB_LAYOUT_OPEN) t.vt_node.push_back(t.off[l] + i);
        const int nvt = (int)t.vt_node.size();
        t.vt_c.resize(nvt);
        t.vt_nb.resize((size_t)8 * nvt);
        t.vt_lev.resize((size_t)8 * nvt);
        parallelFor(nvt, [&](long j0, long j1) {
            for (int j = (int)j0; j < (int)j1; ++j) {
                const int g = t.vt_node[j], l = t.level[g], i = g - t.off[l];
                const Level& Lv = o.lv[l];
                const int x = Lv.x[i], y = Lv.y[i], z = Lv.z[i];
                t.vt_c[j] = num.tilt[l][i];
                for (int k = 0; k < 8; ++k) {
                    lcoor_t c = tiltNeighbor(o, l, x, y, z, k);
                    int32_t code = VNB_ZERO;
                    uint8_t lev = (uint8_t)l;
                    if (c != -1) {
                        const int cl = lc_level(c);
                        lev = (uint8_t)cl;
                        code = VNB_P0;
                        if (octValid(o.lv[cl], lc_x(c), lc_y(c), lc_z(c))) {
                            int ci = o.lv[cl].get(lc_x(c), lc_y(c), lc_z(c));
                            if (ci >= 0 && num.oct[cl][ci] >= 0) code = num.oct[cl][ci];
                        }
                    }
                    t.vt_nb[(size_t)k * nvt + j] = code;
                    t.vt_lev[(size_t)k * nvt + j] = lev;
                }
            }
        }, 4 * kRowSlice);
    }
    tm.mark("velocity tables");
}

}  // namespace astg
