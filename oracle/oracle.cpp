// ORACLE - CPU restatement of the reference's projection (TEST INFRASTRUCTURE).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library, as the checker; the product never links or calls it.
//
// Restated from the reference (file:line into the AST-Grid sources):
//   orc_project      OctreeOmniGridSolver3::project (src/octree_omnigrid_solver3.cpp:350-679):
//                    octagon rows 367-477, tilt rows 479-552, half-tilt rows 554-649,
//                    b *= rho/dt 651, solve 672, updateVelocityByPressure 681-847.
//                    Node lookups go through per-level hash maps (as OctreeLayer3,
//                    include/octree_layer3.h:40-44), the matrix is a sorted row list with
//                    set/add semantics (include/pcgsolver/sparse_matrix.h:69-103).
//   orc_mic_factor   factor_modified_incomplete_cholesky0 (include/pcgsolver/pcg_solver.h:89-190),
//                    right-looking by columns exactly as the reference orders it.
//   orc_pcg          PCGSolver::solve, Eigen overload (pcg_solver.h:366-426), with
//                    solve_lower / solve_lower_transpose_in_place (211-253).
// Inputs are the reference-form topology (node lists, tilt state, numbering,
// weights, half-tilt list) produced by the product's topology builder; that
// builder is pinned bit-exactly against the reference's own dumps by the CPU
// tests, and this restatement is pinned against the reference's assembled
// systems, factors and solutions (tests/test_oracle.py).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

const double SQ3 = 1.7320508075688772;
const double EPS = 2.220446049250313e-16;

// constants.h:85-118 (restated from their geometric meaning)
const int NB6[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
const int FACE6[6][3] = {{1, 0, 0}, {0, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 1}, {0, 0, 0}};
const int CORNER8[8][3] = {{1, 1, 1}, {0, 1, 1}, {0, 0, 1}, {1, 0, 1}, {1, 1, 0}, {0, 1, 0}, {0, 0, 0}, {1, 0, 0}};
const int TNB8[8][3] = {{0, 0, 0}, {-1, 0, 0}, {-1, -1, 0}, {0, -1, 0},
                        {0, 0, -1}, {-1, 0, -1}, {-1, -1, -1}, {0, -1, -1}};
const int HSLOT[6][5] = {{0, 3, 4, 7, 1}, {0, 1, 4, 5, 2}, {0, 1, 2, 3, 5}, {1, 2, 5, 6, 0}, {2, 3, 6, 7, 4}, {4, 5, 6, 7, 3}};

int64_t key(int x, int y, int z) { return ((int64_t)x << 40) | ((int64_t)y << 20) | (int64_t)z; }

struct Lv {
    int nnode, nleaf, res;
    double h;
    const int *xyz, *flags, *tbits, *thd, *poct, *ptilt;
    const double *te, *wmac, *wtilt;
    double *umac, *utilt;
    std::unordered_map<int64_t, int> map;
};

struct Rows {
    std::vector<std::vector<int>> idx;
    std::vector<std::vector<double>> val;
    void set(int i, int j, double v, bool add) {
        auto& I = idx[i];
        auto& V = val[i];
        size_t k = 0;
        for (; k < I.size(); ++k) {
            if (I[k] == j) {
                V[k] = add ? V[k] + v : v;
                return;
            }
            if (I[k] > j) break;
        }
        I.insert(I.begin() + k, j);
        V.insert(V.begin() + k, v);
    }
};

struct Model {
    int L, closed;
    std::vector<Lv> lv;
    int find(int l, int x, int y, int z) const {
        if (l < 0 || l >= L) return -1;
        const Lv& a = lv[l];
        if (x < 0 || y < 0 || z < 0 || x >= a.res || y >= a.res || z >= a.res) return -1;
        auto it = a.map.find(key(x, y, z));
        return it == a.map.end() ? -1 : it->second;
    }
    int need(int l, int x, int y, int z) const {
        int i = find(l, x, y, z);
        if (i < 0) throw std::runtime_error("oracle: node lookup failed");
        return i;
    }
    bool ghost(int l, int i) const { return (lv[l].flags[i] & 2) != 0; }
    bool octValid(int l, int x, int y, int z) const {
        int r = lv[l].res;
        return x >= 0 && y >= 0 && z >= 0 && x <= r - 2 && y <= r - 2 && z <= r - 2;
    }
    // getUntil (src/octree_grid_layout3.cpp:51-69); returns level or -1, coordinates in/out
    int until(int l, int& x, int& y, int& z, bool fine) const {
        int r = lv[l].res;
        if (x < 0 || y < 0 || z < 0 || x >= r || y >= r || z >= r) return -1;
        for (;;) {
            int i = find(l, x, y, z);
            if (i >= 0 && !ghost(l, i)) return l;
            if (fine) { x <<= 1; y <<= 1; z <<= 1; --l; }
            else { x >>= 1; y >>= 1; z >>= 1; ++l; }
            if (l < 0 || l >= L) return -1;
        }
    }
    // tiltNeighborOctagon3 (src/octree_util.cpp:172-194)
    int tiltNb(int l, int x, int y, int z, int k, int& ox, int& oy, int& oz) const {
        int nx = x + TNB8[k][0], ny = y + TNB8[k][1], nz = z + TNB8[k][2];
        int r = lv[l].res;
        if (nx < 0 || ny < 0 || nz < 0 || nx >= r || ny >= r || nz >= r) return -1;
        int ax = nx, ay = ny, az = nz;
        int cl = until(l, ax, ay, az, true);
        if (cl < 0) {
            ax = nx; ay = ny; az = nz;
            cl = until(l, ax, ay, az, false);
        }
        if (cl < 0) return -1;
        if (cl < l) {
            int m4 = k % 4;
            az += k >= 4 ? 1 : 0;
            ay += m4 >= 2 ? 1 : 0;
            ax += (m4 == 1 || m4 == 2) ? 1 : 0;
        }
        ox = ax; oy = ay; oz = az;
        return cl;
    }
    // isClosedDomain(i) (include/omni_boundary_solver3.h:164-182)
    bool closedFace(int i) const {
        static const int bit[6] = {2, 1, 32, 16, 8, 4};
        return (closed & bit[i]) != 0;
    }
};

// ---------------------------------------------------------------- assembly
void assemble(const Model& M, const int* half, int nh, int n, double dt, double rho, Rows& A, std::vector<double>& b) {
    A.idx.assign(n, {});
    A.val.assign(n, {});
    b.assign(n, 0.0);
    for (int l = 0; l < M.L; ++l) {
        const Lv& a = M.lv[l];
        const double h = a.h;
        // octagon rows (367-477)
        for (int i = 0; i < a.nleaf; ++i) {
            const int x = a.xyz[3 * i], y = a.xyz[3 * i + 1], z = a.xyz[3 * i + 2];
            const int c = a.poct[i];
            if (c < 0 || !M.octValid(l, x, y, z)) continue;
            double center = 0, tb = 0;
            for (int f = 0; f < 6; ++f) {
                const int axis = f / 2;
                const int nx = x + NB6[f][0], ny = y + NB6[f][1], nz = z + NB6[f][2];
                const bool nv = M.octValid(l, nx, ny, nz);
                if (nv && M.ghost(l, M.need(l, nx, ny, nz))) continue;
                double area = h * h;
                for (int m = 0; m <= 1; ++m)
                    for (int q = 0; q <= 1; ++q) {
                        int t[3] = {x, y, z};
                        t[axis] += (f & 1) ? 0 : 1;
                        t[(axis + 1) % 3] += m;
                        t[(axis + 2) % 3] += q;
                        const int ti = M.need(l, t[0], t[1], t[2]);
                        area -= a.te[ti] * a.te[ti] * 0.5;
                    }
                const int fi = M.need(l, x + FACE6[f][0], y + FACE6[f][1], z + FACE6[f][2]);
                double tC = 0, tK = 0;
                if (nv) {
                    const double term = a.wmac[3 * fi + axis];
                    const int t = a.poct[M.need(l, nx, ny, nz)];
                    if (t >= 0) {
                        tC = term * area / h;
                        tK = area * term;
                        A.set(c, t, -tC, false);
                    }
                } else if (!M.closedFace(f)) {
                    tK = area;
                    tC = tK / h;
                }
                center += tC;
                tb += ((f & 1) ? 1 : -1) * tK * a.umac[3 * fi + axis];
            }
            for (int k = 0; k < 8; ++k) {
                const int ci = M.need(l, x + CORNER8[k][0], y + CORNER8[k][1], z + CORNER8[k][2]);
                const int bits = a.tbits[ci];
                if ((bits & 2) && !(bits & 32)) continue;
                const int re = ((k & 1) ? 8 : 6) - k;
                const double area = a.te[ci] * a.te[ci] * SQ3 * 0.5;
                const double term = a.wtilt[8 * ci + re];
                double tC = 0, tK = 0;
                const int t = a.ptilt[ci];
                if (t >= 0) {
                    tC = term * area / (h * SQ3 * 0.5);
                    tK = area * term;
                    A.set(c, t, -tC, true);
                }
                center += tC;
                tb += (k < 4 ? -1 : 1) * tK * a.utilt[8 * ci + re];
            }
            if (std::fabs(center) < EPS) {
                center = 1;
                tb = 0;
            }
            A.set(c, c, center, false);
            b[c] = tb;
        }
        // tilt rows (479-552)
        for (int i = 0; i < a.nleaf; ++i) {
            if (!(a.tbits[i] & 64)) continue;
            const int c = a.ptilt[i];
            if (c < 0) continue;
            const int x = a.xyz[3 * i], y = a.xyz[3 * i + 1], z = a.xyz[3 * i + 2];
            const double area = a.te[i] * a.te[i] * SQ3 * 0.5;
            double center = 0, tb = 0;
            for (int k = 0; k < 8; ++k) {
                int ox, oy, oz;
                const int cl = M.tiltNb(l, x, y, z, k, ox, oy, oz);
                double tC = 0, tK = 0;
                if (cl < 0) {
                    if (M.octValid(l, x + TNB8[k][0], y + TNB8[k][1], z + TNB8[k][2])) continue;
                    tC = area / (h * SQ3 / 3.0);
                    tK = area;
                } else if (M.octValid(cl, ox, oy, oz)) {
                    const double hn = M.lv[cl].h;
                    const int t = M.lv[cl].poct[M.need(cl, ox, oy, oz)];
                    if (t >= 0) {
                        tC = a.wtilt[8 * i + k] * area / (hn * SQ3 * 0.5);
                        tK = area * a.wtilt[8 * i + k];
                        A.set(c, t, -tC, false);
                    }
                } else {
                    tC = area / (h * SQ3 / 3.0);
                    tK = area;
                }
                center += tC;
                tb += (k < 4 ? -1 : 1) * tK * a.utilt[8 * i + k];
            }
            if (std::fabs(center) < EPS) center = 1;
            A.set(c, c, center, false);
            b[c] = tb;
        }
    }
    // half-tilt rows (554-649), serial in list order
    for (int q = 0; q < nh; ++q) {
        const int l = half[4 * q], x = half[4 * q + 1], y = half[4 * q + 2], z = half[4 * q + 3];
        const Lv& a = M.lv[l];
        const int i = M.need(l, x, y, z);
        const int c = a.ptilt[i], hd = a.thd[i];
        double area = a.te[i] * a.te[i] * 2;
        const double h = a.h;
        double center = 0, tb = 0;
        int co[3] = {x >> 1, y >> 1, z >> 1};
        co[hd % 3] -= hd < 3 ? 1 : 0;
        const int os = HSLOT[hd][4];
        const double wf = a.wtilt[8 * i + os], uf = a.utilt[8 * i + os];
        if (l + 1 < M.L && M.octValid(l + 1, co[0], co[1], co[2])) {
            const int t = M.lv[l + 1].poct[M.need(l + 1, co[0], co[1], co[2])];
            if (t >= 0 && c >= 0) {
                A.set(t, t, wf * area / h, true);
                center += wf * area / h;
                A.set(t, c, -wf * area / h, true);
                A.set(c, t, -wf * area / h, true);
                b[t] += (hd < 3 ? -1 : 1) * uf * wf * area;
                tb += (hd < 3 ? 1 : -1) * uf * wf * area;
            }
        } else if (c >= 0) {
            const int ax = (hd % 3) * 2 + (hd >= 3);
            center += M.closedFace(ax) ? 0.0 : 1000.0;
            tb += (hd < 3 ? 1 : -1) * uf * area;
        }
        if (c < 0) continue;
        area = a.te[i] * a.te[i] * SQ3 * 0.5;
        for (int m = 0; m < 4; ++m) {
            const int k = HSLOT[hd][m];
            const int nx = x + TNB8[k][0], ny = y + TNB8[k][1], nz = z + TNB8[k][2];
            if (!M.octValid(l, nx, ny, nz)) throw std::runtime_error("oracle: outer half-tilt neighbour");
            double tC = 0, tK = 0;
            const int t = a.poct[M.need(l, nx, ny, nz)];
            if (t >= 0) {
                tC = area * a.wtilt[8 * i + k] / (h * SQ3 * 0.5);
                tK = area * a.wtilt[8 * i + k];
                A.set(c, t, -tC, false);
            }
            center += tC;
            tb += (k < 4 ? -1 : 1) * tK * a.utilt[8 * i + k];
        }
        A.set(c, c, center, true);
        b[c] = tb;
    }
    const double s = rho / dt;
    for (auto& v : b) v *= s;
}

// ---------------------------------------------------------------- MIC(0)
struct Factor {
    std::vector<double> invdiag, value;
    std::vector<unsigned> rowindex, colstart;
};

void micFactor(const Rows& A, double tau, double ratio, Factor& F) {
    const unsigned n = (unsigned)A.idx.size();
    F.invdiag.assign(n, 0.0);
    F.colstart.assign(n + 1, 0);
    F.value.clear();
    F.rowindex.clear();
    std::vector<double> adiag(n, 0.0);
    for (unsigned i = 0; i < n; ++i) {
        F.colstart[i] = (unsigned)F.rowindex.size();
        for (size_t j = 0; j < A.idx[i].size(); ++j) {
            const unsigned c = (unsigned)A.idx[i][j];
            if (c > i) {
                F.rowindex.push_back(c);
                F.value.push_back(A.val[i][j]);
            } else if (c == i) {
                F.invdiag[i] = adiag[i] = A.val[i][j];
            }
        }
    }
    F.colstart[n] = (unsigned)F.rowindex.size();
    for (unsigned k = 0; k < n; ++k) {
        if (adiag[k] == 0) continue;
        F.invdiag[k] = F.invdiag[k] < ratio * adiag[k] ? 1 / std::sqrt(adiag[k]) : 1 / std::sqrt(F.invdiag[k]);
        const unsigned c0 = F.colstart[k], c1 = F.colstart[k + 1];
        for (unsigned p = c0; p < c1; ++p) F.value[p] *= F.invdiag[k];
        for (unsigned p = c0; p < c1; ++p) {
            const unsigned j = F.rowindex[p];
            const double mult = F.value[p];
            double missing = 0;
            unsigned a = c0;
            size_t bb = 0;
            const auto& rj = A.idx[j];
            while (a < c1 && F.rowindex[a] < j) {
                while (bb < rj.size()) {
                    if ((unsigned)rj[bb] < F.rowindex[a]) ++bb;
                    else if ((unsigned)rj[bb] == F.rowindex[a]) break;
                    else {
                        missing += F.value[a];
                        break;
                    }
                }
                ++a;
            }
            if (a < c1 && F.rowindex[a] == j) F.invdiag[j] -= mult * F.value[a];
            ++a;
            unsigned b2 = F.colstart[j];
            while (a < c1 && b2 < F.colstart[j + 1]) {
                if (F.rowindex[b2] < F.rowindex[a]) ++b2;
                else if (F.rowindex[b2] == F.rowindex[a]) {
                    F.value[b2] -= mult * F.value[a];
                    ++a;
                    ++b2;
                } else {
                    missing += F.value[a];
                    ++a;
                }
            }
            while (a < c1) missing += F.value[a++];
            F.invdiag[j] -= tau * mult * missing;
        }
    }
}

void micApply(const Factor& F, const std::vector<double>& r, std::vector<double>& z) {
    const unsigned n = (unsigned)F.invdiag.size();
    z = r;
    for (unsigned i = 0; i < n; ++i) {
        z[i] *= F.invdiag[i];
        for (unsigned p = F.colstart[i]; p < F.colstart[i + 1]; ++p) z[F.rowindex[p]] -= F.value[p] * z[i];
    }
    for (unsigned i = n; i-- > 0;) {
        for (unsigned p = F.colstart[i]; p < F.colstart[i + 1]; ++p) z[i] -= F.value[p] * z[F.rowindex[p]];
        z[i] *= F.invdiag[i];
    }
}

// ---------------------------------------------------------------- PCG
bool pcg(const Rows& A, const std::vector<double>& b, std::vector<double>& x, double tolf, int maxit, double tau,
         double ratio, double& res, int& iters) {
    const size_t n = b.size();
    x.assign(n, 0.0);
    std::vector<double> r = b, z, s, q(n);
    res = 0;
    for (double v : r) res = std::max(res, std::fabs(v));
    if (res == 0) {
        iters = 0;
        return true;
    }
    const double tol = tolf * res;
    Factor F;
    micFactor(A, tau, ratio, F);
    micApply(F, r, z);
    double rho = 0;
    for (size_t i = 0; i < n; ++i) rho += z[i] * r[i];
    if (rho == 0 || rho != rho) {
        iters = 0;
        return false;
    }
    s = z;
    int it;
    for (it = 0; it < maxit; ++it) {
        for (size_t i = 0; i < n; ++i) {
            double t = 0;
            for (size_t k = 0; k < A.idx[i].size(); ++k) t += A.val[i][k] * s[A.idx[i][k]];
            q[i] = t;
        }
        double sq = 0;
        for (size_t i = 0; i < n; ++i) sq += s[i] * q[i];
        const double alpha = rho / sq;
        res = 0;
        for (size_t i = 0; i < n; ++i) {
            x[i] += alpha * s[i];
            r[i] -= alpha * q[i];
            res = std::max(res, std::fabs(r[i]));
        }
        if (res <= tol) {
            iters = it + 1;
            return true;
        }
        micApply(F, r, z);
        double rn = 0;
        for (size_t i = 0; i < n; ++i) rn += z[i] * r[i];
        const double beta = rn / rho;
        for (size_t i = 0; i < n; ++i) s[i] = z[i] + beta * s[i];
        rho = rn;
    }
    iters = it;
    return false;
}

// ---------------------------------------------------------------- velocity
void correct(Model& M, const int* half, int nh, const std::vector<double>& p, double dt, double rho) {
    auto P = [&](int i) { return i >= 0 ? p[i] : 0.0; };
    for (int l = 0; l < M.L; ++l) {
        Lv& a = M.lv[l];
        const double h = a.h;
        // MAC faces (690-737)
        for (int i = 0; i < a.nleaf; ++i) {
            const int x = a.xyz[3 * i], y = a.xyz[3 * i + 1], z = a.xyz[3 * i + 2];
            const int c = M.octValid(l, x, y, z) ? a.poct[i] : -1;
            const double p0 = P(c);
            for (int ax = 0; ax < 3; ++ax) {
                int v[3] = {x, y, z};
                bool ok = true;
                for (int d = 0; d < 3; ++d) ok = ok && v[d] < a.res - (d == ax ? 0 : 1);
                if (!ok) continue;
                if (a.wmac[3 * i + ax] <= 0) {
                    a.umac[3 * i + ax] = 0;
                    continue;
                }
                v[ax] -= 1;
                double p1 = 0;
                if (M.octValid(l, v[0], v[1], v[2])) p1 = P(a.poct[M.need(l, v[0], v[1], v[2])]);
                a.umac[3 * i + ax] = a.umac[3 * i + ax] - dt / rho * (p0 - p1) / h;
            }
        }
        // open tilts (739-786)
        for (int i = 0; i < a.nleaf; ++i) {
            if (!(a.tbits[i] & 64)) continue;
            const int x = a.xyz[3 * i], y = a.xyz[3 * i + 1], z = a.xyz[3 * i + 2];
            const double p0 = P(a.ptilt[i]);
            for (int k = 0; k < 8; ++k) {
                int ox, oy, oz;
                const int cl = M.tiltNb(l, x, y, z, k, ox, oy, oz);
                double& u = a.utilt[8 * i + k];
                if (a.wtilt[8 * i + k] > 0 && cl >= 0) {
                    double p1 = 0;
                    if (M.octValid(cl, ox, oy, oz)) p1 = P(M.lv[cl].poct[M.need(cl, ox, oy, oz)]);
                    u = u - dt / rho * (k < 4 ? 1 : -1) / (M.lv[cl].h * SQ3 * 0.5) * (p1 - p0);
                } else {
                    u = 0;
                }
            }
        }
    }
    // half-tilts (788-846)
    for (int q = 0; q < nh; ++q) {
        const int l = half[4 * q], x = half[4 * q + 1], y = half[4 * q + 2], z = half[4 * q + 3];
        Lv& a = M.lv[l];
        const int i = M.need(l, x, y, z);
        const int hd = a.thd[i];
        const double p0 = P(a.ptilt[i]);
        int co[3] = {x >> 1, y >> 1, z >> 1};
        co[hd % 3] -= hd < 3 ? 1 : 0;
        const int fs = HSLOT[hd][4];
        double* u = &a.utilt[8 * i];
        if (a.wtilt[8 * i + fs] > 0) {
            const double p1 = P(M.lv[l + 1].poct[M.need(l + 1, co[0], co[1], co[2])]);
            u[fs] -= dt / rho * (hd < 3 ? 1 : -1) / a.h * (p0 - p1);
        } else {
            u[fs] = 0;
        }
        for (int k = 0; k < 8; ++k) {
            bool listed = false;
            for (int m = 0; m < 4; ++m) listed = listed || HSLOT[hd][m] == k;
            if (listed && a.wtilt[8 * i + k] > 0) {
                const int nx = x + TNB8[k][0], ny = y + TNB8[k][1], nz = z + TNB8[k][2];
                double p1 = 0;
                if (M.octValid(l, nx, ny, nz)) p1 = P(a.poct[M.need(l, nx, ny, nz)]);
                u[k] = u[k] - dt / rho * (k < 4 ? 1 : -1) / (a.h * SQ3 * 0.5) * (p1 - p0);
            } else if (k != fs) {
                u[k] = 0;
            }
        }
    }
}

Rows fromCsr(int n, const int* rs, const int* col, const double* val) {
    Rows A;
    A.idx.assign(n, {});
    A.val.assign(n, {});
    for (int i = 0; i < n; ++i)
        for (int k = rs[i]; k < rs[i + 1]; ++k) {
            A.idx[i].push_back(col[k]);
            A.val[i].push_back(val[k]);
        }
    return A;
}

struct Result {
    Rows A;
    std::vector<double> b, p;
    int iters = 0;
    double res = 0;
    bool ok = false;
};

std::string g_err;

}  // namespace

extern "C" {

struct OrcLevel {
    int nnode, nleaf, res;
    double h;
    const int *xyz, *flags, *tilt_bits, *tilt_hd, *pid_oct, *pid_tilt;
    const double *tilt_e, *w_mac, *w_tilt;
    double *u_mac, *u_tilt;
};

const char* orc_last_error() { return g_err.c_str(); }

// Full projection; velocities (u_mac n*3, u_tilt n*8 per level) updated in place.
void* orc_project(int L, int closed, const OrcLevel* levels, int nhalf, const int* half, int n, double dt, double rho,
                  double tol, int maxit, double tau, double ratio) {
    try {
        Model M;
        M.L = L;
        M.closed = closed;
        M.lv.resize(L);
        for (int l = 0; l < L; ++l) {
            const OrcLevel& s = levels[l];
            Lv& a = M.lv[l];
            a.nnode = s.nnode;
            a.nleaf = s.nleaf;
            a.res = s.res;
            a.h = s.h;
            a.xyz = s.xyz;
            a.flags = s.flags;
            a.tbits = s.tilt_bits;
            a.thd = s.tilt_hd;
            a.poct = s.pid_oct;
            a.ptilt = s.pid_tilt;
            a.te = s.tilt_e;
            a.wmac = s.w_mac;
            a.wtilt = s.w_tilt;
            a.umac = s.u_mac;
            a.utilt = s.u_tilt;
            a.map.reserve((size_t)s.nnode * 2);
            for (int i = 0; i < s.nnode; ++i) a.map[key(s.xyz[3 * i], s.xyz[3 * i + 1], s.xyz[3 * i + 2])] = i;
        }
        Result* R = new Result();
        assemble(M, half, nhalf, n, dt, rho, R->A, R->b);
        R->ok = pcg(R->A, R->b, R->p, tol, maxit, tau, ratio, R->res, R->iters);
        correct(M, half, nhalf, R->p, dt, rho);
        return R;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

long orc_result_nnz(void* h) {
    long s = 0;
    for (auto& r : static_cast<Result*>(h)->A.idx) s += (long)r.size();
    return s;
}

void orc_result_get(void* h, int* rowstart, int* col, double* val, double* b, double* p, int* iters, double* res,
                    int* ok) {
    Result* R = static_cast<Result*>(h);
    const size_t n = R->b.size();
    long k = 0;
    rowstart[0] = 0;
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < R->A.idx[i].size(); ++j) {
            col[k] = R->A.idx[i][j];
            val[k] = R->A.val[i][j];
            ++k;
        }
        rowstart[i + 1] = (int)k;
        b[i] = R->b[i];
        p[i] = R->p[i];
    }
    *iters = R->iters;
    *res = R->res;
    *ok = R->ok ? 1 : 0;
}

void orc_result_free(void* h) { delete static_cast<Result*>(h); }

// MIC(0) factor of a CSR matrix (reference column form). Returns nnz_L.
long orc_mic_factor(int n, const int* rs, const int* col, const double* val, double tau, double ratio,
                    double* invdiag, unsigned* colstart, unsigned* rowindex, double* value) {
    Rows A = fromCsr(n, rs, col, val);
    Factor F;
    micFactor(A, tau, ratio, F);
    if (invdiag) {
        std::memcpy(invdiag, F.invdiag.data(), (size_t)n * 8);
        std::memcpy(colstart, F.colstart.data(), (size_t)(n + 1) * 4);
        std::memcpy(rowindex, F.rowindex.data(), F.rowindex.size() * 4);
        std::memcpy(value, F.value.data(), F.value.size() * 8);
    }
    return (long)F.value.size();
}

void orc_mic_apply(int n, const int* rs, const int* col, const double* val, double tau, double ratio, const double* r,
                   double* z) {
    Rows A = fromCsr(n, rs, col, val);
    Factor F;
    micFactor(A, tau, ratio, F);
    std::vector<double> rr(r, r + n), zz;
    micApply(F, rr, zz);
    std::memcpy(z, zz.data(), (size_t)n * 8);
}

// MIC(0)-PCG on a CSR matrix. Returns 1 if converged.
int orc_pcg(int n, const int* rs, const int* col, const double* val, const double* b, double* x, double tol, int maxit,
            double tau, double ratio, double* res, int* iters) {
    Rows A = fromCsr(n, rs, col, val);
    std::vector<double> bb(b, b + n), xx;
    bool ok = pcg(A, bb, xx, tol, maxit, tau, ratio, *res, *iters);
    std::memcpy(x, xx.data(), (size_t)n * 8);
    return ok ? 1 : 0;
}

}  // extern "C"
