// Post-projection chain, point sampling and semi-Lagrangian advection on the
// AST octree (see advect.h for the reference map). One thread per node, DOF
// or query point; every lookup goes through the dense per-level index maps.
//
// Arithmetic follows the reference expression by expression (interpolation
// orders, the 1/sqrt(3) macro written as a division, clamping rules of the
// two barycentric helpers) so results agree to rounding.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "advect.h"
#include "device_common.h"
#include "linrec.h"

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);                 \
            return ASTG_ERR_HIP;                                                     \
        }                                                                            \
    } while (0)

namespace astg {

namespace {

constexpr int BS = 256;
constexpr int kMaxLevels = 8;
constexpr double kSqrt3 = 1.7320508075688772;
constexpr double kEps = 2.220446049250313e-16;  // numeric_limits<double>::epsilon

inline int nblocks(long n) { return (int)((n + BS - 1) / BS); }

__constant__ int c_L, c_R, c_G;
__constant__ double c_h0;
__constant__ int c_res[kMaxLevels];
__constant__ int c_moff[kMaxLevels];
__constant__ double c_hl[kMaxLevels];
// least-squares dual-graph matrices [x][y][z][n0][n1][n2][n3] (3 x 7 each)
__constant__ double c_dual[128 * 21];
__constant__ double c_dualTilt[12];  // 3 x 4
__constant__ double c_slip[3];       // centre of the slip normals (omni_boundary_solver3.h:27,57,77)

// OCT_NB_TILT_DXDYDZ (tilt corner k of a cell), restated from the corner order
__constant__ int c_nbt[8][3] = {{1, 1, 1}, {0, 1, 1}, {0, 0, 1}, {1, 0, 1},
                                {1, 1, 0}, {0, 1, 0}, {0, 0, 0}, {1, 0, 0}};
// OCT_NB_DXDYDZ (6 axis neighbours)
__constant__ int c_nb6[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
// HALF_TILT_SLOT_INDEX
__constant__ int c_halfSlot[6][5] = {{0, 3, 4, 7, 1}, {0, 1, 4, 5, 2}, {0, 1, 2, 3, 5},
                                     {1, 2, 5, 6, 0}, {2, 3, 6, 7, 4}, {4, 5, 6, 7, 3}};
// OCTREE_SAMPLE_MARKER: 0x1000 = void axis; nibbles x,y,z: 1 = +, 2 = -
__constant__ int c_marker[8][8] = {
    {0x1000, 0x011, 0x001, 0x101, 0x110, 0x010, 0x1000, 0x100},
    {0x011, 0x1000, 0x201, 0x001, 0x010, 0x210, 0x200, 0x1000},
    {0x001, 0x201, 0x1000, 0x021, 0x1000, 0x200, 0x220, 0x020},
    {0x101, 0x001, 0x021, 0x1000, 0x100, 0x1000, 0x020, 0x120},
    {0x110, 0x010, 0x1000, 0x100, 0x1000, 0x012, 0x002, 0x102},
    {0x010, 0x210, 0x200, 0x1000, 0x012, 0x1000, 0x202, 0x002},
    {0x1000, 0x200, 0x220, 0x020, 0x002, 0x202, 0x1000, 0x022},
    {0x100, 0x1000, 0x020, 0x120, 0x102, 0x002, 0x022, 0x1000}};

typedef AdvTopo Topo;
typedef AdvVel Vel;

__device__ inline int nd(const Topo& T, int l, int x, int y, int z) {
    if (l < 0 || l >= c_L) return -1;
    const int r = c_res[l];
    if (x < 0 || y < 0 || z < 0 || x >= r || y >= r || z >= r) return -1;
    return T.map[c_moff[l] + ((size_t)z * r + y) * r + x];
}
__device__ inline bool ghost(const Topo& T, int g) { return (T.nf[g] & NF_GHOST) != 0; }
__device__ inline uint8_t tbits(const Topo& T, int g) { return g >= 0 ? T.tb[g] : (uint8_t)TB_CLOSED; }

// getUntil (octree_grid_layout3.cpp:51-69): first non-ghost node walking to
// finer (toFinest) or coarser levels; -1 if none.
__device__ int getUntil(const Topo& T, int l, int x, int y, int z, bool toFinest) {
    const int r = c_res[l];
    if (x < 0 || y < 0 || z < 0 || x >= r || y >= r || z >= r) return -1;
    for (;;) {
        const int g = nd(T, l, x, y, z);
        if (g >= 0 && !ghost(T, g)) return g;
        if (toFinest) {
            x <<= 1;
            y <<= 1;
            z <<= 1;
            l -= 1;
        } else {
            x >>= 1;
            y >>= 1;
            z >>= 1;
            l += 1;
        }
        if (l < 0 || l >= c_L) return -1;
    }
}

// ---- sampler arithmetic, generic over the position type P ------------------
// P = double for the forward samplers, P = D3 (value and gradient with
// respect to the sampled position) for the adjoint's position derivatives.
// Field values are always double: every sampler is linear in them, with
// weights that depend on the position only.
struct D3 {
    double v, d[3];
    __device__ D3() {}
    __device__ D3(double x) : v(x), d{0.0, 0.0, 0.0} {}
};
__device__ inline double val(double x) { return x; }
__device__ inline double val(const D3& x) { return x.v; }
__device__ inline D3 operator+(const D3& a, const D3& b) {
    D3 r;
    r.v = a.v + b.v;
    for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
__device__ inline D3 operator-(const D3& a, const D3& b) {
    D3 r;
    r.v = a.v - b.v;
    for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
__device__ inline D3 operator*(const D3& a, const D3& b) {
    D3 r;
    r.v = a.v * b.v;
    for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
__device__ inline D3 operator/(const D3& a, const D3& b) {
    D3 r;
    r.v = a.v / b.v;
    for (int k = 0; k < 3; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r;
}
__device__ inline D3 operator+(const D3& a, double b) { return a + D3(b); }
__device__ inline D3 operator+(double a, const D3& b) { return D3(a) + b; }
__device__ inline D3 operator-(const D3& a, double b) { return a - D3(b); }
__device__ inline D3 operator-(double a, const D3& b) { return D3(a) - b; }
__device__ inline D3 operator*(const D3& a, double b) {
    D3 r;
    r.v = a.v * b;
    for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] * b;
    return r;
}
__device__ inline D3 operator*(double a, const D3& b) { return b * a; }
__device__ inline D3 operator/(const D3& a, double b) {
    D3 r;
    r.v = a.v / b;
    for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] / b;
    return r;
}
__device__ inline D3& operator+=(D3& a, const D3& b) { return a = a + b; }
__device__ inline D3& operator-=(D3& a, double b) { return a = a - b; }
__device__ inline D3& operator*=(D3& a, double b) { return a = a * b; }
// min / max keep the operand the value picks (ties: the first)
__device__ inline double pmin(double a, double b) { return fmin(a, b); }
__device__ inline double pmax(double a, double b) { return fmax(a, b); }
__device__ inline D3 pmin(const D3& a, const D3& b) { return b.v < a.v ? b : a; }
__device__ inline D3 pmax(const D3& a, const D3& b) { return b.v > a.v ? b : a; }

// Loads of field values. PlainLd reads memory; the adjoint's probe policy
// returns a one-hot weight per load instead (transposed scatters).
struct PlainLd {
    __device__ double operator()(const double* base, size_t i) const { return base[i]; }
};

// getBarycentric (util.h:259-282) with iLow = 0
template <class P>
__device__ inline void bary(P x, int iHigh, int& i, P& f) {
    const double s = floor(val(x));
    i = (int)s;
    if (iHigh == 0) {
        i = 0;
        f = P(0.0);
    } else if (i < 0) {
        i = 0;
        f = P(0.0);
    } else if (i > iHigh - 1) {
        i = iHigh - 1;
        f = P(1.0);
    } else {
        f = x - s;
    }
}
// getBaryCentric, vector form (util.h:244-257): truncation, clamps on i and f
template <class P>
__device__ inline void baryVec(P x, int high, int& i, P& f) {
    i = (int)val(x);
    i = max(i, 0);
    i = min(i, high - 1);
    f = x - (double)i;
    f = pmax(f, P(0.0));
    f = pmin(f, P(1.0));
}

template <class V, class F>
__device__ inline auto lerp(const V& v0, const V& v1, const F& f) {
    return v0 * (1 - f) + v1 * f;
}
template <class V, class F>
__device__ inline auto bilerp(const V& v00, const V& v01, const V& v11, const V& v10, const F& fx, const F& fy) {
    return lerp(lerp(v00, v10, fx), lerp(v01, v11, fx), fy);
}
template <class V, class F>
__device__ inline auto trilerp(const V& f000, const V& f100, const V& f010, const V& f110, const V& f001,
                               const V& f101, const V& f011, const V& f111, const F& tx, const F& ty, const F& fz) {
    return lerp(bilerp(f000, f010, f110, f100, tx, ty), bilerp(f001, f011, f111, f101, tx, ty), fz);
}

// levelCoordinate (octree_grid_layout3.cpp:16-41): level of the leaf holding p
__device__ __forceinline__ int locate(const Topo& T, double px, double py, double pz) {
    int xi, yi, zi;
    double f;
    bary(px / c_h0, c_R, xi, f);
    bary(py / c_h0, c_R, yi, f);
    bary(pz / c_h0, c_R, zi, f);
    for (int l = 0; l < c_L; ++l) {
        const int g = nd(T, l, xi, yi, zi);
        if (g >= 0 && !ghost(T, g)) return l;
        xi >>= 1;
        yi >>= 1;
        zi >>= 1;
    }
    return -1;
}

// OctreeFaceCenteredGrid3::sampleX/Y/Z (octree_face_centered_grid3.cpp:47-113)
template <class P, class LD>
__device__ __forceinline__ P macAxisR(const Topo& T, LD& ld, const double* mac, int a, int l, P rx, P ry, P rz) {
    const int r = c_res[l];
    const P rr[3] = {rx, ry, rz};
    int i0[3], i1[3];
    P f[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d == a) {
            bary(rr[d], r - 1, i0[d], f[d]);
            i1[d] = min(i0[d] + 1, r - 1);
        } else {
            bary(rr[d] - 0.5, r - 2, i0[d], f[d]);
            i1[d] = min(i0[d] + 1, r - 2);
        }
    }
    auto U = [&](int x, int y, int z) {
        const int g = nd(T, l, x, y, z);
        return g >= 0 ? ld(mac, (size_t)a * c_G + g) : 0.0;
    };
    return trilerp(U(i0[0], i0[1], i0[2]), U(i1[0], i0[1], i0[2]), U(i0[0], i1[1], i0[2]), U(i1[0], i1[1], i0[2]),
                   U(i0[0], i0[1], i1[2]), U(i1[0], i0[1], i1[2]), U(i0[0], i1[1], i1[2]), U(i1[0], i1[1], i1[2]),
                   f[0], f[1], f[2]);
}

// OctreeFaceCenteredGrid3::sample (octree_face_centered_grid3.cpp:36-45)
template <class P, class LD>
__device__ __forceinline__ bool macSampleR(const Topo& T, LD& ld, const double* mac, P px, P py, P pz, P* out) {
    const int l = locate(T, val(px), val(py), val(pz));
    if (l < 0) {
        out[0] = out[1] = out[2] = P(NAN);
        return false;
    }
    const double dh = c_hl[l];
    const P rx = px / dh, ry = py / dh, rz = pz / dh;
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = macAxisR(T, ld, mac, a, l, rx, ry, rz);
    return true;
}
__device__ inline bool macSample(const Topo& T, const double* mac, double px, double py, double pz, double* out) {
    PlainLd ld;
    return macSampleR(T, ld, mac, px, py, pz, out);
}

// CellCenteredGrid::sample (cell_centered_grid.h:40-54) of the boundary SDF:
// cell-centred trilinear with clamped indices (getBaryCentric, vector form).
// No collider: the reference's constant positive SDF.
template <class P>
__device__ __forceinline__ P sdfSampleT(const Topo& T, P px, P py, P pz) {
    if (!T.sdf) return P(1e4);
    int i0, i1, i2;
    P f0, f1, f2;
    baryVec(px / c_h0 - 0.5, c_R - 1, i0, f0);
    baryVec(py / c_h0 - 0.5, c_R - 1, i1, f1);
    baryVec(pz / c_h0 - 0.5, c_R - 1, i2, f2);
    const int j0 = min(i0 + 1, c_R - 1), j1 = min(i1 + 1, c_R - 1), j2 = min(i2 + 1, c_R - 1);
    const size_t R = c_R;
    auto g = [&](int a, int b, int c) { return T.sdf[((size_t)c * R + b) * R + a]; };
    return trilerp(g(i0, i1, i2), g(j0, i1, i2), g(i0, j1, i2), g(j0, j1, i2), g(i0, i1, j2), g(j0, i1, j2),
                   g(i0, j1, j2), g(j0, j1, j2), f0, f1, f2);
}
__device__ inline double sdfSample(const Topo& T, double px, double py, double pz) {
    return sdfSampleT<double>(T, px, py, pz);
}

// cell-centred trilinear at a given level (sampleGivenRelativePosition,
// octree_cell_centered_grid3.h:156-177); NC values per node
template <int NC, class P, class LD>
__device__ __forceinline__ void cellRelT(const Topo& T, LD& ld, const double* cell, int l, P rx, P ry, P rz, P* out) {
    const int r = c_res[l];
    int x0, y0, z0;
    P fx, fy, fz;
    bary(rx - 0.5, r - 2, x0, fx);
    bary(ry - 0.5, r - 2, y0, fy);
    bary(rz - 0.5, r - 2, z0, fz);
    const int x1 = min(x0 + 1, r - 2), y1 = min(y0 + 1, r - 2), z1 = min(z0 + 1, r - 2);
    const int g[8] = {nd(T, l, x0, y0, z0), nd(T, l, x1, y0, z0), nd(T, l, x0, y1, z0), nd(T, l, x1, y1, z0),
                      nd(T, l, x0, y0, z1), nd(T, l, x1, y0, z1), nd(T, l, x0, y1, z1), nd(T, l, x1, y1, z1)};
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = g[k] >= 0 ? ld(cell, (size_t)NC * g[k] + c) : 0.0;
        out[c] = trilerp(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], fx, fy, fz);
    }
}

// OctreeCellCenteredGrid3::sample (octree_cell_centered_grid3.h:97-123)
template <int NC, class P, class LD>
__device__ __forceinline__ void cellSampleT(const Topo& T, LD& ld, const double* cell, P px, P py, P pz, P* out) {
    const int l = locate(T, val(px), val(py), val(pz));
    if (l < 0) {
        for (int c = 0; c < NC; ++c) out[c] = P(NAN);
        return;
    }
    const double dh = c_hl[l];
    cellRelT<NC>(T, ld, cell, l, px / dh, py / dh, pz / dh, out);
}
template <int NC>
__device__ inline void cellSample(const Topo& T, const double* cell, double px, double py, double pz, double* out) {
    PlainLd ld;
    cellSampleT<NC>(T, ld, cell, px, py, pz, out);
}

template <class LD>
__device__ inline double cellvT(const Topo& T, LD& ld, const double* cell, int NC, int l, int x, int y, int z, int c) {
    const int g = nd(T, l, x, y, z);
    return g >= 0 ? ld(cell, (size_t)NC * g + c) : 0.0;
}
__device__ inline double cellv(const Topo& T, const double* cell, int NC, int l, int x, int y, int z, int c) {
    PlainLd ld;
    return cellvT(T, ld, cell, NC, l, x, y, z, c);
}

// getNeighborCubeValue (octree_util.h:124-185), one corner at a time: corner i
// of the cube around the vertex of region dir holds the vertex value (i ==
// dir), the cell centre (the opposite corner, where the reference's marker is
// 0x1000 too), a half-tilt or closed-by-T-joint vertex value at a T-joint
// (the last such match of the axis loop wins, as in the reference's
// overwrites), or the mean of the centre and its neighbours along the
// marker's axes. Returned by value, so the eight corners live in registers.
template <int NC>
struct CubeVal {
    double c[NC];
};
template <int NC, class LD>
__device__ __forceinline__ CubeVal<NC> cubeCorner(const Topo& T, LD& ld, const double* cell, const double* vert,
                                                  int l, int xi, int yi, int zi, int dir, int i, int gt,
                                                  const CubeVal<NC>& centerV, bool tjoint) {
    CubeVal<NC> v;
    if (i == dir) {
#pragma unroll
        for (int c = 0; c < NC; ++c) v.c[c] = gt >= 0 ? ld(vert, (size_t)NC * gt + c) : 0.0;
        return v;
    }
    const int marker = c_marker[dir][i];
    if (marker == 0x1000) return centerV;  // the opposite corner: the cell centre
    const int m0 = (marker >> 8) & 0xf, m1 = (marker >> 4) & 0xf, m2 = marker & 0xf;
    if (tjoint && l > 0) {
        bool flag = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int mk = k == 0 ? m0 : k == 1 ? m1 : m2;
            const int mk1 = k == 0 ? m1 : k == 1 ? m2 : m0;
            const int mk2 = k == 0 ? m2 : k == 1 ? m0 : m1;
            if (mk != 0 && mk1 == 0 && mk2 == 0) {  // face node
                // the finer level's node at the face centre on axis k
                const int fx = 2 * (xi + (k == 0 && mk == 0x1)) + (k != 0);
                const int fy = 2 * (yi + (k == 1 && mk == 0x1)) + (k != 1);
                const int fz = 2 * (zi + (k == 2 && mk == 0x1)) + (k != 2);
                const int g = nd(T, l - 1, fx, fy, fz);
                if (g >= 0 && (T.tb[g] & TB_HALF)) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) v.c[c] = ld(vert, (size_t)NC * g + c);
                    flag = true;
                }
            }
            if (mk == 0 && mk1 != 0 && mk2 != 0) {  // edge node
                const int id0 = (xi << 1) + (dir == 0 || dir == 3 || dir == 4 || dir == 7);
                const int id1 = (yi << 1) + ((dir % 4) < 2);
                const int id2 = (zi << 1) + (dir < 4);
                const int g = nd(T, l - 1, id0 + c_nbt[i][0], id1 + c_nbt[i][1], id2 + c_nbt[i][2]);
                if (g >= 0 && !ghost(T, g) && (T.tb[g] & TB_CLOSED_BY_TJ)) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) v.c[c] = ld(vert, (size_t)NC * g + c);
                    flag = true;
                }
            }
        }
        if (flag) return v;
    }
    const int r = c_res[l];
    double tv[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) tv[c] = centerV.c[c];
    double count = 1;
    for (int x = -(m0 == 0x2); x <= +(m0 == 0x1); ++x)
        for (int y = -(m1 == 0x2); y <= +(m1 == 0x1); ++y)
            for (int z = -(m2 == 0x2); z <= +(m2 == 0x1); ++z) {
                if (x == 0 && y == 0 && z == 0) continue;
                // OctreeCellCenteredGrid3::clamp: indices into [0, res-2]
                const int cx = min(max(0, xi + x), r - 2), cy = min(max(0, yi + y), r - 2),
                          cz = min(max(0, zi + z), r - 2);
#pragma unroll
                for (int c = 0; c < NC; ++c) tv[c] += cellvT(T, ld, cell, NC, l, cx, cy, cz, c);
                count += 1;
            }
#pragma unroll
    for (int c = 0; c < NC; ++c) v.c[c] = tv[c] / count;
    return v;
}

template <class P>
__device__ inline int region3D(const P& x, const P& y, const P& z) {
    int dir = val(z) >= 0.5 ? 0 : 4;
    dir += (val(x) >= 0.5) ? (val(y) >= 0.5 ? 0 : 3) : (val(y) >= 0.5 ? 1 : 2);
    return dir;
}

// OctreeOmniGrid3<T>::sampleSub (octree_omni_grid3.h:242-286)
template <int NC, class P, class LD>
__device__ __forceinline__ void sampleSubT(const Topo& T, LD& ld, const double* cell, const double* vert, P px, P py, P pz, P* out) {
    const int l = locate(T, val(px), val(py), val(pz));
    if (l < 0) {
        for (int c = 0; c < NC; ++c) out[c] = P(NAN);
        return;
    }
    const double dh = c_hl[l];
    const int r = c_res[l];
    const P rx = px / dh, ry = py / dh, rz = pz / dh;
    int xi, yi, zi;
    P fx, fy, fz;
    bary(rx, r - 1, xi, fx);
    bary(ry, r - 1, yi, fy);
    bary(rz, r - 1, zi, fz);
    const int reg = region3D(fx, fy, fz);
    const int gt = nd(T, l, xi + c_nbt[reg][0], yi + c_nbt[reg][1], zi + c_nbt[reg][2]);
    const uint8_t b = tbits(T, gt);
    if (!(b & TB_HALF) && !(b & TB_CLOSED_BY_TJ)) {
        if (b & TB_CLOSED) {
            cellRelT<NC>(T, ld, cell, l, rx, ry, rz, out);
            return;
        }
    }
    const bool tj = (b & TB_TJOINT) != 0;
    CubeVal<NC> cv;
    const int gc = nd(T, l, xi, yi, zi);
#pragma unroll
    for (int c = 0; c < NC; ++c) cv.c[c] = gc >= 0 ? ld(cell, (size_t)NC * gc + c) : 0.0;
    const CubeVal<NC> v0 = cubeCorner<NC>(T, ld, cell, vert, l, xi, yi, zi, reg, 0, gt, cv, tj),
                      v1 = cubeCorner<NC>(T, ld, cell, vert, l, xi, yi, zi, reg, 1, gt, cv, tj),
                      v2 = cubeCorner<NC>(T, ld, cell, vert, l, xi, yi, zi, reg, 2, gt, cv, tj),
                      v3 = cubeCorner<NC>(T, ld, cell, vert, l, xi, yi, zi, reg, 3, gt, cv, tj),
                      v4 = cubeCorner<NC>(T, ld, cell, vert, l, xi, yi, zi, reg, 4, gt, cv, tj),
                      v5 = cubeCorner<NC>(T, ld, cell, vert, l, xi, yi, zi, reg, 5, gt, cv, tj),
                      v6 = cubeCorner<NC>(T, ld, cell, vert, l, xi, yi, zi, reg, 6, gt, cv, tj),
                      v7 = cubeCorner<NC>(T, ld, cell, vert, l, xi, yi, zi, reg, 7, gt, cv, tj);
    switch (reg) {
        case 0: fx -= 0.5, fy -= 0.5, fz -= 0.5; break;
        case 1: fy -= 0.5, fz -= 0.5; break;
        case 2: fz -= 0.5; break;
        case 3: fx -= 0.5, fz -= 0.5; break;
        case 4: fx -= 0.5, fy -= 0.5; break;
        case 5: fy -= 0.5; break;
        case 6: break;
        case 7: fx -= 0.5; break;
    }
    fx *= 2;
    fy *= 2;
    fz *= 2;
#pragma unroll
    for (int c = 0; c < NC; ++c)
        out[c] = trilerp(v6.c[c], v7.c[c], v5.c[c], v4.c[c], v2.c[c], v3.c[c], v1.c[c], v0.c[c], fx, fy, fz);
}
template <int NC>
__device__ inline void sampleSub(const Topo& T, const double* cell, const double* vert, double px, double py, double pz,
                                 double* out) {
    PlainLd ld;
    sampleSubT<NC>(T, ld, cell, vert, px, py, pz, out);
}

// level-transition test shared by the two correction-based samplers
__device__ inline bool nearTransition(const Topo& T, int l, int xi, int yi, int zi) {
    const int g0 = nd(T, l, xi, yi, zi);
    return (g0 >= 0 && ghost(T, g0)) || (l == 0 && nd(T, l + 1, xi >> 1, yi >> 1, zi >> 1) >= 0) ||
           (l == 1 && nd(T, l - 1, (xi << 1) + 1, (yi << 1) + 1, (zi << 1) + 1) >= 0) ||
           (l == 1 && nd(T, l - 1, xi << 1, yi << 1, zi << 1) >= 0);
}

// OctreeOmniFaceCenteredGrid3::sample, correction-based branch
// (octree_omni_face_centered_grid3.h:37-80)
template <class P, class LD>
__device__ __forceinline__ void velSampleT(const Topo& T, const Vel& V, LD& ld, P px, P py, P pz, P* out) {
    const int l = locate(T, val(px), val(py), val(pz));
    if (l < 0) {
        out[0] = out[1] = out[2] = P(NAN);
        return;
    }
    const double dh = c_hl[l];
    const int r = c_res[l];
    int xi, yi, zi;
    P fx, fy, fz;
    baryVec(px / dh - 0.5, r - 1, xi, fx);
    baryVec(py / dh - 0.5, r - 1, yi, fy);
    baryVec(pz / dh - 0.5, r - 1, zi, fz);
    const int gt = nd(T, l, xi + 1, yi + 1, zi + 1);
    const uint8_t b = tbits(T, gt);
    if (nearTransition(T, l, xi, yi, zi)) {
        sampleSubT<3>(T, ld, V.doct, V.dtilt, px, py, pz, out);
        return;
    }
    P v[3];
    macSampleR(T, ld, V.mac, px, py, pz, v);
    if (!(b & TB_HALF) && (b & TB_CLOSED)) {
        out[0] = v[0], out[1] = v[1], out[2] = v[2];
        return;
    }
    P mt[3];
    macSampleR(T, ld, V.mac, P((double)(xi + 1) * dh), P((double)(yi + 1) * dh), P((double)(zi + 1) * dh), mt);
    const P w = pmin(pmin(pmin(fx, fy), fz), pmin(pmin(1 - fx, 1 - fy), 1 - fz));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const P diff = (gt >= 0 ? ld(V.dtilt, 3 * (size_t)gt + c) : 0.0) - mt[c];
        out[c] = v[c] + diff * 2 * w;
    }
}
__device__ inline void velSample(const Topo& T, const Vel& V, double px, double py, double pz, double* out) {
    PlainLd ld;
    velSampleT(T, V, ld, px, py, pz, out);
}

// OctreeOmniGrid3<REAL>::sample, correction-based branch (octree_omni_grid3.h:146-191)
template <class P, class LD>
__device__ __forceinline__ P scalarSampleT(const Topo& T, LD& ld, const double* cell, const double* vert, P px, P py, P pz) {
    const int l = locate(T, val(px), val(py), val(pz));
    if (l < 0) return P(NAN);
    const double dh = c_hl[l];
    const int r = c_res[l];
    int xi, yi, zi;
    P fx, fy, fz;
    bary(px / dh - 0.5, r - 2, xi, fx);
    bary(py / dh - 0.5, r - 2, yi, fy);
    bary(pz / dh - 0.5, r - 2, zi, fz);
    const int x1 = min(xi + 1, r - 2), y1 = min(yi + 1, r - 2), z1 = min(zi + 1, r - 2);
    const double o000 = cellvT(T, ld, cell, 1, l, xi, yi, zi, 0), o100 = cellvT(T, ld, cell, 1, l, x1, yi, zi, 0),
                 o010 = cellvT(T, ld, cell, 1, l, xi, y1, zi, 0), o110 = cellvT(T, ld, cell, 1, l, x1, y1, zi, 0),
                 o001 = cellvT(T, ld, cell, 1, l, xi, yi, z1, 0), o101 = cellvT(T, ld, cell, 1, l, x1, yi, z1, 0),
                 o011 = cellvT(T, ld, cell, 1, l, xi, y1, z1, 0), o111 = cellvT(T, ld, cell, 1, l, x1, y1, z1, 0);
    const P ret = trilerp(o000, o100, o010, o110, o001, o101, o011, o111, fx, fy, fz);
    const int gt = nd(T, l, xi + 1, yi + 1, zi + 1);
    const uint8_t b = tbits(T, gt);
    if (nearTransition(T, l, xi, yi, zi)) {
        P s;
        sampleSubT<1>(T, ld, cell, vert, px, py, pz, &s);
        return s;
    }
    if ((b & TB_CLOSED) && !(b & TB_HALF)) return ret;
    const double diff = (gt >= 0 ? ld(vert, (size_t)gt) : 0.0) -
                        (1 / 8.0) * (o000 + o100 + o010 + o110 + o001 + o101 + o011 + o111);
    const P w = pmin(pmin(pmin(fx, fy), fz), pmin(pmin(1 - fx, 1 - fy), 1 - fz));
    return ret + diff * 2.0 * w;
}
__device__ inline double scalarSample(const Topo& T, const double* cell, const double* vert, double px, double py,
                                      double pz) {
    PlainLd ld;
    return scalarSampleT(T, ld, cell, vert, px, py, pz);
}

// BackTrace, midpoint branch (advection.h:45-62): where the boundary SDF
// changes sign between p and the traced point, the point is pulled back to
// the blend w p + (1 - w) ret, w = |phi1| / (|phi0| + |phi1|).
// The back-trace and, with `sampleEnd`, the velocity at its end point: the
// velocity samples run in one loop that is not unrolled, so a kernel holds one
// inlined copy of the sampler (a call would pass its arrays through scratch
// memory). Step 0 samples at p, step 1 at the midpoint, step 2 (sampleEnd) at
// the end point; `mid` receives the midpoint when not null.
__device__ __forceinline__ void traceSteps(const Topo& T, const Vel& V, const double* p, double dt, bool sampleEnd,
                                           double* out, double* vEnd, double* mid) {
    double pos[3] = {p[0], p[1], p[2]}, v[3];
    const int steps = sampleEnd ? 3 : 2;
#pragma unroll 1
    for (int s = 0; s < steps; ++s) {
        velSample(T, V, pos[0], pos[1], pos[2], v);
        if (s == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) pos[c] = p[c] - v[c] * dt * 0.5;
            if (mid)
                for (int c = 0; c < 3; ++c) mid[c] = pos[c];
        } else if (s == 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) pos[c] = p[c] - v[c] * dt;
            if (T.sdf) {
                const double phi0 = sdfSample(T, p[0], p[1], p[2]);
                const double phi1 = sdfSample(T, pos[0], pos[1], pos[2]);
                if (phi0 * phi1 < 0.0) {
                    const double w = fabs(phi1) / (fabs(phi0) + fabs(phi1));
#pragma unroll
                    for (int c = 0; c < 3; ++c) pos[c] = w * p[c] + (1.0 - w) * pos[c];
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = pos[c];
    if (sampleEnd)
#pragma unroll
        for (int c = 0; c < 3; ++c) vEnd[c] = v[c];
}
__device__ __forceinline__ void backTrace(const Topo& T, const Vel& V, const double* p, double dt, double* out) {
    traceSteps(T, V, p, dt, false, out, nullptr, nullptr);
}

__device__ inline void nodeXYZ(const Topo& T, int g, int& x, int& y, int& z) {
    x = T.nx[g];
    y = T.ny[g];
    z = T.nz[g];
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// constrainVelocity(MAC) without collider: normal faces on closed walls -> 0
// The linear passes below are bodies run by k_run (forward) or recorded
// for their transposes (linrec.h): field reads and writes go through the
// policy P.
struct BWall {
    Topo T;
    const uint8_t* level;
    double* mac;
    int closed;
    template <class Pol>
    __device__ void operator()(int g, Pol& P) const {
        const int l = level[g], r = c_res[l];
        int x, y, z;
        nodeXYZ(T, g, x, y, z);
        if (((closed & ASTG_WALL_XPOS) && x == r - 1) || ((closed & ASTG_WALL_XNEG) && x == 0)) P.st(mac, g, 0.0);
        if (((closed & ASTG_WALL_YPOS) && y == r - 1) || ((closed & ASTG_WALL_YNEG) && y == 0))
            P.st(mac, (size_t)c_G + g, 0.0);
        if (((closed & ASTG_WALL_ZPOS) && z == r - 1) || ((closed & ASTG_WALL_ZNEG) && z == 0))
            P.st(mac, 2 * (size_t)c_G + g, 0.0);
    }
};

// ---- collider (omni_boundary_solver3.h:26-100, wipeBoundary 1240-1292)
// Slip constraint of the MAC faces inside the collider (sdf <= 0), in the
// reference's visiting order: face k (faces in order: levels ascending, leaf
// order, axis x, y, z) samples the MAC velocity at its position reading the
// new values of the inside faces before it and the old values of every other
// face, and keeps u - (u . n) n along its axis, n the unit vector from the
// slip centre (0 if the face sits on the centre). A face publishes its value
// with an agent-scope store over a signalling-NaN sentinel; the faces after
// it poll for it. Persistent (cooperative launch): block b takes chunks b,
// b + grid, ... so the lowest unfinished face always belongs to a running lane.
struct SlipReader {
    const double* mac;
    const int* ord;
    const double* pub;
    int k;
    bool ok;
    // i = axis * G + node, as the MAC sampler addresses the array
    __device__ double operator()(const double* base, size_t i) {
        const int a = (int)(i / (size_t)c_G), g = (int)(i - (size_t)a * c_G);
        const int o = ord[(size_t)g * 3 + a];
        if (o >= 0 && o < k) {
            const unsigned long long b = Sentinel<double>::load(&pub[o]);
            if (b == Sentinel<double>::value) {
                ok = false;
                return 0.0;
            }
            return Sentinel<double>::val(b);
        }
        return mac[(size_t)a * c_G + g];
    }
};

__device__ inline void facePos(const Topo& T, const uint8_t* level, int g, int a, double* p) {
    const double h = c_hl[level[g]];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    p[0] = x * h + (a == 0 ? 0.0 : 0.5) * h;
    p[1] = y * h + (a == 1 ? 0.0 : 0.5) * h;
    p[2] = z * h + (a == 2 ? 0.0 : 0.5) * h;
}

__global__ void __launch_bounds__(BS) k_slip_mac(Topo T, const uint8_t* level, const double* mac, const int* ord,
                                                 const int* faceG, const uint8_t* faceA, int nf, double* pub,
                                                 int* err) {
    const int nchunks = (nf + BS - 1) / BS;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int k = c * BS + threadIdx.x;
        bool pending = k < nf;
        double p[3] = {0, 0, 0};
        int a = 0;
        if (pending) {
            a = faceA[k];
            facePos(T, level, faceG[k], a, p);
        }
        int backoff = 1;
        for (int spins = 0;; ++spins) {
            if (pending) {
                SlipReader rd{mac, ord, pub, k, true};
                double v[3];
                macSampleR(T, rd, mac, p[0], p[1], p[2], v);
                if (rd.ok) {
                    double n[3] = {p[0] - c_slip[0], p[1] - c_slip[1], p[2] - c_slip[2]};
                    const double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                    double val = 0.0;
                    if (!(len < kEps)) {
                        n[0] /= len;
                        n[1] /= len;
                        n[2] /= len;
                        const double d = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
                        val = v[a] - d * n[a];
                    }
                    Sentinel<double>::store(&pub[k], val);
                    pending = false;
                } else if (spins > kSpinLimit) {
                    Sentinel<double>::store(&pub[k], (double)NAN);
                    atomicOr(err, 1);
                    pending = false;
                }
            }
            if (!__any(pending)) break;
            for (int q = 0; q < backoff; ++q) __builtin_amdgcn_s_sleep(1);
            backoff = backoff < 4 ? 2 * backoff : 4;
        }
    }
}

__global__ void k_fill_sent64(double* a, int n) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) Sentinel<double>::storeSent(&a[i]);
}

__global__ void k_slip_scatter(const int* faceG, const uint8_t* faceA, int nf, const double* pub, double* mac) {
    const int k = blockIdx.x * BS + threadIdx.x;
    if (k < nf) mac[(size_t)faceA[k] * c_G + faceG[k]] = pub[k];
}

// constrainVelocity(tilt, MAC) (omni_boundary_solver3.h:26-50) as the
// reference's code behaves: the projected value is assigned to a shadowing
// local 3-vector, so the only change to the tilt velocity is a slot set to 0
// when its position coincides with the slip centre (open tilts whose centre
// lies within h of the collider, slots inside it).
__global__ void k_slip_tilt(Topo T, const uint8_t* level, const int* leafList, int n, double* tilt) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const int g = leafList[t];
    if (T.tb[g] & TB_CLOSED) return;
    const double h = c_hl[level[g]];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    const double c[3] = {x * h, y * h, z * h};
    if (!(sdfSample(T, c[0], c[1], c[2]) - h <= 0)) return;
    const double e = h * 0.5;
    for (int k = 0; k < 8; ++k) {
        const int k4 = k & 0x3;
        const double p[3] = {c[0] + ((k4 == 0 || k4 == 3) ? 1 : -1) * e / 3.0, c[1] + (k4 < 2 ? 1 : -1) * e / 3.0,
                             c[2] + (k < 4 ? 1 : -1) * e / 3.0};
        if (!(sdfSample(T, p[0], p[1], p[2]) <= 0)) continue;
        const double d[3] = {p[0] - c_slip[0], p[1] - c_slip[1], p[2] - c_slip[2]};
        if (sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) < kEps) tilt[(size_t)k * c_G + g] = 0.0;
    }
}

// constrainVelocity(dual grid) (omni_boundary_solver3.h:52-68): the dual
// vertex vectors inside the collider keep their tangential part.
__global__ void k_slip_dual(Topo T, const uint8_t* level, const int* leafList, int n, double* dtilt) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const int g = leafList[t];
    const double h = c_hl[level[g]];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    const double p[3] = {x * h, y * h, z * h};
    if (!(sdfSample(T, p[0], p[1], p[2]) <= 0)) return;
    double nrm[3] = {p[0] - c_slip[0], p[1] - c_slip[1], p[2] - c_slip[2]};
    const double len = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    double* v = dtilt + 3 * (size_t)g;
    if (len < kEps) {
        v[0] = v[1] = v[2] = 0.0;
        return;
    }
    nrm[0] /= len;
    nrm[1] /= len;
    nrm[2] /= len;
    const double d = v[0] * nrm[0] + v[1] * nrm[1] + v[2] * nrm[2];
    const double o[3] = {v[0] - d * nrm[0], v[1] - d * nrm[1], v[2] - d * nrm[2]};
    v[0] = o[0];
    v[1] = o[1];
    v[2] = o[2];
}

// wipeBoundary: density and heat at leaf cell centres and leaf vertices (and
// half-tilt vertices, list kind 1) where the SDF is negative become 0.
__global__ void k_wipe(Topo T, const uint8_t* level, const int* list, int n, int vertexOnly, double* dC, double* dV,
                       double* hC, double* hV) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const int g = list[t];
    const double h = c_hl[level[g]];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    if (!vertexOnly && sdfSample(T, (x + 0.5) * h, (y + 0.5) * h, (z + 0.5) * h) < 0) {
        dC[g] = 0.0;
        hC[g] = 0.0;
    }
    if (sdfSample(T, x * h, y * h, z * h) < 0) {
        dV[g] = 0.0;
        hV[g] = 0.0;
    }
}

struct BMacGhostClear {
    const int* ghostList;
    double* mac;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
        const int g = ghostList[t];
        P.st(mac, g, 0.0);
        P.st(mac, (size_t)c_G + g, 0.0);
        P.st(mac, 2 * (size_t)c_G + g, 0.0);
    }
};

// a policy's loads as the samplers' load functor
template <class Pol>
struct PolLd {
    Pol& P;
    __device__ double operator()(const double* base, size_t i) const { return P.ld(base, i); }
};

// ghost MAC faces resampled at their positions; fromCoarse selects the pass
struct BMacGhostSample {
    Topo T;
    const uint8_t* level;
    const int* ghostList;
    double* mac;
    int fromCoarse;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
        const int g = ghostList[t];
        if (((T.nf[g] & NF_FROM_COARSE) != 0) != (fromCoarse != 0)) return;
        const int l = level[g];
        const double h = c_hl[l];
        int x, y, z;
        nodeXYZ(T, g, x, y, z);
        const double bx = x * h, by = y * h, bz = z * h;
        double v[3];
        PolLd<Pol> ld{P};
        for (int a = 0; a < 3; ++a) {
            const double px = bx + (a == 0 ? 0.0 : 0.5) * h, py = by + (a == 1 ? 0.0 : 0.5) * h,
                         pz = bz + (a == 2 ? 0.0 : 0.5) * h;
            macSampleR(T, ld, (const double*)mac, px, py, pz, v);
            P.st(mac, (size_t)a * c_G + g, v[a]);
        }
    }
};

// updateGhostFromHalfTilt: coarse face <- half-tilt face slot (one writer per face)
struct BMacFromHalf {
    const int *halfG, *target, *axis, *slot;
    const double* tilt;
    double* mac;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
        P.st(mac, (size_t)axis[t] * c_G + target[t], P.ld(tilt, (size_t)slot[t] * c_G + halfG[t]));
    }
};

// OctreeOmniTiltGrid3::updateGhost, half-tilts kept
struct BTiltGhost {
    Topo T;
    const uint8_t* level;
    const int* ghostList;
    const uint8_t* inHalf;
    double* tilt;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
        const int g = ghostList[t];
        if ((T.tb[g] & TB_HALF) || inHalf[g]) return;
        const int l = level[g];
        int x, y, z;
        nodeXYZ(T, g, x, y, z);
        int src;
        if (T.nf[g] & NF_FROM_COARSE) {
            if ((x & 1) || (y & 1) || (z & 1)) {
                for (int k = 0; k < 8; ++k) P.st(tilt, (size_t)k * c_G + g, 0.0);
                return;
            }
            src = nd(T, l + 1, x >> 1, y >> 1, z >> 1);
        } else {
            src = nd(T, l - 1, x << 1, y << 1, z << 1);
        }
        for (int k = 0; k < 8; ++k)
            P.st(tilt, (size_t)k * c_G + g, src >= 0 ? P.ld(tilt, (size_t)k * c_G + src) : 0.0);
    }
};

// updateDualGrid, tilt part
struct BDualTilt {
    Topo T;
    const int* leafList;
    const double* tilt;
    double* dtilt;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
        const int g = leafList[t];
        const size_t o = 3 * (size_t)g;
        if (T.tb[g] & TB_CLOSED) {
            P.st(dtilt, o, 0.0);
            P.st(dtilt, o + 1, 0.0);
            P.st(dtilt, o + 2, 0.0);
            return;
        }
        double gv[8];
        for (int k = 0; k < 8; ++k) gv[k] = P.ld(tilt, (size_t)k * c_G + g);
        const double v[4] = {(gv[0] + gv[6]) * 0.5, (gv[1] + gv[7]) * 0.5, (gv[2] + gv[4]) * 0.5,
                             (gv[3] + gv[5]) * 0.5};
        for (int r = 0; r < 3; ++r) {
            double s = 0;
            for (int j = 0; j < 4; ++j) s += c_dualTilt[r * 4 + j] * v[j];
            P.st(dtilt, o + r, s);
        }
    }
};

__device__ inline void dualApply(int idx, const double* v, double* o) {
    const double* M = c_dual + idx * 21;
    for (int r = 0; r < 3; ++r) {
        double s = 0;
        for (int j = 0; j < 7; ++j) s += M[r * 7 + j] * v[j];
        o[r] = s;
    }
}
__device__ inline int dualIndex(int x, int y, int z, int n0, int n1, int n2, int n3) {
    return ((((((x * 2 + y) * 2 + z) * 2 + n0) * 2 + n1) * 2 + n2) * 2 + n3);
}

// updateDualGrid, octagon part (octree_omni_face_centered_grid3.cpp:50-128)
struct BDualOct {
    Topo T;
    const uint8_t* level;
    const int* leafList;
    const double* mac;
    const double* tilt;
    double* doct;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
    const int g = leafList[t];
    const int l = level[g], r = c_res[l];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    const size_t o = 3 * (size_t)g;
    if (!(x < r - 1 && y < r - 1 && z < r - 1)) {
        P.st(doct, o, 0.0);
        P.st(doct, o + 1, 0.0);
        P.st(doct, o + 2, 0.0);
        return;
    }
    const double dh = c_hl[l];
    const double s3h = kSqrt3 * dh, hs3h = s3h * 0.5;
    int nj[8], tg[8];
    for (int i = 0; i < 8; ++i) {
        tg[i] = nd(T, l, x + c_nbt[i][0], y + c_nbt[i][1], z + c_nbt[i][2]);
        const uint8_t b = tbits(T, tg[i]);
        const double e = tg[i] >= 0 ? T.te[tg[i]] : 0.0;
        nj[i] = ((b & TB_TJOINT) && fabs(e - dh) < kEps) || (b & TB_HALF);
    }
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    int cnt[7] = {0, 0, 0, 0, 0, 0, 0};
    auto MX = [&](int a, int xx, int yy, int zz) {
        const int q = nd(T, l, xx, yy, zz);
        return q >= 0 ? P.ld(mac, (size_t)a * c_G + q) : 0.0;
    };
    if (!((nj[1] && nj[6]) || (nj[2] && nj[5]))) v[0] += MX(0, x, y, z), cnt[0]++;
    if (!((nj[0] && nj[7]) || (nj[3] && nj[4]))) v[0] += MX(0, x + 1, y, z), cnt[0]++;
    if (!((nj[2] && nj[7]) || (nj[3] && nj[6]))) v[1] += MX(1, x, y, z), cnt[1]++;
    if (!((nj[0] && nj[5]) || (nj[1] && nj[4]))) v[1] += MX(1, x, y + 1, z), cnt[1]++;
    if (!((nj[4] && nj[6]) || (nj[5] && nj[7]))) v[2] += MX(2, x, y, z), cnt[2]++;
    if (!((nj[0] && nj[2]) || (nj[1] && nj[3]))) v[2] += MX(2, x, y, z + 1), cnt[2]++;
    for (int i = 0; i < 4; ++i) {
        const int slot = i, reSlot = ((i & 1) ? 8 : 6) - i, vs = i + 3;
        const int gn = tg[slot], gr = tg[reSlot];
        const uint8_t bn = tbits(T, gn), br = tbits(T, gr);
        const double en = gn >= 0 ? T.te[gn] : 0.0, er = gr >= 0 ? T.te[gr] : 0.0;
        if (!(bn & TB_CLOSED) || (bn & TB_HALF)) {
            v[vs] += gn >= 0 ? P.ld(tilt, (size_t)reSlot * c_G + gn) : 0.0;
            cnt[vs]++;
        }
        if (!(br & TB_CLOSED) || (br & TB_HALF)) {
            const double tv = gr >= 0 ? P.ld(tilt, (size_t)slot * c_G + gr) : 0.0;
            // INV_SQRT3 is the unparenthesised macro 1.0 / SQRT3: e * INV_SQRT3 == e * 1.0 / SQRT3
            v[vs] = cnt[vs] == 0 ? tv
                                 : ((v[vs] * (hs3h - er * 1.0 / kSqrt3) + tv * (hs3h - en * 1.0 / kSqrt3)) /
                                    (s3h - en * 1.0 / kSqrt3 - er * 1.0 / kSqrt3));
            cnt[vs]++;
        }
    }
    for (int i = 0; i < 3; ++i) v[i] = v[i] / (cnt[i] + kEps);
    const int idx = dualIndex(cnt[0] > 0, cnt[1] > 0, cnt[2] > 0, cnt[3] > 1 || nj[0] || nj[6],
                              cnt[4] > 1 || nj[1] || nj[7], cnt[5] > 1 || nj[2] || nj[4], cnt[6] > 1 || nj[3] || nj[5]);
    double oo[3];
    dualApply(idx, v, oo);
    for (int k = 0; k < 3; ++k) P.st(doct, o + k, oo[k]);
    }
};

// updateDualGrid, half-tilt part
struct BDualHalf {
    Topo T;
    const int* halfList;
    const double* tilt;
    double* dtilt;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
        const int g = halfList[t];
        const int dir = T.thd[g];
        const int axis = dir % 3;
        double v[7] = {0, 0, 0, 0, 0, 0, 0};
        v[axis] = P.ld(tilt, (size_t)c_halfSlot[dir][4] * c_G + g);
        for (int i = 0; i < 4; ++i) {
            const int slot = c_halfSlot[dir][i];
            const int reSlot = ((slot & 1) ? 8 : 6) - slot;
            const int id = slot < reSlot ? slot : reSlot;
            v[id + 3] = P.ld(tilt, (size_t)slot * c_G + g);
        }
        double oo[3];
        dualApply(dualIndex(axis == 0, axis == 1, axis == 2, 1, 1, 1, 1), v, oo);
        for (int k = 0; k < 3; ++k) P.st(dtilt, 3 * (size_t)g + k, oo[k]);
    }
};

// ---- omni grid ghost chain (octree_omni_grid3.h:136-143), NC values per node
template <int NC>
struct BCellGhostFine {
    Topo T;
    const uint8_t* level;
    const int* ghostList;
    double* cell;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
        const int g = ghostList[t];
        if (T.nf[g] & NF_FROM_COARSE) return;
        const int l = level[g];
        int x, y, z;
        nodeXYZ(T, g, x, y, z);
        x <<= 1;
        y <<= 1;
        z <<= 1;
        PolLd<Pol> ld{P};
        for (int c = 0; c < NC; ++c) {
            double v = 0;
            for (int i = 0; i <= 1; ++i)
                for (int j = 0; j <= 1; ++j)
                    for (int k = 0; k <= 1; ++k) v += cellvT(T, ld, cell, NC, l - 1, x + i, y + j, z + k, c);
            P.st(cell, (size_t)NC * g + c, v * (1 / 8.0));
        }
    }
};

template <int NC>
struct BCellGhostCoarse {
    Topo T;
    const uint8_t* level;
    const int* ghostList;
    double* cell;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
        const int g = ghostList[t];
        if (!(T.nf[g] & NF_FROM_COARSE)) return;
        const int l = level[g];
        const double h = c_hl[l];
        int x, y, z;
        nodeXYZ(T, g, x, y, z);
        double out[NC];
        PolLd<Pol> ld{P};
        cellSampleT<NC>(T, ld, (const double*)cell, x * h + 0.5 * h, y * h + 0.5 * h, z * h + 0.5 * h, out);
        for (int c = 0; c < NC; ++c) P.st(cell, (size_t)NC * g + c, out[c]);
    }
};

template <int NC>
struct BVertGhost {
    Topo T;
    const uint8_t* level;
    const int* ghostList;
    const uint8_t* inHalf;
    double* vert;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
        const int g = ghostList[t];
        if (inHalf[g]) return;
        const int l = level[g];
        int x, y, z;
        nodeXYZ(T, g, x, y, z);
        if (T.nf[g] & NF_FROM_COARSE) {
            double v[NC];
            for (int c = 0; c < NC; ++c) v[c] = 0;
            int count = 0;
            for (int dx = 0; dx <= (x & 1); ++dx)
                for (int dy = 0; dy <= (y & 1); ++dy)
                    for (int dz = 0; dz <= (z & 1); ++dz) {
                        const int q = getUntil(T, l + 1, (x >> 1) + dx, (y >> 1) + dy, (z >> 1) + dz, true);
                        if (q >= 0) {
                            for (int c = 0; c < NC; ++c) v[c] += P.ld(vert, (size_t)NC * q + c);
                            count += 1;
                        }
                    }
            for (int c = 0; c < NC; ++c) P.st(vert, (size_t)NC * g + c, v[c] * (1.0 / (count + kEps)));
        } else {
            const int q = getUntil(T, l - 1, x << 1, y << 1, z << 1, true);
            for (int c = 0; c < NC; ++c)
                P.st(vert, (size_t)NC * g + c, q >= 0 ? P.ld(vert, (size_t)NC * q + c) : 0.0);
        }
    }
};

// updateCellGhost (octree_omni_grid3.h:287-302)
template <int NC>
struct BCellGhostTJ {
    Topo T;
    const uint8_t* level;
    const int* ghostList;
    const double* vert;
    double* cell;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
        const int g = ghostList[t];
        if (!(T.nf[g] & NF_FROM_COARSE)) return;
        const int l = level[g], r = c_res[l];
        int x, y, z;
        nodeXYZ(T, g, x, y, z);
        if (!(x < r - 1 && y < r - 1 && z < r - 1)) return;
        const int gd = nd(T, l, x + (x & 1), y + (y & 1), z + (z & 1));
        const int gp = nd(T, l + 1, x >> 1, y >> 1, z >> 1);
        for (int c = 0; c < NC; ++c) {
            const double a = gd >= 0 ? P.ld(vert, (size_t)NC * gd + c) : 0.0;
            const double b = gp >= 0 ? P.ld(cell, (size_t)NC * gp + c) : 0.0;
            P.st(cell, (size_t)NC * g + c, 0.5 * (a + b));
        }
    }
};

// getTiltFromOctagonNeighbor3 (octree_util.h:94-112)
template <int NC, class LD>
__device__ void tiltFromOctagons(const Topo& T, LD& ld, const double* cell, int l, int x, int y, int z, double* out) {
    const int r = c_res[l];
    for (int c = 0; c < NC; ++c) out[c] = cellvT(T, ld, cell, NC, l, x, y, z, c);
    double count = 1;
    for (int m = -1; m <= 0; ++m)
        for (int nn = -1; nn <= 0; ++nn)
            for (int k = -1; k <= 0; ++k) {
                if (m == 0 && nn == 0 && k == 0) continue;
                const int xx = x + m, yy = y + nn, zz = z + k;
                if (xx >= 0 && yy >= 0 && zz >= 0 && xx < r - 1 && yy < r - 1 && zz < r - 1 &&
                    nd(T, l, xx, yy, zz) >= 0) {
                    for (int c = 0; c < NC; ++c) out[c] += cellvT(T, ld, cell, NC, l, xx, yy, zz, c);
                    count += 1;
                }
            }
    for (int c = 0; c < NC; ++c) out[c] = out[c] / count;
}

// interpolateClosedJoint (octree_omni_grid3.h:305-337), T-joint part
template <int NC>
struct BClosedJoint {
    Topo T;
    const uint8_t* level;
    const int* tjList;
    const double* cell;
    double* vert;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
        const int g = tjList[t];
        int l = level[g];
        int x, y, z;
        nodeXYZ(T, g, x, y, z);
        if (T.te[g] < c_hl[l]) {
            l -= 1;
            x <<= 1;
            y <<= 1;
            z <<= 1;
        }
        PolLd<Pol> ld{P};
        for (int i = 0; i < 6; ++i) {
            const int xx = x + c_nb6[i][0], yy = y + c_nb6[i][1], zz = z + c_nb6[i][2];
            const int q = nd(T, l, xx, yy, zz);
            if (q >= 0) {
                double v[NC];
                tiltFromOctagons<NC>(T, ld, cell, l, xx, yy, zz, v);
                for (int c = 0; c < NC; ++c) P.st(vert, (size_t)NC * q + c, v[c]);
            }
        }
    }
};

template <int NC>
struct BClosedByHalf {
    Topo T;
    const uint8_t* level;
    const int* list;
    const double* cell;
    double* vert;
    int* bad;
    template <class Pol>
    __device__ void operator()(int t, Pol& P) const {
        const int g = list[t];
        if (!((T.tb[g] & TB_CLOSED) && T.thd[g] == -1)) {
            if (Pol::kAux) atomicOr(bad, 1);
            return;
        }
        int x, y, z;
        nodeXYZ(T, g, x, y, z);
        double v[NC];
        PolLd<Pol> ld{P};
        tiltFromOctagons<NC>(T, ld, cell, level[g], x, y, z, v);
        for (int c = 0; c < NC; ++c) P.st(vert, (size_t)NC * g + c, v[c]);
    }
};

// ---- advection (octree_omnigrid_solver3.cpp:850-914, advection.h:195-239)
// kind 0: octagon DOF (cell centre), 1: tilt / half-tilt DOF (vertex)
__global__ void k_adv_scalar(Topo T, const uint8_t* level, Vel V, const int* list, int n, int kind, double dt,
                             const double* d0c, const double* d0v, const double* h0c, const double* h0v, double* d1,
                             double* h1) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const int g = list[t];
    const int l = level[g];
    const double h = c_hl[l];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    double p[3] = {x * h, y * h, z * h};
    if (kind == 0) {
        p[0] += 0.5 * h;
        p[1] += 0.5 * h;
        p[2] += 0.5 * h;
    }
    double q[3];
    backTrace(T, V, p, dt, q);
#pragma unroll 1
    for (int f = 0; f < 2; ++f) {  // density, then heat: one inlined sampler
        const double r = scalarSample(T, f ? h0c : d0c, f ? h0v : d0v, q[0], q[1], q[2]);
        (f ? h1 : d1)[g] = r;
    }
}

__global__ void k_adv_mac(Topo T, const uint8_t* level, Vel V, const int* leafList, int n, double dt, double* mac1) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= 3 * n) return;
    const int a = t / n, g = leafList[t - a * n];
    const int l = level[g];
    const double h = c_hl[l];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    const double p[3] = {x * h + (a == 0 ? 0.0 : 0.5) * h, y * h + (a == 1 ? 0.0 : 0.5) * h,
                         z * h + (a == 2 ? 0.0 : 0.5) * h};
    double q[3], v[3];
    traceSteps(T, V, p, dt, true, q, v, nullptr);
    mac1[(size_t)a * c_G + g] = v[a];
}

// positionTilt (octree_omni_tilt_grid3.h:33-40)
__device__ inline void tiltPos(const double* c, int k, double e, double* p) {
    const int k4 = k & 0x3;
    p[0] = c[0] + ((k4 == 0 || k4 == 3) ? 1 : -1) * e / 3.0;
    p[1] = c[1] + (k4 < 2 ? 1 : -1) * e / 3.0;
    p[2] = c[2] + (k < 4 ? 1 : -1) * e / 3.0;
}
// TILT3_UNIT_DIR_UP[k]: the tilt's offset direction from its centre (the
// signs of positionTilt), turned to point up (+z), times 1/sqrt(3)
__device__ inline void tiltDir(int k, double* d) {
    const double s = 1.0 / kSqrt3;
    const int k4 = k & 0x3;
    const int up = k < 4 ? 1 : -1;
    d[0] = ((k4 == 0 || k4 == 3) ? up : -up) * s;
    d[1] = (k4 < 2 ? up : -up) * s;
    d[2] = s;
}

// T samples the source fields; teDst is the tilt size of the destination grid
// (they differ in the step in which adjustE changed the tilts).
__global__ void k_adv_tilt(Topo T, const double* teDst, const uint8_t* level, Vel V, const int* list, int n, double dt,
                           double* tilt1) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= 8 * n) return;
    const int g = list[t >> 3], k = t & 7;
    const int l = level[g];
    const double h = c_hl[l];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    const double c[3] = {x * h, y * h, z * h};
    double p[3], q[3], v[3], d[3];
    tiltPos(c, k, teDst[g], p);
    traceSteps(T, V, p, dt, true, q, v, nullptr);
    tiltDir(k, d);
    tilt1[(size_t)k * c_G + g] = v[0] * d[0] + v[1] * d[1] + v[2] * d[2];
}

// One lane per (half-tilt, slot): the five slots' back-traces are independent
// chains of gathers; the first slot's lane also clears the three slots a
// half-tilt does not carry.
__global__ void k_adv_half(Topo T, const double* teDst, const uint8_t* level, Vel V, const int* halfList, int n,
                           double dt, double* tilt1) {
    const long t = (long)blockIdx.x * BS + threadIdx.x;
    if (t >= 5L * n) return;
    const int g = halfList[t / 5], i = (int)(t % 5);
    const int dir = T.thd[g];
    const int l = level[g];
    const double h = c_hl[l];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    const double c[3] = {x * h, y * h, z * h};
    const int k = c_halfSlot[dir][i];
    double p[3], q[3], v[3], d[3];
    if (i == 4) {
        p[0] = c[0], p[1] = c[1], p[2] = c[2];
        // HALF_TILT_XYZ_DIR_POSITIVE[k]
        const int ax = (k == 0 || k == 1) ? 0 : (k == 2 || k == 4) ? 1 : 2;
        d[0] = ax == 0, d[1] = ax == 1, d[2] = ax == 2;
    } else {
        tiltPos(c, k, teDst[g], p);
        tiltDir(k, d);
    }
    traceSteps(T, V, p, dt, true, q, v, nullptr);
    tilt1[(size_t)k * c_G + g] = v[0] * d[0] + v[1] * d[1] + v[2] * d[2];
    if (i == 0) {
        unsigned used = 0;
        for (int j = 0; j < 5; ++j) used |= 1u << c_halfSlot[dir][j];
        for (int kk = 0; kk < 8; ++kk)
            if (!((used >> kk) & 1u)) tilt1[(size_t)kk * c_G + g] = 0.0;
    }
}

// adjustE's decision (octree_omnigrid_solver3.cpp:1131-1149) for every leaf
// tilt: an open one closes when its density leaves (dThr, maxThr); a closed
// one opens when the velocity at the tilt exceeds vThr and the density
// sampled at the point it came from lies inside (dThr, maxThr). Fixed tilts
// keep their state.
__global__ void k_adjust_e(Topo T, const uint8_t* level, Vel V, const double* dCell, const double* dVert,
                           const int* leaf, int n, double dThr, double dt, double vThr, double maxThr,
                           uint8_t* open) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const int g = leaf[t];
    const uint8_t b = T.tb[g];
    uint8_t o = (b & TB_LAYOUT_OPEN) ? 1 : 0;
    if (!(b & TB_LAYOUT_FIXED)) {
        if (o) {
            const double d = dVert[g];
            if (d < dThr || d > maxThr) o = 0;
        } else {
            const double h = c_hl[level[g]];
            int x, y, z;
            nodeXYZ(T, g, x, y, z);
            const double p[3] = {x * h, y * h, z * h};
            double v[3];
            velSample(T, V, p[0], p[1], p[2], v);
            if (sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) > vThr) {
                const double d = scalarSample(T, dCell, dVert, p[0] - v[0] * dt, p[1] - v[1] * dt, p[2] - v[2] * dt);
                if (d > dThr && d < maxThr) o = 1;
            }
        }
    }
    open[t] = o;
}

// ---------------------------------------------------------------------------
// sphere source and buoyancy
// ---------------------------------------------------------------------------
__device__ inline bool inSphere(double px, double py, double pz, double cx, double cy, double cz, double r) {
    const double dx = px - cx, dy = py - cy, dz = pz - cz;
    return sqrt(dx * dx + dy * dy + dz * dz) < r;
}

// Smoke source of the octree smoke example (example/octree_smoke_3d/main.cpp:
// 23-81, sphereSource), with its centre, radius, values and inflow speed as
// parameters: inside the sphere, density and heat on cell and vertex DOFs, +z
// velocity on MAC z-faces, and speed/sqrt(3) on all 8 slots of every leaf tilt
// whose centre is inside.
// VEC = false: the example's source, rising at `speed` (v = (0, 0, speed):
// z-faces and tilt slots only). VEC = true: an inflow vector v on every MAC
// face and tilt slot inside the sphere.
template <bool VEC>
__global__ void k_source(Topo T, const uint8_t* level, const int* leaf, int n, double cx, double cy, double cz,
                         double r, double dval, double hval, double vx, double vy, double vz, double* dC, double* dV,
                         double* hC, double* hV, double* mac, double* tilt) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const int g = leaf[t];
    const double h = c_hl[level[g]];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    const double X = x * h, Y = y * h, Z = z * h;
    if (inSphere(X + 0.5 * h, Y + 0.5 * h, Z + 0.5 * h, cx, cy, cz, r)) {
        dC[g] = dval;
        hC[g] = hval;
    }
    if (inSphere(X, Y, Z, cx, cy, cz, r)) {
        dV[g] = dval;
        hV[g] = hval;
        for (int k = 0; k < 8; ++k) {
            if (VEC) {
                double d[3];
                tiltDir(k, d);
                tilt[(size_t)k * c_G + g] = vx * d[0] + vy * d[1] + vz * d[2];
            } else {
                tilt[(size_t)k * c_G + g] = vz / kSqrt3;
            }
        }
    }
    if (VEC) {
        if (inSphere(X, Y + 0.5 * h, Z + 0.5 * h, cx, cy, cz, r)) mac[g] = vx;
        if (inSphere(X + 0.5 * h, Y, Z + 0.5 * h, cx, cy, cz, r)) mac[(size_t)c_G + g] = vy;
    }
    if (inSphere(X + 0.5 * h, Y + 0.5 * h, Z, cx, cy, cz, r)) mac[2 * (size_t)c_G + g] = vz;
}

// The same source on half-tilts: density and heat at the vertex, and the four
// tilt slots of the half-tilt (plus its face slot: for the rising source only
// when the face is a z-face).
template <bool VEC>
__global__ void k_source_half(Topo T, const uint8_t* level, const int* half, int n, double cx, double cy, double cz,
                              double r, double dval, double hval, double vx, double vy, double vz, double* dV,
                              double* hV, double* tilt) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const int g = half[t];
    const double h = c_hl[level[g]];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    if (!inSphere(x * h, y * h, z * h, cx, cy, cz, r)) return;
    dV[g] = dval;
    hV[g] = hval;
    const int dir = T.thd[g];
    double nv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        const int k = c_halfSlot[dir][i];
        if (VEC) {
            double d[3];
            tiltDir(k, d);
            nv[k] = vx * d[0] + vy * d[1] + vz * d[2];
        } else {
            nv[k] = vz / kSqrt3;
        }
    }
    if (VEC) {
        const int ax = dir % 3;
        nv[c_halfSlot[dir][4]] = ax == 0 ? vx : ax == 1 ? vy : vz;
    } else if (dir % 3 == 2) {
        nv[c_halfSlot[dir][4]] = vz;
    }
    for (int k = 0; k < 8; ++k) tilt[(size_t)k * c_G + g] = nv[k];
}

// Per-block sums of one level's cell values (mean temperature of addForce).
__global__ void k_sum_range(const double* a, int off, int n, double* part) {
    double v = 0.0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) v += a[off + i];
    v = wave_sum(v);
    __shared__ double sm[BS / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < BS / 64; ++w) s += sm[w];
        part[blockIdx.x] = s;
    }
}

// addForce (octree_omnigrid_solver3.cpp:917-972): buoyancy f = fd * density +
// ft * (heat - ambT) sampled at each DOF, added along +z over dt.
__device__ inline double buoyancy(const Topo& T, const double* dC, const double* dV, const double* hC,
                                  const double* hV, double fd, double ft, double ambT, const double* p) {
    return fd * scalarSample(T, dC, dV, p[0], p[1], p[2]) + ft * (scalarSample(T, hC, hV, p[0], p[1], p[2]) - ambT);
}
__global__ void k_force_mac(Topo T, const uint8_t* level, const int* leaf, int n, const double* dC, const double* dV,
                            const double* hC, const double* hV, double fd, double ft, double ambT, double dt,
                            double* mac) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const int g = leaf[t];
    const int l = level[g], res = c_res[l];
    const double h = c_hl[l];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    if (!(x <= res - 2 && y <= res - 2 && z <= res - 1)) return;  // isValidZ
    const double p[3] = {x * h + 0.5 * h, y * h + 0.5 * h, z * h};
    double* u = mac + 2 * (size_t)c_G + g;
    *u = *u + dt * buoyancy(T, dC, dV, hC, hV, fd, ft, ambT, p);
}
__global__ void k_force_tilt(Topo T, const uint8_t* level, const int* openTilt, int n, const double* dC,
                             const double* dV, const double* hC, const double* hV, double fd, double ft, double ambT,
                             double dt, double* tilt) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= 8 * n) return;
    const int g = openTilt[t >> 3], k = t & 7;
    const double h = c_hl[level[g]];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    const double c[3] = {x * h, y * h, z * h};
    double p[3], d[3];
    tiltPos(c, k, T.te[g], p);
    tiltDir(k, d);
    double* u = tilt + (size_t)k * c_G + g;
    *u = *u + dt * (d[2] * buoyancy(T, dC, dV, hC, hV, fd, ft, ambT, p));
}
__global__ void k_force_half(Topo T, const uint8_t* level, const int* half, int n, const double* dC,
                             const double* dV, const double* hC, const double* hV, double fd, double ft, double ambT,
                             double dt, double* tilt) {
    // one lane per (half-tilt, slot), as in k_adv_half
    const long t = (long)blockIdx.x * BS + threadIdx.x;
    if (t >= 5L * n) return;
    const int g = half[t / 5], i = (int)(t % 5);
    const int dir = T.thd[g];
    const double h = c_hl[level[g]];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    const double c[3] = {x * h, y * h, z * h};
    const int k = c_halfSlot[dir][i];
    double p[3], dz;
    if (i == 4) {
        p[0] = c[0], p[1] = c[1], p[2] = c[2];
        dz = (k == 0 || k == 1 || k == 2 || k == 4) ? 0.0 : 1.0;  // HALF_TILT_XYZ_DIR_POSITIVE[k].z
    } else {
        double d[3];
        tiltPos(c, k, T.te[g], p);
        tiltDir(k, d);
        dz = d[2];
    }
    double* u = tilt + (size_t)k * c_G + g;
    *u = *u + dt * (buoyancy(T, dC, dV, hC, hV, fd, ft, ambT, p) * dz);
}

__global__ void k_scale2(long n, double s, double* a, double* b) {
    const long i = (long)blockIdx.x * BS + threadIdx.x;
    if (i < n) {
        a[i] *= s;
        b[i] *= s;
    }
}

__global__ void k_points(Topo T, Vel V, const double* cell, const double* vert, int what, int n, const double* xyz,
                         double dt, double* out) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const double* p = xyz + 3 * (size_t)t;
    if (what == 0) velSample(T, V, p[0], p[1], p[2], out + 3 * (size_t)t);
    else if (what == 1) macSample(T, V.mac, p[0], p[1], p[2], out + 3 * (size_t)t);
    else if (what == 2) backTrace(T, V, p, dt, out + 3 * (size_t)t);
    else if (what == 5) out[t] = sdfSample(T, p[0], p[1], p[2]);
    else out[t] = scalarSample(T, cell, vert, p[0], p[1], p[2]);
}

// Staging of astg_sample_points_noalloc: query points (3 per point) then
// results (3 per point), at most kNoAllocPoints points. One copy per GPU, held
// until the process ends; every call on a GPU shares it.
static __device__ __align__(256) double scratch[(size_t)kNoAllocPoints * 6];

// ---------------------------------------------------------------------------
// adjoint of the advection (frozen topology, no collider)
// ---------------------------------------------------------------------------
// Every sampler is linear in the field values with weights that depend on the
// sampled position. The transpose of a sampler at a point is applied by
// probing: run k of the sampler returns 1 for its k-th load and 0 for the
// others, so its output is that load's weight, which is scattered (atomic
// add) into the adjoint of the loaded entry. The derivative with respect to
// the position comes from the same sampler evaluated on D3 positions.
struct ProbeLd {
    int target;
    int cnt = 0;
    const double* tb = nullptr;
    size_t ti = 0;
    __device__ explicit ProbeLd(int k) : target(k) {}
    __device__ double operator()(const double* b, size_t i) {
        const int c = cnt++;
        if (c == target) {
            tb = b;
            ti = i;
            return 1.0;
        }
        return 0.0;
    }
};

// Adjoint arrays of the sampled fields: the forward array a load came from ->
// the adjoint array the weight goes to.
struct AdjMap {
    const double* src[4];
    double* adj[4];
    int n;
    __device__ double* of(const double* b) const {
        for (int r = 0; r < n; ++r)
            if (src[r] == b) return adj[r];
        return nullptr;
    }
};

__device__ inline D3 seed(double v, int k) {
    D3 r(v);
    r.d[k] = 1.0;
    return r;
}

// Scatters sum_c ubar[c] * d out[c] / d field into the adjoints, out = U(p).
__device__ __noinline__ void velSampleTranspose(const Topo& T, const Vel& V, const AdjMap& A, const double* p, const double* ubar) {
    if (ubar[0] == 0.0 && ubar[1] == 0.0 && ubar[2] == 0.0) return;
    double o[3];
    ProbeLd c0(-1);
    velSampleT(T, V, c0, p[0], p[1], p[2], o);
    for (int k = 0; k < c0.cnt; ++k) {
        ProbeLd pk(k);
        velSampleT(T, V, pk, p[0], p[1], p[2], o);
        const double w = ubar[0] * o[0] + ubar[1] * o[1] + ubar[2] * o[2];
        if (w != 0.0 && isfinite(w)) {
            double* a = A.of(pk.tb);
            if (a) atomicAdd(a + pk.ti, w);
        }
    }
}

// J_U(p)^T ubar: the position gradient of ubar . U(p).
__device__ __noinline__ void velSampleGradT(const Topo& T, const Vel& V, const double* p, const double* ubar, double* pbar) {
    PlainLd ld;
    D3 o[3];
    velSampleT(T, V, ld, seed(p[0], 0), seed(p[1], 1), seed(p[2], 2), o);
    for (int k = 0; k < 3; ++k) {
        double s = 0.0;
        for (int c = 0; c < 3; ++c) s += ubar[c] * o[c].d[k];
        pbar[k] = isfinite(s) ? s : 0.0;
    }
}

// Back-trace of the forward (advection.h:45-62, no collider) and its adjoint:
// given qbar = d loss / d q, scatters the velocity-field adjoints of both
// velocity samples (at p and at the midpoint).
__device__ __noinline__ void backTraceAdjoint(const Topo& T, const Vel& V, const AdjMap& A, const double* p, double dt,
                                 const double* mid, const double* qbar) {
    double v2bar[3], mbar[3], v1bar[3];
    for (int c = 0; c < 3; ++c) v2bar[c] = -dt * qbar[c];
    velSampleTranspose(T, V, A, mid, v2bar);
    velSampleGradT(T, V, mid, v2bar, mbar);
    for (int c = 0; c < 3; ++c) v1bar[c] = -dt * 0.5 * mbar[c];
    velSampleTranspose(T, V, A, p, v1bar);
}

// The forward back-trace with its midpoint (no collider on this path)
__device__ __noinline__ void traceFwd(const Topo& T, const Vel& V, const double* p, double dt, double* mid, double* q) {
    traceSteps(T, V, p, dt, false, q, nullptr, mid);
}

// Adjoint of the density advection (k_adv_scalar): DOF g received
// S(rho; q(g)); rbar = the adjoint of that value.
// skipHalf: the list's half-tilt nodes are left to the half-tilt pass, which
// writes the same entries after this one (its write is the one that stays).
__global__ void k_adj_adv_scalar(Topo T, const uint8_t* level, Vel V, AdjMap A, const int* list, int n, int kind,
                                 double dt, const double* d0c, const double* d0v, const double* rbarAdv, double* rbarC,
                                 double* rbarV, const uint8_t* skipHalf) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const int g = list[t];
    if (skipHalf && skipHalf[g]) return;
    const double ab = rbarAdv[g];
    if (ab == 0.0) return;
    const int l = level[g];
    const double h = c_hl[l];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    double p[3] = {x * h, y * h, z * h};
    if (kind == 0) {
        p[0] += 0.5 * h;
        p[1] += 0.5 * h;
        p[2] += 0.5 * h;
    }
    double mid[3], q[3];
    traceFwd(T, V, p, dt, mid, q);
    // S^T at q into the density adjoints
    {
        ProbeLd c0(-1);
        scalarSampleT(T, c0, d0c, d0v, q[0], q[1], q[2]);
        for (int k = 0; k < c0.cnt; ++k) {
            ProbeLd pk(k);
            const double w = ab * scalarSampleT(T, pk, d0c, d0v, q[0], q[1], q[2]);
            if (w != 0.0 && isfinite(w)) atomicAdd((pk.tb == d0c ? rbarC : rbarV) + pk.ti, w);
        }
    }
    // position gradient at q, then the back-trace
    PlainLd ld;
    const D3 sv = scalarSampleT(T, ld, d0c, d0v, seed(q[0], 0), seed(q[1], 1), seed(q[2], 2));
    double qbar[3];
    for (int k = 0; k < 3; ++k) qbar[k] = isfinite(sv.d[k]) ? ab * sv.d[k] : 0.0;
    backTraceAdjoint(T, V, A, p, dt, mid, qbar);
}

// Adjoint of the velocity advection: kind 0 MAC face (axis = sub), 1 tilt
// slot (slot = sub), 2 half-tilt slot (slot = sub, face slot at i == 4).
__device__ void advVelDof(const Topo& T, const double* teDst, const uint8_t* level, const Vel& V, const AdjMap& A,
                          int g, int kind, int sub, double dt, double ab) {
    const int l = level[g];
    const double h = c_hl[l];
    int x, y, z;
    nodeXYZ(T, g, x, y, z);
    double p[3], d[3] = {0, 0, 0};
    if (kind == 0) {
        const int a = sub;
        p[0] = x * h + (a == 0 ? 0.0 : 0.5) * h;
        p[1] = y * h + (a == 1 ? 0.0 : 0.5) * h;
        p[2] = z * h + (a == 2 ? 0.0 : 0.5) * h;
        d[a] = 1.0;
    } else {
        const double c[3] = {x * h, y * h, z * h};
        const int dir = T.thd[g];
        const int k = kind == 1 ? sub : c_halfSlot[dir][sub];
        if (kind == 2 && sub == 4) {
            p[0] = c[0], p[1] = c[1], p[2] = c[2];
            const int ax = (k == 0 || k == 1) ? 0 : (k == 2 || k == 4) ? 1 : 2;
            d[ax] = 1.0;
        } else {
            tiltPos(c, k, teDst[g], p);
            tiltDir(k, d);
        }
    }
    double mid[3], q[3];
    traceFwd(T, V, p, dt, mid, q);
    double ubar[3] = {ab * d[0], ab * d[1], ab * d[2]}, qbar[3];
    velSampleTranspose(T, V, A, q, ubar);
    velSampleGradT(T, V, q, ubar, qbar);
    backTraceAdjoint(T, V, A, p, dt, mid, qbar);
}

__global__ void k_adj_adv_mac(Topo T, const double* teDst, const uint8_t* level, Vel V, AdjMap A, const int* leafList,
                              int n, double dt, const double* ubarMac) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= 3 * n) return;
    const int a = t / n, g = leafList[t - a * n];
    const double ab = ubarMac[(size_t)a * c_G + g];
    if (ab != 0.0) advVelDof(T, teDst, level, V, A, g, 0, a, dt, ab);
}

// Half-tilt nodes among the velocity tilts are left to k_adj_adv_half: the
// forward's half-tilt pass overwrites every slot k_adv_tilt wrote for them.
__global__ void k_adj_adv_tilt(Topo T, const double* teDst, const uint8_t* level, Vel V, AdjMap A, const int* list,
                               int n, double dt, const double* ubarTilt, const uint8_t* inHalf) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= 8 * n) return;
    const int g = list[t >> 3], k = t & 7;
    if (inHalf[g]) return;
    const double ab = ubarTilt[(size_t)k * c_G + g];
    if (ab != 0.0) advVelDof(T, teDst, level, V, A, g, 1, k, dt, ab);
}

__global__ void k_adj_adv_half(Topo T, const double* teDst, const uint8_t* level, Vel V, AdjMap A, const int* halfList,
                               int n, double dt, const double* ubarTilt) {
    const long t = (long)blockIdx.x * BS + threadIdx.x;
    if (t >= 5L * n) return;
    const int g = halfList[t / 5], i = (int)(t % 5);
    const int k = c_halfSlot[T.thd[g]][i];
    const double ab = ubarTilt[(size_t)k * c_G + g];
    if (ab != 0.0) advVelDof(T, teDst, level, V, A, g, 2, i, dt, ab);
}

// Entries an advection step writes (1) or carries over (0): MAC of leaves,
// tilt slots of velocity tilts and half-tilts, density of leaf cells and of
// open-tilt and half-tilt vertices.
__global__ void k_mark(const int* list, int n, int width, int G, uint8_t* mask) {
    const long t = (long)blockIdx.x * BS + threadIdx.x;
    if (t >= (long)width * n) return;
    const int k = (int)(t / n), g = list[t - (long)k * n];
    mask[(size_t)k * G + g] = 1;
}

// out = in where the entry is carried over (mask 0), 0 where written.
__global__ void k_keep_unwritten(long n, const uint8_t* mask, const double* in, double* out) {
    const long i = (long)blockIdx.x * BS + threadIdx.x;
    if (i < n) out[i] = mask[i] ? 0.0 : in[i];
}

__global__ void k_scale1(long n, double s, double* a) {
    const long i = (long)blockIdx.x * BS + threadIdx.x;
    if (i < n) a[i] *= s;
}

// (M^T M)^-1 M^T of the rows selected by the 7 flags, in long double with
// partial pivoting (Constants::init, constants.cpp:170-256)
void dualMatrices(double* out) {
    const long double s = 1.0L / 1.7320508075688772L;
    const long double dirs[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {s, s, s}, {-s, s, s}, {-s, -s, s}, {s, -s, s}};
    for (int idx = 0; idx < 128; ++idx) {
        double* o = out + idx * 21;
        for (int q = 0; q < 21; ++q) o[q] = 0.0;
        int sel[7], cnt = 0;
        for (int j = 0; j < 7; ++j)
            if ((idx >> (6 - j)) & 1) sel[cnt++] = j;
        if (cnt <= 1) continue;
        long double A[3][3] = {}, I[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int q = 0; q < cnt; ++q) A[a][b] += dirs[sel[q]][a] * dirs[sel[q]][b];
        for (int c = 0; c < 3; ++c) {
            int pr = c;
            for (int r = c + 1; r < 3; ++r)
                if (fabsl(A[r][c]) > fabsl(A[pr][c])) pr = r;
            for (int k = 0; k < 3; ++k) {
                std::swap(A[c][k], A[pr][k]);
                std::swap(I[c][k], I[pr][k]);
            }
            const long double d = A[c][c];
            for (int k = 0; k < 3; ++k) {
                A[c][k] /= d;
                I[c][k] /= d;
            }
            for (int r = 0; r < 3; ++r)
                if (r != c) {
                    const long double f = A[r][c];
                    for (int k = 0; k < 3; ++k) {
                        A[r][k] -= f * A[c][k];
                        I[r][k] -= f * I[c][k];
                    }
                }
        }
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < cnt; ++q) {
                long double v = 0;
                for (int k = 0; k < 3; ++k) v += I[r][k] * dirs[sel[q]][k];
                o[r * 7 + sel[q]] = (double)v;
            }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int AdvectDevice::upload(const Octree& o, const ProjectTables& t, bool keepFields, std::string& err) {
    diffReady_ = false;  // recorded passes belong to the previous topology
    if (keepFields) {
        // a tilt update keeps the node sets; anything else would leave the
        // index maps, node lists and coordinates on the device stale
        bool same = t.G == G_ && o.L == L_ && o.R == R_ && (int)nnode_.size() == o.L;
        for (int l = 0; same && l < o.L; ++l)
            same = o.lv[l].nnode == nnode_[l] && o.lv[l].nleaf == nleaf_[l] && t.off[l] == off_[l];
        if (!same) {
            err = "tilt update changed the node sets";
            return ASTG_ERR_STATE;
        }
    }
    if (keepFields && !sampValid_) {
        // the fields keep the tilt state they were written with until the next
        // advect has sampled them
        HIPCHK(sampTe_.alloc((size_t)G_ * 8));
        HIPCHK(sampTb_.alloc((size_t)G_));
        HIPCHK(sampThd_.alloc((size_t)G_));
        HIPCHK(hipMemcpyAsync(sampTe_.p, te_.p, (size_t)G_ * 8, hipMemcpyDeviceToDevice, stream_));
        HIPCHK(hipMemcpyAsync(sampTb_.p, tb_.p, (size_t)G_, hipMemcpyDeviceToDevice, stream_));
        HIPCHK(hipMemcpyAsync(sampThd_.p, thd_.p, (size_t)G_, hipMemcpyDeviceToDevice, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        sampValid_ = true;
    }
    if (!keepFields) sampValid_ = false;
    if (o.L > kMaxLevels) {
        err = "advection supports at most 8 levels";
        return ASTG_ERR_ARG;
    }
    L_ = o.L;
    R_ = o.R;
    h0_ = o.h0;
    closed_ = o.closed_flags;
    G_ = t.G;
    off_ = t.off;
    nleaf_.assign(L_, 0);
    nnode_.assign(L_, 0);
    std::vector<int> res(L_), moff(L_);
    std::vector<double> hl(L_);
    long mtot = 0;
    for (int l = 0; l < L_; ++l) {
        res[l] = o.lv[l].res;
        hl[l] = o.lv[l].h;
        nleaf_[l] = o.lv[l].nleaf;
        nnode_[l] = o.lv[l].nnode;
        moff[l] = (int)mtot;
        mtot += (long)res[l] * res[l] * res[l];
    }
    if (mtot >= (1L << 31)) {
        err = "index maps too large";
        return ASTG_ERR_ARG;
    }
    cRes_ = res;
    cMoff_ = moff;
    cHl_ = hl;
    if (cDual_.empty()) {
        cDual_.resize(128 * 21);
        dualMatrices(cDual_.data());
    }
    for (int k = 0; k < 3; ++k) cSlip_[k] = o.sdf.slip[k];
    std::lock_guard<std::recursive_mutex> lk(constMutex());
    {
        int rc = bindConstants(true, err);
        if (rc) return rc;
    }
    // A tilt update (keepFields) keeps the node sets: the index maps, node
    // coordinates and flags, leaf and ghost lists and levels on the device
    // stay as they are; only the tilt-state dependent lists are rebuilt.
    const bool sameNodes = keepFields;
    std::vector<int> map(sameNodes ? 0 : (size_t)mtot);
    std::vector<uint16_t>& nx = hnx_;
    std::vector<uint16_t>& ny = hny_;
    std::vector<uint16_t>& nz = hnz_;
    if (!sameNodes) {
        nx.assign(G_, 0);
        ny.assign(G_, 0);
        nz.assign(G_, 0);
    }
    std::vector<uint8_t> nf(sameNodes ? 0 : G_), inHalf(G_, 0);
    std::vector<int> leaf, ghostL, openTilt, half, tj, cbh;
    for (int l = 0; l < L_; ++l) {
        const Level& Lv = o.lv[l];
        if (!sameNodes)
            for (size_t q = 0; q < Lv.map.size(); ++q) map[moff[l] + q] = Lv.map[q] >= 0 ? Lv.map[q] + off_[l] : -1;
        for (int i = 0; i < Lv.nnode; ++i) {
            const int g = off_[l] + i;
            if (!sameNodes) {
                nx[g] = Lv.x[i];
                ny[g] = Lv.y[i];
                nz[g] = Lv.z[i];
                nf[g] = Lv.flag[i];
            }
            if (i < Lv.nleaf) {
                leaf.push_back(g);
                const uint8_t b = o.tilt[l][i].bits;
                if (b & TB_LAYOUT_OPEN) openTilt.push_back(g);
                if (b & TB_TJOINT) tj.push_back(g);
            } else {
                ghostL.push_back(g);
            }
        }
    }
    for (lcoor_t c : o.half) {
        const int g = o.node(lc_level(c), lc_x(c), lc_y(c), lc_z(c));
        if (g < 0) continue;
        half.push_back(g + off_[lc_level(c)]);
        inHalf[g + off_[lc_level(c)]] = 1;
    }
    for (lcoor_t c : o.closed_by_half) {
        const int g = o.node(lc_level(c), lc_x(c), lc_y(c), lc_z(c));
        if (g >= 0) cbh.push_back(g + off_[lc_level(c)]);
    }
    // velocity tilt DOFs: non-closed leaf tilts (tiltE state)
    std::vector<int> velTilt;
    for (int l = 0; l < L_; ++l)
        for (int i = 0; i < o.lv[l].nleaf; ++i)
            if (!(o.tilt[l][i].bits & TB_CLOSED)) velTilt.push_back(off_[l] + i);
    // coarse MAC faces overwritten from half-tilts: the last half-tilt in list order wins
    std::vector<int> hmG, hmT, hmA, hmS;
    {
        std::vector<int> winner;  // index into half, per (target, axis)
        std::vector<long> key;
        for (size_t q = 0; q < half.size(); ++q) {
            const int g = half[q];
            const int l = t.level[g];
            const int dir = t.tilt_hd[g];
            const int axis = dir % 3;
            const int tgt = o.node(l + 1, nx[g] >> 1, ny[g] >> 1, nz[g] >> 1);
            if (tgt < 0) continue;
            key.push_back(((long)(tgt + off_[l + 1]) * 3 + axis) * (long)half.size() + (long)q);
        }
        std::sort(key.begin(), key.end());
        for (size_t q = 0; q < key.size(); ++q) {
            const long fa = key[q] / (long)half.size();
            if (q + 1 < key.size() && key[q + 1] / (long)half.size() == fa) continue;  // a later writer exists
            const int hq = (int)(key[q] % (long)half.size());
            const int g = half[hq];
            const int dir = t.tilt_hd[g];
            hmG.push_back(g);
            hmT.push_back((int)(fa / 3));
            hmA.push_back((int)(fa % 3));
            hmS.push_back(kHalfSlot[dir][4]);
        }
    }
    nLeaf_ = (int)leaf.size();
    nGhost_ = (int)ghostL.size();
    nHalf_ = (int)half.size();
    nTJ_ = (int)tj.size();
    nCBH_ = (int)cbh.size();
    nOpenTilt_ = (int)openTilt.size();
    auto up = [&](DevBuf& d, const void* src, size_t bytes) -> hipError_t {
        hipError_t e = d.alloc(bytes ? bytes : 4);
        if (e != hipSuccess || !bytes) return e;
        return hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, stream_);
    };
    if (!sameNodes) {
        HIPCHK(up(map_, map.data(), map.size() * 4));
        HIPCHK(up(nx_, nx.data(), (size_t)G_ * 2));
        HIPCHK(up(ny_, ny.data(), (size_t)G_ * 2));
        HIPCHK(up(nz_, nz.data(), (size_t)G_ * 2));
        HIPCHK(up(nflag_, nf.data(), (size_t)G_));
        HIPCHK(up(leaf_, leaf.data(), leaf.size() * 4));
        HIPCHK(up(ghost_, ghostL.data(), ghostL.size() * 4));
        HIPCHK(level_.alloc((size_t)G_));
        HIPCHK(hipMemcpyAsync(level_.p, t.level.data(), (size_t)G_, hipMemcpyHostToDevice, stream_));
    }
    HIPCHK(up(te_, t.tilt_e.data(), (size_t)G_ * 8));
    HIPCHK(up(tb_, t.tilt_bits.data(), (size_t)G_));
    HIPCHK(up(thd_, t.tilt_hd.data(), (size_t)G_));
    HIPCHK(up(inHalf_, inHalf.data(), (size_t)G_));
    HIPCHK(up(half_, half.data(), half.size() * 4));
    HIPCHK(up(tj_, tj.data(), tj.size() * 4));
    HIPCHK(up(cbh_, cbh.data(), cbh.size() * 4));
    HIPCHK(up(openTilt_, openTilt.data(), openTilt.size() * 4));
    HIPCHK(up(velTilt_, velTilt.data(), velTilt.size() * 4));
    nVelTilt_ = (int)velTilt.size();
    nHalfMac_ = (int)hmG.size();
    // collider: the SDF grid and the MAC faces inside it (sdf <= 0 at the face
    // position, isValidX/Y/Z), in the reference's visiting order
    hasSdf_ = o.sdf.present();
    if (hasSdf_ && !sameNodes) {
        if (o.sdf.R != R_ || (long)o.sdf.phi.size() != (long)R_ * R_ * R_) {
            err = "boundary SDF does not match the finest resolution";
            return ASTG_ERR_ARG;
        }
        HIPCHK(up(sdf_, o.sdf.phi.data(), o.sdf.phi.size() * 8));
        std::vector<int> fg, ord((size_t)G_ * 3, -1);
        std::vector<uint8_t> fa;
        for (int l = 0; l < L_; ++l) {
            const Level& Lv = o.lv[l];
            const double h = Lv.h;
            const int r = Lv.res;
            for (int i = 0; i < Lv.nleaf; ++i) {
                const int x = Lv.x[i], y = Lv.y[i], z = Lv.z[i];
                for (int a = 0; a < 3; ++a) {
                    const bool valid = (a == 0 ? x <= r - 1 : x <= r - 2) && (a == 1 ? y <= r - 1 : y <= r - 2) &&
                                       (a == 2 ? z <= r - 1 : z <= r - 2);
                    if (!valid) continue;
                    const double px = x * h + (a == 0 ? 0.0 : 0.5) * h, py = y * h + (a == 1 ? 0.0 : 0.5) * h,
                                 pz = z * h + (a == 2 ? 0.0 : 0.5) * h;
                    if (!(o.sdf.sample(px, py, pz) <= 0)) continue;
                    const int g = off_[l] + i;
                    ord[(size_t)g * 3 + a] = (int)fg.size();
                    fg.push_back(g);
                    fa.push_back((uint8_t)a);
                }
            }
        }
        nSlip_ = (int)fg.size();
        HIPCHK(up(slipFaceG_, fg.data(), fg.size() * 4));
        HIPCHK(up(slipFaceA_, fa.data(), fa.size()));
        HIPCHK(up(slipOrd_, ord.data(), ord.size() * 4));
        HIPCHK(slipPub_.alloc((size_t)std::max(nSlip_, 1) * 8));
        HIPCHK(slipErr_.alloc(4));
        int perCu = 0, dev = 0, cus = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_slip_mac, BS, 0));
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        slipGrid_ = std::max(1, std::min(perCu * cus, (nSlip_ + BS - 1) / BS));
    }
    if (!hasSdf_) nSlip_ = 0;
    HIPCHK(up(hmG_, hmG.data(), hmG.size() * 4));
    HIPCHK(up(hmT_, hmT.data(), hmT.size() * 4));
    HIPCHK(up(hmA_, hmA.data(), hmA.size() * 4));
    HIPCHK(up(hmS_, hmS.data(), hmS.size() * 4));
    if (keepFields) {
        HIPCHK(hipStreamSynchronize(stream_));
        return ASTG_OK;
    }
    // fields: both buffers of each scalar start at zero, as the reference's
    // temporaries do
    for (int f = 0; f < 2; ++f)
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < 2; ++k) {
                HIPCHK(scal_[f][b][k].alloc((size_t)G_ * 8));
                HIPCHK(hipMemsetAsync(scal_[f][b][k].p, 0, (size_t)G_ * 8, stream_));
            }
    cur_[0] = cur_[1] = 0;
    HIPCHK(dualOct_.alloc((size_t)G_ * 24));
    HIPCHK(dualTilt_.alloc((size_t)G_ * 24));
    HIPCHK(hipMemsetAsync(dualOct_.p, 0, (size_t)G_ * 24, stream_));
    HIPCHK(hipMemsetAsync(dualTilt_.p, 0, (size_t)G_ * 24, stream_));
    HIPCHK(bad_.alloc(4));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

int AdvectDevice::bindConstants(bool force, std::string& err) {
    if (!takeConstants(CT_ADVECT, constOwner_, force)) return ASTG_OK;
    const double s = 1.0 / 1.7320508075688772;
    double dt4[12] = {s, -s, -s, s, s, s, -s, -s, s, s, s, s};
    for (double& v : dt4) v *= 3.0 / 4.0;
    int LL = L_, RR = R_, GG = G_;
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_L), &LL, sizeof(int), 0, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_R), &RR, sizeof(int), 0, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_G), &GG, sizeof(int), 0, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_h0), &h0_, sizeof(double), 0, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_res), cRes_.data(), sizeof(int) * L_, 0, hipMemcpyHostToDevice,
                                  stream_));
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_moff), cMoff_.data(), sizeof(int) * L_, 0, hipMemcpyHostToDevice,
                                  stream_));
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_hl), cHl_.data(), sizeof(double) * L_, 0, hipMemcpyHostToDevice,
                                  stream_));
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_dual), cDual_.data(), sizeof(double) * cDual_.size(), 0,
                                  hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_dualTilt), dt4, sizeof(dt4), 0, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_slip), cSlip_, sizeof(double) * 3, 0, hipMemcpyHostToDevice, stream_));
    return ASTG_OK;
}

namespace {
std::mutex g_scratchMu;
double* g_scratchPtr[kMaxDevices] = {};
}  // namespace

int noallocScratchInit(int device, std::string& err) {
    if (device < 0 || device >= kMaxDevices) {
        err = "device ordinal beyond the scratch table";
        return ASTG_ERR_ARG;
    }
    std::lock_guard<std::mutex> g(g_scratchMu);
    if (g_scratchPtr[device]) return ASTG_OK;
    // the caller made `device` current; the symbol's address is this GPU's copy
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratchPtr[device] = static_cast<double*>(p);
    return ASTG_OK;
}

double* noallocScratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_scratchMu);
    return g_scratchPtr[device];
}

int AdvectDevice::samplePointsStaged(ProjectDevice& pd, int what, int n, const double* xyz, double dt, double* out,
                                     double* stage, std::string& err) {
    std::lock_guard<std::recursive_mutex> lk_(constMutex());
    if (int rc_ = bindConstants(false, err)) return rc_;
    if (what < 0 || what > 5 || n < 0) {
        err = "bad query";
        return ASTG_ERR_ARG;
    }
    if (n > kNoAllocPoints) {
        err = "at most " + std::to_string(kNoAllocPoints) + " points per call";
        return ASTG_ERR_ARG;
    }
    if (!stage) {
        err = "scratch not set up on this device";
        return ASTG_ERR_STATE;
    }
    if (n == 0) return ASTG_OK;
    const int w = what <= 2 ? 3 : 1;
    double* dPts = stage;
    double* dOut = stage + (size_t)kNoAllocPoints * 3;
    HIPCHK(hipMemcpyAsync(dPts, xyz, (size_t)n * 24, hipMemcpyHostToDevice, stream_));
    const int f = what == 4 ? 1 : 0;
    k_points<<<nblocks(n), BS, 0, stream_>>>(topo(), vel(pd), scal_[f][cur_[f]][0].as<double>(),
                                             scal_[f][cur_[f]][1].as<double>(), what, n, dPts, dt, dOut);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dOut, (size_t)n * 8 * w, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

AdvTopo AdvectDevice::topo() const {
    AdvTopo T;
    T.map = map_.as<int>();
    T.nx = static_cast<const uint16_t*>(nx_.p);
    T.ny = static_cast<const uint16_t*>(ny_.p);
    T.nz = static_cast<const uint16_t*>(nz_.p);
    T.nf = static_cast<const uint8_t*>(nflag_.p);
    T.te = te_.as<double>();
    T.tb = static_cast<const uint8_t*>(tb_.p);
    T.thd = static_cast<const int8_t*>(thd_.p);
    T.sdf = hasSdf_ ? sdf_.as<double>() : nullptr;
    return T;
}

AdvTopo AdvectDevice::sampTopo() const {
    AdvTopo T = topo();
    if (sampValid_) {
        T.te = sampTe_.as<double>();
        T.tb = static_cast<const uint8_t*>(sampTb_.p);
        T.thd = static_cast<const int8_t*>(sampThd_.p);
    }
    return T;
}

int AdvectDevice::decideTilts(ProjectDevice& pd, double dThr, double dt, double vThr, double maxThr,
                              std::vector<uint8_t>& open, std::string& err) {
    std::lock_guard<std::recursive_mutex> lk_(constMutex());
    if (int rc_ = bindConstants(false, err)) return rc_;
    open.assign(nLeaf_, 0);
    if (!nLeaf_) return ASTG_OK;
    HIPCHK(open_.alloc((size_t)nLeaf_));
    const int d = cur_[0];
    k_adjust_e<<<nblocks(nLeaf_), BS, 0, stream_>>>(sampTopo(), static_cast<const uint8_t*>(level_.p), vel(pd),
                                                    scal_[0][d][0].as<double>(), scal_[0][d][1].as<double>(),
                                                    leaf_.as<int>(), nLeaf_, dThr, dt, vThr, maxThr,
                                                    open_.as<uint8_t>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(open.data(), open_.p, (size_t)nLeaf_, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

int AdvectDevice::setScalar(int field, int level, const double* oct, const double* tilt, std::string& err) {
    if (field < 0 || field > 1 || level < 0 || level >= L_) {
        err = "bad field or level";
        return ASTG_ERR_ARG;
    }
    const size_t o = (size_t)off_[level], n = (size_t)nnode_[level];
    HIPCHK(hipMemcpyAsync(scal_[field][cur_[field]][0].as<double>() + o, oct, n * 8, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyAsync(scal_[field][cur_[field]][1].as<double>() + o, tilt, n * 8, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

int AdvectDevice::getScalar(int field, int level, double* oct, double* tilt, std::string& err) {
    if (field < 0 || field > 1 || level < 0 || level >= L_) {
        err = "bad field or level";
        return ASTG_ERR_ARG;
    }
    const size_t o = (size_t)off_[level], n = (size_t)nnode_[level];
    HIPCHK(hipMemcpyAsync(oct, scal_[field][cur_[field]][0].as<double>() + o, n * 8, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(tilt, scal_[field][cur_[field]][1].as<double>() + o, n * 8, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

int AdvectDevice::getDual(int level, double* oct3, double* tilt3, std::string& err) {
    if (level < 0 || level >= L_) {
        err = "bad level";
        return ASTG_ERR_ARG;
    }
    const size_t o = (size_t)off_[level], n = (size_t)nnode_[level];
    HIPCHK(hipMemcpyAsync(oct3, dualOct_.as<double>() + 3 * o, n * 24, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(tilt3, dualTilt_.as<double>() + 3 * o, n * 24, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

int AdvectDevice::omniGhost(double* cell, double* vert, int nc, std::string& err) {
    const Topo T = topo();
    const uint8_t* lev = static_cast<const uint8_t*>(level_.p);
    const int* gl = ghost_.as<int>();
    const uint8_t* ih = static_cast<const uint8_t*>(inHalf_.p);
    const int gb = nblocks(nGhost_);
    if (nc == 1) {
        if (nGhost_) {
            k_run<<<gb, BS, 0, stream_>>>(BCellGhostFine<1>{T, lev, gl, cell}, nGhost_);
            k_run<<<gb, BS, 0, stream_>>>(BCellGhostCoarse<1>{T, lev, gl, cell}, nGhost_);
            k_run<<<gb, BS, 0, stream_>>>(BVertGhost<1>{T, lev, gl, ih, vert}, nGhost_);
            k_run<<<gb, BS, 0, stream_>>>(BCellGhostTJ<1>{T, lev, gl, vert, cell}, nGhost_);
        }
        if (nTJ_) k_run<<<nblocks(nTJ_), BS, 0, stream_>>>(BClosedJoint<1>{T, lev, tj_.as<int>(), cell, vert}, nTJ_);
        if (nCBH_)
            k_run<<<nblocks(nCBH_), BS, 0, stream_>>>(BClosedByHalf<1>{T, lev, cbh_.as<int>(), cell, vert,
                                                                         bad_.as<int>()}, nCBH_);
    } else {
        if (nGhost_) {
            k_run<<<gb, BS, 0, stream_>>>(BCellGhostFine<3>{T, lev, gl, cell}, nGhost_);
            k_run<<<gb, BS, 0, stream_>>>(BCellGhostCoarse<3>{T, lev, gl, cell}, nGhost_);
            k_run<<<gb, BS, 0, stream_>>>(BVertGhost<3>{T, lev, gl, ih, vert}, nGhost_);
            k_run<<<gb, BS, 0, stream_>>>(BCellGhostTJ<3>{T, lev, gl, vert, cell}, nGhost_);
        }
        if (nTJ_) k_run<<<nblocks(nTJ_), BS, 0, stream_>>>(BClosedJoint<3>{T, lev, tj_.as<int>(), cell, vert}, nTJ_);
        if (nCBH_)
            k_run<<<nblocks(nCBH_), BS, 0, stream_>>>(BClosedByHalf<3>{T, lev, cbh_.as<int>(), cell, vert,
                                                                         bad_.as<int>()}, nCBH_);
    }
    HIPCHK(hipGetLastError());
    return ASTG_OK;
}

int AdvectDevice::constrainAndGhost(ProjectDevice& pd, std::string& err, bool constrain) {
    std::lock_guard<std::recursive_mutex> lk_(constMutex());
    if (int rc_ = bindConstants(false, err)) return rc_;
    const Topo T = topo();
    const uint8_t* lev = static_cast<const uint8_t*>(level_.p);
    double* mac = pd.uMac();
    double* tilt = pd.uTilt();
    // constrainVelocity: MAC slip on the collider, then the closed walls, then
    // the tilt slots (omni_boundary_solver3.h:26-161)
    if (constrain) {
        if (hasSdf_ && nSlip_) {
            int rc = slipMac(pd, err);
            if (rc) return rc;
        }
        k_run<<<nblocks(G_), BS, 0, stream_>>>(BWall{T, lev, mac, closed_}, G_);
        if (hasSdf_ && nLeaf_) k_slip_tilt<<<nblocks(nLeaf_), BS, 0, stream_>>>(T, lev, leaf_.as<int>(), nLeaf_, tilt);
    }
    // velocity ghosts (octree_omni_face_centered_grid3.h:97-120)
    if (nGhost_) {
        k_run<<<nblocks(nGhost_), BS, 0, stream_>>>(BMacGhostClear{ghost_.as<int>(), mac}, nGhost_);
        k_run<<<nblocks(nGhost_), BS, 0, stream_>>>(BMacGhostSample{T, lev, ghost_.as<int>(), mac, 0}, nGhost_);
    }
    if (nHalfMac_)
        k_run<<<nblocks(nHalfMac_), BS, 0, stream_>>>(
            BMacFromHalf{hmG_.as<int>(), hmT_.as<int>(), hmA_.as<int>(), hmS_.as<int>(), tilt, mac}, nHalfMac_);
    if (nGhost_) {
        k_run<<<nblocks(nGhost_), BS, 0, stream_>>>(BMacGhostSample{T, lev, ghost_.as<int>(), mac, 1}, nGhost_);
        k_run<<<nblocks(nGhost_), BS, 0, stream_>>>(
            BTiltGhost{T, lev, ghost_.as<int>(), static_cast<const uint8_t*>(inHalf_.p), tilt}, nGhost_);
    }
    HIPCHK(hipGetLastError());
    return ASTG_OK;
}

int AdvectDevice::postProject(ProjectDevice& pd, std::string& err, bool constrain) {
    std::lock_guard<std::recursive_mutex> lk_(constMutex());
    if (int rc_ = bindConstants(false, err)) return rc_;
    const Topo T = topo();
    const uint8_t* lev = static_cast<const uint8_t*>(level_.p);
    double* mac = pd.uMac();
    double* tilt = pd.uTilt();
    HIPCHK(hipMemsetAsync(bad_.p, 0, 4, stream_));
    int rc = constrainAndGhost(pd, err, constrain);
    if (rc) return rc;
    // dual grid
    if (nLeaf_) {
        k_run<<<nblocks(nLeaf_), BS, 0, stream_>>>(BDualTilt{T, leaf_.as<int>(), tilt, dualTilt_.as<double>()},
                                                    nLeaf_);
        k_run<<<nblocks(nLeaf_), BS, 0, stream_>>>(
            BDualOct{T, lev, leaf_.as<int>(), mac, tilt, dualOct_.as<double>()}, nLeaf_);
    }
    if (nHalf_)
        k_run<<<nblocks(nHalf_), BS, 0, stream_>>>(BDualHalf{T, half_.as<int>(), tilt, dualTilt_.as<double>()},
                                                    nHalf_);
    // constrainVelocity(dual grid): the vertex vectors inside the collider
    if (constrain && hasSdf_ && nLeaf_)
        k_slip_dual<<<nblocks(nLeaf_), BS, 0, stream_>>>(T, lev, leaf_.as<int>(), nLeaf_, dualTilt_.as<double>());
    HIPCHK(hipGetLastError());
    rc = omniGhost(dualOct_.as<double>(), dualTilt_.as<double>(), 3, err);
    if (rc) return rc;
    return checkBad(err);
}

int AdvectDevice::slipMac(ProjectDevice& pd, std::string& err) {
    const Topo T = topo();
    const uint8_t* lev = static_cast<const uint8_t*>(level_.p);
    double* mac = pd.uMac();
    const int* ord = slipOrd_.as<int>();
    const int* fg = slipFaceG_.as<int>();
    const uint8_t* fa = static_cast<const uint8_t*>(slipFaceA_.p);
    double* pub = slipPub_.as<double>();
    int* e = slipErr_.as<int>();
    int nf = nSlip_;
    k_fill_sent64<<<nblocks(nf), BS, 0, stream_>>>(pub, nf);
    HIPCHK(hipMemsetAsync(e, 0, 4, stream_));
    // every block must be resident at once (faces wait on earlier faces)
    void* args[] = {(void*)&T, (void*)&lev, (void*)&mac, (void*)&ord, (void*)&fg, (void*)&fa, (void*)&nf, (void*)&pub,
                    (void*)&e};
    HIPCHK(hipLaunchCooperativeKernel((const void*)k_slip_mac, dim3(slipGrid_), dim3(BS), args, 0, stream_));
    k_slip_scatter<<<nblocks(nf), BS, 0, stream_>>>(fg, fa, nf, pub, mac);
    HIPCHK(hipGetLastError());
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, e, 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    if (bad) {
        err = "slip constraint: a face waited past its spin limit";
        return ASTG_ERR_HIP;
    }
    return ASTG_OK;
}

int AdvectDevice::wipeBoundary(std::string& err) {
    std::lock_guard<std::recursive_mutex> lk_(constMutex());
    if (int rc_ = bindConstants(false, err)) return rc_;
    if (!hasSdf_) return ASTG_OK;
    const Topo T = topo();
    const uint8_t* lev = static_cast<const uint8_t*>(level_.p);
    double* dC = scal_[0][cur_[0]][0].as<double>();
    double* dV = scal_[0][cur_[0]][1].as<double>();
    double* hC = scal_[1][cur_[1]][0].as<double>();
    double* hV = scal_[1][cur_[1]][1].as<double>();
    if (nLeaf_) k_wipe<<<nblocks(nLeaf_), BS, 0, stream_>>>(T, lev, leaf_.as<int>(), nLeaf_, 0, dC, dV, hC, hV);
    if (nHalf_) k_wipe<<<nblocks(nHalf_), BS, 0, stream_>>>(T, lev, half_.as<int>(), nHalf_, 1, dC, dV, hC, hV);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

int AdvectDevice::checkBad(std::string& err) {
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, bad_.p, 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    if (bad) {
        err = "a tilt closed by a half-tilt is open or has a half direction";
        return ASTG_ERR_TOPOLOGY;
    }
    return ASTG_OK;
}

int AdvectDevice::scalarGhost(int field, std::string& err) {
    std::lock_guard<std::recursive_mutex> lk_(constMutex());
    if (int rc_ = bindConstants(false, err)) return rc_;
    if (field < 0 || field > 1) {
        err = "bad field";
        return ASTG_ERR_ARG;
    }
    HIPCHK(hipMemsetAsync(bad_.p, 0, 4, stream_));
    int rc = omniGhost(scal_[field][cur_[field]][0].as<double>(), scal_[field][cur_[field]][1].as<double>(), 1, err);
    if (rc) return rc;
    return checkBad(err);
}

AdvVel AdvectDevice::vel(ProjectDevice& pd) const {
    AdvVel V;
    V.mac = pd.uMac();
    V.tilt = pd.uTilt();
    V.doct = dualOct_.as<double>();
    V.dtilt = dualTilt_.as<double>();
    return V;
}

int AdvectDevice::advect(ProjectDevice& pd, double dt, std::string& err) {
    std::lock_guard<std::recursive_mutex> lk_(constMutex());
    if (int rc_ = bindConstants(false, err)) return rc_;
    const Topo T = sampTopo();  // sources: the tilt state they were written with
    const uint8_t* lev = static_cast<const uint8_t*>(level_.p);
    const Vel V = vel(pd);
    const int d = cur_[0], h = cur_[1];
    double* d1c = scal_[0][1 - d][0].as<double>();
    double* d1v = scal_[0][1 - d][1].as<double>();
    double* h1c = scal_[1][1 - h][0].as<double>();
    double* h1v = scal_[1][1 - h][1].as<double>();
    const double* d0c = scal_[0][d][0].as<double>();
    const double* d0v = scal_[0][d][1].as<double>();
    const double* h0c = scal_[1][h][0].as<double>();
    const double* h0v = scal_[1][h][1].as<double>();
    if (nLeaf_)
        k_adv_scalar<<<nblocks(nLeaf_), BS, 0, stream_>>>(T, lev, V, leaf_.as<int>(), nLeaf_, 0, dt, d0c, d0v, h0c,
                                                          h0v, d1c, h1c);
    if (nOpenTilt_)
        k_adv_scalar<<<nblocks(nOpenTilt_), BS, 0, stream_>>>(T, lev, V, openTilt_.as<int>(), nOpenTilt_, 1, dt, d0c,
                                                              d0v, h0c, h0v, d1v, h1v);
    if (nHalf_)
        k_adv_scalar<<<nblocks(nHalf_), BS, 0, stream_>>>(T, lev, V, half_.as<int>(), nHalf_, 1, dt, d0c, d0v, h0c,
                                                          h0v, d1v, h1v);
    HIPCHK(hipGetLastError());
    cur_[0] = 1 - d;
    cur_[1] = 1 - h;
    int rc = scalarGhost(0, err);
    if (rc) return rc;
    rc = scalarGhost(1, err);
    if (rc) return rc;
    // velocity (advection.h:195-239) into the temporaries, then swap
    if (nLeaf_)
        k_adv_mac<<<nblocks(3L * nLeaf_), BS, 0, stream_>>>(T, lev, V, leaf_.as<int>(), nLeaf_, dt, pd.tMac());
    if (nVelTilt_)
        k_adv_tilt<<<nblocks(8L * nVelTilt_), BS, 0, stream_>>>(T, te_.as<double>(), lev, V, velTilt_.as<int>(),
                                                                 nVelTilt_, dt, pd.tTilt());
    if (nHalf_)
        k_adv_half<<<nblocks(5L * nHalf_), BS, 0, stream_>>>(T, te_.as<double>(), lev, V, half_.as<int>(), nHalf_, dt,
                                                         pd.tTilt());
    HIPCHK(hipGetLastError());
    pd.swapVelocity();
    // decay (density->scale, heat->scale)
    const long G = G_;
    k_scale2<<<nblocks(G), BS, 0, stream_>>>(G, 0.999, scal_[0][cur_[0]][0].as<double>(),
                                             scal_[0][cur_[0]][1].as<double>());
    k_scale2<<<nblocks(G), BS, 0, stream_>>>(G, 0.999, scal_[1][cur_[1]][0].as<double>(),
                                             scal_[1][cur_[1]][1].as<double>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    sampValid_ = false;  // the fields now carry the current tilt state
    return ASTG_OK;
}

// ---- differentiable step --------------------------------------------------
int AdvectDevice::applyList(std::vector<PassRecord>& recs, std::string& err) {
    for (auto it = recs.rbegin(); it != recs.rend(); ++it) {
        int rc = applyTransposed(*it, work_.as<double>(), wtmp_.as<double>(), stream_, err);
        if (rc) return rc;
    }
    return ASTG_OK;
}

int AdvectDevice::diffPrepare(ProjectDevice& pd, PcgDevice& pcg, double dt, double rho, std::string& err) {
    std::lock_guard<std::recursive_mutex> lk_(constMutex());
    if (int rc_ = bindConstants(false, err)) return rc_;
    if (hasSdf_) {
        err = "the differentiable step has no form with a collider";
        return ASTG_ERR_STATE;
    }
    const long long G = G_, n = pcg.n();
    double* dC = scal_[0][cur_[0]][0].as<double>();
    double* dV = scal_[0][cur_[0]][1].as<double>();
    double* doct = dualOct_.as<double>();
    double* dtil = dualTilt_.as<double>();
    StateMap M;
    const double* bases[8] = {pd.uMac(), pd.uTilt(), doct, dtil, dC, dV, pcg.rhs(), pcg.solution()};
    const long long lens[8] = {3 * G, 8 * G, 3 * G, 3 * G, G, G, n, n};
    long long off = 0;
    M.n = 8;
    for (int r = 0; r < 8; ++r) {
        M.base[r] = bases[r];
        M.len[r] = lens[r];
        M.off[r] = off;
        off += lens[r];
    }
    if (off >= (1LL << 31)) {
        err = "differentiable step: state exceeds 32-bit indices";
        return ASTG_ERR_ARG;
    }
    offB_ = M.off[6];
    offP_ = M.off[7];
    workLen_ = off;
    HIPCHK(work_.alloc((size_t)workLen_ * 8));
    HIPCHK(wtmp_.alloc((size_t)workLen_ * 8));
    HIPCHK(wadv_.alloc((size_t)13 * G * 8));
    const Topo T = topo();
    const uint8_t* lev = static_cast<const uint8_t*>(level_.p);
    const int* gl = ghost_.as<int>();
    const uint8_t* ih = static_cast<const uint8_t*>(inHalf_.p);
    double* mac = pd.uMac();
    double* tilt = pd.uTilt();
    int rc;
    auto rec = [&](auto body, int cnt, std::vector<PassRecord>& v) {
        v.emplace_back();
        return recordPass(body, cnt, M, workLen_, recClaim_, recScratch_, stream_, v.back(), err);
    };
    recCG_.clear();
    recDual_.clear();
    recG3_.clear();
    recG1_.clear();
    recCG_.reserve(6);
    recDual_.reserve(3);
    recG3_.reserve(6);
    recG1_.reserve(6);
    if ((rc = rec(BWall{T, lev, mac, closed_}, G_, recCG_)) ||
        (rc = rec(BMacGhostClear{gl, mac}, nGhost_, recCG_)) ||
        (rc = rec(BMacGhostSample{T, lev, gl, mac, 0}, nGhost_, recCG_)) ||
        (rc = rec(BMacFromHalf{hmG_.as<int>(), hmT_.as<int>(), hmA_.as<int>(), hmS_.as<int>(), tilt, mac}, nHalfMac_,
                  recCG_)) ||
        (rc = rec(BMacGhostSample{T, lev, gl, mac, 1}, nGhost_, recCG_)) ||
        (rc = rec(BTiltGhost{T, lev, gl, ih, tilt}, nGhost_, recCG_)))
        return rc;
    if ((rc = rec(BDualTilt{T, leaf_.as<int>(), tilt, dtil}, nLeaf_, recDual_)) ||
        (rc = rec(BDualOct{T, lev, leaf_.as<int>(), mac, tilt, doct}, nLeaf_, recDual_)) ||
        (rc = rec(BDualHalf{T, half_.as<int>(), tilt, dtil}, nHalf_, recDual_)))
        return rc;
    if ((rc = rec(BCellGhostFine<3>{T, lev, gl, doct}, nGhost_, recG3_)) ||
        (rc = rec(BCellGhostCoarse<3>{T, lev, gl, doct}, nGhost_, recG3_)) ||
        (rc = rec(BVertGhost<3>{T, lev, gl, ih, dtil}, nGhost_, recG3_)) ||
        (rc = rec(BCellGhostTJ<3>{T, lev, gl, dtil, doct}, nGhost_, recG3_)) ||
        (rc = rec(BClosedJoint<3>{T, lev, tj_.as<int>(), doct, dtil}, nTJ_, recG3_)) ||
        (rc = rec(BClosedByHalf<3>{T, lev, cbh_.as<int>(), doct, dtil, bad_.as<int>()}, nCBH_, recG3_)))
        return rc;
    if ((rc = rec(BCellGhostFine<1>{T, lev, gl, dC}, nGhost_, recG1_)) ||
        (rc = rec(BCellGhostCoarse<1>{T, lev, gl, dC}, nGhost_, recG1_)) ||
        (rc = rec(BVertGhost<1>{T, lev, gl, ih, dV}, nGhost_, recG1_)) ||
        (rc = rec(BCellGhostTJ<1>{T, lev, gl, dV, dC}, nGhost_, recG1_)) ||
        (rc = rec(BClosedJoint<1>{T, lev, tj_.as<int>(), dC, dV}, nTJ_, recG1_)) ||
        (rc = rec(BClosedByHalf<1>{T, lev, cbh_.as<int>(), dC, dV, bad_.as<int>()}, nCBH_, recG1_)))
        return rc;
    if ((rc = pd.recordPasses(M, workLen_, pcg.values(), pcg.rhs(), pcg.devRowstart(), pcg.solution(), rho, dt,
                              recClaim_, recScratch_, recAsm_, recVel_, err)))
        return rc;
    // entries advect writes
    HIPCHK(wMac_.alloc((size_t)3 * G));
    HIPCHK(wTilt_.alloc((size_t)8 * G));
    HIPCHK(wCell_.alloc((size_t)G));
    HIPCHK(wVert_.alloc((size_t)G));
    HIPCHK(hipMemsetAsync(wMac_.p, 0, (size_t)3 * G, stream_));
    HIPCHK(hipMemsetAsync(wTilt_.p, 0, (size_t)8 * G, stream_));
    HIPCHK(hipMemsetAsync(wCell_.p, 0, (size_t)G, stream_));
    HIPCHK(hipMemsetAsync(wVert_.p, 0, (size_t)G, stream_));
    auto mark = [&](const DevBuf& list, int cnt, int width, DevBuf& m) {
        if (cnt) k_mark<<<nblocks((long)width * cnt), BS, 0, stream_>>>(list.as<int>(), cnt, width, G_, m.as<uint8_t>());
    };
    mark(leaf_, nLeaf_, 3, wMac_);
    mark(velTilt_, nVelTilt_, 8, wTilt_);
    mark(half_, nHalf_, 8, wTilt_);
    mark(leaf_, nLeaf_, 1, wCell_);
    mark(openTilt_, nOpenTilt_, 1, wVert_);
    mark(half_, nHalf_, 1, wVert_);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    diffReady_ = true;
    return ASTG_OK;
}

int AdvectDevice::diffGetState(ProjectDevice& pd, double* dst, std::string& err) {
    const size_t G = G_;
    const double* src[6] = {pd.uMac(), pd.uTilt(), dualOct_.as<double>(), dualTilt_.as<double>(),
                            scal_[0][cur_[0]][0].as<double>(), scal_[0][cur_[0]][1].as<double>()};
    const size_t len[6] = {3 * G, 8 * G, 3 * G, 3 * G, G, G};
    size_t off = 0;
    for (int r = 0; r < 6; ++r) {
        HIPCHK(hipMemcpyAsync(dst + off, src[r], len[r] * 8, hipMemcpyDeviceToDevice, stream_));
        off += len[r];
    }
    return ASTG_OK;
}

int AdvectDevice::diffSetState(ProjectDevice& pd, const double* src, std::string& err) {
    const size_t G = G_;
    double* dst[6] = {pd.uMac(), pd.uTilt(), dualOct_.as<double>(), dualTilt_.as<double>(),
                      scal_[0][cur_[0]][0].as<double>(), scal_[0][cur_[0]][1].as<double>()};
    const size_t len[6] = {3 * G, 8 * G, 3 * G, 3 * G, G, G};
    size_t off = 0;
    for (int r = 0; r < 6; ++r) {
        HIPCHK(hipMemcpyAsync(dst[r], src + off, len[r] * 8, hipMemcpyDeviceToDevice, stream_));
        off += len[r];
    }
    sampValid_ = false;
    return ASTG_OK;
}

int AdvectDevice::advectCarry(ProjectDevice& pd, double dt, std::string& err) {
    const size_t b = (size_t)G_ * 8;
    // the destinations start as copies of the sources: entries advect does
    // not write keep their value
    for (int f = 0; f < 2; ++f)
        for (int k = 0; k < 2; ++k)
            HIPCHK(hipMemcpyAsync(scal_[f][1 - cur_[f]][k].p, scal_[f][cur_[f]][k].p, b, hipMemcpyDeviceToDevice,
                                  stream_));
    HIPCHK(hipMemcpyAsync(pd.tMac(), pd.uMac(), 3 * b, hipMemcpyDeviceToDevice, stream_));
    HIPCHK(hipMemcpyAsync(pd.tTilt(), pd.uTilt(), 8 * b, hipMemcpyDeviceToDevice, stream_));
    return advect(pd, dt, err);
}

int AdvectDevice::diffAdjointAdvect(ProjectDevice& pd, double dt, std::string& err) {
    std::lock_guard<std::recursive_mutex> lk_(constMutex());
    if (int rc_ = bindConstants(false, err)) return rc_;
    if (!diffReady_) {
        err = "differentiable step not prepared";
        return ASTG_ERR_STATE;
    }
    const long G = G_;
    double* x = work_.as<double>();
    double* xMac = x;
    double* xTilt = x + 3 * G;
    double* xDoct = x + 11 * G;
    double* xDtilt = x + 14 * G;
    double* xCell = x + 17 * G;
    double* xVert = x + 18 * G;
    double* a = wadv_.as<double>();  // adjoints of the advected values: MAC 3G, tilt 8G, cell G, vertex G
    // the decay, then the density ghosts
    k_scale1<<<nblocks(2 * G), BS, 0, stream_>>>(2 * G, 0.999, xCell);
    int rc = applyList(recG1_, err);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(a, x, (size_t)11 * G * 8, hipMemcpyDeviceToDevice, stream_));
    HIPCHK(hipMemcpyAsync(a + 11 * G, xCell, (size_t)2 * G * 8, hipMemcpyDeviceToDevice, stream_));
    // entries the step carried over pass their adjoint straight through
    k_keep_unwritten<<<nblocks(3 * G), BS, 0, stream_>>>(3 * G, wMac_.as<uint8_t>(), a, xMac);
    k_keep_unwritten<<<nblocks(8 * G), BS, 0, stream_>>>(8 * G, wTilt_.as<uint8_t>(), a + 3 * G, xTilt);
    k_keep_unwritten<<<nblocks(G), BS, 0, stream_>>>(G, wCell_.as<uint8_t>(), a + 11 * G, xCell);
    k_keep_unwritten<<<nblocks(G), BS, 0, stream_>>>(G, wVert_.as<uint8_t>(), a + 12 * G, xVert);
    const Topo T = sampTopo();
    const uint8_t* lev = static_cast<const uint8_t*>(level_.p);
    const Vel V = vel(pd);
    AdjMap A;
    A.n = 3;
    A.src[0] = V.mac;
    A.adj[0] = xMac;
    A.src[1] = V.doct;
    A.adj[1] = xDoct;
    A.src[2] = V.dtilt;
    A.adj[2] = xDtilt;
    const double* d0c = scal_[0][cur_[0]][0].as<double>();
    const double* d0v = scal_[0][cur_[0]][1].as<double>();
    const int BSA = 64;  // heavy lanes: small blocks spread them over the CUs
    auto nbA = [&](long n) { return (int)((n + BSA - 1) / BSA); };
    (void)nbA;
    const uint8_t* ih = static_cast<const uint8_t*>(inHalf_.p);
    if (nLeaf_)
        k_adj_adv_scalar<<<nblocks(nLeaf_), BS, 0, stream_>>>(T, lev, V, A, leaf_.as<int>(), nLeaf_, 0, dt, d0c, d0v,
                                                              a + 11 * G, xCell, xVert, nullptr);
    if (nOpenTilt_)
        k_adj_adv_scalar<<<nblocks(nOpenTilt_), BS, 0, stream_>>>(T, lev, V, A, openTilt_.as<int>(), nOpenTilt_, 1,
                                                                  dt, d0c, d0v, a + 12 * G, xCell, xVert, ih);
    if (nHalf_)
        k_adj_adv_scalar<<<nblocks(nHalf_), BS, 0, stream_>>>(T, lev, V, A, half_.as<int>(), nHalf_, 1, dt, d0c, d0v,
                                                              a + 12 * G, xCell, xVert, nullptr);
    if (nLeaf_)
        k_adj_adv_mac<<<nblocks(3L * nLeaf_), BS, 0, stream_>>>(T, te_.as<double>(), lev, V, A, leaf_.as<int>(),
                                                                nLeaf_, dt, a);
    if (nVelTilt_)
        k_adj_adv_tilt<<<nblocks(8L * nVelTilt_), BS, 0, stream_>>>(T, te_.as<double>(), lev, V, A,
                                                                     velTilt_.as<int>(), nVelTilt_, dt, a + 3 * G, ih);
    if (nHalf_)
        k_adj_adv_half<<<nblocks(5L * nHalf_), BS, 0, stream_>>>(T, te_.as<double>(), lev, V, A, half_.as<int>(),
                                                                 nHalf_, dt, a + 3 * G);
    HIPCHK(hipGetLastError());
    return ASTG_OK;
}

int AdvectDevice::diffAdjointPost(std::string& err) {
    int rc = applyList(recG3_, err);
    if (rc) return rc;
    rc = applyList(recDual_, err);
    if (rc) return rc;
    return applyList(recCG_, err);
}

int AdvectDevice::diffAdjointPre(std::string& err) { return applyList(recCG_, err); }

int AdvectDevice::mirrorTemporaries(ProjectDevice& pd, const bool scalars[2], bool velocity, std::string& err) {
    const size_t b = (size_t)G_ * 8;
    for (int f = 0; f < 2; ++f) {
        if (!scalars[f]) continue;
        const int c = cur_[f];
        for (int k = 0; k < 2; ++k)
            HIPCHK(hipMemcpyAsync(scal_[f][1 - c][k].p, scal_[f][c][k].p, b, hipMemcpyDeviceToDevice, stream_));
    }
    if (velocity) {
        HIPCHK(hipMemcpyAsync(pd.tMac(), pd.uMac(), 3 * b, hipMemcpyDeviceToDevice, stream_));
        HIPCHK(hipMemcpyAsync(pd.tTilt(), pd.uTilt(), 8 * b, hipMemcpyDeviceToDevice, stream_));
    }
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

int AdvectDevice::applySource(ProjectDevice& pd, double cx, double cy, double cz, double r, double dval, double hval,
                              double speed, bool ghost, std::string& err) {
    return applySourceV(pd, cx, cy, cz, r, dval, hval, nullptr, speed, ghost, err);
}

int AdvectDevice::applySourceV(ProjectDevice& pd, double cx, double cy, double cz, double r, double dval,
                               double hval, const double* v, double speed, bool ghost, std::string& err) {
    std::lock_guard<std::recursive_mutex> lk_(constMutex());
    if (int rc_ = bindConstants(false, err)) return rc_;
    const Topo T = topo();
    const uint8_t* lev = static_cast<const uint8_t*>(level_.p);
    double* dC = scal_[0][cur_[0]][0].as<double>();
    double* dV = scal_[0][cur_[0]][1].as<double>();
    double* hC = scal_[1][cur_[1]][0].as<double>();
    double* hV = scal_[1][cur_[1]][1].as<double>();
    const double vx = v ? v[0] : 0.0, vy = v ? v[1] : 0.0, vz = v ? v[2] : speed;
    if (nLeaf_) {
        if (v)
            k_source<true><<<nblocks(nLeaf_), BS, 0, stream_>>>(T, lev, leaf_.as<int>(), nLeaf_, cx, cy, cz, r, dval,
                                                                hval, vx, vy, vz, dC, dV, hC, hV, pd.uMac(),
                                                                pd.uTilt());
        else
            k_source<false><<<nblocks(nLeaf_), BS, 0, stream_>>>(T, lev, leaf_.as<int>(), nLeaf_, cx, cy, cz, r, dval,
                                                                 hval, vx, vy, vz, dC, dV, hC, hV, pd.uMac(),
                                                                 pd.uTilt());
    }
    if (nHalf_) {
        if (v)
            k_source_half<true><<<nblocks(nHalf_), BS, 0, stream_>>>(T, lev, half_.as<int>(), nHalf_, cx, cy, cz, r,
                                                                     dval, hval, vx, vy, vz, dV, hV, pd.uTilt());
        else
            k_source_half<false><<<nblocks(nHalf_), BS, 0, stream_>>>(T, lev, half_.as<int>(), nHalf_, cx, cy, cz, r,
                                                                      dval, hval, vx, vy, vz, dV, hV, pd.uTilt());
    }
    HIPCHK(hipGetLastError());
    if (ghost) {
        int rc = scalarGhost(0, err);
        if (rc) return rc;
        rc = scalarGhost(1, err);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

int AdvectDevice::addForce(ProjectDevice& pd, double dt, double fd, double ft, std::string& err) {
    std::lock_guard<std::recursive_mutex> lk_(constMutex());
    if (int rc_ = bindConstants(false, err)) return rc_;
    const Topo T = topo();
    const uint8_t* lev = static_cast<const uint8_t*>(level_.p);
    const double* dC = scal_[0][cur_[0]][0].as<double>();
    const double* dV = scal_[0][cur_[0]][1].as<double>();
    const double* hC = scal_[1][cur_[1]][0].as<double>();
    const double* hV = scal_[1][cur_[1]][1].as<double>();
    // ambient temperature: mean of the cell heat over every node of every
    // level, nodes of levels above 0 counted 8 times (as addForce weighs them)
    const int nb = 64;
    HIPCHK(out_.alloc((size_t)L_ * nb * 8));
    for (int l = 0; l < L_; ++l)
        k_sum_range<<<nb, BS, 0, stream_>>>(hC, off_[l], nnode_[l], out_.as<double>() + (size_t)l * nb);
    std::vector<double> part((size_t)L_ * nb);
    HIPCHK(hipMemcpyAsync(part.data(), out_.p, part.size() * 8, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    double sum = 0.0, count = 0.0;
    for (int l = 0; l < L_; ++l) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += part[(size_t)l * nb + b];
        const double w = l == 0 ? 1.0 : 8.0;
        sum += s * w;
        count += (double)nnode_[l] * w;
    }
    const double ambT = count > 0 ? sum / count : 0.0;
    if (nLeaf_)
        k_force_mac<<<nblocks(nLeaf_), BS, 0, stream_>>>(T, lev, leaf_.as<int>(), nLeaf_, dC, dV, hC, hV, fd, ft, ambT,
                                                         dt, pd.uMac());
    if (nOpenTilt_)
        k_force_tilt<<<nblocks(8L * nOpenTilt_), BS, 0, stream_>>>(T, lev, openTilt_.as<int>(), nOpenTilt_, dC, dV, hC,
                                                                   hV, fd, ft, ambT, dt, pd.uTilt());
    if (nHalf_)
        k_force_half<<<nblocks(5L * nHalf_), BS, 0, stream_>>>(T, lev, half_.as<int>(), nHalf_, dC, dV, hC, hV, fd, ft, ambT,
                                                          dt, pd.uTilt());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

int AdvectDevice::samplePoints(ProjectDevice& pd, int what, int n, const double* xyz, double dt, double* out,
                               std::string& err) {
    std::lock_guard<std::recursive_mutex> lk_(constMutex());
    if (int rc_ = bindConstants(false, err)) return rc_;
    if (what < 0 || what > 5 || n < 0) {
        err = "bad query";
        return ASTG_ERR_ARG;
    }
    if (n == 0) return ASTG_OK;
    const int w = what <= 2 ? 3 : 1;
    HIPCHK(pts_.alloc((size_t)n * 24));
    HIPCHK(out_.alloc((size_t)n * 8 * w));
    HIPCHK(hipMemcpyAsync(pts_.p, xyz, (size_t)n * 24, hipMemcpyHostToDevice, stream_));
    const int f = what == 4 ? 1 : 0;
    k_points<<<nblocks(n), BS, 0, stream_>>>(topo(), vel(pd), scal_[f][cur_[f]][0].as<double>(),
                                             scal_[f][cur_[f]][1].as<double>(), what, n, pts_.as<double>(), dt,
                                             out_.as<double>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, out_.p, (size_t)n * 8 * w, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

double AdvectDevice::advectBytes() const {
    // compulsory traffic: each source field and the velocity read once, each
    // destination written once, plus the topology arrays
    const double g = G_;
    return g * (2 * 2 * 8 + 2 * 2 * 8) + g * (3 + 8) * 8 * 2 + g * 6 * 8 + g * (2 * 3 + 1 + 8 + 1 + 1);
}

}  // namespace astg
