// Uniform-layout AST projection on the device: numbering, pattern, assembly,
// velocity correction and wall constraint (the reference's
// ast::ASTSolver<d>::project, include/ast_solver.h:322-498, with
// CORRECTION_MAC and no edge-node unknowns, ast_general.h:11-15), the solve
// through PcgDevice.
//
// Every kernel takes one lane per row, face or tilt node. The arithmetic keeps
// the reference's operation order with no contraction, so the system matches
// a strict-IEEE restatement bit for bit.
#include "ast_uniform.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <cstring>

#include "pcg_common.h"

#pragma clang fp contract(off)

namespace astg {

namespace {

inline int nblk(long long n) { return (int)((n + BS - 1) / BS); }

// Corner tilts of a cell (constants.h:71-73; ast_constant.h:62-65), the cells
// of a tilt (constants.h:75-77; ast_constant.h:67-70), and the MAC faces
// (mac_direction: +x -x +y -y +z -z) each corner tilt cuts
// (ast_constant.h:38-56).
template <int D>
__device__ inline int octNbTilt(int k, int a) {
    constexpr int t2[4][2] = {{1, 1}, {1, 0}, {0, 0}, {0, 1}};
    constexpr int t3[8][3] = {{1, 1, 1}, {1, 0, 1}, {0, 0, 1}, {0, 1, 1}, {1, 1, 0}, {1, 0, 0}, {0, 0, 0}, {0, 1, 0}};
    return D == 2 ? t2[k][a] : t3[k][a];
}
template <int D>
__device__ inline int tiltNbOct(int k, int a) {
    constexpr int t2[4][2] = {{0, 0}, {0, -1}, {-1, -1}, {-1, 0}};
    constexpr int t3[8][3] = {{0, 0, 0},   {0, -1, 0},   {-1, -1, 0},  {-1, 0, 0},
                              {0, 0, -1}, {0, -1, -1}, {-1, -1, -1}, {-1, 0, -1}};
    return D == 2 ? t2[k][a] : t3[k][a];
}
template <int D>
__device__ inline int tiltNbMac(int k, int i) {
    constexpr int t2[4][2] = {{0, 2}, {0, 3}, {1, 3}, {1, 2}};
    constexpr int t3[8][3] = {{0, 2, 4}, {0, 3, 4}, {1, 3, 4}, {1, 2, 4}, {0, 2, 5}, {0, 3, 5}, {1, 3, 5}, {1, 2, 5}};
    return D == 2 ? t2[k][i] : t3[k][i];
}
// oppositeTiltIdx (ast_util.h:27-30)
template <int D>
__device__ inline int oppositeTilt(int k) {
    return D == 2 ? (k + 2) % 4 : (((k & 1) ? 8 : 6) - k);
}
// Slots in increasing column order: the corner tilts of a cell and the cells
// of a tilt both ascend in box order (z, then y, then x) in this order.
template <int D>
__device__ inline int sortedSlot(int j) {
    constexpr int s2[4] = {2, 1, 3, 0};
    constexpr int s3[8] = {6, 5, 7, 4, 2, 1, 3, 0};
    return D == 2 ? s2[j] : s3[j];
}

__device__ inline long long tiltLin(const AstGeom& g, int x, int y, int z) {
    return x + (long long)(g.res[0] + 1) * (y + (long long)(g.res[1] + 1) * z);
}
__device__ inline int cellLin(const AstGeom& g, int x, int y, int z) {
    return x + g.res[0] * (y + g.res[1] * z);
}
__device__ inline long long macLin(const AstGeom& g, int ax, int x, int y, int z) {
    const long long mx = g.res[0] + (ax == 0), my = g.res[1] + (ax == 1);
    return x + mx * (y + my * z);
}

// Boundary tilt nodes stay closed (initData 79-87).
__global__ void k_ast_flags(AstGeom g, const uint8_t* open, int* flag) {
    const long long t = (long long)blockIdx.x * BS + threadIdx.x;
    if (t >= g.nt) return;
    const long long tx = g.res[0] + 1, ty = g.res[1] + 1;
    const int x = (int)(t % tx), y = (int)((t / tx) % ty), z = (int)(t / (tx * ty));
    bool in = x > 0 && x < g.res[0] && y > 0 && y < g.res[1];
    if (g.d == 3) in = in && z > 0 && z < g.res[2];
    flag[t] = (open[t] && in) ? 1 : 0;
}

__global__ void k_ast_number(AstGeom g, const int* flag, const int* scan, int* tid, int* tnode) {
    const long long t = (long long)blockIdx.x * BS + threadIdx.x;
    if (t >= g.nt) return;
    if (flag[t]) {
        tid[t] = g.nc + scan[t];
        tnode[scan[t]] = (int)t;
    } else {
        tid[t] = -1;
    }
}

// One row of the system in increasing column order. Cell rows follow
// project 327-402 (corner tilts, then the 2d faces, then the diagonal), tilt
// rows 403-428; the right-hand side is returned unscaled.
template <int D>
struct Row {
    int cnt = 0;
    int col[D == 2 ? 9 : 15];
    double val[D == 2 ? 9 : 15];
    double b = 0.0;
};

template <int D>
__device__ inline void cellRow(const AstGeom& g, int c, const int* tid, const double* const* mac, const double* tu,
                               Row<D>& R) {
    constexpr int kT = D == 2 ? 4 : 8;
    const int x = c % g.res[0], y = (c / g.res[0]) % g.res[1], z = D == 3 ? c / (g.res[0] * g.res[1]) : 0;
    const int q[3] = {x, y, z};
    double area[2 * D];
#pragma unroll
    for (int k = 0; k < 2 * D; ++k) area[k] = 0.0;
    double center = 0.0, tb = 0.0;
    int tcol[kT];
    const double tcoef = g.tarea / g.tod;
#pragma unroll
    for (int k = 0; k < kT; ++k) {
        const long long t = tiltLin(g, x + octNbTilt<D>(k, 0), y + octNbTilt<D>(k, 1),
                                    D == 3 ? z + octNbTilt<D>(k, 2) : 0);
        const int id = tid[t];
        tcol[k] = id;
        if (id < 0) continue;
#pragma unroll
        for (int i = 0; i < D; ++i) area[tiltNbMac<D>(k, i)] -= g.cut;
        center += tcoef;
        const double sign = k <= (D == 2 ? 1 : 3) ? -1.0 : 1.0;
        tb += sign * g.tarea * tu[t * kT + oppositeTilt<D>(k)];
    }
    bool fopen[2 * D];
    int fcol[2 * D];
    double fcoef[2 * D];
#pragma unroll
    for (int k = 0; k < 2 * D; ++k) {
        area[k] += g.full;
        const int ax = k >> 1;
        const bool up = (k & 1) == 0;
        const int nq = q[ax] + (up ? 1 : -1);
        const bool valid = nq >= 0 && nq < g.res[ax];
        fopen[k] = false;
        if (!(area[k] > DBL_EPSILON)) continue;
        const int f[3] = {x + (ax == 0 && up), y + (ax == 1 && up), z + (ax == 2 && up)};
        const double u = mac[ax][macLin(g, ax, f[0], f[1], f[2])];
        const double s = up ? -1.0 : 1.0;
        const double coeff = area[k] / g.dx;
        if (valid) {
            fopen[k] = true;
            fcoef[k] = coeff;
            fcol[k] = c + (up ? 1 : -1) * (ax == 0 ? 1 : ax == 1 ? g.res[0] : g.res[0] * g.res[1]);
            center += coeff;
            tb += s * area[k] * u;
        } else if (ax == 1) {  // the y walls are Dirichlet faces (project 389-395)
            center += coeff;
            tb += s * area[k] * u;
        }
    }
    if (center < DBL_EPSILON) center = 1.0;  // a zero diagonal (project 397-400)
    int n = 0;
    // faces below the diagonal: -z, -y, -x; then the diagonal; then +x, +y, +z
#pragma unroll
    for (int j = 2 * D - 1; j >= 1; j -= 2)
        if (fopen[j]) {
            R.col[n] = fcol[j];
            R.val[n++] = -fcoef[j];
        }
    R.col[n] = c;
    R.val[n++] = center;
#pragma unroll
    for (int j = 0; j < 2 * D; j += 2)
        if (fopen[j]) {
            R.col[n] = fcol[j];
            R.val[n++] = -fcoef[j];
        }
#pragma unroll
    for (int j = 0; j < kT; ++j) {
        const int k = sortedSlot<D>(j);
        if (tcol[k] >= 0) {
            R.col[n] = tcol[k];
            R.val[n++] = -tcoef;
        }
    }
    R.cnt = n;
    R.b = tb;
}

template <int D>
__device__ inline void tiltRow(const AstGeom& g, int r, const int* tnode, const double* tu, Row<D>& R) {
    constexpr int kT = D == 2 ? 4 : 8;
    const long long t = tnode[r - g.nc];
    const long long tx = g.res[0] + 1, ty = g.res[1] + 1;
    const int x = (int)(t % tx), y = (int)((t / tx) % ty), z = D == 3 ? (int)(t / (tx * ty)) : 0;
    const double coeff = g.tarea / g.tod;
    double center = 0.0, tb = 0.0;
#pragma unroll
    for (int k = 0; k < kT; ++k) {
        center += coeff;
        const double sign = k <= (D == 2 ? 1 : 3) ? -1.0 : 1.0;
        tb += sign * g.tarea * tu[t * kT + k];
    }
#pragma unroll
    for (int j = 0; j < kT; ++j) {
        const int k = sortedSlot<D>(j);
        R.col[j] = cellLin(g, x + tiltNbOct<D>(k, 0), y + tiltNbOct<D>(k, 1), D == 3 ? z + tiltNbOct<D>(k, 2) : 0);
        R.val[j] = -coeff;
    }
    R.col[kT] = r;
    R.val[kT] = center;
    R.cnt = kT + 1;
    R.b = tb;
}

struct MacPtrs {
    const double* p[3];
};

template <int D>
__global__ void k_ast_len(AstGeom g, int n, const int* tid, MacPtrs mac, const double* tu, int* len) {
    const int r = blockIdx.x * BS + threadIdx.x;
    if (r > n) return;
    if (r == n) {
        len[n] = 0;
        return;
    }
    if (r >= g.nc) {
        len[r] = (D == 2 ? 4 : 8) + 1;
        return;
    }
    Row<D> R;
    cellRow<D>(g, r, tid, mac.p, tu, R);
    len[r] = R.cnt;
}

// VALUES = false: columns into col; true: values into val and the right-hand
// side, scaled by rho / dt (project 471), into b.
template <int D, bool VALUES>
__global__ void k_ast_fill(AstGeom g, int n, const int* tid, const int* tnode, MacPtrs mac, const double* tu,
                           const int* rs, int* col, double* val, double* b, double scale) {
    const int r = blockIdx.x * BS + threadIdx.x;
    if (r >= n) return;
    Row<D> R;
    if (r < g.nc)
        cellRow<D>(g, r, tid, mac.p, tu, R);
    else
        tiltRow<D>(g, r, tnode, tu, R);
    const int p0 = rs[r];
    if (VALUES) {
        for (int j = 0; j < R.cnt; ++j) val[p0 + j] = R.val[j];
        b[r] = R.b * scale;
    } else {
        for (int j = 0; j < R.cnt; ++j) col[p0 + j] = R.col[j];
    }
}

// Largest |b| as ordered bits (non-negative doubles order as integers): a
// grid-stride pass, one block maximum and one atomic per block.
constexpr int kMaxBlocks = 1024;
__global__ void k_ast_absmax(int n, const double* b, unsigned long long* out) {
    double v = 0.0;
    for (long long i = (long long)blockIdx.x * BS + threadIdx.x; i < n; i += (long long)gridDim.x * BS)
        v = nan_max(v, fabs(b[i]));
    v = block_reduce<true>(v);
    if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(v));
}
__global__ void k_ast_zero_u64(unsigned long long* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}
__global__ void k_ast_zero(long long n, double* p) {
    const long long i = (long long)blockIdx.x * BS + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Block-colour sort key of every row: colour * blocks + block, the block of
// a cell row its own, of a tilt row that of its lower corner cell (tilt
// rows are interior nodes, so that cell exists).
template <int D>
__global__ void k_ast_block_keys(AstGeom g, int n, const int* tnode, int B, int* keys) {
    const int r = blockIdx.x * BS + threadIdx.x;
    if (r >= n) return;
    int x, y, z;
    if (r < g.nc) {
        x = r % g.res[0];
        y = (r / g.res[0]) % g.res[1];
        z = D == 3 ? r / (g.res[0] * g.res[1]) : 0;
    } else {
        const long long t = tnode[r - g.nc];
        const long long tx = g.res[0] + 1, ty = g.res[1] + 1;
        x = (int)(t % tx) - 1;
        y = (int)((t / tx) % ty) - 1;
        z = D == 3 ? (int)(t / (tx * ty)) - 1 : 0;
    }
    const int bx = x / B, by = y / B, bz = z / B;
    const int nbx = (g.res[0] + B - 1) / B, nby = (g.res[1] + B - 1) / B, nbz = D == 3 ? (g.res[2] + B - 1) / B : 1;
    const int colour = (bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2);
    keys[r] = colour * (nbx * nby * nbz) + bx + nbx * (by + nby * bz);
}

// correctVelocityByPressure, tilt part (ast_solver.h:286-296).
template <int D>
__global__ void k_ast_correct_tilt(AstGeom g, int nOpen, const int* tnode, const double* p, double* tu, double dt,
                                   double rho) {
    constexpr int kT = D == 2 ? 4 : 8;
    const int j = blockIdx.x * BS + threadIdx.x;
    if (j >= nOpen) return;
    const long long t = tnode[j];
    const long long tx = g.res[0] + 1, ty = g.res[1] + 1;
    const int x = (int)(t % tx), y = (int)((t / tx) % ty), z = D == 3 ? (int)(t / (tx * ty)) : 0;
    const double pc = p[g.nc + j];
#pragma unroll
    for (int k = 0; k < kT; ++k) {
        const int nb = cellLin(g, x + tiltNbOct<D>(k, 0), y + tiltNbOct<D>(k, 1), D == 3 ? z + tiltNbOct<D>(k, 2) : 0);
        const double s = k < kT / 2 ? -1.0 : 1.0;
        tu[t * kT + k] += s * dt / rho * (p[nb] - pc) / g.tod;
    }
}

// correctVelocityByPressure, MAC part (ast_solver.h:309-319): every face of
// axis ax, pressure 0 outside the domain.
__global__ void k_ast_correct_mac(AstGeom g, int ax, const double* p, double* u, double dt, double rho) {
    const long long f = (long long)blockIdx.x * BS + threadIdx.x;
    if (f >= g.nm[ax]) return;
    const long long mx = g.res[0] + (ax == 0), my = g.res[1] + (ax == 1);
    int q[3] = {(int)(f % mx), (int)((f / mx) % my), (int)(f / (mx * my))};
    const bool v1 = q[ax] < g.res[ax];
    const bool v0 = q[ax] >= 1;
    const double p1 = v1 ? p[cellLin(g, q[0], q[1], q[2])] : 0.0;
    q[ax] -= 1;
    const double p0 = v0 ? p[cellLin(g, q[0], q[1], q[2])] : 0.0;
    u[f] -= dt / rho * (p1 - p0) / g.dx;
}

// constrainVelocity (ast_solver.h:623-643): the x walls closed; in 3D also
// the y walls, whose loop runs x over [0, res_y) (the caller keeps
// res_y <= res_x, so every such face exists).
__global__ void k_ast_constrain(AstGeom g, double* ux, double* uy) {
    const long long i = (long long)blockIdx.x * BS + threadIdx.x;
    const int nx = g.res[0], ny = g.res[1], nz = g.res[2];
    if (g.d == 2) {
        if (i >= ny) return;
        ux[0 + i * (nx + 1)] = 0.0;
        ux[nx + i * (nx + 1)] = 0.0;
        return;
    }
    if (i >= (long long)ny * nz) return;
    const int a = (int)(i % ny), z = (int)(i / ny);
    // x walls at (0, a, z) and (nx, a, z)
    ux[0 + (long long)(nx + 1) * (a + (long long)ny * z)] = 0.0;
    ux[nx + (long long)(nx + 1) * (a + (long long)ny * z)] = 0.0;
    // y walls at (a, 0, z) and (a, ny, z)
    uy[a + (long long)nx * (0 + (long long)(ny + 1) * z)] = 0.0;
    uy[a + (long long)nx * (ny + (long long)(ny + 1) * z)] = 0.0;
}

}  // namespace

AstUniform::~AstUniform() {
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
}

int AstUniform::init(int dim, const int* res, double dx, std::string& err) {
    ready_ = false;
    if (dim != 2 && dim != 3) {
        err = "dimension must be 2 or 3";
        return ASTG_ERR_ARG;
    }
    if (!(dx > 0)) {
        err = "dx must be positive";
        return ASTG_ERR_ARG;
    }
    AstGeom g;
    g.d = dim;
    long long nc = 1, nt = 1;
    for (int a = 0; a < dim; ++a) {
        if (res[a] < 2) {
            err = "every resolution must be at least 2";
            return ASTG_ERR_ARG;
        }
        g.res[a] = res[a];
        nc *= res[a];
        nt *= res[a] + 1;
    }
    if (dim == 3 && res[1] > res[0]) {
        // constrainVelocity runs x over [0, res_y) on the y walls; past res_x
        // the reference writes outside its face grid
        err = "3D resolution needs res_y <= res_x (the reference's wall constraint)";
        return ASTG_ERR_ARG;
    }
    // unknowns nc + open tilts and the largest pattern must index as int
    const long long kT = dim == 2 ? 4 : 8;
    if (nc + nt >= (1LL << 31) || nc * (2 * dim + kT + 1) + nt * (kT + 1) >= (1LL << 31)) {
        err = "grid too large for 32-bit row and column indices";
        return ASTG_ERR_ARG;
    }
    g.nc = (int)nc;
    g.nt = nt;
    for (int a = 0; a < dim; ++a) {
        long long m = 1;
        for (int b = 0; b < dim; ++b) m *= res[b] + (a == b);
        g.nm[a] = m;
    }
    g.kT = (int)kT;
    g.dx = dx;
    g.half = dx / 2.0;
    g.tod = dim == 2 ? g.half * 1.4142135623730951 : g.half * 1.7320508075688772;
    // SQRT3 * 0.5 * squared(half_dx), left to right as the reference writes it
    g.tarea = dim == 2 ? g.half * 1.4142135623730951 : 1.7320508075688772 * 0.5 * (g.half * g.half);
    g.cut = dim == 2 ? g.half : 0.5 * (g.half * g.half);
    g.full = dim == 2 ? dx : dx * dx;
    g_ = g;
    HIPCHK(open_.alloc((size_t)nt));
    HIPCHK(tid_.alloc((size_t)nt * 4));
    HIPCHK(tnode_.alloc((size_t)nt * 4));
    HIPCHK(flag_.alloc((size_t)nt * 4));
    HIPCHK(scan_.alloc((size_t)nt * 4));
    for (int a = 0; a < dim; ++a) {
        HIPCHK(mac_[a].alloc((size_t)g.nm[a] * 8));
        k_ast_zero<<<nblk(g.nm[a]), BS, 0, stream_>>>(g.nm[a], mac_[a].as<double>());
    }
    HIPCHK(tilt_.alloc((size_t)nt * kT * 8));
    k_ast_zero<<<nblk(nt * kT), BS, 0, stream_>>>(nt * kT, tilt_.as<double>());
    HIPCHK(bmax_.alloc(8));
    HIPCHK(hipMemsetAsync(open_.p, 0, (size_t)nt, stream_));
    for (auto& e : ev_)
        if (!e) HIPCHK(hipEventCreate(&e));
    patternReady_ = false;
    blkReady_ = false;
    solved_ = false;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    ready_ = true;
    return ASTG_OK;
}

int AstUniform::setLayout(const uint8_t* open, std::string& err) {
    HIPCHK(hipMemcpyAsync(open_.p, open, (size_t)g_.nt, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    patternReady_ = false;
    blkReady_ = false;
    solved_ = false;
    return ASTG_OK;
}

int AstUniform::setVelocity(const double* const* mac, const double* tilt, std::string& err) {
    for (int a = 0; a < g_.d; ++a)
        HIPCHK(hipMemcpyAsync(mac_[a].p, mac[a], (size_t)g_.nm[a] * 8, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyAsync(tilt_.p, tilt, (size_t)g_.nt * g_.kT * 8, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

int AstUniform::getVelocity(double* const* mac, double* tilt, std::string& err) {
    for (int a = 0; a < g_.d; ++a)
        HIPCHK(hipMemcpyAsync(mac[a], mac_[a].p, (size_t)g_.nm[a] * 8, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(tilt, tilt_.p, (size_t)g_.nt * g_.kT * 8, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

static int astRefMerge(int dim);

// Numbering (updateIndex 599-621) and the pattern; the solver's pattern set-up.
int AstUniform::ensurePattern(std::string& err) {
    if (patternReady_) return ASTG_OK;
    const long long nt = g_.nt;
    k_ast_flags<<<nblk(nt), BS, 0, stream_>>>(g_, open_.as<uint8_t>(), flag_.as<int>());
    size_t tb = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flag_.as<int>(), scan_.as<int>(), (int)nt, stream_));
    HIPCHK(tmp_.alloc(tb));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp_.p, tb, flag_.as<int>(), scan_.as<int>(), (int)nt, stream_));
    k_ast_number<<<nblk(nt), BS, 0, stream_>>>(g_, flag_.as<int>(), scan_.as<int>(), tid_.as<int>(), tnode_.as<int>());
    int last[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&last[0], scan_.as<int>() + nt - 1, 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(&last[1], flag_.as<int>() + nt - 1, 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    const int n = g_.nc + last[0] + last[1];
    MacPtrs m{{mac_[0].as<double>(), mac_[1].as<double>(), g_.d == 3 ? mac_[2].as<double>() : nullptr}};
    HIPCHK(len_.alloc(((size_t)n + 1) * 4));
    HIPCHK(rs_.alloc(((size_t)n + 1) * 4));
    if (g_.d == 2)
        k_ast_len<2><<<nblk((long long)n + 1), BS, 0, stream_>>>(g_, n, tid_.as<int>(), m, tilt_.as<double>(),
                                                                  len_.as<int>());
    else
        k_ast_len<3><<<nblk((long long)n + 1), BS, 0, stream_>>>(g_, n, tid_.as<int>(), m, tilt_.as<double>(),
                                                                  len_.as<int>());
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, len_.as<int>(), rs_.as<int>(), n + 1, stream_));
    HIPCHK(tmp_.alloc(tb));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp_.p, tb, len_.as<int>(), rs_.as<int>(), n + 1, stream_));
    int nnz = 0;
    HIPCHK(hipMemcpyAsync(&nnz, rs_.as<int>() + n, 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    HIPCHK(col_.alloc((size_t)nnz * 4));
    if (g_.d == 2)
        k_ast_fill<2, false><<<nblk(n), BS, 0, stream_>>>(g_, n, tid_.as<int>(), tnode_.as<int>(), m,
                                                           tilt_.as<double>(), rs_.as<int>(), col_.as<int>(), nullptr,
                                                           nullptr, 0.0);
    else
        k_ast_fill<3, false><<<nblk(n), BS, 0, stream_>>>(g_, n, tid_.as<int>(), tnode_.as<int>(), m,
                                                           tilt_.as<double>(), rs_.as<int>(), col_.as<int>(), nullptr,
                                                           nullptr, 0.0);
    HIPCHK(hipGetLastError());
    pcg_.setMergeLevels(astRefMerge(g_.d));
    int rc = pcg_.setPatternDevice(n, nnz, rs_.as<int>(), col_.as<int>(), err);
    if (rc) return rc;
    n_ = n;
    nnz_ = nnz;
    patternReady_ = true;
    return ASTG_OK;
}

// Block size in cells, and levels per sweep group of the block-colour solver:
// pairs below a million rows (a few thousand rows per level, where the
// substituted dependencies cost less than the dependent step they save),
// single levels above. ASTG_AST_BLOCK / ASTG_AST_MERGE override them. Measured
// (mixed form, ms per projection, blocks of 4 / 8 / 16 cells, single levels /
// pairs): 2D 192x384 39.7 / 26.6, 66.3 / 39.9, 118.5 / 66.9; 2D 384x768
// 89.7 / 70.4, 143.2 / 94.5, 253.2 / 149.9; 3D 64^3 19.3 / 13.6,
// 32.8 / 21.5, 46.0 / 29.0; 3D 128^3 44.9 / 51.1, 60.8 / 59.4, 90.2 / 67.7.
static int astBlock() {
    const char* e = std::getenv("ASTG_AST_BLOCK");
    const int v = e ? std::atoi(e) : 0;
    return v > 0 ? v : 4;
}
// Levels per sweep group of the reference-numbered solver: in 2D the levels
// are a few hundred rows (the chain is nx + ny deep), so deep groups pay:
// 192x384 FP64, ms per projection at 2 / 3 / 4 / 6 / 8 levels per group:
// 342 / 235 / 180 / 136 / 133 (384x768: 1342 / - / 694 / 503 / 480), the same
// iterations. In 3D the default pairs stay (128^3: 86 / 102 / 141 at 2 / 3 / 4).
// ASTG_AST_REF_MERGE overrides it.
static int astRefMerge(int dim) {
    const char* e = std::getenv("ASTG_AST_REF_MERGE");
    const int v = e ? std::atoi(e) : 0;
    return v > 0 ? v : (dim == 2 ? 8 : 2);
}
static int astMerge(int n) {
    const char* e = std::getenv("ASTG_AST_MERGE");
    const int v = e ? std::atoi(e) : 0;
    return v > 0 ? v : (n < (1 << 20) ? 2 : 1);
}

int AstUniform::ensureBlockOrder(std::string& err) {
    if (blkReady_) return ASTG_OK;
    HIPCHK(keys_.alloc((size_t)n_ * 4));
    if (g_.d == 2)
        k_ast_block_keys<2><<<nblk(n_), BS, 0, stream_>>>(g_, n_, tnode_.as<int>(), astBlock(), keys_.as<int>());
    else
        k_ast_block_keys<3><<<nblk(n_), BS, 0, stream_>>>(g_, n_, tnode_.as<int>(), astBlock(), keys_.as<int>());
    HIPCHK(hipGetLastError());
    int rc = blk_.build(stream_, n_, nnz_, keys_.as<int>(), pcg_.devRowstart(), pcg_.devCol(), err, true);
    if (rc) return rc;
    if (!pcgBlk_) pcgBlk_.reset(new PcgDevice(stream_));
    pcgBlk_->setMergeLevels(astMerge(n_));
    rc = pcgBlk_->setPatternDevice(n_, nnz_, blk_.rowstart.as<int>(), blk_.col.as<int>(), err);
    if (rc) return rc;
    blkReady_ = true;
    return ASTG_OK;
}

int AstUniform::project(double dt, double rho, const astg_pcg_params& prmIn, astg_pcg_result& res,
                        std::string& err) {
    if (!(dt > 0) || !(rho > 0)) {
        err = "dt and rho must be positive";
        return ASTG_ERR_ARG;
    }
    if (prmIn.ordering < ASTG_ORDER_AUTO || prmIn.ordering > ASTG_ORDER_BLOCK_COLOUR) {
        err = "unknown ordering";
        return ASTG_ERR_ARG;
    }
    const bool blocks = prmIn.ordering == ASTG_ORDER_BLOCK_COLOUR ||
                        (prmIn.ordering == ASTG_ORDER_AUTO && prmIn.precision == ASTG_PREC_MIXED);
    (void)hipEventRecord(ev_[0], stream_);
    int rc = ensurePattern(err);
    if (rc) return rc;
    const int n = n_;
    MacPtrs m{{mac_[0].as<double>(), mac_[1].as<double>(), g_.d == 3 ? mac_[2].as<double>() : nullptr}};
    const double scale = rho / dt;
    if (g_.d == 2)
        k_ast_fill<2, true><<<nblk(n), BS, 0, stream_>>>(g_, n, tid_.as<int>(), tnode_.as<int>(), m,
                                                          tilt_.as<double>(), pcg_.devRowstart(), nullptr,
                                                          pcg_.values(), pcg_.rhs(), scale);
    else
        k_ast_fill<3, true><<<nblk(n), BS, 0, stream_>>>(g_, n, tid_.as<int>(), tnode_.as<int>(), m,
                                                          tilt_.as<double>(), pcg_.devRowstart(), nullptr,
                                                          pcg_.values(), pcg_.rhs(), scale);
    k_ast_zero_u64<<<1, 64, 0, stream_>>>(bmax_.as<unsigned long long>());
    k_ast_absmax<<<std::min(nblk(n), kMaxBlocks), BS, 0, stream_>>>(n, pcg_.rhs(), bmax_.as<unsigned long long>());
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, bmax_.p, 8, hipMemcpyDeviceToHost, stream_));
    (void)hipEventRecord(ev_[1], stream_);
    HIPCHK(hipStreamSynchronize(stream_));
    double bmax;
    std::memcpy(&bmax, &bits, 8);
    if (!(bmax <= DBL_MAX)) {
        err = "PCG: the right-hand side holds a NaN or infinity";
        return ASTG_ERR_NOT_CONVERGED;
    }
    std::memset(&res, 0, sizeof(res));
    res.n = n;
    res.nnz = nnz_;
    if (bmax == 0.0) {
        // the reference returns at once with a zero solution (pcg_solver.h:311-315)
        k_ast_zero<<<nblk(n), BS, 0, stream_>>>(n, pcg_.solution());
        res.converged = 1;
        (void)hipEventRecord(ev_[2], stream_);
        (void)hipEventRecord(ev_[3], stream_);
    } else {
        astg_pcg_params prm = prmIn;
        prm.norm = ASTG_NORM_INF;
        prm.ordering = ASTG_ORDER_REFERENCE;
        prm.tol = prmIn.tol / bmax;  // the reference's test is ||r||_inf <= tol, not relative
        // ASTG_TAU_AUTO: the reference's 0.97 in FP64, IC(0) in the mixed form
        const double tau = prm.mic_tau >= 0 ? prm.mic_tau : (prm.precision == ASTG_PREC_MIXED ? 0.0 : 0.97);
        PcgDevice* S = &pcg_;
        if (blocks) {
            if ((rc = ensureBlockOrder(err))) return rc;
            S = pcgBlk_.get();
            if ((rc = gatherPermuted(stream_, nnz_, blk_.src.as<int>(), pcg_.values(), S->values(), n,
                                     blk_.perm.as<int>(), pcg_.rhs(), S->rhs(), err)))
                return rc;
        }
        rc = S->factor(tau, prm.min_diag_ratio, err);
        if (rc) return rc;
        (void)hipEventRecord(ev_[2], stream_);
        rc = S->solve(prm, res, err);
        if (rc) return rc;
        if (blocks && (rc = scatterPermuted(stream_, n, blk_.perm.as<int>(), S->solution(), pcg_.solution(), err)))
            return rc;
        (void)hipEventRecord(ev_[3], stream_);
    }
    const double* p = pcg_.solution();
    const int nOpen = n - g_.nc;
    if (nOpen > 0) {
        if (g_.d == 2)
            k_ast_correct_tilt<2><<<nblk(nOpen), BS, 0, stream_>>>(g_, nOpen, tnode_.as<int>(), p,
                                                                    tilt_.as<double>(), dt, rho);
        else
            k_ast_correct_tilt<3><<<nblk(nOpen), BS, 0, stream_>>>(g_, nOpen, tnode_.as<int>(), p,
                                                                    tilt_.as<double>(), dt, rho);
    }
    for (int a = 0; a < g_.d; ++a)
        k_ast_correct_mac<<<nblk(g_.nm[a]), BS, 0, stream_>>>(g_, a, p, mac_[a].as<double>(), dt, rho);
    const long long walls = g_.d == 2 ? g_.res[1] : (long long)g_.res[1] * g_.res[2];
    k_ast_constrain<<<nblk(walls), BS, 0, stream_>>>(g_, mac_[0].as<double>(), mac_[1].as<double>());
    HIPCHK(hipGetLastError());
    (void)hipEventRecord(ev_[4], stream_);
    HIPCHK(hipEventSynchronize(ev_[4]));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev_[0], ev_[1]);
    res.ms_assemble = ms;
    (void)hipEventElapsedTime(&ms, ev_[1], ev_[2]);
    res.ms_factor = ms;
    (void)hipEventElapsedTime(&ms, ev_[2], ev_[3]);
    res.ms_solve = ms;
    (void)hipEventElapsedTime(&ms, ev_[3], ev_[4]);
    res.ms_velocity = ms;
    (void)hipEventElapsedTime(&ms, ev_[0], ev_[4]);
    res.ms_total = ms;
    solved_ = true;
    if (!res.converged) {
        err = "PCG did not reach the tolerance";
        return ASTG_ERR_NOT_CONVERGED;
    }
    return ASTG_OK;
}

int AstUniform::system(int* rowstart, int* col, double* val, double* b, double* p, std::string& err) {
    if (!solved_) {
        err = "no projection on the current layout yet";
        return ASTG_ERR_STATE;
    }
    if (rowstart) HIPCHK(hipMemcpyAsync(rowstart, pcg_.devRowstart(), ((size_t)n_ + 1) * 4, hipMemcpyDeviceToHost, stream_));
    if (col) HIPCHK(hipMemcpyAsync(col, pcg_.devCol(), (size_t)nnz_ * 4, hipMemcpyDeviceToHost, stream_));
    if (val) HIPCHK(hipMemcpyAsync(val, pcg_.values(), (size_t)nnz_ * 8, hipMemcpyDeviceToHost, stream_));
    if (b) HIPCHK(hipMemcpyAsync(b, pcg_.rhs(), (size_t)n_ * 8, hipMemcpyDeviceToHost, stream_));
    if (p) HIPCHK(hipMemcpyAsync(p, pcg_.solution(), (size_t)n_ * 8, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

int AstUniform::tiltIndex(int* tid, std::string& err) {
    int rc = ensurePattern(err);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(tid, tid_.p, (size_t)g_.nt * 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

}  // namespace astg
