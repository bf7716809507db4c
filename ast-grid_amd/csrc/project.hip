// Device assembly of the projection system and velocity correction.
//
// One lane per matrix row, reading the SoA stencil tables built on the host
// (project_host.cpp). The arithmetic of each coefficient follows the
// reference's expression trees in OctreeOmniGridSolver3::project
// (src/octree_omnigrid_solver3.cpp:367-651) and updateVelocityByPressure
// (681-847); set/add semantics of PCG::SparseMatrix are reproduced per slot.
#include <cmath>
#include <cstring>

#include "device_common.h"
#include "linrec.h"
#include "project.h"

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
constexpr double kSqrt3 = 1.7320508075688772;  // reference SQRT3 (macro.h:17)
constexpr double kEps = 2.220446049250313e-16; // std::numeric_limits<double>::epsilon()

__constant__ double c_h[16];
__constant__ int c_faceCorner[6][4];
__constant__ int c_faceNode[6];
__constant__ int c_halfSlot[6][5];

inline int blocks(long n) { return (int)((n + BS - 1) / BS); }
}  // namespace

// ---------------------------------------------------------------------------
// octagon rows (project 367-477)
// ---------------------------------------------------------------------------
// The assembly and velocity-correction kernels are bodies run by k_run
// (forward) or recorded for their transposes (linrec.h): velocities, b and p
// go through the policy P; matrix values are written only by the forward.
struct BAsmOct {
    int no;
    const int *row, *corner, *face;
    const uint8_t* pos;
    const int* rs;
    const double* tilt_e;
    const uint8_t* tilt_bits;
    const int* pid_tilt;
    const double *w_mac, *w_tilt, *u_mac, *u_tilt;
    const uint8_t* level;
    long G;
    double *A, *b;
    template <class Pol>
    __device__ void operator()(int j, Pol& P) const {
    const int r = row[j];
    int cg[8];
    double e[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        cg[k] = corner[(size_t)k * no + j];
        e[k] = tilt_e[cg[k]];
    }
    const int g = cg[6];
    const double h = c_h[level[g]];
    const int base = rs[r], len = rs[r + 1] - base;
    if (Pol::kAux)
        for (int p = 0; p < len; ++p) A[base + p] = 0.0;
    double center = 0.0, tb = 0.0;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int code = face[(size_t)f * no + j];
        if (code == FACE_SKIP) continue;
        const int axis = f >> 1;
        double area = h * h;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const double ek = e[c_faceCorner[f][m]];
            area -= ek * ek * 0.5;
        }
        const int fn = cg[c_faceNode[f]];
        double tCenter, tCoeff;
        if (code >= 0) {
            const double term = w_mac[axis * G + fn];
            tCenter = term * area / h;
            tCoeff = area * term;
            if (Pol::kAux) A[base + pos[(size_t)f * no + j]] = -tCenter;
        } else if (code == FACE_EXCLUDED || code == FACE_CLOSED) {
            tCenter = 0.0;
            tCoeff = 0.0;
        } else {  // open wall: Dirichlet
            tCoeff = area;
            tCenter = tCoeff / h;
        }
        center += tCenter;
        tb += ((f & 1) ? 1 : -1) * tCoeff * P.ld(u_mac, axis * G + fn);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int gk = cg[k];
        const uint8_t bits = tilt_bits[gk];
        if ((bits & TB_CLOSED) && !(bits & TB_HALF)) continue;
        const int re = ((k & 1) ? 8 : 6) - k;
        const double area = e[k] * e[k] * kSqrt3 * 0.5;
        const double term = w_tilt[re * G + gk];
        const int t = pid_tilt[gk];
        double tCenter = 0.0, tCoeff = 0.0;
        if (t >= 0) {
            tCenter = term * area / (h * kSqrt3 * 0.5);
            tCoeff = area * term;
            if (Pol::kAux) A[base + pos[(size_t)(6 + k) * no + j]] += -tCenter;
        }
        center += tCenter;
        tb += (k < 4 ? -1 : 1) * tCoeff * P.ld(u_tilt, re * G + gk);
    }
    if (fabs(center) < kEps) {
        center = 1.0;
        tb = 0.0;
    }
    if (Pol::kAux) A[base + pos[(size_t)14 * no + j]] = center;
    P.st(b, r, tb);
    }
};

// ---------------------------------------------------------------------------
// tilt rows (project 479-552)
// ---------------------------------------------------------------------------
struct BAsmTilt {
    int nt;
    const int *row, *node, *nb;
    const uint8_t *nblev, *pos;
    const int* rs;
    const double *tilt_e, *w_tilt, *u_tilt;
    const uint8_t* level;
    long G;
    double *A, *b;
    template <class Pol>
    __device__ void operator()(int j, Pol& P) const {
    const int r = row[j], g = node[j];
    const double h = c_h[level[g]];
    const int base = rs[r], len = rs[r + 1] - base;
    if (Pol::kAux)
        for (int p = 0; p < len; ++p) A[base + p] = 0.0;
    const double e = tilt_e[g];
    const double area = e * e * kSqrt3 * 0.5;
    double center = 0.0, tb = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int code = nb[(size_t)k * nt + j];
        if (code == TNB_NONE) continue;
        const double tw = w_tilt[k * G + g];
        double tCenter, tCoeff;
        if (code == TNB_DIRICHLET) {
            tCenter = area / (h * kSqrt3 / 3.0);
            tCoeff = area;
        } else if (code >= 0) {
            const double hn = c_h[nblev[(size_t)k * nt + j]];
            tCenter = tw * area / (hn * kSqrt3 * 0.5);
            tCoeff = area * tw;
            if (Pol::kAux) A[base + pos[(size_t)k * nt + j]] = -tCenter;
        } else {
            tCenter = 0.0;
            tCoeff = 0.0;
        }
        center += tCenter;
        tb += (k < 4 ? -1 : 1) * tCoeff * P.ld(u_tilt, k * G + g);
    }
    if (fabs(center) < kEps) center = 1.0;
    if (Pol::kAux) A[base + pos[(size_t)8 * nt + j]] = center;
    P.st(b, r, tb);
    }
};

// ---------------------------------------------------------------------------
// half-tilt face terms added into coarse octagon rows (project 572-581), in
// halfGhostIdx order per target row
// ---------------------------------------------------------------------------
struct BAsmHalfCoarse {
    int nrows;
    const int *hrow, *hstart, *hq;
    const uint8_t *pdiag, *poff;
    const int *hnode, *rs;
    const double* tilt_e;
    const int8_t* tilt_hd;
    const double *w_tilt, *u_tilt;
    const uint8_t* level;
    long G;
    double *A, *b;
    template <class Pol>
    __device__ void operator()(int j, Pol& P) const {
    const int r = hrow[j], base = rs[r];
    double br = P.ld(b, r);
    for (int idx = hstart[j]; idx < hstart[j + 1]; ++idx) {
        const int q = hq[idx], g = hnode[q], hd = tilt_hd[g];
        const double h = c_h[level[g]];
        const double e = tilt_e[g];
        const double area = e * e * 2;
        const int os = c_halfSlot[hd][4];
        const double w = w_tilt[os * G + g], u = P.ld(u_tilt, os * G + g);
        if (Pol::kAux) {
            A[base + pdiag[idx]] += w * area / h;
            A[base + poff[idx]] += -w * area / h;
        }
        br += (hd < 3 ? -1 : 1) * u * w * area;
    }
    P.st(b, r, br);
    }
};

// half-tilt rows (project 556-649), occurrences in halfGhostIdx order
struct BAsmHalfRows {
    int nrows;
    const int *hrow, *hstart, *hq;
    const uint8_t* hpos;
    const int *hnode, *ht;
    const uint8_t* hclosed;
    const int *hnb, *rs;
    const double* tilt_e;
    const int8_t* tilt_hd;
    const double *w_tilt, *u_tilt;
    const uint8_t* level;
    long G;
    double *A, *b;
    template <class Pol>
    __device__ void operator()(int j, Pol& P) const {
    const int c = hrow[j], base = rs[c], len = rs[c + 1] - base;
    if (Pol::kAux)
        for (int p = 0; p < len; ++p) A[base + p] = 0.0;
    double bc = 0.0;
    for (int idx = hstart[j]; idx < hstart[j + 1]; ++idx) {
        const int q = hq[idx], g = hnode[q], hd = tilt_hd[g];
        const uint8_t* pp = hpos + (size_t)6 * idx;
        const double h = c_h[level[g]];
        const double e = tilt_e[g];
        double area = e * e * 2;
        const int os = c_halfSlot[hd][4];
        const double wf = w_tilt[os * G + g], uf = P.ld(u_tilt, os * G + g);
        double center = 0.0, tb = 0.0;
        const int t = ht[q];
        if (t != -20) {
            if (t >= 0) {
                center += wf * area / h;
                if (Pol::kAux) A[base + pp[0]] += -wf * area / h;
                tb += (hd < 3 ? 1 : -1) * uf * wf * area;
            }
        } else {
            const double tCenter = hclosed[q] ? 0.0 : 1000.0;
            center += tCenter;
            tb += (hd < 3 ? 1 : -1) * uf * area;
        }
        area = e * e * kSqrt3 * 0.5;
        for (int m = 0; m < 4; ++m) {
            const int k = c_halfSlot[hd][m];
            const int code = hnb[(size_t)4 * q + m];
            const double tw = w_tilt[k * G + g];
            double tCenter = 0.0, tCoeff = 0.0;
            if (code >= 0) {
                tCenter = area * tw / (h * kSqrt3 * 0.5);
                tCoeff = area * tw;
                if (Pol::kAux) A[base + pp[1 + m]] = -tCenter;
            }
            center += tCenter;
            tb += (k < 4 ? -1 : 1) * tCoeff * P.ld(u_tilt, k * G + g);
        }
        if (Pol::kAux) A[base + pp[5]] += center;
        bc = tb;
    }
    P.st(b, c, bc);
    }
};

struct BScale {
    double* b;
    double s;
    template <class Pol>
    __device__ void operator()(int i, Pol& P) const { P.st(b, i, P.ld(b, i) * s); }
};

// ---------------------------------------------------------------------------
// velocity correction (updateVelocityByPressure 681-847)
// ---------------------------------------------------------------------------
struct BVelMac {
    int nvm;
    const int *node, *cidx, *nb;
    const double *w_mac, *p;
    const uint8_t* level;
    long G;
    double dt, rho;
    double* u_mac;
    template <class Pol>
    __device__ void operator()(int j, Pol& P) const {
        const int g = node[j], c = cidx[j];
        const double h = c_h[level[g]];
        const double p0 = c >= 0 ? P.ld(p, c) : 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int code = nb[(size_t)a * nvm + j];
            if (code == VNB_SKIP) continue;
            const size_t ui = a * G + g;
            if (w_mac[a * G + g] <= 0) {
                P.st(u_mac, ui, 0.0);
            } else {
                const double p1 = code >= 0 ? P.ld(p, code) : 0.0;
                P.st(u_mac, ui, P.ld(u_mac, ui) - dt / rho * (p0 - p1) / h);
            }
        }
    }
};

struct BVelTilt {
    int nvt;
    const int *node, *cidx, *nb;
    const uint8_t* lev;
    const double *w_tilt, *p;
    long G;
    double dt, rho;
    double* u_tilt;
    template <class Pol>
    __device__ void operator()(int j, Pol& P) const {
        const int g = node[j], c = cidx[j];
        const double p0 = c >= 0 ? P.ld(p, c) : 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int code = nb[(size_t)k * nvt + j];
            const size_t ui = k * G + g;
            if (w_tilt[k * G + g] > 0 && code != VNB_ZERO) {
                const double hn = c_h[lev[(size_t)k * nvt + j]];
                const double p1 = code >= 0 ? P.ld(p, code) : 0.0;
                P.st(u_tilt, ui, P.ld(u_tilt, ui) - dt / rho * (k < 4 ? 1 : -1) / (hn * kSqrt3 * 0.5) * (p1 - p0));
            } else {
                P.st(u_tilt, ui, 0.0);
            }
        }
    }
};

struct BVelHalf {
    int nh;
    const int *hnode, *hc, *ht, *hnb;
    const int8_t* tilt_hd;
    const double *w_tilt, *p;
    const uint8_t* level;
    long G;
    double dt, rho;
    double* u_tilt;
    template <class Pol>
    __device__ void operator()(int q, Pol& P) const {
        const int g = hnode[q], c = hc[q], hd = tilt_hd[g];
        const double h = c_h[level[g]];
        const double p0 = c >= 0 ? P.ld(p, c) : 0.0;
        const int fs = c_halfSlot[hd][4];
        {
            const size_t ui = fs * G + g;
            if (w_tilt[fs * G + g] > 0) {
                const int t = ht[q];
                const double p1 = t >= 0 ? P.ld(p, t) : 0.0;
                P.st(u_tilt, ui, P.ld(u_tilt, ui) - dt / rho * (hd < 3 ? 1 : -1) / h * (p0 - p1));
            } else {
                P.st(u_tilt, ui, 0.0);
            }
        }
        int m = 0;
        for (int k = 0; k < 8; ++k) {
            if (k == fs) continue;
            const size_t ui = k * G + g;
            const bool slot = (m < 4 && c_halfSlot[hd][m] == k);
            if (slot) {
                const int code = hnb[(size_t)4 * q + m];
                ++m;
                if (w_tilt[k * G + g] > 0) {
                    const double p1 = code >= 0 ? P.ld(p, code) : 0.0;
                    P.st(u_tilt, ui, P.ld(u_tilt, ui) - dt / rho * (k < 4 ? 1 : -1) / (h * kSqrt3 * 0.5) * (p1 - p0));
                    continue;
                }
            }
            P.st(u_tilt, ui, 0.0);
        }
    }
};

struct AsmPasses {
    BAsmOct oct;
    BAsmTilt tilt;
    BAsmHalfCoarse hc;
    BAsmHalfRows hr;
    BScale scale;
};
struct VelPasses {
    BVelMac mac;
    BVelTilt tilt;
    BVelHalf half;
};

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
template <typename T>
static int up(DevBuf& d, const std::vector<T>& v, hipStream_t s, std::string& err) {
    HIPCHK(d.alloc(v.size() * sizeof(T) + 16));
    if (!v.empty()) HIPCHK(hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return ASTG_OK;
}

int ProjectDevice::upload(const ProjectTables& t, const Octree& o, std::string& err) {
    T_ = &t;
    int rc;
#define UP(buf, vec) \
    if ((rc = up(buf, vec, stream_, err))) return rc;
    UP(level_, t.level);
    UP(tilt_e_, t.tilt_e);
    UP(tilt_bits_, t.tilt_bits);
    UP(tilt_hd_, t.tilt_hd);
    UP(pid_tilt_, t.pid_tilt);
    UP(w_mac_, t.w_mac);
    UP(w_tilt_, t.w_tilt);
    UP(oct_row_, t.oct_row);
    UP(oct_corner_, t.oct_corner);
    UP(oct_face_, t.oct_face);
    UP(oct_pos_, t.oct_pos);
    UP(tilt_row_, t.tilt_row);
    UP(tilt_node_, t.tilt_node);
    UP(tilt_nb_, t.tilt_nb);
    UP(tilt_nblev_, t.tilt_nblev);
    UP(tilt_pos_, t.tilt_pos);
    UP(half_node_, t.half_node);
    UP(half_c_, t.half_c);
    UP(half_t_, t.half_t);
    UP(half_closed_, t.half_closed);
    UP(half_nb_, t.half_nb);
    UP(hc_row_, t.hc_row);
    UP(hc_start_, t.hc_start);
    UP(hc_q_, t.hc_q);
    UP(hc_pos_diag_, t.hc_pos_diag);
    UP(hc_pos_off_, t.hc_pos_off);
    UP(hr_row_, t.hr_row);
    UP(hr_start_, t.hr_start);
    UP(hr_q_, t.hr_q);
    UP(hr_pos_, t.hr_pos);
    UP(vm_node_, t.vm_node);
    UP(vm_c_, t.vm_c);
    UP(vm_nb_, t.vm_nb);
    UP(vt_node_, t.vt_node);
    UP(vt_c_, t.vt_c);
    UP(vt_nb_, t.vt_nb);
    UP(vt_lev_, t.vt_lev);
#undef UP
    // The live velocity and the advection's output pair it is swapped with,
    // all four at exactly 3G and 8G doubles, allocated when the node count
    // changes (a new topology: the caller clears them). A tilt update keeps
    // the node sets, so the live fields must stay where they are: a buffer
    // reallocated here would drop the velocity.
    if (velG_ != (long long)t.G) {
        for (DevBuf* d : {&u_mac_, &t_mac_}) {
            d->release();
            HIPCHK(d->alloc((size_t)3 * t.G * 8));
        }
        for (DevBuf* d : {&u_tilt_, &t_tilt_}) {
            d->release();
            HIPCHK(d->alloc((size_t)8 * t.G * 8));
        }
        velG_ = (long long)t.G;
    }
    double h[16] = {0};
    for (int l = 0; l < o.L && l < 16; ++l) h[l] = o.lv[l].h;
    int faceCorner[6][4];
    for (int f = 0; f < 6; ++f) {
        const int axis = f / 2;
        int m = 0;
        for (int a = 0; a <= 1; ++a)
            for (int bb = 0; bb <= 1; ++bb) {
                int off[3] = {0, 0, 0};
                off[axis] = (f & 1) ? 0 : 1;
                off[(axis + 1) % 3] += a;
                off[(axis + 2) % 3] += bb;
                for (int k = 0; k < 8; ++k)
                    if (kOctNbTilt[k][0] == off[0] && kOctNbTilt[k][1] == off[1] && kOctNbTilt[k][2] == off[2])
                        faceCorner[f][m] = k;
                ++m;
            }
    }
    int faceNode[6];
    for (int f = 0; f < 6; ++f)
        for (int k = 0; k < 8; ++k)
            if (kOctNbTilt[k][0] == kOctNbFace[f][0] && kOctNbTilt[k][1] == kOctNbFace[f][1] &&
                kOctNbTilt[k][2] == kOctNbFace[f][2])
                faceNode[f] = k;
    std::memcpy(hConst_, h, sizeof(h));
    std::lock_guard<std::recursive_mutex> lk(constMutex());
    if ((rc = bindConstants(true, err))) return rc;
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_faceCorner), faceCorner, sizeof(faceCorner), 0, hipMemcpyHostToDevice,
                                  stream_));
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_faceNode), faceNode, sizeof(faceNode), 0, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_halfSlot), kHalfSlot, sizeof(int) * 30, 0, hipMemcpyHostToDevice,
                                  stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ASTG_OK;
}

int ProjectDevice::bindConstants(bool force, std::string& err) {
    if (!takeConstants(CT_PROJECT, constOwner_, force)) return ASTG_OK;
    HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_h), hConst_, sizeof(hConst_), 0, hipMemcpyHostToDevice, stream_));
    return ASTG_OK;
}

int ProjectDevice::assemble(double* A, double* b, const int* rs, double rho, double dt, std::string& err) {
    std::lock_guard<std::recursive_mutex> lk(constMutex());
    if (int rc = bindConstants(false, err)) return rc;
    const ProjectTables& t = *T_;
    const long G = t.G;
    AsmPasses ps;
    asmPasses(A, b, rs, rho, dt, ps);
    if (ps.oct.no) k_run<<<blocks(ps.oct.no), BS, 0, stream_>>>(ps.oct, ps.oct.no);
    if (ps.tilt.nt) k_run<<<blocks(ps.tilt.nt), BS, 0, stream_>>>(ps.tilt, ps.tilt.nt);
    if (ps.hc.nrows) k_run<<<blocks(ps.hc.nrows), BS, 0, stream_>>>(ps.hc, ps.hc.nrows);
    if (ps.hr.nrows) k_run<<<blocks(ps.hr.nrows), BS, 0, stream_>>>(ps.hr, ps.hr.nrows);
    k_run<<<blocks(t.n), BS, 0, stream_>>>(ps.scale, t.n);
    HIPCHK(hipGetLastError());
    return ASTG_OK;
}

void ProjectDevice::asmPasses(double* A, double* b, const int* rs, double rho, double dt, AsmPasses& ps) {
    const ProjectTables& t = *T_;
    const long G = t.G;
    ps.oct = BAsmOct{t.n_oct, oct_row_.as<int>(), oct_corner_.as<int>(), oct_face_.as<int>(), oct_pos_.as<uint8_t>(),
                     rs, tilt_e_.as<double>(), tilt_bits_.as<uint8_t>(), pid_tilt_.as<int>(), w_mac_.as<double>(),
                     w_tilt_.as<double>(), u_mac_.as<double>(), u_tilt_.as<double>(), level_.as<uint8_t>(), G, A, b};
    ps.tilt = BAsmTilt{t.n_tilt, tilt_row_.as<int>(), tilt_node_.as<int>(), tilt_nb_.as<int>(),
                       tilt_nblev_.as<uint8_t>(), tilt_pos_.as<uint8_t>(), rs, tilt_e_.as<double>(),
                       w_tilt_.as<double>(), u_tilt_.as<double>(), level_.as<uint8_t>(), G, A, b};
    ps.hc = BAsmHalfCoarse{(int)t.hc_row.size(), hc_row_.as<int>(), hc_start_.as<int>(), hc_q_.as<int>(),
                           hc_pos_diag_.as<uint8_t>(), hc_pos_off_.as<uint8_t>(), half_node_.as<int>(), rs,
                           tilt_e_.as<double>(), tilt_hd_.as<int8_t>(), w_tilt_.as<double>(), u_tilt_.as<double>(),
                           level_.as<uint8_t>(), G, A, b};
    ps.hr = BAsmHalfRows{t.n_half_rows, hr_row_.as<int>(), hr_start_.as<int>(), hr_q_.as<int>(), hr_pos_.as<uint8_t>(),
                         half_node_.as<int>(), half_t_.as<int>(), half_closed_.as<uint8_t>(), half_nb_.as<int>(), rs,
                         tilt_e_.as<double>(), tilt_hd_.as<int8_t>(), w_tilt_.as<double>(), u_tilt_.as<double>(),
                         level_.as<uint8_t>(), G, A, b};
    ps.scale = BScale{b, rho / dt};
}

int ProjectDevice::correct(const double* p, double rho, double dt, std::string& err) {
    std::lock_guard<std::recursive_mutex> lk(constMutex());
    if (int rc = bindConstants(false, err)) return rc;
    const ProjectTables& t = *T_;
    const long G = t.G;
    (void)G;
    VelPasses ps;
    velPasses(p, rho, dt, ps);
    if (ps.mac.nvm) k_run<<<blocks(ps.mac.nvm), BS, 0, stream_>>>(ps.mac, ps.mac.nvm);
    if (ps.tilt.nvt) k_run<<<blocks(ps.tilt.nvt), BS, 0, stream_>>>(ps.tilt, ps.tilt.nvt);
    if (ps.half.nh) k_run<<<blocks(ps.half.nh), BS, 0, stream_>>>(ps.half, ps.half.nh);
    HIPCHK(hipGetLastError());
    return ASTG_OK;
}

void ProjectDevice::velPasses(const double* p, double rho, double dt, VelPasses& ps) {
    const ProjectTables& t = *T_;
    const long G = t.G;
    ps.mac = BVelMac{(int)t.vm_node.size(), vm_node_.as<int>(), vm_c_.as<int>(), vm_nb_.as<int>(), w_mac_.as<double>(),
                     p, level_.as<uint8_t>(), G, dt, rho, u_mac_.as<double>()};
    ps.tilt = BVelTilt{(int)t.vt_node.size(), vt_node_.as<int>(), vt_c_.as<int>(), vt_nb_.as<int>(),
                       vt_lev_.as<uint8_t>(), w_tilt_.as<double>(), p, G, dt, rho, u_tilt_.as<double>()};
    ps.half = BVelHalf{t.nh, half_node_.as<int>(), half_c_.as<int>(), half_t_.as<int>(), half_nb_.as<int>(),
                       tilt_hd_.as<int8_t>(), w_tilt_.as<double>(), p, level_.as<uint8_t>(), G, dt, rho,
                       u_tilt_.as<double>()};
}

int ProjectDevice::recordPasses(const StateMap& M, long long stateLen, double* A, double* b, const int* rs,
                                const double* p, double rho, double dt, DevBuf& claim, DevBuf& scratch,
                                std::vector<PassRecord>& asmRec, std::vector<PassRecord>& velRec, std::string& err) {
    std::lock_guard<std::recursive_mutex> lk(constMutex());
    if (int rc = bindConstants(false, err)) return rc;
    const ProjectTables& t = *T_;
    AsmPasses a;
    asmPasses(A, b, rs, rho, dt, a);
    VelPasses v;
    velPasses(p, rho, dt, v);
    asmRec.clear();
    asmRec.resize(5);
    velRec.clear();
    velRec.resize(3);
    int rc;
    if ((rc = recordPass(a.oct, a.oct.no, M, stateLen, claim, scratch, stream_, asmRec[0], err)) ||
        (rc = recordPass(a.tilt, a.tilt.nt, M, stateLen, claim, scratch, stream_, asmRec[1], err)) ||
        (rc = recordPass(a.hc, a.hc.nrows, M, stateLen, claim, scratch, stream_, asmRec[2], err)) ||
        (rc = recordPass(a.hr, a.hr.nrows, M, stateLen, claim, scratch, stream_, asmRec[3], err)) ||
        (rc = recordPass(a.scale, t.n, M, stateLen, claim, scratch, stream_, asmRec[4], err)) ||
        (rc = recordPass(v.mac, v.mac.nvm, M, stateLen, claim, scratch, stream_, velRec[0], err)) ||
        (rc = recordPass(v.tilt, v.tilt.nvt, M, stateLen, claim, scratch, stream_, velRec[1], err)) ||
        (rc = recordPass(v.half, v.half.nh, M, stateLen, claim, scratch, stream_, velRec[2], err)))
        return rc;
    return ASTG_OK;
}

int ProjectDevice::snapshot(bool restore, std::string& err) {
    const size_t G = T_->G;
    if (!restore) {
        HIPCHK(snap_mac_.alloc(3 * G * 8 + 16));
        HIPCHK(snap_tilt_.alloc(8 * G * 8 + 16));
        HIPCHK(hipMemcpyAsync(snap_mac_.p, u_mac_.p, 3 * G * 8, hipMemcpyDeviceToDevice, stream_));
        HIPCHK(hipMemcpyAsync(snap_tilt_.p, u_tilt_.p, 8 * G * 8, hipMemcpyDeviceToDevice, stream_));
        haveSnap_ = true;
        return ASTG_OK;
    }
    if (!haveSnap_) {
        err = "no snapshot saved";
        return ASTG_ERR_STATE;
    }
    HIPCHK(hipMemcpyAsync(u_mac_.p, snap_mac_.p, 3 * G * 8, hipMemcpyDeviceToDevice, stream_));
    HIPCHK(hipMemcpyAsync(u_tilt_.p, snap_tilt_.p, 8 * G * 8, hipMemcpyDeviceToDevice, stream_));
    return ASTG_OK;
}

// Algorithmic bytes of one assembly: tables, the node data each row gathers
// (counted once per row), A values and b written.
double ProjectDevice::assembleBytes() const {
    const ProjectTables& t = *T_;
    const double no = t.n_oct, nt = t.n_tilt;
    const double oct = no * (4 + 8 * 4 + 6 * 4 + 15) + no * 8 * (8 + 1 + 4 + 8 + 8) + no * 6 * 16;
    const double tilt = nt * (8 + 8 * 4 + 8 + 9) + nt * (8 + 8 * 16);
    const double out = (double)t.col.size() * 8 + (double)t.n * 8 * 2;
    return oct + tilt + out;
}

}  // namespace astg
