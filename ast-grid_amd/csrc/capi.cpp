// C ABI (include/astg.h). No exception crosses this boundary: every entry
// point catches, records the message on the context and returns a status.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/astg.h"
#include "octree.h"
#include "timing.h"
#include "pcg.h"
#include "reorder.h"
#include "frame_io.h"
#include "advect.h"
#include "project.h"
#include "project_host.h"
#include "scene.h"

struct astg_ctx {
    int device = -1;
    hipStream_t stream = nullptr;
    std::string err;
    std::unique_ptr<astg::Octree> tree;
    astg::Numbering num;
    astg::Weights w;
    astg::ProjectTables tables;
    bool tablesReady = false;
    bool deviceReady = false;
    std::vector<std::vector<double>> hostMac, hostTilt;  // host-only contexts
    std::unique_ptr<astg::PcgDevice> pcg;                 // projection system
    std::unique_ptr<astg::PcgDevice> pcgBlk;              // the same system in block-colour order (reorder.h)
    astg::BlockOrderDevice blk;  // permutation and permuted pattern, device-built
    bool blkReady = false;
    bool factoredBlk = false;  // the last factorisation was of the block-colour copy
    std::unique_ptr<astg::ProjectDevice> proj;
    std::unique_ptr<astg::AdvectDevice> adv;              // post-projection step and advection
    std::unique_ptr<astg::PcgDevice> csr;                 // CSR seam
    std::vector<int> csrRs, csrCol;
    std::unique_ptr<astg::PcgDevice> batch;               // batched CSR solves (stacked systems)
    std::vector<int> batchRs, batchCol, batchSeg;
    int batchMerge = 0;                                   // levels per sweep group of the cached pattern
    astg_batch_tuning batchTune = {2, 8, 8, 0.0};         // launch shape of batched solves
    astg::PcgDevice* lastPcg = nullptr;
    hipEvent_t ev[6] = {};
    bool tiltAdaptivity = false;                          // enableTiltAdaptivity
};

namespace {

int fail(astg_ctx* c, int code, const std::string& m) {
    if (c) c->err = m;
    return code;
}

template <typename F>
int guarded(astg_ctx* c, F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(c, ASTG_ERR_ARG, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(c, ASTG_ERR_TOPOLOGY, e.what());
    }
}

int runProject(astg_ctx* c, double dt, double rho, const astg_pcg_params& prm, astg_pcg_result* result);
int needAdvect(astg_ctx* c);

int needDevice(astg_ctx* c) {
    if (c->device < 0) return fail(c, ASTG_ERR_STATE, "context has no device (created with device < 0)");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, ASTG_ERR_HIP, "hipSetDevice failed");
    return ASTG_OK;
}

int ensureTables(astg_ctx* c) {
    if (!c->tree) return fail(c, ASTG_ERR_STATE, "no topology set");
    if (!c->tablesReady) {
        astg::numberUnknowns(*c->tree, c->num);
        astg::computeWeights(*c->tree, c->w);
        astg::buildProjectTables(*c->tree, c->num, c->w, c->tables);
        c->tablesReady = true;
        c->deviceReady = false;
        c->blkReady = false;
        if (c->device < 0) {
            c->hostMac.assign(c->tree->L, {});
            c->hostTilt.assign(c->tree->L, {});
            for (int l = 0; l < c->tree->L; ++l) {
                c->hostMac[l].assign((size_t)c->tree->lv[l].nnode * 3, 0.0);
                c->hostTilt[l].assign((size_t)c->tree->lv[l].nnode * 8, 0.0);
            }
        }
    }
    if (c->device >= 0 && !c->deviceReady) {
        int rc = needDevice(c);
        if (rc) return rc;
        if (!c->pcg) c->pcg.reset(new astg::PcgDevice(c->stream));
        if (!c->proj) c->proj.reset(new astg::ProjectDevice(c->stream));
        rc = c->proj->upload(c->tables, *c->tree, c->err);
        if (rc) return rc;
        if (!c->adv) c->adv.reset(new astg::AdvectDevice(c->stream));
        rc = c->adv->upload(*c->tree, c->tables, false, c->err);
        if (rc) return rc;
        rc = c->pcg->setPattern(c->tables.n, c->tables.rowstart.data(), c->tables.col.data(), c->err);
        if (rc) return rc;
        // fields, and the advection's output buffers, start at zero
        const size_t G = c->tables.G;
        std::vector<double> z((size_t)8 * G, 0.0);
        if (hipMemcpyAsync(c->proj->uMac(), z.data(), 3 * G * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(c->proj->uTilt(), z.data(), 8 * G * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(c->proj->tMac(), z.data(), 3 * G * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(c->proj->tTilt(), z.data(), 8 * G * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return fail(c, ASTG_ERR_HIP, "field upload failed");
        c->deviceReady = true;
    }
    return ASTG_OK;
}

// Tilt update with the fields kept: numbering, weights, tables, device
// topology and the projection pattern follow the new tilt state; the node
// sets, and so every field array, are unchanged. On any failure the cached
// tables and device state are dropped (tablesReady = deviceReady = false), so
// the next call rebuilds them from the tree's tilt state (fields restart at
// zero) instead of running on tables and buffers of two different states.
int retopologizeSteps(astg_ctx* c);
int retopologize(astg_ctx* c) {
    int rc;
    try {
        rc = retopologizeSteps(c);
    } catch (...) {
        c->tablesReady = false;
        c->deviceReady = false;
        throw;
    }
    if (rc) {
        c->tablesReady = false;
        c->deviceReady = false;
    }
    return rc;
}
int retopologizeSteps(astg_ctx* c) {
    astg::PhaseTimer tm("retopologize");
    long long na0;
    double ma0;
    astg::allocStats(na0, ma0);
    c->blkReady = false;  // the block-colour system follows the new numbering on its next use
    astg::numberUnknowns(*c->tree, c->num);
    tm.mark("numbering");
    astg::computeWeights(*c->tree, c->w);
    tm.mark("weights");
    astg::buildProjectTables(*c->tree, c->num, c->w, c->tables);
    tm.mark("tables");
    int rc = c->proj->upload(c->tables, *c->tree, c->err);
    if (rc) return rc;
    tm.mark("project upload");
    rc = c->adv->upload(*c->tree, c->tables, true, c->err);
    if (rc) return rc;
    tm.mark("advect upload");
    rc = c->pcg->setPattern(c->tables.n, c->tables.rowstart.data(), c->tables.col.data(), c->err);
    tm.mark("solver pattern");
    long long na1;
    double ma1;
    astg::allocStats(na1, ma1);
    tm.note("device allocations", (double)(na1 - na0), ma1 - ma0);
    return rc;
}

int adjustTilts(astg_ctx* c, double dThr, double dt, double vThr, double maxThr) {
    astg::PhaseTimer tm("adjustE");
    std::vector<uint8_t> open;
    int rc = c->adv->decideTilts(*c->proj, dThr, dt, vThr, maxThr, open, c->err);
    if (rc) return rc;
    tm.mark("decide");
    const bool changed = c->tree->setTiltOpenAll(open.data());
    tm.mark("commit");
    if (!changed) return ASTG_OK;
    return retopologize(c);
}

int checkLevel(astg_ctx* c, int level) {
    if (!c || !c->tree) return fail(c, ASTG_ERR_STATE, "no topology set");
    if (level < 0 || level >= c->tree->L) return fail(c, ASTG_ERR_ARG, "level out of range");
    return ASTG_OK;
}

}  // namespace

extern "C" {

void astg_pcg_params_default(astg_pcg_params* p) {
    if (!p) return;
    p->tol = 1e-5;
    p->max_iterations = 3000;
    p->norm = ASTG_NORM_INF;
    p->mic_tau = ASTG_TAU_AUTO;
    p->min_diag_ratio = 0.25;
    p->precision = ASTG_PREC_FP64;
    p->max_outer = 0;
    p->ordering = ASTG_ORDER_AUTO;
}

int astg_create(astg_ctx** out, int device, uint32_t flags) {
    (void)flags;
    if (!out) return ASTG_ERR_ARG;
    *out = nullptr;
    auto* c = new (std::nothrow) astg_ctx();
    if (!c) return ASTG_ERR_ARG;
    c->device = device;
    if (device >= 0) {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || device >= count) {
            delete c;
            return ASTG_ERR_HIP;
        }
        if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
            delete c;
            return ASTG_ERR_HIP;
        }
        // this GPU's address of the no-alloc staging array (once per device)
        if (astg::noallocScratchInit(device, c->err) != ASTG_OK) {
            (void)hipStreamDestroy(c->stream);
            delete c;
            return ASTG_ERR_HIP;
        }
        for (auto& e : c->ev)
            if (hipEventCreate(&e) != hipSuccess) {
                delete c;
                return ASTG_ERR_HIP;
            }
    }
    *out = c;
    return ASTG_OK;
}

int astg_destroy(astg_ctx* c) {
    if (!c) return ASTG_ERR_ARG;
    if (c->device >= 0) {
        (void)hipSetDevice(c->device);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        // every device buffer (and pinned host block) is released while the
        // stream and the runtime are still alive, before the stream goes
        c->adv.reset();
        c->pcg.reset();
        c->pcgBlk.reset();
        c->blk = astg::BlockOrderDevice();
        c->proj.reset();
        c->csr.reset();
        c->batch.reset();
        c->lastPcg = nullptr;
        for (auto& e : c->ev)
            if (e) (void)hipEventDestroy(e);
        if (c->stream) (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return ASTG_OK;
}

const char* astg_last_error(const astg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int astg_split_spheres(int R, int levels, int nspheres, const double* spheres, const double* band, int* split_out) {
    if (R <= 0 || levels <= 0 || levels > 15 || (R >> (levels - 1)) < 2 || (R % (1 << (levels - 1))) != 0 ||
        nspheres < 0 || (nspheres && !spheres) || !band || !split_out)
        return ASTG_ERR_ARG;
    try {
        std::vector<int> split;
        astg::splitSpheres(R, levels, spheres, nspheres, band, split);
        astg::balanceSplit(R, levels, split);
        std::memcpy(split_out, split.data(), split.size() * sizeof(int));
    } catch (...) {
        return ASTG_ERR_ARG;
    }
    return ASTG_OK;
}

int astg_set_topology_split(astg_ctx* c, int R, double h0, int levels, const int* split, int closed_flags) {
    if (!c) return ASTG_ERR_ARG;
    if (R <= 0 || levels <= 0 || levels > 15 || !split || h0 <= 0 || (R % (1 << (levels - 1))) != 0 ||
        R > (1 << 16))
        return fail(c, ASTG_ERR_ARG, "bad topology arguments");
    for (size_t i = 0, n = (size_t)R * R * R; i < n; ++i)
        if (split[i] < 0 || split[i] >= levels) return fail(c, ASTG_ERR_ARG, "split map value out of range");
    return guarded(c, [&] {
        std::unique_ptr<astg::Octree> t(new astg::Octree());
        t->build(split, R, h0, levels, closed_flags);
        if (!t->err.empty()) return fail(c, ASTG_ERR_TOPOLOGY, t->err);
        c->tree = std::move(t);
        c->tablesReady = false;
        c->deviceReady = false;
        return ASTG_OK;
    });
}

int astg_set_tilt_open(astg_ctx* c, int level, const unsigned char* mask) {
    int rc = checkLevel(c, level);
    if (rc) return rc;
    if (!mask) return fail(c, ASTG_ERR_ARG, "null mask");
    c->tree->setTiltOpen(level, mask);
    c->tablesReady = false;
    c->deviceReady = false;
    return ASTG_OK;
}

int astg_set_boundary_sdf(astg_ctx* c, const double* sdf, const double* slip_center) {
    if (!c) return ASTG_ERR_ARG;
    if (!c->tree) return fail(c, ASTG_ERR_STATE, "no topology set");
    return guarded(c, [&] {
        astg::BoundarySdf b;
        if (sdf) {
            const size_t n = (size_t)c->tree->R * c->tree->R * c->tree->R;
            for (size_t i = 0; i < n; ++i)
                if (!std::isfinite(sdf[i])) return fail(c, ASTG_ERR_ARG, "boundary SDF holds a non-finite value");
            b.R = c->tree->R;
            b.h0 = c->tree->h0;
            b.phi.assign(sdf, sdf + n);
            if (slip_center)
                for (int k = 0; k < 3; ++k) b.slip[k] = slip_center[k];
        }
        c->tree->sdf = std::move(b);
        // numbering, weights, tables and every device array follow
        c->tablesReady = false;
        c->deviceReady = false;
        return ASTG_OK;
    });
}

int astg_sphere_sdf(int R, double h0, double cx, double cy, double cz, double radius, double* out) {
    if (R <= 0 || !(h0 > 0) || !(radius >= 0) || !out) return ASTG_ERR_ARG;
    for (int z = 0; z < R; ++z)
        for (int y = 0; y < R; ++y)
            for (int x = 0; x < R; ++x) {
                const double dx = (x + 0.5) * h0 - cx, dy = (y + 0.5) * h0 - cy, dz = (z + 0.5) * h0 - cz;
                out[((size_t)z * R + y) * R + x] = std::sqrt(dx * dx + dy * dy + dz * dz) - radius;
            }
    return ASTG_OK;
}

int astg_level_info(astg_ctx* c, int level, int* nnode, int* nleaf) {
    int rc = checkLevel(c, level);
    if (rc) return rc;
    if (nnode) *nnode = c->tree->lv[level].nnode;
    if (nleaf) *nleaf = c->tree->lv[level].nleaf;
    return ASTG_OK;
}

int astg_get_nodes(astg_ctx* c, int level, int* xyz, int* flags) {
    int rc = checkLevel(c, level);
    if (rc) return rc;
    const auto& L = c->tree->lv[level];
    for (int i = 0; i < L.nnode; ++i) {
        if (xyz) {
            xyz[3 * i] = L.x[i];
            xyz[3 * i + 1] = L.y[i];
            xyz[3 * i + 2] = L.z[i];
        }
        if (flags) flags[i] = L.flag[i];
    }
    return ASTG_OK;
}

int astg_get_tilt(astg_ctx* c, int level, double* e, int* bits, int* half_dir) {
    int rc = checkLevel(c, level);
    if (rc) return rc;
    const auto& T = c->tree->tilt[level];
    for (size_t i = 0; i < T.size(); ++i) {
        if (e) e[i] = T[i].e;
        if (bits) bits[i] = T[i].bits;
        if (half_dir) half_dir[i] = T[i].half_dir;
    }
    return ASTG_OK;
}

int astg_get_half_tilts(astg_ctx* c, int* lxyz, int* count) {
    if (!c || !c->tree) return fail(c, ASTG_ERR_STATE, "no topology set");
    const auto& H = c->tree->half;
    if (count) *count = (int)H.size();
    if (lxyz)
        for (size_t i = 0; i < H.size(); ++i) {
            lxyz[4 * i] = astg::lc_level(H[i]);
            lxyz[4 * i + 1] = astg::lc_x(H[i]);
            lxyz[4 * i + 2] = astg::lc_y(H[i]);
            lxyz[4 * i + 3] = astg::lc_z(H[i]);
        }
    return ASTG_OK;
}

int astg_get_numbering(astg_ctx* c, int level, int* oct, int* tilt, int* n_total) {
    int rc = checkLevel(c, level);
    if (rc) return rc;
    return guarded(c, [&] {
        if (!c->tablesReady) {
            astg::numberUnknowns(*c->tree, c->num);
            astg::computeWeights(*c->tree, c->w);
            astg::buildProjectTables(*c->tree, c->num, c->w, c->tables);
            c->tablesReady = true;
            c->deviceReady = false;
        }
        const size_t nn = c->tree->lv[level].nnode;
        if (oct) std::memcpy(oct, c->num.oct[level].data(), nn * 4);
        if (tilt) std::memcpy(tilt, c->num.tilt[level].data(), nn * 4);
        if (n_total) *n_total = c->num.n;
        return ASTG_OK;
    });
}

int astg_get_weights(astg_ctx* c, int level, double* mac, double* tilt8) {
    int rc = checkLevel(c, level);
    if (rc) return rc;
    astg::Weights w;
    astg::computeWeights(*c->tree, w);
    if (mac) std::memcpy(mac, w.mac[level].data(), w.mac[level].size() * 8);
    if (tilt8) std::memcpy(tilt8, w.tilt[level].data(), w.tilt[level].size() * 8);
    return ASTG_OK;
}

int astg_set_velocity(astg_ctx* c, int level, const double* mac, const double* tilt8) {
    int rc = checkLevel(c, level);
    if (rc) return rc;
    return guarded(c, [&] {
        int r = ensureTables(c);
        if (r) return r;
        const int nn = c->tree->lv[level].nnode;
        if (c->device < 0) {
            if (mac) std::memcpy(c->hostMac[level].data(), mac, (size_t)nn * 3 * 8);
            if (tilt8) std::memcpy(c->hostTilt[level].data(), tilt8, (size_t)nn * 8 * 8);
            return ASTG_OK;
        }
        const size_t G = c->tables.G, off = c->tables.off[level];
        std::vector<double> buf((size_t)nn);
        for (int a = 0; a < 3 && mac; ++a) {
            for (int i = 0; i < nn; ++i) buf[i] = mac[(size_t)3 * i + a];
            if (hipMemcpy(c->proj->uMac() + a * G + off, buf.data(), (size_t)nn * 8, hipMemcpyHostToDevice) !=
                hipSuccess)
                return fail(c, ASTG_ERR_HIP, "velocity upload failed");
        }
        for (int k = 0; k < 8 && tilt8; ++k) {
            for (int i = 0; i < nn; ++i) buf[i] = tilt8[(size_t)8 * i + k];
            if (hipMemcpy(c->proj->uTilt() + k * G + off, buf.data(), (size_t)nn * 8, hipMemcpyHostToDevice) !=
                hipSuccess)
                return fail(c, ASTG_ERR_HIP, "velocity upload failed");
        }
        return ASTG_OK;
    });
}

int astg_get_velocity(astg_ctx* c, int level, double* mac, double* tilt8) {
    int rc = checkLevel(c, level);
    if (rc) return rc;
    return guarded(c, [&] {
        int r = ensureTables(c);
        if (r) return r;
        const int nn = c->tree->lv[level].nnode;
        if (c->device < 0) {
            if (mac) std::memcpy(mac, c->hostMac[level].data(), (size_t)nn * 3 * 8);
            if (tilt8) std::memcpy(tilt8, c->hostTilt[level].data(), (size_t)nn * 8 * 8);
            return ASTG_OK;
        }
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, ASTG_ERR_HIP, "stream sync failed");
        const size_t G = c->tables.G, off = c->tables.off[level];
        std::vector<double> buf((size_t)nn);
        for (int a = 0; a < 3 && mac; ++a) {
            if (hipMemcpy(buf.data(), c->proj->uMac() + a * G + off, (size_t)nn * 8, hipMemcpyDeviceToHost) !=
                hipSuccess)
                return fail(c, ASTG_ERR_HIP, "velocity download failed");
            for (int i = 0; i < nn; ++i) mac[(size_t)3 * i + a] = buf[i];
        }
        for (int k = 0; k < 8 && tilt8; ++k) {
            if (hipMemcpy(buf.data(), c->proj->uTilt() + k * G + off, (size_t)nn * 8, hipMemcpyDeviceToHost) !=
                hipSuccess)
                return fail(c, ASTG_ERR_HIP, "velocity download failed");
            for (int i = 0; i < nn; ++i) tilt8[(size_t)8 * i + k] = buf[i];
        }
        return ASTG_OK;
    });
}

int astg_project(astg_ctx* c, double dt, double rho, const astg_pcg_params* params, astg_pcg_result* result) {
    if (!c) return ASTG_ERR_ARG;
    int rc = needDevice(c);
    if (rc) return rc;
    if (!(dt > 0) || !(rho > 0)) return fail(c, ASTG_ERR_ARG, "dt and rho must be positive");
    astg_pcg_params prm;
    astg_pcg_params_default(&prm);
    if (params) prm = *params;
    return guarded(c, [&] { return runProject(c, dt, rho, prm, result); });
}

}  // extern "C"

namespace {
// Cube size of the block-colour ordering in finest cells, and the dependency
// levels per sweep group of its solver (ASTG_BLOCK_CELLS / ASTG_BLOCK_MERGE
// override them for experiments). Measured at 256^3 (mixed form, ms per
// projection): 8^3 cubes unmerged 164.8, 4^3 unmerged 173.0, 8^3 in level
// pairs 188.9, 4^3 in pairs 180.9, 2^3 in pairs 234.7; with about 50 k rows
// per level, substituting a level's dependencies costs more than the
// dependent step it saves.
int blockCells() {
    const char* e = std::getenv("ASTG_BLOCK_CELLS");
    const int v = e ? std::atoi(e) : 0;
    return v > 0 ? v : 8;
}
int blockMerge() {
    const char* e = std::getenv("ASTG_BLOCK_MERGE");
    const int v = e ? std::atoi(e) : 0;
    return v > 0 ? v : 1;
}

// The MIC(0) modification parameter a solve uses: the caller's, or with
// ASTG_TAU_AUTO the reference's 0.97 in FP64 (its factor and iterates) and
// plain IC(0) in the mixed form, which needs about half the iterations of
// MIC(0) at 0.97 on these systems (C2: 176 against 363 in block colours).
double micTau(const astg_pcg_params& prm) {
    if (prm.mic_tau >= 0) return prm.mic_tau;
    return prm.precision == ASTG_PREC_MIXED ? 0.0 : 0.97;
}

// ASTG_ORDER_AUTO: the block-colour form for the mixed precision, the
// reference numbering for FP64. The block-colour copy is built on the device
// (C2: 26 ms against a 79 ms solve), so it pays from the first solve on a
// pattern, also in a run whose tilts change every step.
bool blockOrdered(const astg_pcg_params& prm) {
    if (prm.ordering == ASTG_ORDER_BLOCK_COLOUR) return true;
    if (prm.ordering == ASTG_ORDER_REFERENCE) return false;
    return prm.precision == ASTG_PREC_MIXED;
}

// The block-colour copy of the projection system: permutation, pattern and
// solver plans, built on first use after a topology or tilt change. Keys on
// host threads, the sort and the permuted pattern on the device from A's
// device pattern; ASTG_CHECK_BLOCK_ORDER=1 also builds the host form
// (buildBlockOrder) and fails unless both agree entry for entry.
int ensureBlockOrder(astg_ctx* c) {
    if (c->blkReady) return ASTG_OK;
    astg::PhaseTimer tm("block order");
    const int n = c->pcg->n();
    const long long nnz = c->pcg->nnz();
    std::vector<int> keys;
    astg::blockKeys(*c->tree, c->num, blockCells(), keys);
    tm.mark("keys");
    int rc = c->blk.build(c->stream, n, nnz, keys.data(), c->pcg->devRowstart(), c->pcg->devCol(), c->err);
    if (rc) return rc;
    tm.mark("permutation and pattern");
    if (const char* g = std::getenv("ASTG_CHECK_BLOCK_ORDER"); g && std::atoi(g) != 0) {
        astg::BlockOrder h;
        astg::buildBlockOrder(*c->tree, c->num, c->tables.rowstart, c->tables.col, blockCells(), h);
        std::vector<int> perm(n), rs(n + 1), col(nnz), src(nnz);
        if (hipMemcpyAsync(perm.data(), c->blk.perm.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(rs.data(), c->blk.rowstart.p, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost, c->stream) !=
                hipSuccess ||
            hipMemcpyAsync(col.data(), c->blk.col.p, (size_t)nnz * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(src.data(), c->blk.src.p, (size_t)nnz * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return fail(c, ASTG_ERR_HIP, "block-order check: download failed");
        if (perm != h.perm || rs != h.rowstart || col != h.col || src != h.src)
            return fail(c, ASTG_ERR_HIP, "block order: the device build differs from the host construction");
        tm.mark("host check");
    }
    if (!c->pcgBlk) c->pcgBlk.reset(new astg::PcgDevice(c->stream));
    c->pcgBlk->setMergeLevels(blockMerge());
    rc = c->pcgBlk->setPatternDevice(n, nnz, c->blk.rowstart.as<int>(), c->blk.col.as<int>(), c->err);
    if (rc) return rc;
    tm.mark("solver pattern");
    c->blkReady = true;
    return ASTG_OK;
}

// Solves the assembled projection system (c->pcg's values and right-hand
// side, reference numbering) into c->pcg's solution vector, in the ordering
// the parameters select. `refactor` factors the values first; without it the
// factor of the last call with the same parameters is reused (the adjoint
// solves after a forward projection). `evFactored` (may be null) is recorded
// between factorisation and solve.
int solveAssembled(astg_ctx* c, const astg_pcg_params& prm, astg_pcg_result& res, bool refactor,
                   hipEvent_t evFactored) {
    if (prm.ordering < ASTG_ORDER_AUTO || prm.ordering > ASTG_ORDER_BLOCK_COLOUR)
        return fail(c, ASTG_ERR_ARG, "unknown ordering");
    int r;
    // a solve on the last factor keeps that factor's ordering
    const bool blk = refactor ? blockOrdered(prm) : c->factoredBlk;
    if (!blk) {
        if (refactor) {
            if ((r = c->pcg->factor(micTau(prm), prm.min_diag_ratio, c->err))) return r;
            c->factoredBlk = false;
        }
        if (evFactored) (void)hipEventRecord(evFactored, c->stream);
        c->lastPcg = c->pcg.get();
        return c->pcg->solve(prm, res, c->err);
    }
    if ((r = ensureBlockOrder(c))) return r;
    astg::PcgDevice& B = *c->pcgBlk;
    const int n = c->pcg->n();
    if ((r = astg::gatherPermuted(c->stream, c->pcg->nnz(), c->blk.src.as<int>(), refactor ? c->pcg->values() : nullptr,
                                  B.values(), n, c->blk.perm.as<int>(), c->pcg->rhs(), B.rhs(), c->err)))
        return r;
    if (refactor) {
        if ((r = B.factor(micTau(prm), prm.min_diag_ratio, c->err))) return r;
        c->factoredBlk = true;
    }
    if (evFactored) (void)hipEventRecord(evFactored, c->stream);
    c->lastPcg = &B;
    if ((r = B.solve(prm, res, c->err))) return r;
    return astg::scatterPermuted(c->stream, n, c->blk.perm.as<int>(), B.solution(), c->pcg->solution(), c->err);
}

int runProject(astg_ctx* c, double dt, double rho, const astg_pcg_params& prm, astg_pcg_result* result) {
    {
        int r = ensureTables(c);
        if (r) return r;
        if (c->tables.half_duplicates)
            return fail(c, ASTG_ERR_TOPOLOGY, "half-tilt list holds duplicates (the reference races on them)");
        // the solver's pattern must be the tables' (a failed tilt update drops both)
        if (c->pcg->n() != c->tables.n || c->pcg->nnz() != (long long)c->tables.col.size())
            return fail(c, ASTG_ERR_STATE, "solver pattern does not match the topology tables");
        astg_pcg_result res;
        std::memset(&res, 0, sizeof(res));
        hipEvent_t* ev = c->ev;
        (void)hipEventRecord(ev[0], c->stream);
        r = c->proj->assemble(c->pcg->values(), c->pcg->rhs(), c->pcg->devRowstart(), rho, dt, c->err);
        if (r) return r;
        (void)hipEventRecord(ev[1], c->stream);
        r = solveAssembled(c, prm, res, true, ev[2]);
        if (r) return r;
        (void)hipEventRecord(ev[3], c->stream);
        r = c->proj->correct(c->pcg->solution(), rho, dt, c->err);
        if (r) return r;
        (void)hipEventRecord(ev[4], c->stream);
        if (hipEventSynchronize(ev[4]) != hipSuccess) return fail(c, ASTG_ERR_HIP, "projection failed on the device");
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        res.ms_assemble = ms;
        (void)hipEventElapsedTime(&ms, ev[1], ev[2]);
        res.ms_factor = ms;
        (void)hipEventElapsedTime(&ms, ev[2], ev[3]);
        res.ms_solve = ms;
        (void)hipEventElapsedTime(&ms, ev[3], ev[4]);
        res.ms_velocity = ms;
        (void)hipEventElapsedTime(&ms, ev[0], ev[4]);
        res.ms_total = ms;
        if (result) *result = res;
        if (!res.converged) return fail(c, ASTG_ERR_NOT_CONVERGED, "PCG did not reach the tolerance");
        return ASTG_OK;
    }
}
}  // namespace

extern "C" {

int astg_field_layout(astg_ctx* c, long long* G, long long* offsets) {
    if (!c) return ASTG_ERR_ARG;
    if (!c->tree) return fail(c, ASTG_ERR_STATE, "no topology set");
    long long g = 0;
    for (int l = 0; l < c->tree->L; ++l) {
        if (offsets) offsets[l] = g;
        g += c->tree->lv[l].nnode;
    }
    if (offsets) offsets[c->tree->L] = g;
    if (G) *G = g;
    return ASTG_OK;
}

namespace {
// Copies between the caller's fields and the context's, ordered after the
// caller's stream and before its later work (device pointers), or blocking
// (host pointers). toContext: upload.
int moveFields(astg_ctx* c, const astg_fields* f, bool toContext) {
    if (!f) return fail(c, ASTG_ERR_ARG, "null fields");
    int rc = needAdvect(c);
    if (rc) return rc;
    const size_t G = c->tables.G;
    struct Pair {
        double* user;
        double* mine;
        size_t n;
    } pairs[6] = {{f->mac, c->proj->uMac(), 3 * G},
                  {f->tilt, c->proj->uTilt(), 8 * G},
                  {f->density_cell, c->adv->scalarPtr(0, 0), G},
                  {f->density_vert, c->adv->scalarPtr(0, 1), G},
                  {f->heat_cell, c->adv->scalarPtr(1, 0), G},
                  {f->heat_vert, c->adv->scalarPtr(1, 1), G}};
    if (!f->device_pointers) {
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, ASTG_ERR_HIP, "stream sync failed");
        for (const Pair& p : pairs) {
            if (!p.user) continue;
            if (hipMemcpy(toContext ? p.mine : p.user, toContext ? p.user : p.mine, p.n * 8,
                          toContext ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost) != hipSuccess)
                return fail(c, ASTG_ERR_HIP, "field copy failed");
        }
        return ASTG_OK;
    }
    hipStream_t us = static_cast<hipStream_t>(f->stream);
    // the context's stream waits for the caller's queued work ...
    if (hipEventRecord(c->ev[5], us) != hipSuccess || hipStreamWaitEvent(c->stream, c->ev[5], 0) != hipSuccess)
        return fail(c, ASTG_ERR_HIP, "stream ordering failed");
    for (const Pair& p : pairs) {
        if (!p.user) continue;
        hipPointerAttribute_t at;
        if (hipPointerGetAttributes(&at, p.user) != hipSuccess || at.type != hipMemoryTypeDevice ||
            at.device != c->device)
            return fail(c, ASTG_ERR_ARG, "a field pointer is not device memory of the context's GPU");
        if (hipMemcpyAsync(toContext ? p.mine : p.user, toContext ? p.user : p.mine, p.n * 8,
                           hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            return fail(c, ASTG_ERR_HIP, "field copy failed");
    }
    // ... and the caller's later work waits for the copies
    if (hipEventRecord(c->ev[5], c->stream) != hipSuccess || hipStreamWaitEvent(us, c->ev[5], 0) != hipSuccess)
        return fail(c, ASTG_ERR_HIP, "stream ordering failed");
    if (toContext) c->adv->fieldsCurrent();
    return ASTG_OK;
}
}  // namespace

int astg_fields_upload(astg_ctx* c, const astg_fields* f) {
    if (!c) return ASTG_ERR_ARG;
    return guarded(c, [&] { return moveFields(c, f, true); });
}

int astg_fields_download(astg_ctx* c, const astg_fields* f) {
    if (!c) return ASTG_ERR_ARG;
    return guarded(c, [&] { return moveFields(c, f, false); });
}

int astg_project_fields(astg_ctx* c, double dt, double rho, const astg_pcg_params* params, astg_fields* inout,
                        astg_pcg_result* result) {
    if (!c) return ASTG_ERR_ARG;
    if (!inout) return fail(c, ASTG_ERR_ARG, "null fields");
    astg_fields vel = *inout;  // the projection reads and writes the velocity only
    vel.density_cell = vel.density_vert = vel.heat_cell = vel.heat_vert = nullptr;
    int rc = astg_fields_upload(c, &vel);
    if (rc) return rc;
    rc = astg_project(c, dt, rho, params, result);
    if (rc && rc != ASTG_ERR_NOT_CONVERGED) return rc;
    const int rc2 = astg_fields_download(c, &vel);
    return rc2 ? rc2 : rc;
}

// ---- differentiable step ---------------------------------------------------
namespace {
// Copies n doubles between caller memory (host or device, in the caller's
// stream order) and device memory of the context.
int callerCopy(astg_ctx* c, double* dst, const double* src, size_t n, bool toContext, int devicePtrs, void* stream) {
    if (devicePtrs) {
        hipStream_t us = static_cast<hipStream_t>(stream);
        if (hipEventRecord(c->ev[5], toContext ? us : c->stream) != hipSuccess ||
            hipStreamWaitEvent(toContext ? c->stream : us, c->ev[5], 0) != hipSuccess)
            return fail(c, ASTG_ERR_HIP, "stream ordering failed");
        if (!toContext &&
            (hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
             hipEventRecord(c->ev[5], c->stream) != hipSuccess || hipStreamWaitEvent(us, c->ev[5], 0) != hipSuccess))
            return fail(c, ASTG_ERR_HIP, "copy failed");
        if (toContext && hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            return fail(c, ASTG_ERR_HIP, "copy failed");
        return ASTG_OK;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess ||
        hipMemcpy(dst, src, n * 8, toContext ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost) != hipSuccess)
        return fail(c, ASTG_ERR_HIP, "copy failed");
    return ASTG_OK;
}

int diffReady(astg_ctx* c, double dt, double rho) {
    int rc = needAdvect(c);
    if (rc) return rc;
    if (c->adv->hasCollider()) return fail(c, ASTG_ERR_STATE, "the differentiable step has no collider form");
    if (!c->adv->diffReady()) {
        rc = c->adv->diffPrepare(*c->proj, *c->pcg, dt, rho, c->err);
        if (rc) return rc;
    }
    return ASTG_OK;
}

// The forward of one differentiable step on the context's fields.
int diffForward(astg_ctx* c, double dt, double rho, const astg_pcg_params& prm, astg_pcg_result* result,
                bool advect) {
    int rc = c->adv->constrainAndGhost(*c->proj, c->err);
    if (rc) return rc;
    rc = runProject(c, dt, rho, prm, result);
    if (rc) return rc;
    rc = c->adv->postProject(*c->proj, c->err);
    if (rc) return rc;
    return advect ? c->adv->advectCarry(*c->proj, dt, c->err) : ASTG_OK;
}
}  // namespace

int astg_diff_state_size(astg_ctx* c, long long* n) {
    if (!c || !n) return ASTG_ERR_ARG;
    int rc = needAdvect(c);
    if (rc) return rc;
    *n = c->adv->diffStateLen();
    return ASTG_OK;
}

int astg_diff_get_state(astg_ctx* c, double* state, int device_pointers, void* stream) {
    if (!c || !state) return ASTG_ERR_ARG;
    int rc = needAdvect(c);
    if (rc) return rc;
    return guarded(c, [&] {
        const size_t n = (size_t)c->adv->diffStateLen();
        astg::DevBuf stage;
        double* dst = state;
        if (!device_pointers) {
            if (stage.alloc(n * 8) != hipSuccess) return fail(c, ASTG_ERR_HIP, "allocation failed");
            dst = stage.as<double>();
        } else if (hipEventRecord(c->ev[5], static_cast<hipStream_t>(stream)) != hipSuccess ||
                   hipStreamWaitEvent(c->stream, c->ev[5], 0) != hipSuccess) {
            return fail(c, ASTG_ERR_HIP, "stream ordering failed");
        }
        int r = c->adv->diffGetState(*c->proj, dst, c->err);
        if (r) return r;
        if (!device_pointers) return callerCopy(c, state, dst, n, false, 0, nullptr);
        if (hipEventRecord(c->ev[5], c->stream) != hipSuccess ||
            hipStreamWaitEvent(static_cast<hipStream_t>(stream), c->ev[5], 0) != hipSuccess)
            return fail(c, ASTG_ERR_HIP, "stream ordering failed");
        return ASTG_OK;
    });
}

int astg_diff_set_state(astg_ctx* c, const double* state, int device_pointers, void* stream) {
    if (!c || !state) return ASTG_ERR_ARG;
    int rc = needAdvect(c);
    if (rc) return rc;
    return guarded(c, [&] {
        const size_t n = (size_t)c->adv->diffStateLen();
        astg::DevBuf stage;
        const double* src = state;
        if (!device_pointers) {
            if (stage.alloc(n * 8) != hipSuccess) return fail(c, ASTG_ERR_HIP, "allocation failed");
            int r = callerCopy(c, stage.as<double>(), state, n, true, 0, nullptr);
            if (r) return r;
            src = stage.as<double>();
        } else if (hipEventRecord(c->ev[5], static_cast<hipStream_t>(stream)) != hipSuccess ||
                   hipStreamWaitEvent(c->stream, c->ev[5], 0) != hipSuccess) {
            return fail(c, ASTG_ERR_HIP, "stream ordering failed");
        }
        int r = c->adv->diffSetState(*c->proj, src, c->err);
        if (r) return r;
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, ASTG_ERR_HIP, "stream sync failed");
        return ASTG_OK;
    });
}

int astg_diff_step(astg_ctx* c, double dt, double rho, const astg_pcg_params* params, astg_pcg_result* result) {
    if (!c) return ASTG_ERR_ARG;
    if (!(dt > 0) || !(rho > 0)) return fail(c, ASTG_ERR_ARG, "dt and rho must be positive");
    astg_pcg_params prm;
    astg_pcg_params_default(&prm);
    if (params) prm = *params;
    return guarded(c, [&] {
        int rc = diffReady(c, dt, rho);
        if (rc) return rc;
        return diffForward(c, dt, rho, prm, result, true);
    });
}

int astg_diff_step_adjoint(astg_ctx* c, double dt, double rho, const astg_pcg_params* params, const double* state_in,
                           const double* adj_out, double* adj_in, int device_pointers, void* stream) {
    if (!c) return ASTG_ERR_ARG;
    if (!state_in || !adj_out || !adj_in) return fail(c, ASTG_ERR_ARG, "null array");
    if (!(dt > 0) || !(rho > 0)) return fail(c, ASTG_ERR_ARG, "dt and rho must be positive");
    astg_pcg_params prm;
    astg_pcg_params_default(&prm);
    if (params) prm = *params;
    return guarded(c, [&] {
        int rc = diffReady(c, dt, rho);
        if (rc) return rc;
        astg::AdvectDevice& adv = *c->adv;
        const size_t ns = (size_t)adv.diffStateLen();
        // the step's forward up to the advection, from its input state
        rc = astg_diff_set_state(c, state_in, device_pointers, stream);
        if (rc) return rc;
        astg_pcg_result res;
        rc = diffForward(c, dt, rho, prm, &res, false);
        if (rc) return rc;
        // the work vector: the output adjoint, then zeros for b and p
        double* x = adv.diffWork();
        const size_t n = (size_t)c->pcg->n();
        if (hipMemsetAsync(x + ns, 0, 2 * n * 8, c->stream) != hipSuccess)
            return fail(c, ASTG_ERR_HIP, "clear failed");
        if ((rc = callerCopy(c, x, adj_out, ns, true, device_pointers, stream))) return rc;
        // advect, post-projection chain, velocity correction (gathers p's adjoint)
        if ((rc = adv.diffAdjointAdvect(*c->proj, dt, c->err)) || (rc = adv.diffAdjointPost(c->err)) ||
            (rc = adv.diffAdjointVel(c->err)))
            return rc;
        // b's adjoint = A^-1 p's adjoint (A is symmetric)
        if (hipMemcpyAsync(c->pcg->rhs(), x + adv.diffOffP(), n * 8, hipMemcpyDeviceToDevice, c->stream) !=
                hipSuccess ||
            hipMemsetAsync(x + adv.diffOffP(), 0, n * 8, c->stream) != hipSuccess)
            return fail(c, ASTG_ERR_HIP, "copy failed");
        astg_pcg_result r2;
        std::memset(&r2, 0, sizeof(r2));
        rc = solveAssembled(c, prm, r2, false, nullptr);
        if (rc) return rc;
        if (!r2.converged) return fail(c, ASTG_ERR_NOT_CONVERGED, "adjoint PCG did not reach the tolerance");
        if (hipMemcpyAsync(x + adv.diffOffB(), c->pcg->solution(), n * 8, hipMemcpyDeviceToDevice, c->stream) !=
            hipSuccess)
            return fail(c, ASTG_ERR_HIP, "copy failed");
        // assembly of b, then the constraint and ghosts before the projection
        if ((rc = adv.diffAdjointAsm(c->err)) || (rc = adv.diffAdjointPre(c->err))) return rc;
        rc = callerCopy(c, adj_in, x, ns, false, device_pointers, stream);
        if (rc) return rc;
        if (!device_pointers && hipStreamSynchronize(c->stream) != hipSuccess)
            return fail(c, ASTG_ERR_HIP, "stream sync failed");
        return ASTG_OK;
    });
}

int astg_diff_apply(astg_ctx* c, int op, int transpose, double dt, double rho, const astg_pcg_params* params,
                    const double* x, double* y) {
    if (!c) return ASTG_ERR_ARG;
    if (!x || !y || op < 0 || op > 4) return fail(c, ASTG_ERR_ARG, "bad operator or null array");
    astg_pcg_params prm;
    astg_pcg_params_default(&prm);
    if (params) prm = *params;
    return guarded(c, [&] {
        int rc = diffReady(c, dt, rho);
        if (rc) return rc;
        astg::AdvectDevice& adv = *c->adv;
        const size_t ns = (size_t)adv.diffStateLen();
        if (!transpose) {
            rc = astg_diff_set_state(c, x, 0, nullptr);
            if (rc) return rc;
            astg_pcg_result res;
            switch (op) {
                case ASTG_OP_CONSTRAIN: rc = adv.constrainAndGhost(*c->proj, c->err); break;
                case ASTG_OP_POST: rc = adv.postProject(*c->proj, c->err); break;
                case ASTG_OP_SCALAR_GHOST: rc = adv.scalarGhost(0, c->err); break;
                case ASTG_OP_PROJECT: rc = runProject(c, dt, rho, prm, &res); break;
                case ASTG_OP_ADVECT: rc = adv.advectCarry(*c->proj, dt, c->err); break;
            }
            if (rc) return rc;
            return astg_diff_get_state(c, y, 0, nullptr);
        }
        // transpose: the work vector holds x (b and p regions cleared)
        double* w = adv.diffWork();
        const size_t n = (size_t)c->pcg->n();
        if (hipMemsetAsync(w + ns, 0, 2 * n * 8, c->stream) != hipSuccess) return fail(c, ASTG_ERR_HIP, "clear failed");
        if ((rc = callerCopy(c, w, x, ns, true, 0, nullptr))) return rc;
        switch (op) {
            case ASTG_OP_CONSTRAIN: rc = adv.diffAdjointPre(c->err); break;
            case ASTG_OP_POST: rc = adv.diffAdjointPost(c->err); break;
            case ASTG_OP_SCALAR_GHOST: rc = adv.diffAdjointGhost1(c->err); break;
            case ASTG_OP_ADVECT: rc = adv.diffAdjointAdvect(*c->proj, dt, c->err); break;
            case ASTG_OP_PROJECT: {
                // the projection's factor must be current: assemble and factor first
                astg_pcg_result res;
                double* tmp = adv.diffTmp();
                if (hipMemcpyAsync(tmp, w, ns * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
                    return fail(c, ASTG_ERR_HIP, "copy failed");
                rc = runProject(c, dt, rho, prm, &res);
                if (rc) return rc;
                if (hipMemcpyAsync(w, tmp, ns * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
                    hipMemsetAsync(w + ns, 0, 2 * n * 8, c->stream) != hipSuccess)
                    return fail(c, ASTG_ERR_HIP, "copy failed");
                if ((rc = adv.diffAdjointVel(c->err))) return rc;
                if (hipMemcpyAsync(c->pcg->rhs(), w + adv.diffOffP(), n * 8, hipMemcpyDeviceToDevice, c->stream) !=
                        hipSuccess ||
                    hipMemsetAsync(w + adv.diffOffP(), 0, n * 8, c->stream) != hipSuccess)
                    return fail(c, ASTG_ERR_HIP, "copy failed");
                astg_pcg_result r2;
                std::memset(&r2, 0, sizeof(r2));
                if ((rc = solveAssembled(c, prm, r2, false, nullptr))) return rc;
                if (hipMemcpyAsync(w + adv.diffOffB(), c->pcg->solution(), n * 8, hipMemcpyDeviceToDevice,
                                   c->stream) != hipSuccess)
                    return fail(c, ASTG_ERR_HIP, "copy failed");
                rc = adv.diffAdjointAsm(c->err);
                break;
            }
        }
        if (rc) return rc;
        return callerCopy(c, y, w, ns, false, 0, nullptr);
    });
}

int astg_get_system(astg_ctx* c, int* rowstart, int* col, double* val, double* b, double* p) {
    if (!c) return ASTG_ERR_ARG;
    if (c->device < 0) {
        // host-only context: the pattern is available, values are not
        if (val || b || p) return fail(c, ASTG_ERR_STATE, "values need a device context");
        int rc = guarded(c, [&] { return ensureTables(c); });
        if (rc) return rc;
        if (rowstart) std::memcpy(rowstart, c->tables.rowstart.data(), (c->tables.n + 1) * 4);
        if (col) std::memcpy(col, c->tables.col.data(), c->tables.col.size() * 4);
        return ASTG_OK;
    }
    int rc = needDevice(c);
    if (rc) return rc;
    if (!c->tablesReady || !c->deviceReady) return fail(c, ASTG_ERR_STATE, "no assembled system");
    const auto& t = c->tables;
    if (rowstart) std::memcpy(rowstart, t.rowstart.data(), (t.n + 1) * 4);
    if (col) std::memcpy(col, t.col.data(), t.col.size() * 4);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, ASTG_ERR_HIP, "stream sync failed");
    if (val && hipMemcpy(val, c->pcg->values(), t.col.size() * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(c, ASTG_ERR_HIP, "download failed");
    if (b && hipMemcpy(b, c->pcg->rhs(), (size_t)t.n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(c, ASTG_ERR_HIP, "download failed");
    if (p && hipMemcpy(p, c->pcg->solution(), (size_t)t.n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(c, ASTG_ERR_HIP, "download failed");
    return ASTG_OK;
}

int astg_get_mic_factor(astg_ctx* c, double* invdiag, unsigned* colstart, unsigned* rowindex, double* value,
                        long long* nnz_l) {
    if (!c) return ASTG_ERR_ARG;
    int rc = needDevice(c);
    if (rc) return rc;
    if (!c->lastPcg) return fail(c, ASTG_ERR_STATE, "no factor computed yet");
    std::vector<double> inv, val;
    std::vector<unsigned> cs, ri;
    if (c->lastPcg->downloadFactor(inv, cs, ri, val)) return fail(c, ASTG_ERR_HIP, "factor download failed");
    if (nnz_l) *nnz_l = (long long)val.size();
    if (invdiag) std::memcpy(invdiag, inv.data(), inv.size() * 8);
    if (colstart) std::memcpy(colstart, cs.data(), cs.size() * 4);
    if (rowindex) std::memcpy(rowindex, ri.data(), ri.size() * 4);
    if (value) std::memcpy(value, val.data(), val.size() * 8);
    return ASTG_OK;
}

int astg_pcg_solve_csr(astg_ctx* c, int n, const int* rowstart, const int* colindex, const double* value,
                       const double* b, double* x, const astg_pcg_params* params, astg_pcg_result* result) {
    if (!c) return ASTG_ERR_ARG;
    if (n <= 0 || !rowstart || !colindex || !value || !b || !x) return fail(c, ASTG_ERR_ARG, "null or empty input");
    int rc = needDevice(c);
    if (rc) return rc;
    astg_pcg_params prm;
    astg_pcg_params_default(&prm);
    if (params) prm = *params;
    return guarded(c, [&] {
        const long long nnz = rowstart[n];
        if (rowstart[0] != 0 || nnz < 0) return fail(c, ASTG_ERR_ARG, "bad rowstart");
        for (int i = 0; i < n; ++i)
            if (rowstart[i + 1] < rowstart[i]) return fail(c, ASTG_ERR_ARG, "rowstart not monotone");
        if (!c->csr) c->csr.reset(new astg::PcgDevice(c->stream));
        bool same = (int)c->csrRs.size() == n + 1 && (long long)c->csrCol.size() == nnz &&
                    std::memcmp(c->csrRs.data(), rowstart, (size_t)(n + 1) * 4) == 0 &&
                    std::memcmp(c->csrCol.data(), colindex, (size_t)nnz * 4) == 0;
        if (!same) {
            c->csrRs.clear();
            c->csrCol.clear();
            int r = c->csr->setPattern(n, rowstart, colindex, c->err);
            if (r) return r;
            c->csrRs.assign(rowstart, rowstart + n + 1);
            c->csrCol.assign(colindex, colindex + nnz);
        }
        astg_pcg_result res;
        std::memset(&res, 0, sizeof(res));
        hipEvent_t* ev = c->ev;
        if (hipMemcpyAsync(c->csr->values(), value, (size_t)nnz * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(c->csr->rhs(), b, (size_t)n * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return fail(c, ASTG_ERR_HIP, "upload failed");
        (void)hipEventRecord(ev[0], c->stream);
        int r = c->csr->factor(micTau(prm), prm.min_diag_ratio, c->err);
        if (r) return r;
        (void)hipEventRecord(ev[1], c->stream);
        r = c->csr->solve(prm, res, c->err);
        if (r) return r;
        (void)hipEventRecord(ev[2], c->stream);
        if (hipMemcpyAsync(x, c->csr->solution(), (size_t)n * 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return fail(c, ASTG_ERR_HIP, "download failed");
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        res.ms_factor = ms;
        (void)hipEventElapsedTime(&ms, ev[1], ev[2]);
        res.ms_solve = ms;
        (void)hipEventElapsedTime(&ms, ev[0], ev[2]);
        res.ms_total = ms;
        c->lastPcg = c->csr.get();
        if (result) *result = res;
        if (!res.converged) return fail(c, ASTG_ERR_NOT_CONVERGED, "PCG did not reach the tolerance");
        return ASTG_OK;
    });
}

namespace {

// Stacked block-diagonal form of a batch: system k owns rows [seg[k], seg[k+1]),
// padded with identity rows (zero right-hand side) to a multiple of 256.
struct StackedBatch {
    std::vector<int> seg, rs, col;
    std::vector<double> val, b;
    int N = 0;
    long long nnz = 0;
};

int stackBatch(astg_ctx* c, int count, const int* n, const int* const* rowstart, const int* const* colindex,
               const double* const* value, const double* const* b, StackedBatch& S) {
    constexpr int kChunk = 256;
    S.seg.assign(count + 1, 0);
    long long rows = 0, nnzAll = 0;
    for (int k = 0; k < count; ++k) {
        if (n[k] <= 0 || !rowstart[k] || !colindex[k] || !value[k] || !b[k])
            return fail(c, ASTG_ERR_ARG, "system " + std::to_string(k) + ": null or empty input");
        const int* rs = rowstart[k];
        if (rs[0] != 0) return fail(c, ASTG_ERR_ARG, "system " + std::to_string(k) + ": bad rowstart");
        for (int i = 0; i < n[k]; ++i)
            if (rs[i + 1] < rs[i])
                return fail(c, ASTG_ERR_ARG, "system " + std::to_string(k) + ": rowstart not monotone");
        const long long padded = ((long long)n[k] + kChunk - 1) / kChunk * kChunk;
        nnzAll += rs[n[k]] + (padded - n[k]);
        rows += padded;
        if (rows >= (1LL << 31) - kChunk || nnzAll >= (1LL << 31))
            return fail(c, ASTG_ERR_ARG, "batch exceeds 32-bit row or entry indices");
        S.seg[k + 1] = (int)rows;
    }
    S.N = (int)rows;
    S.nnz = nnzAll;
    S.rs.assign((size_t)S.N + 1, 0);
    S.col.assign((size_t)nnzAll, 0);
    S.val.assign((size_t)nnzAll, 0.0);
    S.b.assign((size_t)S.N, 0.0);
    long long p = 0;
    for (int k = 0; k < count; ++k) {
        const int* rs = rowstart[k];
        const int* cl = colindex[k];
        const int r0 = S.seg[k], nk = n[k];
        for (int i = 0; i < nk; ++i) {
            S.rs[(size_t)r0 + i] = (int)p;
            for (int q = rs[i]; q < rs[i + 1]; ++q, ++p) {
                const int j = cl[q];
                if (j < 0 || j >= nk)
                    return fail(c, ASTG_ERR_ARG, "system " + std::to_string(k) + ": column index out of range");
                S.col[(size_t)p] = r0 + j;
                S.val[(size_t)p] = value[k][q];
            }
        }
        std::memcpy(S.b.data() + r0, b[k], (size_t)nk * 8);
        for (int i = r0 + nk; i < S.seg[k + 1]; ++i, ++p) {  // identity padding rows
            S.rs[(size_t)i] = (int)p;
            S.col[(size_t)p] = i;
            S.val[(size_t)p] = 1.0;
        }
    }
    S.rs[(size_t)S.N] = (int)p;
    return ASTG_OK;
}

// Pattern-only set-up of the stacked system with `merge` levels per sweep
// group, skipped when the cached pattern is the same.
int batchPattern(astg_ctx* c, const StackedBatch& S, int merge, bool cache) {
    if (!c->batch) {
        c->batch.reset(new astg::PcgDevice(c->stream));
        c->batch->setRowOrder(true);  // per-system streaming kernels work on stacked rows
    }
    if (c->batchMerge == merge && c->batchSeg == S.seg && c->batchRs == S.rs && c->batchCol == S.col)
        return ASTG_OK;
    c->batchRs.clear();
    c->batchCol.clear();
    c->batchSeg.clear();
    c->batchMerge = 0;
    c->batch->setMergeLevels(merge);
    int r = c->batch->setPattern(S.N, S.rs.data(), S.col.data(), c->err);
    if (r) return r;
    if ((r = c->batch->setBatch(S.seg, c->err))) return r;
    if (cache) {
        c->batchRs = S.rs;
        c->batchCol = S.col;
        c->batchSeg = S.seg;
        c->batchMerge = merge;
    }
    return ASTG_OK;
}

// Uploads values and right-hand side, factors and solves; device times in ms.
int batchRun(astg_ctx* c, const StackedBatch& S, const astg_pcg_params& prm, std::vector<astg_pcg_result>& res,
             float& msFactor, float& msSolve) {
    hipEvent_t* ev = c->ev;
    if (hipMemcpyAsync(c->batch->values(), S.val.data(), (size_t)S.nnz * 8, hipMemcpyHostToDevice, c->stream) !=
            hipSuccess ||
        hipMemcpyAsync(c->batch->rhs(), S.b.data(), (size_t)S.N * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return fail(c, ASTG_ERR_HIP, "upload failed");
    (void)hipEventRecord(ev[0], c->stream);
    int r = c->batch->factor(micTau(prm), prm.min_diag_ratio, c->err);
    if (r) return r;
    (void)hipEventRecord(ev[1], c->stream);
    r = c->batch->solveBatch(prm, res, c->err);
    if (r) return r;
    (void)hipEventRecord(ev[2], c->stream);
    if (hipEventSynchronize(ev[2]) != hipSuccess) return fail(c, ASTG_ERR_HIP, "batch solve failed");
    (void)hipEventElapsedTime(&msFactor, ev[0], ev[1]);
    (void)hipEventElapsedTime(&msSolve, ev[1], ev[2]);
    return ASTG_OK;
}

bool validTuning(const astg_batch_tuning& t) {
    return t.merge_levels >= 1 && t.merge_levels <= 8 && t.stream_blocks_per_cu >= 0 &&
           t.stream_blocks_per_cu <= 64 && (t.lanes_per_row == 4 || t.lanes_per_row == 8 || t.lanes_per_row == 16);
}

}  // namespace

int astg_pcg_solve_batch(astg_ctx* c, int count, const int* n, const int* const* rowstart,
                         const int* const* colindex, const double* const* value, const double* const* b,
                         double* const* x, const astg_pcg_params* params, astg_pcg_result* results) {
    if (!c) return ASTG_ERR_ARG;
    if (count <= 0 || !n || !rowstart || !colindex || !value || !b || !x)
        return fail(c, ASTG_ERR_ARG, "null or empty batch");
    for (int k = 0; k < count; ++k)
        if (!x[k]) return fail(c, ASTG_ERR_ARG, "system " + std::to_string(k) + ": null output");
    int rc = needDevice(c);
    if (rc) return rc;
    astg_pcg_params prm;
    astg_pcg_params_default(&prm);
    if (params) prm = *params;
    return guarded(c, [&] {
        StackedBatch S;
        int r = stackBatch(c, count, n, rowstart, colindex, value, b, S);
        if (r) return r;
        if ((r = batchPattern(c, S, c->batchTune.merge_levels, true))) return r;
        c->batch->setBatchTuning(c->batchTune.stream_blocks_per_cu, c->batchTune.lanes_per_row);
        std::vector<astg_pcg_result> res;
        float msF = 0, msS = 0;
        if ((r = batchRun(c, S, prm, res, msF, msS))) return r;
        const double* xd = c->batch->solution();
        for (int k = 0; k < count; ++k)
            if (hipMemcpyAsync(x[k], xd + S.seg[k], (size_t)n[k] * 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
                return fail(c, ASTG_ERR_HIP, "download failed");
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, ASTG_ERR_HIP, "download failed");
        int unconverged = 0;
        for (int k = 0; k < count; ++k) {
            astg_pcg_result& R = res[k];
            R.n = n[k];
            R.nnz = rowstart[k][n[k]];
            R.ms_factor = msF;
            R.ms_solve = msS;
            R.ms_total = msF + msS;
            if (!R.converged) ++unconverged;
            if (results) results[k] = R;
        }
        if (unconverged)
            return fail(c, ASTG_ERR_NOT_CONVERGED,
                        std::to_string(unconverged) + " system(s) did not reach the tolerance");
        return ASTG_OK;
    });
}

void astg_batch_tuning_default(astg_batch_tuning* t) {
    if (!t) return;
    t->merge_levels = 2;
    t->stream_blocks_per_cu = 8;
    t->lanes_per_row = 8;
    t->ms_per_iteration = 0.0;
}

int astg_pcg_set_batch_tuning(astg_ctx* c, const astg_batch_tuning* t) {
    if (!c) return ASTG_ERR_ARG;
    if (!t || !validTuning(*t)) return fail(c, ASTG_ERR_ARG, "invalid batch tuning");
    c->batchTune = *t;
    return ASTG_OK;
}

int astg_pcg_tune_batch(astg_ctx* c, int count, const int* n, const int* const* rowstart, const int* const* colindex,
                        const double* const* value, const double* const* b, const astg_pcg_params* params,
                        int iterations, astg_batch_tuning* best) {
    if (!c) return ASTG_ERR_ARG;
    if (count <= 0 || !n || !rowstart || !colindex || !value || !b || !best)
        return fail(c, ASTG_ERR_ARG, "null or empty batch");
    if (iterations < 2 || iterations > 1000) return fail(c, ASTG_ERR_ARG, "iterations must be in [2, 1000]");
    int rc = needDevice(c);
    if (rc) return rc;
    astg_pcg_params prm;
    astg_pcg_params_default(&prm);
    if (params) prm = *params;
    // a fixed amount of work per configuration: no system converges
    prm.tol = 0.0;
    prm.max_iterations = iterations;
    prm.max_outer = 1;
    return guarded(c, [&] {
        StackedBatch S;
        int r = stackBatch(c, count, n, rowstart, colindex, value, b, S);
        if (r) return r;
        static const int kMerge[] = {1, 2, 3};
        static const int kBlocks[] = {4, 8, 16};
        static const int kLanes[] = {4, 8, 16};
        astg_batch_tuning t;
        astg_batch_tuning_default(&t);
        double bestMs = 1e300;
        for (int q : kMerge) {
            if ((r = batchPattern(c, S, q, false))) return r;
            for (int bl : kBlocks)
                for (int ln : kLanes) {
                    c->batch->setBatchTuning(bl, ln);
                    std::vector<astg_pcg_result> res;
                    float msF = 0, msS = 0;
                    double ms = 1e300;
                    for (int rep = 0; rep < 3; ++rep) {  // first run warms; best of the other two
                        if ((r = batchRun(c, S, prm, res, msF, msS))) return r;
                        if (rep) ms = std::min(ms, (double)msS);
                    }
                    if (ms < bestMs) {
                        bestMs = ms;
                        t.merge_levels = q;
                        t.stream_blocks_per_cu = bl;
                        t.lanes_per_row = ln;
                    }
                }
        }
        t.ms_per_iteration = bestMs / iterations;
        c->batchTune = t;
        if ((r = batchPattern(c, S, t.merge_levels, true))) return r;
        *best = t;
        return ASTG_OK;
    });
}

int astg_fields_snapshot(astg_ctx* c, int restore) {
    if (!c) return ASTG_ERR_ARG;
    int rc = needDevice(c);
    if (rc) return rc;
    return guarded(c, [&] {
        int r = ensureTables(c);
        if (r) return r;
        return c->proj->snapshot(restore != 0, c->err);
    });
}

int astg_kernel_time(astg_ctx* c, int which, int reps, double* avg_ms, double* bytes) {
    if (!c || reps <= 0 || !avg_ms || !bytes) return fail(c, ASTG_ERR_ARG, "bad arguments");
    int rc = needDevice(c);
    if (rc) return rc;
    if (!c->deviceReady || !c->lastPcg) return fail(c, ASTG_ERR_STATE, "run astg_project first");
    return guarded(c, [&] {
        double ms = 0, b = 0;
        if (which == ASTG_K_ASSEMBLE) {
            hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
            double tot = 0;
            for (int i = 0; i < reps; ++i) {
                (void)hipEventRecord(e0, c->stream);
                int r = c->proj->assemble(c->pcg->values(), c->pcg->rhs(), c->pcg->devRowstart(), 2.0, 1.0, c->err);
                if (r) return r;
                (void)hipEventRecord(e1, c->stream);
                if (hipEventSynchronize(e1) != hipSuccess) return fail(c, ASTG_ERR_HIP, "assembly failed");
                float f = 0;
                (void)hipEventElapsedTime(&f, e0, e1);
                tot += f;
            }
            ms = tot / reps;
            b = c->proj->assembleBytes();
        } else {
            int r = c->lastPcg->benchKernel(which, reps, ms, b, c->err);
            if (r) return r;
        }
        *avg_ms = ms;
        *bytes = b;
        return ASTG_OK;
    });
}

int astg_sweep_timeline(astg_ctx* c, int which, long long* t_out, int* level_out) {
    if (!c || !t_out || !level_out) return fail(c, ASTG_ERR_ARG, "bad arguments");
    int rc = needDevice(c);
    if (rc) return rc;
    if (!c->deviceReady || !c->lastPcg) return fail(c, ASTG_ERR_STATE, "run astg_project first");
    return guarded(c, [&] {
        std::vector<long long> t;
        std::vector<int> lev;
        int r = c->lastPcg->sweepTimeline(which, t, lev, c->err);
        if (r) return r;
        std::copy(t.begin(), t.end(), t_out);
        std::copy(lev.begin(), lev.end(), level_out);
        return (int)ASTG_OK;
    });
}

// ---- post-projection step and advection ---------------------------------
namespace {
int needAdvect(astg_ctx* c) {
    if (!c) return ASTG_ERR_ARG;
    int rc = needDevice(c);
    if (rc) return rc;
    return guarded(c, [&] { return ensureTables(c); });
}
}  // namespace

int astg_set_scalar(astg_ctx* c, int field, int level, const double* oct, const double* tilt) {
    int rc = checkLevel(c, level);
    if (rc) return rc;
    if (!oct || !tilt) return fail(c, ASTG_ERR_ARG, "null field array");
    rc = needAdvect(c);
    if (rc) return rc;
    return c->adv->setScalar(field, level, oct, tilt, c->err);
}

int astg_get_scalar(astg_ctx* c, int field, int level, double* oct, double* tilt) {
    int rc = checkLevel(c, level);
    if (rc) return rc;
    if (!oct || !tilt) return fail(c, ASTG_ERR_ARG, "null field array");
    rc = needAdvect(c);
    if (rc) return rc;
    return c->adv->getScalar(field, level, oct, tilt, c->err);
}

int astg_post_project(astg_ctx* c) {
    int rc = needAdvect(c);
    if (rc) return rc;
    return c->adv->postProject(*c->proj, c->err);
}

int astg_scalar_update_ghost(astg_ctx* c, int field) {
    int rc = needAdvect(c);
    if (rc) return rc;
    return c->adv->scalarGhost(field, c->err);
}

int astg_get_dual(astg_ctx* c, int level, double* oct3, double* tilt3) {
    int rc = checkLevel(c, level);
    if (rc) return rc;
    if (!oct3 || !tilt3) return fail(c, ASTG_ERR_ARG, "null array");
    rc = needAdvect(c);
    if (rc) return rc;
    return c->adv->getDual(level, oct3, tilt3, c->err);
}

int astg_sample_points(astg_ctx* c, int what, int n, const double* xyz, double dt, double* out) {
    if (n < 0 || (n > 0 && (!xyz || !out))) return fail(c, ASTG_ERR_ARG, "bad point arrays");
    int rc = needAdvect(c);
    if (rc) return rc;
    return c->adv->samplePoints(*c->proj, what, n, xyz, dt, out, c->err);
}

int astg_sample_points_noalloc(astg_ctx* c, int what, int n, const double* xyz, double dt, double* out) {
    if (!c) return ASTG_ERR_ARG;
    if (n < 0 || (n > 0 && (!xyz || !out))) return fail(c, ASTG_ERR_ARG, "bad point arrays");
    int rc = needDevice(c);
    if (rc) return rc;
    // no set-up here: that would allocate
    if (!c->tablesReady || !c->deviceReady || !c->adv || !c->proj)
        return fail(c, ASTG_ERR_STATE, "fields not on the device yet (set or project a velocity first)");
    return c->adv->samplePointsStaged(*c->proj, what, n, xyz, dt, out, astg::noallocScratch(c->device), c->err);
}

int astg_apply_sphere_source(astg_ctx* c, double cx, double cy, double cz, double radius, double density,
                             double heat, double speed, int update_ghost) {
    if (!(radius > 0)) return fail(c, ASTG_ERR_ARG, "radius must be positive");
    int rc = needAdvect(c);
    if (rc) return rc;
    return c->adv->applySource(*c->proj, cx, cy, cz, radius, density, heat, speed, update_ghost != 0, c->err);
}

int astg_apply_sphere_source_velocity(astg_ctx* c, double cx, double cy, double cz, double radius, double density,
                                      double heat, const double* velocity, int update_ghost) {
    if (!(radius > 0)) return fail(c, ASTG_ERR_ARG, "radius must be positive");
    if (!velocity) return fail(c, ASTG_ERR_ARG, "null velocity");
    int rc = needAdvect(c);
    if (rc) return rc;
    return c->adv->applySourceV(*c->proj, cx, cy, cz, radius, density, heat, velocity, 0.0, update_ghost != 0,
                                c->err);
}

int astg_split_distance(int R, int levels, int dist_shift, const float* dist, const double* band, int* split_out) {
    if (R <= 0 || levels <= 0 || levels > 15 || (R >> (levels - 1)) < 2 || (R % (1 << (levels - 1))) != 0 || !dist ||
        (levels > 1 && !band) || !split_out || dist_shift < 0 || dist_shift >= levels)
        return ASTG_ERR_ARG;
    try {
        std::vector<int> split;
        astg::splitDistance(R, levels, dist_shift, dist, band, split);
        astg::balanceSplit(R, levels, split);
        std::memcpy(split_out, split.data(), split.size() * sizeof(int));
    } catch (...) {
        return ASTG_ERR_ARG;
    }
    return ASTG_OK;
}

int astg_wipe_boundary(astg_ctx* c) {
    int rc = needAdvect(c);
    if (rc) return rc;
    return c->adv->wipeBoundary(c->err);
}

int astg_add_force(astg_ctx* c, double dt, double smoke_density_factor, double temperature_factor) {
    if (!(dt >= 0)) return fail(c, ASTG_ERR_ARG, "dt must be non-negative");
    int rc = needAdvect(c);
    if (rc) return rc;
    return c->adv->addForce(*c->proj, dt, smoke_density_factor, temperature_factor, c->err);
}

int astg_adjust_e(astg_ctx* c, double d_threshold, double dt, double v_threshold, double max_threshold) {
    if (!(dt >= 0)) return fail(c, ASTG_ERR_ARG, "dt must be non-negative");
    int rc = needAdvect(c);
    if (rc) return rc;
    return guarded(c, [&] { return adjustTilts(c, d_threshold, dt, v_threshold, max_threshold); });
}

int astg_set_tilt_adaptivity(astg_ctx* c, int on) {
    if (!c) return ASTG_ERR_ARG;
    c->tiltAdaptivity = on != 0;
    return ASTG_OK;
}

int astg_advect(astg_ctx* c, double dt) {
    if (!(dt > 0)) return fail(c, ASTG_ERR_ARG, "dt must be positive");
    int rc = needAdvect(c);
    if (rc) return rc;
    // advect(dt) opens with adjustE(0.02, dt, EPSILON, 1000) when tilt
    // adaptivity is on (octree_omnigrid_solver3.cpp:852-853)
    if (c->tiltAdaptivity) {
        rc = guarded(c, [&] {
            return adjustTilts(c, 0.02, dt, std::numeric_limits<double>::epsilon(), 1000.0);
        });
        if (rc) return rc;
    }
    return c->adv->advect(*c->proj, dt, c->err);
}

int astg_step(astg_ctx* c, double dt, double rho, const astg_pcg_params* params, astg_pcg_result* result) {
    if (!(dt > 0) || !(rho > 0)) return fail(c, ASTG_ERR_ARG, "dt and rho must be positive");
    int rc = needAdvect(c);
    if (rc) return rc;
    astg_pcg_params prm;
    astg_pcg_params_default(&prm);
    if (params) prm = *params;
    // onUpdate (octree_omnigrid_solver3.cpp:222-249) without addForce: constrain
    // and ghost, project, constrain, ghost, dual grid, advect
    rc = c->adv->constrainAndGhost(*c->proj, c->err);
    if (rc) return rc;
    rc = guarded(c, [&] { return runProject(c, dt, rho, prm, result); });
    if (rc) return rc;
    rc = c->adv->postProject(*c->proj, c->err);
    if (rc) return rc;
    return astg_advect(c, dt);
}

}  // extern "C"

int astg_write_frame(astg_ctx* c, const char* layout, const char* tilt, const char* density, const char* heat,
                     const char* velocity) {
    if (!c) return ASTG_ERR_ARG;
    if (!c->tree) return fail(c, ASTG_ERR_STATE, "no topology set");
    return guarded(c, [&] {
        const astg::Octree& o = *c->tree;
        if (layout && !astg::writeLayoutFile(layout, o, c->err)) return ASTG_ERR_IO;
        if (tilt && !astg::writeTiltFile(tilt, o, c->err)) return ASTG_ERR_IO;
        const char* sc[2] = {density, heat};
        for (int f = 0; f < 2; ++f) {
            if (!sc[f]) continue;
            astg::OmniScalarFrame fr;
            fr.oct.assign(o.L, {});
            fr.tilt.assign(o.L, {});
            for (int l = 0; l < o.L; ++l) {
                fr.oct[l].resize(o.lv[l].nnode);
                fr.tilt[l].resize(o.lv[l].nnode);
                int r = astg_get_scalar(c, f, l, fr.oct[l].data(), fr.tilt[l].data());
                if (r) return r;
            }
            if (!astg::writeScalarFile(sc[f], fr, c->err)) return ASTG_ERR_IO;
        }
        if (velocity) {
            astg::OmniVelocityFrame fr;
            fr.mac.assign(o.L, {});
            fr.tilt8.assign(o.L, {});
            for (int l = 0; l < o.L; ++l) {
                fr.mac[l].resize((size_t)o.lv[l].nnode * 3);
                fr.tilt8[l].resize((size_t)o.lv[l].nnode * 8);
                int r = astg_get_velocity(c, l, fr.mac[l].data(), fr.tilt8[l].data());
                if (r) return r;
            }
            if (!astg::writeVelocityFile(velocity, fr, c->err)) return ASTG_ERR_IO;
        }
        return ASTG_OK;
    });
}

int astg_read_frame(astg_ctx* c, const char* layout, const char* tilt, const char* density, const char* heat,
                    const char* velocity) {
    if (!c) return ASTG_ERR_ARG;
    if (!c->tree) return fail(c, ASTG_ERR_STATE, "no topology set");
    if ((density || heat || velocity) && c->device < 0)
        return fail(c, ASTG_ERR_STATE, "fields need a device context");
    return guarded(c, [&] {
        astg::Octree& o = *c->tree;
        if (layout && !astg::checkLayoutFile(layout, o, c->err)) return ASTG_ERR_IO;
        // every file is read and checked before any state changes
        astg::Octree staged;
        const bool newTilt = tilt != nullptr;
        std::vector<std::vector<astg::TiltNode>> tiltState;
        if (newTilt) {
            staged.L = o.L;
            staged.lv = o.lv;
            staged.tilt = o.tilt;
            if (!astg::readTiltFile(tilt, staged, c->err)) return ASTG_ERR_IO;
            tiltState.swap(staged.tilt);
        }
        const char* sc[2] = {density, heat};
        astg::OmniScalarFrame sf[2];
        for (int f = 0; f < 2; ++f)
            if (sc[f] && !astg::readScalarFile(sc[f], o, sf[f], c->err)) return ASTG_ERR_IO;
        astg::OmniVelocityFrame vf;
        if (velocity && !astg::readVelocityFile(velocity, o, vf, c->err)) return ASTG_ERR_IO;
        if (newTilt) {
            o.tilt.swap(tiltState);
            if (c->deviceReady) {
                // a failed update puts the previous tilt state back (the
                // cached tables are dropped and rebuilt on the next call)
                int r;
                try {
                    r = retopologize(c);
                } catch (...) {
                    o.tilt.swap(tiltState);
                    throw;
                }
                if (r) {
                    o.tilt.swap(tiltState);
                    return r;
                }
            } else {
                c->tablesReady = false;
            }
        }
        if (c->device < 0) return ASTG_OK;
        for (int f = 0; f < 2; ++f)
            if (sc[f])
                for (int l = 0; l < o.L; ++l) {
                    int r = astg_set_scalar(c, f, l, sf[f].oct[l].data(), sf[f].tilt[l].data());
                    if (r) return r;
                }
        if (velocity) {
            for (int l = 0; l < o.L; ++l) {
                int r = astg_set_velocity(c, l, vf.mac[l].data(), vf.tilt8[l].data());
                if (r) return r;
            }
            // restart's velocity ghosts, dual grid and dual ghosts, no constraint
            int r = c->adv->postProject(*c->proj, c->err, false);
            if (r) return r;
        }
        // restart loads the advection temporaries from the same files
        const bool sl[2] = {density != nullptr, heat != nullptr};
        if (int r = c->adv->mirrorTemporaries(*c->proj, sl, velocity != nullptr, c->err)) return r;
        if (newTilt || density || heat || velocity) c->adv->fieldsCurrent();
        return ASTG_OK;
    });
}
