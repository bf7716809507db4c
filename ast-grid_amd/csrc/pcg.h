// Device MIC(0)-preconditioned conjugate gradients for the projection system.
//
// Algorithm: the reference's PCGSolver (include/pcgsolver/pcg_solver.h:366-426)
// with the reference's MIC(0) factor (pcg_solver.h:89-190) and triangular
// solves (pcg_solver.h:211-253), restated for a GPU:
//   * the dependency levels of tril(A) are computed on the device by a
//     sync-free pass;
//   * the factorisation is level-synchronous (one launch per level, one row per
//     lane, left-looking per row with the reference's accumulation order);
//   * the triangular sweeps are sync-free persistent kernels over merged
//     stencils (sweep_merge.h): each row publishes its value with an
//     agent-scope atomic store over a signalling-NaN sentinel that dependants
//     poll;
//   * dots, axpys and the convergence test are fused into the SpMV / update /
//     sweep kernels with wave64 reductions; scalars stay on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/astg.h"
#include "device_common.h"

namespace astg {
class PhaseTimer;

// Device allocations made through DevBuf so far in this process, and the
// time spent in them (hipFree of the old buffer included).
void allocStats(long long& count, double& ms);

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;  // capacity: at least what the last alloc asked for
    // One owner per allocation: copies are not allowed, moves hand it over.
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf();
    hipError_t alloc(size_t b);
    void release();  // frees now (the destructor does the same)
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

// Rows in padded level order (or padded super-level order, see
// sweep_merge.h), plus one gate row per 64-row segment.
struct SweepPlan {
    const int* order;  // P entries, -1 = padding
    const int* gateF;  // per segment: position of a row of group G-dist (forward), -1 none
    const int* gateB;  // per segment: position of a row of group G+dist (backward), -1 none
    int P;
    int nchunks;       // P / 256
    int pollCap;       // largest sleep between polling rounds (units of 64 cycles)
    int gateCap;       // largest sleep between gate polls (units of 128 cycles)
    long long* trace;  // optional per-row timeline (3 per row), null in normal runs
    // batched form (null for one system): the system of every 256-row chunk
    // of the stacked rows, and the per-system state (BatchSlot layout)
    const int* chunkSys;
    const double* sysState;
    // 1: the vectors r and z and the pivots are stored in sweep order (index
    // = sweep position), and forward r-term keys are sweep positions; 0: row order
    int vecPos;
};

// Merged sweep stencil of one direction, by sweep position (sweep_merge.h).
template <typename T>
struct StencilView {
    const int* start;  // P + 1
    const int* nr;     // r-terms of the row (the first entries)
    const int* idx;    // per entry: r-term key, y-term sweep position
    const T* coef;     // per entry
};
struct DevStencil {
    DevBuf start, nr, idx, ustart, upd, coef, coef32;
    long long E = 0;  // entries (r-terms plus y-terms) over all rows
};

// Per-block partials of the reductions, consumed by the next kernel.
struct Parts {
    double* sq;    // gs partials of s.q
    double* res;   // gs partials of |r| (max) or r.r (sum)
    double* rho0;  // gb partials of rho (even iterations)
    double* rho1;  // gb partials of rho (odd iterations)
    int gs, gb;
};

// Batched form: the layout of K stacked systems and their per-system state.
enum BatchSlot {
    B_BNORM = 0,  // ||b|| of the system (inner right-hand side)
    B_TOL,        // absolute tolerance of the inner solve
    B_RES,        // last residual norm
    B_DONE,       // 0 running; 1 converged, 2 breakdown, 3 iteration limit, 4 not solved this time,
                  // 5 NaN or infinity in b or the residual
    B_ITER,       // iterations performed
    B_RHO,        // z.r
    B_ALPHA,      // rho / s.q
    B_BETA,       // rho_new / rho
    B_TOLREL,     // relative tolerance of the inner solve (set by the host)
    B_MAXIT,      // iteration limit of the inner solve (set by the host)
    B_SKIP,       // nonzero: the system takes no part in the inner solve (set by the host)
    B_SCALE,      // mixed form: 1 / outer residual norm (inner right-hand side scaling)
    B_OUTER,      // mixed form: FP64 outer residual norm
    B_COUNT = 16
};
struct BatchView {
    const int* chunkSeg;  // per 256-row chunk: its system
    const int* segChunk;  // K + 1: first chunk of each system
    double* st;           // K x B_COUNT
    double* part;         // one partial per chunk
    int* ndone;           // systems finished in the current inner solve
    int K, nchunks;
};

class PcgDevice {
public:
    explicit PcgDevice(hipStream_t s) : stream_(s) {}
    ~PcgDevice();

    // Uploads a CSR pattern (host arrays) and builds everything that depends on
    // the pattern only, on the device: diagonal / upper split, transpose map
    // and structural checks, dependency levels, the padded level and sweep
    // orders, the symbolic plans and, for single-system solves, A's pattern
    // permuted to sweep order.
    int setPattern(int n, const int* rowstart, const int* col, std::string& err);
    // The same from a pattern already in device memory (rowstart n+1, col
    // nnz, on this object's stream): no host copy is made until one is needed.
    int setPatternDevice(int n, long long nnz, const int* dRowstart, const int* dCol, std::string& err);
    // Device pointer to the FP64 value array (nnz) for in-place assembly.
    double* values() { return val_.as<double>(); }
    double* rhs() { return b_.as<double>(); }
    double* solution() { return x_.as<double>(); }
    int n() const { return n_; }
    long long nnz() const { return nnz_; }
    int depth() const { return depth_; }
    const int* devRowstart() const { return rs_.as<int>(); }
    const int* devCol() const { return col_.as<int>(); }

    // MIC(0) factorisation of the current values (FP64).
    int factor(double tau, double ratio, std::string& err);
    // Solves with b = rhs(), x = solution(); both device FP64 in row order
    // (the solve gathers b into sweep order and scatters x back).
    // (Both solves retry once with cooperative sweep launches if an ordinary
    // launch left a sweep block waiting for one that was not resident.)
    int solve(const astg_pcg_params& prm, astg_pcg_result& res, std::string& err);

    // Times `reps` launches of one kernel (ASTG_K_*) with events on the stream.
    int benchKernel(int which, int reps, double& avgMs, double& bytes, std::string& err);
    long long nnzL() const { return (nnz_ - n_) / 2; }

    // Timeline of one sweep launch (ASTG_K_FWD64 / ASTG_K_BWD64): per row the
    // device wall clock (100 MHz) when its wavefront passed the gate, when all
    // its dependencies were seen, and after it published its value; plus the
    // dependency level of every row.
    int sweepTimeline(int which, std::vector<long long>& t, std::vector<int>& lev, std::string& err);
    // Reference-form copy of the factor (host arrays), for parity checks.
    int downloadFactor(std::vector<double>& invdiag, std::vector<unsigned>& colstart, std::vector<unsigned>& rowindex,
                       std::vector<double>& value);
    // Batched form (pcg_batch.hip). The pattern set by setPattern is the
    // block-diagonal stack of K systems; system k owns rows
    // [segRow[k], segRow[k+1]), every segRow a multiple of 256 (rows past the
    // system's own are identity rows with a zero right-hand side).
    int setBatch(const std::vector<int>& segRow, std::string& err);
    int batchCount() const { return K_; }
    // Launch shape of the batched streaming kernels: blocks per compute unit
    // (0: the single-system default) and SpMV lanes per row (4, 8 or 16).
    void setBatchTuning(int blocksPerCu, int lanesPerRow);
    // Keep the PCG vectors in row order (the batched form, whose streaming
    // kernels work per system on stacked rows). Takes effect at setPattern.
    void setRowOrder(bool on) { rowOrderForced_ = on; }
    // Levels per merged sweep group used by the next setPattern (0: default).
    void setMergeLevels(int q) { mergeLevels_ = q; }
    int mergeLevels() const { return merge_; }
    // Solves every system of the stack with its own PCG scalars, convergence
    // test and iteration count (rhs(), solution(), values() hold the stack).
    int solveBatch(const astg_pcg_params& prm, std::vector<astg_pcg_result>& res, std::string& err);
    bool cooperativeSweeps() const { return coop_; }

private:
    // x: FP64 solution (the FP32 iteration accumulates its steps in FP64).
    // b64, replaceEvery: FP32 iteration only; every replaceEvery iterations
    // the recurrence residual becomes b64 - A x in FP64 (0: never).
    template <typename T>
    int pcgInner(const T* A, const T* invd, const T* b, double* x, double tolRel, int maxIt, int norm, int& iters,
                 double& resOut, bool& conv, std::string& err, const double* b64, int replaceEvery);
    // Sweep k of a solve (0: the first preconditioning, it + 1: iteration it)
    // writes y buffer k & 1; the backward sweep re-arms buffer (k + 1) & 1.
    template <typename T>
    int launchFwd(const T* invd, const T* r, int k, int mode, int it, int maxIt, int norm, double tolRel,
                  std::string& err);
    template <typename T>
    int launchBwd(const T* invd, const T* r, T* z, int k, int parityOut, std::string& err);
    // Batched inner PCG over every system not marked B_SKIP (state set by the
    // host beforehand); per-system results stay in the device state.
    template <typename T>
    int pcgBatchInner(const T* A, const T* invd, const T* b, double* x, int maxIt, int norm, std::string& err,
                      const double* b64, int replaceEvery);
    int batchState(std::vector<double>& st, bool toDevice, std::string& err);
    BatchView bview() const;
    int batchGrid() const;
    bool isSweepKernel(const void* f) const;  // k_fwd / k_bwd_dot (cooperative graph nodes)
    int solveOnce(const astg_pcg_params& prm, astg_pcg_result& res, std::string& err);
    int solveBatchOnce(const astg_pcg_params& prm, std::vector<astg_pcg_result>& res, std::string& err);
    // After a spin-limit failure under ordinary sweep launches: cooperative
    // launches from now on, error bits cleared; true if the solve should be retried.
    bool fallBackToCooperative(std::string& err);
    // symbolic plans (factorisation programs, both merged stencils) built on
    // the device from the uploaded pattern and sweep order (plan_device.hip);
    // the host planners serve when hostPlan_ is set
    int planDevice(int maxY, std::string& err);
    // Sweep-ordered solve: permuted CSR, forward r-term keys as positions
    // (once per pattern); values, pivots and rhs gathered per solve.
    int permuteSetup(std::string& err);
    int prepareSolve(std::string& err);
    int vlen() const { return perm_ ? PS_ : n_; }  // length of the PCG vectors
    const int* sRs() const { return perm_ ? prs_.as<int>() : rs_.as<int>(); }
    const int* sCol() const { return perm_ ? pcol_.as<int>() : col_.as<int>(); }
    double* sVal64() const { return perm_ ? valP_.as<double>() : val_.as<double>(); }
    double* sInvd64() const { return perm_ ? invdP_.as<double>() : invd_.as<double>(); }
    double* sB() const { return perm_ ? bP_.as<double>() : b_.as<double>(); }
    double* sX() const { return perm_ ? xP_.as<double>() : x_.as<double>(); }
    template <typename T>
    void launchSpmv(int n, const int* rs, const int* col, const T* A, const T* s, T* q, const Parts& Pt,
                    const double* scal);
    int launchMic(double tau, double ratio, std::string& err);
    int launchMerge(std::string& err);
    int convert32(std::string& err);
    template <typename T>
    StencilView<T> view(const DevStencil& s) const;
    template <typename T>
    T* ybuf(int k) const { return y_.as<T>() + (size_t)(k & 1) * PS_; }
    int syncScalars(std::string& err);
    int checkErrors(const char* where, std::string& err);
    template <typename T>
    int fwdGrid() const { return sizeof(T) == 8 ? gF64_ : gF32_; }
    template <typename T>
    int bwdGrid() const { return sizeof(T) == 8 ? gB64_ : gB32_; }
    SweepPlan splan() const;  // sweeps: rows in super-level order
    Parts parts(int gb) const;

    hipStream_t stream_;
    int n_ = 0;
    long long nnz_ = 0;
    int depth_ = 0;
    std::vector<int> hrs_, hcol_;
    bool hostPatternValid_ = true;  // hrs_ / hcol_ hold the current pattern
    int ensureHostPattern(std::string& err);
    int setPatternImpl(int n, long long nnz, const int* dRowstart, const int* dCol, PhaseTimer& tm,
                       std::string& err);
    DevBuf rs_, col_, lend_, ubeg_, tpos_, order_, lev_;  // pattern split, level order (factorisation)
    std::vector<int> levRange_;             // per level: [first, last + 1) position in the level order
    DevBuf micStart_, micProg_, live_;      // factorisation programs (mic_plan.h), eliminated columns
    DevBuf sorder_, sgateF_, sgateB_;       // sweep order (super-levels) and its gates
    DevStencil stF_, stB_;                  // merged forward / backward stencils
    DevBuf val_, F_, invd_;                 // FP64 A, factor (symmetric copy), 1/pivot
    DevBuf val32_, invd32_;                 // FP32 copies for the mixed form
    DevBuf b_, x_;                          // FP64 rhs / solution
    DevBuf r_, z_, s_, q_;                  // work vectors, vlen() entries (sized for FP64; reused as FP32)
    DevBuf y_, zp_;                         // sweep vectors in sweep order (2 x PS and PS)
    DevBuf r64_, d64_, rhs64_;              // mixed form outer vectors
    int mixedReplace_ = 100;                // mixed form: residual replacement period
    double mixedEta_ = 0.9;                 // mixed form: inner tolerance / outer target
    DevBuf scal_, cnt_, part_;              // device scalars, counters, partials
    double* hscal_ = nullptr;               // pinned mirror of the scalars
    long long* hplan_ = nullptr;            // pinned: set-up counters and plan sizes
    DevBuf sortKeys_, sortIdx_, sortRows_;  // set-up: level sort keys, row indices, rows in level order
    DevBuf lvlStart_, grpStart_, grpFirst_; // set-up: first sorted row of each level; group layout of an order
    DevBuf planLen_, planPos_, planCnt_, planTmp_;  // set-up: plan counts, sweep positions, scan scratch
    int P_ = 0;                             // padded length of the level order
    int PS_ = 0;                            // padded length of the sweep order
    int merge_ = 2;                         // levels per super-level of the sweeps
    int cus_ = 0;                           // compute units
    int gs_ = 0;                            // grid of the streaming kernels
    int gb_ = 0;                            // largest backward-sweep grid (rho partial slots)
    int gF64_ = 0, gB64_ = 0, gF32_ = 0, gB32_ = 0;  // persistent grids
    int gateDist_ = 2;                      // sweep gates: groups back (forward) / ahead
    int pollCap_ = 1;
    int gateCap_ = 8;
    long long* trace_ = nullptr;            // set only while a timeline is recorded
    int K_ = 0;                             // batched form: systems in the stack (0: single system)
    DevBuf chunkSeg_, segChunk_, bst_, bpart_, bcnt_;  // batched form: layout, state, partials, counter
    int blocksPerCu_ = 0;                   // batched form: streaming grid (0: gs_)
    int lpr_ = 8;                           // batched form: SpMV lanes per row
    int mergeLevels_ = 0;                   // requested levels per sweep group (0: default)
    bool hostPlan_ = false;                 // ASTG_HOST_PLAN: symbolic plans on the host
    bool perm_ = false;                     // PCG vectors in sweep order (single-system solves)
    int spmvLpr_ = 4;                       // ASTG_SPMV_LPR: lanes per row of the single-system SpMV
    int spmvRows_ = 2;                      // ASTG_SPMV_ROWS: rows in flight per lane group (1, 2 or 4; 4 lanes per row)
    bool rowOrderForced_ = false;           // setRowOrder (batched form)
    DevBuf pos_;                            // row -> sweep position
    DevBuf prs_, pcol_, pmap_;              // CSR in sweep order: rowstart, columns as positions, source entry
    DevBuf valP_, invdP_, bP_, xP_;         // values, pivots, rhs and solution in sweep order
    DevBuf permLen_;                        // set-up: row lengths in sweep order
    bool batchGraph_ = true;                // batched form: iterations replayed from hipGraphs
    bool coop_ = false;                     // sweeps through the cooperative launch API (ASTG_COOP=1)
    bool spinFailed_ = false;               // a sweep hit its spin limit since the last solve started
    bool forceSpinError_ = false;           // test hook (ASTG_TEST_SPIN_ERROR_ONCE): one forced spin error
    bool f32Valid_ = false;                 // FP32 copies match the current factor (convert32 since factor)
};

}  // namespace astg
