// Checker-side harness around the reference's uniform-layout AST solver
// (TEST INFRASTRUCTURE ONLY).
//
// Compiled by oracle/ref.mk, together with the reference's own sources, into
// oracle/_ref/libastgref.so. Nothing in the product links or loads it:
// tests/golden/make_golden_ast.py writes fixtures with it, the CPU tests pin
// the oracle's restatement against it when it is present, and bench.py may time
// its project() as the CPU baseline of the uniform-layout projection.
//
// It instantiates ast::ASTSolver<d> (include/ast_solver.h:20-673) as the
// reference's pybind module does (python_bindings/ast_solver.cc:43-54), sets
// the tilt layout and the velocity field from caller arrays and runs the
// reference's own project(dt) (ast_solver.h:322-498: updateIndex, assembly,
// MIC(0)-PCG, correctVelocityByPressure, constrainVelocity). project() and the
// solver's fields are private members of ASTSolver, so that one header is
// read with `private` spelled `public`; every header it includes is included
// first, unchanged.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include <spdlog/spdlog.h>

#include "solver.h"
#include "ast_grid.h"
#include "ast_face_grid.h"
#include "ast_iterator.h"
#include "ast_advection.h"
#include "ast_util.h"
#include "ast_grid_pool.h"
#include "viewer/image_dump.h"
#include "pcgsolver/pcg_solver.h"

#define private public
#include "ast_solver.h"
#undef private

namespace {

std::string g_err;

struct AstBase {
    virtual ~AstBase() {}
    virtual void setLayout(const uint8_t* open) = 0;
    virtual void setVelocity(const double* const* mac, const double* tilt) = 0;
    virtual void getVelocity(double* const* mac, double* tilt) = 0;
    virtual void project(double dt, double rho, int* n, long long* nnz, bool report) = 0;
    virtual void system(int* rs, int* col, double* val, double* b, double* x) = 0;
    virtual void numbering(int* cell, int* tilt) = 0;
    virtual void solveInfo(int* iters, double* err) = 0;
};

template <int d>
struct AstImpl : AstBase {
    using IndexT = Eigen::Matrix<int, d, 1>;
    static constexpr int kT = d == 2 ? 4 : 8;
    ast::ASTSolver<d> s;
    IndexT res;
    int iters = 0;
    double err = 0;
    bool same = false;

    static IndexT toIndex(const int* r) {
        IndexT v;
        for (int a = 0; a < d; ++a) v[a] = r[a];
        return v;
    }
    AstImpl(const int* r, double dx) : s(toIndex(r), dx, true, "astg_ast_harness"), res(toIndex(r)) {
        s.initialize();
    }
    // box iteration, x fastest (box_iterator.h), over [0, hi]
    template <class F>
    static void box(const IndexT& hi, F f) {
        for (auto it = ast::BoxIterator<d>(IndexT::Zero(), hi); it.valid(); it.next()) f(it.index());
    }
    void setLayout(const uint8_t* open) override {
        long i = 0;
        box(res, [&](const IndexT& c) { s.layout->getTiltNode(c).open = open[i++] != 0; });
        s.layout->constrainFixedTiltNode();
        s.layout->propogateNodeInfo();
    }
    template <int axis>
    void macCopy(const double* src, double* dst) {
        IndexT hi = res - IndexT::Ones();
        hi[axis] += 1;
        long i = 0;
        box(hi, [&](const IndexT& c) {
            auto& v = s.velocity->template get<ast::node_type::cell_node, axis>(c);
            if (src) v = src[i]; else dst[i] = v;
            ++i;
        });
    }
    void tiltCopy(const double* src, double* dst) {
        long i = 0;
        box(res, [&](const IndexT& c) {
            auto& u = s.velocity->template get<ast::node_type::vertex_node>(c);
            for (int k = 0; k < kT; ++k) {
                if (src) u[k] = src[i * kT + k]; else dst[i * kT + k] = u[k];
            }
            ++i;
        });
    }
    void setVelocity(const double* const* mac, const double* tilt) override {
        macCopy<0>(mac[0], nullptr);
        macCopy<1>(mac[1], nullptr);
        if constexpr (d == 3) macCopy<2>(mac[2], nullptr);
        tiltCopy(tilt, nullptr);
    }
    void getVelocity(double* const* mac, double* tilt) override {
        macCopy<0>(nullptr, mac[0]);
        macCopy<1>(nullptr, mac[1]);
        if constexpr (d == 3) macCopy<2>(nullptr, mac[2]);
        tiltCopy(nullptr, tilt);
    }
    void project(double dt, double rho, int* n, long long* nnz, bool report) override {
        s.rho = rho;
        s.project(dt);
        *n = s.total_index_size;
        long long z = 0;
        for (int i = 0; i < (int)s.A.n; ++i) z += (long long)s.A.index[i].size();
        *nnz = z;
        same = false;
        if (!report) return;
        // project() keeps no record of its solve: the same solver settings
        // (ASTSolver::initialize, ast_solver.h:40) on the same A and b report it
        PCGSolver<double> again;
        again.set_solver_parameters(SOLVER_ACCURACY, 5000, 0.97, 0.95);
        std::vector<double> x(s.vRes.size(), 0.0);  // solve() does not size its result
        again.solve(s.A, s.b, x, err, iters);
        same = x == s.vRes;
    }
    void system(int* rs, int* col, double* val, double* b, double* x) override {
        long long p = 0;
        for (int i = 0; i < (int)s.A.n; ++i) {
            rs[i] = (int)p;
            for (size_t j = 0; j < s.A.index[i].size(); ++j, ++p) {
                col[p] = (int)s.A.index[i][j];
                val[p] = s.A.value[i][j];
            }
        }
        rs[s.A.n] = (int)p;
        for (size_t i = 0; i < s.b.size(); ++i) {
            b[i] = s.b[i];
            x[i] = s.vRes[i];
        }
    }
    void numbering(int* cell, int* tilt) override {
        long i = 0;
        box(res - IndexT::Ones(), [&](const IndexT& c) { cell[i++] = s.pIndex->template get<ast::node_type::cell_node>(c); });
        i = 0;
        box(res, [&](const IndexT& c) { tilt[i++] = s.pIndex->template get<ast::node_type::vertex_node>(c); });
    }
    void solveInfo(int* it, double* e) override {
        *it = same ? iters : -1;
        *e = err;
    }
};

template <class F>
int guard(F f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
    } catch (...) {
        g_err = "unknown exception";
    }
    return 1;
}

}  // namespace

extern "C" {

const char* ast_ref_last_error() { return g_err.c_str(); }

void* ast_ref_create(int d, const int* res, double dx) {
    AstBase* h = nullptr;
    if (guard([&] {
            spdlog::set_level(spdlog::level::off);
            if (d == 2) h = new AstImpl<2>(res, dx);
            else if (d == 3) h = new AstImpl<3>(res, dx);
            else throw std::runtime_error("dimension must be 2 or 3");
        }))
        return nullptr;
    return h;
}

void ast_ref_destroy(void* h) { delete static_cast<AstBase*>(h); }

int ast_ref_set_layout(void* h, const uint8_t* open) {
    return guard([&] { static_cast<AstBase*>(h)->setLayout(open); });
}

int ast_ref_set_velocity(void* h, const double* mac0, const double* mac1, const double* mac2, const double* tilt) {
    const double* m[3] = {mac0, mac1, mac2};
    return guard([&] { static_cast<AstBase*>(h)->setVelocity(m, tilt); });
}

int ast_ref_velocity(void* h, double* mac0, double* mac1, double* mac2, double* tilt) {
    double* m[3] = {mac0, mac1, mac2};
    return guard([&] { static_cast<AstBase*>(h)->getVelocity(m, tilt); });
}

// report != 0 also re-solves the system to report the solve's iterations
// (ast_ref_solve_info); leave it 0 to time project() alone.
int ast_ref_project(void* h, double dt, double rho, int* n, long long* nnz, int report) {
    return guard([&] { static_cast<AstBase*>(h)->project(dt, rho, n, nnz, report != 0); });
}

int ast_ref_system(void* h, int* rs, int* col, double* val, double* b, double* x) {
    return guard([&] { static_cast<AstBase*>(h)->system(rs, col, val, b, x); });
}

int ast_ref_numbering(void* h, int* cell, int* tilt) {
    return guard([&] { static_cast<AstBase*>(h)->numbering(cell, tilt); });
}

int ast_ref_solve_info(void* h, int* iters, double* err) {
    return guard([&] { static_cast<AstBase*>(h)->solveInfo(iters, err); });
}

}  // extern "C"
