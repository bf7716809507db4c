// astg_cuda_solver.hpp - header-only C++ shim with the shape of the
// reference's CUDA solver seam, over the C ABI in astg.h.
//
// Replaces the `CudaSolver` member used under CUDA_SOLVER
// (include/octree_omnigrid_solver3.h:101-105, include/ast_solver.h:666-668)
// and its call
//     fixedA.construct_from_matrix(A);
//     cudaSolver.solve(totalSize, fixedA.rowstart, fixedA.colindex, fixedA.value, b, vRes);
//     err = cudaSolver.solver_error; iter = cudaSolver.solver_iters;
// (src/octree_omnigrid_solver3.cpp:664-669). A maintainer adds
//     using CudaSolver = astg::HipSolver;
// and links libastg.so; the calling code stays as it is.
//
// Conventions kept from the call site: no return value, x is fully
// overwritten, the residual and iteration count are read from public members.
// Invalid input and device errors throw std::runtime_error; a solve that does
// not reach the tolerance leaves the last iterate in x (the reference logs the
// error and iteration count and carries on).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "astg.h"

namespace astg {

class HipSolver {
public:
    double solver_error = 0.0;  // final residual, ||r||_inf by default (absolute)
    int solver_iters = 0;       // PCG iterations of the last solve
    astg_pcg_params params;     // tol 1e-5, 3000 iterations, tau ASTG_TAU_AUTO (0.97 in FP64), ratio 0.25, FP64
    astg_pcg_result last{};     // full result of the last solve

    explicit HipSolver(int device = 0) {
        astg_pcg_params_default(&params);
        check(astg_create(&ctx_, device, 0u), "astg_create");
    }
    ~HipSolver() {
        if (ctx_) astg_destroy(ctx_);
    }
    HipSolver(const HipSolver&) = delete;
    HipSolver& operator=(const HipSolver&) = delete;

    void solve(int n, const std::vector<int>& rowstart, const std::vector<int>& colindex,
               const std::vector<double>& value, const std::vector<double>& b, std::vector<double>& x) {
        if (n < 0 || rowstart.size() != static_cast<size_t>(n) + 1 || b.size() < static_cast<size_t>(n))
            throw std::runtime_error("HipSolver::solve: inconsistent sizes");
        const size_t nnz = static_cast<size_t>(rowstart[n]);
        if (colindex.size() < nnz || value.size() < nnz)
            throw std::runtime_error("HipSolver::solve: CSR arrays shorter than rowstart[n]");
        x.resize(n);
        const int rc = astg_pcg_solve_csr(ctx_, n, rowstart.data(), colindex.data(), value.data(), b.data(),
                                          x.data(), &params, &last);
        if (rc != ASTG_OK && rc != ASTG_ERR_NOT_CONVERGED) check(rc, "astg_pcg_solve_csr");
        solver_error = last.residual;
        solver_iters = last.iterations;
    }

private:
    void check(int rc, const char* what) {
        if (rc == ASTG_OK) return;
        std::string msg = std::string(what) + " failed (" + std::to_string(rc) + ")";
        if (ctx_) msg += ": " + std::string(astg_last_error(ctx_));
        throw std::runtime_error(msg);
    }
    astg_ctx* ctx_ = nullptr;
};

}  // namespace astg
