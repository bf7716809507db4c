// Drop-in use of the CUDA-solver seam from C++: reads a CSR system, solves it
// through astg::HipSolver exactly as the reference's project() calls its
// CudaSolver (src/octree_omnigrid_solver3.cpp:664-669), writes the solution.
//
//   cuda_solver_shim <in.bin> <out.bin> [device]
//
// in.bin : int32 n, int32 nnz, int32 rowstart[n+1], int32 col[nnz],
//          float64 val[nnz], float64 b[n]
// out.bin: float64 x[n], float64 solver_error, int32 solver_iters
//
// Persistent sweeps use cooperative launches, which hold the device's
// global-wave-sync resource for the process: run one solver process per GPU.
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <vector>

#include "astg_cuda_solver.hpp"

// The same call sequence for in-process use (ctypes): returns 0 on success.
extern "C" int astg_example_shim_solve(int n, const int* rowstart, const int* col, const double* val,
                                       const double* b, double tol, double* x, double* error, int* iters) {
    try {
        const int nnz = rowstart[n];
        std::vector<int> rs(rowstart, rowstart + n + 1), ci(col, col + nnz);
        std::vector<double> v(val, val + nnz), rhs(b, b + n), sol;
        astg::HipSolver cudaSolver(0);
        cudaSolver.params.tol = tol;
        cudaSolver.solve(n, rs, ci, v, rhs, sol);
        for (int i = 0; i < n; ++i) x[i] = sol[i];
        *error = cudaSolver.solver_error;
        *iters = cudaSolver.solver_iters;
        return 0;
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}

#ifndef ASTG_SHIM_NO_MAIN
template <class T>
static bool readN(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return fread(v.data(), sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s in.bin out.bin [device]\n", argv[0]);
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) {
        perror(argv[1]);
        return 2;
    }
    int hdr[2];
    std::vector<int> rowstart, col;
    std::vector<double> val, b, x;
    bool ok = fread(hdr, sizeof(int), 2, f) == 2 && hdr[0] >= 0 && hdr[1] >= 0;
    ok = ok && readN(f, rowstart, (size_t)hdr[0] + 1) && readN(f, col, hdr[1]) && readN(f, val, hdr[1]) &&
         readN(f, b, hdr[0]);
    fclose(f);
    if (!ok) {
        fprintf(stderr, "%s: truncated input\n", argv[1]);
        return 2;
    }
    try {
        astg::HipSolver cudaSolver(argc > 3 ? atoi(argv[3]) : 0);
        cudaSolver.params.tol = 1e-6;
        cudaSolver.solve(hdr[0], rowstart, col, val, b, x);
        FILE* o = fopen(argv[2], "wb");
        if (!o) {
            perror(argv[2]);
            return 2;
        }
        fwrite(x.data(), sizeof(double), x.size(), o);
        fwrite(&cudaSolver.solver_error, sizeof(double), 1, o);
        fwrite(&cudaSolver.solver_iters, sizeof(int), 1, o);
        fclose(o);
        printf("n=%d iters=%d error=%.3e\n", hdr[0], cudaSolver.solver_iters, cudaSolver.solver_error);
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
#endif
