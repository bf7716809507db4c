// Block-colour ordering of the projection system, an added form for the
// mixed-precision solve (the FP64 form keeps the reference's numbering).
//
// The domain is cut into cubes of `blockCells` finest cells; cubes get one of
// 8 colours by the parity of their coordinates (no two cubes that share a
// face, edge or corner have the same colour), and the unknowns are ordered by
// colour, then cube, then their reference index. Every coupling of the
// octree system joins unknowns at most one cell of their level apart, so for
// cubes at least as wide as the coarsest cells that touch them, unknowns of
// two different cubes of one colour are not coupled: the lower-triangular
// sweeps of MIC(0) in this order have one chain of dependency levels per
// colour, each as deep as the deepest cube of that colour, instead of one
// chain across the domain. At 256^3 (4 levels, 3.5 M unknowns), MIC(0)-PCG
// to 1e-6 in FP64 (depth / iterations): reference numbering 262 / 267; cubes
// of 2^3 cells 193 / 306, 4^3 91 / 313, 8^3 67 / 353, 16^3 130 / 359.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "octree.h"
#include "pcg.h"

namespace astg {

struct BlockOrder {
    int blockCells = 0;
    int n = 0;
    std::vector<int> perm;            // new row -> row of the reference numbering
    std::vector<int> rowstart, col;   // pattern of P A P^T, each row sorted by column
    std::vector<int> src;             // entry of P A P^T -> entry of A (CSR order)
};

// Positions come from the numbering: an octagon row at its cell centre, a tilt
// or half-tilt row at its vertex. `rowstart`/`col` are A's CSR pattern (rows
// sorted by column). Throws std::runtime_error on inconsistent input.
void buildBlockOrder(const Octree& o, const Numbering& num, const std::vector<int>& rowstart,
                     const std::vector<int>& col, int blockCells, BlockOrder& out);

// The sort key of every row of the reference numbering: colour * cubes + cube
// of the row's position (the first node, in level and node order, that holds
// the row), 8 * cubes for a row without one. Host worker pool. buildBlockOrder
// sorts the rows by these keys, stably.
void blockKeys(const Octree& o, const Numbering& num, int blockCells, std::vector<int>& keys);

// buildBlockOrder on the device from the keys (uploaded here) and A's device
// pattern (reference numbering, rows sorted by column): the same perm,
// rowstart, col and src, in device buffers. Rows of more than 64 entries are
// an error. On `stream`; returns after it has finished. keysOnDevice: `keys`
// points to device memory (copied on the device) instead of host memory.
struct BlockOrderDevice {
    int n = 0;
    long long nnz = 0;
    DevBuf keys, sortedKeys, iota, perm, inv, rowstart, col, src, tmp, flag;
    int build(hipStream_t stream, int n, long long nnz, const int* hostKeys, const int* rs, const int* colA,
              std::string& err, bool keysOnDevice = false);
};

// Device side (reorder.hip), on `stream`:
//   valOut[e] = valIn[src[e]] for e < nnz (skipped when valIn is null),
//   bOut[i] = bIn[perm[i]] for i < n;
int gatherPermuted(hipStream_t stream, long long nnz, const int* src, const double* valIn, double* valOut, int n,
                   const int* perm, const double* bIn, double* bOut, std::string& err);
//   xOut[perm[i]] = xIn[i] for i < n.
int scatterPermuted(hipStream_t stream, int n, const int* perm, const double* xIn, double* xOut, std::string& err);

}  // namespace astg
