// Level merging for the sync-free triangular sweeps (host-side planning).
//
// A forward-sweep row i computes y_i = d_i (r_i - sum_k F_ik y_k) and has to
// wait for every y_k it reads. Grouping the dependency levels in runs of q
// ("super-levels") and substituting, in row i's formula, each dependency j of
// i's own super-level by j's formula (y_j = d_j (r_j - sum_m F_jm y_m)) gives
//     y_i = d_i (r_i - sum_j R_ij r_j - sum_m W_im y_m)
// where every remaining y_m lies in an earlier super-level, so the sweep makes
// depth/q dependent steps instead of depth. Keys are substituted from the
// highest level down, so each has received all of its contributions when it is
// substituted; the identity holds after every single substitution, which lets
// a row stop early. The backward sweep z = L^-T y is the mirror image: upper
// dependencies, y in place of r, lowest level first.
//
// The plan is symbolic and depends on the pattern only: per row it fixes the
// substituted keys (r-terms, in substitution order), the remaining keys
// (y-terms, ascending row index -- the reference's accumulation order when
// nothing is substituted) and a slot program that the device replays on every
// new factor to obtain the coefficients R and W.
#pragma once

#include <cstdint>
#include <vector>

namespace astg {

struct SweepStencil {
    std::vector<int> start;    // per sweep position (P + 1): first entry of the row
    std::vector<int> nr;       // per sweep position: number of r-terms (the first entries)
    std::vector<int> idx;      // per entry. r-term: row of the key (forward) or its sweep
                               // position (backward); y-term: sweep position of the key
    std::vector<int> ustart;   // per sweep position (P + 1): first slot target in upd
    std::vector<uint8_t> upd;  // slot targets: the row's own dependencies in pattern order,
                               // then for each r-term the dependencies of its key
    long long rTerms = 0, yTerms = 0, kept = 0;
};

// Most dependencies of a row on one side of the diagonal, and so most y-terms
// of a merged row: the sweep kernels hold a row's dependencies in registers.
constexpr int kMaxDeps = 16;
// Most r-terms per row; with at most maxY y-terms, a row's slots fit in a byte.
constexpr int kMaxRTerms = 64;

// dir 0: forward (lower dependencies [rs, lend)); dir 1: backward (upper
// dependencies [ubeg, rs+1)). lev: dependency level per row; order: rows in
// sweep order (P entries, -1 = padding); pos: row -> sweep position.
// A substitution that would leave more than maxY y-terms is not made (the key
// stays a y-term of the same super-level, at an earlier sweep position).
void planStencil(int dir, int n, const int* rs, const int* lend, const int* ubeg, const int* col, const int* lev,
                 int q, int maxY, const int* order, int P, const int* pos, SweepStencil& st);

}  // namespace astg
