// Symbolic plan of the MIC(0) factorisation (host side, pattern only).
//
// Row k of the factorisation (pcg_solver.h:129-189, restated left-looking per
// row) needs, for every lower entry (k,m), the entries of column m below the
// diagonal that are missing from row k's pattern (the modified-Cholesky
// compensation), and for every upper entry (k,i), the columns m common to
// lower(k) and lower(i). Both are pattern-only, so they are listed once per
// pattern, one segment per matrix entry; on the device every entry of a row
// is handled by its own lane and gathers its operands independently.
//
// Segment of entry position a (CSR order, [start[a], start[a+1]) in prog):
//   lower entry (k,m):  positions p of column m below the diagonal (upper
//                       entries of row m, ascending) that are not in row k's
//                       pattern;
//   upper entry (k,i):  pairs (j, c), ascending m: lower entry rs[k] + j of
//                       row k and lower entry c of row i share the column m;
//   diagonal:           empty.
#pragma once

#include <vector>

namespace astg {

void planMic(int n, const int* rs, const int* lend, const int* ubeg, const int* col, std::vector<long long>& start,
             std::vector<int>& prog);

}  // namespace astg
