// Block-colour ordering of the projection system (host part).
#include "reorder.h"

#include <algorithm>
#include <stdexcept>

#include "parallel.h"

namespace astg {

void buildBlockOrder(const Octree& o, const Numbering& num, const std::vector<int>& rowstart,
                     const std::vector<int>& col, int blockCells, BlockOrder& out) {
    const int n = num.n;
    if ((int)rowstart.size() != n + 1) throw std::runtime_error("block order: pattern does not match the numbering");
    if (blockCells < 1) throw std::runtime_error("block order: cube size must be positive");
    // doubled finest-grid coordinates of every row (cell centres are odd)
    std::vector<int> px(n, -1), py(n, -1), pz(n, -1);
    for (int l = 0; l < o.L; ++l) {
        const Level& Lv = o.lv[l];
        const int s = 2 << l;
        for (int i = 0; i < Lv.nnode; ++i) {
            const int a = num.oct[l][i], t = num.tilt[l][i];
            if (a >= 0 && px[a] < 0) {
                px[a] = Lv.x[i] * s + (s >> 1);
                py[a] = Lv.y[i] * s + (s >> 1);
                pz[a] = Lv.z[i] * s + (s >> 1);
            }
            if (t >= 0 && px[t] < 0) {
                px[t] = Lv.x[i] * s;
                py[t] = Lv.y[i] * s;
                pz[t] = Lv.z[i] * s;
            }
        }
    }
    // cube and colour of every row; a row without a position (none in a
    // consistent numbering) goes last in reference order
    const int w = 2 * blockCells;
    const int nb = (o.R + blockCells) / blockCells + 1;  // cubes per axis, the boundary vertices included
    const long long cubes = (long long)nb * nb * nb;
    std::vector<long long> key(n);
    for (int i = 0; i < n; ++i) {
        if (px[i] < 0) {
            key[i] = 8 * cubes;
            continue;
        }
        const int bx = px[i] / w, by = py[i] / w, bz = pz[i] / w;
        const int colour = (bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2);
        key[i] = (long long)colour * cubes + ((long long)bz * nb + by) * nb + bx;
    }
    // stable counting sort by key: the reference index breaks ties
    out.perm.resize(n);
    {
        std::vector<int> idx(n);
        for (int i = 0; i < n; ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key[a] < key[b]; });
        out.perm = std::move(idx);
    }
    std::vector<int> inv(n);
    for (int i = 0; i < n; ++i) inv[out.perm[i]] = i;
    // P A P^T: row i is A's row perm[i], columns renamed and sorted
    out.rowstart.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        const int r = out.perm[i];
        out.rowstart[i + 1] = out.rowstart[i] + (rowstart[r + 1] - rowstart[r]);
    }
    const size_t nnz = (size_t)out.rowstart[n];
    if (nnz != col.size()) throw std::runtime_error("block order: pattern size mismatch");
    out.col.resize(nnz);
    out.src.resize(nnz);
    parallelFor(n, [&](long lo, long hi) {
        std::pair<int, int> e[64];
        for (int i = lo; i < hi; ++i) {
            const int r = out.perm[i];
            const int len = rowstart[r + 1] - rowstart[r];
            if (len > 64) throw std::runtime_error("block order: row wider than 64 entries");
            for (int k = 0; k < len; ++k) e[k] = {inv[col[rowstart[r] + k]], rowstart[r] + k};
            std::sort(e, e + len);
            for (int k = 0; k < len; ++k) {
                out.col[out.rowstart[i] + k] = e[k].first;
                out.src[out.rowstart[i] + k] = e[k].second;
            }
        }
    });
    out.blockCells = blockCells;
    out.n = n;
}

}  // namespace astg
