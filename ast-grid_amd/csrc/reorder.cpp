// Block-colour ordering of the projection system (host part).
#include "reorder.h"

#include <algorithm>
#include <climits>
#include <stdexcept>

#include "parallel.h"

namespace astg {

void blockKeys(const Octree& o, const Numbering& num, int blockCells, std::vector<int>& keys) {
    const int n = num.n;
    if (blockCells < 1) throw std::runtime_error("block order: cube size must be positive");
    // first node holding each row: 2 * (global node index) + 0 for an
    // octagon row, + 1 for a tilt row (a row is of one kind only)
    std::vector<int> off(o.L + 1, 0);
    for (int l = 0; l < o.L; ++l) off[l + 1] = off[l] + o.lv[l].nnode;
    if ((long long)off[o.L] * 2 + 1 >= INT_MAX) throw std::runtime_error("block order: too many nodes");
    std::vector<int> first(n, INT_MAX);
    auto amin = [](int* p, int v) {
        int cur = __atomic_load_n(p, __ATOMIC_RELAXED);
        while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
    };
    for (int l = 0; l < o.L; ++l)
        parallelFor(o.lv[l].nnode, [&](long i0, long i1) {
            for (long i = i0; i < i1; ++i) {
                const int g = off[l] + (int)i, a = num.oct[l][i], t = num.tilt[l][i];
                if (a >= 0) amin(&first[a], 2 * g);
                if (t >= 0) amin(&first[t], 2 * g + 1);
            }
        });
    const int w = 2 * blockCells;
    const int nb = (o.R + blockCells) / blockCells + 1;
    const long long cubes = (long long)nb * nb * nb;
    if (8 * cubes + 1 >= INT_MAX) throw std::runtime_error("block order: too many cubes");
    keys.resize(n);
    parallelFor(n, [&](long r0, long r1) {
        for (long r = r0; r < r1; ++r) {
            const int e = first[r];
            if (e == INT_MAX) {
                keys[r] = (int)(8 * cubes);
                continue;
            }
            const int g = e >> 1;
            const int l = (int)(std::upper_bound(off.begin(), off.end(), g) - off.begin()) - 1;
            const int i = g - off[l];
            const Level& Lv = o.lv[l];
            const int s = 2 << l, h = (e & 1) ? 0 : (s >> 1);
            const int px = Lv.x[i] * s + h, py = Lv.y[i] * s + h, pz = Lv.z[i] * s + h;
            const int bx = px / w, by = py / w, bz = pz / w;
            const int colour = (bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2);
            keys[r] = (int)(colour * cubes + ((long long)bz * nb + by) * nb + bx);
        }
    });
}

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
