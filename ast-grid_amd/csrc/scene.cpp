#include "scene.h"

#include <algorithm>
#include <cmath>

namespace astg {

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

double u01(uint64_t seed, uint64_t id) {
    uint64_t v = splitmix64(splitmix64(seed) ^ (id * 0xD1B54A32D192ED03ull));
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}

double normal01(uint64_t seed, uint64_t id) {
    double a = u01(seed, 2 * id), b = u01(seed, 2 * id + 1);
    if (a < 1e-300) a = 1e-300;
    return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
}

void splitSpheres(int R, int levels, const double* sph, int n, const double* band, std::vector<int>& split) {
    split.assign((size_t)R * R * R, levels - 1);
    auto dist = [&](double px, double py, double pz) {
        double d = 1e30;
        for (int s = 0; s < n; ++s) {
            double dx = px - sph[4 * s], dy = py - sph[4 * s + 1], dz = pz - sph[4 * s + 2];
            d = std::min(d, std::sqrt(dx * dx + dy * dy + dz * dz) - sph[4 * s + 3]);
        }
        return d;
    };
    const double h0 = 1.0 / R;
    for (int l = levels - 1; l >= 1; --l) {
        const int b = 1 << l, res = R >> l;
        for (int z = 0; z < res; ++z)
            for (int y = 0; y < res; ++y)
                for (int x = 0; x < res; ++x) {
                    size_t c0 = ((size_t)(z * b) * R + y * b) * R + x * b;
                    if (split[c0] != l) continue;
                    double d = dist((x + 0.5) * b * h0, (y + 0.5) * b * h0, (z + 0.5) * b * h0);
                    if (d > band[l - 1]) continue;
                    for (int k = 0; k < b; ++k)
                        for (int j = 0; j < b; ++j)
                            for (int i = 0; i < b; ++i)
                                split[((size_t)(z * b + k) * R + y * b + j) * R + x * b + i] = l - 1;
                }
    }
}

void splitDistance(int R, int levels, int shift, const float* dist, const double* band, std::vector<int>& split) {
    split.assign((size_t)R * R * R, levels - 1);
    // smallest distance over every aligned level-l block from the lattice's
    // level up; a block finer than the lattice takes its lattice cell's value
    std::vector<std::vector<float>> mn(levels);
    const int dr = R >> shift;
    mn[shift].assign(dist, dist + (size_t)dr * dr * dr);
    for (int l = shift + 1; l < levels; ++l) {
        const int res = R >> l, pr = R >> (l - 1);
        mn[l].assign((size_t)res * res * res, 0.f);
        for (int z = 0; z < res; ++z)
            for (int y = 0; y < res; ++y)
                for (int x = 0; x < res; ++x) {
                    float m = 1e30f;
                    for (int k = 0; k < 8; ++k) {
                        const int cx = 2 * x + (k & 1), cy = 2 * y + ((k >> 1) & 1), cz = 2 * z + (k >> 2);
                        m = std::min(m, mn[l - 1][((size_t)cz * pr + cy) * pr + cx]);
                    }
                    mn[l][((size_t)z * res + y) * res + x] = m;
                }
    }
    for (int l = levels - 1; l >= 1; --l) {
        const int b = 1 << l, res = R >> l;
        for (int z = 0; z < res; ++z)
            for (int y = 0; y < res; ++y)
                for (int x = 0; x < res; ++x) {
                    size_t c0 = ((size_t)(z * b) * R + y * b) * R + x * b;
                    if (split[c0] != l) continue;
                    const float m = l >= shift ? mn[l][((size_t)z * res + y) * res + x]
                                               : mn[shift][((size_t)(z >> (shift - l)) * dr + (y >> (shift - l))) * dr +
                                                           (x >> (shift - l))];
                    if (m > band[l - 1]) continue;
                    for (int k = 0; k < b; ++k)
                        for (int j = 0; j < b; ++j)
                            for (int i = 0; i < b; ++i)
                                split[((size_t)(z * b + k) * R + y * b + j) * R + x * b + i] = l - 1;
                }
    }
}

int balanceSplit(int R, int levels, std::vector<int>& split) {
    int passes = 0;
    for (;;) {
        ++passes;
        // min / max level inside every aligned level-l block
        std::vector<std::vector<int8_t>> mn(levels), mx(levels);
        for (int l = 0; l < levels; ++l) {
            const int res = R >> l;
            mn[l].assign((size_t)res * res * res, 127);
            mx[l].assign((size_t)res * res * res, -1);
        }
        for (int z = 0; z < R; ++z)
            for (int y = 0; y < R; ++y)
                for (int x = 0; x < R; ++x) {
                    int v = split[((size_t)z * R + y) * R + x];
                    for (int l = 0; l < levels; ++l) {
                        const int res = R >> l;
                        size_t c = ((size_t)(z >> l) * res + (y >> l)) * res + (x >> l);
                        if (v < mn[l][c]) mn[l][c] = (int8_t)v;
                        if (v > mx[l][c]) mx[l][c] = (int8_t)v;
                    }
                }
        bool changed = false;
        auto refineBlock = [&](int l, int bx, int by, int bz, int to) {
            const int b = 1 << l;
            for (int k = 0; k < b; ++k)
                for (int j = 0; j < b; ++j)
                    for (int i = 0; i < b; ++i) {
                        int& v = split[((size_t)(bz * b + k) * R + by * b + j) * R + bx * b + i];
                        if (v > to) v = to;
                    }
            changed = true;
        };
        for (int l = levels - 1; l >= 0; --l) {
            const int res = R >> l, b = 1 << l;
            for (int z = 0; z < res; ++z)
                for (int y = 0; y < res; ++y)
                    for (int x = 0; x < res; ++x) {
                        size_t c = ((size_t)z * res + y) * res + x;
                        if (mn[l][c] != l || mx[l][c] != l) continue;  // not a level-l leaf
                        for (int dz = -1; dz <= 1; ++dz)
                            for (int dy = -1; dy <= 1; ++dy)
                                for (int dx = -1; dx <= 1; ++dx) {
                                    int nx = x + dx, ny = y + dy, nz = z + dz;
                                    if (nx < 0 || ny < 0 || nz < 0 || nx >= res || ny >= res || nz >= res) continue;
                                    size_t q = ((size_t)nz * res + ny) * res + nx;
                                    if (mn[l][q] < l - 1) {
                                        // a much finer leaf sits in the neighbour block: refine this leaf
                                        refineBlock(l, x, y, z, l - 1);
                                        mn[l][c] = mx[l][c] = (int8_t)(l - 1);
                                        goto next_leaf;
                                    }
                                    if (mx[l][q] > l + 1) {
                                        // the neighbour block lies inside a much coarser leaf: split that leaf once
                                        int L2 = mx[l][q];
                                        int s = L2 - l;
                                        refineBlock(L2, nx >> s, ny >> s, nz >> s, L2 - 1);
                                    }
                                }
                    next_leaf:;
                        (void)b;
                    }
        }
        if (!changed) break;
    }
    return passes;
}

}  // namespace astg
