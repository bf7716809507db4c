// Split-map generation for seeded sphere scenes.
//
// A split map holds, for every finest cell of an R^3 grid, the level of the
// octree leaf covering it (the reference fills the same grid through its
// splitMapSetup callback and refineSplitMapByObstacleSdf, octree_util.h:17-47).
#pragma once

#include <cstdint>
#include <vector>

namespace astg {

// Refines top-down: a level-l leaf whose centre lies within band[l-1] (absolute
// distance, negative inside) of the sphere union becomes 2x2x2 level-(l-1)
// leaves. spheres: n x (cx, cy, cz, r) in the unit cube.
void splitSpheres(int R, int levels, const double* spheres, int n, const double* band, std::vector<int>& split);

// Refines top-down: a level-l leaf becomes 2x2x2 level-(l-1) leaves when the
// smallest of dist over its cells is within band[l-1]. dist: (R >> shift)^3
// values at the centres of cells 2^shift finest cells wide (x fastest; shift
// below levels), e.g. the distance to a region that must be finest (<= 0
// inside it); a block finer than those cells takes its cell's value.
void splitDistance(int R, int levels, int shift, const float* dist, const double* band, std::vector<int>& split);

// Enforces the neighbourhood condition the reference's layer construction
// relies on: for every leaf of level l, each of its 26 level-l neighbour blocks
// is covered only by leaves of level l-1, l or l+1 (so that every such block
// exists as a leaf or ghost node). Refines until the condition holds.
// Returns the number of refinement passes.
int balanceSplit(int R, int levels, std::vector<int>& split);

// Splitmix64-based counter RNG: identical streams on host and device.
uint64_t splitmix64(uint64_t x);
double u01(uint64_t seed, uint64_t id);            // uniform [0,1)
double normal01(uint64_t seed, uint64_t id);       // standard normal (Box-Muller)

}  // namespace astg
