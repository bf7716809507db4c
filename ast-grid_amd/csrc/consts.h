// Ownership of the per-topology kernel constants.
//
// The octree kernels read level resolutions, spacings and offsets from
// __constant__ memory, which exists once per GPU in the process, not once per
// context. Every context that launches those kernels first takes ownership
// of its table on the current GPU; when ownership changes hands the GPU is
// synchronised (kernels of the previous owner may still be reading the old
// values) and the new owner uploads its values. Calls that launch such
// kernels hold the process-wide lock for their duration, so contexts used
// from several threads take turns instead of racing on the tables.
#pragma once

#include <cstdint>
#include <mutex>

namespace astg {

enum ConstTable { CT_PROJECT = 0, CT_ADVECT = 1, CT_COUNT };

std::recursive_mutex& constMutex();
// A process-unique owner id (never reused, unlike addresses).
uint64_t newConstOwner();
// True when `owner` must upload its values of `table` on the current device:
// ownership changed (the device has been synchronised) or `force` is set.
bool takeConstants(int table, uint64_t owner, bool force);

}  // namespace astg
