// Frames on disk in the reference's binary formats (src/viewer/grid_dump.cpp):
// the layout (dumpOctreeLayout3dBinary / readOctreeLayout3dBinary, 280-377),
// the tilt state (dumpOctreeOmniTiltE3dBinary, 448-490), omni scalar grids
// (379-446) and the omni face-centred velocity (492-614). Host byte order,
// REAL = double, no header, as the reference writes them.
#pragma once

#include <string>
#include <vector>

#include "octree.h"

namespace astg {

// Per-level field arrays in node order (leaves, then ghosts).
struct OmniScalarFrame {
    std::vector<std::vector<double>> oct, tilt;  // per level: nnode values each
};
struct OmniVelocityFrame {
    std::vector<std::vector<double>> mac;    // per level: 3 per node (x, y, z interleaved)
    std::vector<std::vector<double>> tilt8;  // per level: 8 per node
};

bool writeLayoutFile(const std::string& path, const Octree& o, std::string& err);
// Reads a layout file and checks that it describes the octree's layers: the
// same nodes with the same indices, coordinates and flags, the same leaf and
// ghost orders and the same half-tilt list.
bool checkLayoutFile(const std::string& path, const Octree& o, std::string& err);

bool writeTiltFile(const std::string& path, const Octree& o, std::string& err);
// Replaces the tilt state (TiltENode: e, largest, closed, fixed, T-joint,
// closed by T-joint, half, half direction) of every node; the numbering's own
// open / fixed bits are kept, as the reference's restart keeps tilt_layout.
bool readTiltFile(const std::string& path, Octree& o, std::string& err);

bool writeScalarFile(const std::string& path, const OmniScalarFrame& f, std::string& err);
bool readScalarFile(const std::string& path, const Octree& o, OmniScalarFrame& f, std::string& err);

bool writeVelocityFile(const std::string& path, const OmniVelocityFrame& f, std::string& err);
bool readVelocityFile(const std::string& path, const Octree& o, OmniVelocityFrame& f, std::string& err);

}  // namespace astg
