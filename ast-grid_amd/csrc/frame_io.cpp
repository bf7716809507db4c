// Reference binary frame formats (see frame_io.h).
//
// Records, as the reference writes them with sizeof of its structs (x86-64):
//   LayerNode3 (include/octree_layer3.h:11-20), 24 bytes:
//     size_t index | uint16 x, y, z | 2 pad | uint32 flag | 4 pad
//   TiltENode (include/tilt_e_node.h:15-23), 24 bytes:
//     double e | bool largest, closed, fixed, T_joint, closed_by_T_joint,
//     half_tilt | 2 pad | int half_direction | 4 pad
// Layout map keys and leaf / ghost lists are PACK_COOR3 (no level bits,
// macro.h:39); half-tilt entries are PACK_LCOOR3 (with level, macro.h:40).
// Padding bytes are written as zero. The reference's map may also hold
// default records (index 0, coordinates 0, flag 0) left by lookups of absent
// coordinates through operator[]; they are accepted and skipped on reading
// and never written.
#include "frame_io.h"

#include <cstdint>
#include <cstring>
#include <fstream>

namespace astg {

namespace {

struct LayerNodeRec {
    uint64_t index;
    uint16_t x, y, z, pad0;
    uint32_t flag, pad1;
};
static_assert(sizeof(LayerNodeRec) == 24, "LayerNode3 record is 24 bytes");

struct TiltRec {
    double e;
    uint8_t largest, closed, fixed, tjoint, closedByTJ, half, pad0[2];
    int32_t halfDir;
    uint32_t pad1;
};
static_assert(sizeof(TiltRec) == 24, "TiltENode record is 24 bytes");

inline int64_t coor3(int x, int y, int z) { return ((int64_t)x << 40) | ((int64_t)y << 20) | (int64_t)z; }

class Out {
public:
    explicit Out(const std::string& p) : f_(p, std::ios::binary) {}
    bool ok() const { return (bool)f_; }
    template <typename T>
    void put(const T& v) { f_.write(reinterpret_cast<const char*>(&v), sizeof(T)); }
    void putDoubles(const double* v, size_t n) { f_.write(reinterpret_cast<const char*>(v), n * 8); }
    bool close() {
        f_.close();
        return !f_.fail();
    }

private:
    std::ofstream f_;
};

class In {
public:
    explicit In(const std::string& p) : f_(p, std::ios::binary) {}
    bool ok() const { return (bool)f_; }
    template <typename T>
    bool get(T& v) { return (bool)f_.read(reinterpret_cast<char*>(&v), sizeof(T)); }
    bool getDoubles(double* v, size_t n) { return (bool)f_.read(reinterpret_cast<char*>(v), n * 8); }
    bool atEnd() { return f_.peek() == std::char_traits<char>::eof(); }

private:
    std::ifstream f_;
};

bool failWith(std::string& err, const std::string& path, const std::string& what) {
    err = path + ": " + what;
    return false;
}

// One per-level array block: int levels, then per level int size + values.
bool readLevels(In& in, const std::string& path, const Octree& o, int perNode, std::vector<std::vector<double>>& out,
                std::string& err, const char* what) {
    int L = 0;
    if (!in.get(L)) return failWith(err, path, std::string("truncated (") + what + ")");
    if (L != o.L) return failWith(err, path, std::string(what) + ": level count differs from the layout");
    out.assign(L, {});
    for (int l = 0; l < L; ++l) {
        int size = 0;
        if (!in.get(size)) return failWith(err, path, std::string("truncated (") + what + ")");
        if (size != o.lv[l].nnode) return failWith(err, path, std::string(what) + ": level size differs from the layout");
        out[l].resize((size_t)size * perNode);
        if (!in.getDoubles(out[l].data(), out[l].size()))
            return failWith(err, path, std::string("truncated (") + what + ")");
    }
    return true;
}

void writeLevels(Out& out, const std::vector<std::vector<double>>& v, int perNode) {
    out.put((int)v.size());
    for (const auto& a : v) {
        out.put((int)(a.size() / perNode));
        out.putDoubles(a.data(), a.size());
    }
}

}  // namespace

bool writeLayoutFile(const std::string& path, const Octree& o, std::string& err) {
    Out out(path);
    if (!out.ok()) return failWith(err, path, "cannot open for writing");
    out.put((int)o.L);
    for (int l = 0; l < o.L; ++l) {
        const Level& Lv = o.lv[l];
        out.put((int)Lv.nnode);  // the map table: every node, in node order
        for (int i = 0; i < Lv.nnode; ++i) {
            LayerNodeRec r;
            std::memset(&r, 0, sizeof r);
            r.index = (uint64_t)i;
            r.x = Lv.x[i];
            r.y = Lv.y[i];
            r.z = Lv.z[i];
            r.flag = Lv.flag[i];
            out.put(coor3(Lv.x[i], Lv.y[i], Lv.z[i]));
            out.put(r);
        }
        out.put((int)(Lv.nnode - Lv.nleaf));  // ghost list
        for (int i = Lv.nleaf; i < Lv.nnode; ++i) out.put(coor3(Lv.x[i], Lv.y[i], Lv.z[i]));
        out.put((int)Lv.nleaf);  // leaf list
        for (int i = 0; i < Lv.nleaf; ++i) out.put(coor3(Lv.x[i], Lv.y[i], Lv.z[i]));
    }
    out.put((int)o.half.size());
    for (lcoor_t h : o.half) out.put((int64_t)h);
    if (!out.close()) return failWith(err, path, "write failed");
    return true;
}

bool checkLayoutFile(const std::string& path, const Octree& o, std::string& err) {
    In in(path);
    if (!in.ok()) return failWith(err, path, "cannot open");
    int L = 0;
    if (!in.get(L)) return failWith(err, path, "truncated");
    if (L != o.L) return failWith(err, path, "level count differs from the context's layout");
    for (int l = 0; l < L; ++l) {
        const Level& Lv = o.lv[l];
        const std::string lev = "level " + std::to_string(l) + ": ";
        int size = 0;
        if (!in.get(size)) return failWith(err, path, "truncated");
        if (size < Lv.nnode) return failWith(err, path, lev + "node count differs");
        std::vector<uint8_t> seen((size_t)Lv.nnode, 0);
        int nodes = 0;
        for (int q = 0; q < size; ++q) {
            int64_t key = 0;
            LayerNodeRec r;
            if (!in.get(key) || !in.get(r)) return failWith(err, path, "truncated");
            // the reference's map also holds default records (flag 0) that its
            // operator[] lookups of absent coordinates inserted; they are no nodes
            if (r.flag == 0) {
                if (r.index != 0 || r.x != 0 || r.y != 0 || r.z != 0)
                    return failWith(err, path, lev + "record without flags is not a default record");
                continue;
            }
            ++nodes;
            if (r.index >= (uint64_t)Lv.nnode) return failWith(err, path, lev + "node index out of range");
            const int i = (int)r.index;
            if (seen[i]) return failWith(err, path, lev + "node index repeated");
            seen[i] = 1;
            if (key != coor3(r.x, r.y, r.z) || r.x != Lv.x[i] || r.y != Lv.y[i] || r.z != Lv.z[i] ||
                r.flag != Lv.flag[i])
                return failWith(err, path, lev + "node " + std::to_string(i) + " differs");
        }
        if (nodes != Lv.nnode) return failWith(err, path, lev + "node count differs");
        int ng = 0;
        if (!in.get(ng)) return failWith(err, path, "truncated");
        if (ng != Lv.nnode - Lv.nleaf) return failWith(err, path, lev + "ghost count differs");
        for (int q = 0; q < ng; ++q) {
            int64_t c = 0;
            const int i = Lv.nleaf + q;
            if (!in.get(c)) return failWith(err, path, "truncated");
            if (c != coor3(Lv.x[i], Lv.y[i], Lv.z[i])) return failWith(err, path, lev + "ghost order differs");
        }
        int nl = 0;
        if (!in.get(nl)) return failWith(err, path, "truncated");
        if (nl != Lv.nleaf) return failWith(err, path, lev + "leaf count differs");
        for (int i = 0; i < nl; ++i) {
            int64_t c = 0;
            if (!in.get(c)) return failWith(err, path, "truncated");
            if (c != coor3(Lv.x[i], Lv.y[i], Lv.z[i])) return failWith(err, path, lev + "leaf order differs");
        }
    }
    int nh = 0;
    if (!in.get(nh)) return failWith(err, path, "truncated");
    if (nh != (int)o.half.size()) return failWith(err, path, "half-tilt count differs");
    for (int q = 0; q < nh; ++q) {
        int64_t c = 0;
        if (!in.get(c)) return failWith(err, path, "truncated");
        if (c != (int64_t)o.half[q]) return failWith(err, path, "half-tilt order differs");
    }
    if (!in.atEnd()) return failWith(err, path, "trailing bytes");
    return true;
}

bool writeTiltFile(const std::string& path, const Octree& o, std::string& err) {
    Out out(path);
    if (!out.ok()) return failWith(err, path, "cannot open for writing");
    out.put((int)o.L);
    for (int l = 0; l < o.L; ++l) {
        out.put((int)o.lv[l].nnode);
        for (int i = 0; i < o.lv[l].nnode; ++i) {
            const TiltNode& t = o.tilt[l][i];
            TiltRec r;
            std::memset(&r, 0, sizeof r);
            r.e = t.e;
            r.largest = (t.bits & TB_LARGEST) ? 1 : 0;
            r.closed = (t.bits & TB_CLOSED) ? 1 : 0;
            r.fixed = (t.bits & TB_FIXED) ? 1 : 0;
            r.tjoint = (t.bits & TB_TJOINT) ? 1 : 0;
            r.closedByTJ = (t.bits & TB_CLOSED_BY_TJ) ? 1 : 0;
            r.half = (t.bits & TB_HALF) ? 1 : 0;
            r.halfDir = t.half_dir;
            out.put(r);
        }
    }
    if (!out.close()) return failWith(err, path, "write failed");
    return true;
}

bool readTiltFile(const std::string& path, Octree& o, std::string& err) {
    In in(path);
    if (!in.ok()) return failWith(err, path, "cannot open");
    int L = 0;
    if (!in.get(L)) return failWith(err, path, "truncated");
    if (L != o.L) return failWith(err, path, "level count differs from the context's layout");
    std::vector<std::vector<TiltNode>> t(L);
    for (int l = 0; l < L; ++l) {
        int size = 0;
        if (!in.get(size)) return failWith(err, path, "truncated");
        if (size != o.lv[l].nnode) return failWith(err, path, "level size differs from the layout");
        t[l] = o.tilt[l];
        for (int i = 0; i < size; ++i) {
            TiltRec r;
            if (!in.get(r)) return failWith(err, path, "truncated");
            if (r.halfDir < -128 || r.halfDir > 127) return failWith(err, path, "half direction out of range");
            // a half-tilt's face direction indexes the six faces (HALF_TILT_SLOT_INDEX)
            if (r.half && (r.halfDir < 0 || r.halfDir > 5))
                return failWith(err, path, "half-tilt with a half direction outside 0..5");
            uint8_t b = t[l][i].bits & (TB_LAYOUT_OPEN | TB_LAYOUT_FIXED);
            if (r.largest) b |= TB_LARGEST;
            if (r.closed) b |= TB_CLOSED;
            if (r.fixed) b |= TB_FIXED;
            if (r.tjoint) b |= TB_TJOINT;
            if (r.closedByTJ) b |= TB_CLOSED_BY_TJ;
            if (r.half) b |= TB_HALF;
            t[l][i].e = r.e;
            t[l][i].bits = b;
            t[l][i].half_dir = (int8_t)r.halfDir;
        }
    }
    if (!in.atEnd()) return failWith(err, path, "trailing bytes");
    o.tilt.swap(t);
    return true;
}

bool writeScalarFile(const std::string& path, const OmniScalarFrame& f, std::string& err) {
    Out out(path);
    if (!out.ok()) return failWith(err, path, "cannot open for writing");
    writeLevels(out, f.oct, 1);
    writeLevels(out, f.tilt, 1);
    if (!out.close()) return failWith(err, path, "write failed");
    return true;
}

bool readScalarFile(const std::string& path, const Octree& o, OmniScalarFrame& f, std::string& err) {
    In in(path);
    if (!in.ok()) return failWith(err, path, "cannot open");
    if (!readLevels(in, path, o, 1, f.oct, err, "octagon values")) return false;
    if (!readLevels(in, path, o, 1, f.tilt, err, "tilt values")) return false;
    if (!in.atEnd()) return failWith(err, path, "trailing bytes");
    return true;
}

bool writeVelocityFile(const std::string& path, const OmniVelocityFrame& f, std::string& err) {
    Out out(path);
    if (!out.ok()) return failWith(err, path, "cannot open for writing");
    // ux, uy, uz as three per-level blocks, then the 8 tilt slots per node
    for (int a = 0; a < 3; ++a) {
        out.put((int)f.mac.size());
        for (const auto& m : f.mac) {
            const size_t nn = m.size() / 3;
            out.put((int)nn);
            for (size_t i = 0; i < nn; ++i) out.put(m[3 * i + a]);
        }
    }
    writeLevels(out, f.tilt8, 8);
    if (!out.close()) return failWith(err, path, "write failed");
    return true;
}

bool readVelocityFile(const std::string& path, const Octree& o, OmniVelocityFrame& f, std::string& err) {
    In in(path);
    if (!in.ok()) return failWith(err, path, "cannot open");
    f.mac.assign(o.L, {});
    for (int l = 0; l < o.L; ++l) f.mac[l].assign((size_t)o.lv[l].nnode * 3, 0.0);
    static const char* kAxis[3] = {"ux", "uy", "uz"};
    for (int a = 0; a < 3; ++a) {
        std::vector<std::vector<double>> comp;
        if (!readLevels(in, path, o, 1, comp, err, kAxis[a])) return false;
        for (int l = 0; l < o.L; ++l)
            for (size_t i = 0; i < comp[l].size(); ++i) f.mac[l][3 * i + a] = comp[l][i];
    }
    if (!readLevels(in, path, o, 8, f.tilt8, err, "tilt velocity")) return false;
    if (!in.atEnd()) return failWith(err, path, "trailing bytes");
    return true;
}

}  // namespace astg
