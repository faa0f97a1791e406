// Test infrastructure only: the CPU oracle.  Nothing outside tests/, the smoke
// check and bench.py's cpu_baseline leg may link or call this code.
//
// Restatement of DualContouringMesher::mesh (src/voxel/meshing/
// dual_contouring.rs:23-83): edge crossings with interpolated gradient
// normals (:249-387), one feature point per sign-changing cube by the damped
// particle iteration (:176-247, :429-458), two triangles per sign-changing
// edge whose four cells exist (:93-174), scaled by the voxel size, degenerate
// triangles dropped (triangle3.rs:272-275).  Sequential: leaves in origin
// order, voxels x, y, z, edges X, Y, Z -- one of the orders the reference's
// parallel visit can produce.
#pragma once
#include <map>
#include <stdexcept>
#include <vector>

#include "vdb.hpp"

namespace oracle {

struct DcError : std::runtime_error {
    int code;  // 1: tile (todo!), 2: gradient without neighbours (unreachable!)
    DcError(int c, const char* m) : std::runtime_error(m), code(c) {}
};

struct IntPoint {
    Vec3f point, normal;
};

// dual_contouring.rs:389-423 and utils.rs:86-95
constexpr int kCubeOffsets[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
constexpr int kEdgeOffsets[12][4] = {{0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 2}, {1, 0, 0, 1},
                                     {1, 0, 0, 2}, {0, 0, 1, 0}, {0, 0, 1, 1}, {0, 1, 0, 0},
                                     {0, 1, 0, 2}, {1, 0, 1, 1}, {0, 1, 1, 0}, {1, 1, 0, 2}};
constexpr int kCellOffsets[3][4][3] = {{{0, 0, 0}, {0, 0, -1}, {0, -1, -1}, {0, -1, 0}},
                                       {{0, 0, 0}, {-1, 0, 0}, {-1, 0, -1}, {0, 0, -1}},
                                       {{0, -1, 0}, {-1, -1, 0}, {-1, 0, 0}, {0, 0, 0}}};

class DualContouring {
public:
    explicit DualContouring(const VolumeGrid& g) : g_(g) {}

    std::vector<Vec3f> mesh(float voxel_size) {
        std::vector<const LeafNode*> leaves;
        bool tile = false;
        g_.visit_leafs([&](Vec3i, int, float) { tile = true; }, [&](const LeafNode& l) { leaves.push_back(&l); });
        if (tile) throw DcError(1, "dual contouring meets a tile (dual_contouring.rs:146 todo!)");
        std::sort(leaves.begin(), leaves.end(), [](const LeafNode* a, const LeafNode* b) { return a->origin < b->origin; });
        for (const LeafNode* l : leaves) edges(*l);
        for (const LeafNode* l : leaves) cells(*l);
        std::vector<Vec3f> verts;
        for (const LeafNode* l : leaves) triangulate(*l, verts);
        // dual_contouring.rs:62-79
        std::vector<Vec3f> out;
        for (size_t i = 0; i + 2 < verts.size(); i += 3) {
            const Vec3f v0 = mul(verts[i], voxel_size), v1 = mul(verts[i + 1], voxel_size),
                        v2 = mul(verts[i + 2], voxel_size);
            if (norm_sq(cross(sub(v1, v0), sub(v2, v0))) == 0.0f) continue;
            out.push_back(v0);
            out.push_back(v1);
            out.push_back(v2);
        }
        return out;
    }

private:
    const VolumeGrid& g_;
    std::map<Vec3i, IntPoint> ints_[3];
    std::map<Vec3i, Vec3f> cells_;

    static Vec3i step(const Vec3i& v, int axis, int64_t d) {
        Vec3i r = v;
        r[axis] += d;
        return r;
    }

    // dual_contouring.rs:332-349
    float grad(const Vec3i& p, int axis) const {
        const float* vp = g_.at(p);
        const float* vl = g_.at(step(p, axis, -1));
        const float* vr = g_.at(step(p, axis, 1));
        if (vp && vl && vr) return (*vr - *vl) * 0.5f;
        if (vp && vr) return *vr - *vp;
        if (vp && vl) return *vp - *vl;
        throw DcError(2, "dual contouring gradient without neighbours (dual_contouring.rs:347 unreachable!)");
    }

    // dual_contouring.rs:257-330
    void intersection(const Vec3i& v1, int dir) {
        const Vec3i v2 = step(v1, dir, 1);
        const float* a = g_.at(v1);
        const float* b = g_.at(v2);
        if (!a || !b) return;
        const float v1v = *a, v2v = *b;
        if (std::signbit(v1v) == std::signbit(v2v)) return;
        const float t = (v1v == v2v) ? 0.5f : v1v / (v1v - v2v);
        Vec3f point{(float)v1.x, (float)v1.y, (float)v1.z};
        point[dir] = point[dir] + t;
        float n[3];
        for (int ax = 0; ax < 3; ++ax) {
            const float g1 = grad(v1, ax);
            const float g2 = grad(v2, ax);
            n[ax] = (1.0f - t) * g1 + t * g2;
        }
        const Vec3f nv{n[0], n[1], n[2]};
        const float len = std::sqrt(norm_sq(nv));
        ints_[dir][v1] = IntPoint{point, Vec3f{nv.x / len, nv.y / len, nv.z / len}};
    }

    void edges(const LeafNode& l) {
        for (int x = 0; x < 8; ++x)
            for (int y = 0; y < 8; ++y)
                for (int z = 0; z < 8; ++z) {
                    const Vec3i v{l.origin.x + x, l.origin.y + y, l.origin.z + z};
                    for (int d = 0; d < 3; ++d) intersection(v, d);
                }
    }

public:
    // dual_contouring.rs:433-458
    static Vec3f feature_point(const std::vector<IntPoint>& pts) {
        const float cnt = (float)pts.size();
        Vec3f c{0.0f, 0.0f, 0.0f};
        for (const IntPoint& p : pts) c = add(c, p.point);
        c = Vec3f{c.x / cnt, c.y / cnt, c.z / cnt};
        for (int i = 0; i < 50; ++i) {
            Vec3f f{0.0f, 0.0f, 0.0f};
            for (const IntPoint& p : pts) {
                const Vec3f n = p.normal;
                const float d = dot(n, sub(c, p.point));
                f = add(f, mul(mul(n, -1.0f), d));
            }
            const float damping = 1.0f - (float)i / 50.0f;
            const Vec3f step_ = mul(f, damping);
            c = add(c, Vec3f{step_.x / cnt, step_.y / cnt, step_.z / cnt});
            if (norm_sq(f) < 1e-6f) break;
        }
        return c;
    }

private:
    // dual_contouring.rs:176-247
    void cells(const LeafNode& l) {
        std::vector<IntPoint> pts;
        for (int x = 0; x < 8; ++x)
            for (int y = 0; y < 8; ++y)
                for (int z = 0; z < 8; ++z) {
                    const Vec3i o{l.origin.x + x, l.origin.y + y, l.origin.z + z};
                    float vals[8];
                    bool all = true;
                    for (int i = 0; i < 8 && all; ++i) {
                        const float* v = g_.at({o.x + kCubeOffsets[i][0], o.y + kCubeOffsets[i][1], o.z + kCubeOffsets[i][2]});
                        if (!v) all = false;
                        else vals[i] = *v;
                    }
                    if (!all) continue;
                    bool same = true;
                    for (int i = 1; i < 8; ++i) same &= std::signbit(vals[i]) == std::signbit(vals[0]);
                    if (same) continue;
                    pts.clear();
                    for (const auto& e : kEdgeOffsets) {
                        const Vec3i p{o.x + e[0], o.y + e[1], o.z + e[2]};
                        auto it = ints_[e[3]].find(p);
                        if (it != ints_[e[3]].end()) pts.push_back(it->second);
                    }
                    cells_[o] = feature_point(pts);
                }
    }

    // dual_contouring.rs:93-174
    void triangulate(const LeafNode& l, std::vector<Vec3f>& out) {
        for (int x = 0; x < 8; ++x)
            for (int y = 0; y < 8; ++y)
                for (int z = 0; z < 8; ++z) {
                    const Vec3i v{l.origin.x + x, l.origin.y + y, l.origin.z + z};
                    const float* vv = l.at(v);
                    if (!vv) continue;
                    for (int d = 0; d < 3; ++d) {
                        const float* w = g_.at(step(v, d, 1));
                        if (!w) continue;
                        if (std::signbit(*vv) == std::signbit(*w)) continue;
                        Vec3f p[4];
                        bool ok = true;
                        for (int k = 0; k < 4 && ok; ++k) {
                            auto it = cells_.find({v.x + kCellOffsets[d][k][0], v.y + kCellOffsets[d][k][1],
                                                   v.z + kCellOffsets[d][k][2]});
                            if (it == cells_.end()) ok = false;
                            else p[k] = it->second;
                        }
                        if (!ok) continue;
                        Vec3f f[6] = {p[0], p[1], p[2], p[2], p[3], p[0]};
                        if (std::signbit(*vv)) {
                            std::swap(f[1], f[2]);
                            std::swap(f[4], f[5]);
                        }
                        out.insert(out.end(), f, f + 6);
                    }
                }
    }
};

}  // namespace oracle
