// Test infrastructure only: the CPU oracle.  Nothing outside tests/, the smoke
// check and bench.py's cpu_baseline leg may link or call this code.
//
// Restatement of src/voxel/mesh_to_volume.rs with the reference's parallel
// decomposition kept: distances per sub-triangle in parallel, a serial merge
// into the tree, then signs per leaf in parallel with one global lock.
#include "mesh_to_volume.hpp"

#include <cmath>
#include <mutex>

#include "parallel.hpp"

namespace oracle {

// triangle3.rs:117-124
float tri_max_side(const Tri& t) {
    float ab = norm_sq(sub(t.b, t.a));
    float ac = norm_sq(sub(t.c, t.a));
    float bc = norm_sq(sub(t.c, t.b));
    return std::sqrt(rmax(rmax(ab, ac), bc));
}

// triangle3.rs:317-382 (Ericson, Real-Time Collision Detection 5.1.5)
template <class T>
V3<T> closest_point_t(const V3<T>& A, const V3<T>& B, const V3<T>& C, const V3<T>& p) {
    const T zero = T(0);
    V3<T> ab = sub(B, A), ac = sub(C, A), ap = sub(p, A);
    T d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= zero && d2 <= zero) return A;
    V3<T> bp = sub(p, B);
    T d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= zero && d4 <= d3) return B;
    T vc = d1 * d4 - d3 * d2;
    if (vc <= zero && d1 >= zero && d3 <= zero) {
        T v = d1 / (d1 - d3);
        return add(A, mul(ab, v));
    }
    V3<T> cp = sub(p, C);
    T d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= zero && d5 <= d6) return C;
    T vb = d5 * d2 - d1 * d6;
    if (vb <= zero && d2 >= zero && d6 <= zero) {
        T w = d2 / (d2 - d6);
        return add(A, mul(ac, w));
    }
    T va = d3 * d6 - d5 * d4;
    if (va <= zero && (d4 - d3) >= zero && (d5 - d6) >= zero) {
        T w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return add(B, mul(sub(C, B), w));
    }
    T denom = T(1) / ((va + vb) + vc);
    T v = vb * denom, w = vc * denom;
    return add(add(A, mul(ab, v)), mul(ac, w));
}

Vec3f closest_point(const Tri& t, const Vec3f& p) { return closest_point_t(t.a, t.b, t.c, p); }

float point_tri_distance(const Tri& t, const Vec3f& p) { return norm(sub(closest_point(t, p), p)); }

double point_tri_distance_f64(const Tri& t, const Vec3f& pf) {
    auto D = [](const Vec3f& v) { return Vec3d{v.x, v.y, v.z}; };
    Vec3d p = D(pf);
    return norm(sub(closest_point_t(D(t.a), D(t.b), D(t.c), p), p));
}

// mesh_to_volume.rs:75-116
void subdivide_triangle(const Tri& tri, float voxel_size, std::vector<Tri>& out) {
    float num_subs = std::floor(tri_max_side(tri) / voxel_size);
    float num_subs_inv = 1.0f / num_subs;
    if (num_subs < 2.0f) {
        out.push_back(tri);
        return;
    }
    Vec3f s1 = mul(sub(tri.b, tri.a), num_subs_inv);
    Vec3f s2 = mul(sub(tri.c, tri.b), num_subs_inv);
    Vec3f a = tri.a;
    uint64_t n = f2u(num_subs);
    for (uint64_t i = 0; i < n; ++i) {
        Vec3f b = add(a, s1);
        Vec3f c = add(b, s2);
        Vec3f a_s = add(a, s2), b_s = add(b, s2), c_s = add(c, s2);
        Vec3f a_prev = a;
        for (uint64_t k = 0; k < i; ++k) {
            out.push_back({a_prev, b_s, a_s});
            out.push_back({a_s, b_s, c_s});
            a_prev = a_s;
            a_s = add(a_s, s2);
            b_s = add(b_s, s2);
            c_s = add(c_s, s2);
        }
        out.push_back({a, b, c});
        a = add(a, s1);
    }
}

// mesh_to_volume.rs:124-141
VoxelBox dilated_box(const Tri& t, float inv_voxel, int64_t band) {
    Box3f bb = tri_bbox(t);
    VoxelBox b;
    for (int k = 0; k < 3; ++k) {
        b.min[k] = f2i(std::floor(bb.min[k] * inv_voxel)) - band;
        b.max[k] = f2i(std::ceil(bb.max[k] * inv_voxel)) + band;
    }
    if (b.max.x == b.min.x || b.max.y == b.min.y || b.max.z == b.min.z) {
        for (int k = 0; k < 3; ++k) {
            b.min[k] -= 1;
            b.max[k] += 1;
        }
    }
    return b;
}

MeshToVolume::MeshToVolume() { set_voxel_size(1.0f); }

void MeshToVolume::set_voxel_size(float s) {
    voxel_size = s;
    inverse_voxel_size = 1.0f / s;
}

std::unique_ptr<VolumeGrid> MeshToVolume::convert(const std::vector<Tri>& mesh) {
    subdivided.clear();
    for (auto& t : mesh) subdivide_triangle(t, voxel_size, subdivided);
    if (subdivided.empty()) return nullptr;
    wn = WindingNumbers::from_mesh(mesh);
    auto df = unsigned_distance_field();
    return signs(*df);
}

// mesh_to_volume.rs:118-196.  Sub-triangles are processed in chunks so that the
// per-triangle distance vectors never all live at once; the merge visits the
// chunks in sub-triangle order, so the tree receives the same insert sequence.
std::unique_ptr<VolumeGrid> MeshToVolume::unsigned_distance_field() {
    auto df = std::make_unique<VolumeGrid>();
    const size_t CHUNK = 1 << 16;
    std::vector<VoxelBox> boxes;
    std::vector<std::vector<float>> dists;
    for (size_t base = 0; base < subdivided.size(); base += CHUNK) {
        size_t n = std::min(CHUNK, subdivided.size() - base);
        boxes.assign(n, VoxelBox{});
        dists.assign(n, {});
        par_for(n, threads, [&](size_t k) {
            const Tri& tri = subdivided[base + k];
            VoxelBox b = dilated_box(tri, inverse_voxel_size, band_width);
            boxes[k] = b;
            auto& d = dists[k];
            d.reserve((size_t)b.volume());
            for (int64_t x = b.min.x; x <= b.max.x; ++x) {
                float xw = (float)x * voxel_size;
                for (int64_t y = b.min.y; y <= b.max.y; ++y) {
                    float yw = (float)y * voxel_size;
                    for (int64_t z = b.min.z; z <= b.max.z; ++z) {
                        float zw = (float)z * voxel_size;
                        d.push_back(point_tri_distance(tri, {xw, yw, zw}));
                    }
                }
            }
        });
        for (size_t k = 0; k < n; ++k) {  // serial merge, mesh_to_volume.rs:170-195
            const VoxelBox& b = boxes[k];
            const auto& d = dists[k];
            size_t i = 0;
            for (int64_t x = b.min.x; x <= b.max.x; ++x)
                for (int64_t y = b.min.y; y <= b.max.y; ++y)
                    for (int64_t z = b.min.z; z <= b.max.z; ++z) {
                        Vec3i idx{x, y, z};
                        const float* cur = df->at(idx);
                        float cd = cur ? *cur : INFINITY;
                        if (d[i] < cd) df->insert(idx, d[i]);
                        ++i;
                    }
        }
    }
    return df;
}

// mesh_to_volume.rs:198-292: per leaf in parallel, one global lock per voxel.
std::unique_ptr<VolumeGrid> MeshToVolume::signs(const VolumeGrid& df) {
    auto out = std::make_unique<VolumeGrid>();
    std::mutex lock;
    std::vector<const LeafNode*> leaves;
    df.visit_leafs([](Vec3i, int, float) {}, [&](const LeafNode& l) { leaves.push_back(&l); });
    par_for(leaves.size(), threads, [&](size_t li) {
        const LeafNode& node = *leaves[li];
        Vec3i o = node.origin;
        for (int64_t x = o.x; x < o.x + 8; ++x)
            for (int64_t y = o.y; y < o.y + 8; ++y)
                for (int64_t z = o.z; z < o.z + 8; ++z) {
                    Vec3i idx{x, y, z};
                    const float* v = node.at(idx);
                    if (!v) continue;
                    Vec3f p{(float)x * voxel_size, (float)y * voxel_size, (float)z * voxel_size};
                    float w = wn.approximate(p, 2.0f);
                    float dist = (w < 0.2f) ? std::copysign(*v, 1.0f) : std::copysign(*v, -1.0f);
                    std::lock_guard<std::mutex> g(lock);
                    out->insert(idx, dist);
                }
    }, 1);
    return out;
}

}  // namespace oracle
