// Test infrastructure only: the CPU oracle.  Nothing outside tests/, the smoke
// check and bench.py's cpu_baseline leg may link or call this code.
#pragma once
#include <memory>
#include <vector>

#include "parallel.hpp"
#include "vdb.hpp"
#include "winding.hpp"

namespace oracle {

struct VoxelBox {
    Vec3i min{0, 0, 0}, max{0, 0, 0};
    int64_t volume() const { return (max.x - min.x + 1) * (max.y - min.y + 1) * (max.z - min.z + 1); }
};

float tri_max_side(const Tri& t);
Vec3f closest_point(const Tri& t, const Vec3f& p);
float point_tri_distance(const Tri& t, const Vec3f& p);
double point_tri_distance_f64(const Tri& t, const Vec3f& p);
void subdivide_triangle(const Tri& tri, float voxel_size, std::vector<Tri>& out);
VoxelBox dilated_box(const Tri& t, float inv_voxel, int64_t band);

// mesh_to_volume.rs:17-50, defaults :223-236
struct MeshToVolume {
    int64_t band_width = 0;
    float voxel_size = 1.0f;
    float inverse_voxel_size = 1.0f;
    int threads = default_threads();
    std::vector<Tri> subdivided;
    WindingNumbers wn;

    MeshToVolume();
    void set_voxel_size(float s);
    std::unique_ptr<VolumeGrid> convert(const std::vector<Tri>& mesh);  // :52-73; nullptr = None
    std::unique_ptr<VolumeGrid> unsigned_distance_field();
    std::unique_ptr<VolumeGrid> signs(const VolumeGrid& df);
};

}  // namespace oracle
