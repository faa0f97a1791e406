// Test infrastructure only: the CPU oracle.  Nothing outside tests/, the smoke
// check and bench.py's cpu_baseline leg may link or call this code.
//
// Restatement of src/spatial_partitioning/aabb_tree.rs: the top-down AABB tree
// with the `Area` (12-bucket SAH) partition (:67-80, :118-126, :157-271,
// :419-514) and the `winding_numbers` module (:571-817): order-2 dipole
// coefficients per node and the recursive fast winding number.
#pragma once
#include <vector>

#include "vec.hpp"

namespace oracle {

struct Tri {
    Vec3f a, b, c;
};

struct Box3f {
    Vec3f min, max;
    static Box3f empty() {  // box3.rs:22-27
        return {{3.40282347e+38f, 3.40282347e+38f, 3.40282347e+38f},
                {-3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f}};
    }
    void union_box(const Box3f& o) {  // box3.rs:65-70
        max = max2(max, o.max);
        min = min2(min, o.min);
    }
    void union_point(const Vec3f& p) {
        max = max2(max, p);
        min = min2(min, p);
    }
    Vec3f center() const { return divs(add(min, max), 2.0f); }  // box3.rs:40-42
    float size(int i) const { return max[i] - min[i]; }
    float area() const {  // box3.rs:125-128
        Vec3f d = sub(max, min);
        return 2.0f * ((d.x * d.y + d.x * d.z) + d.y * d.z);
    }
    Vec3f offset(const Vec3f& p) const {  // box3.rs:185-201
        Vec3f o = sub(p, min);
        if (max.x > min.x) o.x /= max.x - min.x;
        if (max.y > min.y) o.y /= max.y - min.y;
        if (max.z > min.z) o.z /= max.z - min.z;
        return o;
    }
};

// triangle3.rs:307-315
inline Box3f tri_bbox(const Tri& t) { return {min2(t.c, min2(t.a, t.b)), max2(t.c, max2(t.a, t.b))}; }

struct BvhNode {
    bool leaf;
    size_t left, right;  // objects [left,right) for a leaf, child nodes for a branch
    Box3f bbox;
};

struct DipoleData {  // aabb_tree.rs:719-726
    Vec3f order1;
    Mat3f order2;
    float radius;
    Vec3f center;
};

class WindingNumbers {
public:
    // aabb_tree.rs:636-644: AABBTree::from_mesh -> min 3 objects per leaf -> Area
    static WindingNumbers from_mesh(const std::vector<Tri>& tris);
    // aabb_tree.rs:654-660
    float approximate(const Vec3f& p, float accuracy_scale) const;
    // FP64 evaluation over the same tree: same traversal decisions as the f32
    // path, every term evaluated in double.  Used to report error, not parity.
    double approximate_f64(const Vec3f& p, float accuracy_scale) const;

    std::vector<BvhNode> nodes;  // root is last (post-order)
    std::vector<Tri> objects;
    std::vector<Box3f> object_boxes;
    std::vector<DipoleData> data;

    // node visits of approximate(p): internal nodes opened, far evaluations, leaves summed
    void count_visits(size_t node, const Vec3f& p, float acc, uint64_t counts[3]) const;

private:
    float fast_wn(size_t node, const Vec3f& p, float acc) const;
    double fast_wn_f64(size_t node, const Vec3f& p, float acc) const;
};

float solid_angle(const Tri& t, const Vec3f& q);         // aabb_tree.rs:582-615
double solid_angle_f64(const Tri& t, const Vec3f& q);
float winding_number_exact(const std::vector<Tri>& tris, const Vec3f& q);  // brute force Σ Ω / 4π

}  // namespace oracle
