// Test infrastructure only: the CPU oracle.  Nothing outside tests/, the smoke
// check and bench.py's cpu_baseline leg may link or call this code.
//
// Restatement of Volume::offset (src/voxel/volume/mod.rs:95-108): remove_if
// |v| > 2h, FastSweeping (src/voxel/fast_sweep.rs:10-287) with its leaf queue
// (:289-343), stencil (:345-387) and sweep orders (:612-684), Godunov update
// compute_distance (:509-537) with sort3 (src/helpers/utils.rs:4-16), removal
// of empty branches, then v -= D on every active value
// (visitors/value_mut_visitor.rs:4-32).  Sequential, like the reference.
#pragma once
#include <functional>
#include <queue>
#include <set>
#include <stdexcept>

#include "vdb.hpp"

namespace oracle {

// helpers/utils.rs:4-16
inline void sort3(float& a, float& b, float& c) {
    if (a > c) std::swap(a, c);
    if (a > b) std::swap(a, b);
    if (b > c) std::swap(b, c);
}

// fast_sweep.rs:509-537, f32 throughout
inline float fs_compute_distance(float a1, float a2, float a3, float h) {
    sort3(a1, a2, a3);
    const float s1 = a1 + h;
    if (std::fabs(s1) <= a2) return s1;
    const float a1a2_sum = a1 + a2;
    const float h_sq = h * h;
    const float two_h_sq = h_sq + h_sq;
    const float a1a2_diff = a1 - a2;
    const float a1a2_diff_sq = a1a2_diff * a1a2_diff;
    const float s2 = (a1a2_sum + std::sqrt(two_h_sq - a1a2_diff_sq)) * 0.5f;
    if (std::fabs(s2) <= a3) return s2;
    const float a1a2a3_sum = a1a2_sum + a3;
    const float three_h_sq = two_h_sq + h_sq;
    const float a1a3_diff = a1 - a3;
    const float a1a3_diff_sq = a1a3_diff * a1a3_diff;
    const float a2a3_diff = a2 - a3;
    const float a2a3_diff_sq = a2a3_diff * a2a3_diff;
    return (a1a2a3_sum + std::sqrt(three_h_sq - a1a2_diff_sq - a1a3_diff_sq - a2a3_diff_sq)) * (1.0f / 3.0f);
}

struct FastSweepError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// One sweep direction: +1 / -1 per axis, in the order of fast_sweep.rs:36-57.
struct SweepDir {
    int sx, sy, sz;
};
constexpr SweepDir kSweeps[8] = {{1, 1, 1},  {-1, 1, 1},  {1, -1, 1},  {-1, -1, 1},
                                 {1, 1, -1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};

class FastSweep {
public:
    FastSweep(float grid_spacing, float limit)
        : h_(grid_spacing), limit_abs_(std::fabs(limit)), sweep_neg_(std::signbit(limit)) {}

    // fast_sweep.rs:31-65
    void run(VolumeGrid& sdf) {
        frozen_ = sdf.clone();
        // CollectLeafIndicesVisitor: every leaf and every active tile origin
        std::vector<Vec3i> origins;
        sdf.visit_leafs([&](Vec3i o, int, float) { origins.push_back(o); },
                        [&](const LeafNode& l) { origins.push_back(l.origin); });
        for (const SweepDir& d : kSweeps) origins = sweep(sdf, d, origins);
        sdf.remove_empty_branches();
    }

private:
    float h_, limit_abs_;
    bool sweep_neg_;
    std::unique_ptr<VolumeGrid> frozen_;

    // Order of the leaf queue: per axis ascending for a positive sweep
    // direction, descending for a negative one, x then y then z
    // (fast_sweep.rs:612-684 with the reversed max-heap).
    static bool before(const SweepDir& d, const Vec3i& a, const Vec3i& b) {
        if (a.x != b.x) return d.sx > 0 ? a.x < b.x : a.x > b.x;
        if (a.y != b.y) return d.sy > 0 ? a.y < b.y : a.y > b.y;
        if (a.z != b.z) return d.sz > 0 ? a.z < b.z : a.z > b.z;
        return false;
    }

    struct Stencil {
        std::unique_ptr<LeafNode> left, right, front, back, top, bottom, center;
        const LeafNode* frozen;
        Vec3i min, max;
        // fast_sweep.rs:359-387
        const float* at(const Vec3i& i) const {
            if (i.z < min.z) return bottom->at(i);
            if (i.z >= max.z) return top->at(i);
            if (i.y < min.y) return front->at(i);
            if (i.y >= max.y) return back->at(i);
            if (i.x < min.x) return left->at(i);
            if (i.x >= max.x) return right->at(i);
            return center->at(i);
        }
    };

    static std::unique_ptr<LeafNode> take_or_empty(VolumeGrid& sdf, const Vec3i& o) {
        auto l = sdf.take_leaf_at(o);
        return l ? std::move(l) : std::make_unique<LeafNode>(o);
    }

    // fast_sweep.rs:67-104 plus the queue of :289-343: returns the origins
    // popped in this sweep, the next sweep's queue.
    std::vector<Vec3i> sweep(VolumeGrid& sdf, const SweepDir& d, const std::vector<Vec3i>& start) {
        auto later = [&](const Vec3i& a, const Vec3i& b) { return before(d, b, a); };
        std::priority_queue<Vec3i, std::vector<Vec3i>, decltype(later)> heap(later);
        std::set<Vec3i> existing;
        for (const Vec3i& o : start)
            if (existing.insert(o).second) heap.push(o);
        std::vector<Vec3i> removed;
        const int64_t S = LeafNode::DIM;
        while (!heap.empty()) {
            const Vec3i o = heap.top();
            heap.pop();
            removed.push_back(o);
            Stencil st;
            st.left = take_or_empty(sdf, {o.x - S, o.y, o.z});
            st.right = take_or_empty(sdf, {o.x + S, o.y, o.z});
            st.front = take_or_empty(sdf, {o.x, o.y - S, o.z});
            st.back = take_or_empty(sdf, {o.x, o.y + S, o.z});
            st.top = take_or_empty(sdf, {o.x, o.y, o.z + S});
            st.bottom = take_or_empty(sdf, {o.x, o.y, o.z - S});
            st.center = sdf.take_leaf_at(o);
            if (!st.center) throw FastSweepError("leaf queue holds a tile origin (fast_sweep.rs:100 unwrap)");
            st.frozen = frozen_->leaf_at(o);
            st.min = o;
            st.max = {o.x + S, o.y + S, o.z + S};
            sweep_stencil(st, d);
            // fast_sweep.rs:186-273: push the +direction face neighbours that
            // hold a value of the sweep sign within the limit
            auto face_has = [&](int axis) {
                const int64_t f = (axis == 0 ? d.sx : axis == 1 ? d.sy : d.sz) > 0 ? S - 1 : 0;
                for (int64_t a = 0; a < S; ++a)
                    for (int64_t b = 0; b < S; ++b) {
                        Vec3i l = axis == 0 ? Vec3i{f, a, b} : axis == 1 ? Vec3i{a, f, b} : Vec3i{a, b, f};
                        const float* v = st.center->at({o.x + l.x, o.y + l.y, o.z + l.z});
                        if (v && std::signbit(*v) == sweep_neg_ && std::fabs(*v) < limit_abs_) return true;
                    }
                return false;
            };
            const bool ix = face_has(0), iy = face_has(1), iz = face_has(2);
            auto push = [&](const Vec3i& n) {
                if (existing.insert(n).second) heap.push(n);
            };
            if (ix) push({o.x + d.sx * S, o.y, o.z});
            if (iy) push({o.x, o.y + d.sy * S, o.z});
            if (iz) push({o.x, o.y, o.z + d.sz * S});
            sdf.insert_leaf_at(std::move(st.top));
            sdf.insert_leaf_at(std::move(st.bottom));
            sdf.insert_leaf_at(std::move(st.left));
            sdf.insert_leaf_at(std::move(st.right));
            sdf.insert_leaf_at(std::move(st.front));
            sdf.insert_leaf_at(std::move(st.back));
            sdf.insert_leaf_at(std::move(st.center));
        }
        return removed;
    }

    // fast_sweep.rs:106-118: x outer, y, z inner, each in the sweep direction
    void sweep_stencil(Stencil& st, const SweepDir& d) {
        const int64_t S = LeafNode::DIM;
        for (int64_t i = 0; i < S; ++i)
            for (int64_t j = 0; j < S; ++j)
                for (int64_t k = 0; k < S; ++k) {
                    const int64_t x = d.sx > 0 ? i : S - 1 - i;
                    const int64_t y = d.sy > 0 ? j : S - 1 - j;
                    const int64_t z = d.sz > 0 ? k : S - 1 - k;
                    sweep_voxel({st.min.x + x, st.min.y + y, st.min.z + z}, st);
                }
    }

    // utils.rs:64-73 option_min_by with fast_sweep.rs:137-140 cmp_abs: the
    // first argument unless the second has a strictly smaller magnitude.
    static const float* min_abs(const float* a, const float* b) {
        if (a && b) return (std::fabs(*a) > std::fabs(*b)) ? b : a;
        return a ? a : b;
    }

    // fast_sweep.rs:120-184
    void sweep_voxel(const Vec3i& i, Stencil& st) const {
        if (st.frozen && st.frozen->at(i)) return;
        const float* dx = min_abs(st.at({i.x + 1, i.y, i.z}), st.at({i.x - 1, i.y, i.z}));
        const float* dy = min_abs(st.at({i.x, i.y + 1, i.z}), st.at({i.x, i.y - 1, i.z}));
        const float* dz = min_abs(st.at({i.x, i.y, i.z + 1}), st.at({i.x, i.y, i.z - 1}));
        const float* first = dx ? dx : (dy ? dy : dz);
        if (!first) return;
        if (std::signbit(*first) != sweep_neg_) return;
        const float d1 = dx ? *dx : F32_MAX, d2 = dy ? *dy : F32_MAX, d3 = dz ? *dz : F32_MAX;
        const float d_new_abs = fs_compute_distance(std::fabs(d1), std::fabs(d2), std::fabs(d3), h_);
        const float d_new = std::copysign(d_new_abs, sweep_neg_ ? -1.0f : 1.0f);
        if (d_new_abs > limit_abs_) return;
        const float* old = st.center->at(i);
        const float d_old = old ? *old : F32_MAX;
        if (d_new_abs < std::fabs(d_old)) st.center->insert(i, d_new);
    }
};

// Volume::offset (volume/mod.rs:95-108)
inline void volume_offset(VolumeGrid& grid, float voxel_size, float distance) {
    const float two_h = voxel_size * 2.0f;
    grid.remove_if([&](float v) { return std::fabs(v) > two_h; });
    float ext = std::fabs(distance) + voxel_size + voxel_size;
    ext = std::copysign(ext, std::signbit(distance) ? -1.0f : 1.0f);
    FastSweep fs(voxel_size, ext);
    fs.run(grid);
    grid.visit_values_mut([&](float& v) { v -= distance; });
}

}  // namespace oracle
