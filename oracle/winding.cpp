// Test infrastructure only: the CPU oracle.  Nothing outside tests/, the smoke
// check and bench.py's cpu_baseline leg may link or call this code.
#include "winding.hpp"

#include <algorithm>
#include <cmath>

namespace oracle {

namespace {

constexpr float PI_F = 3.14159274f;  // std::f32::consts::PI
constexpr size_t MIN_OBJECTS_PER_LEAF = 3;  // aabb_tree.rs:639
constexpr size_t MAX_DEPTH = 40;            // aabb_tree.rs:70
constexpr size_t NUM_BUCKETS = 12;          // aabb_tree.rs:442

struct Builder {
    std::vector<Tri>& objs;
    std::vector<Box3f>& boxes;
    std::vector<BvhNode>& nodes;

    // aabb_tree.rs:254-271
    size_t leaf_node(size_t first, size_t last) {
        Box3f bb = boxes[first];
        for (size_t i = first + 1; i < last; ++i) bb.union_box(boxes[i]);
        nodes.push_back({true, first, last, bb});
        return nodes.size() - 1;
    }

    // aabb_tree.rs:419-514 (Area::split) on objects[first,last) sorted along `axis`
    bool area_split(size_t first, size_t last, int axis, const Box3f& objects_bbox, size_t& split_at) {
        if (first == last) return false;
        Box3f centroid_bounds = Box3f::empty();
        for (size_t i = first; i < last; ++i) centroid_bounds.union_point(boxes[i].center());
        const float nb = (float)NUM_BUCKETS;
        auto bucket_of = [&](const Box3f& b) {
            Vec3f c = b.center();
            uint64_t tmp = f2u(nb * centroid_bounds.offset(c)[axis]);
            return (size_t)std::min<uint64_t>(tmp, NUM_BUCKETS - 1);
        };
        size_t count[NUM_BUCKETS] = {0};
        Box3f bbox[NUM_BUCKETS];
        for (auto& b : bbox) b = Box3f::empty();
        for (size_t i = first; i < last; ++i) {
            size_t k = bucket_of(boxes[i]);
            count[k] += 1;
            bbox[k].union_box(boxes[i]);
        }
        float costs[NUM_BUCKETS - 1];
        for (size_t i = 0; i < NUM_BUCKETS - 1; ++i) {
            Box3f b0 = Box3f::empty(), b1 = Box3f::empty();
            size_t c0 = 0, c1 = 0;
            for (size_t j = 0; j <= i; ++j) { b0.union_box(bbox[j]); c0 += count[j]; }
            for (size_t j = i + 1; j < NUM_BUCKETS; ++j) { b1.union_box(bbox[j]); c1 += count[j]; }
            float f0 = (float)c0, f1 = (float)c1;
            if (f0 == 0.0f || f1 == 0.0f) costs[i] = INFINITY;
            else costs[i] = 0.125f + (f0 * b0.area() + f1 * b1.area()) / objects_bbox.area();
        }
        // Iterator::min_by keeps the first of equal minima
        size_t best = 0;
        for (size_t i = 1; i < NUM_BUCKETS - 1; ++i)
            if (costs[i] < costs[best]) best = i;
        float leaf_cost = (float)(last - first);
        if (costs[best] < leaf_cost) {
            for (size_t i = first; i < last; ++i)
                if (bucket_of(boxes[i]) > best) { split_at = i; return true; }
            return false;
        }
        return false;
    }

    // aabb_tree.rs:237-251: stable sort by bbox centre along axis, then partition
    bool sort_and_try(size_t first, size_t last, int axis, const Box3f& bb, size_t& split_at) {
        std::vector<size_t> idx(last - first);
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = first + i;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t p, size_t q) {
            return boxes[p].center()[axis] < boxes[q].center()[axis];
        });
        std::vector<Tri> to(idx.size());
        std::vector<Box3f> tb(idx.size());
        for (size_t i = 0; i < idx.size(); ++i) { to[i] = objs[idx[i]]; tb[i] = boxes[idx[i]]; }
        std::copy(to.begin(), to.end(), objs.begin() + first);
        std::copy(tb.begin(), tb.end(), boxes.begin() + first);
        return area_split(first, last, axis, bb, split_at);
    }

    // aabb_tree.rs:201-235
    bool split(size_t first, size_t last, size_t& split_at) {
        Box3f bb = boxes[first];
        for (size_t i = first; i < last; ++i) bb.union_box(boxes[i]);
        int axes[3] = {0, 1, 2};
        float sz[3] = {bb.size(0), bb.size(1), bb.size(2)};
        std::stable_sort(axes, axes + 3, [&](int p, int q) { return sz[q] < sz[p]; });
        for (int a = 0; a < 3; ++a)
            if (sort_and_try(first, last, axes[a], bb, split_at)) return true;
        return false;
    }

    // aabb_tree.rs:157-199
    size_t build(size_t first, size_t last, size_t depth) {
        if (depth >= MAX_DEPTH || last - first <= MIN_OBJECTS_PER_LEAF) return leaf_node(first, last);
        size_t split_at = 0;
        if (!split(first, last, split_at)) return leaf_node(first, last);
        size_t l = build(first, split_at, depth + 1);
        size_t r = build(split_at, last, depth + 1);
        Box3f bb = nodes[l].bbox;
        bb.union_box(nodes[r].bbox);
        nodes.push_back({false, l, r, bb});
        return nodes.size() - 1;
    }
};

struct InitData {  // aabb_tree.rs:711-717
    Vec3f awn, awc;
    Mat3f order1_sum;
    float total_area;
    Vec3f center;
};

// triangle3.rs:261-269, 278-280, 113-115
bool tri_normal(const Tri& t, Vec3f& n) {
    Vec3f cr = cross(sub(t.b, t.a), sub(t.c, t.a));
    if (norm_sq(cr) == 0.0f) return false;
    n = divs(cr, norm(cr));  // nalgebra normalize: v / |v|
    return true;
}
float tri_area(const Tri& t) { return norm(cross(sub(t.b, t.a), sub(t.c, t.a))) * 0.5f; }
Vec3f tri_center(const Tri& t) { return divs(add(add(t.a, t.b), t.c), 3.0f); }

InitData node_data(const WindingNumbers& w, size_t idx, std::vector<DipoleData>& out);

InitData leaf_data(const WindingNumbers& w, const BvhNode& n) {  // aabb_tree.rs:747-778
    InitData d{{0, 0, 0}, {0, 0, 0}, mat_zero(), 0.0f, {0, 0, 0}};
    for (size_t t = n.left; t < n.right; ++t) {
        const Tri& tri = w.objects[t];
        Vec3f nn;
        if (!tri_normal(tri, nn)) continue;
        float area = tri_area(tri);
        d.total_area += area;
        d.awn = add(d.awn, smul(area, nn));
        Vec3f c = tri_center(tri);
        d.order1_sum = mat_add(d.order1_sum, outer(smul(area, c), nn));
        d.awc = add(d.awc, smul(area, c));
    }
    d.center = divs(d.awc, d.total_area);
    return d;
}

InitData branch_data(const WindingNumbers& w, const BvhNode& n, std::vector<DipoleData>& out) {
    InitData l = node_data(w, n.left, out);  // aabb_tree.rs:780-801
    InitData r = node_data(w, n.right, out);
    InitData d;
    d.order1_sum = mat_add(l.order1_sum, r.order1_sum);
    d.awn = add(l.awn, r.awn);
    d.awc = add(l.awc, r.awc);
    d.total_area = l.total_area + r.total_area;
    d.center = divs(add(l.awc, r.awc), d.total_area);
    return d;
}

InitData node_data(const WindingNumbers& w, size_t idx, std::vector<DipoleData>& out) {
    const BvhNode& n = w.nodes[idx];  // aabb_tree.rs:728-745
    InitData d = n.leaf ? leaf_data(w, n) : branch_data(w, n, out);
    float dmin = norm_sq(sub(n.bbox.min, d.center));
    float dmax = norm_sq(sub(n.bbox.max, d.center));
    float radius = std::sqrt(rmax(dmin, dmax));
    out[idx] = {d.awn, mat_sub(d.order1_sum, outer(d.center, d.awn)), radius, d.center};
    return d;
}

// aabb_tree.rs:803-816
void hessians(const Vec3f& dipole, const Vec3f& q, Vec3f& ord1, Mat3f& ord2) {
    Vec3f r = sub(dipole, q);
    float r2 = norm_sq(r);
    float r1 = std::sqrt(r2);
    float r3 = r2 * r1;
    float den = 4.0f * PI_F * r3;
    float inv = 1.0f / den;
    ord1 = mul(r, inv);
    float r5 = r3 * r2;
    float den2 = 4.0f * PI_F * r5;
    Mat3f rr = outer(smul(3.0f, r), r);
    for (int c = 0; c < 3; ++c)
        for (int rw = 0; rw < 3; ++rw) {
            float id = (rw == c ? 1.0f : 0.0f) * inv;
            ord2.at(rw, c) = id - rr.at(rw, c) / den2;
        }
}

}  // namespace

float solid_angle(const Tri& t, const Vec3f& q) {
    Vec3f qa = sub(t.a, q), qb = sub(t.b, q), qc = sub(t.c, q);
    float la = norm(qa), lb = norm(qb), lc = norm(qc);
    if (la == 0.0f || lb == 0.0f || lc == 0.0f) return 0.0f;
    qa = divs(qa, la);
    qb = divs(qb, lb);
    qc = divs(qc, lc);
    float num = dot(qa, cross(sub(qb, qa), sub(qc, qa)));
    if (num == 0.0f) return 0.0f;
    float den = ((1.0f + dot(qa, qb)) + dot(qa, qc)) + dot(qb, qc);
    return std::atan2(num, den) * 2.0f;
}

double solid_angle_f64(const Tri& t, const Vec3f& qf) {
    auto D = [](const Vec3f& v) { return Vec3d{v.x, v.y, v.z}; };
    Vec3d q = D(qf);
    Vec3d qa = sub(D(t.a), q), qb = sub(D(t.b), q), qc = sub(D(t.c), q);
    double la = norm(qa), lb = norm(qb), lc = norm(qc);
    if (la == 0.0 || lb == 0.0 || lc == 0.0) return 0.0;
    qa = divs(qa, la);
    qb = divs(qb, lb);
    qc = divs(qc, lc);
    double num = dot(qa, cross(sub(qb, qa), sub(qc, qa)));
    if (num == 0.0) return 0.0;
    double den = ((1.0 + dot(qa, qb)) + dot(qa, qc)) + dot(qb, qc);
    return std::atan2(num, den) * 2.0;
}

float winding_number_exact(const std::vector<Tri>& tris, const Vec3f& q) {
    float wn = 0.0f;  // aabb_tree.rs:617-628 over all triangles
    for (auto& t : tris) wn += solid_angle(t, q);
    return wn / (4.0f * PI_F);
}

WindingNumbers WindingNumbers::from_mesh(const std::vector<Tri>& tris) {
    WindingNumbers w;
    w.objects = tris;
    w.object_boxes.resize(tris.size());
    for (size_t i = 0; i < tris.size(); ++i) w.object_boxes[i] = tri_bbox(tris[i]);
    if (!tris.empty()) {
        Builder b{w.objects, w.object_boxes, w.nodes};
        b.build(0, tris.size(), 1);
        w.data.resize(w.nodes.size());
        node_data(w, w.nodes.size() - 1, w.data);
    }
    return w;
}

float WindingNumbers::fast_wn(size_t idx, const Vec3f& p, float acc) const {
    const DipoleData& d = data[idx];  // aabb_tree.rs:662-691
    float dist = norm(sub(p, d.center));
    if (dist > d.radius * acc) {
        Vec3f o1;
        Mat3f o2;
        hessians(d.center, p, o1, o2);
        return dot(d.order1, o1) + mat_dot(d.order2, o2);
    }
    const BvhNode& n = nodes[idx];
    if (n.leaf) {
        float wn = 0.0f;
        for (size_t t = n.left; t < n.right; ++t) wn += solid_angle(objects[t], p);
        return wn / (4.0f * PI_F);
    }
    float l = fast_wn(n.left, p, acc);
    float r = fast_wn(n.right, p, acc);
    return l + r;
}

void WindingNumbers::count_visits(size_t idx, const Vec3f& p, float acc, uint64_t counts[3]) const {
    const DipoleData& d = data[idx];
    if (norm(sub(p, d.center)) > d.radius * acc) {
        ++counts[1];
        return;
    }
    const BvhNode& n = nodes[idx];
    if (n.leaf) {
        ++counts[2];
        return;
    }
    ++counts[0];
    count_visits(n.left, p, acc, counts);
    count_visits(n.right, p, acc, counts);
}

float WindingNumbers::approximate(const Vec3f& p, float acc) const {
    if (nodes.empty()) return 0.0f;
    return fast_wn(nodes.size() - 1, p, acc);
}

double WindingNumbers::fast_wn_f64(size_t idx, const Vec3f& p, float acc) const {
    const DipoleData& d = data[idx];
    float dist = norm(sub(p, d.center));
    if (dist > d.radius * acc) {
        Vec3d r{(double)d.center.x - p.x, (double)d.center.y - p.y, (double)d.center.z - p.z};
        double r2 = norm_sq(r), r1 = std::sqrt(r2), r3 = r2 * r1, r5 = r3 * r2;
        double k = 1.0 / (4.0 * M_PI);
        double w = (d.order1.x * r.x + d.order1.y * r.y + d.order1.z * r.z) * k / r3;
        double tr = 0.0, rr = 0.0;
        for (int c = 0; c < 3; ++c) {
            tr += d.order2.at(c, c);
            for (int rw = 0; rw < 3; ++rw) rr += (double)d.order2.at(rw, c) * r[rw] * r[c];
        }
        w += k * (tr / r3 - 3.0 * rr / r5);
        return w;
    }
    const BvhNode& n = nodes[idx];
    if (n.leaf) {
        double wn = 0.0;
        for (size_t t = n.left; t < n.right; ++t) wn += solid_angle_f64(objects[t], p);
        return wn / (4.0 * M_PI);
    }
    return fast_wn_f64(n.left, p, acc) + fast_wn_f64(n.right, p, acc);
}

double WindingNumbers::approximate_f64(const Vec3f& p, float acc) const {
    if (nodes.empty()) return 0.0;
    return fast_wn_f64(nodes.size() - 1, p, acc);
}

}  // namespace oracle
