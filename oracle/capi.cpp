// Test infrastructure only: the CPU oracle.  Nothing outside tests/, the smoke
// check and bench.py's cpu_baseline leg may link or call this code.
//
// C entry points of the oracle (loaded by tests through ctypes).  Volumes are
// exported canonically: leaves sorted by origin (x, then y, then z), an active
// mask of 8 u64 words per leaf with bit (i % 64) of word (i / 64) standing for
// voxel offset i = (x&7)<<6 | (y&7)<<3 | (z&7), 512 f32 values per leaf, and
// active tiles as (origin, edge length, value).
#include <algorithm>
#include <array>
#include <thread>
#include <atomic>
#include <chrono>
#include <cstring>
#include <cstdlib>
#include <unordered_map>

#include "dual_contouring.hpp"
#include "marching_cubes.hpp"
#include "fast_sweep.hpp"
#include "mesh_to_volume.hpp"
#include "parallel.hpp"

using namespace oracle;

struct or_volume {
    std::unique_ptr<VolumeGrid> grid;
    float voxel_size;
    // canonical export cache
    std::vector<const LeafNode*> leaves;
    struct TileRec { Vec3i o; int size; float v; };
    std::vector<TileRec> tiles;
    void index() {
        leaves.clear();
        tiles.clear();
        grid->visit_leafs([&](Vec3i o, int s, float v) { tiles.push_back({o, s, v}); },
                          [&](const LeafNode& l) { leaves.push_back(&l); });
        std::sort(leaves.begin(), leaves.end(), [](const LeafNode* a, const LeafNode* b) { return a->origin < b->origin; });
        std::sort(tiles.begin(), tiles.end(), [](const TileRec& a, const TileRec& b) {
            if (a.o != b.o) return a.o < b.o;
            return a.size < b.size;
        });
    }
};

static std::vector<Tri> to_tris(const float* xyz, size_t n) {
    std::vector<Tri> t(n);
    for (size_t i = 0; i < n; ++i) {
        const float* p = xyz + 9 * i;
        t[i] = {{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}};
    }
    return t;
}

extern "C" {

// MeshToVolume::default().with_voxel_size(h).with_narrow_band_width(band).convert(mesh)
// Returns NULL where the reference returns None.  `seconds` (nullable) receives
// the wall time of convert() alone.
or_volume* or_mesh_to_volume(const float* tris, size_t n, float voxel_size, int64_t band, int threads, double* seconds) {
    MeshToVolume m;
    m.set_voxel_size(voxel_size);
    m.band_width = band;
    if (threads > 0) m.threads = threads;
    auto mesh = to_tris(tris, n);
    auto t0 = std::chrono::steady_clock::now();
    auto g = m.convert(mesh);
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (!g) return nullptr;
    auto* v = new or_volume{std::move(g), voxel_size, {}, {}};
    v->index();
    return v;
}

// Unsigned pass only (mesh_to_volume.rs:118-196), for stage-wise parity.
or_volume* or_unsigned_distance(const float* tris, size_t n, float voxel_size, int64_t band, int threads) {
    MeshToVolume m;
    m.set_voxel_size(voxel_size);
    m.band_width = band;
    if (threads > 0) m.threads = threads;
    auto mesh = to_tris(tris, n);
    for (auto& t : mesh) subdivide_triangle(t, voxel_size, m.subdivided);
    if (m.subdivided.empty()) return nullptr;
    auto* v = new or_volume{m.unsigned_distance_field(), voxel_size, {}, {}};
    v->index();
    return v;
}

void or_volume_free(or_volume* v) { delete v; }

size_t or_volume_leaf_count(const or_volume* v) { return v->leaves.size(); }
size_t or_volume_tile_count(const or_volume* v) { return v->tiles.size(); }

void or_volume_export(const or_volume* v, int32_t* origins, uint64_t* masks, float* values) {
    for (size_t i = 0; i < v->leaves.size(); ++i) {
        const LeafNode* l = v->leaves[i];
        if (origins) {
            origins[3 * i + 0] = (int32_t)l->origin.x;
            origins[3 * i + 1] = (int32_t)l->origin.y;
            origins[3 * i + 2] = (int32_t)l->origin.z;
        }
        if (masks)
            for (int w = 0; w < 8; ++w) {
                uint64_t m = 0;
                for (int b = 0; b < 64; ++b)
                    if (l->value_mask.at(w * 64 + b)) m |= uint64_t(1) << b;
                masks[8 * i + w] = m;
            }
        if (values) std::memcpy(values + 512 * i, l->values, 512 * sizeof(float));
    }
}

void or_volume_export_tiles(const or_volume* v, int32_t* origins, int32_t* sizes, float* values) {
    for (size_t i = 0; i < v->tiles.size(); ++i) {
        const auto& t = v->tiles[i];
        origins[3 * i + 0] = (int32_t)t.o.x;
        origins[3 * i + 1] = (int32_t)t.o.y;
        origins[3 * i + 2] = (int32_t)t.o.z;
        sizes[i] = t.size;
        values[i] = t.v;
    }
}

// Sub-triangles of mesh_to_volume.rs:75-116; returns the count, fills `out`
// (n_sub*9 floats) when non-null.
size_t or_subdivide(const float* tris, size_t n, float voxel_size, float* out) {
    std::vector<Tri> sub;
    for (size_t i = 0; i < n; ++i) {
        const float* p = tris + 9 * i;
        subdivide_triangle({{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}}, voxel_size, sub);
    }
    if (out) std::memcpy(out, sub.data(), sub.size() * sizeof(Tri));
    return sub.size();
}

// Σ over sub-triangles of the dilated-box voxel count (the (voxel, sub-triangle)
// pair count of mesh_to_volume.rs:124-164).
uint64_t or_pair_count(const float* tris, size_t n, float voxel_size, int64_t band) {
    std::vector<Tri> sub;
    for (size_t i = 0; i < n; ++i) {
        const float* p = tris + 9 * i;
        subdivide_triangle({{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}}, voxel_size, sub);
    }
    float inv = 1.0f / voxel_size;
    uint64_t total = 0;
    for (auto& t : sub) total += (uint64_t)dilated_box(t, inv, band).volume();
    return total;
}

// Winding numbers at query points: reference tree, accuracy scale `acc`
// (f32 and FP64 re-evaluation), plus the brute-force f32 sum when `exact` is non-null.
void or_winding_numbers(const float* tris, size_t n, const float* pts, size_t m, float acc, float* out_f32,
                        double* out_f64, float* exact, int threads) {
    auto mesh = to_tris(tris, n);
    WindingNumbers w = WindingNumbers::from_mesh(mesh);
    par_for(m, threads > 0 ? threads : default_threads(), [&](size_t i) {
        Vec3f p{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        if (out_f32) out_f32[i] = w.approximate(p, acc);
        if (out_f64) out_f64[i] = w.approximate_f64(p, acc);
        if (exact) exact[i] = winding_number_exact(mesh, p);
    });
}

// Node visits of the reference tree's approximate() summed over m points:
// counts = (internal nodes opened, far evaluations, leaves summed).
void or_winding_visit_counts(const float* tris, size_t n, const float* pts, size_t m, float acc, uint64_t* counts) {
    auto mesh = to_tris(tris, n);
    WindingNumbers w = WindingNumbers::from_mesh(mesh);
    counts[0] = counts[1] = counts[2] = 0;
    for (size_t i = 0; i < m; ++i) {
        if (w.nodes.empty()) break;
        w.count_visits(w.nodes.size() - 1, {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, acc, counts);
    }
}

// The reference-tree BVH of a soup (AABBTree::from_mesh + min 3 objects per
// leaf + top_down::<Area>, aabb_tree.rs:636-644), nodes in push order (root
// last).  Returns the node count; when the arrays are non-null (sized by a
// first call with NULLs) fills leaf flags, left/right (object range of a leaf,
// child nodes of a branch), node boxes (6 floats), the reordered objects
// (9 floats each) and the per-node dipole centre and radius.
size_t or_bvh_export(const float* tris, size_t n, uint8_t* leaf, uint64_t* left, uint64_t* right, float* bbox,
                     float* objects, float* centre, float* radius) {
    auto mesh = to_tris(tris, n);
    WindingNumbers w = WindingNumbers::from_mesh(mesh);
    if (!leaf) return w.nodes.size();
    for (size_t i = 0; i < w.nodes.size(); ++i) {
        const BvhNode& nd = w.nodes[i];
        leaf[i] = nd.leaf ? 1 : 0;
        left[i] = nd.left;
        right[i] = nd.right;
        const float b[6] = {nd.bbox.min.x, nd.bbox.min.y, nd.bbox.min.z, nd.bbox.max.x, nd.bbox.max.y, nd.bbox.max.z};
        std::copy(b, b + 6, bbox + 6 * i);
        if (centre) {
            centre[3 * i] = w.data[i].center.x;
            centre[3 * i + 1] = w.data[i].center.y;
            centre[3 * i + 2] = w.data[i].center.z;
        }
        if (radius) radius[i] = w.data[i].radius;
    }
    for (size_t t = 0; t < w.objects.size(); ++t) {
        const Tri& o = w.objects[t];
        const float v[9] = {o.a.x, o.a.y, o.a.z, o.b.x, o.b.y, o.b.z, o.c.x, o.c.y, o.c.z};
        std::copy(v, v + 9, objects + 9 * t);
    }
    return w.nodes.size();
}

// find_feature_point (dual_contouring.rs:429-458) of n intersection points
// (points n*3, unit normals n*3, voxel units).
void or_dc_feature_point(const float* points, const float* normals, size_t n, float* out) {
    std::vector<IntPoint> pts(n);
    for (size_t i = 0; i < n; ++i)
        pts[i] = {{points[3 * i], points[3 * i + 1], points[3 * i + 2]},
                  {normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]}};
    const Vec3f c = DualContouring::feature_point(pts);
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}

float or_solid_angle(const float* tri, const float* q) {
    return solid_angle({{tri[0], tri[1], tri[2]}, {tri[3], tri[4], tri[5]}, {tri[6], tri[7], tri[8]}}, {q[0], q[1], q[2]});
}

void or_closest_point(const float* tri, const float* q, float* out) {
    Vec3f c = closest_point({{tri[0], tri[1], tri[2]}, {tri[3], tri[4], tri[5]}, {tri[6], tri[7], tri[8]}}, {q[0], q[1], q[2]});
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}

// FP64 unsigned distance of every active voxel of `v` (canonical leaf order,
// 512 slots per leaf; inactive slots get +inf): min over exactly the
// (voxel, sub-triangle) pairs of the reference pair set, evaluated in double.
void or_unsigned_distance_f64(const or_volume* v, const float* tris, size_t n, int64_t band, double* out, int threads) {
    std::vector<Tri> sub;
    float h = v->voxel_size;
    for (size_t i = 0; i < n; ++i) {
        const float* p = tris + 9 * i;
        subdivide_triangle({{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}}, h, sub);
    }
    struct KeyHash {
        size_t operator()(const Vec3i& k) const {
            return (size_t)(k.x * 73856093) ^ (size_t)(k.y * 19349663) ^ (size_t)(k.z * 83492791);
        }
    };
    std::unordered_map<Vec3i, size_t, KeyHash> leaf_index;
    for (size_t i = 0; i < v->leaves.size(); ++i) leaf_index[v->leaves[i]->origin] = i;
    size_t slots = v->leaves.size() * 512;
    std::vector<std::atomic<uint64_t>> best(slots);
    double inf = INFINITY;
    uint64_t inf_bits;
    std::memcpy(&inf_bits, &inf, 8);
    for (auto& b : best) b.store(inf_bits, std::memory_order_relaxed);
    float inv = 1.0f / h;
    par_for(sub.size(), threads > 0 ? threads : default_threads(), [&](size_t k) {
        const Tri& t = sub[k];
        VoxelBox b = dilated_box(t, inv, band);
        for (int64_t x = b.min.x; x <= b.max.x; ++x)
            for (int64_t y = b.min.y; y <= b.max.y; ++y)
                for (int64_t z = b.min.z; z <= b.max.z; ++z) {
                    auto it = leaf_index.find({x & ~int64_t(7), y & ~int64_t(7), z & ~int64_t(7)});
                    if (it == leaf_index.end()) continue;
                    size_t slot = it->second * 512 + (size_t)(((x & 7) << 6) | ((y & 7) << 3) | (z & 7));
                    double d = point_tri_distance_f64(t, {(float)x * h, (float)y * h, (float)z * h});
                    uint64_t bits;
                    std::memcpy(&bits, &d, 8);
                    uint64_t cur = best[slot].load(std::memory_order_relaxed);
                    while (bits < cur && !best[slot].compare_exchange_weak(cur, bits)) {
                    }
                }
    }, 16);
    for (size_t s = 0; s < slots; ++s) {
        uint64_t bits = best[s].load();
        std::memcpy(&out[s], &bits, 8);
    }
}

// Active voxel set alone: the union over sub-triangles of the dilated boxes
// (mesh_to_volume.rs:124-168), as canonical (origin, mask) pairs.  Returns
// the leaf count; fills the arrays when non-null (call twice).
struct ActiveSet {
    std::vector<Vec3i> origins;
    std::vector<std::array<uint64_t, 8>> masks;
};
static ActiveSet g_active;

size_t or_active_set(const float* tris, size_t n, float voxel_size, int64_t band, int threads) {
    std::vector<Tri> sub;
    for (size_t i = 0; i < n; ++i) {
        const float* p = tris + 9 * i;
        subdivide_triangle({{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}}, voxel_size, sub);
    }
    float inv = 1.0f / voxel_size;
    struct KeyHash {
        size_t operator()(const Vec3i& k) const {
            return (size_t)(k.x * 73856093) ^ (size_t)(k.y * 19349663) ^ (size_t)(k.z * 83492791);
        }
    };
    int T = threads > 0 ? threads : default_threads();
    std::vector<std::unordered_map<Vec3i, std::array<uint64_t, 8>, KeyHash>> part(T);
    std::atomic<size_t> next{0};
    auto work = [&](int tid) {
        auto& mp = part[tid];
        for (;;) {
            size_t s0 = next.fetch_add(4096);
            if (s0 >= sub.size()) return;
            size_t s1 = std::min(sub.size(), s0 + 4096);
            for (size_t k = s0; k < s1; ++k) {
                VoxelBox b = dilated_box(sub[k], inv, band);
                for (int64_t lx = b.min.x >> 3; lx <= b.max.x >> 3; ++lx)
                    for (int64_t ly = b.min.y >> 3; ly <= b.max.y >> 3; ++ly)
                        for (int64_t lz = b.min.z >> 3; lz <= b.max.z >> 3; ++lz) {
                            auto& m = mp[{lx * 8, ly * 8, lz * 8}];
                            int64_t x0 = std::max(b.min.x, lx * 8), x1 = std::min(b.max.x, lx * 8 + 7);
                            int64_t y0 = std::max(b.min.y, ly * 8), y1 = std::min(b.max.y, ly * 8 + 7);
                            int64_t z0 = std::max(b.min.z, lz * 8) - lz * 8, z1 = std::min(b.max.z, lz * 8 + 7) - lz * 8;
                            uint64_t zbits = ((uint64_t(1) << (z1 + 1)) - 1) & ~((uint64_t(1) << z0) - 1);
                            for (int64_t x = x0; x <= x1; ++x)
                                for (int64_t y = y0; y <= y1; ++y) {
                                    int off = (int)(((x & 7) << 6) | ((y & 7) << 3));
                                    m[off >> 6] |= zbits << (off & 63);
                                }
                        }
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    for (int t = 1; t < T; ++t)
        for (auto& kv : part[t]) {
            auto& m = part[0][kv.first];
            for (int w = 0; w < 8; ++w) m[w] |= kv.second[w];
        }
    g_active.origins.clear();
    g_active.masks.clear();
    std::vector<std::pair<Vec3i, std::array<uint64_t, 8>>> all(part[0].begin(), part[0].end());
    std::sort(all.begin(), all.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (auto& kv : all) {
        g_active.origins.push_back(kv.first);
        g_active.masks.push_back(kv.second);
    }
    return g_active.origins.size();
}

void or_active_set_export(int32_t* origins, uint64_t* masks) {
    for (size_t i = 0; i < g_active.origins.size(); ++i) {
        for (int k = 0; k < 3; ++k) origins[3 * i + k] = (int32_t)g_active.origins[i][k];
        for (int w = 0; w < 8; ++w) masks[8 * i + w] = g_active.masks[i][w];
    }
}

// Unsigned distance at sample voxels (m x 3 int32 indices): the reference's
// f32 value (min over the sub-triangles whose dilated box holds the voxel,
// mesh_to_volume.rs:147-195) and the same min evaluated in FP64.  +inf when
// no box holds the voxel.
void or_sample_unsigned(const float* tris, size_t n, float voxel_size, int64_t band, const int32_t* vox, size_t m,
                        float* out32, double* out64, int threads) {
    std::vector<Tri> sub;
    for (size_t i = 0; i < n; ++i) {
        const float* p = tris + 9 * i;
        subdivide_triangle({{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}}, voxel_size, sub);
    }
    struct KeyHash {
        size_t operator()(const Vec3i& k) const {
            return (size_t)(k.x * 73856093) ^ (size_t)(k.y * 19349663) ^ (size_t)(k.z * 83492791);
        }
    };
    std::unordered_map<Vec3i, std::vector<size_t>, KeyHash> by_leaf;
    for (size_t i = 0; i < m; ++i)
        by_leaf[{(int64_t)vox[3 * i] & ~int64_t(7), (int64_t)vox[3 * i + 1] & ~int64_t(7), (int64_t)vox[3 * i + 2] & ~int64_t(7)}]
            .push_back(i);
    std::vector<std::atomic<uint32_t>> b32(m);
    std::vector<std::atomic<uint64_t>> b64(m);
    float finf = INFINITY;
    double dinf = INFINITY;
    uint32_t fi;
    uint64_t di;
    std::memcpy(&fi, &finf, 4);
    std::memcpy(&di, &dinf, 8);
    for (size_t i = 0; i < m; ++i) {
        b32[i].store(fi);
        b64[i].store(di);
    }
    float inv = 1.0f / voxel_size;
    float h = voxel_size;
    par_for(sub.size(), threads > 0 ? threads : default_threads(), [&](size_t k) {
        const Tri& t = sub[k];
        VoxelBox b = dilated_box(t, inv, band);
        for (int64_t lx = b.min.x >> 3; lx <= b.max.x >> 3; ++lx)
            for (int64_t ly = b.min.y >> 3; ly <= b.max.y >> 3; ++ly)
                for (int64_t lz = b.min.z >> 3; lz <= b.max.z >> 3; ++lz) {
                    auto it = by_leaf.find({lx * 8, ly * 8, lz * 8});
                    if (it == by_leaf.end()) continue;
                    for (size_t s : it->second) {
                        int64_t x = vox[3 * s], y = vox[3 * s + 1], z = vox[3 * s + 2];
                        if (x < b.min.x || x > b.max.x || y < b.min.y || y > b.max.y || z < b.min.z || z > b.max.z) continue;
                        Vec3f p{(float)x * h, (float)y * h, (float)z * h};
                        float d = point_tri_distance(t, p);
                        double dd = point_tri_distance_f64(t, p);
                        uint32_t bits;
                        std::memcpy(&bits, &d, 4);
                        uint32_t cur = b32[s].load();
                        while (bits < cur && !b32[s].compare_exchange_weak(cur, bits)) {
                        }
                        uint64_t bb;
                        std::memcpy(&bb, &dd, 8);
                        uint64_t c64 = b64[s].load();
                        while (bb < c64 && !b64[s].compare_exchange_weak(c64, bb)) {
                        }
                    }
                }
    }, 256);
    for (size_t i = 0; i < m; ++i) {
        uint32_t a = b32[i].load();
        uint64_t b = b64[i].load();
        std::memcpy(&out32[i], &a, 4);
        std::memcpy(&out64[i], &b, 8);
    }
}

// ---------------------------------------------------------------- volume ops
extern "C++" {
// Volume::from_fn (volume/mod.rs:40-72): every grid point of the index box
// floor(min / h) ..= ceil(max / h), p = idx as f32 * h, kept when
// |f(p)| <= (width + 1) as f32 * h.
template <class F>
static or_volume* from_fn(float h, const Vec3f& mn, const Vec3f& mx, int width, F f) {
    auto g = std::make_unique<VolumeGrid>();
    const float nbw = (float)(width + 1) * h;
    int64_t lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = f2i(std::floor(mn[k] / h));
        hi[k] = f2i(std::ceil(mx[k] / h));
    }
    for (int64_t x = lo[0]; x <= hi[0]; ++x)
        for (int64_t y = lo[1]; y <= hi[1]; ++y)
            for (int64_t z = lo[2]; z <= hi[2]; ++z) {
                Vec3f p{(float)x * h, (float)y * h, (float)z * h};
                float v = f(p);
                if (std::fabs(v) > nbw) continue;
                g->insert({x, y, z}, v);
            }
    auto* r = new or_volume{std::move(g), h, {}, {}};
    r->index();
    return r;
}

// box3.rs:115-122
static bool box_contains(const Vec3f& mn, const Vec3f& mx, const Vec3f& p) {
    return p.x >= mn.x && p.x <= mx.x && p.y >= mn.y && p.y <= mx.y && p.z >= mn.z && p.z <= mx.z;
}

// box3.rs:96-112
static float box_squared_distance(const Vec3f& mn, const Vec3f& mx, const Vec3f& p) {
    float sq = 0.0f;
    for (int i = 0; i < 3; ++i) {
        const float v = p[i];
        if (v < mn[i]) sq += (mn[i] - v) * (mn[i] - v);
        if (v > mx[i]) sq += (v - mx[i]) * (v - mx[i]);
    }
    return sq;
}
}  // extern "C++"

// VolumeBuilder::sphere (builder.rs:21-30)
or_volume* or_volume_sphere(float radius, float cx, float cy, float cz, float h) {
    const float band_offset = 1.0f * h;  // band_width = 1, `band_width as f32 * voxel_size`
    const float offset = radius + band_offset;
    const Vec3f o{cx, cy, cz};
    const Vec3f mn{cx + (-offset), cy + (-offset), cz + (-offset)};
    const Vec3f mx{cx + offset, cy + offset, cz + offset};
    return from_fn(h, mn, mx, 1, [&](const Vec3f& p) { return norm(sub(p, o)) - radius; });
}

// VolumeBuilder::cuboid (builder.rs:32-51): inside, minus the distance to the
// nearest face (f32::min chain); outside, the distance to the box.
or_volume* or_volume_cuboid(const float* mn_, const float* mx_, float h) {
    const Vec3f mn{mn_[0], mn_[1], mn_[2]}, mx{mx_[0], mx_[1], mx_[2]};
    const float offset = 1.0f * h;
    const Vec3f gmn{mn.x + (-offset), mn.y + (-offset), mn.z + (-offset)};
    const Vec3f gmx{mx.x + offset, mx.y + offset, mx.z + offset};
    return from_fn(h, gmn, gmx, 1, [&](const Vec3f& p) {
        if (box_contains(mn, mx, p))
            return -std::fmin(std::fmin(std::fmin(std::fmin(std::fmin(p.x - mn.x, mx.x - p.x), p.y - mn.y), mx.y - p.y),
                                        p.z - mn.z),
                              mx.z - p.z);
        return std::sqrt(box_squared_distance(mn, mx, p));
    });
}

// VolumeBuilder::iwp (builder.rs:53-76): the IWP triply periodic surface of
// period cell_size inside [min, max], the distance to the box grown by one
// voxel outside it; band width 2.  cos is libm's cosf, as f32::cos on Linux.
or_volume* or_volume_iwp(const float* mn_, const float* mx_, float cell_size, float h) {
    const Vec3f mn{mn_[0], mn_[1], mn_[2]}, mx{mx_[0], mx_[1], mx_[2]};
    const float inv = 1.0f / cell_size;
    const Vec3f imn{mn.x + (-h), mn.y + (-h), mn.z + (-h)};
    const Vec3f imx{mx.x + h, mx.y + h, mx.z + h};
    return from_fn(h, imn, imx, 2, [&](const Vec3f& p) {
        const float x = p.x * inv, y = p.y * inv, z = p.z * inv;
        const float cx = std::cos(x), cy = std::cos(y), cz = std::cos(z);
        const float v = -(cx + cy + cz - 0.51f * (cx * cy + cy * cz + cz * cx) - 1.0f);
        if (!box_contains(mn, mx, p)) return std::sqrt(box_squared_distance(imn, imx, p));
        return v * cell_size;
    });
}

// An empty VolumeGrid receiving `insert(idx, value)` for each voxel in order
// (TreeNode::insert: root_node/tree_node.rs, internal_node/tree_node.rs:54-79,
// leaf_node/tree_node.rs:46-50); later inserts overwrite earlier ones.
or_volume* or_volume_from_voxels(const int32_t* idx, const float* values, size_t n, float h) {
    auto g = std::make_unique<VolumeGrid>();
    for (size_t i = 0; i < n; ++i) g->insert({idx[3 * i], idx[3 * i + 1], idx[3 * i + 2]}, values[i]);
    auto* r = new or_volume{std::move(g), h, {}, {}};
    r->index();
    return r;
}

// A grid holding n given leaves as they are: origin (x, y, z per leaf), the
// 512-bit active mask (bit i % 64 of word i / 64 is voxel offset i) and all
// 512 values, inactive ones included; the leaves are placed with
// insert_leaf_at (internal_node/tree_node.rs:280-298).  Lets the tests run the
// oracle's CSG and offset on operands another implementation produced.
or_volume* or_volume_from_leaves(const int32_t* origins, const uint64_t* masks, const float* values, size_t n,
                                 float h) {
    auto g = std::make_unique<VolumeGrid>();
    for (size_t l = 0; l < n; ++l) {
        auto leaf = std::make_unique<LeafNode>(Vec3i{origins[3 * l], origins[3 * l + 1], origins[3 * l + 2]});
        for (int o = 0; o < LeafNode::SIZE; ++o) {
            leaf->values[o] = values[l * LeafNode::SIZE + o];
            if ((masks[l * 8 + (o >> 6)] >> (o & 63)) & 1ull) leaf->value_mask.on(o);
        }
        g->insert_leaf_at(std::move(leaf));
    }
    auto* r = new or_volume{std::move(g), h, {}, {}};
    r->index();
    return r;
}

// DualContouringMesher::default().with_voxel_size(h).mesh(volume)
// (dual_contouring.rs:23-83).  *out receives n_tris*9 floats (free with
// or_free).  Returns 0, 1 where the reference reaches todo!() (a tile) or 2
// where it reaches unreachable!() (a crossing voxel without neighbours).
int or_dual_contour(const or_volume* a, float h, float** out, size_t* n_tris) {
    *out = nullptr;
    *n_tris = 0;
    try {
        DualContouring dc(*a->grid);
        std::vector<Vec3f> v = dc.mesh(h);
        *n_tris = v.size() / 3;
        *out = (float*)std::malloc(sizeof(float) * 3 * (v.size() ? v.size() : 1));
        for (size_t i = 0; i < v.size(); ++i) {
            (*out)[3 * i] = v[i].x;
            (*out)[3 * i + 1] = v[i].y;
            (*out)[3 * i + 2] = v[i].z;
        }
    } catch (const DcError& e) {
        return e.code;
    }
    return 0;
}

// MarchingCubesMesher::default().with_voxel_size(h).mesh(volume)
// (marching_cubes.rs:43-63).  *out receives n_tris*9 floats (free with or_free).
int or_marching_cubes(const or_volume* a, float h, float** out, size_t* n_tris) {
    std::vector<Vec3f> v = mc::mesh(*a->grid, h);
    *n_tris = v.size() / 3;
    *out = (float*)std::malloc(sizeof(float) * 3 * (v.size() ? v.size() : 1));
    for (size_t i = 0; i < v.size(); ++i) {
        (*out)[3 * i] = v[i].x;
        (*out)[3 * i + 1] = v[i].y;
        (*out)[3 * i + 2] = v[i].z;
    }
    return 0;
}

// The triangles of one cube as crossing-edge triples (12 = centre) in the
// emitted vertex order, for the per-cube checks: c = 8 clamped corner values;
// returns the triangle count.
int or_mc_cube_tiling(const float* c, int* out_edges) {
    std::vector<mc::Tri> t;
    int id = 0;
    for (int i = 0; i < 8; ++i)
        if (c[i] < 0.0f) id |= 1 << i;
    mc::handle_cube(c, id, t);
    for (size_t i = 0; i < t.size(); ++i)
        for (int k = 0; k < 3; ++k) out_edges[3 * i + k] = t[i][k];
    return (int)t.size();
}

void or_free(void* p) { std::free(p); }

// fast_sweep.rs:509-537 (exposed for its closed-form checks)
float or_fs_compute_distance(float a1, float a2, float a3, float h) { return fs_compute_distance(a1, a2, a3, h); }

// Tile densification (an extension with no reference counterpart, for dual
// contouring volumes that hold active tiles): every active tile becomes the
// leaves it covers, all 512 entries active and equal to the tile value,
// inserted with insert_leaf_at (internal_node/tree_node.rs:280-298).
void or_volume_densify_tiles(or_volume* a) {
    struct T { Vec3i o; int size; float v; };
    std::vector<T> tiles;
    a->grid->visit_leafs([&](Vec3i o, int s, float v) { tiles.push_back({o, s, v}); }, [&](const LeafNode&) {});
    for (const T& t : tiles)
        for (int64_t x = 0; x < t.size; x += 8)
            for (int64_t y = 0; y < t.size; y += 8)
                for (int64_t z = 0; z < t.size; z += 8) {
                    auto leaf = std::make_unique<LeafNode>(Vec3i{t.o.x + x, t.o.y + y, t.o.z + z});
                    leaf->fill(t.v);
                    a->grid->insert_leaf_at(std::move(leaf));
                }
    a->index();
}

// Volume::offset (volume/mod.rs:95-108).  Returns 0, or 1 where the reference
// panics (an active tile inside the band reaches the leaf queue).
int or_volume_offset(or_volume* a, float distance) {
    try {
        volume_offset(*a->grid, a->voxel_size, distance);
    } catch (const FastSweepError&) {
        a->index();
        return 1;
    }
    a->index();
    return 0;
}

// Volume::union / intersect / subtract (volume/mod.rs:74-93): both operands
// flood filled, then the tree CSG.  `b` is consumed (freed) as in Rust.
// op: 0 union, 1 subtract, 2 intersect.
void or_volume_csg(or_volume* a, or_volume* b, int op) {
    a->grid->flood_fill();
    b->grid->flood_fill();
    if (op == 0) a->grid->csg_union(std::move(b->grid));
    else if (op == 1) a->grid->csg_subtract(std::move(b->grid));
    else a->grid->csg_intersect(std::move(b->grid));
    delete b;
    a->index();
}

// FloodFill::flood_fill of the whole grid (root_node/flood_fill.rs:8-42)
void or_volume_flood_fill(or_volume* a) {
    a->grid->flood_fill();
    a->index();
}

or_volume* or_volume_clone(const or_volume* a) {
    auto* r = new or_volume{a->grid->clone(), a->voxel_size, {}, {}};
    r->index();
    return r;
}

// Sign of the flood-filled value at voxel indices (root_node/flood_fill.rs:56-63):
// 1 negative, 0 positive.
void or_volume_signs_at(const or_volume* a, const int32_t* idx, size_t m, uint8_t* out) {
    for (size_t i = 0; i < m; ++i)
        out[i] = a->grid->sign_at({idx[3 * i], idx[3 * i + 1], idx[3 * i + 2]}) == Sign::Negative;
}

// ---------------------------------------------------------------- reference unit tests
// Restatements of the trees used by the reference's own unit tests, so that
// tests can compare the oracle with the answers those tests assert.

// leaf_node/csg.rs:53-132: static_vdb!(f32, 1) leaves with full masks.
void or_reftest_leaf_csg(int op, const float* a, const float* b, float* out, int* mask_full) {
    using L = Leaf<1>;
    auto n1 = std::make_unique<L>(Vec3i{0, 0, 0});
    auto n2 = std::make_unique<L>(Vec3i{0, 0, 0});
    for (int i = 0; i < 8; ++i) {
        n1->values[i] = a[i];
        n2->values[i] = b[i];
    }
    n1->value_mask.on_all();
    n2->value_mask.on_all();
    if (op == 0) n1->csg_union(std::move(n2));
    else if (op == 1) n1->csg_subtract(std::move(n2));
    else n1->csg_intersect(std::move(n2));
    for (int i = 0; i < 8; ++i) out[i] = n1->values[i];
    *mask_full = n1->value_mask.is_full();
}

// leaf_node/flood_fill.rs:82-113 on static_vdb!(f32, 2): returns the 64 signs.
void or_reftest_leaf_flood_fill(int which, uint8_t* neg_out) {
    using L = Leaf<2>;
    L n(Vec3i{0, 0, 0});
    if (which == 0) n.insert({2, 2, 2}, 1.0f);
    else if (which == 1) n.insert({2, 2, 2}, -1.0f);
    else
        for (int y = 0; y < 4; ++y)
            for (int z = 0; z < 4; ++z) {
                n.insert({1, y, z}, -1.0f);
                n.insert({2, y, z}, 1.0f);
            }
    n.flood_fill();
    for (int i = 0; i < 64; ++i) neg_out[i] = std::signbit(n.values[i]);
}

// internal_node/flood_fill.rs:128-193 on static_vdb!(f32, 2, 1): signs over the
// node's 8^3 voxels, x-major.
void or_reftest_internal_flood_fill(int which, uint8_t* neg_out) {
    using I = Internal<Leaf<1>, 2>;
    auto n = std::make_unique<I>(Vec3i{0, 0, 0});
    const int R = 8;
    if (which == 0) n->insert({3, 2, 1}, 1.0f);
    else if (which == 1) n->insert({3, 3, 3}, -1.0f);
    else if (which == 2) {
        for (int y = 0; y < R; ++y)
            for (int z = 0; z < R; ++z) {
                n->insert({4, y, z}, -1.0f);
                n->insert({5, y, z}, 1.0f);
            }
    } else {
        for (int x = 2; x < 2 + 2; ++x)
            for (int y = 0; y < R; ++y)
                for (int z = 0; z < R; ++z) n->insert({x, y, z}, -1.0f);
        for (int x = 4; x < 4 + 2; ++x)
            for (int y = 0; y < R; ++y)
                for (int z = 0; z < R; ++z) n->insert({x, y, z}, 1.0f);
    }
    n->flood_fill();
    int k = 0;
    for (int x = 0; x < R; ++x)
        for (int y = 0; y < R; ++y)
            for (int z = 0; z < R; ++z) neg_out[k++] = n->sign_at({x, y, z}) == Sign::Negative;
}

// root_node/flood_fill.rs:74-95 on dynamic_vdb!(f32, 1): signs over the
// region (0,4,4)..(2,6,12), x-major.
void or_reftest_root_flood_fill(uint8_t* neg_out) {
    Root<Leaf<1>> t;
    for (int x = 0; x < 2; ++x)
        for (int y = 4; y < 6; ++y)
            for (int z = 4; z < 6; ++z) t.insert({x, y, z}, -1.0f);
    for (int x = 0; x < 2; ++x)
        for (int y = 4; y < 6; ++y)
            for (int z = 10; z < 12; ++z) t.insert({x, y, z}, -1.0f);
    t.flood_fill();
    int k = 0;
    for (int x = 0; x < 2; ++x)
        for (int y = 4; y < 6; ++y)
            for (int z = 4; z < 12; ++z) neg_out[k++] = t.sign_at({x, y, z}) == Sign::Negative;
}

}  // extern "C"
