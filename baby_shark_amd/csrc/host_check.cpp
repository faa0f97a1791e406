// The library's host C++ (the tree topology of topology.cpp and the STL
// reader/writer of stl.cpp) driven on the CPU alone, for the sanitizer build
// (`make -C baby_shark_amd/csrc host_check`: AddressSanitizer and
// UndefinedBehaviorSanitizer, no GPU code).  It runs every Topology operation
// the device paths call -- build from leaves, upper-level flood fill, CSG for
// the three operations, collection, visit order, sign lookup, the offset
// rebuild, tile densification and compaction -- on sphere shells and mixed
// tiles, plus STL round trips and error paths, and checks the invariants that
// do not need the device's leaf values.  Exit status 0 when every check holds.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <string>
#include <vector>

#include "stl.hpp"
#include "topology.hpp"

using namespace bs;

namespace {

int g_fail = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            ++g_fail;                                                   \
        }                                                               \
    } while (0)

using Origin = std::array<int32_t, 3>;

// Leaf origins of the shell |p - c| - r within 1.5 leaves (voxel units).
std::vector<Origin> shell(double cx, double cy, double cz, double r) {
    std::vector<Origin> out;
    const int lo = (int)std::floor((std::min({cx, cy, cz}) - r - 16) / 8.0), hi = (int)std::ceil((std::max({cx, cy, cz}) + r + 16) / 8.0);
    for (int i = lo; i <= hi; ++i)
        for (int j = lo; j <= hi; ++j)
            for (int k = lo; k <= hi; ++k) {
                const double x = i * 8 + 4 - cx, y = j * 8 + 4 - cy, z = k * 8 + 4 - cz;
                if (std::fabs(std::sqrt(x * x + y * y + z * z) - r) < 12.0) out.push_back({i * 8, j * 8, k * 8});
            }
    return out;
}

// Leaf flood-fill results the device would report: sign of the sphere
// function at the leaf's first and last voxel.
void set_leaf_signs(Topology& t, double cx, double cy, double cz, double r) {
    for (size_t i = 0; i < t.leaves.size(); ++i) {
        LeafRef& L = t.leaves[i];
        auto neg = [&](double x, double y, double z) {
            return std::sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz)) < r;
        };
        L.first_neg = neg(L.x, L.y, L.z);
        L.last_neg = neg(L.x + 7, L.y + 7, L.z + 7);
        L.empty = 0;
    }
}

std::unique_ptr<Topology> sphere_topo(double cx, double cy, double cz, double r, std::vector<TileRec> tiles = {}) {
    auto t = Topology::from_leaves(shell(cx, cy, cz, r), tiles);
    set_leaf_signs(*t, cx, cy, cz, r);
    t->flood_fill_upper();
    return t;
}

void check_topology() {
    // leaves spread over several root blocks, in arbitrary order, with tiles of both sizes
    std::vector<TileRec> tiles = {{-4096, 0, 0, 8, -3.0f}, {4096, 4096, 0, 128, 2.5f}, {0, -1024, 2048, 8, 1.0f}};  // away from every shell
    for (double r : {20.0, 75.0, 300.0}) {
        auto a = sphere_topo(0, 0, 0, r, tiles);
        const size_t na = a->leaves.size();
        CHECK(na > 0);
        std::vector<int64_t> units;
        std::vector<TileRec> vt;
        a->visit_units(units, vt);
        size_t leaf_units = 0;
        for (int64_t u : units) leaf_units += u >= 0;
        CHECK(leaf_units == na);
        CHECK(vt.size() == tiles.size());
        bool neg = false;
        CHECK(a->locate(0, 0, 0, &neg) < 0);  // the centre is inside a tile of the flood-filled tree
        for (int op = 0; op < 3; ++op) {
            auto x = sphere_topo(0, 0, 0, r, tiles);
            auto y = sphere_topo(r * 0.6, 5, -7, r * 0.8);
            x->csg(std::move(*y), (CsgOp)op);
            std::vector<int32_t> refs;
            std::vector<TileRec> out_tiles;
            x->collect(refs, out_tiles);
            for (int32_t ref : refs) CHECK(ref >= 0 && (size_t)ref < x->leaves.size());
            std::vector<int32_t> idx(x->leaves.size(), -1);
            for (size_t k = 0; k < refs.size(); ++k) idx[refs[k]] = (int32_t)k;
            auto c = x->compact(idx);
            std::vector<int32_t> refs2;
            std::vector<TileRec> tiles2;
            c->collect(refs2, tiles2);
            CHECK(refs2.size() == refs.size());
            for (size_t k = 0; k < refs2.size(); ++k) CHECK(refs2[k] == (int32_t)k);
        }
        // the offset rebuild: drop every third leaf; the output holds every
        // kept leaf (a sweep never empties one) and a grown shell
        std::vector<uint32_t> keep(na);
        std::set<Origin> out_set;
        for (size_t i = 0; i < na; ++i) {
            keep[i] = i % 3 != 0;
            if (keep[i]) out_set.insert({a->leaves[i].x, a->leaves[i].y, a->leaves[i].z});
        }
        for (const Origin& o : shell(0, 0, 0, r + 20)) out_set.insert(o);
        const std::vector<Origin> grown(out_set.begin(), out_set.end());
        auto b = a->after_offset(keep, grown, 2.0f);
        std::vector<int32_t> refs;
        std::vector<TileRec> bt;
        b->collect(refs, bt);
        CHECK(refs.size() == grown.size());
        // tiles to leaves
        auto d = a->clone();
        const size_t before = d->leaves.size();
        std::vector<TileRec> made = d->densify_tiles();
        size_t expect = 0;
        for (const TileRec& t : tiles) expect += t.size == 8 ? 1 : 4096;
        CHECK(made.size() == expect);
        CHECK(d->leaves.size() == before + made.size());
    }
}

void check_stl(const std::string& dir) {
    for (const char* name : {"box.stl", "box2.stl", "bunny.stl"}) {
        std::vector<float> xyz;
        const std::string path = dir + "/" + name;
        CHECK(stl_read(path.c_str(), xyz).empty());
        CHECK(!xyz.empty() && xyz.size() % 9 == 0);
        const std::string tmp = std::string("/tmp/bs_host_check_") + name;
        CHECK(stl_write(tmp.c_str(), xyz.data(), xyz.size() / 9).empty());
        std::vector<float> back;
        CHECK(stl_read(tmp.c_str(), back).empty());
        CHECK(back == xyz);
        std::remove(tmp.c_str());
    }
    std::vector<float> xyz;
    CHECK(!stl_read("/nonexistent/file.stl", xyz).empty());
    // a file cut inside its last face
    const std::string cut = "/tmp/bs_host_check_cut.stl";
    FILE* f = std::fopen((dir + "/box.stl").c_str(), "rb");
    FILE* g = std::fopen(cut.c_str(), "wb");
    CHECK(f && g);
    if (f && g) {
        std::vector<unsigned char> b(84 + 50 * 12);
        const size_t got = std::fread(b.data(), 1, b.size(), f);
        std::fwrite(b.data(), 1, got - 20, g);
    }
    if (f) std::fclose(f);
    if (g) std::fclose(g);
    xyz.clear();
    CHECK(!stl_read(cut.c_str(), xyz).empty());
    std::remove(cut.c_str());
}

}  // namespace

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : "tests/golden";
    check_topology();
    check_stl(dir);
    if (g_fail) {
        std::fprintf(stderr, "%d checks failed\n", g_fail);
        return 1;
    }
    std::printf("host checks passed\n");
    return 0;
}
