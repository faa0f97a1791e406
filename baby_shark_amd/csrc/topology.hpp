// Host-side topology of a volume: the upper levels of the reference's grid
// `dynamic_vdb!(f32, par 5, 4, 3)` (src/voxel/volume/mod.rs:8) -- an ordered
// root map of 4096^3 blocks, 32^3-slot and 16^3-slot internal nodes -- kept as
// flat slot arrays (kind, tile value, child index).  Leaves (8^3 values and a
// mask) live in device memory; a slot of a 16^3 node refers to a leaf by index.
//
// The slot rules of flood fill and CSG follow internal_node/{flood_fill,csg}.rs
// and root_node/{flood_fill,csg}.rs; each is cited where implemented.
#pragma once
#include <array>
#include <cstdlib>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <utility>
#include <vector>

namespace bs {

enum : uint8_t { SLOT_EMPTY = 0, SLOT_TILE = 1, SLOT_BRANCH = 2 };
enum : uint8_t { LEAF_COPY_A = 0, LEAF_COPY_B = 1, LEAF_UNION = 2, LEAF_SUBTRACT = 3, LEAF_INTERSECT = 4 };
enum CsgOp : int { CSG_UNION = 0, CSG_SUBTRACT = 1, CSG_INTERSECT = 2 };

constexpr float F32_MAX_V = 3.40282347e+38f;

// An active tile of an internal node: origin, edge length (8 or 128), value.
struct TileRec {
    int32_t x, y, z, size;
    float value;
};

// A leaf of the volume being built: where its 512 values come from.  Plain
// data (no default member values), so a large table is allocated without a
// serial initialising pass; leaf_ref gives a fresh record.
struct LeafRef {
    int32_t x, y, z;  // origin
    int32_t a, b;     // source leaf in operand A / operand B (-1: none)
    uint8_t op;       // LEAF_*
    uint8_t neg;      // negate every value (flip_signs)
    uint8_t first_neg, last_neg, empty;  // after the leaf flood fill
};
inline LeafRef leaf_ref(int32_t x, int32_t y, int32_t z, int32_t a) {
    return LeafRef{x, y, z, a, -1, LEAF_COPY_A, 0, 0, 0, 0};
}

// Allocator whose value-less construct leaves a trivially constructible
// element uninitialised: resizing a table of n records touches no memory, and
// the threads that fill it are the first to touch its pages.
template <class T>
struct DefaultInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = DefaultInitAlloc<U>;
    };
    DefaultInitAlloc() = default;
    template <class U>
    DefaultInitAlloc(const DefaultInitAlloc<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... args) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(args)...);
    }
};
using LeafVec = std::vector<LeafRef, DefaultInitAlloc<LeafRef>>;

// Leaf references in segments, index i running over them in order: a CSG
// operand's references join the result by moving its segments in, without
// copying them.
struct LeafStore {
    std::vector<LeafVec> seg;
    size_t size() const {
        size_t n = 0;
        for (const auto& s : seg) n += s.size();
        return n;
    }
    LeafRef& operator[](size_t i) {
        for (auto& s : seg) {
            if (i < s.size()) return s[i];
            i -= s.size();
        }
        std::abort();  // an index past the end is a bug of the caller
    }
    const LeafRef& operator[](size_t i) const { return const_cast<LeafStore&>(*this)[i]; }
    // one segment of n uninitialised references, every one of which the
    // caller writes (the store must be empty)
    void resize(size_t n) {
        seg.clear();
        seg.emplace_back(n);
    }
    void push_back(const LeafRef& r) {
        if (seg.empty()) seg.emplace_back();
        seg.back().push_back(r);
    }
};

// Runs f(0) .. f(n-1) on the library's persistent host thread pool (the
// caller works too); returns when every call has finished.  f must not throw.
void host_parallel_for(size_t n, const std::function<void(size_t)>& f);
// The number of threads host_parallel_for may use (caller included).
size_t host_threads();

template <int LOG2>
struct SlotNode {
    static constexpr int DIM = 1 << LOG2;
    static constexpr int SIZE = 1 << (3 * LOG2);
    int32_t ox, oy, oz;
    std::vector<uint8_t> kind;
    std::vector<float> tile;
    std::vector<int32_t> child;
    SlotNode(int32_t x, int32_t y, int32_t z) : ox(x), oy(y), oz(z), kind(SIZE, SLOT_EMPTY), tile(SIZE, 0.0f), child(SIZE, -1) {}
};
using Node2 = SlotNode<4>;  // 16^3 slots of 8^3 leaves: covers 128^3
using Node1 = SlotNode<5>;  // 32^3 slots of Node2: covers 4096^3

struct Topology {
    using Key = std::array<int32_t, 3>;  // ordered (x, y, z), root_node/mod.rs:55-67
    std::map<Key, int32_t> root;         // -> Node1 index
    std::vector<Node1> n1;
    std::vector<Node2> n2;
    LeafStore leaves;

    // Builds the tree of a leaf set sorted or not, plus active tiles
    // (origin, edge 8 or 128, value).
    static std::unique_ptr<Topology> from_leaves(const std::vector<std::array<int32_t, 3>>& origins,
                                                 const std::vector<TileRec>& tiles);
    std::unique_ptr<Topology> clone() const { return std::make_unique<Topology>(*this); }

    // Flood fill of the internal levels and the root (leaves already filled,
    // their first/last signs in `leaves`): internal_node/flood_fill.rs:17-75,
    // root_node/flood_fill.rs:8-42.
    void flood_fill_upper();

    // Merges `other` into this topology (its leaves become B-side references,
    // its nodes are appended) and applies the CSG of root_node/csg.rs:9-58 /
    // internal_node/csg.rs:68-162 with this tree as `self`.
    void csg(Topology&& other, CsgOp op);

    // Leaves (in tree order) and active tiles of the result.
    void collect(std::vector<int32_t>& leaf_refs, std::vector<TileRec>& tiles) const;

    // The serial visit order of visit_leafs (root in key order, slots in offset
    // order): units[i] >= 0 is a leaf index, -1 - k is active tile tiles[k].
    void visit_units(std::vector<int64_t>& units, std::vector<TileRec>& tiles) const;

    // sign_at down to the leaf level (root_node/flood_fill.rs:56-63,
    // internal_node/flood_fill.rs:110-117): the index of the leaf holding
    // voxel (x, y, z), or -1 with *neg set from the slot's tile value (active
    // or not) or, for a missing root entry, positive.
    int32_t locate(int32_t x, int32_t y, int32_t z, bool* neg) const;

    // The tree after Volume::offset (volume/mod.rs:95-108): remove_if turns
    // active tiles with |v| > two_h inactive and drops the leaves whose index
    // has keep[i] == 0, empty nodes are pruned (internal_node/tree_node.rs:
    // 312-337, root_node/tree_node.rs:150-158), then the result's leaves
    // (origins in output order; leaf r is device leaf r) are inserted, creating
    // fresh nodes where needed, and empty nodes are pruned again
    // (remove_empty_branches, fast_sweep.rs:61-64).  out_origins must hold
    // every kept leaf (a sweep never empties a leaf that has an active voxel).
    std::unique_ptr<Topology> after_offset(const std::vector<uint32_t>& keep,
                                           const std::vector<std::array<int32_t, 3>>& out_origins, float two_h) const;

    // Every active tile becomes the leaves it covers (a Node1-level tile gets a
    // fresh 16^3 node of leaves): the new leaves are appended to `leaves` and
    // returned as (origin, edge 8, tile value) in index order.
    std::vector<TileRec> densify_tiles();

    // The reachable part of the tree with leaf reference r renumbered to
    // new_index[r] (a plain copy of the device leaf at that index).  Moves the
    // nodes out: this topology is empty afterwards.
    std::unique_ptr<Topology> compact(const std::vector<int32_t>& new_index);
};

}  // namespace bs
