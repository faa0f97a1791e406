// Test infrastructure only: the CPU oracle.  Nothing outside tests/, the smoke
// check and bench.py's cpu_baseline leg may link or call this code.
//
// Restatement of baby_shark's sparse volume tree (src/voxel/{mod.rs,init.rs},
// leaf_node/*, internal_node/*, root_node/*), the `dynamic_vdb!(f32, par 5, 4, 3)`
// grid of src/voxel/volume/mod.rs:8:
//   Root   : ordered map RootKey(index & !4095) -> Internal<5>   (root_node/mod.rs:13-32)
//   L1     : Internal, 32^3 children of 128^3                    (internal_node/mod.rs:16-29)
//   L2     : Internal, 16^3 children of 8^3
//   Leaf   : 8^3 f32 values + active mask                        (leaf_node/mod.rs:9-19)
// Masks are kept in index order; baby_shark stores bit i MSB-first inside a
// usize word (data_structures/bitset.rs:150-153), which changes no observable
// result because every query (`at`, `find_first_on`) is by index.
// Child slots hold both a branch pointer and a tile value; baby_shark overlays
// them in a union (internal_node/mod.rs:219-222).  The overlay is only visible
// through inactive slots that were never flood filled, which no path reads.
#pragma once
#include <algorithm>
#include <bitset>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "vec.hpp"

namespace oracle {

constexpr float F32_MAX = 3.40282347e+38f;

enum class Sign { Positive, Negative };

// value/f32.rs:9-32
inline Sign sign_of(float v) { return std::signbit(v) ? Sign::Negative : Sign::Positive; }
inline float with_sign(float v, Sign s) { return std::copysign(v, s == Sign::Negative ? -1.0f : 1.0f); }
inline float far_with_sign(Sign s) { return with_sign(F32_MAX, s); }

// utils.rs:40-61: ties and unordered pairs return the second argument.
inline float partial_min(float a, float b) { return (a < b) ? a : b; }
inline float partial_max(float a, float b) { return (a > b) ? a : b; }

template <int N>
struct Bits {
    std::bitset<N> b;
    bool at(int i) const { return b[i]; }
    void on(int i) { b[i] = true; }
    void off(int i) { b[i] = false; }
    void set(int i, bool v) { b[i] = v; }
    bool is_empty() const { return b.none(); }
    bool is_full() const { return b.all(); }
    void on_all() { b.set(); }
    void off_all() { b.reset(); }
    int find_first_on() const {
        for (int i = 0; i < N; ++i)
            if (b[i]) return i;
        return -1;
    }
};

template <int LOG2>
struct Leaf {
    static constexpr int BRANCHING = LOG2;
    static constexpr int BRANCHING_TOTAL = LOG2;
    static constexpr int DIM = 1 << LOG2;
    static constexpr int SIZE = 1 << (3 * LOG2);
    static constexpr bool IS_LEAF = true;
    using LeafT = Leaf;

    Vec3i origin;
    Bits<SIZE> value_mask;
    float values[SIZE];

    explicit Leaf(Vec3i o) : origin(o) {
        for (int i = 0; i < SIZE; ++i) values[i] = 0.0f;  // f32::default()
    }
    static int resolution() { return DIM; }
    // leaf_node/mod.rs:30-36
    static int offset(const Vec3i& i) {
        int64_t m = DIM - 1;
        return (int)(((i.x & m) << (2 * LOG2)) + ((i.y & m) << LOG2) + (i.z & m));
    }
    static Vec3i local_of(int off) {
        return {off >> (2 * LOG2), (off >> LOG2) & (DIM - 1), off & (DIM - 1)};
    }

    const float* at(const Vec3i& i) const {
        int o = offset(i);
        return value_mask.at(o) ? &values[o] : nullptr;
    }
    float* at_mut(const Vec3i& i) {
        int o = offset(i);
        return value_mask.at(o) ? &values[o] : nullptr;
    }
    void insert(const Vec3i& i, float v) {
        int o = offset(i);
        value_mask.on(o);
        values[o] = v;
    }
    void remove(const Vec3i& i) { value_mask.off(offset(i)); }
    bool is_empty() const { return value_mask.is_empty(); }
    void fill(float v) {
        value_mask.on_all();
        for (int i = 0; i < SIZE; ++i) values[i] = v;
    }
    std::unique_ptr<Leaf> clone() const { return std::make_unique<Leaf>(*this); }
    template <class F> void visit_leafs(F&& f) const { f(*this); }
    template <class F> void visit_leafs_mut(F&& f) { f(*this); }
    template <class F> void visit_values_mut(F&& f) {
        for (int i = 0; i < SIZE; ++i)
            if (value_mask.at(i)) f(values[i]);
    }
    const Leaf* leaf_at(const Vec3i&) const { return this; }
    template <class P> void remove_if(P&& pred) {  // leaf_node/tree_node.rs:202-211
        for (int i = 0; i < SIZE; ++i)
            if (value_mask.at(i) && pred(values[i])) value_mask.off(i);
    }
    void remove_empty_branches() {}

    // ---- flood fill (leaf_node/flood_fill.rs:12-76)
    void flood_fill() {
        if (value_mask.is_full()) return;
        int first = value_mask.find_first_on();
        if (first < 0) return;
        Sign i = sign_of(values[first]);
        for (int x = 0; x < DIM; ++x) {
            int x00 = x << (2 * LOG2);
            if (value_mask.at(x00)) i = sign_of(values[x00]);
            Sign j = i;
            for (int y = 0; y < DIM; ++y) {
                int xy0 = x00 + (y << LOG2);
                if (value_mask.at(xy0)) j = sign_of(values[xy0]);
                Sign k = j;
                for (int z = 0; z < DIM; ++z) {
                    int xyz = xy0 + z;
                    if (value_mask.at(xyz)) {
                        k = sign_of(values[xyz]);
                    } else {
                        values[xyz] = far_with_sign(k);
                    }
                }
            }
        }
    }
    void fill_with_sign(Sign s) {
        for (int i = 0; i < SIZE; ++i) {
            if (!value_mask.at(i)) values[i] = F32_MAX;
            values[i] = with_sign(values[i], s);
        }
    }
    Sign first_value_sign() const { return sign_of(values[0]); }
    Sign last_value_sign() const { return sign_of(values[SIZE - 1]); }
    Sign sign_at(const Vec3i& i) const { return sign_of(values[offset(i)]); }

    // ---- CSG (leaf_node/csg.rs:17-45), over every entry including inactive ones
    void csg_union(std::unique_ptr<Leaf> o) {
        for (int i = 0; i < SIZE; ++i) values[i] = partial_min(values[i], o->values[i]);
        value_mask.b |= o->value_mask.b;
    }
    void csg_subtract(std::unique_ptr<Leaf> o) {
        for (int i = 0; i < SIZE; ++i) values[i] = partial_max(values[i], -o->values[i]);
        value_mask.b |= o->value_mask.b;
    }
    void csg_intersect(std::unique_ptr<Leaf> o) {
        for (int i = 0; i < SIZE; ++i) values[i] = partial_max(values[i], o->values[i]);
        value_mask.b |= o->value_mask.b;
    }
    void flip_signs() {
        for (int i = 0; i < SIZE; ++i) values[i] = -values[i];
    }
};

// internal_node/{mod.rs,tree_node.rs,flood_fill.rs,csg.rs}
template <class Child, int LOG2>
struct Internal {
    static constexpr int BRANCHING = LOG2;
    static constexpr int BRANCHING_TOTAL = LOG2 + Child::BRANCHING_TOTAL;
    static constexpr int DIM = 1 << LOG2;
    static constexpr int SIZE = 1 << (3 * LOG2);
    static constexpr bool IS_LEAF = false;
    using LeafT = typename Child::LeafT;
    using ChildT = Child;

    Vec3i origin;
    Bits<SIZE> child_mask;
    Bits<SIZE> value_mask;
    std::vector<std::unique_ptr<Child>> branch;
    std::vector<float> tile;

    explicit Internal(Vec3i o) : origin(o), branch(SIZE), tile(SIZE, 0.0f) {}
    Internal(const Internal&) = delete;

    static int resolution() { return 1 << BRANCHING_TOTAL; }
    // internal_node/mod.rs:130-136
    static int offset(const Vec3i& i) {
        int64_t m = (int64_t(1) << BRANCHING_TOTAL) - 1;
        int ct = Child::BRANCHING_TOTAL;
        return (int)((((i.x & m) >> ct) << (2 * LOG2)) + (((i.y & m) >> ct) << LOG2) + ((i.z & m) >> ct));
    }
    // internal_node/mod.rs:42-50,139-148
    Vec3i offset_to_global_index(int off) const {
        int64_t x = off >> (2 * LOG2);
        int64_t y = (off >> LOG2) & (DIM - 1);
        int64_t z = off & (DIM - 1);
        int ct = Child::BRANCHING_TOTAL;
        return {(x << ct) + origin.x, (y << ct) + origin.y, (z << ct) + origin.z};
    }
    bool is_branch(int o) const { return child_mask.at(o); }
    bool is_tile(int o) const { return !child_mask.at(o) && value_mask.at(o); }

    // internal_node/mod.rs:52-67
    Child* add_branch(int o) {
        if (child_mask.at(o)) return branch[o].get();
        child_mask.on(o);
        value_mask.off(o);
        branch[o] = std::make_unique<Child>(offset_to_global_index(o));
        return branch[o].get();
    }
    std::unique_ptr<Child> remove_branch(int o) {
        if (!child_mask.at(o)) return nullptr;
        child_mask.off(o);
        return std::move(branch[o]);
    }
    // internal_node/mod.rs:81-93 (the tile case only clears the value bit)
    void remove_child(int o) {
        if (child_mask.at(o)) {
            child_mask.off(o);
            branch[o].reset();
        } else if (value_mask.at(o)) {
            value_mask.off(o);
        }
    }

    const float* at(const Vec3i& i) const {
        int o = offset(i);
        if (child_mask.at(o)) return branch[o]->at(i);
        if (value_mask.at(o)) return &tile[o];
        return nullptr;
    }
    float* at_mut(const Vec3i& i) {
        int o = offset(i);
        if (child_mask.at(o)) return branch[o]->at_mut(i);
        if (value_mask.at(o)) return &tile[o];
        return nullptr;
    }
    // internal_node/tree_node.rs:54-79
    void insert(const Vec3i& i, float v) {
        int o = offset(i);
        if (child_mask.at(o)) {
            branch[o]->insert(i, v);
        } else if (value_mask.at(o)) {
            float tv = tile[o];
            if (tv == v) return;
            Child* b = add_branch(o);
            b->fill(tv);
            b->insert(i, v);
        } else {
            add_branch(o)->insert(i, v);
        }
    }
    // internal_node/tree_node.rs:81-103
    void remove(const Vec3i& i) {
        int o = offset(i);
        if (child_mask.at(o)) {
            branch[o]->remove(i);
            if (branch[o]->is_empty()) remove_branch(o);
        } else if (value_mask.at(o)) {
            float tv = tile[o];
            Child* b = add_branch(o);
            b->fill(tv);
            b->remove(i);
        }
    }
    bool is_empty() const { return child_mask.is_empty() && value_mask.is_empty(); }
    void clear() {
        for (int o = 0; o < SIZE; ++o) branch[o].reset();
        child_mask.off_all();
        value_mask.off_all();
    }
    void fill(float v) {
        clear();
        value_mask.on_all();
        for (int o = 0; o < SIZE; ++o) tile[o] = v;
    }
    std::unique_ptr<Internal> clone() const {
        auto c = std::make_unique<Internal>(origin);
        c->child_mask = child_mask;
        c->value_mask = value_mask;
        c->tile = tile;
        for (int o = 0; o < SIZE; ++o)
            if (child_mask.at(o)) c->branch[o] = branch[o]->clone();
        return c;
    }
    // visit order of internal_node/tree_node.rs:228-239 (children in offset order)
    template <class FT, class FD> void visit_leafs(FT&& ft, FD&& fd) const {
        for (int o = 0; o < SIZE; ++o) {
            if (child_mask.at(o)) {
                visit_child(*branch[o], ft, fd);
            } else if (value_mask.at(o)) {
                ft(offset_to_global_index(o), Child::resolution(), tile[o]);
            }
        }
    }
    template <class FT, class FD> static void visit_child(const Child& c, FT& ft, FD& fd) {
        if constexpr (Child::IS_LEAF) fd(c);
        else c.visit_leafs(ft, fd);
    }
    template <class FD> void visit_leaves_mut(FD&& fd) {
        for (int o = 0; o < SIZE; ++o) {
            if (!child_mask.at(o)) continue;
            if constexpr (Child::IS_LEAF) fd(*branch[o]);
            else branch[o]->visit_leaves_mut(fd);
        }
    }
    template <class F> void visit_values_mut(F&& f) {
        for (int o = 0; o < SIZE; ++o) {
            if (child_mask.at(o)) branch[o]->visit_values_mut(f);
            else if (value_mask.at(o)) f(tile[o]);
        }
    }
    const LeafT* leaf_at(const Vec3i& i) const {
        int o = offset(i);
        if (!child_mask.at(o)) return nullptr;
        return branch[o]->leaf_at(i);
    }
    // internal_node/tree_node.rs:263-278
    std::unique_ptr<LeafT> take_leaf_at(const Vec3i& i) {
        int o = offset(i);
        if (!child_mask.at(o)) return nullptr;
        if constexpr (Child::IS_LEAF) return remove_branch(o);
        else return branch[o]->take_leaf_at(i);
    }
    // internal_node/tree_node.rs:280-298
    void insert_leaf_at(std::unique_ptr<LeafT> leaf) {
        Vec3i idx = leaf->origin;
        int o = offset(idx);
        value_mask.off(o);
        if constexpr (Child::IS_LEAF) {
            remove_branch(o);
            child_mask.on(o);
            value_mask.off(o);
            branch[o] = std::move(leaf);
        } else {
            if (child_mask.at(o)) branch[o]->insert_leaf_at(std::move(leaf));
            else add_branch(o)->insert_leaf_at(std::move(leaf));
        }
    }
    // internal_node/tree_node.rs:300-310
    void remove_empty_branches() {
        for (int o = 0; o < SIZE; ++o) {
            if (!child_mask.at(o)) continue;
            branch[o]->remove_empty_branches();
            if (branch[o]->is_empty()) remove_branch(o);
        }
    }
    // internal_node/tree_node.rs:312-337
    template <class P> void remove_if(P&& pred) {
        for (int o = 0; o < SIZE; ++o) {
            if (child_mask.at(o)) {
                branch[o]->remove_if(pred);
                if (branch[o]->is_empty()) remove_branch(o);
            } else if (value_mask.at(o)) {
                if (pred(tile[o])) value_mask.off(o);
            }
        }
    }

    // ---- flood fill (internal_node/flood_fill.rs:17-117)
    void flood_fill() {
        if (value_mask.is_full()) return;
        for (int o = 0; o < SIZE; ++o)
            if (child_mask.at(o)) branch[o]->flood_fill();
        int fv = value_mask.find_first_on();
        int fb = child_mask.find_first_on();
        Sign i;
        if (fv >= 0 && fb >= 0) i = (fv <= fb) ? sign_of(tile[fv]) : branch[fb]->first_value_sign();
        else if (fv >= 0) i = sign_of(tile[fv]);
        else if (fb >= 0) i = branch[fb]->first_value_sign();
        else return;
        auto slot_sign = [&](int o, Sign& s) {
            if (child_mask.at(o)) { s = branch[o]->last_value_sign(); return true; }
            if (value_mask.at(o)) { s = sign_of(tile[o]); return true; }
            return false;
        };
        for (int x = 0; x < DIM; ++x) {
            int x00 = x << (2 * LOG2);
            slot_sign(x00, i);
            Sign j = i;
            for (int y = 0; y < DIM; ++y) {
                int xy0 = x00 + (y << LOG2);
                slot_sign(xy0, j);
                Sign k = j;
                for (int z = 0; z < DIM; ++z) {
                    int xyz = xy0 + z;
                    if (!slot_sign(xyz, k)) tile[xyz] = far_with_sign(k);
                }
            }
        }
    }
    void fill_with_sign(Sign s) {
        if (value_mask.is_full()) return;
        for (int o = 0; o < SIZE; ++o) {
            if (child_mask.at(o)) branch[o]->fill_with_sign(s);
            else if (value_mask.at(o)) tile[o] = with_sign(tile[o], s);
            else tile[o] = far_with_sign(s);
        }
    }
    Sign first_value_sign() const {
        if (child_mask.at(0)) return branch[0]->first_value_sign();
        return sign_of(tile[0]);
    }
    Sign sign_at(const Vec3i& i) const {  // internal_node/flood_fill.rs:110-117
        int o = offset(i);
        if (child_mask.at(o)) return branch[o]->sign_at(i);
        return sign_of(tile[o]);
    }
    // internal_node/flood_fill.rs:103-107 returns the branch's *first* sign.
    Sign last_value_sign() const {
        if (child_mask.at(SIZE - 1)) return branch[SIZE - 1]->first_value_sign();
        return sign_of(tile[SIZE - 1]);
    }

    // ---- CSG (internal_node/csg.rs:20-162)
    bool is_inside_tile(int o) const { return !child_mask.at(o) && sign_of(tile[o]) == Sign::Negative; }
    bool is_outside_tile(int o) const { return !child_mask.at(o) && sign_of(tile[o]) == Sign::Positive; }
    void take_child(Internal& other, int o) {
        if (other.child_mask.at(o)) {
            other.child_mask.set(o, child_mask.at(o));
            other.value_mask.set(o, value_mask.at(o));
            child_mask.on(o);
            value_mask.off(o);
            std::swap(branch[o], other.branch[o]);
            std::swap(tile[o], other.tile[o]);
        }
    }
    void make_child_inside(int o) {
        remove_child(o);
        value_mask.on(o);
        tile[o] = -F32_MAX;
    }
    void csg_union(std::unique_ptr<Internal> other) {
        for (int o = 0; o < SIZE; ++o) {
            if (is_inside_tile(o)) continue;
            if (other->is_inside_tile(o)) { make_child_inside(o); continue; }
            if (is_outside_tile(o)) { take_child(*other, o); continue; }
            if (child_mask.at(o) && other->child_mask.at(o)) {
                auto ob = other->remove_branch(o);
                branch[o]->csg_union(std::move(ob));
                if (branch[o]->is_empty()) make_child_inside(o);
            } else {
                other->remove_child(o);
            }
        }
    }
    void csg_subtract(std::unique_ptr<Internal> other) {
        for (int o = 0; o < SIZE; ++o) {
            if (is_outside_tile(o) || other->is_outside_tile(o)) continue;
            if (other->is_inside_tile(o)) { remove_child(o); continue; }
            if (is_inside_tile(o)) {
                take_child(*other, o);
                if (child_mask.at(o)) branch[o]->flip_signs();
                else if (value_mask.at(o)) tile[o] = -tile[o];
                continue;
            }
            if (child_mask.at(o) && other->child_mask.at(o)) {
                auto ob = other->remove_branch(o);
                branch[o]->csg_subtract(std::move(ob));
            } else {
                other->remove_child(o);
            }
        }
    }
    void csg_intersect(std::unique_ptr<Internal> other) {
        for (int o = 0; o < SIZE; ++o) {
            if (is_outside_tile(o) || other->is_inside_tile(o)) continue;
            if (other->is_outside_tile(o)) { remove_child(o); continue; }
            if (is_inside_tile(o)) { take_child(*other, o); continue; }
            if (child_mask.at(o) && other->child_mask.at(o)) {
                auto ob = other->remove_branch(o);
                branch[o]->csg_intersect(std::move(ob));
            } else {
                other->remove_child(o);
            }
        }
    }
    void flip_signs() {
        for (int o = 0; o < SIZE; ++o) {
            if (child_mask.at(o)) branch[o]->flip_signs();
            else if (value_mask.at(o)) tile[o] = -tile[o];
        }
    }
};

// root_node/{mod.rs,tree_node.rs,flood_fill.rs,csg.rs}: ordered map keyed by
// `index & !(resolution-1)` compared lexicographically (x, y, z).
template <class Child>
struct Root {
    using LeafT = typename Child::LeafT;
    using ChildT = Child;
    std::map<Vec3i, std::unique_ptr<Child>> root;

    static Vec3i root_key(const Vec3i& i) {
        int64_t m = ~((int64_t(1) << Child::BRANCHING_TOTAL) - 1);
        return {i.x & m, i.y & m, i.z & m};
    }
    const float* at(const Vec3i& i) const {
        auto it = root.find(root_key(i));
        return it == root.end() ? nullptr : it->second->at(i);
    }
    float* at_mut(const Vec3i& i) {
        auto it = root.find(root_key(i));
        return it == root.end() ? nullptr : it->second->at_mut(i);
    }
    Child* child_for(const Vec3i& i) {
        Vec3i k = root_key(i);
        auto it = root.find(k);
        if (it == root.end()) it = root.emplace(k, std::make_unique<Child>(k)).first;
        return it->second.get();
    }
    void insert(const Vec3i& i, float v) { child_for(i)->insert(i, v); }
    void remove(const Vec3i& i) {
        auto it = root.find(root_key(i));
        if (it == root.end()) return;
        it->second->remove(i);
        if (it->second->is_empty()) root.erase(it);
    }
    bool is_empty() const {
        for (auto& kv : root)
            if (!kv.second->is_empty()) return false;
        return true;
    }
    void clear() { root.clear(); }
    std::unique_ptr<Root> clone() const {
        auto c = std::make_unique<Root>();
        for (auto& kv : root) c->root.emplace(kv.first, kv.second->clone());
        return c;
    }
    template <class FT, class FD> void visit_leafs(FT&& ft, FD&& fd) const {
        for (auto& kv : root) kv.second->visit_leafs(ft, fd);
    }
    template <class FD> void visit_leaves_mut(FD&& fd) {
        for (auto& kv : root) kv.second->visit_leaves_mut(fd);
    }
    template <class F> void visit_values_mut(F&& f) {
        for (auto& kv : root) kv.second->visit_values_mut(f);
    }
    const LeafT* leaf_at(const Vec3i& i) const {
        auto it = root.find(root_key(i));
        return it == root.end() ? nullptr : it->second->leaf_at(i);
    }
    std::unique_ptr<LeafT> take_leaf_at(const Vec3i& i) {
        auto it = root.find(root_key(i));
        return it == root.end() ? nullptr : it->second->take_leaf_at(i);
    }
    void insert_leaf_at(std::unique_ptr<LeafT> leaf) {
        Vec3i idx = leaf->origin;
        child_for(idx)->insert_leaf_at(std::move(leaf));
    }
    void remove_empty_branches() {
        for (auto it = root.begin(); it != root.end();) {
            it->second->remove_empty_branches();
            if (it->second->is_empty()) it = root.erase(it);
            else ++it;
        }
    }
    template <class P> void remove_if(P&& pred) {
        for (auto it = root.begin(); it != root.end();) {
            it->second->remove_if(pred);
            if (it->second->is_empty()) it = root.erase(it);
            else ++it;
        }
    }
    Sign sign_at(const Vec3i& i) const {  // root_node/flood_fill.rs:56-63
        auto it = root.find(root_key(i));
        return it == root.end() ? Sign::Positive : it->second->sign_at(i);
    }
    // root_node/flood_fill.rs:8-42
    void flood_fill() {
        if (root.empty()) return;
        for (auto& kv : root) kv.second->flood_fill();
        std::vector<Vec3i> keys;
        for (auto& kv : root) keys.push_back(kv.first);
        int64_t res = Child::resolution();
        for (size_t n = 0; n + 1 < keys.size(); ++n) {
            const Vec3i& a = keys[n];
            const Vec3i& b = keys[n + 1];
            if (a.x != b.x || a.y != b.y || b.z == a.z + res) continue;
            Sign as = root[a]->last_value_sign();
            Sign bs = root[b]->first_value_sign();
            if (as != Sign::Negative || bs != Sign::Negative) continue;
            Vec3i t = {a.x, a.y, a.z + res};
            while (t.z < b.z) {
                Vec3i k = root_key(t);
                auto node = std::make_unique<Child>(k);
                node->fill_with_sign(Sign::Negative);
                root[k] = std::move(node);
                t.z += res;
            }
        }
    }
    // root_node/csg.rs:9-58
    void csg_union(std::unique_ptr<Root> other) {
        std::vector<Vec3i> keys;
        for (auto& kv : root) keys.push_back(kv.first);
        for (auto& kv : other->root) keys.push_back(kv.first);
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        for (auto& k : keys) {
            auto it = root.find(k);
            auto ot = other->root.find(k);
            if (ot == other->root.end()) continue;
            auto n2 = std::move(ot->second);
            other->root.erase(ot);
            if (it != root.end()) it->second->csg_union(std::move(n2));
            else root.emplace(k, std::move(n2));
        }
    }
    void csg_subtract(std::unique_ptr<Root> other) {
        for (auto& kv : other->root) {
            auto it = root.find(kv.first);
            if (it != root.end()) it->second->csg_subtract(std::move(kv.second));
        }
    }
    void csg_intersect(std::unique_ptr<Root> other) {
        for (auto it = root.begin(); it != root.end();) {
            if (other->root.find(it->first) == other->root.end()) it = root.erase(it);
            else ++it;
        }
        for (auto& kv : root) {
            auto ot = other->root.find(kv.first);
            if (ot != other->root.end()) kv.second->csg_intersect(std::move(ot->second));
        }
    }
    void flip_signs() {
        for (auto& kv : root) kv.second->flip_signs();
    }
};

// The grid of src/voxel/volume/mod.rs:8.
using LeafNode = Leaf<3>;
using Node2 = Internal<LeafNode, 4>;
using Node1 = Internal<Node2, 5>;
using VolumeGrid = Root<Node1>;

}  // namespace oracle
