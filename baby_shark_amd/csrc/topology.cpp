// Host-side topology: slot rules of flood fill and CSG for the two internal
// levels and the root of the volume grid.  See topology.hpp.
#include "topology.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <numeric>
#include <thread>

namespace bs {

namespace {

// A persistent pool of host threads (up to 15 beside the caller), so that the
// many short parallel loops of one operation do not each start threads.  Jobs
// from concurrent callers queue; every caller works on its own job and returns
// only when no helper is still inside it.
struct PoolJob {
    const std::function<void(size_t)>* f = nullptr;
    size_t n = 0, helpers_max = 0;
    std::atomic<size_t> next{0};
    size_t helpers = 0, finished = 0;  // guarded by the pool's mutex
};

class WorkerPool {
public:
    // never destroyed: its threads sleep until the process ends
    static WorkerPool& get() {
        static WorkerPool* p = new WorkerPool();
        return *p;
    }
    size_t threads() const { return workers_ + 1; }
    void run(size_t n, size_t max_threads, const std::function<void(size_t)>& f) {
        PoolJob job;
        job.f = &f;
        job.n = n;
        job.helpers_max = max_threads - 1;
        {
            std::lock_guard<std::mutex> l(m_);
            jobs_.push_back(&job);
        }
        cv_.notify_all();
        work(job);
        std::unique_lock<std::mutex> l(m_);
        // every item is handed out: no further helper may take the job
        auto it = std::find(jobs_.begin(), jobs_.end(), &job);
        if (it != jobs_.end()) jobs_.erase(it);
        done_.wait(l, [&] { return job.finished == job.helpers; });
    }

private:
    WorkerPool() {
        const size_t hw = std::max<size_t>(1, std::min<size_t>(16, std::thread::hardware_concurrency()));
        workers_ = hw - 1;
        for (size_t k = 0; k < workers_; ++k) std::thread([this] { loop(); }).detach();
    }
    static void work(PoolJob& j) {
        for (size_t i; (i = j.next.fetch_add(1, std::memory_order_relaxed)) < j.n;) (*j.f)(i);
    }
    void loop() {
        std::unique_lock<std::mutex> l(m_);
        for (;;) {
            cv_.wait(l, [&] { return !jobs_.empty(); });
            PoolJob* j = jobs_.front();
            ++j->helpers;
            if (j->helpers >= j->helpers_max || j->next.load(std::memory_order_relaxed) >= j->n) jobs_.pop_front();
            l.unlock();
            work(*j);
            l.lock();
            ++j->finished;
            done_.notify_all();
        }
    }
    std::mutex m_;
    std::condition_variable cv_, done_;
    std::deque<PoolJob*> jobs_;
    size_t workers_ = 0;
};

// Runs f(i) for i in [0, n) on up to 16 host threads, at least `grain` items
// per thread (serially when that leaves one).
template <class F>
void parallel_for(size_t n, F&& f, size_t grain = 8) {
    WorkerPool& pool = WorkerPool::get();
    const size_t T = std::min(pool.threads(), n / std::max<size_t>(grain, 1));
    if (T <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    const std::function<void(size_t)> fn = [&](size_t i) { f(i); };
    pool.run(n, T, fn);
}

}  // namespace

void host_parallel_for(size_t n, const std::function<void(size_t)>& f) { parallel_for(n, f, 1); }
size_t host_threads() { return WorkerPool::get().threads(); }

namespace {

constexpr int32_t L1_SPAN = 4096;  // voxels covered by a Node1
constexpr int32_t L2_SPAN = 128;   // voxels covered by a Node2

inline bool neg(float v) { return std::signbit(v); }
inline float far_value(bool negative) { return negative ? -F32_MAX_V : F32_MAX_V; }

// internal_node/mod.rs:130-136 with BRANCHING 5 (children of 128) and 4 (children of 8)
inline int l1_slot(int32_t x, int32_t y, int32_t z) {
    return (((x & (L1_SPAN - 1)) >> 7) << 10) | (((y & (L1_SPAN - 1)) >> 7) << 5) | ((z & (L1_SPAN - 1)) >> 7);
}
inline int l2_slot(int32_t x, int32_t y, int32_t z) {
    return (((x & (L2_SPAN - 1)) >> 3) << 8) | (((y & (L2_SPAN - 1)) >> 3) << 4) | ((z & (L2_SPAN - 1)) >> 3);
}

template <int LOG2>
inline void slot_origin(const SlotNode<LOG2>& n, int o, int child_span, int32_t& x, int32_t& y, int32_t& z) {
    constexpr int D = SlotNode<LOG2>::DIM;
    x = n.ox + (o / (D * D)) * child_span;
    y = n.oy + ((o / D) % D) * child_span;
    z = n.oz + (o % D) * child_span;
}

// first/last value sign of a node (internal_node/flood_fill.rs:94-108): the
// "last" sign of a branch in the last slot is that branch's *first* sign.
template <int LOG2, class ChildFirst>
bool first_neg(const SlotNode<LOG2>& n, ChildFirst child_first) {
    return n.kind[0] == SLOT_BRANCH ? child_first(n.child[0]) : neg(n.tile[0]);
}
template <int LOG2, class ChildFirst>
bool last_neg(const SlotNode<LOG2>& n, ChildFirst child_first) {
    constexpr int L = SlotNode<LOG2>::SIZE - 1;
    return n.kind[L] == SLOT_BRANCH ? child_first(n.child[L]) : neg(n.tile[L]);
}

// Scan-line sign propagation over the slots of one node
// (internal_node/flood_fill.rs:17-75).  child_first / child_last give the
// first / last value sign of a branch child.
template <int LOG2, class ChildFirst, class ChildLast>
void fill_node(SlotNode<LOG2>& n, ChildFirst child_first, ChildLast child_last) {
    constexpr int D = SlotNode<LOG2>::DIM, N = SlotNode<LOG2>::SIZE;
    int fv = -1, fb = -1;
    bool full = true;
    for (int o = 0; o < N; ++o) {
        const uint8_t k = n.kind[o];
        full &= k == SLOT_TILE;
        if (k == SLOT_TILE && fv < 0) fv = o;
        if (k == SLOT_BRANCH && fb < 0) fb = o;
    }
    if (full) return;
    bool s;
    if (fv >= 0 && fb >= 0) s = (fv <= fb) ? neg(n.tile[fv]) : child_first(n.child[fb]);
    else if (fv >= 0) s = neg(n.tile[fv]);
    else if (fb >= 0) s = child_first(n.child[fb]);
    else return;
    auto present = [&](int o, bool& sign) {
        if (n.kind[o] == SLOT_BRANCH) {
            sign = child_last(n.child[o]);
            return true;
        }
        if (n.kind[o] == SLOT_TILE) {
            sign = neg(n.tile[o]);
            return true;
        }
        return false;
    };
    bool i = s;
    for (int x = 0; x < D; ++x) {
        const int x00 = x * D * D;
        present(x00, i);
        bool j = i;
        for (int y = 0; y < D; ++y) {
            const int xy0 = x00 + y * D;
            present(xy0, j);
            bool k = j;
            for (int z = 0; z < D; ++z) {
                const int o = xy0 + z;
                if (!present(o, k)) n.tile[o] = far_value(k);
            }
        }
    }
}

}  // namespace

std::unique_ptr<Topology> Topology::from_leaves(const std::vector<std::array<int32_t, 3>>& origins,
                                                const std::vector<TileRec>& tiles) {
    auto t = std::make_unique<Topology>();
    Key last_key{INT32_MIN, 0, 0};
    int32_t last_n1 = -1;
    auto node1_for = [&](int32_t x, int32_t y, int32_t z) -> Node1& {
        Key k{x & ~(L1_SPAN - 1), y & ~(L1_SPAN - 1), z & ~(L1_SPAN - 1)};
        if (k == last_key) return t->n1[last_n1];
        auto it = t->root.find(k);
        if (it == t->root.end()) {
            t->n1.emplace_back(k[0], k[1], k[2]);
            it = t->root.emplace(k, (int32_t)t->n1.size() - 1).first;
        }
        last_key = k;
        last_n1 = it->second;
        return t->n1[it->second];
    };
    auto node2_for = [&](int32_t x, int32_t y, int32_t z) -> Node2& {
        Node1& a = node1_for(x, y, z);
        const int o = l1_slot(x, y, z);
        if (a.kind[o] != SLOT_BRANCH) {
            int32_t nx, ny, nz;
            slot_origin(a, o, L2_SPAN, nx, ny, nz);
            const int32_t idx = (int32_t)t->n2.size();
            t->n2.emplace_back(nx, ny, nz);
            Node1& a2 = node1_for(x, y, z);  // n1 may not move, but re-fetch for clarity
            a2.kind[o] = SLOT_BRANCH;
            a2.child[o] = idx;
            return t->n2.back();
        }
        return t->n2[a.child[o]];
    };
    t->leaves.resize(origins.size());
    LeafVec& recs = t->leaves.seg[0];
    const size_t n = origins.size();
    bool sorted = n >= 65536;
    for (size_t i = 1; sorted && i < n; ++i) sorted = origins[i - 1] < origins[i];
    if (sorted) {
        // Leaves sorted by origin: those of one 128-voxel x-slab are contiguous
        // and own the 32^3-node slots of that slab, so slabs are built in
        // parallel.  Nodes are numbered by first appearance, as serially: the
        // few distinct 4096^3 keys are listed per chunk in order of first
        // appearance, in parallel, and created chunk by chunk.
        {
            constexpr size_t CH = 32768;
            const size_t n_ch = (n + CH - 1) / CH;
            std::vector<std::vector<Key>> seen(n_ch);
            parallel_for(n_ch, [&](size_t c) {
                std::vector<Key>& ks = seen[c];
                Key last{INT32_MIN, 0, 0};
                for (size_t i = c * CH, e = std::min(n, (c + 1) * CH); i < e; ++i) {
                    const Key k{origins[i][0] & ~(L1_SPAN - 1), origins[i][1] & ~(L1_SPAN - 1),
                                origins[i][2] & ~(L1_SPAN - 1)};
                    if (k == last) continue;
                    last = k;
                    if (std::find(ks.begin(), ks.end(), k) == ks.end()) ks.push_back(k);
                }
            }, 1);
            for (const auto& ks : seen)
                for (const Key& k : ks) node1_for(k[0], k[1], k[2]);
        }
        std::vector<size_t> cut{0};
        for (size_t i = 1; i < n; ++i)
            if ((origins[i][0] >> 7) != (origins[i - 1][0] >> 7)) cut.push_back(i);
        cut.push_back(n);
        const size_t n_slab = cut.size() - 1;
        std::vector<std::vector<Node2>> local(n_slab);
        std::vector<std::vector<std::pair<int32_t, int>>> made(n_slab);  // (32^3 node, slot) per local node
        parallel_for(n_slab, [&](size_t sl) {
            Key lk{INT32_MIN, 0, 0};
            int32_t ln1 = -1;
            for (size_t i = cut[sl]; i < cut[sl + 1]; ++i) {
                const auto& g = origins[i];
                const Key k{g[0] & ~(L1_SPAN - 1), g[1] & ~(L1_SPAN - 1), g[2] & ~(L1_SPAN - 1)};
                if (k != lk) {
                    ln1 = t->root.find(k)->second;
                    lk = k;
                }
                Node1& a = t->n1[ln1];
                const int o1 = l1_slot(g[0], g[1], g[2]);
                if (a.kind[o1] != SLOT_BRANCH) {
                    int32_t nx, ny, nz;
                    slot_origin(a, o1, L2_SPAN, nx, ny, nz);
                    a.kind[o1] = SLOT_BRANCH;
                    a.child[o1] = (int32_t)local[sl].size();
                    local[sl].emplace_back(nx, ny, nz);
                    made[sl].push_back({ln1, o1});
                }
                Node2& b = local[sl][a.child[o1]];
                const int o = l2_slot(g[0], g[1], g[2]);
                b.kind[o] = SLOT_BRANCH;
                b.child[o] = (int32_t)i;
                recs[i] = leaf_ref(g[0], g[1], g[2], (int32_t)i);
            }
        }, 1);
        size_t total = 0;
        for (const auto& l : local) total += l.size();
        t->n2.reserve(total);
        for (size_t sl = 0; sl < n_slab; ++sl) {
            const int32_t base = (int32_t)t->n2.size();
            for (const auto& m : made[sl]) t->n1[m.first].child[m.second] += base;
            for (auto& nd : local[sl]) t->n2.push_back(std::move(nd));
        }
    }
    for (size_t i = 0; !sorted && i < n; ++i) {
        const auto& g = origins[i];
        Node2& b = node2_for(g[0], g[1], g[2]);
        const int o = l2_slot(g[0], g[1], g[2]);
        b.kind[o] = SLOT_BRANCH;
        b.child[o] = (int32_t)i;
        recs[i] = leaf_ref(g[0], g[1], g[2], (int32_t)i);
    }
    for (const auto& tl : tiles) {
        if (tl.size == L2_SPAN) {
            Node1& a = node1_for(tl.x, tl.y, tl.z);
            const int o = l1_slot(tl.x, tl.y, tl.z);
            a.kind[o] = SLOT_TILE;
            a.tile[o] = tl.value;
        } else {
            Node2& b = node2_for(tl.x, tl.y, tl.z);
            const int o = l2_slot(tl.x, tl.y, tl.z);
            b.kind[o] = SLOT_TILE;
            b.tile[o] = tl.value;
        }
    }
    return t;
}

void Topology::flood_fill_upper() {
    auto leaf_first = [&](int32_t li) { return leaves[li].first_neg != 0; };
    auto leaf_last = [&](int32_t li) { return leaves[li].last_neg != 0; };
    // nodes of one level are independent: each reads its children, writes itself
    parallel_for(n2.size(), [&](size_t i) { fill_node(n2[i], leaf_first, leaf_last); }, 1);
    auto n2_first = [&](int32_t k) { return first_neg(n2[k], leaf_first); };
    auto n2_last = [&](int32_t k) { return last_neg(n2[k], leaf_first); };
    parallel_for(n1.size(), [&](size_t i) { fill_node(n1[i], n2_first, n2_last); }, 1);
    // root_node/flood_fill.rs:14-41: negative tiles between two negative
    // children on one z line
    std::vector<Key> keys;
    for (auto& kv : root) keys.push_back(kv.first);
    for (size_t i = 0; i + 1 < keys.size(); ++i) {
        const Key& a = keys[i];
        const Key& b = keys[i + 1];
        if (a[0] != b[0] || a[1] != b[1] || b[2] == a[2] + L1_SPAN) continue;
        const bool a_neg = last_neg(n1[root[a]], n2_first);
        const bool b_neg = first_neg(n1[root[b]], n2_first);
        if (!a_neg || !b_neg) continue;
        for (int32_t z = a[2] + L1_SPAN; z < b[2]; z += L1_SPAN) {
            n1.emplace_back(a[0], a[1], z);
            Node1& fresh = n1.back();
            std::fill(fresh.tile.begin(), fresh.tile.end(), -F32_MAX_V);  // fill_with_sign(Negative)
            root[Key{a[0], a[1], z}] = (int32_t)n1.size() - 1;
        }
    }
}

namespace {

struct Csg {
    Topology& t;
    CsgOp op;

    template <int LOG2>
    static bool inside(const SlotNode<LOG2>& n, int o) { return n.kind[o] != SLOT_BRANCH && neg(n.tile[o]); }
    template <int LOG2>
    static bool outside(const SlotNode<LOG2>& n, int o) { return n.kind[o] != SLOT_BRANCH && !neg(n.tile[o]); }
    // internal_node/mod.rs:81-93
    template <int LOG2>
    static void remove_child(SlotNode<LOG2>& n, int o) {
        if (n.kind[o] != SLOT_EMPTY) n.kind[o] = SLOT_EMPTY;
    }
    // internal_node/csg.rs:42-48
    template <int LOG2>
    static void make_inside(SlotNode<LOG2>& n, int o) {
        n.kind[o] = SLOT_TILE;
        n.tile[o] = -F32_MAX_V;
    }
    // internal_node/csg.rs:31-39: move other's branch into self
    template <int LOG2>
    static void take(SlotNode<LOG2>& s, SlotNode<LOG2>& other, int o) {
        if (other.kind[o] != SLOT_BRANCH) return;
        std::swap(s.child[o], other.child[o]);
        std::swap(s.tile[o], other.tile[o]);
        other.kind[o] = s.kind[o];
        s.kind[o] = SLOT_BRANCH;
    }

    void flip_leaf(int32_t li) { t.leaves[li].neg ^= 1; }
    // internal_node/csg.rs:154-162 on a Node2
    void flip_node2(int32_t k) {
        Node2& n = t.n2[k];
        for (int o = 0; o < Node2::SIZE; ++o) {
            if (n.kind[o] == SLOT_BRANCH) flip_leaf(n.child[o]);
            else if (n.kind[o] == SLOT_TILE) n.tile[o] = -n.tile[o];
        }
    }

    // leaf_node/csg.rs:17-45: the combination is recorded, the device does it
    void leaf_pair(int32_t sa, int32_t ob) {
        LeafRef& r = t.leaves[sa];
        const LeafRef& b = t.leaves[ob];
        r.b = b.b;
        r.op = op == CSG_UNION ? LEAF_UNION : (op == CSG_SUBTRACT ? LEAF_SUBTRACT : LEAF_INTERSECT);
        r.empty = r.empty && b.empty;
    }
    bool node2_empty(int32_t k) const {
        const Node2& n = t.n2[k];
        for (int o = 0; o < Node2::SIZE; ++o)
            if (n.kind[o] != SLOT_EMPTY) return false;
        return true;
    }

    // internal_node/csg.rs:68-152 over the slots of one node; `pair` combines
    // two branch children and returns true when the result is empty.
    template <int LOG2, class Pair, class Flip>
    void slots(SlotNode<LOG2>& s, SlotNode<LOG2>& o_, Pair pair, Flip flip) {
        for (int o = 0; o < SlotNode<LOG2>::SIZE; ++o) {
            if (op == CSG_UNION) {
                if (inside(s, o)) continue;
                if (inside(o_, o)) {
                    make_inside(s, o);
                    continue;
                }
                if (outside(s, o)) {
                    take(s, o_, o);
                    continue;
                }
                if (s.kind[o] == SLOT_BRANCH && o_.kind[o] == SLOT_BRANCH)
                    if (pair(s.child[o], o_.child[o], o)) make_inside(s, o);
            } else if (op == CSG_SUBTRACT) {
                if (outside(s, o) || outside(o_, o)) continue;
                if (inside(o_, o)) {
                    remove_child(s, o);
                    continue;
                }
                if (inside(s, o)) {
                    take(s, o_, o);
                    if (s.kind[o] == SLOT_BRANCH) flip(s.child[o]);
                    else if (s.kind[o] == SLOT_TILE) s.tile[o] = -s.tile[o];
                    continue;
                }
                if (s.kind[o] == SLOT_BRANCH && o_.kind[o] == SLOT_BRANCH) pair(s.child[o], o_.child[o], o);
            } else {
                if (outside(s, o) || inside(o_, o)) continue;
                if (outside(o_, o)) {
                    remove_child(s, o);
                    continue;
                }
                if (inside(s, o)) {
                    take(s, o_, o);
                    continue;
                }
                if (s.kind[o] == SLOT_BRANCH && o_.kind[o] == SLOT_BRANCH) pair(s.child[o], o_.child[o], o);
            }
        }
    }

    bool node2_pair(int32_t a, int32_t b) {
        slots(t.n2[a], t.n2[b],
              [&](int32_t la, int32_t lb, int) {
                  leaf_pair(la, lb);
                  return t.leaves[la].empty != 0;
              },
              [&](int32_t li) { flip_leaf(li); });
        return node2_empty(a);
    }

    // The 16^3-node pairs of a 32^3 node touch disjoint slots and leaves, so
    // they are collected during the slot pass and combined in parallel; a
    // pair is the last step of its slot, so union's make_child_inside of an
    // emptied pair is applied afterwards with the same effect.
    void node1_pair(int32_t a, int32_t b) {
        std::vector<std::array<int32_t, 3>> pairs;  // (slot, node of self, node of other)
        std::vector<int32_t> flips;                 // 16^3 nodes moved in and negated
        slots(t.n1[a], t.n1[b],
              [&](int32_t ka, int32_t kb, int o) {
                  pairs.push_back({o, ka, kb});
                  return false;
              },
              [&](int32_t k) { flips.push_back(k); });
        // a negation touches only its own node and leaves, and no later step
        // of the slot pass reads them: applied afterwards, in parallel
        parallel_for(flips.size(), [&](size_t i) { flip_node2(flips[i]); }, 1);
        Node1& s = t.n1[a];
        std::vector<uint8_t> empty(pairs.size(), 0);
        parallel_for(pairs.size(), [&](size_t i) { empty[i] = node2_pair(pairs[i][1], pairs[i][2]); });
        if (op == CSG_UNION) {
            for (size_t i = 0; i < pairs.size(); ++i)
                if (empty[i]) make_inside(s, pairs[i][0]);
        }
    }
};

}  // namespace

void Topology::csg(Topology&& other, CsgOp op) {
    // other's pools are appended; its leaves become B-side references
    const int32_t n1_base = (int32_t)n1.size(), n2_base = (int32_t)n2.size(), leaf_base = (int32_t)leaves.size();
    for (auto& sg : other.leaves.seg) {
        parallel_for((sg.size() + 4095) / 4096, [&](size_t chunk) {
            const size_t e = std::min(sg.size(), (chunk + 1) * 4096);
            for (size_t i = chunk * 4096; i < e; ++i) {
                LeafRef& r = sg[i];
                r.b = r.a;
                r.a = -1;
                r.op = LEAF_COPY_B;
            }
        });
        leaves.seg.push_back(std::move(sg));
    }
    parallel_for(other.n2.size(), [&](size_t i) {
        Node2& n = other.n2[i];
        for (int o = 0; o < Node2::SIZE; ++o)
            if (n.kind[o] == SLOT_BRANCH) n.child[o] += leaf_base;
    });
    n2.reserve(n2.size() + other.n2.size());
    for (auto& n : other.n2) n2.push_back(std::move(n));
    for (auto& n : other.n1) {
        for (int o = 0; o < Node1::SIZE; ++o)
            if (n.kind[o] == SLOT_BRANCH) n.child[o] += n2_base;
        n1.push_back(std::move(n));
    }
    std::map<Key, int32_t> oroot;
    for (auto& kv : other.root) oroot[kv.first] = kv.second + n1_base;
    Csg c{*this, op};
    if (op == CSG_UNION) {  // root_node/csg.rs:9-26
        for (auto& kv : oroot) {
            auto it = root.find(kv.first);
            if (it != root.end()) c.node1_pair(it->second, kv.second);
            else root[kv.first] = kv.second;
        }
    } else if (op == CSG_SUBTRACT) {  // root_node/csg.rs:28-37
        for (auto& kv : oroot) {
            auto it = root.find(kv.first);
            if (it != root.end()) c.node1_pair(it->second, kv.second);
        }
    } else {  // root_node/csg.rs:39-54
        for (auto it = root.begin(); it != root.end();) {
            if (!oroot.count(it->first)) it = root.erase(it);
            else ++it;
        }
        for (auto& kv : root) c.node1_pair(kv.second, oroot[kv.first]);
    }
}

void Topology::collect(std::vector<int32_t>& leaf_refs, std::vector<TileRec>& tiles) const {
    // the 32^3-level items in tree order (a tile, or a 16^3 node whose slots
    // are gathered in parallel), then concatenated in that order
    struct Item {
        int32_t node;  // 16^3 node, or -1 for a tile
        TileRec tile;
    };
    std::vector<Item> items;
    for (const auto& kv : root) {
        const Node1& a = n1[kv.second];
        for (int o1 = 0; o1 < Node1::SIZE; ++o1) {
            if (a.kind[o1] == SLOT_TILE) {
                int32_t x, y, z;
                slot_origin(a, o1, L2_SPAN, x, y, z);
                items.push_back({-1, {x, y, z, L2_SPAN, a.tile[o1]}});
            } else if (a.kind[o1] == SLOT_BRANCH) {
                items.push_back({a.child[o1], {}});
            }
        }
    }
    std::vector<std::vector<int32_t>> refs(items.size());
    std::vector<std::vector<TileRec>> tl(items.size());
    parallel_for(items.size(), [&](size_t k) {
        if (items[k].node < 0) return;
        const Node2& b = n2[items[k].node];
        // filled in locals and moved once: pushing into refs[k] directly makes
        // the threads write the neighbouring vectors' headers, one cache line
        std::vector<int32_t> r;
        std::vector<TileRec> t;
        for (int o2 = 0; o2 < Node2::SIZE; ++o2) {
            if (b.kind[o2] == SLOT_BRANCH) {
                r.push_back(b.child[o2]);
            } else if (b.kind[o2] == SLOT_TILE) {
                int32_t x, y, z;
                slot_origin(b, o2, 8, x, y, z);
                t.push_back({x, y, z, 8, b.tile[o2]});
            }
        }
        refs[k] = std::move(r);
        tl[k] = std::move(t);
    });
    size_t nr = 0;
    for (const auto& r : refs) nr += r.size();
    leaf_refs.reserve(leaf_refs.size() + nr);
    for (size_t k = 0; k < items.size(); ++k) {
        if (items[k].node < 0) {
            tiles.push_back(items[k].tile);
            continue;
        }
        leaf_refs.insert(leaf_refs.end(), refs[k].begin(), refs[k].end());
        tiles.insert(tiles.end(), tl[k].begin(), tl[k].end());
    }
}

// internal_node/tree_node.rs:228-239, root_node/tree_node.rs:109-113
void Topology::visit_units(std::vector<int64_t>& units, std::vector<TileRec>& tiles) const {
    // The order: roots in key order, Node1 slots, then Node2 slots.  The
    // Node1 slots are listed serially (a Node2 index, or a 128^3 tile as
    // -1 - slot); the Node2 scans, the bulk of the work, count and then fill
    // their share of the output in parallel.
    struct Seg {
        int32_t n1i, o1, n2i;  // n2i < 0: a 128^3 tile at (n1i, o1)
        uint64_t u0, t0;       // first unit and first tile of this segment
    };
    std::vector<Seg> segs;
    for (const auto& kv : root) {
        const Node1& a = n1[kv.second];
        for (int o1 = 0; o1 < Node1::SIZE; ++o1) {
            if (a.kind[o1] == SLOT_TILE)
                segs.push_back({kv.second, o1, -1, 0, 0});
            else if (a.kind[o1] == SLOT_BRANCH)
                segs.push_back({kv.second, o1, a.child[o1], 0, 0});
        }
    }
    std::vector<uint32_t> n_units(segs.size()), n_tiles(segs.size());
    host_parallel_for(segs.size(), [&](size_t i) {
        if (segs[i].n2i < 0) {
            n_units[i] = n_tiles[i] = 1;
            return;
        }
        const Node2& b = n2[segs[i].n2i];
        uint32_t nu = 0, nt = 0;
        for (int o2 = 0; o2 < Node2::SIZE; ++o2) {
            nu += b.kind[o2] != SLOT_EMPTY;
            nt += b.kind[o2] == SLOT_TILE;
        }
        n_units[i] = nu;
        n_tiles[i] = nt;
    });
    const uint64_t u_base = units.size(), t_base = tiles.size();
    uint64_t nu = 0, nt = 0;
    for (size_t i = 0; i < segs.size(); ++i) {
        segs[i].u0 = nu;
        segs[i].t0 = nt;
        nu += n_units[i];
        nt += n_tiles[i];
    }
    units.resize(u_base + nu);
    tiles.resize(t_base + nt);
    host_parallel_for(segs.size(), [&](size_t i) {
        const Seg& g = segs[i];
        uint64_t u = u_base + g.u0, t = t_base + g.t0;
        const Node1& a = n1[g.n1i];
        if (g.n2i < 0) {
            int32_t x, y, z;
            slot_origin(a, g.o1, L2_SPAN, x, y, z);
            units[u] = -1 - (int64_t)t;
            tiles[t] = {x, y, z, L2_SPAN, a.tile[g.o1]};
            return;
        }
        const Node2& b = n2[g.n2i];
        for (int o2 = 0; o2 < Node2::SIZE; ++o2) {
            if (b.kind[o2] == SLOT_BRANCH) {
                units[u++] = b.child[o2];
            } else if (b.kind[o2] == SLOT_TILE) {
                int32_t x, y, z;
                slot_origin(b, o2, 8, x, y, z);
                units[u++] = -1 - (int64_t)t;
                tiles[t++] = {x, y, z, 8, b.tile[o2]};
            }
        }
    });
}

int32_t Topology::locate(int32_t x, int32_t y, int32_t z, bool* neg_out) const {
    auto it = root.find(Key{x & ~(L1_SPAN - 1), y & ~(L1_SPAN - 1), z & ~(L1_SPAN - 1)});
    if (it == root.end()) {
        *neg_out = false;
        return -1;
    }
    const Node1& a = n1[it->second];
    const int o1 = l1_slot(x, y, z);
    if (a.kind[o1] != SLOT_BRANCH) {
        *neg_out = neg(a.tile[o1]);
        return -1;
    }
    const Node2& b = n2[a.child[o1]];
    const int o2 = l2_slot(x, y, z);
    if (b.kind[o2] != SLOT_BRANCH) {
        *neg_out = neg(b.tile[o2]);
        return -1;
    }
    return b.child[o2];
}

std::unique_ptr<Topology> Topology::after_offset(const std::vector<uint32_t>& keep,
                                                const std::vector<std::array<int32_t, 3>>& out_origins,
                                                float two_h) const {
    // the nodes only: the leaf records are replaced below
    auto t = std::make_unique<Topology>();
    t->root = root;
    t->n1 = n1;
    t->n2 = n2;
    // remove_if: tiles beyond the band go inactive, emptied leaves go
    for (auto& n : t->n1)
        for (int o = 0; o < Node1::SIZE; ++o)
            if (n.kind[o] == SLOT_TILE && std::fabs(n.tile[o]) > two_h) n.kind[o] = SLOT_EMPTY;
    for (auto& n : t->n2)
        for (int o = 0; o < Node2::SIZE; ++o) {
            if (n.kind[o] == SLOT_TILE && std::fabs(n.tile[o]) > two_h) n.kind[o] = SLOT_EMPTY;
            if (n.kind[o] == SLOT_BRANCH) {
                const int32_t li = n.child[o];
                if (li < 0 || (size_t)li >= keep.size() || !keep[li]) n.kind[o] = SLOT_EMPTY;
                n.child[o] = -1;  // renumbered below
            }
        }
    auto prune = [&]() {
        for (auto it = t->root.begin(); it != t->root.end();) {
            Node1& a = t->n1[it->second];
            bool a_empty = true;
            for (int o1 = 0; o1 < Node1::SIZE; ++o1) {
                if (a.kind[o1] == SLOT_BRANCH) {
                    const Node2& b = t->n2[a.child[o1]];
                    bool b_empty = true;
                    for (int o2 = 0; o2 < Node2::SIZE && b_empty; ++o2) b_empty = b.kind[o2] == SLOT_EMPTY;
                    if (b_empty) a.kind[o1] = SLOT_EMPTY;
                }
                a_empty &= a.kind[o1] == SLOT_EMPTY;
            }
            if (a_empty) it = t->root.erase(it);
            else ++it;
        }
    };
    prune();
    // the new leaves: nodes first, created serially in leaf order for the
    // leaves whose 128^3 slot is no branch yet (the numbering of a one-by-one
    // insertion), then every leaf's 8^3 slot and record in parallel (each leaf
    // owns one slot)
    const size_t n_out = out_origins.size();
    t->leaves.resize(n_out);
    constexpr size_t CHUNK = 8192;
    const size_t n_chunks = (n_out + CHUNK - 1) / CHUNK;
    auto find_n1 = [&](const Topology& tt, int32_t x, int32_t y, int32_t z) -> int32_t {
        auto it = tt.root.find(Key{x & ~(L1_SPAN - 1), y & ~(L1_SPAN - 1), z & ~(L1_SPAN - 1)});
        return it == tt.root.end() ? -1 : it->second;
    };
    std::vector<uint8_t> need(n_out);
    host_parallel_for(n_chunks, [&](size_t c) {
        for (size_t i = c * CHUNK, e = std::min(n_out, i + CHUNK); i < e; ++i) {
            const int32_t x = out_origins[i][0], y = out_origins[i][1], z = out_origins[i][2];
            const int32_t ai = find_n1(*t, x, y, z);
            need[i] = ai < 0 || t->n1[ai].kind[l1_slot(x, y, z)] != SLOT_BRANCH;
        }
    });
    for (size_t i = 0; i < n_out; ++i) {
        if (!need[i]) continue;
        const int32_t x = out_origins[i][0], y = out_origins[i][1], z = out_origins[i][2];
        const Key k{x & ~(L1_SPAN - 1), y & ~(L1_SPAN - 1), z & ~(L1_SPAN - 1)};
        auto it = t->root.find(k);
        if (it == t->root.end()) {
            t->n1.emplace_back(k[0], k[1], k[2]);
            it = t->root.emplace(k, (int32_t)t->n1.size() - 1).first;
        }
        const int32_t ai = it->second;
        const int o1 = l1_slot(x, y, z);
        if (t->n1[ai].kind[o1] != SLOT_BRANCH) {
            int32_t nx, ny, nz;
            slot_origin(t->n1[ai], o1, L2_SPAN, nx, ny, nz);
            t->n2.emplace_back(nx, ny, nz);
            t->n1[ai].kind[o1] = SLOT_BRANCH;
            t->n1[ai].child[o1] = (int32_t)t->n2.size() - 1;
        }
    }
    LeafVec& recs = t->leaves.seg[0];
    host_parallel_for(n_chunks, [&](size_t c) {
        for (size_t i = c * CHUNK, e = std::min(n_out, i + CHUNK); i < e; ++i) {
            const int32_t x = out_origins[i][0], y = out_origins[i][1], z = out_origins[i][2];
            const Node1& a = t->n1[find_n1(*t, x, y, z)];
            Node2& b = t->n2[a.child[l1_slot(x, y, z)]];
            const int o2 = l2_slot(x, y, z);
            b.kind[o2] = SLOT_BRANCH;
            b.child[o2] = (int32_t)i;
            recs[i] = leaf_ref(x, y, z, (int32_t)i);
        }
    });
    prune();
    // drop the node records no slot refers to any more
    std::vector<int32_t> keep_leaf(out_origins.size());
    std::iota(keep_leaf.begin(), keep_leaf.end(), 0);
    return t->compact(keep_leaf);
}

std::vector<TileRec> Topology::densify_tiles() {
    std::vector<TileRec> made;
    auto new_leaf = [&](int32_t x, int32_t y, int32_t z, float v) {
        const int32_t idx = (int32_t)leaves.size();
        leaves.push_back(leaf_ref(x, y, z, idx));
        made.push_back({x, y, z, 8, v});
        return idx;
    };
    std::vector<int32_t> n1_ids;
    for (const auto& kv : root) n1_ids.push_back(kv.second);
    for (int32_t ai : n1_ids) {
        for (int o1 = 0; o1 < Node1::SIZE; ++o1) {
            const uint8_t k1 = n1[ai].kind[o1];
            if (k1 == SLOT_TILE) {
                const float v = n1[ai].tile[o1];
                int32_t nx, ny, nz;
                slot_origin(n1[ai], o1, L2_SPAN, nx, ny, nz);
                Node2 nb(nx, ny, nz);
                for (int o2 = 0; o2 < Node2::SIZE; ++o2) {
                    int32_t lx, ly, lz;
                    slot_origin(nb, o2, 8, lx, ly, lz);
                    nb.kind[o2] = SLOT_BRANCH;
                    nb.child[o2] = new_leaf(lx, ly, lz, v);
                }
                n2.push_back(std::move(nb));
                n1[ai].kind[o1] = SLOT_BRANCH;
                n1[ai].child[o1] = (int32_t)n2.size() - 1;
            } else if (k1 == SLOT_BRANCH) {
                const int32_t bi = n1[ai].child[o1];
                for (int o2 = 0; o2 < Node2::SIZE; ++o2) {
                    if (n2[bi].kind[o2] != SLOT_TILE) continue;
                    int32_t lx, ly, lz;
                    slot_origin(n2[bi], o2, 8, lx, ly, lz);
                    const int32_t li = new_leaf(lx, ly, lz, n2[bi].tile[o2]);
                    n2[bi].kind[o2] = SLOT_BRANCH;
                    n2[bi].child[o2] = li;
                }
            }
        }
    }
    return made;
}

std::unique_ptr<Topology> Topology::compact(const std::vector<int32_t>& new_index) {
    auto t = std::make_unique<Topology>();
    int32_t max_leaf = -1;
    for (int32_t v : new_index) max_leaf = std::max(max_leaf, v);
    t->leaves.resize((size_t)(max_leaf + 1));
    // nodes in tree order (moves only), then each 16^3 node's leaves in
    // parallel: every leaf sits in exactly one slot
    for (const auto& kv : root) {
        Node1 na = std::move(n1[kv.second]);
        for (int o1 = 0; o1 < Node1::SIZE; ++o1) {
            if (na.kind[o1] != SLOT_BRANCH) continue;
            t->n2.push_back(std::move(n2[na.child[o1]]));
            na.child[o1] = (int32_t)t->n2.size() - 1;
        }
        t->n1.push_back(std::move(na));
        t->root[kv.first] = (int32_t)t->n1.size() - 1;
    }
    parallel_for(t->n2.size(), [&](size_t k) {
        Node2& nb = t->n2[k];
        for (int o2 = 0; o2 < Node2::SIZE; ++o2) {
            if (nb.kind[o2] != SLOT_BRANCH) continue;
            const int32_t r = nb.child[o2];
            const int32_t ni = new_index[r];
            LeafRef& lr = t->leaves[ni];
            lr = leaves[r];
            lr.a = ni;
            lr.b = -1;
            lr.op = LEAF_COPY_A;
            lr.neg = 0;
            nb.child[o2] = ni;
        }
    });
    root.clear();
    n1.clear();
    n2.clear();
    leaves.seg.clear();
    return t;
}

}  // namespace bs
