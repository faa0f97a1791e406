// Volume operations on the device: sampling a sphere (Volume::from_fn for
// VolumeBuilder::sphere), the leaf flood fill, and CSG.
//
// Division of labour.  The leaf payloads (8^3 f32 values + a 512-bit mask per
// leaf) live in device memory and every pass over them is a streaming kernel
// (one wavefront per leaf, 8 values per lane).  The upper tree levels are a
// few hundred slot nodes per volume; their flood fill and CSG slot decisions
// run on the host (topology.cpp) and produce, for each output leaf, a
// reference to its source leaves and the combination to apply.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <chrono>
#include <future>
#include <mutex>
#include <unordered_map>
#include <vector>
#include <cstring>
#include <numeric>
#include <tuple>

#include "common.hpp"

namespace bs {

namespace {

constexpr int WAVES_PER_BLOCK = 4;  // one leaf per wavefront, 256-thread blocks
constexpr float FAR = 3.40282347e+38f;

// Highest set bit of m at or below position `top` (-1 when none).
__device__ inline int last_at_or_below(uint64_t m, int top) {
    const uint64_t keep = (top >= 63) ? ~0ull : ((2ull << top) - 1);
    m &= keep;
    return m ? 63 - __builtin_clzll(m) : -1;
}

}  // namespace

// ---------------------------------------------------------------- leaf arrays
// A volume's leaf arrays (origins, masks, values: 2128 B per leaf) are taken
// from a small per-device pool of released sets.  A mesh-to-volume call that
// is repeated on meshes of similar size then reuses the arrays of the volume
// freed before it instead of paying hipFree (which unmaps and synchronises)
// and hipMalloc for a few hundred MB each time.  Release synchronises the
// device first, as hipFree does, so no kernel still reads a set that is handed
// out again.  The pool keeps at most POOL_SETS sets and POOL_BYTES bytes per
// device; bs_release_cached_memory empties it.
namespace {

constexpr uint64_t LEAF_BYTES = sizeof(int32_t) * 4 + sizeof(uint64_t) * 8 + sizeof(float) * 512;
constexpr int POOL_SETS = 4;
constexpr uint64_t POOL_BYTES = 8ull << 30;
constexpr int POOL_DEVICES = 64;

struct LeafSet {
    int32_t* origins;
    uint64_t* masks;
    float* values;
    uint64_t cap;  // leaves
};

struct LeafPool {
    std::mutex mu;
    std::unordered_map<const float*, uint64_t> live_cap;  // values pointer -> capacity, sets in use
    std::vector<LeafSet> idle[POOL_DEVICES];              // released sets, oldest first
    uint64_t idle_bytes[POOL_DEVICES] = {};
};

LeafPool& leaf_pool() {
    static LeafPool* p = new LeafPool();  // never destroyed: device memory may outlive static teardown
    return *p;
}

void free_set(const LeafSet& s) {
    (void)hipFree(s.origins);
    (void)hipFree(s.masks);
    (void)hipFree(s.values);
}

}  // namespace

void alloc_leaves(bs_volume* v, uint64_t n) {
    v->n_leaves = n;
    v->origins = nullptr;
    v->masks = nullptr;
    v->values = nullptr;
    if (!n) return;
    LeafPool& P = leaf_pool();
    const int dev = v->device;
    if (dev >= 0 && dev < POOL_DEVICES) {
        std::lock_guard<std::mutex> lk(P.mu);
        auto& idle = P.idle[dev];
        int best = -1;
        for (int i = 0; i < (int)idle.size(); ++i)  // smallest set that holds n without wasting more than n
            if (idle[i].cap >= n && idle[i].cap <= 2 * n && (best < 0 || idle[i].cap < idle[best].cap)) best = i;
        if (best >= 0) {
            const LeafSet s = idle[best];
            idle.erase(idle.begin() + best);
            P.idle_bytes[dev] -= s.cap * LEAF_BYTES;
            P.live_cap[s.values] = s.cap;
            v->origins = s.origins;
            v->masks = s.masks;
            v->values = s.values;
            return;
        }
    }
    auto attempt = [&]() -> hipError_t {
        hipError_t e = hipMalloc(&v->origins, sizeof(int32_t) * 4 * n);
        if (e == hipSuccess) e = hipMalloc(&v->masks, sizeof(uint64_t) * 8 * n);
        if (e == hipSuccess) e = hipMalloc(&v->values, sizeof(float) * 512 * n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            (void)hipFree(v->origins);
            (void)hipFree(v->masks);
            (void)hipFree(v->values);
            v->origins = nullptr;
            v->masks = nullptr;
            v->values = nullptr;
        }
        return e;
    };
    hipError_t e = attempt();
    if (e == hipErrorOutOfMemory) {  // give the pooled sets back and try once more
        release_cached_leaves();
        e = attempt();
    }
    BS_HIP(e);
    std::lock_guard<std::mutex> lk(P.mu);
    P.live_cap[v->values] = n;
}

void free_leaves(bs_volume* v, hipStream_t pending) {
    if (v->values || v->origins || v->masks) {
        LeafPool& P = leaf_pool();
        const int dev = v->device;
        bool pooled = false;
        std::vector<LeafSet> evict;
        int cur = -1;
        // Every library call returns with its stream's work done, so only a
        // call that fails part-way (`pending`: its stream) can leave kernels
        // that still touch the arrays; that stream alone is waited for.
        const bool idle = !pending || hipStreamSynchronize(pending) == hipSuccess;
        if (idle && dev >= 0 && dev < POOL_DEVICES && v->origins && v->masks && v->values &&
            hipGetDevice(&cur) == hipSuccess && cur == dev) {
            std::lock_guard<std::mutex> lk(P.mu);
            auto it = P.live_cap.find(v->values);
            if (it != P.live_cap.end()) {
                const LeafSet s{v->origins, v->masks, v->values, it->second};
                P.live_cap.erase(it);
                if (s.cap * LEAF_BYTES <= POOL_BYTES) {
                    auto& idle = P.idle[dev];
                    idle.push_back(s);
                    P.idle_bytes[dev] += s.cap * LEAF_BYTES;
                    while ((int)idle.size() > POOL_SETS || P.idle_bytes[dev] > POOL_BYTES) {
                        evict.push_back(idle.front());
                        P.idle_bytes[dev] -= idle.front().cap * LEAF_BYTES;
                        idle.erase(idle.begin());
                    }
                    pooled = true;
                }
            }
        }
        if (!pooled) {
            {
                std::lock_guard<std::mutex> lk(P.mu);
                if (v->values) P.live_cap.erase(v->values);
            }
            (void)hipFree(v->origins);
            (void)hipFree(v->masks);
            (void)hipFree(v->values);
        }
        for (const LeafSet& s : evict) free_set(s);
    }
    v->origins = nullptr;
    v->masks = nullptr;
    v->values = nullptr;
    v->n_leaves = 0;
}

void release_cached_leaves() {
    LeafPool& P = leaf_pool();
    std::vector<LeafSet> all[POOL_DEVICES];
    {
        std::lock_guard<std::mutex> lk(P.mu);
        for (int d = 0; d < POOL_DEVICES; ++d) {
            all[d].swap(P.idle[d]);
            P.idle_bytes[d] = 0;
        }
    }
    int prev = -1;
    (void)hipGetDevice(&prev);
    for (int d = 0; d < POOL_DEVICES; ++d) {
        if (all[d].empty() || hipSetDevice(d) != hipSuccess) continue;
        for (const LeafSet& s : all[d]) free_set(s);
    }
    if (prev >= 0) (void)hipSetDevice(prev);
}

// The host topology of a volume, built from its leaf origins and active tiles
// the first time an operation needs it.
static std::vector<std::array<int32_t, 3>> leaf_origins(bs_ctx* ctx, const bs_volume* v) {
    std::vector<int32_t> o4(4 * v->n_leaves);
    if (v->n_leaves) {
        BS_HIP(hipMemcpyAsync(o4.data(), v->origins, sizeof(int32_t) * o4.size(), hipMemcpyDeviceToHost,
                              ctx->stream));
        BS_HIP(hipStreamSynchronize(ctx->stream));
    }
    std::vector<std::array<int32_t, 3>> origins(v->n_leaves);
    for (uint64_t i = 0; i < v->n_leaves; ++i) origins[i] = {o4[4 * i], o4[4 * i + 1], o4[4 * i + 2]};
    return origins;
}

Topology& ensure_topology(bs_ctx* ctx, bs_volume* v) {
    if (v->topo) return *v->topo;
    v->topo = Topology::from_leaves(leaf_origins(ctx, v), v->tiles);
    BS_REQUIRE(v->topo->leaves.size() == v->n_leaves, BS_ERR_INTERNAL, "volume has two leaves at one origin");
    return *v->topo;
}

// ensure_topology of two volumes, the second built on another host thread
static void ensure_topologies(bs_ctx* ctx, bs_volume* a, bs_volume* b) {
    std::vector<std::array<int32_t, 3>> oa, ob;
    if (!a->topo) oa = leaf_origins(ctx, a);
    if (!b->topo) ob = leaf_origins(ctx, b);
    auto build_b = std::async(std::launch::async, [&] {
        if (!b->topo) b->topo = Topology::from_leaves(ob, b->tiles);
    });
    if (!a->topo) a->topo = Topology::from_leaves(oa, a->tiles);
    build_b.get();
    BS_REQUIRE(a->topo->leaves.size() == a->n_leaves && b->topo->leaves.size() == b->n_leaves, BS_ERR_INTERNAL,
               "volume has two leaves at one origin");
}

namespace {

// ------------------------------------------------------------------ flood fill
// leaf_node/flood_fill.rs:12-50.  The scan line keeps three running signs:
// i over x (entries (x,0,0)), j over y (entries (x,y,0)), k over z.  The value
// of each is the sign of the last active entry of its line at or before the
// position, else the running sign one level up; the seed is the sign of the
// first active entry in offset order.  Lane (x, y) owns z-row (x, y, 0..7), so
// i and j come from two ballots and k is a serial walk along the row.
// Inactive entries become +-FLT_MAX with the running sign and stay inactive.
// flags[leaf]: bit 0 first value negative, bit 1 last value negative (after
// the fill, leaf_node/flood_fill.rs:62-70), bit 2 mask empty.
__global__ __launch_bounds__(256) void k_leaf_flood_fill(uint64_t n_leaves, const uint64_t* __restrict__ masks,
                                                          float* __restrict__ values, uint8_t* __restrict__ flags) {
    const uint64_t leaf = (uint64_t)blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x >> 6);
    if (leaf >= n_leaves) return;  // uniform per wavefront
    const int lane = threadIdx.x & 63;
    const int x = lane >> 3, y = lane & 7;
    const uint32_t row = (uint32_t)(masks[leaf * 8 + x] >> (8 * y)) & 0xFFu;
    float4* vp = reinterpret_cast<float4*>(values + leaf * 512 + (uint64_t)lane * 8);
    float4 q0 = vp[0], q1 = vp[1];
    float v[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};

    const uint64_t any_rows = __ballot(row != 0);
    const uint64_t full_rows = __ballot(row == 0xFFu);
    if (any_rows != 0 && full_rows != ~0ull) {
        // seed: first active entry in offset order = lowest lane with a row, lowest z in it
        const int seed_lane = __builtin_ctzll(any_rows);
        const int my_first = row ? (int)(__float_as_uint(v[__builtin_ctz(row)]) >> 31) : 0;
        const int seed = __shfl(my_first, seed_lane);
        const int z0_neg = (int)(__float_as_uint(v[0]) >> 31);
        const uint64_t x00 = __ballot(y == 0 && (row & 1u));  // active (x, 0, 0)
        const uint64_t xy0 = __ballot((row & 1u) != 0);       // active (x, y, 0)
        const int src_i = last_at_or_below(x00, x * 8);
        int src_j = last_at_or_below(xy0, lane);
        if (src_j < x * 8) src_j = -1;  // the j line restarts at every x
        const int from_i = __shfl(z0_neg, src_i < 0 ? lane : src_i);
        const int from_j = __shfl(z0_neg, src_j < 0 ? lane : src_j);
        const int i_sign = src_i < 0 ? seed : from_i;
        int k = src_j < 0 ? i_sign : from_j;
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            if (row & (1u << z)) k = (int)(__float_as_uint(v[z]) >> 31);
            else v[z] = k ? -FAR : FAR;
        }
        vp[0] = make_float4(v[0], v[1], v[2], v[3]);
        vp[1] = make_float4(v[4], v[5], v[6], v[7]);
    }
    const int first_neg = __shfl((int)(__float_as_uint(v[0]) >> 31), 0);
    const int last_neg = __shfl((int)(__float_as_uint(v[7]) >> 31), 63);
    if (lane == 0) flags[leaf] = (uint8_t)(first_neg | (last_neg << 1) | ((any_rows == 0) << 2));
}

// ------------------------------------------------------------------ CSG assembly
// One output leaf per wavefront.  ref = (a, b, op, negate): leaf a of operand
// A, leaf b of operand B, and the combination of leaf_node/csg.rs:17-45 over
// all 512 entries, active or not -- union min, subtract max(a, -b), intersect
// max, each with utils.rs partial_min / partial_max (ties and NaN take the
// second argument) -- and the OR of the masks.  A copied leaf keeps its mask.
// `negate` is flip_signs (internal_node/csg.rs:103-111) of a leaf moved from B.
__global__ __launch_bounds__(256) void k_csg_assemble(uint64_t n_out, const int4* __restrict__ refs,
                                                      const int4* __restrict__ origins_in,
                                                      const uint64_t* __restrict__ ma, const float* __restrict__ va,
                                                      const uint64_t* __restrict__ mb, const float* __restrict__ vb,
                                                      int4* __restrict__ origins, uint64_t* __restrict__ masks,
                                                      float* __restrict__ values) {
    const uint64_t leaf = (uint64_t)blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x >> 6);
    if (leaf >= n_out) return;
    const int lane = threadIdx.x & 63;
    const int4 r = refs[leaf];
    const int op = r.z & 0xFF;
    const bool negate = (r.z >> 8) & 1;
    const float4* pa = reinterpret_cast<const float4*>(va + (uint64_t)(uint32_t)r.x * 512 + lane * 8);
    const float4* pb = reinterpret_cast<const float4*>(vb + (uint64_t)(uint32_t)r.y * 512 + lane * 8);
    float o[8];
    uint64_t m = 0;
    if (op == LEAF_COPY_A || op == LEAF_COPY_B) {
        const float4* p = op == LEAF_COPY_A ? pa : pb;
        const float4 q0 = p[0], q1 = p[1];
        o[0] = q0.x, o[1] = q0.y, o[2] = q0.z, o[3] = q0.w, o[4] = q1.x, o[5] = q1.y, o[6] = q1.z, o[7] = q1.w;
        if (lane < 8) m = (op == LEAF_COPY_A ? ma + (uint64_t)(uint32_t)r.x * 8 : mb + (uint64_t)(uint32_t)r.y * 8)[lane];
    } else {
        const float4 a0 = pa[0], a1 = pa[1], b0 = pb[0], b1 = pb[1];
        const float a[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
        const float b[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            if (op == LEAF_UNION) o[z] = a[z] < b[z] ? a[z] : b[z];
            else if (op == LEAF_SUBTRACT) o[z] = a[z] > -b[z] ? a[z] : -b[z];
            else o[z] = a[z] > b[z] ? a[z] : b[z];
        }
        if (lane < 8) m = ma[(uint64_t)(uint32_t)r.x * 8 + lane] | mb[(uint64_t)(uint32_t)r.y * 8 + lane];
    }
    if (negate) {
#pragma unroll
        for (int z = 0; z < 8; ++z) o[z] = -o[z];
    }
    float4* po = reinterpret_cast<float4*>(values + leaf * 512 + (uint64_t)lane * 8);
    po[0] = make_float4(o[0], o[1], o[2], o[3]);
    po[1] = make_float4(o[4], o[5], o[6], o[7]);
    if (lane < 8) masks[leaf * 8 + lane] = m;
    if (lane == 0) origins[leaf] = origins_in[leaf];
}

// ------------------------------------------------------------------ from_fn
// Volume::from_fn (volume/mod.rs:40-72) for the functions of VolumeBuilder
// (builder.rs:21-76).  The reference visits every index of the inclusive box
// [lo, hi]^3, evaluates f at p = idx * h and inserts the value when
// |f| <= (width + 1) * h; a leaf exists when it received a value and its other
// entries keep the default 0.0.  Here one wavefront takes one 8^3 cell of the
// box; a cell whose bound puts every entry outside the band (with a margin far
// above f32 rounding) is skipped without evaluation.  Pass 1 flags cells that
// hold a value, pass 2 writes them at their exclusive-scan position: cells are
// numbered in (x, y, z) order, so the leaves come out in origin order.
struct FnGrid {
    float h, band;
    int32_t lo[3], hi[3];  // inclusive voxel index box
    int32_t c0[3];         // first cell (in units of 8 voxels)
    uint32_t nc[3];        // cells per axis
};

// sample points of cell c inside the index box, per axis [p0, p1] (double)
__device__ inline void cell_span(const FnGrid& g, const int32_t c[3], double p0[3], double p1[3]) {
    for (int k = 0; k < 3; ++k) {
        p0[k] = (double)max(c[k] * 8, g.lo[k]) * (double)g.h;
        p1[k] = (double)min(c[k] * 8 + 7, g.hi[k]) * (double)g.h;
    }
}

// f32 box helpers in the reference's order: box3.rs:115-122, :96-112
__device__ inline bool box_contains(const float mn[3], const float mx[3], float x, float y, float z) {
    return x >= mn[0] && x <= mx[0] && y >= mn[1] && y <= mx[1] && z >= mn[2] && z <= mx[2];
}
__device__ inline float box_sq_dist(const float mn[3], const float mx[3], const float p[3]) {
    float sq = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (p[i] < mn[i]) sq = __fadd_rn(sq, __fmul_rn(__fsub_rn(mn[i], p[i]), __fsub_rn(mn[i], p[i])));
        if (p[i] > mx[i]) sq = __fadd_rn(sq, __fmul_rn(__fsub_rn(p[i], mx[i]), __fsub_rn(p[i], mx[i])));
    }
    return sq;
}
// exact signed distance to a box (double): 1-Lipschitz
__device__ inline double box_sdf(const double mn[3], const double mx[3], const double p[3]) {
    double out2 = 0.0, in = 1e300;
    bool inside = true;
    for (int k = 0; k < 3; ++k) {
        const double d = p[k] < mn[k] ? mn[k] - p[k] : (p[k] > mx[k] ? p[k] - mx[k] : 0.0);
        out2 += d * d;
        inside &= d == 0.0;
        in = fmin(in, fmin(p[k] - mn[k], mx[k] - p[k]));
    }
    return inside ? -in : sqrt(out2);
}

// VolumeBuilder::sphere (builder.rs:21-30): f = |p - o| - r
struct SphereFn {
    float r, ox, oy, oz;
    double margin;
    __device__ bool culled(const FnGrid& g, const int32_t c[3]) const {
        const double o[3] = {ox, oy, oz};
        double p0[3], p1[3];
        cell_span(g, c, p0, p1);
        double dmin2 = 0.0, dmax2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double a = p0[k] - o[k], b = p1[k] - o[k];
            const double near = (a > 0.0) ? a : (b < 0.0 ? -b : 0.0);
            const double far = fmax(fabs(a), fabs(b));
            dmin2 += near * near;
            dmax2 += far * far;
        }
        const double lim = (double)g.band + margin;
        return (sqrt(dmin2) - (double)r > lim) || ((double)r - sqrt(dmax2) > lim);
    }
    // f32 in the reference's order: d = p - o, |d| = sqrt((dx*dx + dy*dy) + dz*dz)
    __device__ float value(const FnGrid& g, int32_t xi, int32_t yi, int32_t zi) const {
        const float dx = __fsub_rn(__fmul_rn((float)xi, g.h), ox);
        const float dy = __fsub_rn(__fmul_rn((float)yi, g.h), oy);
        const float dz = __fsub_rn(__fmul_rn((float)zi, g.h), oz);
        const float n2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
        return __fsub_rn(sqrt_rn(n2), r);
    }
};

// VolumeBuilder::cuboid (builder.rs:32-51): inside, minus the f32::min chain
// of the six face distances; outside, the distance to the box
struct CuboidFn {
    float mn[3], mx[3];
    double margin;
    __device__ bool culled(const FnGrid& g, const int32_t c[3]) const {
        double p0[3], p1[3], ctr[3], bmn[3], bmx[3], r2 = 0.0;
        cell_span(g, c, p0, p1);
        for (int k = 0; k < 3; ++k) {
            ctr[k] = 0.5 * (p0[k] + p1[k]);
            r2 += 0.25 * (p1[k] - p0[k]) * (p1[k] - p0[k]);
            bmn[k] = mn[k];
            bmx[k] = mx[k];
        }
        return fabs(box_sdf(bmn, bmx, ctr)) - sqrt(r2) > (double)g.band + margin;
    }
    __device__ float value(const FnGrid& g, int32_t xi, int32_t yi, int32_t zi) const {
        const float p[3] = {__fmul_rn((float)xi, g.h), __fmul_rn((float)yi, g.h), __fmul_rn((float)zi, g.h)};
        if (box_contains(mn, mx, p[0], p[1], p[2])) {
            float m = fminf(__fsub_rn(p[0], mn[0]), __fsub_rn(mx[0], p[0]));
            m = fminf(m, __fsub_rn(p[1], mn[1]));
            m = fminf(m, __fsub_rn(mx[1], p[1]));
            m = fminf(m, __fsub_rn(p[2], mn[2]));
            m = fminf(m, __fsub_rn(mx[2], p[2]));
            return -m;
        }
        return sqrt_rn(box_sq_dist(mn, mx, p));
    }
};

// VolumeBuilder::iwp (builder.rs:53-76): inside [mn, mx] the IWP surface
// v = -(cx + cy + cz - 0.51 (cx cy + cy cz + cz cx) - 1) times the cell size,
// c* = cos(p * (1 / cell)); outside it the distance to the box grown by one
// voxel (zero in that one-voxel shell).  cos is correctly rounded (through
// double); the reference's is the platform libm's cosf.
struct IwpFn {
    float mn[3], mx[3], imn[3], imx[3], inv, cell;
    double margin;
    __device__ bool culled(const FnGrid& g, const int32_t c[3]) const {
        double p0[3], p1[3], ctr[3], r2 = 0.0;
        cell_span(g, c, p0, p1);
        for (int k = 0; k < 3; ++k) {
            // only cells well inside the sampling box: the function jumps at its faces
            if (p0[k] < (double)mn[k] + margin || p1[k] > (double)mx[k] - margin) return false;
            ctr[k] = 0.5 * (p0[k] + p1[k]);
            r2 += 0.25 * (p1[k] - p0[k]) * (p1[k] - p0[k]);
        }
        const double ic = 1.0 / (double)cell;
        const double a = cos(ctr[0] * ic), b = cos(ctr[1] * ic), d = cos(ctr[2] * ic);
        const double v = -(a + b + d - 0.51 * (a * b + b * d + d * a) - 1.0) * (double)cell;
        // |grad v| <= sqrt(3) * (1 + 2 * 0.51) < 3.5
        return fabs(v) - 3.5 * sqrt(r2) > (double)g.band + margin;
    }
    __device__ float value(const FnGrid& g, int32_t xi, int32_t yi, int32_t zi) const {
        const float p[3] = {__fmul_rn((float)xi, g.h), __fmul_rn((float)yi, g.h), __fmul_rn((float)zi, g.h)};
        if (!box_contains(mn, mx, p[0], p[1], p[2])) return sqrt_rn(box_sq_dist(imn, imx, p));
        const float cx = (float)cos((double)__fmul_rn(p[0], inv));
        const float cy = (float)cos((double)__fmul_rn(p[1], inv));
        const float cz = (float)cos((double)__fmul_rn(p[2], inv));
        const float s = __fadd_rn(__fadd_rn(cx, cy), cz);
        const float q = __fadd_rn(__fadd_rn(__fmul_rn(cx, cy), __fmul_rn(cy, cz)), __fmul_rn(cz, cx));
        const float v = -__fsub_rn(__fsub_rn(s, __fmul_rn(0.51f, q)), 1.0f);
        return __fmul_rn(v, cell);
    }
};

template <class F, bool WRITE>
__global__ __launch_bounds__(256) void k_from_fn(FnGrid g, F f, uint64_t n_cells, uint32_t* __restrict__ cell_flag,
                                                 const uint32_t* __restrict__ cell_off, int4* __restrict__ origins,
                                                 uint64_t* __restrict__ masks, float* __restrict__ values) {
    const uint64_t cell = (uint64_t)blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x >> 6);
    if (cell >= n_cells) return;
    if (WRITE && !cell_flag[cell]) return;
    const uint32_t iz = (uint32_t)(cell % g.nc[2]);
    const uint32_t iy = (uint32_t)((cell / g.nc[2]) % g.nc[1]);
    const uint32_t ix = (uint32_t)(cell / ((uint64_t)g.nc[2] * g.nc[1]));
    const int32_t c[3] = {g.c0[0] + (int32_t)ix, g.c0[1] + (int32_t)iy, g.c0[2] + (int32_t)iz};
    const int lane = threadIdx.x & 63;
    if (!WRITE && f.culled(g, c)) {
        if (lane == 0) cell_flag[cell] = 0;
        return;
    }
    const int x = lane >> 3, y = lane & 7;
    const int32_t xi = c[0] * 8 + x, yi = c[1] * 8 + y;
    const bool xy_in = xi >= g.lo[0] && xi <= g.hi[0] && yi >= g.lo[1] && yi <= g.hi[1];
    uint32_t row = 0;
    float v[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        const int32_t zi = c[2] * 8 + z;
        v[z] = 0.0f;
        if (xy_in && zi >= g.lo[2] && zi <= g.hi[2]) {
            const float fv = f.value(g, xi, yi, zi);
            if (!(fabsf(fv) > g.band)) {
                row |= 1u << z;
                v[z] = fv;
            }
        }
    }
    if (!WRITE) {
        const uint64_t any = __ballot(row != 0);
        if (lane == 0) cell_flag[cell] = any != 0;
        return;
    }
    const uint64_t out = cell_off[cell];
    float4* po = reinterpret_cast<float4*>(values + out * 512 + (uint64_t)lane * 8);
    po[0] = make_float4(v[0], v[1], v[2], v[3]);
    po[1] = make_float4(v[4], v[5], v[6], v[7]);
    const uint64_t w = gather_word(row, y);
    if (y == 0) masks[out * 8 + x] = w;
    if (lane == 0) origins[out] = make_int4(c[0] * 8, c[1] * 8, c[2] * 8, 0);
}

// Sign bit of entry (leaf, offset) for each query q = leaf * 512 + offset.
__global__ void k_gather_signs(uint64_t n, const uint64_t* __restrict__ q, const float* __restrict__ values,
                               uint8_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint8_t)(__float_as_uint(values[q[i]]) >> 31);
}

inline int32_t floor_div8(int64_t v) { return (int32_t)(v >= 0 ? v / 8 : -((-v + 7) / 8)); }

// Runs the leaf flood fill of one volume on the device (values in place) and
// records each leaf's first/last sign and emptiness in its topology.  The
// caller synchronises before reading `flags_host`.
void launch_leaf_flood_fill(bs_ctx* ctx, bs_volume* v, int slot, std::vector<uint8_t>& flags_host) {
    flags_host.assign(v->n_leaves, 0);
    if (!v->n_leaves) return;
    uint8_t* d_flags = ctx->scratch.as<uint8_t>(slot, v->n_leaves);
    k_leaf_flood_fill<<<div_up(v->n_leaves, WAVES_PER_BLOCK), 256, 0, ctx->stream>>>(v->n_leaves, v->masks, v->values,
                                                                                   d_flags);
    BS_HIP(hipGetLastError());
    BS_HIP(hipMemcpyAsync(flags_host.data(), d_flags, v->n_leaves, hipMemcpyDeviceToHost, ctx->stream));
}

// flags[i] of leaf i (the store's segments in order), in parallel chunks
void apply_flags(Topology& t, const std::vector<uint8_t>& flags) {
    size_t base = 0;
    for (auto& sg : t.leaves.seg) {
        const size_t n = std::min(sg.size(), flags.size() - std::min(base, flags.size()));
        constexpr size_t CH = 16384;
        host_parallel_for((n + CH - 1) / CH, [&, base](size_t c) {
            const size_t e = std::min(n, (c + 1) * CH);
            for (size_t i = c * CH; i < e; ++i) {
                LeafRef& r = sg[i];
                const uint8_t f = flags[base + i];
                r.first_neg = f & 1;
                r.last_neg = (f >> 1) & 1;
                r.empty = (f >> 2) & 1;
            }
        });
        base += sg.size();
    }
}

}  // namespace

// FloodFill::flood_fill of a whole grid (root_node/flood_fill.rs:8-42).
void volume_flood_fill(bs_ctx* ctx, bs_volume* v) {
    Topology& t = ensure_topology(ctx, v);
    std::vector<uint8_t> flags;
    launch_leaf_flood_fill(ctx, v, S_FLAGS_A, flags);
    BS_HIP(hipStreamSynchronize(ctx->stream));
    apply_flags(t, flags);
    t.flood_fill_upper();
}

// sign_at for a batch of voxel indices (root_node/flood_fill.rs:56-63): tile
// and root answers come from the host topology, leaf answers from the sign
// bits of the device values.
void volume_signs_at(bs_ctx* ctx, bs_volume* v, const int32_t* idx, uint64_t n, uint8_t* out_neg) {
    const Topology& t = ensure_topology(ctx, v);
    std::vector<uint64_t> q;
    std::vector<uint64_t> where;
    for (uint64_t i = 0; i < n; ++i) {
        const int32_t x = idx[3 * i], y = idx[3 * i + 1], z = idx[3 * i + 2];
        bool neg = false;
        const int32_t leaf = t.locate(x, y, z, &neg);
        if (leaf < 0) {
            out_neg[i] = neg;
            continue;
        }
        BS_REQUIRE((uint64_t)leaf < v->n_leaves, BS_ERR_INTERNAL, "topology refers to a missing leaf");
        q.push_back((uint64_t)leaf * 512 + (uint64_t)(((x & 7) << 6) | ((y & 7) << 3) | (z & 7)));
        where.push_back(i);
    }
    if (q.empty()) return;
    uint64_t* d_q = ctx->scratch.as<uint64_t>(S_MISC0, q.size());
    uint8_t* d_o = ctx->scratch.as<uint8_t>(S_MISC1, q.size());
    std::vector<uint8_t> got(q.size());
    BS_HIP(hipMemcpyAsync(d_q, q.data(), sizeof(uint64_t) * q.size(), hipMemcpyHostToDevice, ctx->stream));
    k_gather_signs<<<div_up(q.size(), 256), 256, 0, ctx->stream>>>(q.size(), d_q, v->values, d_o);
    BS_HIP(hipGetLastError());
    BS_HIP(hipMemcpyAsync(got.data(), d_o, q.size(), hipMemcpyDeviceToHost, ctx->stream));
    BS_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t k = 0; k < q.size(); ++k) out_neg[where[k]] = got[k];
}

// Volume::union / subtract / intersect (volume/mod.rs:74-93): both operands
// are flood filled, then the tree CSG runs with `a` as self.  `a` receives the
// result; `b`'s leaves are consumed (the caller frees its handle).
void volume_csg(bs_ctx* ctx, bs_volume* a, bs_volume* b, CsgOp op) {
    BS_REQUIRE(a != b, BS_ERR_INVALID_ARG, "CSG operands must be two distinct volumes");
    BS_REQUIRE(a->device == ctx->device && b->device == ctx->device, BS_ERR_INVALID_ARG,
               "CSG operands must live on the context's device");
    const auto t_wall = std::chrono::steady_clock::now();
    bs_csg_stats cs{};
    cs.n_a = a->n_leaves;
    cs.n_b = b->n_leaves;
    auto lap = [last = std::chrono::steady_clock::now()]() mutable {
        const auto now = std::chrono::steady_clock::now();
        const double ms = std::chrono::duration<double, std::milli>(now - last).count();
        last = now;
        return ms;
    };
    ensure_topologies(ctx, a, b);
    Topology& ta = *a->topo;
    Topology& tb = *b->topo;
    cs.ms_topology = lap();
    std::vector<uint8_t> fa, fb;
    EventPair ev_flood, ev_asm;
    ev_flood.start(ctx->stream);
    launch_leaf_flood_fill(ctx, a, S_FLAGS_A, fa);
    launch_leaf_flood_fill(ctx, b, S_FLAGS_B, fb);
    ev_flood.stop(ctx->stream);
    BS_HIP(hipStreamSynchronize(ctx->stream));
    cs.ms_flood = ev_flood.ms();
    lap();
    auto t_plan = std::chrono::steady_clock::now();
    auto split = [&t_plan]() {
        const auto now = std::chrono::steady_clock::now();
        const double ms = std::chrono::duration<double, std::milli>(now - t_plan).count();
        t_plan = now;
        return ms;
    };
    {
        auto fill_b = std::async(std::launch::async, [&] {
            apply_flags(tb, fb);
            tb.flood_fill_upper();
        });
        apply_flags(ta, fa);
        ta.flood_fill_upper();
        fill_b.get();
    }
    cs.ms_plan_fill = split();

    std::unique_ptr<Topology> tb_owned = std::move(b->topo);
    ta.csg(std::move(*tb_owned), op);
    tb_owned.reset();

    std::vector<int32_t> refs;
    std::vector<TileRec> tiles;
    ta.collect(refs, tiles);
    cs.ms_plan_tree = split();
    // output leaves in lexicographic origin order, as every volume here keeps
    // them.  Every result leaf is a leaf of an operand (copied, negated or
    // combined with the other operand's leaf at the same origin), and each
    // operand keeps its leaves in that order: the kept leaves of `a` (paired
    // ones included) and the copied leaves of `b` are two sorted lists, and
    // one merge orders the result.
    const size_t n_refs = refs.size();
    const size_t T = std::max<size_t>(1, std::min(host_threads(), (n_refs + 16383) / 16384));
    const size_t chunk = (n_refs + T - 1) / T;
    auto chunk_end = [&](size_t c) { return std::min(n_refs, (c + 1) * chunk); };
    {
        std::vector<int32_t> ra(a->n_leaves, -1), rb(b->n_leaves, -1);
        std::atomic<int> bad_ref{0};
        host_parallel_for(T, [&](size_t c) {
            for (size_t i = c * chunk; i < chunk_end(c); ++i) {
                const LeafRef& u = ta.leaves[refs[i]];
                if (u.op == LEAF_COPY_B) {
                    if (u.b < 0 || (uint64_t)u.b >= b->n_leaves) bad_ref.fetch_or(2);
                    else rb[u.b] = refs[i];
                } else {
                    if (u.a < 0 || (uint64_t)u.a >= a->n_leaves) bad_ref.fetch_or(1);
                    else ra[u.a] = refs[i];
                }
            }
        });
        BS_REQUIRE(!(bad_ref.load() & 1), BS_ERR_INTERNAL, "CSG plan refers to a missing leaf of the first operand");
        BS_REQUIRE(!(bad_ref.load() & 2), BS_ERR_INTERNAL, "CSG plan refers to a missing leaf of the second operand");
        struct Item {
            int32_t x, y, z, ref;
        };
        auto gather = [&](const std::vector<int32_t>& r) {
            std::vector<Item> out;
            out.reserve(r.size());
            for (int32_t k : r)
                if (k >= 0) {
                    const LeafRef& u = ta.leaves[k];
                    out.push_back({u.x, u.y, u.z, k});
                }
            return out;
        };
        std::vector<Item> la, lb;
        {
            auto fb = std::async(std::launch::async, [&] { lb = gather(rb); });
            la = gather(ra);
            fb.get();
        }
        BS_REQUIRE(la.size() + lb.size() == n_refs, BS_ERR_INTERNAL, "CSG plan lists an operand leaf twice");
        auto less = [](const Item& p, const Item& q) {
            return p.x != q.x ? p.x < q.x : (p.y != q.y ? p.y < q.y : p.z < q.z);
        };
        size_t ia = 0, ib = 0, o = 0;
        while (ia < la.size() && ib < lb.size()) refs[o++] = less(lb[ib], la[ia]) ? lb[ib++].ref : la[ia++].ref;
        while (ia < la.size()) refs[o++] = la[ia++].ref;
        while (ib < lb.size()) refs[o++] = lb[ib++].ref;
    }
    const uint64_t n_out = refs.size();
    std::vector<int32_t> new_index(ta.leaves.size(), -1);
    std::vector<int4> host(2 * n_out);
    std::atomic<int> bad{0};
    std::vector<uint64_t> paired(T, 0);
    host_parallel_for(T, [&](size_t c) {
        for (size_t i = c * chunk; i < chunk_end(c); ++i) {
            const LeafRef& r = ta.leaves[refs[i]];
            new_index[refs[i]] = (int32_t)i;
            if (!(r.op == LEAF_COPY_B || (r.a >= 0 && (uint64_t)r.a < a->n_leaves))) bad.fetch_or(1);
            if (!(r.op == LEAF_COPY_A || (r.b >= 0 && (uint64_t)r.b < b->n_leaves))) bad.fetch_or(2);
            host[i] = make_int4(std::max(r.a, 0), std::max(r.b, 0), (int)r.op | ((int)r.neg << 8), 0);
            paired[c] += r.op == LEAF_UNION || r.op == LEAF_SUBTRACT || r.op == LEAF_INTERSECT;
            host[n_out + i] = make_int4(r.x, r.y, r.z, 0);
        }
    });
    BS_REQUIRE(!(bad.load() & 1), BS_ERR_INTERNAL, "CSG plan refers to a missing leaf of the first operand");
    BS_REQUIRE(!(bad.load() & 2), BS_ERR_INTERNAL, "CSG plan refers to a missing leaf of the second operand");
    for (uint64_t p : paired) cs.n_paired += p;

    cs.ms_plan_order = split();
    cs.ms_plan = lap();
    bs_volume out;
    out.device = a->device;
    try {
        alloc_leaves(&out, n_out);
        cs.ms_memory = lap();
        ev_asm.start(ctx->stream);
        if (n_out) {
            int4* d_host = ctx->scratch.as<int4>(S_CSG_PLAN, 2 * n_out);
            BS_HIP(hipMemcpyAsync(d_host, host.data(), sizeof(int4) * host.size(), hipMemcpyHostToDevice, ctx->stream));
            k_csg_assemble<<<div_up(n_out, WAVES_PER_BLOCK), 256, 0, ctx->stream>>>(
                n_out, d_host, d_host + n_out, a->masks, a->values, b->masks, b->values,
                reinterpret_cast<int4*>(out.origins), out.masks, out.values);
            BS_HIP(hipGetLastError());
        }
        ev_asm.stop(ctx->stream);
        BS_HIP(hipStreamSynchronize(ctx->stream));
        cs.ms_assemble = ev_asm.ms();
    } catch (...) {
        free_leaves(&out, ctx->stream);
        throw;
    }
    lap();
    free_leaves(a);
    free_leaves(b);
    cs.ms_memory += lap();
    a->n_leaves = out.n_leaves;
    a->origins = out.origins;
    a->masks = out.masks;
    a->values = out.values;
    out.origins = nullptr;
    out.masks = nullptr;
    out.values = nullptr;
    a->tiles = std::move(tiles);
    a->topo = ta.compact(new_index);
    cs.ms_compact = lap();
    cs.n_out = n_out;
    cs.ms_wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall).count();
    ctx->last_csg = cs;
}

// An empty grid receiving insert(idx[i], values[i]) for i = 0..n-1 in order
// (TreeNode::insert, leaf_node/tree_node.rs:46-50): a leaf exists once any of
// its voxels was inserted, its other entries hold the default 0.0, and a later
// insert of the same voxel overwrites an earlier one.  The leaves are packed
// on the host (this is an import of caller data) and uploaded once.
void volume_from_voxels(bs_ctx* ctx, const int32_t* idx, const float* values, uint64_t n, float h, bs_volume** out) {
    BS_REQUIRE(std::isfinite(h) && h > 0.0f, BS_ERR_INVALID_ARG, "voxel size must be finite and positive");
    std::vector<uint64_t> keys(n);
    for (uint64_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k)
            BS_REQUIRE(idx[3 * i + k] >= -(1 << 23) + 16 && idx[3 * i + k] < (1 << 23) - 16, BS_ERR_INVALID_ARG,
                       "voxel index outside +-(2^23 - 16)");
        keys[i] = leaf_key(idx[3 * i] >> 3, idx[3 * i + 1] >> 3, idx[3 * i + 2] >> 3);
    }
    std::vector<uint64_t> uniq(keys);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    const uint64_t nl = uniq.size();
    std::vector<int32_t> o4(4 * nl);
    std::vector<uint64_t> m(8 * nl, 0);
    std::vector<float> v(512 * nl, 0.0f);
    for (uint64_t l = 0; l < nl; ++l) {
        int lx, ly, lz;
        leaf_key_decode(uniq[l], lx, ly, lz);
        o4[4 * l] = lx * 8;
        o4[4 * l + 1] = ly * 8;
        o4[4 * l + 2] = lz * 8;
    }
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t l = (uint64_t)(std::lower_bound(uniq.begin(), uniq.end(), keys[i]) - uniq.begin());
        const int off = ((idx[3 * i] & 7) << 6) | ((idx[3 * i + 1] & 7) << 3) | (idx[3 * i + 2] & 7);
        m[8 * l + (off >> 6)] |= 1ull << (off & 63);
        v[512 * l + off] = values[i];
    }
    auto* vol = new bs_volume();
    vol->device = ctx->device;
    vol->voxel_size = h;
    try {
        alloc_leaves(vol, nl);
        if (nl) {
            BS_HIP(hipMemcpyAsync(vol->origins, o4.data(), sizeof(int32_t) * o4.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
            BS_HIP(hipMemcpyAsync(vol->masks, m.data(), sizeof(uint64_t) * m.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
            BS_HIP(hipMemcpyAsync(vol->values, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
        }
        BS_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        free_leaves(vol, ctx->stream);
        delete vol;
        throw;
    }
    *out = vol;
}

// Output leaf i: rec[i] = (source leaf of the old arrays or -1, tile value
// bits for a new leaf), org[i] = origin.  A new leaf is all active.
__global__ __launch_bounds__(256) void k_densify_assemble(uint64_t n_out, const int2* __restrict__ rec,
                                                          const int4* __restrict__ org, const uint64_t* __restrict__ om,
                                                          const float* __restrict__ ov, int4* __restrict__ origins,
                                                          uint64_t* __restrict__ masks, float* __restrict__ values) {
    const uint64_t leaf = (uint64_t)blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x >> 6);
    if (leaf >= n_out) return;
    const int lane = threadIdx.x & 63;
    const int2 r = rec[leaf];
    float4* po = reinterpret_cast<float4*>(values + leaf * 512 + (uint64_t)lane * 8);
    if (r.x >= 0) {
        const float4* pi = reinterpret_cast<const float4*>(ov + (uint64_t)(uint32_t)r.x * 512 + lane * 8);
        po[0] = pi[0];
        po[1] = pi[1];
        if (lane < 8) masks[leaf * 8 + lane] = om[(uint64_t)(uint32_t)r.x * 8 + lane];
    } else {
        const float v = __int_as_float(r.y);
        po[0] = make_float4(v, v, v, v);
        po[1] = make_float4(v, v, v, v);
        if (lane < 8) masks[leaf * 8 + lane] = ~0ull;
    }
    if (lane == 0) origins[leaf] = org[leaf];
}

// Active tiles become full leaves of the tile value (an extension with no
// reference counterpart: the reference's dual contouring stops at todo!() on
// tiles, dual_contouring.rs:146).
void volume_densify_tiles(bs_ctx* ctx, bs_volume* v) {
    if (v->tiles.empty()) return;
    Topology& t = ensure_topology(ctx, v);
    const uint64_t n_old = v->n_leaves;
    const std::vector<TileRec> made = t.densify_tiles();
    const uint64_t n_out = n_old + made.size();
    BS_REQUIRE(n_out < (1ull << 31), BS_ERR_INVALID_ARG, "densified volume exceeds 2^31 leaves");
    std::vector<std::pair<uint64_t, int64_t>> order(n_out);  // (leaf key, source: old index or -1-k)
    for (uint64_t i = 0; i < n_old; ++i) {
        const LeafRef& r = t.leaves[i];
        order[i] = {leaf_key(r.x >> 3, r.y >> 3, r.z >> 3), (int64_t)i};
    }
    for (size_t k = 0; k < made.size(); ++k)
        order[n_old + k] = {leaf_key(made[k].x >> 3, made[k].y >> 3, made[k].z >> 3), -1 - (int64_t)k};
    std::sort(order.begin(), order.end());
    std::vector<int2> rec(n_out);
    std::vector<int4> org(n_out);
    std::vector<int32_t> new_index(n_out);
    for (uint64_t i = 0; i < n_out; ++i) {
        const int64_t src = order[i].second;
        const uint64_t ref = src >= 0 ? (uint64_t)src : n_old + (uint64_t)(-1 - src);
        new_index[ref] = (int32_t)i;
        const LeafRef& r = t.leaves[ref];
        org[i] = make_int4(r.x, r.y, r.z, 0);
        if (src >= 0) {
            rec[i] = make_int2((int)src, 0);
        } else {
            float fv = made[(size_t)(-1 - src)].value;
            int bits;
            std::memcpy(&bits, &fv, 4);
            rec[i] = make_int2(-1, bits);
        }
    }
    bs_volume out;
    out.device = v->device;
    try {
        alloc_leaves(&out, n_out);
        int2* d_rec = reinterpret_cast<int2*>(ctx->scratch.get(S_MISC0, sizeof(int2) * n_out));
        int4* d_org = ctx->scratch.as<int4>(S_MISC1, n_out);
        BS_HIP(hipMemcpyAsync(d_rec, rec.data(), sizeof(int2) * n_out, hipMemcpyHostToDevice, ctx->stream));
        BS_HIP(hipMemcpyAsync(d_org, org.data(), sizeof(int4) * n_out, hipMemcpyHostToDevice, ctx->stream));
        k_densify_assemble<<<div_up(n_out, WAVES_PER_BLOCK), 256, 0, ctx->stream>>>(
            n_out, d_rec, d_org, v->masks, v->values, reinterpret_cast<int4*>(out.origins), out.masks, out.values);
        BS_HIP(hipGetLastError());
        BS_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        free_leaves(&out, ctx->stream);
        throw;
    }
    free_leaves(v);
    v->n_leaves = out.n_leaves;
    v->origins = out.origins;
    v->masks = out.masks;
    v->values = out.values;
    out.origins = nullptr;
    out.masks = nullptr;
    out.values = nullptr;
    v->tiles.clear();
    v->topo = t.compact(new_index);
}

// The index box of Volume::from_fn (volume/mod.rs:49-51): floor(min / h) ..=
// ceil(max / h) per axis, in f32 as the reference divides; pmax is the largest
// sample coordinate.  Returns false for an empty box.
static bool fn_grid(FnGrid& g, const float mn[3], const float mx[3], float h, int width, double& pmax) {
    g.h = h;
    g.band = (float)(width + 1) * h;
    bool empty = false;
    pmax = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double lo = std::floor((double)(mn[k] / h)), hi = std::ceil((double)(mx[k] / h));
        BS_REQUIRE(std::fabs(lo) < (double)(1 << 22) && std::fabs(hi) < (double)(1 << 22), BS_ERR_INVALID_ARG,
                   "from_fn index box exceeds +-2^22 voxels");
        g.lo[k] = (int32_t)lo;
        g.hi[k] = (int32_t)hi;
        empty |= g.lo[k] > g.hi[k];
        pmax = std::max(pmax, std::max(std::fabs(lo), std::fabs(hi)) * (double)h);
    }
    for (int k = 0; k < 3; ++k) {
        g.c0[k] = empty ? 0 : floor_div8(g.lo[k]);
        const int32_t c1 = empty ? -1 : floor_div8(g.hi[k]);
        g.nc[k] = (uint32_t)(c1 - g.c0[k] + 1);
    }
    return !empty;
}

// Both passes of k_from_fn over the cells of g.
template <class F>
static bs_volume* build_from_fn(bs_ctx* ctx, const FnGrid& g, const F& f) {
    uint64_t n_cells = (uint64_t)g.nc[0] * g.nc[1] * g.nc[2];
    BS_REQUIRE(n_cells < (1ull << 31), BS_ERR_INVALID_ARG, "from_fn box spans more than 2^31 leaf cells");
    auto* v = new bs_volume();
    v->device = ctx->device;
    v->voxel_size = g.h;
    try {
        if (n_cells) {
            uint32_t* flag = ctx->scratch.as<uint32_t>(S_MISC0, n_cells);
            uint32_t* off = ctx->scratch.as<uint32_t>(S_MISC1, n_cells);
            const unsigned grid = div_up(n_cells, WAVES_PER_BLOCK);
            k_from_fn<F, false><<<grid, 256, 0, ctx->stream>>>(g, f, n_cells, flag, nullptr, nullptr, nullptr, nullptr);
            BS_HIP(hipGetLastError());
            size_t bytes = 0;
            BS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, flag, off, (int)n_cells, ctx->stream));
            void* tmp = ctx->scratch.get(S_CUB, bytes);
            BS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, flag, off, (int)n_cells, ctx->stream));
            uint32_t tail[2] = {0, 0};
            BS_HIP(hipMemcpyAsync(&tail[0], off + n_cells - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
            BS_HIP(hipMemcpyAsync(&tail[1], flag + n_cells - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
            BS_HIP(hipStreamSynchronize(ctx->stream));
            const uint64_t n_leaves = (uint64_t)tail[0] + tail[1];
            alloc_leaves(v, n_leaves);
            if (n_leaves) {
                k_from_fn<F, true><<<grid, 256, 0, ctx->stream>>>(g, f, n_cells, flag, off,
                                                                  reinterpret_cast<int4*>(v->origins), v->masks,
                                                                  v->values);
                BS_HIP(hipGetLastError());
            }
            BS_HIP(hipStreamSynchronize(ctx->stream));
        }
    } catch (...) {
        free_leaves(v, ctx->stream);
        delete v;
        throw;
    }
    return v;
}

static bool all_finite(std::initializer_list<float> xs) {
    for (float x : xs)
        if (!std::isfinite(x)) return false;
    return true;
}

// VolumeBuilder::sphere (builder.rs:21-30) through Volume::from_fn.
void volume_sphere(bs_ctx* ctx, float radius, float cx, float cy, float cz, float h, bs_volume** out) {
    BS_REQUIRE(all_finite({radius, cx, cy, cz}), BS_ERR_INVALID_ARG, "sphere radius and centre must be finite");
    BS_REQUIRE(std::isfinite(h) && h > 0.0f, BS_ERR_INVALID_ARG, "voxel size must be finite and positive");
    // builder.rs:22-25: offset = r + band_width * h with band_width 1; min/max = o -+ offset
    const float offset = radius + 1.0f * h;
    const float c[3] = {cx, cy, cz};
    float mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        mn[k] = c[k] + (-offset);
        mx[k] = c[k] + offset;
    }
    FnGrid g{};
    double pmax = 0.0;
    fn_grid(g, mn, mx, h, 1, pmax);
    // f32 rounding of idx*h, p - o, the norm and the subtraction stays below
    // 8 ulps of the largest magnitude involved; the margin is far above it.
    const double margin = 2.0 * (double)h + 1e-5 * (pmax + std::fabs((double)cx) + std::fabs((double)cy) +
                                                    std::fabs((double)cz) + std::fabs((double)radius));
    *out = build_from_fn(ctx, g, SphereFn{radius, cx, cy, cz, margin});
}

// VolumeBuilder::cuboid (builder.rs:32-51) through Volume::from_fn.
void volume_cuboid(bs_ctx* ctx, const float mn[3], const float mx[3], float h, bs_volume** out) {
    BS_REQUIRE(all_finite({mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]}), BS_ERR_INVALID_ARG,
               "cuboid corners must be finite");
    BS_REQUIRE(std::isfinite(h) && h > 0.0f, BS_ERR_INVALID_ARG, "voxel size must be finite and positive");
    const float offset = 1.0f * h;  // band_width 1
    float gmn[3], gmx[3];
    double bmax = 0.0;
    CuboidFn f{};
    for (int k = 0; k < 3; ++k) {
        gmn[k] = mn[k] + (-offset);
        gmx[k] = mx[k] + offset;
        f.mn[k] = mn[k];
        f.mx[k] = mx[k];
        bmax = std::max({bmax, std::fabs((double)mn[k]), std::fabs((double)mx[k])});
    }
    FnGrid g{};
    double pmax = 0.0;
    fn_grid(g, gmn, gmx, h, 1, pmax);
    f.margin = 2.0 * (double)h + 1e-5 * (pmax + bmax);
    *out = build_from_fn(ctx, g, f);
}

// VolumeBuilder::iwp (builder.rs:53-76) through Volume::from_fn.
void volume_iwp(bs_ctx* ctx, const float mn[3], const float mx[3], float cell_size, float h, bs_volume** out) {
    BS_REQUIRE(all_finite({mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]}), BS_ERR_INVALID_ARG,
               "iwp corners must be finite");
    BS_REQUIRE(std::isfinite(h) && h > 0.0f, BS_ERR_INVALID_ARG, "voxel size must be finite and positive");
    BS_REQUIRE(std::isfinite(cell_size) && cell_size > 0.0f, BS_ERR_INVALID_ARG,
               "iwp cell size must be finite and positive");
    IwpFn f{};
    f.inv = 1.0f / cell_size;
    f.cell = cell_size;
    double bmax = 0.0;
    for (int k = 0; k < 3; ++k) {
        f.mn[k] = mn[k];
        f.mx[k] = mx[k];
        f.imn[k] = mn[k] + (-h);
        f.imx[k] = mx[k] + h;
        bmax = std::max({bmax, std::fabs((double)mn[k]), std::fabs((double)mx[k])});
    }
    FnGrid g{};
    double pmax = 0.0;
    fn_grid(g, f.imn, f.imx, h, 2, pmax);
    // the cos arguments x = p / cell carry ~3 ulps of |x|; v sums six terms of them
    const double xmax = (pmax + bmax) / (double)cell_size;
    f.margin = 2.0 * (double)h + (double)cell_size * (1e-5 + 1e-6 * xmax) + 1e-5 * (pmax + bmax);
    *out = build_from_fn(ctx, g, f);
}

}  // namespace bs
