// Volume::offset on the device (src/voxel/volume/mod.rs:95-108): remove_if
// |v| > 2h, fast sweeping (src/voxel/fast_sweep.rs), removal of empty leaves,
// v -= D on the active values.
//
// Parallel schedule with the reference's exact result.  A sweep in direction
// (sx, sy, sz) pops leaves in lexicographic order with each axis ascending or
// descending by its sign (fast_sweep.rs:612-684), and a leaf reads only its
// six face neighbours: the three upwind ones are already swept, the three
// downwind ones are not.  With K = sx*lx + sy*ly + sz*lz over leaf
// coordinates, face neighbours sit on planes K - 1 and K + 1, so every leaf on
// one plane can be swept at once once plane K - 1 is done.  Inside a leaf the
// x-y-z loop order has the same structure over voxel planes u + v + w
// (u, v, w the voxel coordinates counted in the sweep direction): a voxel
// reads updated values on plane - 1 and old ones on plane + 1.  So a sweep is
// a sequence of plane launches, and inside each leaf a wavefront walks the 22
// voxel planes.
//
// The leaf queue.  Every leaf in the grid at the start of a sweep is in its
// queue (sweep 1 collects every leaf; sweep s + 1 takes the leaves popped in
// sweep s), and a leaf joins when a swept upwind neighbour pushes it
// (fast_sweep.rs:186-273).  The reference also creates empty leaves at the
// other faces of every stencil; they stay empty unless pushed and are removed
// at the end, so only pushed leaves are created here.  A pushed leaf lies on
// the next plane and is swept there.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>

#include "common.hpp"
#include "leafmap.hpp"

namespace bs {

namespace {

constexpr float FS_FAR = 3.40282347e+38f;

struct OffsetState {
    uint32_t pool_count;  // leaves in the pool before the sweeps
    uint32_t fail;        // bit 0: a leaf outside the reachable box (bit 2: past the last plane,
                          // bit 3: a pool leaf off the planes, else off the map), bit 1: pool or hash full
};

struct SweepArgs {
    LeafMap map;
    int4* origins;
    uint64_t* masks;
    uint64_t* frozen;
    float* values;
    uint32_t capacity;
    int sx, sy, sz;
    float h, limit_abs;
    int neg;  // sweep sign
};

// ---------------------------------------------------------------- remove_if
// leaf_node/tree_node.rs:151-160: active entries with |v| > 2h become
// inactive.  keep[leaf] = the leaf still has an active entry.
__global__ __launch_bounds__(256) void k_remove_if(uint64_t n, const uint64_t* __restrict__ masks,
                                                    const float* __restrict__ values, float two_h,
                                                    uint64_t* __restrict__ out_masks, uint32_t* __restrict__ keep) {
    const uint64_t leaf = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (leaf >= n) return;
    const int lane = threadIdx.x & 63;
    const int w = lane >> 3;
    const uint64_t word = masks[leaf * 8 + w];
    const float4* vp = reinterpret_cast<const float4*>(values + leaf * 512 + (uint64_t)lane * 8);
    const float4 q0 = vp[0], q1 = vp[1];
    const float v[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
    uint32_t drop = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) drop |= (uint32_t)(fabsf(v[k]) > two_h) << k;
    const uint64_t bits = (word >> ((lane & 7) * 8)) & 0xFFull;
    uint64_t part = (bits & ~(uint64_t)drop) << ((lane & 7) * 8);
    part |= shfl_xor_u64(part, 1);
    part |= shfl_xor_u64(part, 2);
    part |= shfl_xor_u64(part, 4);
    if ((lane & 7) == 0) out_masks[leaf * 8 + w] = part;
    const uint64_t any = __ballot(part != 0);
    if (lane == 0) keep[leaf] = any != 0;
}

// Copies the kept leaves into the pool (frozen = mask after remove_if) and
// registers them in the leaf map.
__global__ __launch_bounds__(256) void k_pool_init(uint64_t n, const uint32_t* __restrict__ keep,
                                                    const uint32_t* __restrict__ dst, const int32_t* __restrict__ origins,
                                                    const uint64_t* __restrict__ rm_masks,
                                                    const float* __restrict__ values, SweepArgs a, uint32_t* fail) {
    const uint64_t leaf = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (leaf >= n || !keep[leaf]) return;
    const int lane = threadIdx.x & 63;
    const uint32_t s = dst[leaf];
    const float4* src = reinterpret_cast<const float4*>(values + leaf * 512 + (uint64_t)lane * 8);
    float4* out = reinterpret_cast<float4*>(a.values + (uint64_t)s * 512 + (uint64_t)lane * 8);
    out[0] = src[0];
    out[1] = src[1];
    if (lane < 8) {
        const uint64_t m = rm_masks[leaf * 8 + lane];
        a.masks[(uint64_t)s * 8 + lane] = m;
        a.frozen[(uint64_t)s * 8 + lane] = m;
    }
    if (lane == 0) {
        const int4 o = make_int4(origins[leaf * 4], origins[leaf * 4 + 1], origins[leaf * 4 + 2], 0);
        a.origins[s] = o;
        if (!a.map.put(o.x >> 3, o.y >> 3, o.z >> 3, (int32_t)s)) atomicOr(fail, 1u);
    }
}

// ---------------------------------------------------------------- plane buckets
__device__ inline int32_t leaf_plane(const int4 o, int sx, int sy, int sz) {
    return sx * (o.x >> 3) + sy * (o.y >> 3) + sz * (o.z >> 3);
}

__global__ void k_plane_hist(uint32_t n, const int4* __restrict__ origins, int sx, int sy, int sz, int32_t kmin,
                             uint32_t n_planes, uint32_t* __restrict__ hist, uint32_t* fail) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t k = leaf_plane(origins[i], sx, sy, sz) - kmin;
    if (k < 0 || (uint32_t)k >= n_planes) {
        atomicOr(fail, 1u | 8u);
        return;
    }
    atomicAdd(&hist[k], 1u);
}

// order[] = the pool's leaves bucketed by plane; org_order[] their origins in
// the same order (so a sweep reads a leaf's slot and origin side by side).
__global__ void k_plane_scatter(uint32_t n, const int4* __restrict__ origins, int sx, int sy, int sz, int32_t kmin,
                                uint32_t n_planes, uint32_t* __restrict__ cursor, uint32_t* __restrict__ order,
                                int4* __restrict__ org_order) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 o = origins[i];
    const int32_t k = leaf_plane(o, sx, sy, sz) - kmin;
    if (k < 0 || (uint32_t)k >= n_planes) return;
    const uint32_t q = atomicAdd(&cursor[k], 1u);
    order[q] = i;
    org_order[q] = o;
}

// ---------------------------------------------------------------- sweep
// The leaf and the six neighbour layers touching it in one padded 10^3 block
// (leaf-local -1..8 per axis; edge and corner cells are never read), so a
// stencil read is one unconditional LDS load.  st bit 0: active, bit 1: frozen.
constexpr int PB = 10;
__device__ inline int pidx(int x, int y, int z) { return ((x + 1) * PB + (y + 1)) * PB + (z + 1); }

struct LeafLds {
    float val[PB * PB * PB];
    uint8_t st[PB * PB * PB];
};

// option_min_by(at(+), at(-), cmp_abs) (utils.rs:64-73, fast_sweep.rs:137-140):
// the + value unless the - value has the strictly smaller magnitude.
__device__ inline float pick(bool hp, float p, bool hm, float m) {
    return (hp && hm) ? (fabsf(p) > fabsf(m) ? m : p) : (hp ? p : m);
}

// fast_sweep.rs:509-537 with sort3 (helpers/utils.rs:4-16), f32 throughout;
// the three candidate solutions are all computed and one is selected, which
// yields the same value as the reference's early returns.
__device__ inline float compute_distance(float a1, float a2, float a3, float h) {
    float t;
    if (a1 > a3) t = a1, a1 = a3, a3 = t;
    if (a1 > a2) t = a1, a1 = a2, a2 = t;
    if (a2 > a3) t = a2, a2 = a3, a3 = t;
    const float s1 = a1 + h;
    const float a1a2_sum = a1 + a2;
    const float h_sq = h * h;
    const float two_h_sq = h_sq + h_sq;
    const float a1a2_diff = a1 - a2;
    const float a1a2_diff_sq = a1a2_diff * a1a2_diff;
    const float s2 = (a1a2_sum + sqrt_rn(two_h_sq - a1a2_diff_sq)) * 0.5f;
    const float a1a2a3_sum = a1a2_sum + a3;
    const float three_h_sq = two_h_sq + h_sq;
    const float a1a3_diff = a1 - a3;
    const float a1a3_diff_sq = a1a3_diff * a1a3_diff;
    const float a2a3_diff = a2 - a3;
    const float a2a3_diff_sq = a2a3_diff * a2a3_diff;
    const float s3 = (a1a2a3_sum + sqrt_rn(three_h_sq - a1a2_diff_sq - a1a3_diff_sq - a2a3_diff_sq)) * (1.0f / 3.0f);
    return fabsf(s1) <= a2 ? s1 : (fabsf(s2) <= a3 ? s2 : s3);
}

// fast_sweep.rs:120-184 for the voxel at padded index o
__device__ inline void sweep_voxel(LeafLds& L, const SweepArgs& a, int o) {
    const uint8_t so = L.st[o];
    const uint8_t sxp = L.st[o + PB * PB], sxm = L.st[o - PB * PB], syp = L.st[o + PB], sym = L.st[o - PB],
                  szp = L.st[o + 1], szm = L.st[o - 1];
    const float vxp = L.val[o + PB * PB], vxm = L.val[o - PB * PB], vyp = L.val[o + PB], vym = L.val[o - PB],
                vzp = L.val[o + 1], vzm = L.val[o - 1], vo = L.val[o];
    const bool hx = (sxp | sxm) & 1, hy = (syp | sym) & 1, hz = (szp | szm) & 1;
    const float dx = pick(sxp & 1, vxp, sxm & 1, vxm);
    const float dy = pick(syp & 1, vyp, sym & 1, vym);
    const float dz = pick(szp & 1, vzp, szm & 1, vzm);
    const float first = hx ? dx : (hy ? dy : dz);
    const bool go = !(so & 2) && (hx || hy || hz) && (int)(__float_as_uint(first) >> 31) == a.neg;
    if (!go) return;  // a wave with no such voxel skips the solve
    const float d_abs = compute_distance(fabsf(hx ? dx : FS_FAR), fabsf(hy ? dy : FS_FAR), fabsf(hz ? dz : FS_FAR), a.h);
    const float d_old = (so & 1) ? vo : FS_FAR;
    if (!(d_abs > a.limit_abs) && d_abs < fabsf(d_old)) {
        L.val[o] = a.neg ? -fabsf(d_abs) : fabsf(d_abs);  // copysign(d_abs, sweep sign)
        L.st[o] = so | 1;
    }
}

// The leaf's own values and state (this lane's z-row; lane = (lx, ly)) and,
// in lanes 0..5, its six face-neighbour slots, as registers.  stage_issue
// only issues the loads (dense map: the raw cell word, resolved when stored),
// so the leaf's data and its neighbour lookups travel in one round trip.
struct LeafStage {
    float4 v0, v1;
    uint64_t mw, fw;
    int32_t nb;  // lanes 0..5: neighbour slot or map cell word
};

__device__ inline void stage_issue(const SweepArgs& a, uint32_t slot, int4 org, int lane, LeafStage& s) {
    const int lx = lane >> 3;
    const float4* vp = reinterpret_cast<const float4*>(a.values + (uint64_t)slot * 512 + lane * 8);
    s.v0 = vp[0];
    s.v1 = vp[1];
    s.mw = a.masks[(uint64_t)slot * 8 + lx];
    s.fw = a.frozen[(uint64_t)slot * 8 + lx];
    s.nb = -1;
    if (lane < 6) {
        const int ax = lane >> 1, dir = (lane & 1) ? 1 : -1;
        const int32_t nx = (org.x >> 3) + (ax == 0 ? dir : 0), ny = (org.y >> 3) + (ax == 1 ? dir : 0),
                      nz = (org.z >> 3) + (ax == 2 ? dir : 0);
        if (a.map.dense) {
            const uint32_t ix = (uint32_t)(nx - a.map.x0), iy = (uint32_t)(ny - a.map.y0), iz = (uint32_t)(nz - a.map.z0);
            if (ix < a.map.nx && iy < a.map.ny && iz < a.map.nz)
                s.nb = a.map.cells[((uint64_t)ix * a.map.ny + iy) * a.map.nz + iz];
        } else {
            s.nb = a.map.find(nx, ny, nz);
        }
    }
}

__device__ inline void stage_store(LeafLds& L, const LeafStage& s, int lane) {
    const int lx = lane >> 3, ly = lane & 7;
    const uint64_t mw = s.mw >> (ly * 8), fw = s.fw >> (ly * 8);
    const float v[8] = {s.v0.x, s.v0.y, s.v0.z, s.v0.w, s.v1.x, s.v1.y, s.v1.z, s.v1.w};
    const int base = pidx(lx, ly, 0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        L.val[base + k] = v[k];
        L.st[base + k] = (uint8_t)(((mw >> k) & 1) | (((fw >> k) & 1) << 1));
    }
}

// The six neighbour layers next to the leaf, into the padding.  nb: lanes
// 0..5 hold the neighbour slots (-1 none; a cell being created, -2, is no
// leaf yet).  The slots are read from those lanes and every one of the twelve
// loads is issued unconditionally (slot 0 in place of a missing neighbour,
// its result discarded), so they travel in one round trip.
__device__ inline void stage_layers(LeafLds& L, const SweepArgs& a, int lane, int32_t nb) {
    const int i = lane >> 3, j = lane & 7;
    uint8_t on[6];
    float val[6];
    uint64_t mw[6];
    float nv[6];
    int32_t n[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        n[k] = __builtin_amdgcn_readlane(nb, k);
        const int f = (k & 1) ? 0 : 7;  // -side neighbour: its far layer 7; +side: layer 0
        const int o = k < 2 ? ((f << 6) | (i << 3) | j) : (k < 4 ? ((i << 6) | (f << 3) | j) : ((i << 6) | (j << 3) | f));
        const uint64_t ns = n[k] >= 0 ? (uint64_t)n[k] : 0;
        mw[k] = a.masks[ns * 8 + (o >> 6)];
        nv[k] = a.values[ns * 512 + o];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int f = (k & 1) ? 0 : 7;
        const int o = k < 2 ? ((f << 6) | (i << 3) | j) : (k < 4 ? ((i << 6) | (f << 3) | j) : ((i << 6) | (j << 3) | f));
        on[k] = n[k] >= 0 ? (uint8_t)((mw[k] >> (o & 63)) & 1ull) : (uint8_t)0;
        val[k] = on[k] ? nv[k] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int g = (k & 1) ? 8 : -1;  // the padding layer this maps to
        const int po = k < 2 ? pidx(g, i, j) : (k < 4 ? pidx(i, g, j) : pidx(i, j, g));
        L.st[po] = on[k];
        L.val[po] = val[k];
    }
}

// The staged leaf swept (the wavefront over its 22 voxel planes), written
// back, and its pushes (fast_sweep.rs:186-273): the +direction face of each
// axis; lanes 0..2 create the pushed neighbours of the three axes at once.
// A created slot is published without a fence: in this plane a lookup of the
// new leaf reads only its masks and values, which hold the empty leaf
// already, and its origin is read only in a later plane.
__device__ inline void sweep_and_finish(LeafLds& L, const SweepArgs& a, uint32_t slot, int4 org, uint32_t plane,
                                        uint32_t n_planes, uint32_t start, uint32_t* created, OffsetState* st,
                                        int lane) {
    const int lx = lane >> 3, ly = lane & 7;
    // voxel plane p holds the voxels u + v + w = p (u, v, w counted in the
    // sweep direction); they are independent of each other, so lane (u, v)
    // = (lx, ly) takes w = p - u - v when that is a voxel
    const int pu = a.sx > 0 ? lx : 7 - lx, pv = a.sy > 0 ? ly : 7 - ly;
    for (int p = 0; p < 22; ++p) {
        const int w = p - lx - ly;
        if (w >= 0 && w < 8) sweep_voxel(L, a, pidx(pu, pv, a.sz > 0 ? w : 7 - w));
        __syncthreads();
    }
    {
        const int base = pidx(lx, ly, 0);
        float4* vp = reinterpret_cast<float4*>(a.values + (uint64_t)slot * 512 + lane * 8);
        vp[0] = make_float4(L.val[base + 0], L.val[base + 1], L.val[base + 2], L.val[base + 3]);
        vp[1] = make_float4(L.val[base + 4], L.val[base + 5], L.val[base + 6], L.val[base + 7]);
        uint32_t row = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) row |= (uint32_t)(L.st[base + k] & 1) << k;
        const uint64_t w = gather_word(row, ly);
        if (ly == 0) a.masks[(uint64_t)slot * 8 + lx] = w;
    }
    bool push[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int dir = ax == 0 ? a.sx : (ax == 1 ? a.sy : a.sz);
        const int f = dir > 0 ? 7 : 0;
        const int i = lx, j = ly;
        const int o = ax == 0 ? pidx(f, i, j) : (ax == 1 ? pidx(i, f, j) : pidx(i, j, f));
        const float v = L.val[o];
        const bool q = (L.st[o] & 1) && (int)(__float_as_uint(v) >> 31) == a.neg && fabsf(v) < a.limit_abs;
        push[ax] = __ballot(q) != 0;
    }
    if (lane < 3 && (lane == 0 ? push[0] : (lane == 1 ? push[1] : push[2]))) {
        const int ax = lane;
        const int dir = ax == 0 ? a.sx : (ax == 1 ? a.sy : a.sz);
        if (plane + 1 >= n_planes) atomicOr(&st->fail, 1u | 4u);
        const int32_t nx = (org.x >> 3) + (ax == 0 ? dir : 0), ny = (org.y >> 3) + (ax == 1 ? dir : 0),
                      nz = (org.z >> 3) + (ax == 2 ? dir : 0);
        int32_t* cell = a.map.claim(nx, ny, nz, &st->fail);
        if (cell) {
            const uint32_t s = start + atomicAdd(&created[plane + 1], 1u);
            if (s < a.capacity) {
                a.origins[s] = make_int4(nx * 8, ny * 8, nz * 8, 0);
                atomicExch(cell, (int32_t)s);
            } else {
                atomicOr(&st->fail, 2u);  // pool overflow: retried with a larger pool
            }
        }
    }
}

// One leaf plane of one sweep: the leaves of bucket [b0, b1) of `order` plus
// the leaves created while the previous plane was swept.  starts[p + 1] is
// the pool size when plane p begins and created[p + 1] the leaves plane p
// creates (all on plane p + 1), so this plane sweeps the pool range
// [starts[plane], starts[plane] + created[plane]) after its bucket and
// allocates its own new leaves from starts[plane] + created[plane] on.
// One 64-lane block per leaf, looping over the plane's leaves.
__global__ __launch_bounds__(64) void k_sweep_plane(SweepArgs a, const uint32_t* __restrict__ order,
                                                     const int4* __restrict__ org_order,
                                                     const uint32_t* __restrict__ bucket_off, uint32_t plane,
                                                     uint32_t n_planes, uint32_t* starts, uint32_t* created,
                                                     OffsetState* st) {
    __shared__ LeafLds L;
    const uint32_t b0 = bucket_off[plane], b1 = bucket_off[plane + 1];
    const uint32_t n_old = b1 - b0;
    const uint32_t new_lo = starts[plane];
    const uint32_t start = min(new_lo + created[plane], a.capacity);
    const uint32_t n_new = start - new_lo;
    const int lane = threadIdx.x;
    if (blockIdx.x == 0 && lane == 0) starts[plane + 1] = start;
    for (uint32_t item = blockIdx.x; item < n_old + n_new; item += gridDim.x) {
        // a bucket leaf's slot and origin in one round trip, a pushed one's origin
        const uint32_t slot = item < n_old ? order[b0 + item] : new_lo + (item - n_old);
        const int4 org = item < n_old ? org_order[b0 + item] : a.origins[slot];
        LeafStage s;
        stage_issue(a, slot, org, lane, s);
        stage_store(L, s, lane);
        stage_layers(L, a, lane, s.nb >= 0 ? s.nb : -1);  // the padding: no overlap with stage_store
        __syncthreads();
        sweep_and_finish(L, a, slot, org, plane, n_planes, start, created, st, lane);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- output
// Sort keys of the final pool: empty leaves (fast_sweep.rs:61-64
// remove_empty_branches) sort last.
__global__ __launch_bounds__(256) void k_final_keys(uint32_t n, const int4* __restrict__ origins,
                                                     const uint64_t* __restrict__ masks, uint64_t* __restrict__ keys,
                                                     uint32_t* __restrict__ idx, uint32_t* __restrict__ n_live) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t any = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w) any |= masks[(uint64_t)i * 8 + w];
    const int4 o = origins[i];
    keys[i] = any ? leaf_key(o.x >> 3, o.y >> 3, o.z >> 3) : ~0ull;
    idx[i] = i;
    if (any) atomicAdd(n_live, 1u);
}

// Gathers the live leaves in key order and applies v -= D to active values
// (visitors/value_mut_visitor.rs:4-32 visits active values only).
__global__ __launch_bounds__(256) void k_final_gather(uint64_t n_out, const uint32_t* __restrict__ src_idx,
                                                       const int4* __restrict__ porigins,
                                                       const uint64_t* __restrict__ pmasks,
                                                       const float* __restrict__ pvalues, float distance,
                                                       int4* __restrict__ origins, uint64_t* __restrict__ masks,
                                                       float* __restrict__ values) {
    const uint64_t leaf = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (leaf >= n_out) return;
    const int lane = threadIdx.x & 63;
    const uint32_t s = src_idx[leaf];
    const uint64_t word = pmasks[(uint64_t)s * 8 + (lane >> 3)] >> ((lane & 7) * 8);
    const float4* vp = reinterpret_cast<const float4*>(pvalues + (uint64_t)s * 512 + lane * 8);
    const float4 q0 = vp[0], q1 = vp[1];
    float v[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if ((word >> k) & 1) v[k] = v[k] - distance;
    float4* op = reinterpret_cast<float4*>(values + leaf * 512 + (uint64_t)lane * 8);
    op[0] = make_float4(v[0], v[1], v[2], v[3]);
    op[1] = make_float4(v[4], v[5], v[6], v[7]);
    if (lane < 8) masks[leaf * 8 + lane] = pmasks[(uint64_t)s * 8 + lane];
    if (lane == 0) origins[leaf] = porigins[s];
}

constexpr int kDirs[8][3] = {{1, 1, 1},  {-1, 1, 1},  {1, -1, 1},  {-1, -1, 1},
                             {1, 1, -1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};  // fast_sweep.rs:36-57

template <class K, class V>
void sort_pairs_u64(bs_ctx* ctx, const K* k, K* ks, const V* v, V* vs, uint64_t n) {
    size_t bytes = 0;
    BS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k, ks, v, vs, (int)n, 0, 64, ctx->stream));
    void* tmp = ctx->scratch.get(S_CUB, bytes);
    BS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, k, ks, v, vs, (int)n, 0, 64, ctx->stream));
}

void exclusive_scan_u32(bs_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n) {
    size_t bytes = 0;
    BS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, ctx->stream));
    void* tmp = ctx->scratch.get(S_CUB, bytes);
    BS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, ctx->stream));
}

template <class T>
T read_back(bs_ctx* ctx, const T* d) {
    T h{};
    BS_HIP(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    BS_HIP(hipStreamSynchronize(ctx->stream));
    return h;
}

// Device buffers of one offset attempt (released on every exit path).
struct OffsetBuffers {
    std::vector<void*> ptrs;
    template <class T> T* alloc(uint64_t count) {
        void* p = nullptr;
        BS_HIP(hipMalloc(&p, std::max<uint64_t>(count, 1) * sizeof(T)));
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    ~OffsetBuffers() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

}  // namespace

Topology& ensure_topology(bs_ctx* ctx, bs_volume* v);

namespace {

struct OffsetResult {
    uint64_t n_out = 0;
    std::vector<uint32_t> keep;  // per input leaf
    bool retry = false;
};

// One attempt with a pool of `capacity` leaves; retry = the pool or the leaf
// map was too small (nothing of `vol` was modified).
OffsetResult offset_attempt(bs_ctx* ctx, bs_volume* vol, float distance, uint64_t capacity, bool force_hash,
                            bs_volume& out, bs_offset_stats* stats) {
    hipStream_t st = ctx->stream;
    OffsetResult res;
    const uint64_t n = vol->n_leaves;
    const float h = vol->voxel_size;
    const float two_h = h * 2.0f;  // volume/mod.rs:96
    float ext = fabsf(distance) + h + h;  // volume/mod.rs:98-99
    ext = std::copysign(ext, std::signbit(distance) ? -1.0f : 1.0f);
    const float limit_abs = fabsf(ext);
    const int neg = std::signbit(ext) ? 1 : 0;

    OffsetBuffers B;
    // remove_if and compaction of the surviving leaves
    uint64_t* rm_masks = B.alloc<uint64_t>(8 * n);
    uint32_t* keep = B.alloc<uint32_t>(n);
    uint32_t* dst = B.alloc<uint32_t>(n);
    if (n) {
        k_remove_if<<<div_up(n, 4), 256, 0, st>>>(n, vol->masks, vol->values, two_h, rm_masks, keep);
        BS_HIP(hipGetLastError());
        exclusive_scan_u32(ctx, keep, dst, n);
    }
    res.keep.assign(n, 0);
    std::vector<int32_t> o4(4 * n);
    if (n) {
        BS_HIP(hipMemcpyAsync(res.keep.data(), keep, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
        BS_HIP(hipMemcpyAsync(o4.data(), vol->origins, sizeof(int32_t) * 4 * n, hipMemcpyDeviceToHost, st));
        BS_HIP(hipStreamSynchronize(st));
    }
    uint64_t n_keep = 0;
    int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (uint64_t i = 0; i < n; ++i) {
        if (!res.keep[i]) continue;
        ++n_keep;
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], o4[4 * i + k] >> 3);
            hi[k] = std::max(hi[k], o4[4 * i + k] >> 3);
        }
    }
    if (stats) stats->n_kept = n_keep;
    if (n_keep == 0) {
        out.n_leaves = 0;
        return res;
    }
    BS_REQUIRE(capacity >= n_keep && capacity < (1ull << 31), BS_ERR_INVALID_ARG, "offset leaf pool too large");

    // Reach of the extension: a value grows by at least h/sqrt(3) per voxel
    // step (the three-term Godunov update is the smallest increment), so no
    // leaf beyond limit*sqrt(3)/h + 2 voxels of the band is ever pushed.
    const int32_t margin = (int32_t)std::ceil(((double)limit_abs / (double)h * 1.7320508075688772 + 4.0) / 8.0) + 1;
    LeafMap map{};
    uint64_t cells = 1;
    for (int k = 0; k < 3; ++k) {
        lo[k] -= margin;
        hi[k] += margin;
        cells *= (uint64_t)(hi[k] - lo[k] + 1);
    }
    map.dense = !force_hash && cells <= (1ull << 28);
    if (map.dense) {
        map.x0 = lo[0], map.y0 = lo[1], map.z0 = lo[2];
        map.nx = hi[0] - lo[0] + 1, map.ny = hi[1] - lo[1] + 1, map.nz = hi[2] - lo[2] + 1;
        map.cells = B.alloc<int32_t>(cells);
        BS_HIP(hipMemsetAsync(map.cells, 0xFF, sizeof(int32_t) * cells, st));
    } else {
        uint64_t cap = 1024;
        while (cap < 2 * capacity) cap <<= 1;
        map.cap_mask = (uint32_t)(cap - 1);
        map.keys = B.alloc<uint64_t>(cap);
        map.cells = B.alloc<int32_t>(cap);
        BS_HIP(hipMemsetAsync(map.keys, 0xFF, sizeof(uint64_t) * cap, st));
        BS_HIP(hipMemsetAsync(map.cells, 0xFF, sizeof(int32_t) * cap, st));
    }
    if (stats) stats->used_hash = !map.dense;
    if (stats) stats->n_capacity = capacity;

    SweepArgs a{};
    a.map = map;
    a.capacity = (uint32_t)capacity;
    a.origins = B.alloc<int4>(capacity);
    a.masks = B.alloc<uint64_t>(8 * capacity);
    a.frozen = B.alloc<uint64_t>(8 * capacity);
    a.values = B.alloc<float>(512 * capacity);
    a.h = h;
    a.limit_abs = limit_abs;
    a.neg = neg;
    // leaves created by the sweep start empty, with 0.0 values (Leaf::empty)
    BS_HIP(hipMemsetAsync(a.masks, 0, sizeof(uint64_t) * 8 * capacity, st));
    BS_HIP(hipMemsetAsync(a.frozen, 0, sizeof(uint64_t) * 8 * capacity, st));
    BS_HIP(hipMemsetAsync(a.values + 512 * n_keep, 0, sizeof(float) * 512 * (capacity - n_keep), st));
    OffsetState* dstate = B.alloc<OffsetState>(1);
    OffsetState init{(uint32_t)n_keep, 0};
    BS_HIP(hipMemcpyAsync(dstate, &init, sizeof(init), hipMemcpyHostToDevice, st));
    k_pool_init<<<div_up(n, 4), 256, 0, st>>>(n, keep, dst, vol->origins, rm_masks, vol->values, a, &dstate->fail);
    BS_HIP(hipGetLastError());

    // plane range over the whole map box
    uint32_t* order = B.alloc<uint32_t>(capacity);
    int4* org_order = B.alloc<int4>(capacity);

    uint64_t planes_total = 0, launches = 0, swept = 0;
    uint32_t pool = (uint32_t)n_keep;  // pool size at the start of the sweep
    for (int s = 0; s < 8; ++s) {
        a.sx = kDirs[s][0], a.sy = kDirs[s][1], a.sz = kDirs[s][2];
        const OffsetState cur = read_back(ctx, dstate);
        BS_REQUIRE(!(cur.fail & 1u), BS_ERR_INTERNAL,
                   "offset: a pushed leaf left the reachable box (flags " + std::to_string(cur.fail) + ", sweep " +
                       std::to_string(s) + ", pool " + std::to_string(pool) + ")");
        if (cur.fail) {
            res.retry = true;
            return res;
        }
        const uint32_t n_pool = pool;
        if (s > 0) swept += n_pool;  // every leaf in the pool at the end of a sweep was swept in it
        // planes of the map box in this direction
        // planes of the reachable box (the leaves plus the extension margin)
        int32_t kmin = 0, kmax = 0;
        const int sgn[3] = {a.sx, a.sy, a.sz};
        for (int k = 0; k < 3; ++k) {
            kmin += sgn[k] > 0 ? lo[k] : -hi[k];
            kmax += sgn[k] > 0 ? hi[k] : -lo[k];
        }
        const uint32_t n_planes = (uint32_t)(kmax - kmin + 1);
        uint32_t* hist = B.alloc<uint32_t>(n_planes + 1);
        uint32_t* off = B.alloc<uint32_t>(n_planes + 1);
        BS_HIP(hipMemsetAsync(hist, 0, sizeof(uint32_t) * (n_planes + 1), st));
        k_plane_hist<<<div_up(n_pool, 256), 256, 0, st>>>(n_pool, a.origins, a.sx, a.sy, a.sz, kmin, n_planes, hist,
                                                         &dstate->fail);
        BS_HIP(hipGetLastError());
        exclusive_scan_u32(ctx, hist, off, n_planes + 1);
        BS_HIP(hipMemcpyAsync(hist, off, sizeof(uint32_t) * (n_planes + 1), hipMemcpyDeviceToDevice, st));
        k_plane_scatter<<<div_up(n_pool, 256), 256, 0, st>>>(n_pool, a.origins, a.sx, a.sy, a.sz, kmin, n_planes, hist,
                                                            order, org_order);
        BS_HIP(hipGetLastError());
        // host view of the bucket sizes: the first and last plane holding a leaf
        std::vector<uint32_t> hoff(n_planes + 1);
        BS_HIP(hipMemcpyAsync(hoff.data(), off, sizeof(uint32_t) * (n_planes + 1), hipMemcpyDeviceToHost, st));
        BS_HIP(hipStreamSynchronize(st));
        uint32_t first = 0, last = 0, max_bucket = 0;
        bool any = false;
        for (uint32_t p = 0; p < n_planes; ++p) {
            const uint32_t c = hoff[p + 1] - hoff[p];
            if (!c) continue;
            if (!any) first = p;
            any = true;
            last = p;
            max_bucket = std::max(max_bucket, c);
        }
        if (!any) break;
        // starts[p + 1], created[p + 1]: pool size when plane p begins, leaves it creates
        uint32_t* starts = B.alloc<uint32_t>(n_planes + 2);
        uint32_t* created = B.alloc<uint32_t>(n_planes + 2);
        BS_HIP(hipMemsetAsync(created, 0, sizeof(uint32_t) * (n_planes + 2), st));
        BS_HIP(hipMemcpyAsync(starts + first, &pool, sizeof(uint32_t), hipMemcpyHostToDevice, st));
        uint32_t p_end = first;
        for (uint32_t p = first; p < n_planes; ++p) {
            const uint32_t bucket = hoff[p + 1] - hoff[p];
            if (p > last) {
                // past the last bucket: continue only while the previous plane created leaves
                uint32_t c = 0;
                BS_HIP(hipMemcpyAsync(&c, created + p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
                BS_HIP(hipStreamSynchronize(st));
                if (c == 0) break;
            }
            // enough blocks for the bucket and the leaves the previous plane may have created
            const uint32_t prev = p > first ? hoff[p] - hoff[p - 1] : 0;
            const unsigned grid = (unsigned)std::min<uint64_t>(std::max<uint64_t>((uint64_t)bucket + prev + 64, 64), 4096);
            k_sweep_plane<<<grid, 64, 0, st>>>(a, order, org_order, off, p, n_planes, starts, created, dstate);
            BS_HIP(hipGetLastError());
            ++planes_total;
            ++launches;
            p_end = p;
        }
        // pool size after the last plane swept
        uint32_t tail[2] = {0, 0};
        BS_HIP(hipMemcpyAsync(&tail[0], starts + p_end + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        BS_HIP(hipMemcpyAsync(&tail[1], created + p_end + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        BS_HIP(hipStreamSynchronize(st));
        pool = std::min<uint32_t>(tail[0] + tail[1], (uint32_t)capacity);
        if ((uint64_t)tail[0] + tail[1] > capacity) {
            res.retry = true;
            return res;
        }
    }
    const OffsetState fin = read_back(ctx, dstate);
    BS_REQUIRE(!(fin.fail & 1u), BS_ERR_INTERNAL,
               "offset: a pushed leaf left the reachable box (flags " + std::to_string(fin.fail) + ", after the sweeps)");
    if (fin.fail) {
        res.retry = true;
        return res;
    }
    const uint32_t n_pool = pool;
    swept += n_pool;
    if (stats) {
        stats->n_swept = swept;
        stats->n_pool = n_pool;
        stats->n_planes = planes_total;
        stats->n_launches = launches;
    }

    // empty leaves go, the rest sorted by origin, shifted by -D
    uint64_t* keys = B.alloc<uint64_t>(n_pool);
    uint64_t* keys_s = B.alloc<uint64_t>(n_pool);
    uint32_t* idx = B.alloc<uint32_t>(n_pool);
    uint32_t* idx_s = B.alloc<uint32_t>(n_pool);
    uint32_t* n_live = B.alloc<uint32_t>(1);
    BS_HIP(hipMemsetAsync(n_live, 0, sizeof(uint32_t), st));
    k_final_keys<<<div_up(n_pool, 256), 256, 0, st>>>(n_pool, a.origins, a.masks, keys, idx, n_live);
    BS_HIP(hipGetLastError());
    sort_pairs_u64(ctx, keys, keys_s, idx, idx_s, n_pool);
    const uint32_t n_out = read_back(ctx, n_live);
    alloc_leaves(&out, n_out);
    if (n_out) {
        k_final_gather<<<div_up(n_out, 4), 256, 0, st>>>(n_out, idx_s, a.origins, a.masks, a.values, distance,
                                                         reinterpret_cast<int4*>(out.origins), out.masks, out.values);
        BS_HIP(hipGetLastError());
    }
    BS_HIP(hipStreamSynchronize(st));
    res.n_out = n_out;
    return res;
}

}  // namespace

// Volume::offset (volume/mod.rs:95-108) of `vol` in place.
void volume_offset(bs_ctx* ctx, bs_volume* vol, float distance, bs_offset_stats* stats) {
    BS_REQUIRE(std::isfinite(distance), BS_ERR_INVALID_ARG, "offset distance must be finite");
    const auto t_wall = std::chrono::steady_clock::now();
    EventPair ev;
    ev.start(ctx->stream);
    BS_REQUIRE(vol->device == ctx->device, BS_ERR_INVALID_ARG, "volume lives on another device");
    Topology& topo = ensure_topology(ctx, vol);
    const float two_h = vol->voxel_size * 2.0f;
    // fast_sweep.rs:100 unwraps the leaf at every queued origin, and the queue
    // holds every active tile that survives remove_if: the reference panics.
    for (const TileRec& t : vol->tiles)
        BS_REQUIRE(fabsf(t.value) > two_h, BS_ERR_INTERNAL,
                   "offset: an active tile lies within the band; the reference panics (fast_sweep.rs:100)");
    if (stats) *stats = bs_offset_stats{};
    uint64_t capacity = std::max<uint64_t>(4 * vol->n_leaves + 4096, 8192);
    bs_volume out;
    out.device = vol->device;
    OffsetResult r;
    for (int attempt = 0;; ++attempt) {
        try {
            r = offset_attempt(ctx, vol, distance, capacity, (ctx->paths & BS_PATH_FORCE_HASH) != 0, out, stats);
        } catch (...) {
            free_leaves(&out, ctx->stream);
            throw;
        }
        if (!r.retry) break;
        BS_REQUIRE(attempt < 6, BS_ERR_INTERNAL, "offset: leaf pool kept overflowing");
        capacity *= 4;
    }
    // topology: remove_if (tiles, emptied leaves, empty nodes), the new leaves, empty nodes again
    std::vector<std::array<int32_t, 3>> out_origins(out.n_leaves);
    if (out.n_leaves) {
        std::vector<int32_t> o4(4 * out.n_leaves);
        BS_HIP(hipMemcpy(o4.data(), out.origins, sizeof(int32_t) * o4.size(), hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < out.n_leaves; ++i) out_origins[i] = {o4[4 * i], o4[4 * i + 1], o4[4 * i + 2]};
    }
    std::unique_ptr<Topology> t2 = topo.after_offset(r.keep, out_origins, two_h);
    free_leaves(vol);
    vol->n_leaves = out.n_leaves;
    vol->origins = out.origins;
    vol->masks = out.masks;
    vol->values = out.values;
    out.origins = nullptr;
    out.masks = nullptr;
    out.values = nullptr;
    vol->tiles.clear();  // remove_if deactivated every tile (checked above)
    vol->topo = std::move(t2);
    ev.stop(ctx->stream);
    if (stats) {
        stats->n_out = out.n_leaves;
        stats->ms_device = ev.ms();
        stats->ms_wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall).count();
    }
}

}  // namespace bs
