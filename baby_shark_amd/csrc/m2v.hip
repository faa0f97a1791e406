// Mesh -> narrow-band signed distance volume on one MI355X.
//
// Replaces MeshToVolume::convert (src/voxel/mesh_to_volume.rs:52-73):
//   subdivide     mesh_to_volume.rs:75-116   one thread per (triangle, row)
//   leaves        :118-168 active set        dilated boxes -> GPU hash -> compact -> sort
//   distance      :118-196 min over pairs    one workgroup per 8^3 leaf, 4^3 block per wave
//   winding       aabb_tree.rs:571-817       LBVH + order-2 dipoles, wave-coherent traversal
//   sign          mesh_to_volume.rs:238-292  wn < threshold => +|d| else -|d|
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>

#include "common.hpp"
#include "m2v_kernels.hpp"

namespace bs {

namespace {

constexpr float INV_4PI = 0.0795774715459476678844f;
constexpr int BVH_LEAF_SIZE = 4;     // a node covering <= 4 triangles is evaluated as a leaf
constexpr int WN_STACK = 128;        // per-wave traversal stack depth
constexpr int WN_BLOCK = 256;

enum Counter : int {
    C_ERR = 0, C_LEAVES, C_RECHECK, C_NEAR_INTERNAL, C_FAR, C_LEAF_VISITS, C_LEAF_TRIS, C_BOX_LO, C_BOX_HI = C_BOX_LO + 3,
    C_PAIRS = C_BOX_HI + 3, C_COUNT
};
enum ErrBits : unsigned { E_RANGE = 1, E_TOO_MANY_SUBS = 2, E_STACK = 4, E_HASH = 8, E_TREE = 16 };

// ---------------------------------------------------------------------------
// subdivision (mesh_to_volume.rs:75-116)

__device__ inline void load_tri(const float* t, F3& a, F3& b, F3& c) {
    a = {t[0], t[1], t[2]};
    b = {t[3], t[4], t[5]};
    c = {t[6], t[7], t[8]};
}

__global__ void k_subdiv_count(const float* __restrict__ tris, uint64_t n, float h, uint32_t* rows, uint64_t* subs,
                               unsigned* err) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 a, b, c;
    load_tri(tris + 9 * i, a, b, c);
    float num_subs = floorf(__fdiv_rn(tri_max_side(a, b, c), h));
    if (num_subs < 2.0f) {
        rows[i] = 1;
        subs[i] = 1;
        return;
    }
    // `for i in 0..num_subs as usize` (NaN -> 0 rows)
    float lim = 65536.0f;
    if (!(num_subs < lim)) {
        if (num_subs == num_subs) atomicOr(err, E_TOO_MANY_SUBS);
        rows[i] = 0;
        subs[i] = 0;
        return;
    }
    uint32_t ns = (uint32_t)num_subs;
    rows[i] = ns;
    subs[i] = (uint64_t)ns * ns;
}

// One thread per (triangle, row i): a_i is p1 + s1 accumulated i times, then the
// row emits (a_prev, b_s, a_s), (a_s, b_s, c_s) i times and (a, b, c), exactly
// as the reference's sequential loop does.
__global__ void k_subdiv_emit(const float* __restrict__ tris, uint64_t n, float h, const uint32_t* __restrict__ row_off,
                              const uint64_t* __restrict__ sub_off, uint64_t total_rows, float* __restrict__ out) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= total_rows) return;
    // triangle t: row_off[t] <= r < row_off[t+1]
    uint64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint64_t mid = (lo + hi) >> 1;
        if (row_off[mid] <= r) lo = mid;
        else hi = mid;
    }
    uint64_t t = lo;
    uint32_t i = (uint32_t)(r - row_off[t]);
    F3 p1, p2, p3;
    load_tri(tris + 9 * t, p1, p2, p3);
    float* o = out + 9 * sub_off[t];
    auto put = [&](uint64_t k, F3 a, F3 b, F3 c) {
        float* q = o + 9 * k;
        q[0] = a.x; q[1] = a.y; q[2] = a.z;
        q[3] = b.x; q[4] = b.y; q[5] = b.z;
        q[6] = c.x; q[7] = c.y; q[8] = c.z;
    };
    float num_subs = floorf(__fdiv_rn(tri_max_side(p1, p2, p3), h));
    if (num_subs < 2.0f) {
        put(0, p1, p2, p3);
        return;
    }
    float inv = __fdiv_rn(1.0f, num_subs);
    F3 s1 = f3_muls(f3_sub(p2, p1), inv);
    F3 s2 = f3_muls(f3_sub(p3, p2), inv);
    F3 a = p1;
    for (uint32_t k = 0; k < i; ++k) a = f3_add(a, s1);
    F3 b = f3_add(a, s1);
    F3 c = f3_add(b, s2);
    F3 a_s = f3_add(a, s2), b_s = f3_add(b, s2), c_s = f3_add(c, s2);
    F3 a_prev = a;
    uint64_t k0 = (uint64_t)i * i;
    for (uint32_t k = 0; k < i; ++k) {
        put(k0 + 2 * k, a_prev, b_s, a_s);
        put(k0 + 2 * k + 1, a_s, b_s, c_s);
        a_prev = a_s;
        a_s = f3_add(a_s, s2);
        b_s = f3_add(b_s, s2);
        c_s = f3_add(c_s, s2);
    }
    put(k0 + 2 * (uint64_t)i, a, b, c);
}

// ---------------------------------------------------------------------------
// dilated boxes (mesh_to_volume.rs:124-141) and distance records

__device__ inline int floor_div8(int v) { return v >> 3; }

__global__ void k_subtri_setup(const float* __restrict__ sub, uint64_t n_sub, float inv_h, int band,
                               SubRec* __restrict__ recs, uint32_t* __restrict__ cov, unsigned long long* counters,
                               unsigned* err) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long pairs = 0;
    int blo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, bhi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    if (t < n_sub) {
        F3 a, b, c;
        load_tri(sub + 9 * t, a, b, c);
        F3 mn = f3_min2(c, f3_min2(a, b));
        F3 mx = f3_max2(c, f3_max2(a, b));
        float fl[3] = {floorf(mn.x * inv_h), floorf(mn.y * inv_h), floorf(mn.z * inv_h)};
        float cl[3] = {ceilf(mx.x * inv_h), ceilf(mx.y * inv_h), ceilf(mx.z * inv_h)};
        int lo[3], hi[3];
        bool ok = true;
        for (int k = 0; k < 3; ++k) {
            ok &= fabsf(fl[k]) < 8.0e6f && fabsf(cl[k]) < 8.0e6f;
            lo[k] = ok ? (int)fl[k] - band : 0;
            hi[k] = ok ? (int)cl[k] + band : 0;
        }
        if (!ok) atomicOr(err, E_RANGE);
        if (hi[0] == lo[0] || hi[1] == lo[1] || hi[2] == lo[2])
            for (int k = 0; k < 3; ++k) { lo[k] -= 1; hi[k] += 1; }
        SubRec r;
        F3 v[3] = {a, b, c};
        for (int k = 0; k < 3; ++k) {
            r.v[k][0] = v[k].x; r.v[k][1] = v[k].y; r.v[k][2] = v[k].z;
            F3 e = f3_sub(v[(k + 1) % 3], v[k]);
            r.e[k][0] = e.x; r.e[k][1] = e.y; r.e[k][2] = e.z;
            float l2 = e.x * e.x + e.y * e.y + e.z * e.z;
            r.inv_l2[k] = l2 > 0.0f ? 1.0f / l2 : 0.0f;
        }
        F3 e0 = f3_sub(b, a), e2 = f3_sub(c, a);
        F3 nn = {e0.y * e2.z - e0.z * e2.y, e0.z * e2.x - e0.x * e2.z, e0.x * e2.y - e0.y * e2.x};
        float nl2 = nn.x * nn.x + nn.y * nn.y + nn.z * nn.z;
        bool degenerate = !(nl2 > 0.0f);
        float inl = degenerate ? 0.0f : 1.0f / sqrtf(nl2);
        r.n[0] = nn.x * inl; r.n[1] = nn.y * inl; r.n[2] = nn.z * inl;
        for (int k = 0; k < 3; ++k) {
            F3 e = {r.e[k][0], r.e[k][1], r.e[k][2]};
            // inward normal of edge k: n x e_k
            r.m[k][0] = degenerate ? 0.0f : nn.y * e.z - nn.z * e.y;
            r.m[k][1] = degenerate ? 0.0f : nn.z * e.x - nn.x * e.z;
            r.m[k][2] = degenerate ? 0.0f : nn.x * e.y - nn.y * e.x;
        }
        r.bias = degenerate ? -1.0f : 0.0f;
        for (int k = 0; k < 3; ++k) { r.box[k] = lo[k]; r.box[3 + k] = hi[k]; }
        recs[t] = r;
        uint32_t nl = 1;
        for (int k = 0; k < 3; ++k) nl *= (uint32_t)(floor_div8(hi[k]) - floor_div8(lo[k]) + 1);
        cov[t] = nl;
        pairs = (unsigned long long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
        for (int k = 0; k < 3; ++k) { blo[k] = lo[k]; bhi[k] = hi[k]; }
    }
    // wave reductions, then one atomic per wave
    for (int s = 32; s >= 1; s >>= 1) {
        pairs += __shfl_xor(pairs, s);
        for (int k = 0; k < 3; ++k) {
            blo[k] = min(blo[k], __shfl_xor(blo[k], s));
            bhi[k] = max(bhi[k], __shfl_xor(bhi[k], s));
        }
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&counters[C_PAIRS], pairs);
        for (int k = 0; k < 3; ++k) {
            atomicMin((long long*)&counters[C_BOX_LO + k], (long long)blo[k]);
            atomicMax((long long*)&counters[C_BOX_HI + k], (long long)bhi[k]);
        }
    }
}

// ---------------------------------------------------------------------------
// leaf hash-and-compact

template <class F>
__device__ inline void for_each_leaf(const SubRec& r, F&& f) {
    int x0 = floor_div8(r.box[0]), y0 = floor_div8(r.box[1]), z0 = floor_div8(r.box[2]);
    int x1 = floor_div8(r.box[3]), y1 = floor_div8(r.box[4]), z1 = floor_div8(r.box[5]);
    uint32_t j = 0;
    for (int x = x0; x <= x1; ++x)
        for (int y = y0; y <= y1; ++y)
            for (int z = z0; z <= z1; ++z) f(j++, x, y, z);
}

__global__ void k_hash_insert(const SubRec* __restrict__ recs, uint64_t n_sub, uint64_t* keys, uint32_t cap_mask,
                              unsigned* err) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_sub) return;
    const SubRec& r = recs[t];
    bool ok = true;
    for_each_leaf(r, [&](uint32_t, int x, int y, int z) { ok &= hash_insert(keys, cap_mask, leaf_key(x, y, z)); });
    if (!ok) atomicOr(err, E_HASH);
}

__global__ void k_hash_compact(const uint64_t* __restrict__ keys, uint32_t cap, uint64_t* out, unsigned long long* counters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool occ = i < cap && keys[i] != EMPTY_KEY;
    uint64_t m = __ballot(occ);
    unsigned long long base = 0;
    int lane = threadIdx.x & 63;
    if (lane == 0 && m) base = atomicAdd(&counters[C_LEAVES], (unsigned long long)__popcll(m));
    base = __shfl(base, 0);
    if (occ) out[base + __popcll(m & ((1ull << lane) - 1))] = keys[i];
}

__global__ void k_hash_assign(const uint64_t* __restrict__ sorted, uint64_t n_leaves, const uint64_t* __restrict__ keys,
                              uint32_t cap_mask, uint32_t* vals, int32_t* origins, unsigned* err) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_leaves) return;
    uint64_t k = sorted[i];
    uint32_t slot = hash_find(keys, cap_mask, k);
    if (slot == NOT_FOUND) atomicOr(err, E_HASH);
    else vals[slot] = (uint32_t)i;
    int lx, ly, lz;
    leaf_key_decode(k, lx, ly, lz);
    origins[4 * i + 0] = lx * 8;
    origins[4 * i + 1] = ly * 8;
    origins[4 * i + 2] = lz * 8;
    origins[4 * i + 3] = 0;
}

// candidate lists: for every leaf, the sub-triangles whose dilated box touches it
__global__ void k_cand_count(const SubRec* __restrict__ recs, uint64_t n_sub, const uint32_t* __restrict__ cov_off,
                             const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t cap_mask,
                             uint32_t* cand_leaf, uint32_t* cand_cnt, uint32_t n_leaves, unsigned* err) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_sub) return;
    const SubRec& r = recs[t];
    uint32_t base = cov_off[t];
    for_each_leaf(r, [&](uint32_t j, int x, int y, int z) {
        uint32_t slot = hash_find(keys, cap_mask, leaf_key(x, y, z));
        uint32_t lid = slot == NOT_FOUND ? NOT_FOUND : vals[slot];
        if (lid >= n_leaves) {
            atomicOr(err, E_HASH);
            lid = n_leaves;  // the spare counter slot: never read back as a leaf
        }
        cand_leaf[base + j] = lid;
        atomicAdd(&cand_cnt[lid], 1u);
    });
}

__global__ void k_cand_fill(uint64_t n_sub, const uint32_t* __restrict__ cov_off, const uint32_t* __restrict__ cand_leaf,
                            const uint32_t* __restrict__ cand_off, uint32_t* cursor, uint32_t* cand_list,
                            uint32_t n_leaves) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_sub) return;
    for (uint32_t e = cov_off[t]; e < cov_off[t + 1]; ++e) {
        uint32_t lid = cand_leaf[e];
        if (lid >= n_leaves) continue;
        uint32_t pos = atomicAdd(&cursor[lid], 1u);
        cand_list[cand_off[lid] + pos] = (uint32_t)t;
    }
}

// ---------------------------------------------------------------------------
// distance: one workgroup per leaf, wave w owns the 4^3 block
// (w>>2, (w>>1)&1, w&1) * 4, one voxel per lane.  A voxel's candidates are the
// sub-triangles whose dilated box contains it (the reference pair set); the
// wave skips a candidate whose box misses its block, each lane skips one whose
// box misses its voxel.  The arg-min by squared distance (ties -> lower
// sub-triangle index) is re-evaluated with the reference's closest-point
// arithmetic to give the stored value.
__global__ __launch_bounds__(512) void k_distance(const int32_t* __restrict__ origins, const uint32_t* __restrict__ cand_off,
                                                  const uint32_t* __restrict__ cand_list, const SubRec* __restrict__ recs,
                                                  float h, float* __restrict__ values, uint64_t* __restrict__ masks) {
    __shared__ unsigned long long smask[8];
    const int leaf = blockIdx.x;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    if (tid < 8) smask[tid] = 0;
    const int ox = origins[4 * leaf], oy = origins[4 * leaf + 1], oz = origins[4 * leaf + 2];
    const int bx = (wave >> 2) * 4, by = ((wave >> 1) & 1) * 4, bz = (wave & 1) * 4;
    const int lx = bx + (lane >> 4), ly = by + ((lane >> 2) & 3), lz = bz + (lane & 3);
    const int vx = ox + lx, vy = oy + ly, vz = oz + lz;
    const float px = (float)vx * h, py = (float)vy * h, pz = (float)vz * h;
    const int X0 = ox + bx, Y0 = oy + by, Z0 = oz + bz;
    float best = INFINITY;
    uint32_t best_t = 0xFFFFFFFFu;
    const uint32_t c0 = cand_off[leaf], c1 = cand_off[leaf + 1];
    for (uint32_t c = c0; c < c1; ++c) {
        const uint32_t t = __builtin_amdgcn_readfirstlane(cand_list[c]);
        const SubRec& r = recs[t];
        const int b0 = r.box[0], b1 = r.box[1], b2 = r.box[2], b3 = r.box[3], b4 = r.box[4], b5 = r.box[5];
        if (b0 > X0 + 3 || b3 < X0 || b1 > Y0 + 3 || b4 < Y0 || b2 > Z0 + 3 || b5 < Z0) continue;
        const bool inb = (unsigned)(vx - b0) <= (unsigned)(b3 - b0) && (unsigned)(vy - b1) <= (unsigned)(b4 - b1) &&
                         (unsigned)(vz - b2) <= (unsigned)(b5 - b2);
        if (inb) {
            float d2 = fast_dist2(r, px, py, pz);
            if (d2 < best || (d2 == best && t < best_t)) {
                best = d2;
                best_t = t;
            }
        }
    }
    const int off = (lx << 6) | (ly << 3) | lz;
    const bool active = best_t != 0xFFFFFFFFu;
    float v = 0.0f;
    if (active) {
        const SubRec& r = recs[best_t];
        v = ericson_distance({r.v[0][0], r.v[0][1], r.v[0][2]}, {r.v[1][0], r.v[1][1], r.v[1][2]},
                             {r.v[2][0], r.v[2][1], r.v[2][2]}, {px, py, pz});
    }
    values[(uint64_t)leaf * LEAF_VOX + off] = v;
    __syncthreads();
    if (active) atomicOr(&smask[off >> 6], 1ull << (off & 63));
    __syncthreads();
    if (tid < 8) masks[(uint64_t)leaf * 8 + tid] = smask[tid];
}

// active voxel list: slot = leaf * 512 + offset, leaves in order
__global__ void k_leaf_popcount(const uint64_t* __restrict__ masks, uint64_t n_leaves, uint32_t* cnt) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_leaves) return;
    uint32_t c = 0;
    for (int w = 0; w < 8; ++w) c += __popcll(masks[8 * i + w]);
    cnt[i] = c;
}

__global__ void k_active_list(const uint64_t* __restrict__ masks, const uint32_t* __restrict__ act_off, uint32_t* out) {
    const int leaf = blockIdx.x;
    const int tid = threadIdx.x;  // 512 threads, one per voxel
    const int w = tid >> 6, lane = tid & 63;
    const uint64_t m = masks[(uint64_t)leaf * 8 + w];
    uint32_t before = 0;
    for (int k = 0; k < w; ++k) before += __popcll(masks[(uint64_t)leaf * 8 + k]);
    if ((m >> lane) & 1) out[act_off[leaf] + before + __popcll(m & ((1ull << lane) - 1))] = (uint32_t)leaf * 512u + tid;
}

// ---------------------------------------------------------------------------
// LBVH over the original triangles + order-2 dipole data (aabb_tree.rs:711-816)

__device__ inline unsigned ordered_bits(float f) {
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float from_ordered_bits(unsigned u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__global__ void k_centroid_bounds(const float* __restrict__ tris, uint64_t n, unsigned* bounds) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned lo[3] = {~0u, ~0u, ~0u}, hi[3] = {0, 0, 0};
    if (i < n) {
        F3 a, b, c;
        load_tri(tris + 9 * i, a, b, c);
        F3 mn = f3_min2(c, f3_min2(a, b)), mx = f3_max2(c, f3_max2(a, b));
        float ce[3] = {(mn.x + mx.x) * 0.5f, (mn.y + mx.y) * 0.5f, (mn.z + mx.z) * 0.5f};
        for (int k = 0; k < 3; ++k) lo[k] = hi[k] = ordered_bits(ce[k]);
    }
    for (int s = 32; s >= 1; s >>= 1)
        for (int k = 0; k < 3; ++k) {
            lo[k] = min(lo[k], (unsigned)__shfl_xor((int)lo[k], s));
            hi[k] = max(hi[k], (unsigned)__shfl_xor((int)hi[k], s));
        }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k) {
            atomicMin(&bounds[k], lo[k]);
            atomicMax(&bounds[3 + k], hi[k]);
        }
}

__device__ inline uint64_t spread21(uint64_t v) {
    v &= 0x1FFFFF;
    v = (v | v << 32) & 0x1F00000000FFFFull;
    v = (v | v << 16) & 0x1F0000FF0000FFull;
    v = (v | v << 8) & 0x100F00F00F00F00Full;
    v = (v | v << 4) & 0x10C30C30C30C30C3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

__global__ void k_morton(const float* __restrict__ tris, uint64_t n, const unsigned* __restrict__ bounds,
                         uint64_t* codes, uint32_t* idx) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 a, b, c;
    load_tri(tris + 9 * i, a, b, c);
    F3 mn = f3_min2(c, f3_min2(a, b)), mx = f3_max2(c, f3_max2(a, b));
    float ce[3] = {(mn.x + mx.x) * 0.5f, (mn.y + mx.y) * 0.5f, (mn.z + mx.z) * 0.5f};
    uint64_t q[3];
    for (int k = 0; k < 3; ++k) {
        float lo = from_ordered_bits(bounds[k]), hi = from_ordered_bits(bounds[3 + k]);
        float ext = hi - lo;
        float u = ext > 0.0f ? (ce[k] - lo) / ext : 0.0f;
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        q[k] = (uint64_t)fminf(u * 2097152.0f, 2097151.0f);
    }
    codes[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    idx[i] = (uint32_t)i;
}

__device__ inline int lbvh_delta(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((unsigned)(i ^ j));
    return __clzll(a ^ b);
}

// Karras 2012: internal node i of n-1; leaves are n-1+k.  links = (left, right,
// first, count) over the sorted order.
__global__ void k_lbvh_build(const uint64_t* __restrict__ keys, int n, int4* links, int* parent) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (lbvh_delta(keys, n, i, i + 1) - lbvh_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = lbvh_delta(keys, n, i, i - d);
    int lmax = 2;
    while (lbvh_delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (lbvh_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = lbvh_delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (lbvh_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + min(d, 0);
    int first = min(i, j), last = max(i, j);
    int left = (first == gamma) ? (n - 1 + gamma) : gamma;
    int right = (last == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    links[i] = make_int4(left, right, first, last - first + 1);
    parent[left] = i;
    parent[right] = i;
}

// Per-node accumulation of aabb_tree.rs:711-801 (InitData): area-weighted
// normal, area-weighted centre, sum of a*c*n^T, total area, plus the bbox.
struct NodeAcc {
    float awn[3], awc[3], o1[9], area;
    float bmin[3], bmax[3];
};

__device__ inline void acc_add(NodeAcc& r, const NodeAcc& a, const NodeAcc& b) {
    for (int k = 0; k < 3; ++k) {
        r.awn[k] = a.awn[k] + b.awn[k];
        r.awc[k] = a.awc[k] + b.awc[k];
        r.bmin[k] = fminf(a.bmin[k], b.bmin[k]);
        r.bmax[k] = fmaxf(a.bmax[k], b.bmax[k]);
    }
    for (int k = 0; k < 9; ++k) r.o1[k] = a.o1[k] + b.o1[k];
    r.area = a.area + b.area;
}

// Node record (64 B): (c.xyz, (radius*acc)^2), (awn.xyz, trace O2),
// (S00, S11, S22, 2 S01), (2 S02, 2 S12, 0, 0) with S the symmetric part of
// order2 = sum a c n^T - c_d awn^T; only S enters r^T O2 r.
__device__ inline void acc_finalize(const NodeAcc& a, float acc, float4* node) {
    float c[3];
    for (int k = 0; k < 3; ++k) c[k] = a.awc[k] / a.area;
    float o2[9];
    for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) o2[3 * r + q] = a.o1[3 * r + q] - c[r] * a.awn[q];
    float dmin = 0.0f, dmax = 0.0f;
    for (int k = 0; k < 3; ++k) {
        dmin += (a.bmin[k] - c[k]) * (a.bmin[k] - c[k]);
        dmax += (a.bmax[k] - c[k]) * (a.bmax[k] - c[k]);
    }
    float radius = sqrtf(fmaxf(dmin, dmax));
    float ra = radius * acc;
    float far2 = (a.area > 0.0f) ? ra * ra : INFINITY;  // all-degenerate: never far
    node[0] = make_float4(c[0], c[1], c[2], far2);
    node[1] = make_float4(a.awn[0], a.awn[1], a.awn[2], o2[0] + o2[4] + o2[8]);
    node[2] = make_float4(o2[0], o2[4], o2[8], o2[1] + o2[3]);
    node[3] = make_float4(o2[2] + o2[6], o2[5] + o2[7], 0.0f, 0.0f);
}

__global__ void k_lbvh_leaves(const float* __restrict__ tris, int n, const uint32_t* __restrict__ order, float acc,
                              const int4* __restrict__ links, const int* __restrict__ parent, unsigned* flags,
                              NodeAcc* accs, float4* nodes, float* btris, int4* leaf_links, unsigned* err) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    F3 a, b, c;
    load_tri(tris + 9 * (uint64_t)order[k], a, b, c);
    float* bt = btris + 9 * (uint64_t)k;
    bt[0] = a.x; bt[1] = a.y; bt[2] = a.z; bt[3] = b.x; bt[4] = b.y; bt[5] = b.z; bt[6] = c.x; bt[7] = c.y; bt[8] = c.z;
    NodeAcc d;
    F3 e0 = f3_sub(b, a), e1 = f3_sub(c, a);
    F3 cr = {e0.y * e1.z - e0.z * e1.y, e0.z * e1.x - e0.x * e1.z, e0.x * e1.y - e0.y * e1.x};
    float cl2 = cr.x * cr.x + cr.y * cr.y + cr.z * cr.z;
    F3 mn = f3_min2(c, f3_min2(a, b)), mx = f3_max2(c, f3_max2(a, b));
    d.bmin[0] = mn.x; d.bmin[1] = mn.y; d.bmin[2] = mn.z;
    d.bmax[0] = mx.x; d.bmax[1] = mx.y; d.bmax[2] = mx.z;
    for (int q = 0; q < 3; ++q) { d.awn[q] = 0.0f; d.awc[q] = 0.0f; }
    for (int q = 0; q < 9; ++q) d.o1[q] = 0.0f;
    d.area = 0.0f;
    if (cl2 > 0.0f) {  // degenerate triangles are skipped (aabb_tree.rs:757-759)
        float cl = sqrtf(cl2);
        float area = cl * 0.5f;
        float nrm[3] = {cr.x / cl, cr.y / cl, cr.z / cl};
        float ce[3] = {(a.x + b.x + c.x) / 3.0f, (a.y + b.y + c.y) / 3.0f, (a.z + b.z + c.z) / 3.0f};
        d.area = area;
        for (int q = 0; q < 3; ++q) {
            d.awn[q] = area * nrm[q];
            d.awc[q] = area * ce[q];
        }
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) d.o1[3 * r + q] = (area * ce[r]) * nrm[q];
    }
    int node = n - 1 + k;
    accs[node] = d;
    acc_finalize(d, acc, nodes + 4 * (uint64_t)node);
    leaf_links[k] = make_int4(-1, -1, k, 1);
    int p = parent[node];
    for (int level = 0; p >= 0; ++level) {
        if (level > 4096) {  // deeper than any tree of < 2^30 leaves: corrupt parent links
            atomicOr(err, E_TREE);
            return;
        }
        __threadfence();
        unsigned old = atomicAdd(&flags[p], 1u);
        if (old == 0) return;  // the sibling finishes this node
        __threadfence();
        int4 L = links[p];
        NodeAcc l = accs[L.x], r = accs[L.y], s;
        acc_add(s, l, r);
        accs[p] = s;
        acc_finalize(s, acc, nodes + 4 * (uint64_t)p);
        p = parent[p];
    }
}

// ---------------------------------------------------------------------------
// winding numbers at active voxels, wave-coherent traversal

// Solid angle of triangle (a, b, c) seen from the origin (a, b, c already
// relative to the query), aabb_tree.rs:582-615 in the Van Oosterom-Strackee
// form: tan(W/2) = det[a b c] / (|a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a|),
// which equals the reference's normalised expression.  Zero when a vertex
// coincides with the query or the triple product vanishes.
__device__ inline float solid_angle_rel(float ax, float ay, float az, float bx, float by, float bz, float cx, float cy,
                                        float cz) {
    float la = sqrtf(ax * ax + ay * ay + az * az);
    float lb = sqrtf(bx * bx + by * by + bz * bz);
    float lc = sqrtf(cx * cx + cy * cy + cz * cz);
    float det = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    float den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (ax * cx + ay * cy + az * cz) * lb +
                (bx * cx + by * cy + bz * cz) * la;
    float w = 2.0f * atan2f(det, den);
    return (la == 0.0f || lb == 0.0f || lc == 0.0f || det == 0.0f) ? 0.0f : w;
}

__device__ inline double solid_angle_rel64(double ax, double ay, double az, double bx, double by, double bz, double cx,
                                           double cy, double cz) {
    double la = sqrt(ax * ax + ay * ay + az * az);
    double lb = sqrt(bx * bx + by * by + bz * bz);
    double lc = sqrt(cx * cx + cy * cy + cz * cz);
    double det = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    double den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (ax * cx + ay * cy + az * cz) * lb +
                 (bx * cx + by * cy + bz * cz) * la;
    if (la == 0.0 || lb == 0.0 || lc == 0.0 || det == 0.0) return 0.0;
    return 2.0 * atan2(det, den);
}

// Neumaier compensated accumulation
__device__ inline void kb_add(float& s, float& c, float x) {
    float t = s + x;
    c += (fabsf(s) >= fabsf(x)) ? ((s - t) + x) : ((x - t) + s);
    s = t;
}

// order-2 far-field term of a node at query offset r = c - p
__device__ inline float dipole_term(const float4* __restrict__ nd, float rx, float ry, float rz, float r2) {
    float4 B = nd[1], C = nd[2], D = nd[3];
    float inv_r = 1.0f / sqrtf(r2);
    float inv_r2 = inv_r * inv_r;
    float rsr = C.x * rx * rx + C.y * ry * ry + C.z * rz * rz + C.w * rx * ry + D.x * rx * rz + D.y * ry * rz;
    float lin = B.x * rx + B.y * ry + B.z * rz + B.w;
    return INV_4PI * inv_r * inv_r2 * (lin - 3.0f * rsr * inv_r2);
}

__global__ __launch_bounds__(WN_BLOCK) void k_winding(const uint32_t* __restrict__ act, uint32_t n_act,
                                                      const int32_t* __restrict__ origins, float h,
                                                      const float4* __restrict__ nodes, const int4* __restrict__ links,
                                                      const int4* __restrict__ leaf_links, int n_int, int n_nodes,
                                                      const float* __restrict__ btris, float thr, float fp64_band,
                                                      float* __restrict__ values, uint32_t* __restrict__ recheck,
                                                      unsigned long long* counters, unsigned* err) {
    __shared__ int st_node[WN_BLOCK / 64][WN_STACK];
    __shared__ unsigned long long st_mask[WN_BLOCK / 64][WN_STACK];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t gid = blockIdx.x * WN_BLOCK + threadIdx.x;
    const bool valid = gid < n_act;
    const uint32_t slot = valid ? act[gid] : 0u;
    const uint32_t leaf = slot >> 9, off = slot & 511u;
    int vx = 0, vy = 0, vz = 0;
    if (valid) {
        vx = origins[4 * leaf] + (int)(off >> 6);
        vy = origins[4 * leaf + 1] + (int)((off >> 3) & 7);
        vz = origins[4 * leaf + 2] + (int)(off & 7);
    }
    const float px = (float)vx * h, py = (float)vy * h, pz = (float)vz * h;
    const unsigned long long init = __ballot(valid);
    if (init == 0) return;
    float s = 0.0f, cmp = 0.0f;
    unsigned long long n_near_int = 0, n_far = 0, n_leaf = 0, n_tri = 0;
    int sp = 1;
    st_node[wave][0] = 0;  // root: internal node 0 (or leaf 0 when n == 1)
    st_mask[wave][0] = init;
    bool overflow = false;
    // every node is pushed at most once per wave in a tree: more pops than
    // nodes means corrupt links, and the walk stops
    for (int pops = 0; sp > 0; ++pops) {
        if (pops > n_nodes) {
            overflow = true;
            break;
        }
        --sp;
        const int node = __builtin_amdgcn_readfirstlane(st_node[wave][sp]);
        const unsigned long long m = st_mask[wave][sp];
        const bool act_l = (m >> lane) & 1ull;
        const float4* nd = nodes + 4 * (uint64_t)node;
        const float4 A = nd[0];
        const float rx = A.x - px, ry = A.y - py, rz = A.z - pz;
        const float r2 = rx * rx + ry * ry + rz * rz;
        const bool far = r2 > A.w;
        if (act_l && far) kb_add(s, cmp, dipole_term(nd, rx, ry, rz, r2));
        const bool near = act_l && !far;
        const unsigned long long nm = __ballot(near);
        n_far += __popcll(m & __ballot(far));
        if (nm == 0) continue;
        const int4 L = node < n_int ? links[node] : leaf_links[node - n_int];
        if (L.w <= BVH_LEAF_SIZE) {
            n_leaf += __popcll(nm);
            n_tri += (unsigned long long)__popcll(nm) * L.w;
            if (near) {
                float om = 0.0f;
                for (int k = 0; k < L.w; ++k) {
                    const float* t = btris + 9 * (uint64_t)(L.z + k);
                    om += solid_angle_rel(t[0] - px, t[1] - py, t[2] - pz, t[3] - px, t[4] - py, t[5] - pz, t[6] - px,
                                          t[7] - py, t[8] - pz);
                }
                kb_add(s, cmp, om * INV_4PI);
            }
        } else {
            n_near_int += __popcll(nm);
            if (sp + 2 > WN_STACK) {
                overflow = true;
                break;
            }
            st_node[wave][sp] = L.y;
            st_mask[wave][sp] = nm;
            st_node[wave][sp + 1] = L.x;
            st_mask[wave][sp + 1] = nm;
            sp += 2;
        }
    }
    if (overflow && lane == 0) atomicOr(err, E_STACK);
    if (lane == 0) {
        atomicAdd(&counters[C_NEAR_INTERNAL], n_near_int);
        atomicAdd(&counters[C_FAR], n_far);
        atomicAdd(&counters[C_LEAF_VISITS], n_leaf);
        atomicAdd(&counters[C_LEAF_TRIS], n_tri);
    }
    if (!valid) return;
    const float w = s + cmp;
    float* vp = values + slot;
    const float d = *vp;
    *vp = (w < thr) ? copysignf(d, 1.0f) : copysignf(d, -1.0f);
    if (fabsf(w - thr) < fp64_band) {
        uint32_t k = (uint32_t)atomicAdd(&counters[C_RECHECK], 1ull);
        recheck[k] = slot;
    }
}

// FP64 re-evaluation of the voxels whose f32 winding number fell inside the
// ambiguity band around the threshold: same tree, every term in double.
__global__ void k_winding_fp64(const uint32_t* __restrict__ list, uint32_t n, const int32_t* __restrict__ origins, float h,
                               const float4* __restrict__ nodes, const int4* __restrict__ links,
                               const int4* __restrict__ leaf_links, int n_int, int n_nodes,
                               const float* __restrict__ btris, float thr, float* __restrict__ values, unsigned* err) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t slot = list[i];
    const uint32_t leaf = slot >> 9, off = slot & 511u;
    const int vx = origins[4 * leaf] + (int)(off >> 6);
    const int vy = origins[4 * leaf + 1] + (int)((off >> 3) & 7);
    const int vz = origins[4 * leaf + 2] + (int)(off & 7);
    const double px = (float)vx * h, py = (float)vy * h, pz = (float)vz * h;
    int stack[WN_STACK];
    int sp = 0;
    stack[sp++] = 0;
    double w = 0.0;
    for (int pops = 0; sp > 0; ++pops) {
        if (pops > n_nodes) {
            atomicOr(err, E_STACK);
            return;
        }
        const int node = stack[--sp];
        const float4* nd = nodes + 4 * (uint64_t)node;
        const float4 A = nd[0], B = nd[1], C = nd[2], D = nd[3];
        const double rx = A.x - px, ry = A.y - py, rz = A.z - pz;
        const double r2 = rx * rx + ry * ry + rz * rz;
        if (r2 > (double)A.w) {
            double inv_r = 1.0 / sqrt(r2), inv_r2 = inv_r * inv_r;
            double rsr = C.x * rx * rx + C.y * ry * ry + C.z * rz * rz + C.w * rx * ry + D.x * rx * rz + D.y * ry * rz;
            double lin = B.x * rx + B.y * ry + B.z * rz + B.w;
            w += (1.0 / (4.0 * M_PI)) * inv_r * inv_r2 * (lin - 3.0 * rsr * inv_r2);
            continue;
        }
        const int4 L = node < n_int ? links[node] : leaf_links[node - n_int];
        if (L.w <= BVH_LEAF_SIZE) {
            double om = 0.0;
            for (int k = 0; k < L.w; ++k) {
                const float* t = btris + 9 * (uint64_t)(L.z + k);
                om += solid_angle_rel64(t[0] - px, t[1] - py, t[2] - pz, t[3] - px, t[4] - py, t[5] - pz, t[6] - px,
                                        t[7] - py, t[8] - pz);
            }
            w += om / (4.0 * M_PI);
        } else {
            if (sp + 2 > WN_STACK) {
                atomicOr(err, E_STACK);
                return;
            }
            stack[sp++] = L.y;
            stack[sp++] = L.x;
        }
    }
    float* vp = values + slot;
    const float d = *vp;
    *vp = (w < (double)thr) ? copysignf(d, 1.0f) : copysignf(d, -1.0f);
}

// ---------------------------------------------------------------------------
// host driver

template <class T>
T read_scalar(bs_ctx* ctx, const T* dptr) {
    T v;
    BS_HIP(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    BS_HIP(hipStreamSynchronize(ctx->stream));
    return v;
}

template <class TIn, class TOut>
void exclusive_scan(bs_ctx* ctx, const TIn* in, TOut* out, uint64_t n) {
    size_t bytes = 0;
    BS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, ctx->stream));
    void* tmp = ctx->scratch.get(S_CUB, bytes);
    BS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, ctx->stream));
}

struct PhaseTimer {
    bs_ctx* ctx;
    hipEvent_t ev[8];
    int n = 0;
    explicit PhaseTimer(bs_ctx* c) : ctx(c) {
        if (ctx->profiling)
            for (auto& e : ev) BS_HIP(hipEventCreate(&e));
    }
    ~PhaseTimer() {
        if (ctx->profiling)
            for (auto& e : ev) (void)hipEventDestroy(e);
    }
    void mark() {
        if (ctx->profiling && n < 8) BS_HIP(hipEventRecord(ev[n++], ctx->stream));
    }
    double ms(int a, int b) {
        if (!ctx->profiling || b >= n) return 0.0;
        float f = 0;
        BS_HIP(hipEventSynchronize(ev[b]));
        BS_HIP(hipEventElapsedTime(&f, ev[a], ev[b]));
        return f;
    }
};

uint32_t next_pow2(uint64_t v) {
    uint64_t p = 1024;
    while (p < v) p <<= 1;
    if (p > (1ull << 31)) throw Error(BS_ERR_INVALID_ARG, "leaf hash table would exceed 2^31 slots");
    return (uint32_t)p;
}

}  // namespace

void mesh_to_volume_device(bs_ctx* ctx, const float* d_tris, uint64_t n, const bs_m2v_params& prm, bs_volume** out) {
    *out = nullptr;
    BS_REQUIRE(prm.voxel_size > 0.0f && std::isfinite(prm.voxel_size), BS_ERR_INVALID_ARG, "voxel_size must be > 0");
    BS_REQUIRE(prm.band >= 0 && prm.band < (1 << 20), BS_ERR_INVALID_ARG, "band must be in [0, 2^20)");
    if (n == 0) throw Error(BS_ERR_EMPTY, "empty mesh: the reference returns None");
    BS_REQUIRE(n < (1ull << 30), BS_ERR_INVALID_ARG, "more than 2^30 triangles");
    hipStream_t st = ctx->stream;
    auto& S = ctx->scratch;
    const float h = prm.voxel_size;
    const float inv_h = 1.0f / h;  // MeshToVolume::set_voxel_size (mesh_to_volume.rs:48)
    const int band = (int)prm.band;
    PhaseTimer tm(ctx);
    auto wall0 = std::chrono::steady_clock::now();
    tm.mark();

    unsigned long long* counters = S.as<unsigned long long>(S_COUNTERS, C_COUNT);
    unsigned* err = reinterpret_cast<unsigned*>(S.get(S_MISC3, 64));
    {
        unsigned long long init[C_COUNT] = {};
        for (int k = 0; k < 3; ++k) {
            init[C_BOX_LO + k] = (unsigned long long)(long long)INT32_MAX;
            init[C_BOX_HI + k] = (unsigned long long)(long long)INT32_MIN;
        }
        BS_HIP(hipMemcpyAsync(counters, init, sizeof(init), hipMemcpyHostToDevice, st));
        BS_HIP(hipMemsetAsync(err, 0, 64, st));
    }
    const unsigned TB = 256;

    // ---- 1. subdivision
    uint32_t* rows = S.as<uint32_t>(S_ROWS, n + 1);
    uint64_t* subs = S.as<uint64_t>(S_SUBS, n + 1);
    uint32_t* row_off = S.as<uint32_t>(S_ROW_OFF, n + 1);
    uint64_t* sub_off = S.as<uint64_t>(S_SUB_OFF, n + 1);
    BS_HIP(hipMemsetAsync(rows + n, 0, sizeof(uint32_t), st));
    BS_HIP(hipMemsetAsync(subs + n, 0, sizeof(uint64_t), st));
    k_subdiv_count<<<div_up(n, TB), TB, 0, st>>>(d_tris, n, h, rows, subs, err);
    exclusive_scan(ctx, rows, row_off, n + 1);
    exclusive_scan(ctx, subs, sub_off, n + 1);
    uint64_t total_rows = read_scalar(ctx, row_off + n);
    uint64_t n_sub = read_scalar(ctx, sub_off + n);
    if (read_scalar(ctx, err) & E_TOO_MANY_SUBS)
        throw Error(BS_ERR_INVALID_ARG, "a triangle is subdivided into more than 2^32 sub-triangles");
    if (n_sub == 0) throw Error(BS_ERR_EMPTY, "no sub-triangles: the reference returns None");
    BS_REQUIRE(n_sub < (1ull << 32), BS_ERR_INVALID_ARG, "more than 2^32 sub-triangles");
    float* sub = S.as<float>(S_SUBTRIS, 9 * n_sub);
    k_subdiv_emit<<<div_up(total_rows, TB), TB, 0, st>>>(d_tris, n, h, row_off, sub_off, total_rows, sub);
    tm.mark();

    // ---- 2. boxes, records, leaf hash-and-compact
    SubRec* recs = S.as<SubRec>(S_RECS, n_sub);
    uint32_t* cov = S.as<uint32_t>(S_LEAFCNT, n_sub + 1);
    uint32_t* cov_off = S.as<uint32_t>(S_LEAFOFF, n_sub + 1);
    BS_HIP(hipMemsetAsync(cov + n_sub, 0, sizeof(uint32_t), st));
    k_subtri_setup<<<div_up(n_sub, TB), TB, 0, st>>>(sub, n_sub, inv_h, band, recs, cov, counters, err);
    exclusive_scan(ctx, cov, cov_off, n_sub + 1);
    unsigned long long hc[C_COUNT];
    BS_HIP(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, st));
    uint32_t n_cov = read_scalar(ctx, cov_off + n_sub);
    if (read_scalar(ctx, err) & E_RANGE)
        throw Error(BS_ERR_INVALID_ARG, "mesh extent exceeds 8e6 voxels from the origin");
    const uint64_t n_pairs = hc[C_PAIRS];
    uint64_t bbox_leaves = 1;
    for (int k = 0; k < 3; ++k) {
        long long lo = (long long)hc[C_BOX_LO + k] >> 3, hi = (long long)hc[C_BOX_HI + k] >> 3;
        bbox_leaves *= (uint64_t)(hi - lo + 1);
        if (bbox_leaves > (1ull << 40)) bbox_leaves = 1ull << 40;
    }
    uint32_t cap = next_pow2(2 * std::min<uint64_t>(n_cov, bbox_leaves));
    uint64_t* hkeys = S.as<uint64_t>(S_HASH_KEYS, cap);
    uint32_t* hvals = S.as<uint32_t>(S_HASH_VALS, cap);
    BS_HIP(hipMemsetAsync(hkeys, 0xFF, sizeof(uint64_t) * cap, st));
    k_hash_insert<<<div_up(n_sub, TB), TB, 0, st>>>(recs, n_sub, hkeys, cap - 1, err);
    uint64_t* keys = S.as<uint64_t>(S_KEYS, std::min<uint64_t>(n_cov, cap));
    k_hash_compact<<<div_up(cap, TB), TB, 0, st>>>(hkeys, cap, keys, counters);
    const uint64_t n_leaves = read_scalar(ctx, counters + C_LEAVES);
    uint64_t* keys_sorted = S.as<uint64_t>(S_KEYS_SORTED, n_leaves);
    {
        size_t bytes = 0;
        BS_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys, keys_sorted, (int)n_leaves, 0, 63, st));
        void* tmp = S.get(S_CUB, bytes);
        BS_HIP(hipcub::DeviceRadixSort::SortKeys(tmp, bytes, keys, keys_sorted, (int)n_leaves, 0, 63, st));
    }
    // output volume
    auto* vol = new bs_volume();
    vol->device = ctx->device;
    vol->voxel_size = h;
    vol->n_leaves = n_leaves;
    try {
        BS_HIP(hipMalloc(&vol->origins, sizeof(int32_t) * 4 * n_leaves));
        BS_HIP(hipMalloc(&vol->masks, sizeof(uint64_t) * 8 * n_leaves));
        BS_HIP(hipMalloc(&vol->values, sizeof(float) * 512 * n_leaves));
    } catch (...) {
        (void)hipFree(vol->origins);
        (void)hipFree(vol->masks);
        (void)hipFree(vol->values);
        delete vol;
        throw;
    }
    std::unique_ptr<bs_volume, void (*)(bs_volume*)> guard(vol, [](bs_volume* v) {
        (void)hipFree(v->origins);
        (void)hipFree(v->masks);
        (void)hipFree(v->values);
        delete v;
    });
    k_hash_assign<<<div_up(n_leaves, TB), TB, 0, st>>>(keys_sorted, n_leaves, hkeys, cap - 1, hvals, vol->origins, err);
    // candidate lists
    uint32_t* cand_leaf = S.as<uint32_t>(S_CAND_LEAF, n_cov);
    uint32_t* cand_cnt = S.as<uint32_t>(S_CAND_CNT, n_leaves + 1);
    uint32_t* cand_off = S.as<uint32_t>(S_CAND_OFF, n_leaves + 1);
    uint32_t* cursor = S.as<uint32_t>(S_CURSOR, n_leaves);
    uint32_t* cand_list = S.as<uint32_t>(S_CAND_LIST, n_cov);
    BS_HIP(hipMemsetAsync(cand_cnt, 0, sizeof(uint32_t) * (n_leaves + 1), st));
    BS_HIP(hipMemsetAsync(cursor, 0, sizeof(uint32_t) * n_leaves, st));
    k_cand_count<<<div_up(n_sub, TB), TB, 0, st>>>(recs, n_sub, cov_off, hkeys, hvals, cap - 1, cand_leaf, cand_cnt,
                                                   (uint32_t)n_leaves, err);
    exclusive_scan(ctx, cand_cnt, cand_off, n_leaves + 1);
    k_cand_fill<<<div_up(n_sub, TB), TB, 0, st>>>(n_sub, cov_off, cand_leaf, cand_off, cursor, cand_list,
                                                  (uint32_t)n_leaves);
    if (read_scalar(ctx, err) & E_HASH) throw Error(BS_ERR_INTERNAL, "leaf hash table overflow");
    tm.mark();

    // ---- 3. distance
    k_distance<<<(unsigned)n_leaves, 512, 0, st>>>(vol->origins, cand_off, cand_list, recs, h, vol->values, vol->masks);
    tm.mark();

    // ---- 4. active list
    uint32_t* act_cnt = S.as<uint32_t>(S_ACT_CNT, n_leaves + 1);
    uint32_t* act_off = S.as<uint32_t>(S_ACT_OFF, n_leaves + 1);
    BS_HIP(hipMemsetAsync(act_cnt + n_leaves, 0, sizeof(uint32_t), st));
    k_leaf_popcount<<<div_up(n_leaves, TB), TB, 0, st>>>(vol->masks, n_leaves, act_cnt);
    exclusive_scan(ctx, act_cnt, act_off, n_leaves + 1);
    const uint32_t n_act = read_scalar(ctx, act_off + n_leaves);
    uint32_t* act = S.as<uint32_t>(S_ACT_LIST, n_act);
    k_active_list<<<(unsigned)n_leaves, 512, 0, st>>>(vol->masks, act_off, act);

    // ---- 5. LBVH + dipoles over the original mesh
    const int nt = (int)n;
    const int n_int = nt - 1;
    unsigned* bounds = reinterpret_cast<unsigned*>(S.get(S_MISC0, 64));
    {
        unsigned init[6] = {~0u, ~0u, ~0u, 0u, 0u, 0u};
        BS_HIP(hipMemcpyAsync(bounds, init, sizeof(init), hipMemcpyHostToDevice, st));
    }
    k_centroid_bounds<<<div_up(n, TB), TB, 0, st>>>(d_tris, n, bounds);
    uint64_t* codes = S.as<uint64_t>(S_MORTON, n);
    uint64_t* codes_s = S.as<uint64_t>(S_MORTON_SORTED, n);
    uint32_t* idx = S.as<uint32_t>(S_IDX, n);
    uint32_t* idx_s = S.as<uint32_t>(S_IDX_SORTED, n);
    k_morton<<<div_up(n, TB), TB, 0, st>>>(d_tris, n, bounds, codes, idx);
    {
        size_t bytes = 0;
        BS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, codes, codes_s, idx, idx_s, nt, 0, 63, st));
        void* tmp = S.get(S_CUB, bytes);
        BS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, codes, codes_s, idx, idx_s, nt, 0, 63, st));
    }
    const int n_nodes = 2 * nt - 1;
    int4* links = S.as<int4>(S_BVH_LINKS, std::max(n_int, 1));
    int4* leaf_links = S.as<int4>(S_MISC1, nt);
    int* parent = S.as<int>(S_BVH_PARENT, n_nodes);
    unsigned* flags = S.as<unsigned>(S_BVH_FLAGS, std::max(n_int, 1));
    NodeAcc* accs = S.as<NodeAcc>(S_BVH_INIT, n_nodes);
    float4* nodes = S.as<float4>(S_BVH_NODES, 4 * (uint64_t)n_nodes);
    float* btris = S.as<float>(S_BVH_TRIS, 9 * n);
    BS_HIP(hipMemsetAsync(parent, 0xFF, sizeof(int) * n_nodes, st));
    BS_HIP(hipMemsetAsync(flags, 0, sizeof(unsigned) * std::max(n_int, 1), st));
    if (n_int > 0) k_lbvh_build<<<div_up(n_int, TB), TB, 0, st>>>(codes_s, nt, links, parent);
    k_lbvh_leaves<<<div_up(n, TB), TB, 0, st>>>(d_tris, nt, idx_s, prm.wn_accuracy, links, parent, flags, accs, nodes,
                                                btris, leaf_links, err);
    tm.mark();

    // ---- 6. winding numbers and signs
    uint32_t* recheck = S.as<uint32_t>(S_RECHECK, n_act);
    if (n_act > 0)
        k_winding<<<div_up(n_act, WN_BLOCK), WN_BLOCK, 0, st>>>(act, n_act, vol->origins, h, nodes, links, leaf_links,
                                                                n_int, n_nodes, btris, prm.sign_threshold, prm.fp64_band,
                                                                vol->values, recheck, counters, err);
    tm.mark();
    BS_HIP(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, st));
    BS_HIP(hipStreamSynchronize(st));
    const uint32_t n_re = (uint32_t)hc[C_RECHECK];
    if (n_re > 0)
        k_winding_fp64<<<div_up(n_re, 64), 64, 0, st>>>(recheck, n_re, vol->origins, h, nodes, links, leaf_links, n_int,
                                                        n_nodes, btris, prm.sign_threshold, vol->values, err);
    unsigned e = read_scalar(ctx, err);
    if (e & E_STACK) throw Error(BS_ERR_INTERNAL, "winding-number traversal stack overflow");
    if (e & E_TREE) throw Error(BS_ERR_INTERNAL, "BVH parent links are corrupt");
    auto wall1 = std::chrono::steady_clock::now();

    bs_m2v_stats& stt = ctx->last_m2v;
    stt = {};
    stt.n_tris = n;
    stt.n_sub = n_sub;
    stt.n_pairs = n_pairs;
    stt.n_leaves = n_leaves;
    stt.n_active = n_act;
    stt.n_candidates = n_cov;
    stt.n_bvh_nodes = (uint64_t)n_nodes;
    stt.wn_node_visits = hc[C_NEAR_INTERNAL];
    stt.wn_far_evals = hc[C_FAR];
    stt.wn_leaf_tris = hc[C_LEAF_TRIS];
    stt.n_fp64_rechecks = n_re;
    stt.ms_total = std::chrono::duration<double, std::milli>(wall1 - wall0).count();
    stt.ms_subdivide = tm.ms(0, 1);
    stt.ms_leaves = tm.ms(1, 2);
    stt.ms_distance = tm.ms(2, 3);
    stt.ms_bvh = tm.ms(3, 4);
    stt.ms_winding = tm.ms(4, 5);
    *out = guard.release();
}

}  // namespace bs
