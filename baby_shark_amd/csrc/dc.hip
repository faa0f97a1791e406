// Dual contouring on the device (DualContouringMesher::mesh,
// src/voxel/meshing/dual_contouring.rs:23-83).
//
// The reference runs three visitor passes and keeps the edge crossings and
// the cell points in sparse trees behind locks.  Here:
//   * cell pass: one wavefront per leaf stages the leaf and a one-to-two voxel
//     halo (an 11^3 block from the 27 surrounding leaves) in LDS; every lane
//     owns one z-row.  It checks the gradient of every crossing edge (the
//     reference's unreachable!() case), and for every fully active
//     sign-changing cube recomputes the crossings of its edges on the fly
//     (deterministic, so bit-identical to storing them) and runs the damped
//     particle iteration.  Cell points go to a per-leaf 8^3 block plus a
//     512-bit presence mask.
//   * triangle pass: the same staging; each active voxel emits two triangles
//     per sign-changing edge whose four cells exist, scaled by the voxel size,
//     degenerate ones dropped.  A counting launch, a scan and a writing
//     launch give the sequential order: leaves by origin, voxels x-y-z, edges
//     X-Y-Z.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "leafmap.hpp"

namespace bs {

Topology& ensure_topology(bs_ctx* ctx, bs_volume* v);

namespace {

constexpr int HB = 11;  // staged block edge: leaf-local -1 .. 9
constexpr int HB3 = HB * HB * HB;

struct DcArgs {
    LeafMap map;
    const int4* origins;
    const uint64_t* masks;
    const float* values;
    float* cell_pts;        // n_leaves * 512 * 3
    uint64_t* cell_mask;    // n_leaves * 8
    float voxel_size;
    uint32_t* fail;         // bit 0: gradient without neighbours
};

struct Block {
    float val[HB3];
    uint8_t act[HB3];
    int32_t nb[27];
};

__device__ inline int bidx(int x, int y, int z) { return ((x + 1) * HB + (y + 1)) * HB + (z + 1); }

// Stages the values and activity of leaf-local voxels -1..9 on each axis.
__device__ inline void stage_block(Block& B, const DcArgs& a, int4 org, int lane) {
    if (lane < 27) {
        const int dx = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dz = lane % 3 - 1;
        B.nb[lane] = a.map.find((org.x >> 3) + dx, (org.y >> 3) + dy, (org.z >> 3) + dz);
    }
    __syncthreads();
    for (int i = lane; i < HB3; i += 64) {
        const int bx = i / (HB * HB), by = (i / HB) % HB, bz = i % HB;
        const int x = bx - 1, y = by - 1, z = bz - 1;
        const int dx = x >> 3, dy = y >> 3, dz = z >> 3;  // -1, 0 or 1
        const int32_t s = B.nb[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)];
        float v = 0.0f;
        uint8_t on = 0;
        if (s >= 0) {
            const int o = ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
            on = (a.masks[(uint64_t)s * 8 + (o >> 6)] >> (o & 63)) & 1ull;
            if (on) v = a.values[(uint64_t)s * 512 + o];
        }
        B.val[i] = v;
        B.act[i] = on;
    }
    __syncthreads();
}

// dual_contouring.rs:332-349; false where the reference reaches unreachable!()
__device__ inline bool grad(const Block& B, int x, int y, int z, int axis, float& g) {
    const int c = bidx(x, y, z);
    const int step = axis == 0 ? HB * HB : (axis == 1 ? HB : 1);
    const bool hp = B.act[c], hl = B.act[c - step], hr = B.act[c + step];
    if (hp && hl && hr) {
        g = (B.val[c + step] - B.val[c - step]) * 0.5f;
        return true;
    }
    if (hp && hr) {
        g = B.val[c + step] - B.val[c];
        return true;
    }
    if (hp && hl) {
        g = B.val[c] - B.val[c - step];
        return true;
    }
    return false;
}

// dual_contouring.rs:257-330: crossing of edge (v1, v1 + e_dir) at leaf-local
// v1, in voxel units of the global grid.  Both ends active, signs differ.
__device__ inline void crossing(const Block& B, int4 org, int x, int y, int z, int dir, float p[3], float n[3]) {
    const int c1 = bidx(x, y, z);
    const int c2 = c1 + (dir == 0 ? HB * HB : (dir == 1 ? HB : 1));
    const float v1 = B.val[c1], v2 = B.val[c2];
    const float t = (v1 == v2) ? 0.5f : v1 / (v1 - v2);
    const float px = (float)(org.x + x), py = (float)(org.y + y), pz = (float)(org.z + z);
    p[0] = dir == 0 ? px + t : px;
    p[1] = dir == 1 ? py + t : py;
    p[2] = dir == 2 ? pz + t : pz;
    const int x2 = x + (dir == 0), y2 = y + (dir == 1), z2 = z + (dir == 2);
    float nv[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        float g1 = 0.0f, g2 = 0.0f;
        grad(B, x, y, z, ax, g1);  // availability was checked in the edge sweep
        grad(B, x2, y2, z2, ax, g2);
        nv[ax] = (1.0f - t) * g1 + t * g2;
    }
    const float len = sqrt_rn((nv[0] * nv[0] + nv[1] * nv[1]) + nv[2] * nv[2]);
    n[0] = nv[0] / len;
    n[1] = nv[1] / len;
    n[2] = nv[2] / len;
}

// dual_contouring.rs:410-423 (base offset, direction) and utils.rs:86-95
__constant__ int8_t kEdge[12][4] = {{0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 2}, {1, 0, 0, 1}, {1, 0, 0, 2}, {0, 0, 1, 0},
                                    {0, 0, 1, 1}, {0, 1, 0, 0}, {0, 1, 0, 2}, {1, 0, 1, 1}, {0, 1, 1, 0}, {1, 1, 0, 2}};
__constant__ int8_t kCube[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
// dual_contouring.rs:389-408
__constant__ int8_t kCell[3][4][3] = {{{0, 0, 0}, {0, 0, -1}, {0, -1, -1}, {0, -1, 0}},
                                      {{0, 0, 0}, {-1, 0, 0}, {-1, 0, -1}, {0, 0, -1}},
                                      {{0, -1, 0}, {-1, -1, 0}, {-1, 0, 0}, {0, 0, 0}}};

__global__ __launch_bounds__(64) void k_dc_cells(uint64_t n_leaves, DcArgs a) {
    __shared__ Block B;
    __shared__ uint16_t cell_list[512];  // offset x*64 + y*8 + z of each listed cell
    const int lane = threadIdx.x;
    const int x = lane >> 3, y = lane & 7;
    for (uint64_t leaf = blockIdx.x; leaf < n_leaves; leaf += gridDim.x) {
        const int4 org = a.origins[leaf];
        stage_block(B, a, org, lane);
        // edge sweep (dual_contouring.rs:352-387): every crossing edge from a
        // voxel of this leaf needs all six gradients
        bool bad = false;
        for (int z = 0; z < 8; ++z) {
            const int c = bidx(x, y, z);
            if (!B.act[c]) continue;
            const uint32_t s1 = __float_as_uint(B.val[c]) >> 31;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int c2 = c + (d == 0 ? HB * HB : (d == 1 ? HB : 1));
                if (!B.act[c2] || (__float_as_uint(B.val[c2]) >> 31) == s1) continue;
                float g;
                const int x2 = x + (d == 0), y2 = y + (d == 1), z2 = z + (d == 2);
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) bad |= !grad(B, x, y, z, ax, g) || !grad(B, x2, y2, z2, ax, g);
            }
        }
        if (bad) atomicOr(a.fail, 1u);
        // cells (dual_contouring.rs:176-247): the leaf's cells with all eight
        // corners active and mixed signs are listed first, then each lane
        // takes one listed cell, so the feature-point work runs on full waves
        uint32_t present = 0;
        int n_cells = 0;
        for (int z = 0; z < 8; ++z) {
            bool all = true;
            uint32_t neg = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int c = bidx(x + kCube[i][0], y + kCube[i][1], z + kCube[i][2]);
                all &= B.act[c] != 0;
                neg |= (__float_as_uint(B.val[c]) >> 31) << i;
            }
            const bool surface = all && neg != 0 && neg != 0xFFu;
            const unsigned long long b = __ballot(surface);
            if (surface) {
                present |= 1u << z;
                cell_list[n_cells + __popcll(b & ((1ull << lane) - 1))] = (uint16_t)((lane << 3) | z);
            }
            n_cells += __popcll(b);
        }
        __syncthreads();
        for (int k = lane; k < n_cells; k += 64) {
            const int cell = cell_list[k];
            const int x = cell >> 6, y = (cell >> 3) & 7, z = cell & 7;
            // crossings in EDGE_OFFSETS order; all indexing static so they stay in registers
            float P[12][3], N[12][3];
            uint32_t valid = 0;
#pragma unroll
            for (int e = 0; e < 12; ++e) {
                const int ex = x + kEdge[e][0], ey = y + kEdge[e][1], ez = z + kEdge[e][2], d = kEdge[e][3];
                const int c1 = bidx(ex, ey, ez);
                const int c2 = c1 + (d == 0 ? HB * HB : (d == 1 ? HB : 1));
                P[e][0] = P[e][1] = P[e][2] = 0.0f;
                N[e][0] = N[e][1] = N[e][2] = 0.0f;
                if ((__float_as_uint(B.val[c1]) >> 31) == (__float_as_uint(B.val[c2]) >> 31)) continue;
                crossing(B, org, ex, ey, ez, d, P[e], N[e]);
                valid |= 1u << e;
            }
            // find_feature_point (dual_contouring.rs:433-458)
            const float fc = (float)__popc(valid);
            float cx = 0.0f, cy = 0.0f, cz = 0.0f;
#pragma unroll
            for (int e = 0; e < 12; ++e)
                if (valid & (1u << e)) cx = cx + P[e][0], cy = cy + P[e][1], cz = cz + P[e][2];
            cx = cx / fc, cy = cy / fc, cz = cz / fc;
            for (int it = 0; it < 50; ++it) {
                float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
                for (int e = 0; e < 12; ++e) {
                    if (!(valid & (1u << e))) continue;
                    const float d = (N[e][0] * (cx - P[e][0]) + N[e][1] * (cy - P[e][1])) + N[e][2] * (cz - P[e][2]);
                    fx = fx + (N[e][0] * -1.0f) * d;
                    fy = fy + (N[e][1] * -1.0f) * d;
                    fz = fz + (N[e][2] * -1.0f) * d;
                }
                const float damping = 1.0f - (float)it / 50.0f;
                cx = cx + (fx * damping) / fc;
                cy = cy + (fy * damping) / fc;
                cz = cz + (fz * damping) / fc;
                if ((fx * fx + fy * fy) + fz * fz < 1e-6f) break;
            }
            float* out = a.cell_pts + (leaf * 512 + (uint64_t)cell) * 3;
            out[0] = cx;
            out[1] = cy;
            out[2] = cz;
        }
        // cell presence mask: word x, bits y*8+z
        const uint64_t w = gather_word(present, y);
        if (y == 0) a.cell_mask[leaf * 8 + x] = w;
        __syncthreads();
    }
}

// Cell point at leaf-local (x, y, z) in -1..7, through the 27-leaf table.
__device__ inline bool cell_at(const Block& B, const DcArgs& a, int x, int y, int z, float p[3]) {
    const int dx = x >> 3, dy = y >> 3, dz = z >> 3;
    const int32_t s = B.nb[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)];
    if (s < 0) return false;
    const int o = ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
    if (!((a.cell_mask[(uint64_t)s * 8 + (o >> 6)] >> (o & 63)) & 1ull)) return false;
    const float* q = a.cell_pts + ((uint64_t)s * 512 + o) * 3;
    p[0] = q[0];
    p[1] = q[1];
    p[2] = q[2];
    return true;
}

// Triangle3::is_degenerate (triangle3.rs:272-275) of the scaled corners
__device__ inline bool degenerate(const float a[3], const float b[3], const float c[3]) {
    const float ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    const float vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
    const float cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
    return (cx * cx + cy * cy) + cz * cz == 0.0f;
}

// The quad of crossing edge d from voxel (x, y, z) with sign bit s1
// (dual_contouring.rs:93-174, :62-79): its two triangles, scaled, in q, and
// whether each survives the degenerate-triangle filter; false when one of the
// four cells around the edge is missing.
__device__ inline bool edge_quad(const Block& B, const DcArgs& a, int x, int y, int z, int d, uint32_t s1,
                                 float q[6][3], bool keep[2]) {
    float p[4][3];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        ok = ok && cell_at(B, a, x + kCell[d][k][0], y + kCell[d][k][1], z + kCell[d][k][2], p[k]);
    if (!ok) return false;
    // [p0, p1, p2, p2, p3, p0], with 1<->2 and 4<->5 swapped when v is negative
    const int f1 = s1 ? 2 : 1, f2 = s1 ? 1 : 2, f4 = s1 ? 0 : 3, f5 = s1 ? 3 : 0;
    const int f[6] = {0, f1, f2, 2, f4, f5};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int src = f[k];
#pragma unroll
        for (int j = 0; j < 3; ++j)
            q[k][j] = (src == 0 ? p[0][j] : src == 1 ? p[1][j] : src == 2 ? p[2][j] : p[3][j]) * a.voxel_size;
    }
    keep[0] = !degenerate(q[0], q[1], q[2]);
    keep[1] = !degenerate(q[3], q[4], q[5]);
    return true;
}

// Crossing edges from the voxels of this lane's z-row, in the sequential
// order (z, then axis); emit(z, d) is called for each.
template <class F>
__device__ inline void row_edges(const Block& B, int x, int y, F&& emit) {
    for (int z = 0; z < 8; ++z) {
        const int c = bidx(x, y, z);
        if (!B.act[c]) continue;
        const uint32_t s1 = __float_as_uint(B.val[c]) >> 31;
        for (int d = 0; d < 3; ++d) {
            const int c2 = c + (d == 0 ? HB * HB : (d == 1 ? HB : 1));
            if (!B.act[c2] || (__float_as_uint(B.val[c2]) >> 31) == s1) continue;
            emit(z, d);
        }
    }
}

// WRITE = false: triangles per leaf into tri_count.  WRITE = true: the
// triangles at tri_off[leaf].  A leaf's crossing edges are listed in the
// sequential order (z-rows = lanes, then z, then axis), then taken 64 at a
// time, one per lane; each lane's triangles go after those of the edges
// before it (two ballots give the prefix), so the soup keeps that order.
template <bool WRITE>
__global__ __launch_bounds__(64) void k_dc_triangles(uint64_t n_leaves, DcArgs a, uint32_t* __restrict__ tri_count,
                                                     const uint64_t* __restrict__ tri_off, float* __restrict__ out) {
    __shared__ Block B;
    __shared__ uint32_t lane_cnt[64];
    __shared__ uint16_t edges[512 * 3];  // lane << 5 | z << 2 | axis
    const int lane = threadIdx.x;
    const int x = lane >> 3, y = lane & 7;
    const unsigned long long lt = (1ull << lane) - 1;
    for (uint64_t leaf = blockIdx.x; leaf < n_leaves; leaf += gridDim.x) {
        stage_block(B, a, a.origins[leaf], lane);
        uint32_t cnt = 0;
        row_edges(B, x, y, [&](int, int) { ++cnt; });
        lane_cnt[lane] = cnt;
        __syncthreads();
        uint32_t before = 0, n_edges = 0;
        for (int l = 0; l < 64; ++l) {
            before += l < lane ? lane_cnt[l] : 0;
            n_edges += lane_cnt[l];
        }
        row_edges(B, x, y, [&](int z, int d) { edges[before++] = (uint16_t)((lane << 5) | (z << 2) | d); });
        __syncthreads();
        const uint64_t base = WRITE ? tri_off[leaf] : 0;
        uint32_t done = 0;
        for (uint32_t k0 = 0; k0 < n_edges; k0 += 64) {
            const uint32_t k = k0 + (uint32_t)lane;
            float q[6][3];
            bool keep[2] = {false, false};
            if (k < n_edges) {
                const int e = edges[k];
                const int el = e >> 5, ex = el >> 3, ey = el & 7, ez = (e >> 2) & 7, d = e & 3;
                const uint32_t s1 = __float_as_uint(B.val[bidx(ex, ey, ez)]) >> 31;
                if (!edge_quad(B, a, ex, ey, ez, d, s1, q, keep)) keep[0] = keep[1] = false;
            }
            const unsigned long long b0 = __ballot(keep[0]), b1 = __ballot(keep[1]);
            if (WRITE) {
                uint64_t at = base + done + (uint64_t)(__popcll(b0 & lt) + __popcll(b1 & lt));
#pragma unroll
                for (int tr = 0; tr < 2; ++tr) {
                    if (!keep[tr]) continue;
                    float* o = out + at * 9;
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk)
#pragma unroll
                        for (int j = 0; j < 3; ++j) o[kk * 3 + j] = q[3 * tr + kk][j];
                    ++at;
                }
            }
            done += (uint32_t)(__popcll(b0) + __popcll(b1));
        }
        if (!WRITE && lane == 0) tri_count[leaf] = done;
        __syncthreads();
    }
}

void exclusive_scan_u64(bs_ctx* ctx, const uint32_t* in, uint64_t* out, uint64_t n) {
    size_t bytes = 0;
    BS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, ctx->stream));
    void* tmp = ctx->scratch.get(S_CUB, bytes);
    BS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, ctx->stream));
}

}  // namespace

// DualContouringMesher::default().with_voxel_size(h).mesh(volume): the
// triangle soup (n_tris * 9 floats) in device memory owned by the caller
// (hipFree), or nullptr for no triangle.
void dual_contour_device(bs_ctx* ctx, bs_volume* vol, float voxel_size, float** out_dev, uint64_t* n_tris,
                         bs_dc_stats* stats) {
    hipStream_t st = ctx->stream;
    *out_dev = nullptr;
    *n_tris = 0;
    BS_REQUIRE(vol->device == ctx->device, BS_ERR_INVALID_ARG, "volume lives on another device");
    // visit_leafs_par meets every active tile: todo!() (dual_contouring.rs:146,181,354)
    BS_REQUIRE(vol->tiles.empty(), BS_ERR_UNSUPPORTED_TILE,
               "dual contouring of a volume with active tiles: the reference stops at todo!()");
    const uint64_t n = vol->n_leaves;
    if (stats) *stats = bs_dc_stats{};
    if (!n) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (stats) {
        BS_HIP(hipEventCreate(&e0));
        BS_HIP(hipEventCreate(&e1));
        BS_HIP(hipEventRecord(e0, st));
    }

    // leaf map over the leaves' box (one leaf of margin for the halo lookups)
    int32_t lo[3], hi[3];
    leaf_box(ctx, vol->origins, n, lo, hi);
    LeafMap map{};
    uint64_t cells = 1;
    for (int k = 0; k < 3; ++k) cells *= (uint64_t)(hi[k] - lo[k] + 1);
    map.dense = !(ctx->paths & BS_PATH_FORCE_HASH) && cells <= (1ull << 28);
    if (map.dense) {
        map.x0 = lo[0], map.y0 = lo[1], map.z0 = lo[2];
        map.nx = hi[0] - lo[0] + 1, map.ny = hi[1] - lo[1] + 1, map.nz = hi[2] - lo[2] + 1;
        map.cells = ctx->scratch.as<int32_t>(S_DC_MAP, cells);
        BS_HIP(hipMemsetAsync(map.cells, 0xFF, sizeof(int32_t) * cells, st));
    } else {
        uint64_t cap = 1024;
        while (cap < 2 * n) cap <<= 1;
        map.cap_mask = (uint32_t)(cap - 1);
        map.keys = ctx->scratch.as<uint64_t>(S_DC_KEYS, cap);
        map.cells = ctx->scratch.as<int32_t>(S_DC_MAP, cap);
        BS_HIP(hipMemsetAsync(map.keys, 0xFF, sizeof(uint64_t) * cap, st));
        BS_HIP(hipMemsetAsync(map.cells, 0xFF, sizeof(int32_t) * cap, st));
    }
    uint32_t* fail = ctx->scratch.as<uint32_t>(S_DC_FAIL, 1);
    BS_HIP(hipMemsetAsync(fail, 0, sizeof(uint32_t), st));
    k_leafmap_insert<<<div_up(n, 256), 256, 0, st>>>(n, reinterpret_cast<const int4*>(vol->origins), map, fail);
    BS_HIP(hipGetLastError());

    DcArgs a{};
    a.map = map;
    a.origins = reinterpret_cast<const int4*>(vol->origins);
    a.masks = vol->masks;
    a.values = vol->values;
    a.cell_pts = ctx->scratch.as<float>(S_DC_CELLS, n * 512 * 3);
    a.cell_mask = ctx->scratch.as<uint64_t>(S_DC_CELLMASK, n * 8);
    a.voxel_size = voxel_size;
    a.fail = fail;
    const unsigned grid = (unsigned)std::min<uint64_t>(n, 8192);
    k_dc_cells<<<grid, 64, 0, st>>>(n, a);
    BS_HIP(hipGetLastError());
    uint32_t* tri_count = ctx->scratch.as<uint32_t>(S_DC_TRICNT, n);
    uint64_t* tri_off = ctx->scratch.as<uint64_t>(S_DC_TRIOFF, n + 1);
    k_dc_triangles<false><<<grid, 64, 0, st>>>(n, a, tri_count, nullptr, nullptr);
    BS_HIP(hipGetLastError());
    exclusive_scan_u64(ctx, tri_count, tri_off, n);
    uint32_t h_fail = 0, last_cnt = 0;
    uint64_t last_off = 0;
    BS_HIP(hipMemcpyAsync(&h_fail, fail, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    BS_HIP(hipMemcpyAsync(&last_cnt, tri_count + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    BS_HIP(hipMemcpyAsync(&last_off, tri_off + n - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    BS_HIP(hipStreamSynchronize(st));
    BS_REQUIRE(!(h_fail & 1u), BS_ERR_INTERNAL,
               "dual contouring: a crossing voxel has no neighbour along an axis; the reference reaches "
               "unreachable!() (dual_contouring.rs:347)");
    const uint64_t total = last_off + last_cnt;
    if (total) {
        float* out = nullptr;
        BS_HIP(hipMalloc(&out, sizeof(float) * 9 * total));
        k_dc_triangles<true><<<grid, 64, 0, st>>>(n, a, nullptr, tri_off, out);
        hipError_t le = hipGetLastError();
        if (le == hipSuccess) le = hipStreamSynchronize(st);
        if (le != hipSuccess) {
            (void)hipFree(out);
            BS_HIP(le);
        }
        *out_dev = out;
    }
    *n_tris = total;
    if (stats) {
        BS_HIP(hipEventRecord(e1, st));
        BS_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        BS_HIP(hipEventElapsedTime(&ms, e0, e1));
        stats->ms_total = ms;
        stats->n_leaves = n;
        stats->n_tris = total;
        stats->used_hash = !map.dense;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
}

}  // namespace bs
