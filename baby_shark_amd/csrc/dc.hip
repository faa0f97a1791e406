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
#include <mutex>
#include <string>

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

// Opt-in vertex placement (BS_DC_VERTEX_QEF_F64; the reference has only the
// particle iteration): the point minimising sum_i (n_i . (x - p_i))^2 over the
// cell's edge crossings p_i with unit normals n_i, in FP64.  Relative to the
// mass point m (the crossings' mean), A = sum n n^T and b = sum n (n . (p - m));
// A's eigenvectors come from cyclic Jacobi rotations and x = m + sum_k
// (v_k . b / l_k) v_k over the eigenvalues l_k > QEF_TRUNC * l_max (a plane
// fixes one direction, an edge two, a corner three; the rest stays at m).
// The point is clamped to the cell.
constexpr double QEF_TRUNC = 0.1;

// One Jacobi rotation zeroing A[p][q] of the symmetric A (Numerical Recipes'
// form: theta = (A_qq - A_pp) / (2 A_pq), t = sgn(theta) / (|theta| +
// sqrt(theta^2 + 1))), accumulated into the eigenvector columns of V.
template <int P, int Q>
__device__ inline void jacobi_rotate(double (&A)[3][3], double (&V)[3][3]) {
    const double apq = A[P][Q];
    if (apq == 0.0) return;
    const double theta = (A[Q][Q] - A[P][P]) / (2.0 * apq);
    const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
    constexpr int R = 3 - P - Q;  // the third index
    const double arp = A[R][P], arq = A[R][Q];
    A[P][P] -= t * apq;
    A[Q][Q] += t * apq;
    A[P][Q] = A[Q][P] = 0.0;
    A[R][P] = A[P][R] = c * arp - s * arq;
    A[R][Q] = A[Q][R] = s * arp + c * arq;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const double vp = V[r][P], vq = V[r][Q];
        V[r][P] = c * vp - s * vq;
        V[r][Q] = s * vp + c * vq;
    }
}

__device__ inline void qef_point(const float (&P)[12][3], const float (&N)[12][3], uint32_t valid, float cx0,
                                 float cy0, float cz0, float out[3]) {
    double m[3] = {0.0, 0.0, 0.0};
    const double cnt = (double)__popc(valid);
#pragma unroll
    for (int e = 0; e < 12; ++e)
        if (valid & (1u << e))
#pragma unroll
            for (int k = 0; k < 3; ++k) m[k] += (double)P[e][k];
#pragma unroll
    for (int k = 0; k < 3; ++k) m[k] /= cnt;
    double A[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}}, b[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        if (!(valid & (1u << e))) continue;
        const double n0 = N[e][0], n1 = N[e][1], n2 = N[e][2];
        const double d = n0 * ((double)P[e][0] - m[0]) + n1 * ((double)P[e][1] - m[1]) + n2 * ((double)P[e][2] - m[2]);
        A[0][0] += n0 * n0, A[0][1] += n0 * n1, A[0][2] += n0 * n2;
        A[1][1] += n1 * n1, A[1][2] += n1 * n2, A[2][2] += n2 * n2;
        b[0] += n0 * d, b[1] += n1 * d, b[2] += n2 * d;
    }
    A[1][0] = A[0][1], A[2][0] = A[0][2], A[2][1] = A[1][2];
    double V[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    for (int sweep = 0; sweep < 8; ++sweep) {
        jacobi_rotate<0, 1>(A, V);
        jacobi_rotate<0, 2>(A, V);
        jacobi_rotate<1, 2>(A, V);
    }
    const double lmax = fmax(fmax(A[0][0], A[1][1]), A[2][2]);
    double x[3] = {m[0], m[1], m[2]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double l = A[k][k];
        if (!(l > QEF_TRUNC * lmax)) continue;
        const double w = (V[0][k] * b[0] + V[1][k] * b[1] + V[2][k] * b[2]) / l;
#pragma unroll
        for (int r = 0; r < 3; ++r) x[r] += w * V[r][k];
    }
    // the cell spans [c0, c0 + 1] on each axis (voxel units)
    out[0] = (float)fmin(fmax(x[0], (double)cx0), (double)cx0 + 1.0);
    out[1] = (float)fmin(fmax(x[1], (double)cy0), (double)cy0 + 1.0);
    out[2] = (float)fmin(fmax(x[2], (double)cz0), (double)cz0 + 1.0);
}

template <bool QEF>
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
            float* out = a.cell_pts + (leaf * 512 + (uint64_t)cell) * 3;
            if constexpr (QEF) {
                qef_point(P, N, valid, (float)(org.x + x), (float)(org.y + y), (float)(org.z + z), out);
                continue;
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

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Fills n 32-bit words with a kernel (not hipMemsetAsync, so that the call
// stays valid inside a caller's graph capture).
void fill_u32(hipStream_t st, void* p, uint64_t n, uint32_t v) {
    k_fill_u32<<<(unsigned)std::min<uint64_t>(div_up(n, 256), 2048), 256, 0, st>>>(static_cast<uint32_t*>(p), n, v);
    BS_HIP(hipGetLastError());
}

// Device buffers of one dual-contouring call.  The usual entry point takes
// them from the context's scratch arena; bs_dual_contour_into from the
// per-device static workspace below.
struct DcWork {
    LeafMap map;
    uint32_t* fail;
    float* cell_pts;
    uint64_t* cell_mask;
    uint32_t* tri_count;
    uint64_t* tri_off;
    void* scan_tmp;
    size_t scan_bytes;
};

uint64_t hash_capacity(uint64_t n) {
    uint64_t cap = 1024;
    while (cap < 2 * n) cap <<= 1;
    return cap;
}

// Everything up to the triangle count: leaf map, cell points, per-leaf
// triangle counts and their scan.  Returns the soup's triangle count.
uint64_t dc_count(hipStream_t st, const bs_volume* vol, float voxel_size, bool qef, DcWork& w, DcArgs& a) {
    const uint64_t n = vol->n_leaves;
    if (w.map.dense) {
        fill_u32(st, w.map.cells, (uint64_t)w.map.nx * w.map.ny * w.map.nz, 0xFFFFFFFFu);
    } else {
        const uint64_t cap = (uint64_t)w.map.cap_mask + 1;
        fill_u32(st, w.map.keys, 2 * cap, 0xFFFFFFFFu);
        fill_u32(st, w.map.cells, cap, 0xFFFFFFFFu);
    }
    fill_u32(st, w.fail, 1, 0);
    k_leafmap_insert<<<div_up(n, 256), 256, 0, st>>>(n, reinterpret_cast<const int4*>(vol->origins), w.map, w.fail);
    BS_HIP(hipGetLastError());

    a = DcArgs{};
    a.map = w.map;
    a.origins = reinterpret_cast<const int4*>(vol->origins);
    a.masks = vol->masks;
    a.values = vol->values;
    a.cell_pts = w.cell_pts;
    a.cell_mask = w.cell_mask;
    a.voxel_size = voxel_size;
    a.fail = w.fail;
    const unsigned grid = (unsigned)std::min<uint64_t>(n, 8192);
    if (qef)
        k_dc_cells<true><<<grid, 64, 0, st>>>(n, a);
    else
        k_dc_cells<false><<<grid, 64, 0, st>>>(n, a);
    BS_HIP(hipGetLastError());
    k_dc_triangles<false><<<grid, 64, 0, st>>>(n, a, w.tri_count, nullptr, nullptr);
    BS_HIP(hipGetLastError());
    size_t bytes = w.scan_bytes;
    BS_HIP(hipcub::DeviceScan::ExclusiveSum(w.scan_tmp, bytes, w.tri_count, w.tri_off, (int)n, st));
    uint32_t h_fail = 0, last_cnt = 0;
    uint64_t last_off = 0;
    BS_HIP(hipMemcpyAsync(&h_fail, w.fail, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    BS_HIP(hipMemcpyAsync(&last_cnt, w.tri_count + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    BS_HIP(hipMemcpyAsync(&last_off, w.tri_off + n - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    BS_HIP(hipStreamSynchronize(st));
    BS_REQUIRE(!(h_fail & 1u), BS_ERR_INTERNAL,
               "dual contouring: a crossing voxel has no neighbour along an axis; the reference reaches "
               "unreachable!() (dual_contouring.rs:347)");
    return last_off + last_cnt;
}

void dc_write(hipStream_t st, uint64_t n, const DcArgs& a, const DcWork& w, float* out) {
    const unsigned grid = (unsigned)std::min<uint64_t>(n, 8192);
    k_dc_triangles<true><<<grid, 64, 0, st>>>(n, a, nullptr, w.tri_off, out);
    BS_HIP(hipGetLastError());
}

void check_meshable(bs_ctx* ctx, const bs_volume* vol) {
    BS_REQUIRE(vol->device == ctx->device, BS_ERR_INVALID_ARG, "volume lives on another device");
    // visit_leafs_par meets every active tile: todo!() (dual_contouring.rs:146,181,354)
    BS_REQUIRE(vol->tiles.empty(), BS_ERR_UNSUPPORTED_TILE,
               "dual contouring of a volume with active tiles: the reference stops at todo!()");
}

// ---------------------------------------------------------------- workspace of bs_dual_contour_into
// bs_dual_contour_into allocates nothing: its buffers live in this array,
// which each device holds from the library's load until the process ends.
// Sized for DC_NOALLOC_MAX_LEAVES leaves with the hash leaf map.
constexpr uint64_t DC_NOALLOC_MAX_LEAVES = BS_DC_INTO_MAX_LEAVES;
constexpr uint64_t DC_NOALLOC_HASH_CAP = 2 * DC_NOALLOC_MAX_LEAVES;  // power of two >= 1024
constexpr uint64_t DC_NOALLOC_SCAN_BYTES = 1ull << 20;
constexpr uint64_t al256(uint64_t b) { return (b + 255) & ~255ull; }
constexpr uint64_t NA_KEYS = 0;
constexpr uint64_t NA_CELLS = NA_KEYS + al256(8 * DC_NOALLOC_HASH_CAP);
constexpr uint64_t NA_FAIL = NA_CELLS + al256(4 * DC_NOALLOC_HASH_CAP);
constexpr uint64_t NA_PTS = NA_FAIL + 256;
constexpr uint64_t NA_CMASK = NA_PTS + al256(4ull * 512 * 3 * DC_NOALLOC_MAX_LEAVES);
constexpr uint64_t NA_TCNT = NA_CMASK + al256(64ull * DC_NOALLOC_MAX_LEAVES);
constexpr uint64_t NA_TOFF = NA_TCNT + al256(4ull * DC_NOALLOC_MAX_LEAVES);
constexpr uint64_t NA_SCAN = NA_TOFF + al256(8ull * (DC_NOALLOC_MAX_LEAVES + 1));
constexpr uint64_t NA_BYTES = NA_SCAN + DC_NOALLOC_SCAN_BYTES;
static_assert(DC_NOALLOC_HASH_CAP >= 1024 && (DC_NOALLOC_HASH_CAP & (DC_NOALLOC_HASH_CAP - 1)) == 0, "hash capacity");

}  // namespace

// The array itself.  In host code this name is a stand-in, not a device
// address: host code uses the address hipGetSymbolAddress returned for the
// device (dc_noalloc_ptr), never the name.
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {
constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mutex;
unsigned char* g_na_ptr[NA_MAX_DEVICES] = {};
}  // namespace

// Looks up this device's copy of the workspace once, with `device` current.
// Called by bs_ctx_create, so the entry point itself needs no set-up call.
void dc_noalloc_init(int device) {
    BS_REQUIRE(device >= 0 && device < NA_MAX_DEVICES, BS_ERR_INVALID_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (g_na_ptr[device]) return;
    void* p = nullptr;
    BS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BS_REQUIRE(p != nullptr, BS_ERR_HIP, "hipGetSymbolAddress returned no address for the DC workspace");
    g_na_ptr[device] = static_cast<unsigned char*>(p);
}

static unsigned char* dc_noalloc_ptr(int device) {
    std::lock_guard<std::mutex> lock(g_na_mutex);
    BS_REQUIRE(device >= 0 && device < NA_MAX_DEVICES && g_na_ptr[device], BS_ERR_INTERNAL,
               "the DC workspace of this device was not set up (bs_ctx_create does it)");
    return g_na_ptr[device];
}

// DualContouringMesher::default().with_voxel_size(h).mesh(volume): the
// triangle soup (n_tris * 9 floats) in device memory owned by the caller
// (hipFree), or nullptr for no triangle.
void dual_contour_device(bs_ctx* ctx, bs_volume* vol, float voxel_size, float** out_dev, uint64_t* n_tris,
                         bs_dc_stats* stats) {
    hipStream_t st = ctx->stream;
    *out_dev = nullptr;
    *n_tris = 0;
    check_meshable(ctx, vol);
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
    DcWork w{};
    uint64_t cells = 1;
    for (int k = 0; k < 3; ++k) cells *= (uint64_t)(hi[k] - lo[k] + 1);
    w.map.dense = !(ctx->paths & BS_PATH_FORCE_HASH) && cells <= (1ull << 28);
    if (w.map.dense) {
        w.map.x0 = lo[0], w.map.y0 = lo[1], w.map.z0 = lo[2];
        w.map.nx = hi[0] - lo[0] + 1, w.map.ny = hi[1] - lo[1] + 1, w.map.nz = hi[2] - lo[2] + 1;
        w.map.cells = ctx->scratch.as<int32_t>(S_DC_MAP, cells);
    } else {
        const uint64_t cap = hash_capacity(n);
        w.map.cap_mask = (uint32_t)(cap - 1);
        w.map.keys = ctx->scratch.as<uint64_t>(S_DC_KEYS, cap);
        w.map.cells = ctx->scratch.as<int32_t>(S_DC_MAP, cap);
    }
    w.fail = ctx->scratch.as<uint32_t>(S_DC_FAIL, 1);
    w.cell_pts = ctx->scratch.as<float>(S_DC_CELLS, n * 512 * 3);
    w.cell_mask = ctx->scratch.as<uint64_t>(S_DC_CELLMASK, n * 8);
    w.tri_count = ctx->scratch.as<uint32_t>(S_DC_TRICNT, n);
    w.tri_off = ctx->scratch.as<uint64_t>(S_DC_TRIOFF, n + 1);
    BS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, w.scan_bytes, w.tri_count, w.tri_off, (int)n, st));
    w.scan_tmp = ctx->scratch.get(S_CUB, w.scan_bytes);

    DcArgs a{};
    const uint64_t total = dc_count(st, vol, voxel_size, ctx->dc_vertex == BS_DC_VERTEX_QEF_F64, w, a);
    if (total) {
        float* out = nullptr;
        BS_HIP(hipMalloc(&out, sizeof(float) * 9 * total));
        hipError_t le = hipSuccess;
        try {
            dc_write(st, n, a, w, out);
        } catch (...) {
            (void)hipFree(out);
            throw;
        }
        le = hipStreamSynchronize(st);
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
        stats->used_hash = !w.map.dense;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
}

// bs_dual_contour_into: the same soup written to the caller's device buffer
// of capacity_tris triangles, with every intermediate buffer in this
// device's static workspace and no device or host allocation.  *n_tris is
// the soup's triangle count; a soup larger than the capacity writes nothing
// and reports BS_ERR_CAPACITY.  Calls that overlap on one device race on the
// shared workspace.
void dual_contour_into(bs_ctx* ctx, const bs_volume* vol, float voxel_size, float* out_dev, uint64_t capacity_tris,
                       uint64_t* n_tris) {
    hipStream_t st = ctx->stream;
    *n_tris = 0;
    check_meshable(ctx, vol);
    const uint64_t n = vol->n_leaves;
    BS_REQUIRE(n <= DC_NOALLOC_MAX_LEAVES, BS_ERR_CAPACITY,
               "bs_dual_contour_into: the volume has more leaves than the workspace holds (" +
                   std::to_string(n) + " > " + std::to_string(DC_NOALLOC_MAX_LEAVES) + ")");
    if (!n) return;
    unsigned char* base = dc_noalloc_ptr(ctx->device);
    DcWork w{};
    const uint64_t cap = hash_capacity(n);
    w.map.dense = 0;
    w.map.cap_mask = (uint32_t)(cap - 1);
    w.map.keys = reinterpret_cast<uint64_t*>(base + NA_KEYS);
    w.map.cells = reinterpret_cast<int32_t*>(base + NA_CELLS);
    w.fail = reinterpret_cast<uint32_t*>(base + NA_FAIL);
    w.cell_pts = reinterpret_cast<float*>(base + NA_PTS);
    w.cell_mask = reinterpret_cast<uint64_t*>(base + NA_CMASK);
    w.tri_count = reinterpret_cast<uint32_t*>(base + NA_TCNT);
    w.tri_off = reinterpret_cast<uint64_t*>(base + NA_TOFF);
    w.scan_tmp = base + NA_SCAN;
    size_t need = 0;
    BS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, w.tri_count, w.tri_off, (int)n, st));
    BS_REQUIRE(need <= DC_NOALLOC_SCAN_BYTES, BS_ERR_INTERNAL, "scan workspace too small");
    w.scan_bytes = need;

    DcArgs a{};
    const uint64_t total = dc_count(st, vol, voxel_size, ctx->dc_vertex == BS_DC_VERTEX_QEF_F64, w, a);
    *n_tris = total;
    BS_REQUIRE(total <= capacity_tris, BS_ERR_CAPACITY,
               "bs_dual_contour_into: the soup has " + std::to_string(total) +
                   " triangles, more than the output capacity of " + std::to_string(capacity_tris));
    if (total) {
        BS_REQUIRE(out_dev != nullptr, BS_ERR_INVALID_ARG, "null output buffer");
        dc_write(st, n, a, w, out_dev);
        BS_HIP(hipStreamSynchronize(st));
    }
}

}  // namespace bs
