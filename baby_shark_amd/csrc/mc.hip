// Marching cubes on the device (MarchingCubesMesher::mesh,
// src/voxel/meshing/marching_cubes.rs:43-63).
//
// The reference visits the leaves and active tiles serially in tree order and
// runs Custodio et al.'s corrected Marching Cubes 33 on every cube whose eight
// corners have values, with Lewiner et al.'s case tables
// (src/voxel/meshing/lookup_table.rs; here the data of mc_tables.inc, written
// from that file by gen_mc_tables.py).  Each cube takes the tiling that
// handle_cube (marching_cubes.rs:72-243) selects with the same face and
// interior tests, and emits its triangles in add_faces' vertex order
// (:245-272).  The centre vertex is mesher state: compute_c_vertex
// (:919-940) sets it in cases 7.3, 10.2, 12.2, 13.3 and 13.4, and case 6.1.2
// uses vertex 12 without computing it, reading the centre left by the last
// earlier cube (in the serial visit order) that computed one, or zero.  When
// a volume holds a 6.1.2 cube, a summary pass finds each unit's last centre
// and a scan hands every unit the centre it starts with.
//
// Device side: the units (leaves and active tiles) in the reference's visit
// order; one wavefront per unit.  A leaf stages its values and the +1 halo
// (9^3) in LDS; lane (x, y) owns the z-row of cubes with that corner, which
// is also the serial order (x, y, z).  A tile meets the cubes of its six
// faces, (i, j) major, faces left, right, top, bottom, front, back (the
// reference's tile visitor, :970-994, which meets edge cubes more than once);
// cubes with every corner inside the tile are all one value and emit
// nothing, so only cubes touching the tile's far faces are evaluated.  A
// counting launch, a scan over units and a writing launch give the serial
// order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "leafmap.hpp"

namespace bs {

Topology& ensure_topology(bs_ctx* ctx, bs_volume* v);

namespace {

// ------------------------------------------------------------------ cube geometry
// device copies: CUBE_OFFSETS (utils.rs:86-95), edges and their axes, the
// test_face corner order a, b, c, d, the interior_ambiguity lists (corner,
// corner, edge named, in test order), and for interior_ambiguity_verification
// (:532-917) the corner pairs (P, Q) of A, B, C, D of the quadratic
// (qA + dA t)(qC + dC t) - (qB + dB t)(qD + dD t), d = P - Q
__constant__ int8_t c_corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__constant__ int8_t c_edge[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
                                     {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
__constant__ int8_t c_axis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
__constant__ int8_t c_face[7][4] = {{0, 0, 0, 0}, {0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3},
                                    {3, 7, 4, 0}, {0, 3, 2, 1}, {4, 7, 6, 5}};
__constant__ int8_t c_amb[3][4][3] = {{{1, 7, 4}, {0, 6, 5}, {3, 5, 6}, {2, 4, 7}},
                                      {{1, 7, 0}, {2, 4, 1}, {3, 5, 2}, {0, 6, 3}},
                                      {{0, 6, 8}, {1, 7, 9}, {2, 4, 10}, {3, 5, 11}}};
__constant__ int8_t c_ver[12][4][2] = {
    {{0, 1}, {4, 5}, {7, 6}, {3, 2}}, {{3, 2}, {0, 1}, {4, 5}, {7, 6}}, {{2, 3}, {6, 7}, {5, 4}, {1, 0}},
    {{1, 0}, {2, 3}, {6, 7}, {5, 4}}, {{2, 1}, {3, 0}, {7, 4}, {6, 5}}, {{3, 0}, {2, 1}, {6, 5}, {7, 4}},
    {{0, 3}, {4, 7}, {5, 6}, {1, 2}}, {{1, 2}, {0, 3}, {4, 7}, {5, 6}}, {{4, 0}, {7, 3}, {6, 2}, {5, 1}},
    {{5, 1}, {4, 0}, {7, 3}, {6, 2}}, {{6, 2}, {5, 1}, {4, 0}, {7, 3}}, {{7, 3}, {6, 2}, {5, 1}, {4, 0}}};

// ------------------------------------------------------------------ case tables
#define MC_TABLE __constant__ static const
#include "mc_tables.inc"
#undef MC_TABLE

// ------------------------------------------------------------------ device
struct McArgs {
    LeafMap leaves, t8, t128;
    int n8, n128;
    const float* t8_val;
    const float* t128_val;
    const int4* origins;
    const uint64_t* masks;
    const float* values;
    const int64_t* units;
    uint64_t n_units;
    const int4* tile_org;  // x, y, z, size per visited tile
    const float* tile_val;
    uint32_t* flags;        // bit 0: a cube took case 6.1.2
    const float4* v12_in;   // per unit: the centre vertex the unit starts with (xyz)
    float4* unit_cv;        // per unit: its last computed centre (xyz), w = 1 when there is one
    float voxel_size;
};

// VolumeGrid::at: a leaf answers for its active voxels only, an active tile
// for its whole box, anything else has no value
__device__ bool mc_at(const McArgs& a, int x, int y, int z, float& v) {
    const int bx = x >> 3, by = y >> 3, bz = z >> 3;
    int32_t s = a.leaves.find(bx, by, bz);
    if (s >= 0) {
        const int o = ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
        if (!((a.masks[(uint64_t)s * 8 + (o >> 6)] >> (o & 63)) & 1ull)) return false;
        v = a.values[(uint64_t)s * LEAF_VOX + o];
        return true;
    }
    if (a.n8) {
        s = a.t8.find(bx, by, bz);
        if (s >= 0) {
            v = a.t8_val[s];
            return true;
        }
    }
    if (a.n128) {
        s = a.t128.find(x >> 7, y >> 7, z >> 7);
        if (s >= 0) {
            v = a.t128_val[s];
            return true;
        }
    }
    return false;
}

__device__ inline float mc_clamp(float v) { return fabsf(v) < 1e-6f ? copysignf(1e-6f, v) : v; }

__device__ inline bool mc_test_face(const float* c, int face) {
    const int f = face < 0 ? -face : face;
    const float a = c[c_face[f][0]], b = c[c_face[f][1]], cc = c[c_face[f][2]], d = c[c_face[f][3]];
    const float val = a * cc - b * d;
    if (fabsf(val) < 1.1920929e-07f) return face >= 0;
    return (float)face * a * val >= 0.0f;
}

__device__ inline int mc_amb_edge(const float* c, int list, float face) {
    int e = 0;
    for (int k = 0; k < 4; ++k)
        if (c[c_amb[list][k][0]] * face > 0.0f && c[c_amb[list][k][1]] * face > 0.0f) e = c_amb[list][k][2];
    return e;
}

__device__ inline int mc_verify(const float* c, int edge) {
    const float qa = c[c_ver[edge][0][1]], qb = c[c_ver[edge][1][1]], qc = c[c_ver[edge][2][1]],
                qd = c[c_ver[edge][3][1]];
    const float da = c[c_ver[edge][0][0]] - qa, db = c[c_ver[edge][1][0]] - qb, dc = c[c_ver[edge][2][0]] - qc,
                dd = c[c_ver[edge][3][0]] - qd;
    const float a = da * dc - db * dd;
    const float b = qc * da + qa * dc - qd * db - qb * dd;
    if (a > 0.0f) return 1;
    const float t = -b / (2.0f * a);
    if (t < 0.0f || t > 1.0f) return 1;
    const float at = qa + da * t, bt = qb + db * t, ct = qc + dc * t, dt = qd + dd * t;
    const float v = at * ct - bt * dt;
    return v < 0.0f ? 1 : 0;
}

__device__ inline bool mc_interior13(const float* c) {
    const float a = (c[0] - c[1]) * (c[7] - c[6]) - (c[4] - c[5]) * (c[3] - c[2]);
    const float b = c[6] * (c[0] - c[1]) + c[1] * (c[7] - c[6]) - c[2] * (c[4] - c[5]) - c[5] * (c[3] - c[2]);
    const float cc = c[1] * c[6] - c[5] * c[2];
    const float delta = b * b - 4.0f * a * cc;
    const float t1 = (-b + sqrtf(delta)) / (a + a);
    const float t2 = (-b - sqrtf(delta)) / (a + a);
    if (t1 < 1.0f && t1 > 0.0f && t2 < 1.0f && t2 > 0.0f) {
        float x[2], y[2];
        const float ts[2] = {t1, t2};
        for (int k = 0; k < 2; ++k) {
            const float t = ts[k];
            const float at = c[1] + (c[0] - c[1]) * t, bt = c[5] + (c[4] - c[5]) * t;
            const float ct = c[6] + (c[7] - c[6]) * t, dt = c[2] + (c[3] - c[2]) * t;
            x[k] = (at - dt) / (at + ct - bt - dt);
            y[k] = (at - bt) / (at + ct - bt - dt);
        }
        return !(x[0] < 1.0f && x[0] > 0.0f && x[1] < 1.0f && x[1] > 0.0f && y[0] < 1.0f && y[0] > 0.0f &&
                 y[1] < 1.0f && y[1] > 0.0f);
    }
    return true;
}

// Position of the crossing on cube edge e (ComputeEdgeIntersections,
// :1070-1098: t = |v1| / (|v1| + |v2|) from the lower corner).
__device__ inline void mc_vertex(const float* c, int e, int gx, int gy, int gz, float& x, float& y, float& z) {
    const int c0 = c_edge[e][0], c1 = c_edge[e][1];
    const float va = fabsf(c[c0]), vb = fabsf(c[c1]);
    const float t = va / (va + vb);
    x = (float)(gx + c_corner[c0][0]);
    y = (float)(gy + c_corner[c0][1]);
    z = (float)(gz + c_corner[c0][2]);
    const int ax = c_axis[e];
    if (ax == 0)
        x = x + t;
    else if (ax == 1)
        y = y + t;
    else
        z = z + t;
}

// interior_ambiguity's diagonal list of an ambiguous face (:474-530): faces
// 1|3, 2|4 and 5|6 (and 0)
__device__ inline int mc_family(int amb_face) { return (amb_face == 1 || amb_face == 3) ? 0 : (amb_face == 2 || amb_face == 4) ? 1 : 2; }

// test_interior (:404-468)
__device__ bool mc_test_interior(const float* c, int cas, int cfg, int face) {
    const float f = (float)face;
    switch (cas) {
        case 4:
            return mc_verify(c, mc_amb_edge(c, 0, f)) + mc_verify(c, mc_amb_edge(c, 1, f)) +
                       mc_verify(c, mc_amb_edge(c, 2, f)) != 0;
        case 6: return mc_verify(c, mc_amb_edge(c, mc_family(abs(MC_TEST_6[cfg][0])), f)) != 0;
        case 7:
            return mc_verify(c, mc_amb_edge(c, 0, -f)) + mc_verify(c, mc_amb_edge(c, 1, -f)) +
                       mc_verify(c, mc_amb_edge(c, 2, -f)) != 0;
        case 10: return mc_verify(c, mc_amb_edge(c, mc_family(abs(MC_TEST_10[cfg][0])), f)) != 0;
        default:
            return mc_verify(c, mc_amb_edge(c, mc_family(abs(MC_TEST_12[cfg][0])), f)) +
                       mc_verify(c, mc_amb_edge(c, mc_family(abs(MC_TEST_12[cfg][1])), f)) != 0;
    }
}

// The tiling handle_cube selects (:72-243): its entries, their count, and
// kind 1 where compute_c_vertex runs first (7.3, 10.2, 12.2, 13.3, 13.4) or
// kind 2 where vertex 12 is read without it (6.1.2).
struct McSel {
    const int8_t* t;
    int n;
    int kind;
};
#define MC_PICK(tab) (r.t = &(tab)[0], r.n = (int)sizeof(tab))
__device__ McSel mc_select(const float* c, int id) {
    const int cas = MC_CASES[id][0], cfg = MC_CASES[id][1];
    McSel r{nullptr, 0, 0};
    switch (cas) {
        case 1: MC_PICK(MC_TILING_1[cfg]); break;
        case 2: MC_PICK(MC_TILING_2[cfg]); break;
        case 3:
            if (mc_test_face(c, MC_TEST_3[cfg])) MC_PICK(MC_TILING_3_2[cfg]);
            else MC_PICK(MC_TILING_3_1[cfg]);
            break;
        case 4:
            if (mc_test_interior(c, 4, cfg, MC_TEST_4[cfg])) MC_PICK(MC_TILING_4_1[cfg]);
            else MC_PICK(MC_TILING_4_2[cfg]);
            break;
        case 5: MC_PICK(MC_TILING_5[cfg]); break;
        case 6:
            if (mc_test_face(c, MC_TEST_6[cfg][0])) MC_PICK(MC_TILING_6_2[cfg]);
            else if (mc_test_interior(c, 6, cfg, MC_TEST_6[cfg][1])) MC_PICK(MC_TILING_6_1_1[cfg]);
            else MC_PICK(MC_TILING_6_1_2[cfg]), r.kind = 2;
            break;
        case 7: {
            const int sub = (int)mc_test_face(c, MC_TEST_7[cfg][0]) + 2 * (int)mc_test_face(c, MC_TEST_7[cfg][1]) +
                            4 * (int)mc_test_face(c, MC_TEST_7[cfg][2]);
            if (sub == 0) MC_PICK(MC_TILING_7_1[cfg]);
            else if (sub == 1 || sub == 2) MC_PICK(MC_TILING_7_2[cfg][sub - 1]);
            else if (sub == 4) MC_PICK(MC_TILING_7_2[cfg][2]);
            else if (sub == 3) MC_PICK(MC_TILING_7_3[cfg][0]), r.kind = 1;
            else if (sub == 5 || sub == 6) MC_PICK(MC_TILING_7_3[cfg][sub - 4]), r.kind = 1;
            else if (mc_test_interior(c, 7, cfg, MC_TEST_7[cfg][3])) MC_PICK(MC_TILING_7_4_1[cfg]);
            else MC_PICK(MC_TILING_7_4_2[cfg]);
            break;
        }
        case 8: MC_PICK(MC_TILING_8[cfg]); break;
        case 9: MC_PICK(MC_TILING_9[cfg]); break;
        case 10:
            if (mc_test_face(c, MC_TEST_10[cfg][0])) {
                if (mc_test_face(c, MC_TEST_10[cfg][1])) {
                    if (mc_test_interior(c, 10, cfg, -MC_TEST_10[cfg][2])) MC_PICK(MC_TILING_10_1_1_[cfg]);
                    else MC_PICK(MC_TILING_10_1_2[5 - cfg]);
                } else {
                    MC_PICK(MC_TILING_10_2[cfg]), r.kind = 1;
                }
            } else if (mc_test_face(c, MC_TEST_10[cfg][1])) {
                MC_PICK(MC_TILING_10_2_[cfg]), r.kind = 1;
            } else if (mc_test_interior(c, 10, cfg, MC_TEST_10[cfg][2])) {
                MC_PICK(MC_TILING_10_1_1[cfg]);
            } else {
                MC_PICK(MC_TILING_10_1_2[cfg]);
            }
            break;
        case 11: MC_PICK(MC_TILING_11[cfg]); break;
        case 12:
            if (mc_test_face(c, MC_TEST_12[cfg][0])) {
                if (mc_test_face(c, MC_TEST_12[cfg][1])) {
                    if (mc_test_interior(c, 12, cfg, -MC_TEST_12[cfg][2])) MC_PICK(MC_TILING_12_1_1_[cfg]);
                    else MC_PICK(MC_TILING_12_1_2[23 - cfg]);
                } else {
                    MC_PICK(MC_TILING_12_2[cfg]), r.kind = 1;
                }
            } else if (mc_test_face(c, MC_TEST_12[cfg][1])) {
                MC_PICK(MC_TILING_12_2_[cfg]), r.kind = 1;
            } else if (mc_test_interior(c, 12, cfg, MC_TEST_12[cfg][2])) {
                MC_PICK(MC_TILING_12_1_1[cfg]);
            } else {
                MC_PICK(MC_TILING_12_1_2[cfg]);
            }
            break;
        case 13: {
            int sub = 0;
            for (int k = 0; k < 6; ++k) sub |= (int)mc_test_face(c, MC_TEST_13[cfg][k]) << k;
            const int o = MC_SUB_CONFIG_13[sub];
            if (o == 0) MC_PICK(MC_TILING_13_1[cfg]);
            else if (o <= 6) MC_PICK(MC_TILING_13_2[cfg][o - 1]);
            else if (o <= 18) MC_PICK(MC_TILING_13_3[cfg][o - 7]), r.kind = 1;
            else if (o <= 22) MC_PICK(MC_TILING_13_4[cfg][o - 19]), r.kind = 1;
            else if (o <= 26) {
                if (mc_interior13(c)) MC_PICK(MC_TILING_13_5_1[cfg][o - 23]);
                else MC_PICK(MC_TILING_13_5_2[1 - cfg][o - 23]);
            } else if (o <= 38) MC_PICK(MC_TILING_13_3_[cfg][o - 27]), r.kind = 1;
            else if (o <= 44) MC_PICK(MC_TILING_13_2_[cfg][o - 39]);
            else if (o == 45) MC_PICK(MC_TILING_13_1_[cfg]);
            break;
        }
        case 14: MC_PICK(MC_TILING_14[cfg]); break;
        default: break;
    }
    return r;
}
#undef MC_PICK

__device__ inline bool mc_crosses(const float* c, int e) { return (c[c_edge[e][0]] < 0.0f) != (c[c_edge[e][1]] < 0.0f); }

// compute_c_vertex (:919-940): the mean of the crossings, summed in edge order
__device__ inline float3 mc_centre(const float* c, int gx, int gy, int gz) {
    float mx = 0.0f, my = 0.0f, mz = 0.0f;
    int n = 0;
    for (int k = 0; k < 12; ++k) {
        if (!mc_crosses(c, k)) continue;
        float x, y, z;
        mc_vertex(c, k, gx, gy, gz, x, y, z);
        mx = mx + x, my = my + y, mz = mz + z;
        ++n;
    }
    return make_float3(mx / (float)n, my / (float)n, mz / (float)n);
}

// The centre vertex a cube leaves behind: v12 updated where it computes one.
__device__ inline void mc_track(const float* c, int id, int gx, int gy, int gz, float3& v12, bool& has) {
    if (mc_select(c, id).kind == 1) {
        v12 = mc_centre(c, gx, gy, gz);
        has = true;
    }
}

// One cube with a surface (min corner g; c: its eight clamped corner values,
// configuration id; v12: the mesher's centre vertex, updated here): the number
// of kept triangles; with `out`, they are written there, 9 floats each.
__device__ uint32_t mc_cube(const McArgs& a, const float* c, int id, int gx, int gy, int gz, float* out, float3& v12) {
    const McSel sel = mc_select(c, id);
    if (sel.kind == 1) v12 = mc_centre(c, gx, gy, gz);
    if (sel.kind == 2) atomicOr(a.flags, 1u);
    const float h = a.voxel_size;
    uint32_t kept = 0;
    for (int i = 0; i + 2 < sel.n; i += 3) {
        const int ed[3] = {sel.t[i], sel.t[i + 2], sel.t[i + 1]};  // add_faces: entries i, i + 2, i + 1
        float v[9];
        bool ok = true;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            float x = v12.x, y = v12.y, z = v12.z;
            if (ed[q] != 12) {
                ok = ok && mc_crosses(c, ed[q]);  // no intersection: the reference skips the triangle
                mc_vertex(c, ed[q], gx, gy, gz, x, y, z);
            }
            v[3 * q] = x * h, v[3 * q + 1] = y * h, v[3 * q + 2] = z * h;
        }
        if (!ok) continue;
        // Triangle3::is_degenerate (triangle3.rs:272-275)
        const float ux = v[3] - v[0], uy = v[4] - v[1], uz = v[5] - v[2];
        const float wx = v[6] - v[0], wy = v[7] - v[1], wz = v[8] - v[2];
        const float nx = uy * wz - uz * wy, ny = uz * wx - ux * wz, nz = ux * wy - uy * wx;
        if (nx * nx + ny * ny + nz * nz == 0.0f) continue;
        if (out) {
            float* o = out + 9 * (uint64_t)kept;
#pragma unroll
            for (int q = 0; q < 9; ++q) o[q] = v[q];
        }
        ++kept;
    }
    return kept;
}

// The centre each lane starts with, in serial (lane) order: that of the last
// earlier lane that computed one, else `seed`; `last` = the wave's last one
// (seed when none did).
__device__ inline float3 mc_lane_in(bool has, float3 mine, float3 seed, int lane, float3& last) {
    const unsigned long long m = __ballot(has);
    const unsigned long long before = m & ((1ull << lane) - 1);
    const int src = before ? 63 - __clzll((long long)before) : lane;
    float3 r = make_float3(__shfl(mine.x, src, 64), __shfl(mine.y, src, 64), __shfl(mine.z, src, 64));
    if (!before) r = seed;
    if (m) {
        const int hi = 63 - __clzll((long long)m);
        last = make_float3(__shfl(mine.x, hi, 64), __shfl(mine.y, hi, 64), __shfl(mine.z, hi, 64));
    } else {
        last = seed;
    }
    return r;
}

__device__ inline uint32_t wave_exclusive(uint32_t v, int lane, uint32_t& total) {
    uint32_t incl = v;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(incl, d, 64);
        if (lane >= d) incl += u;
    }
    total = __shfl(incl, 63, 64);
    return incl - v;
}

constexpr int MB = 9;  // staged block edge: leaf-local 0 .. 8
constexpr int CS = 9;  // per-lane corner slot stride in LDS

// offset of cube corner k in the staged 9^3 block
__device__ inline int mc_corner_off(int k) {
    return (((k + 1) >> 1) & 1) * (MB * MB) + ((k >> 1) & 1) * MB + ((k >> 2) & 1);
}

// Stages a leaf unit's 9^3 block (its values and the +x/+y/+z neighbour
// blocks: leaf, 8^3 tile or 128^3 tile) in LDS, clamped like Cube::from_voxel.
struct McStage {
    float val[MB * MB * MB];
    uint8_t act[MB * MB * MB];
    int32_t nb_leaf[8];
    float nb_val[8];
    uint8_t nb_kind[8];  // 0 none, 1 leaf, 2 tile
};
__device__ void mc_stage_leaf(const McArgs& a, McStage& S, int64_t unit, int4 org, int lane) {
    if (lane < 8) {
        const int dx = lane >> 2, dy = (lane >> 1) & 1, dz = lane & 1;
        const int bx = (org.x >> 3) + dx, by = (org.y >> 3) + dy, bz = (org.z >> 3) + dz;
        int32_t s = lane == 0 ? (int32_t)unit : a.leaves.find(bx, by, bz);
        uint8_t kind = s >= 0 ? 1 : 0;
        float tv = 0.0f;
        if (s < 0 && a.n8) {
            const int32_t t = a.t8.find(bx, by, bz);
            if (t >= 0) kind = 2, tv = a.t8_val[t];
        }
        if (s < 0 && kind == 0 && a.n128) {
            const int32_t t = a.t128.find(bx >> 4, by >> 4, bz >> 4);
            if (t >= 0) kind = 2, tv = a.t128_val[t];
        }
        S.nb_leaf[lane] = s;
        S.nb_kind[lane] = kind;
        S.nb_val[lane] = tv;
    }
    __syncthreads();
    for (int i = lane; i < MB * MB * MB; i += 64) {
        const int x = i / (MB * MB), y = (i / MB) % MB, z = i % MB;
        const int nb = ((x >> 3) << 2) | ((y >> 3) << 1) | (z >> 3);
        float v = 0.0f;
        bool on = false;
        if (S.nb_kind[nb] == 1) {
            const uint64_t s = (uint64_t)S.nb_leaf[nb];
            const int o = ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
            on = (a.masks[s * 8 + (o >> 6)] >> (o & 63)) & 1ull;
            if (on) v = a.values[s * LEAF_VOX + o];
        } else if (S.nb_kind[nb] == 2) {
            on = true;
            v = S.nb_val[nb];
        }
        S.val[i] = mc_clamp(v);
        S.act[i] = on;
    }
    __syncthreads();
}

// Cube (x, y, z) of a staged leaf: its corners copied to c; false when a
// corner has no value or the cube has no surface.
__device__ inline bool mc_leaf_cube(const McStage& S, int x, int y, int z, float* c, int& id) {
    const int b0 = (x * MB + y) * MB + z;
    id = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int i = b0 + mc_corner_off(k);
        if (!S.act[i]) return false;
        id |= (S.val[i] < 0.0f) << k;
    }
    if (id == 0 || id == 255) return false;
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = S.val[b0 + mc_corner_off(k)];
    return true;
}

// Face cube f of pair (i, j) of a tile of size s (origin to, value tv): its
// local min corner in l, corners in c; false when it is interior to the
// tile, has a corner without a value or has no surface.
__device__ inline bool mc_tile_cube(const McArgs& a, int4 to, int s, float tv, int f, int i, int j, int* l, float* c,
                                    int& id) {
    int lx, ly, lz;
    switch (f) {
        case 0: lx = 0, ly = i, lz = j; break;
        case 1: lx = s - 1, ly = i, lz = j; break;
        case 2: lx = i, ly = j, lz = s - 1; break;
        case 3: lx = i, ly = j, lz = 0; break;
        case 4: lx = i, ly = s - 1, lz = j; break;
        default: lx = i, ly = 0, lz = j; break;
    }
    if (lx < s - 1 && ly < s - 1 && lz < s - 1) return false;  // one value throughout
    id = 0;
    for (int k = 0; k < 8; ++k) {
        const int cx = lx + c_corner[k][0], cy = ly + c_corner[k][1], cz = lz + c_corner[k][2];
        float v = tv;
        if (!(cx < s && cy < s && cz < s)) {
            if (!mc_at(a, to.x + cx, to.y + cy, to.z + cz, v)) return false;
            v = mc_clamp(v);
        }
        c[k] = v;
        id |= (v < 0.0f) << k;
    }
    l[0] = lx, l[1] = ly, l[2] = lz;
    return id != 0 && id != 255;
}

__device__ inline float3 f3(float4 v) { return make_float3(v.x, v.y, v.z); }

// MODE 0: summary (unit_cv[u] = the unit's last computed centre, w = 1 when
// there is one).  MODE 1: count (unit_cnt[u] = kept triangles of unit u,
// lane_cnt[64 u + l] = those of lane l of a leaf unit).  MODE 2: write (the
// triangles at out + 9 * unit_off[u]).  STALE: the volume has a 6.1.2 cube,
// so every cube needs the serial centre: each unit starts from v12_in[u] and
// each lane from the last centre of the lanes before it.
template <int MODE, bool STALE>
__global__ __launch_bounds__(64) void k_mc_units(McArgs a, uint32_t* __restrict__ unit_cnt,
                                                 uint16_t* __restrict__ lane_cnt, const uint64_t* __restrict__ unit_off,
                                                 float* __restrict__ out) {
    constexpr bool WRITE = MODE == 2;
    __shared__ McStage S;
    __shared__ float cv[64 * CS];
    const int lane = threadIdx.x;
    float* mine_c = cv + lane * CS;
    for (uint64_t u = blockIdx.x; u < a.n_units; u += gridDim.x) {
        const int64_t unit = a.units[u];
        uint64_t running = WRITE ? unit_off[u] : 0;
        const float3 seed = STALE ? f3(a.v12_in[u]) : make_float3(0.0f, 0.0f, 0.0f);
        if (unit >= 0) {
            const int4 org = a.origins[unit];
            mc_stage_leaf(a, S, unit, org, lane);
            const int x = lane >> 3, y = lane & 7;
            const int gx = org.x + x, gy = org.y + y;
            float3 v12 = seed;
            if (MODE == 0 || STALE) {  // the lane's last centre, then the one it starts with
                bool has = false;
                float3 last = make_float3(0.0f, 0.0f, 0.0f);
                int id;
                for (int z = 0; z < 8; ++z)
                    if (mc_leaf_cube(S, x, y, z, mine_c, id)) mc_track(mine_c, id, gx, gy, org.z + z, last, has);
                float3 ulast;
                v12 = mc_lane_in(has, last, seed, lane, ulast);
                const bool any = __ballot(has) != 0;  // every lane takes part in the ballot
                if (MODE == 0 && lane == 0) a.unit_cv[u] = make_float4(ulast.x, ulast.y, ulast.z, any ? 1.0f : 0.0f);
            }
            if (MODE != 0) {
                uint32_t base = 0;
                if (WRITE) {
                    uint32_t total;
                    base = wave_exclusive(lane_cnt[u * 64 + lane], lane, total);
                }
                uint32_t mine = 0;
                int id;
                for (int z = 0; z < 8; ++z) {
                    if (!mc_leaf_cube(S, x, y, z, mine_c, id)) continue;
                    mine += mc_cube(a, mine_c, id, gx, gy, org.z + z,
                                    WRITE ? out + 9 * (running + base + mine) : nullptr, v12);
                }
                if (MODE == 1) {
                    uint32_t total;
                    wave_exclusive(mine, lane, total);
                    lane_cnt[u * 64 + lane] = (uint16_t)mine;
                    if (lane == 0) unit_cnt[u] = total;
                }
            }
            __syncthreads();
        } else {
            const int4 to = a.tile_org[-1 - unit];
            const float tv = mc_clamp(a.tile_val[-1 - unit]);
            const int s = to.w;
            const uint32_t pairs = (uint32_t)s * (uint32_t)s;
            uint32_t tile_total = 0;
            float3 carry = seed;  // the centre left by the chunks before
            bool unit_has = false;
            for (uint32_t q0 = 0; q0 < pairs; q0 += 64) {
                const uint32_t q = q0 + lane;
                const int i = (int)(q / (uint32_t)s), j = (int)(q % (uint32_t)s);
                float3 lane_v12 = carry;
                int l[3], id;
                if (MODE == 0 || STALE) {
                    bool has = false;
                    float3 last = make_float3(0.0f, 0.0f, 0.0f);
                    for (int f = 0; f < 6 && q < pairs; ++f)
                        if (mc_tile_cube(a, to, s, tv, f, i, j, l, mine_c, id))
                            mc_track(mine_c, id, to.x + l[0], to.y + l[1], to.z + l[2], last, has);
                    float3 clast;
                    lane_v12 = mc_lane_in(has, last, carry, lane, clast);
                    const bool any = __ballot(has) != 0;  // every lane takes part in the ballot
                    unit_has = unit_has || any;
                    carry = clast;
                }
                if (MODE != 0) {
                    uint32_t mine = 0, base = 0;
                    for (int pass = 0; pass < (WRITE ? 2 : 1); ++pass) {
                        if (pass == 1) {
                            uint32_t total;
                            base = wave_exclusive(mine, lane, total);
                            mine = 0;
                        }
                        float3 v12 = lane_v12;
                        for (int f = 0; f < 6 && q < pairs; ++f) {
                            if (!mc_tile_cube(a, to, s, tv, f, i, j, l, mine_c, id)) continue;
                            mine += mc_cube(a, mine_c, id, to.x + l[0], to.y + l[1], to.z + l[2],
                                            pass == 1 ? out + 9 * (running + base + mine) : nullptr, v12);
                        }
                    }
                    uint32_t total;
                    wave_exclusive(mine, lane, total);
                    running += total;
                    tile_total += total;
                }
            }
            if (MODE == 0 && lane == 0) a.unit_cv[u] = make_float4(carry.x, carry.y, carry.z, unit_has ? 1.0f : 0.0f);
            if (MODE == 1 && lane == 0) unit_cnt[u] = tile_total;
        }
    }
}

// the centre a unit starts with: that of the last earlier unit holding one
struct LastCentre {
    __device__ __host__ float4 operator()(const float4& a, const float4& b) const { return b.w != 0.0f ? b : a; }
};

void mc_scan(bs_ctx* ctx, const uint32_t* in, uint64_t* out, uint64_t n) {
    size_t bytes = 0;
    BS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, ctx->stream));
    void* tmp = ctx->scratch.get(S_CUB, bytes);
    BS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, ctx->stream));
}

// an open-addressing map of block coordinates -> position in `org`
LeafMap tile_map(bs_ctx* ctx, const std::vector<int4>& org, ScratchSlot s_org, ScratchSlot s_cells, ScratchSlot s_keys,
                 uint32_t* fail) {
    hipStream_t st = ctx->stream;
    LeafMap m{};
    m.dense = 0;
    uint64_t cap = 64;
    while (cap < 2 * org.size()) cap <<= 1;
    m.cap_mask = (uint32_t)(cap - 1);
    m.keys = ctx->scratch.as<uint64_t>(s_keys, cap);
    m.cells = ctx->scratch.as<int32_t>(s_cells, cap);
    BS_HIP(hipMemsetAsync(m.keys, 0xFF, sizeof(uint64_t) * cap, st));
    BS_HIP(hipMemsetAsync(m.cells, 0xFF, sizeof(int32_t) * cap, st));
    int4* d = ctx->scratch.as<int4>(s_org, org.size());
    BS_HIP(hipMemcpyAsync(d, org.data(), sizeof(int4) * org.size(), hipMemcpyHostToDevice, st));
    k_leafmap_insert<<<div_up(org.size(), 256), 256, 0, st>>>(org.size(), d, m, fail);
    BS_HIP(hipGetLastError());
    return m;
}

}  // namespace

void marching_cubes_device(bs_ctx* ctx, bs_volume* vol, float voxel_size, float** out_dev, uint64_t* n_tris,
                           bs_dc_stats* stats) {
    hipStream_t st = ctx->stream;
    *out_dev = nullptr;
    *n_tris = 0;
    BS_REQUIRE(vol->device == ctx->device, BS_ERR_INVALID_ARG, "volume lives on another device");
    if (stats) *stats = bs_dc_stats{};
    Topology& topo = ensure_topology(ctx, vol);
    std::vector<int64_t> units;
    std::vector<TileRec> tiles;
    topo.visit_units(units, tiles);
    const uint64_t n_units = units.size(), n = vol->n_leaves;
    if (!n_units) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (stats) {
        BS_HIP(hipEventCreate(&e0));
        BS_HIP(hipEventCreate(&e1));
        BS_HIP(hipEventRecord(e0, st));
    }
    McArgs a{};
    uint32_t* fail = ctx->scratch.as<uint32_t>(S_DC_FAIL, 1);
    BS_HIP(hipMemsetAsync(fail, 0, sizeof(uint32_t), st));
    // leaf map over the leaves' box
    a.leaves.dense = 1;
    a.leaves.nx = a.leaves.ny = a.leaves.nz = 0;
    if (n) {
        int32_t lo[3], hi[3];
        leaf_box(ctx, vol->origins, n, lo, hi);
        uint64_t cells = 1;
        for (int k = 0; k < 3; ++k) cells *= (uint64_t)(hi[k] - lo[k] + 1);
        LeafMap& m = a.leaves;
        m.dense = !(ctx->paths & BS_PATH_FORCE_HASH) && cells <= (1ull << 28);
        if (m.dense) {
            m.x0 = lo[0], m.y0 = lo[1], m.z0 = lo[2];
            m.nx = hi[0] - lo[0] + 1, m.ny = hi[1] - lo[1] + 1, m.nz = hi[2] - lo[2] + 1;
            m.cells = ctx->scratch.as<int32_t>(S_DC_MAP, cells);
            BS_HIP(hipMemsetAsync(m.cells, 0xFF, sizeof(int32_t) * cells, st));
        } else {
            uint64_t cap = 1024;
            while (cap < 2 * n) cap <<= 1;
            m.cap_mask = (uint32_t)(cap - 1);
            m.keys = ctx->scratch.as<uint64_t>(S_DC_KEYS, cap);
            m.cells = ctx->scratch.as<int32_t>(S_DC_MAP, cap);
            BS_HIP(hipMemsetAsync(m.keys, 0xFF, sizeof(uint64_t) * cap, st));
            BS_HIP(hipMemsetAsync(m.cells, 0xFF, sizeof(int32_t) * cap, st));
        }
        k_leafmap_insert<<<div_up(n, 256), 256, 0, st>>>(n, reinterpret_cast<const int4*>(vol->origins), m, fail);
        BS_HIP(hipGetLastError());
    }
    // every active tile of the volume, by size, for lookups; the visited
    // tiles (the same set, in visit order) for the tile units
    std::vector<int4> org8, org128;
    std::vector<float> val8, val128;
    for (const TileRec& t : tiles) {
        BS_REQUIRE(t.size == 8 || t.size == 128, BS_ERR_INTERNAL, "marching cubes: unexpected tile size");
        if (t.size == 8) {
            org8.push_back({t.x, t.y, t.z, 0});
            val8.push_back(t.value);
        } else {
            org128.push_back({t.x >> 4, t.y >> 4, t.z >> 4, 0});  // the map keys on >> 3
            val128.push_back(t.value);
        }
    }
    a.n8 = (int)org8.size();
    a.n128 = (int)org128.size();
    if (a.n8) a.t8 = tile_map(ctx, org8, S_MC_TORG, S_MC_T8, S_MC_T8K, fail);
    if (a.n128) a.t128 = tile_map(ctx, org128, S_MC_TILES, S_MC_T128, S_MC_T128K, fail);
    std::vector<float> tv(tiles.size() + val8.size() + val128.size() + 1);
    std::vector<int4> torg(tiles.size() + 1);
    for (size_t k = 0; k < tiles.size(); ++k) {
        torg[k] = {tiles[k].x, tiles[k].y, tiles[k].z, tiles[k].size};
        tv[k] = tiles[k].value;
    }
    std::copy(val8.begin(), val8.end(), tv.begin() + tiles.size());
    std::copy(val128.begin(), val128.end(), tv.begin() + tiles.size() + val8.size());
    // units, tile records and tile values in one upload
    const size_t b_units = 8 * n_units, b_torg = 16 * torg.size(), b_tv = 4 * tv.size();
    std::vector<uint8_t> up(b_units + b_torg + b_tv);
    std::memcpy(up.data(), units.data(), b_units);
    std::memcpy(up.data() + b_units, torg.data(), b_torg);
    std::memcpy(up.data() + b_units + b_torg, tv.data(), b_tv);
    uint8_t* dup = ctx->scratch.as<uint8_t>(S_MC_UNITS, up.size());
    BS_HIP(hipMemcpyAsync(dup, up.data(), up.size(), hipMemcpyHostToDevice, st));
    a.units = reinterpret_cast<const int64_t*>(dup);
    a.n_units = n_units;
    a.tile_org = reinterpret_cast<const int4*>(dup + b_units);
    a.tile_val = reinterpret_cast<const float*>(dup + b_units + b_torg);
    a.t8_val = a.tile_val + tiles.size();
    a.t128_val = a.t8_val + val8.size();
    a.origins = reinterpret_cast<const int4*>(vol->origins);
    a.masks = vol->masks;
    a.values = vol->values;
    a.voxel_size = voxel_size;
    a.flags = ctx->scratch.as<uint32_t>(S_MC_FLAGS, 1);
    BS_HIP(hipMemsetAsync(a.flags, 0, sizeof(uint32_t), st));

    const unsigned grid = (unsigned)std::min<uint64_t>(n_units, 16384);
    uint32_t* cnt = ctx->scratch.as<uint32_t>(S_DC_TRICNT, n_units);
    uint64_t* off = ctx->scratch.as<uint64_t>(S_DC_TRIOFF, n_units + 1);
    uint16_t* lane_cnt = ctx->scratch.as<uint16_t>(S_MC_LANES, n_units * 64);
    uint32_t h_fail = 0, h_flags = 0, last_cnt = 0;
    uint64_t last_off = 0;
    auto count = [&](bool stale) {
        if (stale)
            k_mc_units<1, true><<<grid, 64, 0, st>>>(a, cnt, lane_cnt, nullptr, nullptr);
        else
            k_mc_units<1, false><<<grid, 64, 0, st>>>(a, cnt, lane_cnt, nullptr, nullptr);
        BS_HIP(hipGetLastError());
        mc_scan(ctx, cnt, off, n_units);
        BS_HIP(hipMemcpyAsync(&h_fail, fail, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        BS_HIP(hipMemcpyAsync(&h_flags, a.flags, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        BS_HIP(hipMemcpyAsync(&last_cnt, cnt + n_units - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        BS_HIP(hipMemcpyAsync(&last_off, off + n_units - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        BS_HIP(hipStreamSynchronize(st));
        BS_REQUIRE(!h_fail, BS_ERR_INTERNAL, "marching cubes: a leaf or tile key could not be registered");
    };
    count(false);
    // a 6.1.2 cube reads the centre vertex an earlier cube left: find each
    // unit's last centre, hand every unit the one it starts with, count again
    const bool stale = (h_flags & 1u) != 0;
    if (stale) {
        a.unit_cv = ctx->scratch.as<float4>(S_MC_CV, n_units);
        float4* v12_in = ctx->scratch.as<float4>(S_MC_V12, n_units);
        k_mc_units<0, false><<<grid, 64, 0, st>>>(a, nullptr, nullptr, nullptr, nullptr);
        BS_HIP(hipGetLastError());
        size_t bytes = 0;
        const float4 none = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        BS_HIP(hipcub::DeviceScan::ExclusiveScan(nullptr, bytes, a.unit_cv, v12_in, LastCentre{}, none, (int)n_units,
                                                 st));
        void* tmp = ctx->scratch.get(S_CUB, bytes);
        BS_HIP(hipcub::DeviceScan::ExclusiveScan(tmp, bytes, a.unit_cv, v12_in, LastCentre{}, none, (int)n_units,
                                                 st));
        a.v12_in = v12_in;
        count(true);
    }
    const uint64_t total = last_off + last_cnt;
    if (total) {
        float* out = nullptr;
        BS_HIP(hipMalloc(&out, sizeof(float) * 9 * total));
        if (stale)
            k_mc_units<2, true><<<grid, 64, 0, st>>>(a, nullptr, lane_cnt, off, out);
        else
            k_mc_units<2, false><<<grid, 64, 0, st>>>(a, nullptr, lane_cnt, off, out);
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
        stats->n_leaves = n;
        stats->n_tris = total;
        stats->used_hash = a.leaves.dense ? 0 : 1;
        stats->ms_total = ms;
        BS_HIP(hipEventDestroy(e0));
        BS_HIP(hipEventDestroy(e1));
    }
}

}  // namespace bs
