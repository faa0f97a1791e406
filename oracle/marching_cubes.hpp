// Test infrastructure only: the CPU oracle.  Nothing outside tests/, the smoke
// check and bench.py's cpu_baseline leg may link or call this code.
//
// Restatement of MarchingCubesMesher (src/voxel/meshing/marching_cubes.rs),
// the corrected Marching Cubes 33 that the reference drives with Lewiner et
// al.'s case tables (src/voxel/meshing/lookup_table.rs, here as the data of
// mc_tables.inc, generated from that file): handle_cube's case selection
// (:72-243) with test_face (:323-342), test_interior and its helpers
// (:404-917) and interior_test_case13 (:269-321); add_faces' vertex order
// (:245-272, entries i, i + 2, i + 1; a triangle with an edge that has no
// crossing is skipped, a degenerate one dropped); the centre vertex as mesher
// state set only by compute_c_vertex (:919-940), which case 6.1.2 does not
// call; the serial visit of leaves and tile faces (:957-1000).
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "vdb.hpp"
#include "vec.hpp"

namespace oracle {
namespace mc {

constexpr float MIN_ABS = 1e-6f;           // MIN_ABS_VERTEX_VALUE, marching_cubes.rs:1131
constexpr float F32_EPS = 1.1920929e-07f;  // f32::EPSILON in test_face
// CUBE_OFFSETS (utils.rs:86-95)
constexpr int CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
// Edge::from_edge_num (lookup_table.rs:54-69): lower corner first; 12 is the centre
constexpr int EDGE_V[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
                               {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
constexpr int EDGE_AXIS[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
// test_face corner order a, b, c, d of faces 1..6 (marching_cubes.rs:327-334)
constexpr int FACE_Q[7][4] = {{0, 0, 0, 0}, {0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3},
                              {3, 7, 4, 0}, {0, 3, 2, 1}, {4, 7, 6, 5}};
// interior_ambiguity (:474-530): faces 1|3, 2|4 and 5|6 share a list of body
// diagonals (corner, corner, edge named), tested in this order, the last match wins
constexpr int FAMILY[7] = {2, 0, 1, 0, 1, 2, 2};
constexpr int AMB[3][4][3] = {{{1, 7, 4}, {0, 6, 5}, {3, 5, 6}, {2, 4, 7}},
                              {{1, 7, 0}, {2, 4, 1}, {3, 5, 2}, {0, 6, 3}},
                              {{0, 6, 8}, {1, 7, 9}, {2, 4, 10}, {3, 5, 11}}};
// interior_ambiguity_verification (:532-917): for each edge the four corner
// pairs (P, Q) of A, B, C, D; the test reads the quadratic
// (qA + dA t)(qC + dC t) - (qB + dB t)(qD + dD t), d = P - Q, at its vertex
constexpr int VER[12][4][2] = {
    {{0, 1}, {4, 5}, {7, 6}, {3, 2}}, {{3, 2}, {0, 1}, {4, 5}, {7, 6}}, {{2, 3}, {6, 7}, {5, 4}, {1, 0}},
    {{1, 0}, {2, 3}, {6, 7}, {5, 4}}, {{2, 1}, {3, 0}, {7, 4}, {6, 5}}, {{3, 0}, {2, 1}, {6, 5}, {7, 4}},
    {{0, 3}, {4, 7}, {5, 6}, {1, 2}}, {{1, 2}, {0, 3}, {4, 7}, {5, 6}}, {{4, 0}, {7, 3}, {6, 2}, {5, 1}},
    {{5, 1}, {4, 0}, {7, 3}, {6, 2}}, {{6, 2}, {5, 1}, {4, 0}, {7, 3}}, {{7, 3}, {6, 2}, {5, 1}, {4, 0}}};

// marching_cubes.rs:323-342
inline bool test_face(const float* c, int face) {
    const int* q = FACE_Q[face < 0 ? -face : face];
    const float a = c[q[0]], b = c[q[1]], cc = c[q[2]], d = c[q[3]];
    const float val = a * cc - b * d;
    if (std::fabs(val) < F32_EPS) return face >= 0;
    return (float)face * a * val >= 0.0f;
}

// marching_cubes.rs:474-530 (family of the ambiguous face, sign carrier `face`)
inline int amb_edge(const float* c, int family, float face) {
    int e = 0;
    for (int k = 0; k < 4; ++k) {
        const int* r = AMB[family][k];
        if (c[r[0]] * face > 0.0f && c[r[1]] * face > 0.0f) e = r[2];
    }
    return e;
}

// marching_cubes.rs:532-917: 1 when the corners are not joined in the interior
inline int verify(const float* c, int edge) {
    const int(*p)[2] = VER[edge];
    const float qa = c[p[0][1]], qb = c[p[1][1]], qc = c[p[2][1]], qd = c[p[3][1]];
    const float da = c[p[0][0]] - qa, db = c[p[1][0]] - qb, dc = c[p[2][0]] - qc, dd = c[p[3][0]] - qd;
    const float a = da * dc - db * dd;
    const float b = qc * da + qa * dc - qd * db - qb * dd;
    if (a > 0.0f) return 1;
    const float t = -b / (2.0f * a);
    if (t < 0.0f || t > 1.0f) return 1;
    const float at = qa + da * t, bt = qb + db * t, ct = qc + dc * t, dt = qd + dd * t;
    const float v = at * ct - bt * dt;
    return v < 0.0f ? 1 : 0;  // v > 0 joined; v == 0 or NaN falls through to 0
}

// interior_test_case13 (marching_cubes.rs:269-321)
inline bool interior13(const float* c) {
    const float a = (c[0] - c[1]) * (c[7] - c[6]) - (c[4] - c[5]) * (c[3] - c[2]);
    const float b = c[6] * (c[0] - c[1]) + c[1] * (c[7] - c[6]) - c[2] * (c[4] - c[5]) - c[5] * (c[3] - c[2]);
    const float cc = c[1] * c[6] - c[5] * c[2];
    const float delta = b * b - 4.0f * a * cc;
    const float t1 = (-b + std::sqrt(delta)) / (a + a);
    const float t2 = (-b - std::sqrt(delta)) / (a + a);
    if (t1 < 1.0f && t1 > 0.0f && t2 < 1.0f && t2 > 0.0f) {
        auto xy = [&](float t, float& x, float& y) {
            const float at = c[1] + (c[0] - c[1]) * t, bt = c[5] + (c[4] - c[5]) * t;
            const float ct = c[6] + (c[7] - c[6]) * t, dt = c[2] + (c[3] - c[2]) * t;
            x = (at - dt) / (at + ct - bt - dt);
            y = (at - bt) / (at + ct - bt - dt);
        };
        float x1, y1, x2, y2;
        xy(t1, x1, y1);
        xy(t2, x2, y2);
        return !(x1 < 1.0f && x1 > 0.0f && x2 < 1.0f && x2 > 0.0f && y1 < 1.0f && y1 > 0.0f && y2 < 1.0f &&
                 y2 > 0.0f);
    }
    return true;
}

using Tri = std::array<int, 3>;

#define MC_TABLE static const
#include "mc_tables.inc"
#undef MC_TABLE

// test_interior (marching_cubes.rs:404-468): `face` carries the sign of the
// corners whose body diagonal is followed
inline bool test_interior(const float* c, int cas, int config, int face) {
    auto ver = [&](int amb_face, int f) { return verify(c, amb_edge(c, FAMILY[amb_face], (float)f)); };
    switch (cas) {
        case 4: return ver(1, face) + ver(2, face) + ver(5, face) != 0;
        case 6: return ver(std::abs(MC_TEST_6[config][0]), face) != 0;
        case 7: return ver(1, -face) + ver(2, -face) + ver(5, -face) != 0;
        case 10: return ver(std::abs(MC_TEST_10[config][0]), face) != 0;
        case 12:
            return ver(std::abs(MC_TEST_12[config][0]), face) + ver(std::abs(MC_TEST_12[config][1]), face) != 0;
        default: std::abort();
    }
}

// handle_cube (marching_cubes.rs:72-243): the tiling the tables select for a
// cube of corner values c and configuration id, appended as triangles in the
// emitted vertex order (add_faces, :245-272, reads entries i, i + 2, i + 1).
// Returns true where the reference computes the centre vertex first
// (compute_c_vertex): 7.3, 10.2, 12.2, 13.3 and 13.4.  Case 6.1.2 also uses
// vertex 12 but does not compute it, so it reads the centre left by an
// earlier cube.
template <size_t N>
inline void add_faces(const int8_t (&t)[N], std::vector<Tri>& out) {
    for (size_t i = 0; i + 2 < N; i += 3) out.push_back({t[i], t[i + 2], t[i + 1]});
}
inline bool handle_cube(const float* c, int id, std::vector<Tri>& out) {
    const int cas = MC_CASES[id][0], cfg = MC_CASES[id][1];
    auto tf = [&](int f) { return test_face(c, f); };
    auto ti = [&](int f) { return test_interior(c, cas, cfg, f); };
    switch (cas) {
        case 0: return false;
        case 1: add_faces(MC_TILING_1[cfg], out); return false;
        case 2: add_faces(MC_TILING_2[cfg], out); return false;
        case 3:
            if (tf(MC_TEST_3[cfg])) add_faces(MC_TILING_3_2[cfg], out);
            else add_faces(MC_TILING_3_1[cfg], out);
            return false;
        case 4:
            if (ti(MC_TEST_4[cfg])) add_faces(MC_TILING_4_1[cfg], out);
            else add_faces(MC_TILING_4_2[cfg], out);
            return false;
        case 5: add_faces(MC_TILING_5[cfg], out); return false;
        case 6:
            if (tf(MC_TEST_6[cfg][0])) add_faces(MC_TILING_6_2[cfg], out);
            else if (ti(MC_TEST_6[cfg][1])) add_faces(MC_TILING_6_1_1[cfg], out);
            else add_faces(MC_TILING_6_1_2[cfg], out);
            return false;
        case 7: {
            const int sub = (int)tf(MC_TEST_7[cfg][0]) + 2 * (int)tf(MC_TEST_7[cfg][1]) + 4 * (int)tf(MC_TEST_7[cfg][2]);
            switch (sub) {
                case 0: add_faces(MC_TILING_7_1[cfg], out); return false;
                case 1: add_faces(MC_TILING_7_2[cfg][0], out); return false;
                case 2: add_faces(MC_TILING_7_2[cfg][1], out); return false;
                case 3: add_faces(MC_TILING_7_3[cfg][0], out); return true;
                case 4: add_faces(MC_TILING_7_2[cfg][2], out); return false;
                case 5: add_faces(MC_TILING_7_3[cfg][1], out); return true;
                case 6: add_faces(MC_TILING_7_3[cfg][2], out); return true;
                default:
                    if (ti(MC_TEST_7[cfg][3])) add_faces(MC_TILING_7_4_1[cfg], out);
                    else add_faces(MC_TILING_7_4_2[cfg], out);
                    return false;
            }
        }
        case 8: add_faces(MC_TILING_8[cfg], out); return false;
        case 9: add_faces(MC_TILING_9[cfg], out); return false;
        case 10:
            if (tf(MC_TEST_10[cfg][0])) {
                if (tf(MC_TEST_10[cfg][1])) {
                    if (ti(-MC_TEST_10[cfg][2])) add_faces(MC_TILING_10_1_1_[cfg], out);
                    else add_faces(MC_TILING_10_1_2[5 - cfg], out);
                    return false;
                }
                add_faces(MC_TILING_10_2[cfg], out);
                return true;
            }
            if (tf(MC_TEST_10[cfg][1])) {
                add_faces(MC_TILING_10_2_[cfg], out);
                return true;
            }
            if (ti(MC_TEST_10[cfg][2])) add_faces(MC_TILING_10_1_1[cfg], out);
            else add_faces(MC_TILING_10_1_2[cfg], out);
            return false;
        case 11: add_faces(MC_TILING_11[cfg], out); return false;
        case 12:
            if (tf(MC_TEST_12[cfg][0])) {
                if (tf(MC_TEST_12[cfg][1])) {
                    if (ti(-MC_TEST_12[cfg][2])) add_faces(MC_TILING_12_1_1_[cfg], out);
                    else add_faces(MC_TILING_12_1_2[23 - cfg], out);
                    return false;
                }
                add_faces(MC_TILING_12_2[cfg], out);
                return true;
            }
            if (tf(MC_TEST_12[cfg][1])) {
                add_faces(MC_TILING_12_2_[cfg], out);
                return true;
            }
            if (ti(MC_TEST_12[cfg][2])) add_faces(MC_TILING_12_1_1[cfg], out);
            else add_faces(MC_TILING_12_1_2[cfg], out);
            return false;
        case 13: {
            int sub = 0;
            for (int k = 0; k < 6; ++k)
                if (tf(MC_TEST_13[cfg][k])) sub += 1 << k;
            const int o = MC_SUB_CONFIG_13[sub];
            if (o == 0) add_faces(MC_TILING_13_1[cfg], out);
            else if (o <= 6) add_faces(MC_TILING_13_2[cfg][o - 1], out);
            else if (o <= 18) { add_faces(MC_TILING_13_3[cfg][o - 7], out); return true; }
            else if (o <= 22) { add_faces(MC_TILING_13_4[cfg][o - 19], out); return true; }
            else if (o <= 26) {
                if (interior13(c)) add_faces(MC_TILING_13_5_1[cfg][o - 23], out);
                else add_faces(MC_TILING_13_5_2[1 - cfg][o - 23], out);
            } else if (o <= 38) { add_faces(MC_TILING_13_3_[cfg][o - 27], out); return true; }
            else if (o <= 44) add_faces(MC_TILING_13_2_[cfg][o - 39], out);
            else if (o == 45) add_faces(MC_TILING_13_1_[cfg], out);
            return false;
        }
        case 14: add_faces(MC_TILING_14[cfg], out); return false;
        default: std::abort();
    }
}

// MarchingCubesMesher::mesh (marching_cubes.rs:43-63): a serial visit of the
// leaves and active tiles in tree order; a tile meets the cubes of its six
// boundary faces, face order left, right, top, bottom, front, back per (i, j)
// (:970-994), so cubes on its edges come more than once.
template <class Grid>
std::vector<Vec3f> mesh(const Grid& grid, float voxel_size) {
    std::vector<Vec3f> verts;
    std::vector<Tri> tris;
    // the centre vertex is mesher state (MarchingCubesMesher::v12, zero in a
    // new mesher, :941-955) that only compute_c_vertex sets: case 6.1.2 reads
    // the one an earlier cube left
    Vec3f v12{0.0f, 0.0f, 0.0f};
    auto handle = [&](int64_t x, int64_t y, int64_t z) {
        float c[8];
        Vec3i idx[8];
        int id = 0;
        for (int i = 0; i < 8; ++i) {  // Cube::from_voxel (:1161-1182)
            idx[i] = {x + CORNER[i][0], y + CORNER[i][1], z + CORNER[i][2]};
            const float* p = grid.at(idx[i]);
            if (!p) return;
            float v = *p;
            if (std::fabs(v) < MIN_ABS) v = std::copysign(MIN_ABS, v);
            c[i] = v;
            if (v < 0.0f) id |= 1 << i;
        }
        tris.clear();
        const bool need_c = handle_cube(c, id, tris);
        if (tris.empty()) return;
        // ComputeEdgeIntersections (:1070-1098): t = |v1| / (|v1| + |v2|) from the lower corner
        Vec3f p[13];
        bool has[13];
        for (int e = 0; e < 12; ++e) {
            const int a = EDGE_V[e][0], b = EDGE_V[e][1];
            has[e] = (c[a] < 0.0f) != (c[b] < 0.0f);
            if (!has[e]) continue;
            const float va = std::fabs(c[a]), vb = std::fabs(c[b]);
            const float t = va / (va + vb);
            Vec3f q{(float)idx[a].x, (float)idx[a].y, (float)idx[a].z};
            q[EDGE_AXIS[e]] = q[EDGE_AXIS[e]] + t;
            p[e] = q;
        }
        if (need_c) {  // compute_c_vertex (:919-940)
            Vec3f s{0.0f, 0.0f, 0.0f};
            int n = 0;
            for (int e = 0; e < 12; ++e)
                if (has[e]) s = add(s, p[e]), ++n;
            v12 = divs(s, (float)n);
        }
        p[12] = v12;
        has[12] = true;
        for (const Tri& t : tris) {  // add_faces (:245-272)
            if (!has[t[0]] || !has[t[1]] || !has[t[2]]) continue;  // no intersection: the triangle is skipped
            const Vec3f v1 = mul(p[t[0]], voxel_size), v2 = mul(p[t[1]], voxel_size), v3 = mul(p[t[2]], voxel_size);
            if (norm_sq(cross(sub(v2, v1), sub(v3, v1))) == 0.0f) continue;
            verts.push_back(v1);
            verts.push_back(v2);
            verts.push_back(v3);
        }
    };
    grid.visit_leafs(
        [&](Vec3i o, int s, float) {
            for (int64_t i = 0; i < s; ++i)
                for (int64_t j = 0; j < s; ++j) {
                    handle(o.x, o.y + i, o.z + j);
                    handle(o.x + s - 1, o.y + i, o.z + j);
                    handle(o.x + i, o.y + j, o.z + s - 1);
                    handle(o.x + i, o.y + j, o.z);
                    handle(o.x + i, o.y + s - 1, o.z + j);
                    handle(o.x + i, o.y, o.z + j);
                }
        },
        [&](const auto& leaf) {
            for (int64_t x = 0; x < 8; ++x)
                for (int64_t y = 0; y < 8; ++y)
                    for (int64_t z = 0; z < 8; ++z) handle(leaf.origin.x + x, leaf.origin.y + y, leaf.origin.z + z);
        });
    return verts;
}

}  // namespace mc
}  // namespace oracle
