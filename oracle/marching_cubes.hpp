// Test infrastructure only: the CPU oracle.  Nothing outside tests/, the smoke
// check and bench.py's cpu_baseline leg may link or call this code.
//
// Restatement of MarchingCubesMesher (src/voxel/meshing/marching_cubes.rs),
// the corrected Marching Cubes 33 that the reference drives with Lewiner et
// al.'s case tables (src/voxel/meshing/lookup_table.rs).  The tables are not
// retyped.  The surface in a cube is derived from the cube itself:
//   * every face with two diagonal corners of each sign is decided with the
//     reference's face test (test_face, marching_cubes.rs:323-342), called
//     with the sign the tables give every tested face (+ when at most four
//     corners are negative, else -) and read as the tables read it;
//   * the contour segments on the six faces chain into closed loops;
//   * the interior tests of cases 4, 6, 7, 10, 12 and 13 (test_interior,
//     interior_ambiguity, interior_ambiguity_verification and
//     interior_test_case13, marching_cubes.rs:253-917) decide whether two loops
//     bound one tube;
//   * a loop of eight or more crossings is fanned around the centre vertex
//     (compute_c_vertex, :919-940), as the tables do in cases 7.3, 10.2, 12.2,
//     13.3 and 13.4; the tube of case 6.1.2 holds the centre vertex too.
// Two behaviours of the reference's table indexing are kept: case 13.5.2
// emits the tiling of the complementary configuration (TILING_13_5_2 indexed
// by the other config, :228-238), and cases 10.1.2 and 12.1.2 reached with both
// face tests true emit their tube reversed (index 5 - config and 23 - config,
// :161,183).  The result has the reference's crossings, topology and
// orientation per cube; which diagonal splits a polygon and the order of the
// triangles within a cube are this restatement's own.
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
// the same faces, corners counter-clockwise seen from outside the cube
constexpr int FACE_CCW[7][4] = {{0, 0, 0, 0}, {1, 5, 4, 0}, {2, 6, 5, 1}, {3, 7, 6, 2},
                                {0, 4, 7, 3}, {0, 3, 2, 1}, {5, 6, 7, 4}};
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

inline int edge_of(int a, int b) {
    for (int e = 0; e < 12; ++e)
        if ((EDGE_V[e][0] == a && EDGE_V[e][1] == b) || (EDGE_V[e][0] == b && EDGE_V[e][1] == a)) return e;
    std::abort();
}

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
using Loop = std::vector<int>;

// Contour loops over the cube faces.  Faces are walked counter-clockwise from
// outside; a segment runs from a positive-to-negative crossing to a
// negative-to-positive one, so a loop's direction makes its fan face the
// positive side, as the reference's triangles do.
inline std::vector<Loop> face_loops(const bool* neg, const bool* negconn) {
    int succ[12];
    for (int& s : succ) s = -1;
    for (int f = 1; f <= 6; ++f) {
        const int* q = FACE_CCW[f];
        bool sg[4];
        int e[4], nx = 0;
        for (int i = 0; i < 4; ++i) {
            sg[i] = neg[q[i]];
            e[i] = edge_of(q[i], q[(i + 1) & 3]);
        }
        for (int i = 0; i < 4; ++i) nx += sg[i] != sg[(i + 1) & 3];
        if (nx == 4) {
            for (int i = 0; i < 4; ++i) {
                if (sg[i] == negconn[f]) continue;  // corners of the joined sign are not cut off
                const int ein = e[(i + 3) & 3], eout = e[i];
                if (sg[i])
                    succ[ein] = eout;
                else
                    succ[eout] = ein;
            }
        } else if (nx == 2) {
            int pn = -1, np = -1;
            for (int i = 0; i < 4; ++i) {
                if (!sg[i] && sg[(i + 1) & 3]) pn = e[i];
                if (sg[i] && !sg[(i + 1) & 3]) np = e[i];
            }
            succ[pn] = np;
        }
    }
    std::vector<Loop> loops;
    bool seen[12] = {};
    for (int s = 0; s < 12; ++s) {
        if (succ[s] < 0 || seen[s]) continue;
        Loop l;
        for (int x = s; !seen[x]; x = succ[x]) {
            seen[x] = true;
            l.push_back(x);
        }
        loops.push_back(l);
    }
    return loops;
}

inline void fan(const Loop& l, std::vector<Tri>& out) {
    for (size_t i = 1; i + 1 < l.size(); ++i) out.push_back({l[0], l[i], l[i + 1]});
}
inline void centre_fan(const Loop& l, std::vector<Tri>& out) {
    for (size_t i = 0; i < l.size(); ++i) out.push_back({l[i], l[(i + 1) % l.size()], 12});
}
// A tube between two loops, both in their boundary direction: A walked
// forward, B backward, each step closing one edge of one loop.  With `centre`
// the rung between the first vertices is split through the centre vertex.
inline void tube(const Loop& A, const Loop& B, bool centre, bool reversed, std::vector<Tri>& out) {
    const size_t n1 = A.size(), n2 = B.size();
    std::vector<int> b(B.rbegin(), B.rend());
    std::vector<Tri> t;
    size_t i = 0, k = 0;
    while (i < n1 || k < n2) {
        if (k >= n2 || (i < n1 && (i + 1) * n2 <= (k + 1) * n1)) {
            t.push_back({A[i % n1], A[(i + 1) % n1], b[k % n2]});
            ++i;
        } else {
            t.push_back({b[(k + 1) % n2], b[k % n2], A[i % n1]});
            ++k;
        }
    }
    for (Tri tr : t) {
        if (centre) {
            int ia = -1, ib = -1;
            for (int j = 0; j < 3; ++j) {
                if (tr[j] == A[0]) ia = j;
                if (tr[j] == b[0]) ib = j;
            }
            if (ia >= 0 && ib >= 0) {
                Tri t1 = tr, t2 = tr;
                t1[ib] = 12;
                t2[ia] = 12;
                out.push_back(reversed ? Tri{t1[0], t1[2], t1[1]} : t1);
                out.push_back(reversed ? Tri{t2[0], t2[2], t2[1]} : t2);
                continue;
            }
        }
        out.push_back(reversed ? Tri{tr[0], tr[2], tr[1]} : tr);
    }
}

inline int loop_corner(const Loop& l) {
    for (int v = 0; v < 8; ++v) {
        bool all = true;
        for (int e : l) all = all && (EDGE_V[e][0] == v || EDGE_V[e][1] == v);
        if (all) return v;
    }
    return -1;
}

// The triangles of one cube as crossing-edge triples (12 = centre vertex), in
// the orientation of the reference's emitted vertex order.  c: the clamped
// corner values (Cube::from_voxel, :1161-1182).
inline void cube_tiling(const float* c, std::vector<Tri>& out) {
    bool neg[8];
    int pc = 0;
    for (int i = 0; i < 8; ++i) {
        neg[i] = c[i] < 0.0f;
        pc += neg[i];
    }
    if (pc == 0 || pc == 8) return;
    const float s0 = pc <= 4 ? 7.0f : -7.0f;
    bool negconn[7] = {};
    int amb_faces[6], n_amb = 0, n_conn = 0;
    for (int f = 1; f <= 6; ++f) {
        const int* q = FACE_CCW[f];
        if (neg[q[0]] != neg[q[1]] && neg[q[1]] != neg[q[2]] && neg[q[2]] != neg[q[3]]) {
            const bool r = test_face(c, s0 > 0.0f ? f : -f);
            negconn[f] = s0 > 0.0f ? r : !r;
            amb_faces[n_amb++] = f;
            n_conn += negconn[f];
        }
    }
    std::vector<Loop> loops = face_loops(neg, negconn);
    // minority corners (the negative ones when there are at most four)
    int k = 0, edges_in = 0;
    bool minor[8];
    for (int i = 0; i < 8; ++i) {
        minor[i] = neg[i] == (pc <= 4);
        k += minor[i];
    }
    for (int e = 0; e < 12; ++e) edges_in += minor[EDGE_V[e][0]] && minor[EDGE_V[e][1]];

    int ta = -1, tb = -1;  // loops joined by a tube
    bool centre = false, reversed = false;
    if (k == 2 && n_amb == 0 && loops.size() == 2) {  // case 4
        const int sum = verify(c, amb_edge(c, 0, s0)) + verify(c, amb_edge(c, 1, s0)) + verify(c, amb_edge(c, 2, s0));
        if (sum == 0) ta = 0, tb = 1;
    } else if (k == 3 && edges_in == 1 && loops.size() == 2) {  // case 6.1
        if (verify(c, amb_edge(c, FAMILY[amb_faces[0]], s0)) == 0) ta = 0, tb = 1, centre = true;
    } else if (k == 3 && edges_in == 0 && n_conn == (pc <= 4 ? 3 : 0)) {  // case 7.4: all three faces join
        const float s = -s0;
        const int sum = verify(c, amb_edge(c, 0, s)) + verify(c, amb_edge(c, 1, s)) + verify(c, amb_edge(c, 2, s));
        if (sum == 0) ta = 0, tb = 1;
    } else if (k == 4 && edges_in == 2 && loops.size() == 2) {  // cases 10.1 and 12.1
        const bool both = negconn[amb_faces[0]];
        const float arg = both ? -7.0f : 7.0f;
        int sum = verify(c, amb_edge(c, FAMILY[amb_faces[0]], arg));
        if (FAMILY[amb_faces[0]] != FAMILY[amb_faces[1]]) sum += verify(c, amb_edge(c, FAMILY[amb_faces[1]], arg));
        if (sum == 0) ta = 0, tb = 1, reversed = both;
    } else if (k == 4 && edges_in == 0 && loops.size() == 3 && n_conn == 3) {  // case 13.5
        if (!interior13(c)) {
            // the complementary configuration's tiling: faces 1..4 of the
            // pattern turned a quarter, signs flipped (TEST_13 face orders)
            const int id = [&] {
                int v = 0;
                for (int i = 0; i < 8; ++i) v |= neg[i] << i;
                return v;
            }();
            bool rp[7] = {};
            if (id == 165) {
                rp[2] = negconn[1], rp[3] = negconn[2], rp[4] = negconn[3], rp[1] = negconn[4];
            } else {
                rp[1] = negconn[2], rp[2] = negconn[3], rp[3] = negconn[4], rp[4] = negconn[1];
            }
            rp[5] = negconn[5], rp[6] = negconn[6];
            bool np[8];
            for (int i = 0; i < 8; ++i) np[i] = !neg[i];
            std::vector<Loop> L = face_loops(np, rp);
            int six = -1, joined = -1, alone = -1;
            for (int j = 0; j < (int)L.size(); ++j) {
                if (L[j].size() == 6)
                    six = j;
                else if (np[loop_corner(L[j])])
                    joined = j;
                else
                    alone = j;
            }
            tube(L[six], L[joined], false, false, out);
            fan(L[alone], out);
            return;
        }
    }
    if (ta >= 0) tube(loops[ta], loops[tb], centre, reversed, out);
    for (int j = 0; j < (int)loops.size(); ++j) {
        if (j == ta || j == tb) continue;
        if (loops[j].size() >= 8)
            centre_fan(loops[j], out);
        else
            fan(loops[j], out);
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
    auto handle = [&](int64_t x, int64_t y, int64_t z) {
        float c[8];
        Vec3i idx[8];
        for (int i = 0; i < 8; ++i) {
            idx[i] = {x + CORNER[i][0], y + CORNER[i][1], z + CORNER[i][2]};
            const float* p = grid.at(idx[i]);
            if (!p) return;
            float v = *p;
            if (std::fabs(v) < MIN_ABS) v = std::copysign(MIN_ABS, v);
            c[i] = v;
        }
        tris.clear();
        cube_tiling(c, tris);
        if (tris.empty()) return;
        // ComputeEdgeIntersections (:1070-1098): t = |v1| / (|v1| + |v2|) from the lower corner
        Vec3f p[13];
        bool has[12];
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
        bool need_c = false;
        for (const Tri& t : tris) need_c = need_c || t[0] == 12 || t[1] == 12 || t[2] == 12;
        if (need_c) {  // compute_c_vertex (:919-940)
            Vec3f s{0.0f, 0.0f, 0.0f};
            int n = 0;
            for (int e = 0; e < 12; ++e)
                if (has[e]) s = add(s, p[e]), ++n;
            p[12] = divs(s, (float)n);
        }
        for (const Tri& t : tris) {  // add_faces (:245-272)
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
