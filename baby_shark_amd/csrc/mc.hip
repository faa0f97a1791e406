// Marching cubes on the device (MarchingCubesMesher::mesh,
// src/voxel/meshing/marching_cubes.rs:43-63).
//
// The reference visits the leaves and active tiles serially in tree order and
// runs Custodio et al.'s corrected Marching Cubes 33 on every cube whose eight
// corners have values, with Lewiner et al.'s case tables
// (src/voxel/meshing/lookup_table.rs).  Those tables are not retyped here.  At
// first use the host derives, for every corner-sign configuration and every
// outcome of its ambiguous-face tests, the surface from the cube itself: the
// contour segments of the six faces chain into loops; two loops become one
// tube where the configuration's interior test says so; a loop of eight or
// more crossings is fanned around the centre vertex; the reference's two
// table-indexing behaviours (13.5.2 takes the complementary configuration's
// tiling, 10.1.2 and 12.1.2 after two true face tests take their tube
// reversed) are kept.  The result matches the reference cube by cube in
// crossings, topology and orientation; the diagonal that splits a polygon and
// the order of the triangles inside a cube are this library's own.
//
// Device side: the units (leaves and active tiles) in the reference's visit
// order; one wavefront per unit.  A leaf stages its values and the +1 halo
// (9^3) in LDS; lane (x, y) owns the z-row of cubes with that corner.  A tile
// meets the cubes of its six faces, (i, j) major, faces left, right, top,
// bottom, front, back (the reference's tile visitor, :970-994, which meets
// edge cubes more than once); cubes with every corner inside the tile are all
// one value and emit nothing, so only cubes touching the tile's far faces are
// evaluated.  A counting launch, a scan over units and a writing launch give
// the serial order.
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
// Edge::from_edge_num (lookup_table.rs:54-69): lower corner first; 12 = centre
constexpr int MC_EDGE[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
                                {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
// cube faces 1..6 (test_face numbering, marching_cubes.rs:327-334), corners
// counter-clockwise seen from outside
constexpr int MC_CCW[7][4] = {{0, 0, 0, 0}, {1, 5, 4, 0}, {2, 6, 5, 1}, {3, 7, 6, 2},
                              {0, 4, 7, 3}, {0, 3, 2, 1}, {5, 6, 7, 4}};
// interior_ambiguity (:474-530): face pairs 1|3, 2|4, 5|6 use list 0, 1, 2
constexpr int MC_FAMILY[7] = {2, 0, 1, 0, 1, 2, 2};

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

// ------------------------------------------------------------------ host tables
// Per configuration id (bit i: corner i negative): the ambiguous faces, and
// per pattern of their outcomes (bit j: the negative corners of the j-th
// ambiguous face are joined across it) the interior test to run and the
// triangles without and with a tube.
enum : uint8_t { TEST_NONE = 0, TEST_VERIFY = 1, TEST_CASE13 = 2 };
struct McTables {
    std::vector<uint8_t> amb_mask;  // 256: bit f-1 for ambiguous face f
    std::vector<uint16_t> base;     // 256: first pattern slot
    std::vector<uint8_t> spec;      // per slot: test | arg negative << 2 | list mask << 3
    std::vector<uint32_t> entry;    // per slot * 2 + tube: offset | count << 24 | centre << 28
    std::vector<uint16_t> tris;     // e0 | e1 << 4 | e2 << 8
};

using HTri = std::array<int, 3>;
using HLoop = std::vector<int>;

int h_edge_of(int a, int b) {
    for (int e = 0; e < 12; ++e)
        if ((MC_EDGE[e][0] == a && MC_EDGE[e][1] == b) || (MC_EDGE[e][0] == b && MC_EDGE[e][1] == a)) return e;
    return -1;
}

// contour loops: faces walked counter-clockwise from outside, a segment from a
// positive-to-negative crossing to a negative-to-positive one
std::vector<HLoop> h_loops(const bool* neg, const bool* joined) {
    int next[12];
    for (int& s : next) s = -1;
    for (int f = 1; f <= 6; ++f) {
        const int* q = MC_CCW[f];
        bool sg[4];
        int e[4], nx = 0;
        for (int i = 0; i < 4; ++i) {
            sg[i] = neg[q[i]];
            e[i] = h_edge_of(q[i], q[(i + 1) & 3]);
        }
        for (int i = 0; i < 4; ++i) nx += sg[i] != sg[(i + 1) & 3];
        if (nx == 4) {
            for (int i = 0; i < 4; ++i) {
                if (sg[i] == joined[f]) continue;
                if (sg[i])
                    next[e[(i + 3) & 3]] = e[i];
                else
                    next[e[i]] = e[(i + 3) & 3];
            }
        } else if (nx == 2) {
            int pn = -1, np = -1;
            for (int i = 0; i < 4; ++i) {
                if (!sg[i] && sg[(i + 1) & 3]) pn = e[i];
                if (sg[i] && !sg[(i + 1) & 3]) np = e[i];
            }
            next[pn] = np;
        }
    }
    std::vector<HLoop> loops;
    bool seen[12] = {};
    for (int s = 0; s < 12; ++s) {
        if (next[s] < 0 || seen[s]) continue;
        HLoop l;
        for (int x = s; !seen[x]; x = next[x]) {
            seen[x] = true;
            l.push_back(x);
        }
        loops.push_back(l);
    }
    return loops;
}

void h_fan(const HLoop& l, std::vector<HTri>& out) {
    for (size_t i = 1; i + 1 < l.size(); ++i) out.push_back({l[0], l[i], l[i + 1]});
}
void h_centre_fan(const HLoop& l, std::vector<HTri>& out) {
    for (size_t i = 0; i < l.size(); ++i) out.push_back({l[i], l[(i + 1) % l.size()], 12});
}
// tube: A forward, B backward, each step closes one loop edge; `centre` splits
// the rung between the first vertices through the centre vertex
void h_tube(const HLoop& A, const HLoop& B, bool centre, bool reversed, std::vector<HTri>& out) {
    const size_t n1 = A.size(), n2 = B.size();
    std::vector<int> b(B.rbegin(), B.rend());
    size_t i = 0, k = 0;
    auto put = [&](HTri t) { out.push_back(reversed ? HTri{t[0], t[2], t[1]} : t); };
    while (i < n1 || k < n2) {
        HTri t;
        if (k >= n2 || (i < n1 && (i + 1) * n2 <= (k + 1) * n1)) {
            t = {A[i % n1], A[(i + 1) % n1], b[k % n2]};
            ++i;
        } else {
            t = {b[(k + 1) % n2], b[k % n2], A[i % n1]};
            ++k;
        }
        int ia = -1, ib = -1;
        for (int j = 0; j < 3; ++j) {
            if (t[j] == A[0]) ia = j;
            if (t[j] == b[0]) ib = j;
        }
        if (centre && ia >= 0 && ib >= 0) {
            HTri t1 = t, t2 = t;
            t1[ib] = 12;
            t2[ia] = 12;
            put(t1);
            put(t2);
        } else {
            put(t);
        }
    }
}

int h_loop_corner(const HLoop& l) {
    for (int v = 0; v < 8; ++v) {
        bool all = true;
        for (int e : l) all = all && (MC_EDGE[e][0] == v || MC_EDGE[e][1] == v);
        if (all) return v;
    }
    return -1;
}

// The surface of configuration `id` with face outcomes `joined`: the interior
// test it needs (spec), and its triangles without (t0) and with (t1) a tube.
void h_configure(int id, const bool* joined, uint8_t& spec, std::vector<HTri>& t0, std::vector<HTri>& t1) {
    spec = TEST_NONE;
    bool neg[8];
    int pc = 0;
    for (int i = 0; i < 8; ++i) {
        neg[i] = (id >> i) & 1;
        pc += neg[i];
    }
    if (pc == 0 || pc == 8) return;
    int amb_faces[6], n_amb = 0, n_join = 0;
    for (int f = 1; f <= 6; ++f) {
        const int* q = MC_CCW[f];
        if (neg[q[0]] != neg[q[1]] && neg[q[1]] != neg[q[2]] && neg[q[2]] != neg[q[3]]) {
            amb_faces[n_amb++] = f;
            n_join += joined[f];
        }
    }
    const std::vector<HLoop> loops = h_loops(neg, joined);
    int k = 0, edges_in = 0;
    bool minor[8];
    for (int i = 0; i < 8; ++i) {
        minor[i] = neg[i] == (pc <= 4);
        k += minor[i];
    }
    for (int e = 0; e < 12; ++e) edges_in += minor[MC_EDGE[e][0]] && minor[MC_EDGE[e][1]];
    const bool s0_neg = pc > 4;
    bool tube = false, centre = false, reversed = false;
    if (k == 2 && n_amb == 0 && loops.size() == 2) {  // case 4 (test_interior, :405-419)
        spec = TEST_VERIFY | (s0_neg << 2) | (7 << 3);
        tube = true;
    } else if (k == 3 && edges_in == 1 && loops.size() == 2) {  // case 6.1 (:421-427)
        spec = TEST_VERIFY | (s0_neg << 2) | ((1 << MC_FAMILY[amb_faces[0]]) << 3);
        tube = centre = true;
    } else if (k == 3 && edges_in == 0 && n_join == (pc <= 4 ? 3 : 0)) {  // case 7.4 (:428-444)
        spec = TEST_VERIFY | ((!s0_neg) << 2) | (7 << 3);
        tube = true;
    } else if (k == 4 && edges_in == 2 && loops.size() == 2) {  // cases 10.1, 12.1 (:445-463)
        const bool both = joined[amb_faces[0]];
        spec = TEST_VERIFY | (both << 2) | (((1 << MC_FAMILY[amb_faces[0]]) | (1 << MC_FAMILY[amb_faces[1]])) << 3);
        tube = true;
        reversed = both;
    } else if (k == 4 && edges_in == 0 && loops.size() == 3 && n_join == 3) {  // case 13.5 (:228-238)
        spec = TEST_CASE13;
        // the complementary configuration's tiling: faces 1..4 of the pattern
        // a quarter turn on, every sign flipped
        bool rp[7] = {};
        if (id == 165) {
            rp[2] = joined[1], rp[3] = joined[2], rp[4] = joined[3], rp[1] = joined[4];
        } else {
            rp[1] = joined[2], rp[2] = joined[3], rp[3] = joined[4], rp[4] = joined[1];
        }
        rp[5] = joined[5], rp[6] = joined[6];
        bool np[8];
        for (int i = 0; i < 8; ++i) np[i] = !neg[i];
        const std::vector<HLoop> L = h_loops(np, rp);
        int six = -1, with = -1, alone = -1;
        for (int j = 0; j < (int)L.size(); ++j) {
            if (L[j].size() == 6)
                six = j;
            else if (np[h_loop_corner(L[j])])
                with = j;
            else
                alone = j;
        }
        h_tube(L[six], L[with], false, false, t1);
        h_fan(L[alone], t1);
    }
    for (const HLoop& l : loops) {
        if (l.size() >= 8)
            h_centre_fan(l, t0);
        else
            h_fan(l, t0);
    }
    if (tube) {
        h_tube(loops[0], loops[1], centre, reversed, t1);
        for (size_t j = 2; j < loops.size(); ++j) h_fan(loops[j], t1);
    }
}

const McTables& mc_tables() {
    static McTables T;
    static std::once_flag once;
    std::call_once(once, [] {
        T.amb_mask.assign(256, 0);
        T.base.assign(256, 0);
        for (int id = 0; id < 256; ++id) {
            bool neg[8];
            for (int i = 0; i < 8; ++i) neg[i] = (id >> i) & 1;
            int amb_faces[6], n_amb = 0;
            for (int f = 1; f <= 6; ++f) {
                const int* q = MC_CCW[f];
                if (neg[q[0]] != neg[q[1]] && neg[q[1]] != neg[q[2]] && neg[q[2]] != neg[q[3]]) {
                    amb_faces[n_amb++] = f;
                    T.amb_mask[id] |= (uint8_t)(1 << (f - 1));
                }
            }
            T.base[id] = (uint16_t)T.spec.size();
            for (int p = 0; p < (1 << n_amb); ++p) {
                bool joined[7] = {};
                for (int j = 0; j < n_amb; ++j) joined[amb_faces[j]] = (p >> j) & 1;
                uint8_t spec;
                std::vector<HTri> t[2];
                h_configure(id, joined, spec, t[0], t[1]);
                T.spec.push_back(spec);
                for (int tube = 0; tube < 2; ++tube) {
                    const std::vector<HTri>& tl = (spec == TEST_NONE || !tube) ? t[0] : t[1];
                    bool centre = false;
                    for (const HTri& tr : tl) centre = centre || tr[0] == 12 || tr[1] == 12 || tr[2] == 12;
                    T.entry.push_back((uint32_t)T.tris.size() | ((uint32_t)tl.size() << 24) | ((uint32_t)centre << 28));
                    for (const HTri& tr : tl) T.tris.push_back((uint16_t)(tr[0] | (tr[1] << 4) | (tr[2] << 8)));
                }
            }
        }
    });
    return T;
}

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
    const uint8_t* amb_mask;
    const uint16_t* base;
    const uint8_t* spec;
    const uint32_t* entry;
    const uint16_t* tris;
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

// One cube with a surface (min corner g; c: its eight clamped corner values,
// in LDS, configuration id): the number of kept triangles; with `out`, they
// are written there, 9 floats each.
__device__ uint32_t mc_cube(const McArgs& a, const float* c, int id, int gx, int gy, int gz, float* out) {
    const int pc = __popc(id);
    const uint8_t am = a.amb_mask[id];
    int p = 0, j = 0;
    for (int f = 1; f <= 6; ++f) {
        if (!((am >> (f - 1)) & 1)) continue;
        const bool r = mc_test_face(c, pc <= 4 ? f : -f);
        p |= (int)(pc <= 4 ? r : !r) << j;
        ++j;
    }
    const int slot = a.base[id] + p;
    const uint8_t sp = a.spec[slot];
    int tube = 0;
    if ((sp & 3) == TEST_VERIFY) {
        const float arg = (sp & 4) ? -7.0f : 7.0f;
        int sum = 0;
        for (int l = 0; l < 3; ++l)
            if ((sp >> (3 + l)) & 1) sum += mc_verify(c, mc_amb_edge(c, l, arg));
        tube = sum == 0;
    } else if ((sp & 3) == TEST_CASE13) {
        tube = !mc_interior13(c);
    }
    const uint32_t e = a.entry[slot * 2 + tube];
    const uint32_t off = e & 0xFFFFFFu, cnt = (e >> 24) & 15u;
    float mx = 0.0f, my = 0.0f, mz = 0.0f;
    if ((e >> 28) & 1u) {  // compute_c_vertex (:919-940): the mean of the crossings, summed in edge order
        int n = 0;
        for (int k = 0; k < 12; ++k) {
            if ((c[c_edge[k][0]] < 0.0f) == (c[c_edge[k][1]] < 0.0f)) continue;
            float x, y, z;
            mc_vertex(c, k, gx, gy, gz, x, y, z);
            mx = mx + x, my = my + y, mz = mz + z;
            ++n;
        }
        mx = mx / (float)n, my = my / (float)n, mz = mz / (float)n;
    }
    const float h = a.voxel_size;
    uint32_t kept = 0;
    for (uint32_t t = 0; t < cnt; ++t) {
        const uint16_t tr = a.tris[off + t];
        float v[9];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int ed = (tr >> (4 * q)) & 15;
            float x = mx, y = my, z = mz;
            if (ed != 12) mc_vertex(c, ed, gx, gy, gz, x, y, z);
            v[3 * q] = x * h, v[3 * q + 1] = y * h, v[3 * q + 2] = z * h;
        }
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

// WRITE = false: unit_cnt[u] = kept triangles of unit u, lane_cnt[64 u + l] =
// those of lane l of a leaf unit.  WRITE = true: the triangles at
// out + 9 * unit_off[u].
template <bool WRITE>
__global__ __launch_bounds__(64) void k_mc_units(McArgs a, uint32_t* __restrict__ unit_cnt,
                                                 uint16_t* __restrict__ lane_cnt, const uint64_t* __restrict__ unit_off,
                                                 float* __restrict__ out) {
    __shared__ float val[MB * MB * MB];
    __shared__ uint8_t act[MB * MB * MB];
    __shared__ float cv[64 * CS];
    __shared__ int32_t nb_leaf[8];
    __shared__ float nb_val[8];
    __shared__ uint8_t nb_kind[8];  // 0 none, 1 leaf, 2 tile
    const int lane = threadIdx.x;
    float* mine_c = cv + lane * CS;
    for (uint64_t u = blockIdx.x; u < a.n_units; u += gridDim.x) {
        const int64_t unit = a.units[u];
        uint64_t running = WRITE ? unit_off[u] : 0;
        if (unit >= 0) {
            const int4 org = a.origins[unit];
            if (lane < 8) {  // the leaf and its +x/+y/+z neighbour blocks
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
                nb_leaf[lane] = s;
                nb_kind[lane] = kind;
                nb_val[lane] = tv;
            }
            __syncthreads();
            for (int i = lane; i < MB * MB * MB; i += 64) {
                const int x = i / (MB * MB), y = (i / MB) % MB, z = i % MB;
                const int nb = ((x >> 3) << 2) | ((y >> 3) << 1) | (z >> 3);
                float v = 0.0f;
                bool on = false;
                if (nb_kind[nb] == 1) {
                    const uint64_t s = (uint64_t)nb_leaf[nb];
                    const int o = ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
                    on = (a.masks[s * 8 + (o >> 6)] >> (o & 63)) & 1ull;
                    if (on) v = a.values[s * LEAF_VOX + o];
                } else if (nb_kind[nb] == 2) {
                    on = true;
                    v = nb_val[nb];
                }
                val[i] = mc_clamp(v);
                act[i] = on;
            }
            __syncthreads();
            const int x = lane >> 3, y = lane & 7;
            uint32_t base = 0;
            if (WRITE) {
                uint32_t total;
                base = wave_exclusive(lane_cnt[u * 64 + lane], lane, total);
            }
            uint32_t mine = 0;
            for (int z = 0; z < 8; ++z) {
                const int b0 = (x * MB + y) * MB + z;
                bool full = true;
                int id = 0;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int i = b0 + mc_corner_off(k);
                    full = full && act[i];
                    id |= (val[i] < 0.0f) << k;
                }
                if (!full || id == 0 || id == 255) continue;
#pragma unroll
                for (int k = 0; k < 8; ++k) mine_c[k] = val[b0 + mc_corner_off(k)];
                const uint32_t kept = mc_cube(a, mine_c, id, org.x + x, org.y + y, org.z + z,
                                              WRITE ? out + 9 * (running + base + mine) : nullptr);
                mine += kept;
            }
            if (!WRITE) {
                uint32_t total;
                wave_exclusive(mine, lane, total);
                lane_cnt[u * 64 + lane] = (uint16_t)mine;
                if (lane == 0) unit_cnt[u] = total;
            }
            __syncthreads();
        } else {
            const int4 to = a.tile_org[-1 - unit];
            const float tv = mc_clamp(a.tile_val[-1 - unit]);
            const int s = to.w;
            const uint32_t pairs = (uint32_t)s * (uint32_t)s;
            uint32_t tile_total = 0;
            for (uint32_t q0 = 0; q0 < pairs; q0 += 64) {
                const uint32_t q = q0 + lane;
                const int i = (int)(q / (uint32_t)s), j = (int)(q % (uint32_t)s);
                uint32_t mine = 0, base = 0;
                for (int pass = 0; pass < (WRITE ? 2 : 1); ++pass) {
                    if (pass == 1) {
                        uint32_t total;
                        base = wave_exclusive(mine, lane, total);
                        mine = 0;
                    }
                    for (int f = 0; f < 6 && q < pairs; ++f) {
                        int lx, ly, lz;
                        switch (f) {
                            case 0: lx = 0, ly = i, lz = j; break;
                            case 1: lx = s - 1, ly = i, lz = j; break;
                            case 2: lx = i, ly = j, lz = s - 1; break;
                            case 3: lx = i, ly = j, lz = 0; break;
                            case 4: lx = i, ly = s - 1, lz = j; break;
                            default: lx = i, ly = 0, lz = j; break;
                        }
                        if (lx < s - 1 && ly < s - 1 && lz < s - 1) continue;  // one value throughout
                        bool full = true;
                        int id = 0;
                        for (int k = 0; k < 8 && full; ++k) {
                            const int cx = lx + c_corner[k][0], cy = ly + c_corner[k][1], cz = lz + c_corner[k][2];
                            float v = tv;
                            if (!(cx < s && cy < s && cz < s)) {
                                full = mc_at(a, to.x + cx, to.y + cy, to.z + cz, v);
                                v = mc_clamp(v);
                            }
                            mine_c[k] = v;
                            id |= (v < 0.0f) << k;
                        }
                        if (!full || id == 0 || id == 255) continue;
                        mine += mc_cube(a, mine_c, id, to.x + lx, to.y + ly, to.z + lz,
                                        pass == 1 ? out + 9 * (running + base + mine) : nullptr);
                    }
                }
                uint32_t total;
                wave_exclusive(mine, lane, total);
                running += total;
                tile_total += total;
            }
            if (!WRITE && lane == 0) unit_cnt[u] = tile_total;
        }
    }
}

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
    // host-built tables, one buffer
    const McTables& T = mc_tables();
    const size_t o_base = 256, o_spec = o_base + 2 * 256, o_entry = (o_spec + T.spec.size() + 3) & ~size_t(3),
                 o_tris = o_entry + 4 * T.entry.size(), tab_bytes = o_tris + 2 * T.tris.size();
    std::vector<uint8_t> tab(tab_bytes, 0);
    std::memcpy(tab.data(), T.amb_mask.data(), 256);
    std::memcpy(tab.data() + o_base, T.base.data(), 2 * 256);
    std::memcpy(tab.data() + o_spec, T.spec.data(), T.spec.size());
    std::memcpy(tab.data() + o_entry, T.entry.data(), 4 * T.entry.size());
    std::memcpy(tab.data() + o_tris, T.tris.data(), 2 * T.tris.size());
    uint8_t* dtab = ctx->scratch.as<uint8_t>(S_MC_TAB, tab_bytes);
    BS_HIP(hipMemcpyAsync(dtab, tab.data(), tab_bytes, hipMemcpyHostToDevice, st));

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
    a.amb_mask = dtab;
    a.base = reinterpret_cast<const uint16_t*>(dtab + o_base);
    a.spec = dtab + o_spec;
    a.entry = reinterpret_cast<const uint32_t*>(dtab + o_entry);
    a.tris = reinterpret_cast<const uint16_t*>(dtab + o_tris);
    a.voxel_size = voxel_size;

    const unsigned grid = (unsigned)std::min<uint64_t>(n_units, 16384);
    uint32_t* cnt = ctx->scratch.as<uint32_t>(S_DC_TRICNT, n_units);
    uint64_t* off = ctx->scratch.as<uint64_t>(S_DC_TRIOFF, n_units + 1);
    uint16_t* lane_cnt = ctx->scratch.as<uint16_t>(S_MC_LANES, n_units * 64);
    k_mc_units<false><<<grid, 64, 0, st>>>(a, cnt, lane_cnt, nullptr, nullptr);
    BS_HIP(hipGetLastError());
    mc_scan(ctx, cnt, off, n_units);
    uint32_t h_fail = 0, last_cnt = 0;
    uint64_t last_off = 0;
    BS_HIP(hipMemcpyAsync(&h_fail, fail, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    BS_HIP(hipMemcpyAsync(&last_cnt, cnt + n_units - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    BS_HIP(hipMemcpyAsync(&last_off, off + n_units - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    BS_HIP(hipStreamSynchronize(st));
    BS_REQUIRE(!h_fail, BS_ERR_INTERNAL, "marching cubes: a leaf or tile key could not be registered");
    const uint64_t total = last_off + last_cnt;
    if (total) {
        float* out = nullptr;
        BS_HIP(hipMalloc(&out, sizeof(float) * 9 * total));
        k_mc_units<true><<<grid, 64, 0, st>>>(a, nullptr, lane_cnt, off, out);
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
