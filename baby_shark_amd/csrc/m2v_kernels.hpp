// Device code of mesh -> narrow-band SDF (src/voxel/mesh_to_volume.rs).
// Everything that decides the active voxel set or a voxel's final value runs
// without FMA contraction and with correctly rounded division and square
// root, in the operation order of the reference, so that sub-triangles, boxes,
// grid points and the selected distance are the reference's f32 values.
#pragma once
#include "common.hpp"

namespace bs {

struct F3 {
    float x, y, z;
};

__device__ inline F3 f3_sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline F3 f3_add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline F3 f3_muls(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
// nalgebra dot of 3-vectors: (x*x' + y*y') + z*z', no fusion.
__device__ inline float f3_dot_exact(F3 a, F3 b) {
#pragma clang fp contract(off)
    float p0 = a.x * b.x;
    float p1 = a.y * b.y;
    float p2 = a.z * b.z;
    return (p0 + p1) + p2;
}
__device__ inline F3 f3_min2(F3 a, F3 b) { return {a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z}; }
__device__ inline F3 f3_max2(F3 a, F3 b) { return {a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z}; }
__device__ inline float rust_max(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}

// Triangle3::max_side (triangle3.rs:117-124)
__device__ inline float tri_max_side(F3 a, F3 b, F3 c) {
    F3 ab = f3_sub(b, a), ac = f3_sub(c, a), bc = f3_sub(c, b);
    float m = rust_max(rust_max(f3_dot_exact(ab, ab), f3_dot_exact(ac, ac)), f3_dot_exact(bc, bc));
    return __fsqrt_rn(m);
}

// Triangle3::closest_point (triangle3.rs:317-382) followed by
// `(closest - grid_point).norm()` (mesh_to_volume.rs:154-155), bit-faithful.
__device__ inline float ericson_distance(F3 A, F3 B, F3 C, F3 p) {
#pragma clang fp contract(off)
    F3 ab = f3_sub(B, A), ac = f3_sub(C, A), ap = f3_sub(p, A);
    float d1 = f3_dot_exact(ab, ap), d2 = f3_dot_exact(ac, ap);
    F3 q;
    if (d1 <= 0.0f && d2 <= 0.0f) {
        q = A;
    } else {
        F3 bp = f3_sub(p, B);
        float d3 = f3_dot_exact(ab, bp), d4 = f3_dot_exact(ac, bp);
        float vc = d1 * d4 - d3 * d2;
        if (d3 >= 0.0f && d4 <= d3) {
            q = B;
        } else if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
            float v = __fdiv_rn(d1, d1 - d3);
            q = f3_add(A, f3_muls(ab, v));
        } else {
            F3 cp = f3_sub(p, C);
            float d5 = f3_dot_exact(ab, cp), d6 = f3_dot_exact(ac, cp);
            float vb = d5 * d2 - d1 * d6;
            if (d6 >= 0.0f && d5 <= d6) {
                q = C;
            } else if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
                float w = __fdiv_rn(d2, d2 - d6);
                q = f3_add(A, f3_muls(ac, w));
            } else {
                float va = d3 * d6 - d5 * d4;
                if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
                    float w = __fdiv_rn(d4 - d3, (d4 - d3) + (d5 - d6));
                    q = f3_add(B, f3_muls(f3_sub(C, B), w));
                } else {
                    float denom = __fdiv_rn(1.0f, (va + vb) + vc);
                    float v = vb * denom, w = vc * denom;
                    q = f3_add(f3_add(A, f3_muls(ab, v)), f3_muls(ac, w));
                }
            }
        }
    }
    F3 d = f3_sub(q, p);
    return __fsqrt_rn(f3_dot_exact(d, d));
}

// Per sub-triangle record for the distance kernel: everything the branch-free
// squared distance needs, so a (voxel, sub-triangle) pair costs only the
// per-voxel arithmetic.  40 words, read wave-uniformly (scalar loads).
struct alignas(16) SubRec {
    float v[3][3];    // vertices a, b, c
    float e[3][3];    // edges b-a, c-b, a-c
    float m[3][3];    // inward edge normals n x e_k (unnormalised)
    float n[3];       // unit normal (0 for a degenerate triangle)
    float inv_l2[3];  // 1/|e_k|^2 (0 for a zero-length edge)
    int box[6];       // dilated voxel box lo.xyz, hi.xyz (inclusive)
    float bias;       // 0, or -1 for a degenerate triangle (never "inside")
};
static_assert(sizeof(SubRec) == 160, "SubRec layout");

// Squared distance from p to the triangle of `r`: the face-plane distance when
// p projects inside, else the nearest of the three clamped edge projections.
// Only used to choose the arg-min sub-triangle; the stored value comes from
// ericson_distance on that sub-triangle.
__device__ inline float fast_dist2(const SubRec& r, float px, float py, float pz) {
    float best_e = 3.4e38f;
    float s[3];
    float face = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float dx = px - r.v[k][0], dy = py - r.v[k][1], dz = pz - r.v[k][2];
        s[k] = __builtin_fmaf(r.m[k][0], dx, __builtin_fmaf(r.m[k][1], dy, __builtin_fmaf(r.m[k][2], dz, r.bias)));
        if (k == 0) face = __builtin_fmaf(r.n[0], dx, __builtin_fmaf(r.n[1], dy, r.n[2] * dz));
        float u = __builtin_fmaf(r.e[k][0], dx, __builtin_fmaf(r.e[k][1], dy, r.e[k][2] * dz));
        float t = __builtin_amdgcn_fmed3f(u * r.inv_l2[k], 0.0f, 1.0f);
        float qx = __builtin_fmaf(-t, r.e[k][0], dx);
        float qy = __builtin_fmaf(-t, r.e[k][1], dy);
        float qz = __builtin_fmaf(-t, r.e[k][2], dz);
        float d2 = __builtin_fmaf(qx, qx, __builtin_fmaf(qy, qy, qz * qz));
        best_e = fminf(best_e, d2);
    }
    bool inside = fminf(fminf(s[0], s[1]), s[2]) >= 0.0f;
    return inside ? face * face : best_e;
}

}  // namespace bs
