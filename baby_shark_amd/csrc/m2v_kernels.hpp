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
    return sqrt_rn(m);
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
            float v = div_rn(d1, d1 - d3);
            q = f3_add(A, f3_muls(ab, v));
        } else {
            F3 cp = f3_sub(p, C);
            float d5 = f3_dot_exact(ab, cp), d6 = f3_dot_exact(ac, cp);
            float vb = d5 * d2 - d1 * d6;
            if (d6 >= 0.0f && d5 <= d6) {
                q = C;
            } else if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
                float w = div_rn(d2, d2 - d6);
                q = f3_add(A, f3_muls(ac, w));
            } else {
                float va = d3 * d6 - d5 * d4;
                if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
                    float w = div_rn(d4 - d3, (d4 - d3) + (d5 - d6));
                    q = f3_add(B, f3_muls(f3_sub(C, B), w));
                } else {
                    float denom = div_rn(1.0f, (va + vb) + vc);
                    float v = vb * denom, w = vc * denom;
                    q = f3_add(f3_add(A, f3_muls(ab, v)), f3_muls(ac, w));
                }
            }
        }
    }
    F3 d = f3_sub(q, p);
    return sqrt_rn(f3_dot_exact(d, d));
}

// Per sub-triangle record for the distance kernel (112 B, read wave-uniformly
// through scalar loads).  The triangle is expressed in its own frame: origin
// a, axes U = (b-a)/|b-a|, V = N x U, N = unit normal; there b = (L0, 0) and
// c = (cu, cv) with cv > 0.  A degenerate sub-triangle (zero area or a zero
// first edge) has flags = 1 and is evaluated with the reference's own
// closest-point arithmetic instead.
struct alignas(16) DistRec {
    float a[3];
    float U[3], V[3], N[3];
    float L0, inv0;       // |b-a|, 1/|b-a|
    float cu, cv;         // c in the frame
    float e1x, e1y, inv1; // c-b in the frame, 1/|c-b|^2
    float inv2;           // 1/|a-c|^2
    int box[6];           // dilated voxel box lo.xyz, hi.xyz (inclusive)
    int flags;
    int pad;
};
static_assert(sizeof(DistRec) == 112, "DistRec layout");

// Every field of a DistRec as a scalar-register operand of an empty asm: all
// of the record's loads are issued before it, one wait for the lot.
__device__ inline void pin_dist_rec(const DistRec& r) {
    asm volatile("" ::"s"(r.a[0]), "s"(r.a[1]), "s"(r.a[2]), "s"(r.U[0]), "s"(r.U[1]), "s"(r.U[2]), "s"(r.V[0]),
                 "s"(r.V[1]), "s"(r.V[2]), "s"(r.N[0]), "s"(r.N[1]), "s"(r.N[2]), "s"(r.L0), "s"(r.inv0), "s"(r.cu),
                 "s"(r.cv), "s"(r.e1x), "s"(r.e1y), "s"(r.inv1), "s"(r.inv2), "s"(r.box[0]), "s"(r.box[1]),
                 "s"(r.box[2]), "s"(r.box[3]), "s"(r.box[4]), "s"(r.box[5]), "s"(r.flags));
}

// The fields of a DistRec that the per-voxel step reads, held in scalar
// registers: plane_rec's empty asm takes each as an SGPR operand, so their
// loads are all issued (and waited for) once per candidate.
struct PlaneRec {
    float a0, U0, V0, N0, L0, inv0, cu, cv, e1x, e1y, inv1, inv2;
};

__device__ inline PlaneRec plane_rec(const DistRec& r) {
    PlaneRec q{r.a[0], r.U[0], r.V[0], r.N[0], r.L0, r.inv0, r.cu, r.cv, r.e1x, r.e1y, r.inv1, r.inv2};
    asm volatile("" ::"s"(q.a0), "s"(q.U0), "s"(q.V0), "s"(q.N0), "s"(q.L0), "s"(q.inv0), "s"(q.cu), "s"(q.cv),
                 "s"(q.e1x), "s"(q.e1y), "s"(q.inv1), "s"(q.inv2));
    return q;
}

// In-plane squared distance from (pu, pv) to the frame triangle (0,0), (L0,0),
// (cu,cv): 0 inside, else the nearest clamped edge projection.  The three edge
// terms are computed for every lane (the empty asm keeps them out of a branch
// on `inside`, which would cost more in mask handling than it saves).
__device__ inline float plane_dist2(const PlaneRec& r, float pu, float pv) {
    const float t0 = __builtin_amdgcn_fmed3f(pu * r.inv0, 0.0f, 1.0f);
    const float qx = __builtin_fmaf(-t0, r.L0, pu);
    const float d0 = __builtin_fmaf(qx, qx, pv * pv);
    const float rx = pu - r.L0;
    const float t1 = __builtin_amdgcn_fmed3f(__builtin_fmaf(rx, r.e1x, pv * r.e1y) * r.inv1, 0.0f, 1.0f);
    const float sx = __builtin_fmaf(-t1, r.e1x, rx), sy = __builtin_fmaf(-t1, r.e1y, pv);
    const float d1 = __builtin_fmaf(sx, sx, sy * sy);
    const float wx = pu - r.cu, wy = pv - r.cv;
    const float t2 = __builtin_amdgcn_fmed3f(-__builtin_fmaf(wx, r.cu, wy * r.cv) * r.inv2, 0.0f, 1.0f);
    const float ux = __builtin_fmaf(t2, r.cu, wx), uy = __builtin_fmaf(t2, r.cv, wy);
    const float d2 = __builtin_fmaf(ux, ux, uy * uy);
    const float f1 = __builtin_fmaf(r.e1x, pv, -r.e1y * rx);   // cross(c-b, q-b)
    const float f2 = __builtin_fmaf(r.cv, wx, -r.cu * wy);     // cross(a-c, q-c)
    const bool inside = fminf(fminf(pv, f1), f2) >= 0.0f;
    float m = fminf(fminf(d0, d1), d2);
    asm volatile("" : "+v"(m));
    return inside ? 0.0f : m;
}

// ---------------------------------------------------------------- winding numbers
// atan(a) = a * P(a^2) on [0, 1]: a degree-6 polynomial in a^2 fitted here by
// iteratively reweighted least squares (max error 3.2e-7 rad in f32), with the
// usual octant reduction.  The sign decision it feeds has a margin of ~0.2.
__device__ inline float fast_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mn * __builtin_amdgcn_rcpf(mx);
    const float s = a * a;
    float p = 0.006811828352510929f;
    p = __builtin_fmaf(p, s, -0.03360431268811226f);
    p = __builtin_fmaf(p, s, 0.0796237587928772f);
    p = __builtin_fmaf(p, s, -0.13233345746994019f);
    p = __builtin_fmaf(p, s, 0.19807815551757812f);
    p = __builtin_fmaf(p, s, -0.3331736922264099f);
    p = __builtin_fmaf(p, s, 0.9999961256980896f);
    float r = p * a;
    r = (ay > ax) ? 1.5707963267948966f - r : r;
    r = (x < 0.0f) ? 3.141592653589793f - r : r;
    return (y < 0.0f) ? -r : r;
}

__device__ inline float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }

}  // namespace bs
