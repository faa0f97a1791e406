// Test infrastructure only: the CPU oracle.  Nothing outside tests/, the smoke
// check and bench.py's cpu_baseline leg may link or call this code.
//
// f32/f64 vector helpers whose every operation follows the evaluation order of
// nalgebra 0.34 as used by baby_shark (Cargo.toml:31): component-wise +,-,*,/,
// dot = (x*x' + y*y') + z*z', cross = (ay*bz - az*by, az*bx - ax*bz,
// ax*by - ay*bx), norm = sqrt(dot(v, v)).  The oracle is compiled with
// -ffp-contract=off so no product is fused into an add, as in Rust.
#pragma once
#include <cmath>
#include <cstdint>

namespace oracle {

template <class T>
struct V3 {
    T x, y, z;
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
using Vec3f = V3<float>;
using Vec3d = V3<double>;

struct Vec3i {
    int64_t x, y, z;
    int64_t operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    int64_t& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    bool operator==(const Vec3i& o) const { return x == o.x && y == o.y && z == o.z; }
    bool operator!=(const Vec3i& o) const { return !(*this == o); }
    bool operator<(const Vec3i& o) const {
        if (x != o.x) return x < o.x;
        if (y != o.y) return y < o.y;
        return z < o.z;
    }
};

template <class T> inline V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> mul(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> smul(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> divs(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) {
    T p0 = a.x * b.x;
    T p1 = a.y * b.y;
    T p2 = a.z * b.z;
    return (p0 + p1) + p2;
}
template <class T> inline T norm_sq(V3<T> a) { return dot(a, a); }
template <class T> inline T norm(V3<T> a) { return std::sqrt(norm_sq(a)); }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// nalgebra_glm::min2/max2: component-wise `if a < b {a} else {b}` (and `>`).
template <class T> inline T min2s(T a, T b) { return a < b ? a : b; }
template <class T> inline T max2s(T a, T b) { return a > b ? a : b; }
template <class T> inline V3<T> min2(V3<T> a, V3<T> b) { return {min2s(a.x, b.x), min2s(a.y, b.y), min2s(a.z, b.z)}; }
template <class T> inline V3<T> max2(V3<T> a, V3<T> b) { return {max2s(a.x, b.x), max2s(a.y, b.y), max2s(a.z, b.z)}; }

// Rust f32::max / f32::min ignore a NaN operand.
inline float rmax(float a, float b) {
    if (std::isnan(a)) return b;
    if (std::isnan(b)) return a;
    return a > b ? a : b;
}
inline float rmin(float a, float b) {
    if (std::isnan(a)) return b;
    if (std::isnan(b)) return a;
    return a < b ? a : b;
}
// Rust `f as isize`: saturating, NaN -> 0.
inline int64_t f2i(float f) {
    if (std::isnan(f)) return 0;
    if (f >= 9.2233720368547758e18f) return INT64_MAX;
    if (f <= -9.2233720368547758e18f) return INT64_MIN;
    return (int64_t)f;
}
// Rust `f as usize`: saturating at 0, NaN -> 0.
inline uint64_t f2u(float f) {
    if (std::isnan(f) || f <= 0.0f) return 0;
    if (f >= 1.8446744073709552e19f) return UINT64_MAX;
    return (uint64_t)f;
}

// Mat3 in column-major storage m[col][row], like nalgebra.
struct Mat3f {
    float m[3][3];  // m[c][r]
    float at(int r, int c) const { return m[c][r]; }
    float& at(int r, int c) { return m[c][r]; }
};
inline Mat3f mat_zero() {
    Mat3f z;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) z.m[c][r] = 0.0f;
    return z;
}
// u * v^T
inline Mat3f outer(Vec3f u, Vec3f v) {
    Mat3f o;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) o.at(r, c) = u[r] * v[c];
    return o;
}
inline Mat3f mat_add(const Mat3f& a, const Mat3f& b) {
    Mat3f o;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) o.m[c][r] = a.m[c][r] + b.m[c][r];
    return o;
}
inline Mat3f mat_sub(const Mat3f& a, const Mat3f& b) {
    Mat3f o;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) o.m[c][r] = a.m[c][r] - b.m[c][r];
    return o;
}
// nalgebra Matrix::dot for a 3x3: `res += a(k,j)*b(k,j)` column by column.
inline float mat_dot(const Mat3f& a, const Mat3f& b) {
    float res = 0.0f;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) res += a.m[c][r] * b.m[c][r];
    return res;
}

}  // namespace oracle
