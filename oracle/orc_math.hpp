// Oracle: CPU restatement of the rbps math layer, templated on the scalar type.
//
// TEST INFRASTRUCTURE ONLY.  This directory is the checker for the GPU
// pipeline; nothing in the product (rbps_amd/, include/) may include or link it.
//
// Every operation keeps the reference's operand order so that the double
// instantiation reproduces the reference bit-for-bit when compiled without
// FMA contraction (-ffp-contract=off), and the float instantiation is the
// arithmetic the GPU kernels perform.
//
// Reference: include/math3d/core.hpp, vec3.hpp, quat.hpp, mat3.hpp, tf.hpp.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>

namespace orc {

template <class T> struct C {
    static constexpr T EPS = T(1e-9);          // core.hpp:49 EPSILON
    static constexpr T ULP_THRESH = T(2.107342e-08);  // quat.hpp:157 Padé branch
};

// std::min / std::max / m3d::clamp semantics (tie and signed-zero behaviour kept).
template <class T> inline T smin(T a, T b) { return (b < a) ? b : a; }
template <class T> inline T smax(T a, T b) { return (a < b) ? b : a; }
template <class T> inline T sclamp(T v, T lo, T hi) { return (v < lo) ? lo : ((v > hi) ? hi : v); }

template <class T> struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    explicit V3(T d) : x(d), y(d), z(d) {}
    V3 operator+(const V3 &r) const { return {x + r.x, y + r.y, z + r.z}; }
    V3 operator-(const V3 &r) const { return {x - r.x, y - r.y, z - r.z}; }
    V3 operator*(T s) const { return {x * s, y * s, z * s}; }
    // vec3.hpp:56-60: division is a multiply by the reciprocal.
    V3 operator/(T s) const { T inv = T(1.0) / s; return {x * inv, y * inv, z * inv}; }
    V3 operator-() const { return {-x, -y, -z}; }
    T &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    const T &operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> inline V3<T> operator*(T s, const V3<T> &v) { return {v.x * s, v.y * s, v.z * s}; }
template <class T> inline T dot(const V3<T> &a, const V3<T> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(const V3<T> &a, const V3<T> &b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T length_sq(const V3<T> &v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
template <class T> inline T length(const V3<T> &v) { return std::sqrt(length_sq(v)); }
// vec3.hpp:196-203
template <class T> inline V3<T> normalize(const V3<T> &v) {
    T l2 = length_sq(v);
    if (l2 < C<T>::EPS) return {T(0), T(0), T(0)};
    T inv = T(1.0) / std::sqrt(l2);
    return v * inv;
}

template <class T> struct Q {
    T w, x, y, z;
    Q() : w(1), x(0), y(0), z(0) {}
    Q(T a, T b, T c, T d) : w(a), x(b), y(c), z(d) {}
    // quat.hpp:91-99 Hamilton product
    Q operator*(const Q &r) const {
        Q o;
        o.w = w * r.w - x * r.x - y * r.y - z * r.z;
        o.x = w * r.x + x * r.w + y * r.z - z * r.y;
        o.y = w * r.y + y * r.w + z * r.x - x * r.z;
        o.z = w * r.z + z * r.w + x * r.y - y * r.x;
        return o;
    }
    Q operator*(T s) const { return {w * s, x * s, y * s, z * s}; }
    Q operator+(const Q &r) const { return {w + r.w, x + r.x, y + r.y, z + r.z}; }
};
// quat.hpp:153-167
template <class T> inline Q<T> normalize(const Q<T> &q) {
    T m2 = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
    if (std::abs(T(1.0) - m2) < C<T>::ULP_THRESH) return q * (T(2.0) / (T(1.0) + m2));
    if (m2 < C<T>::EPS) return {T(1), T(0), T(0), T(0)};
    return q * (T(1.0) / std::sqrt(m2));
}
template <class T> inline Q<T> conj(const Q<T> &q) { return {q.w, -q.x, -q.y, -q.z}; }
// quat.hpp:188-199 cross-product form
template <class T> inline V3<T> rotate(const Q<T> &q, const V3<T> &v) {
    V3<T> u(q.x, q.y, q.z);
    T s = q.w;
    V3<T> uv = cross(u, v);
    V3<T> uuv = cross(u, uv);
    return v + ((uv * s) + uuv) * T(2.0);
}

// Column-major 3x3 (mat3.hpp:29-110)
template <class T> struct M3 {
    V3<T> c[3];
    M3() { c[0] = {1, 0, 0}; c[1] = {0, 1, 0}; c[2] = {0, 0, 1}; }
    V3<T> operator*(const V3<T> &v) const {
        return {c[0].x * v.x + c[1].x * v.y + c[2].x * v.z,
                c[0].y * v.x + c[1].y * v.y + c[2].y * v.z,
                c[0].z * v.x + c[1].z * v.y + c[2].z * v.z};
    }
    M3 operator*(const M3 &r) const { M3 o; for (int k = 0; k < 3; ++k) o.c[k] = (*this) * r.c[k]; return o; }
    M3 transpose() const {
        M3 o;
        o.c[0] = {c[0].x, c[1].x, c[2].x};
        o.c[1] = {c[0].y, c[1].y, c[2].y};
        o.c[2] = {c[0].z, c[1].z, c[2].z};
        return o;
    }
};
// mat3.hpp:166-178
template <class T> inline M3<T> mat3_cast(const Q<T> &q) {
    M3<T> m;
    T x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    T xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
    T yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    T wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    m.c[0] = {T(1) - (yy + zz), xy + wz, xz - wy};
    m.c[1] = {xy - wz, T(1) - (xx + zz), yz + wx};
    m.c[2] = {xz + wy, yz - wx, T(1) - (xx + yy)};
    return m;
}

// Symmetric 3x3 (mat3.hpp:190-300)
template <class T> struct SM {
    T xx, yy, zz, xy, xz, yz;
    SM() : xx(0), yy(0), zz(0), xy(0), xz(0), yz(0) {}
    SM(T a, T b, T c, T d, T e, T f) : xx(a), yy(b), zz(c), xy(d), xz(e), yz(f) {}
    static SM sym(const M3<T> &m) {   // smat3(const mat3&) mat3.hpp:214-218
        return SM(m.c[0].x, m.c[1].y, m.c[2].z, T(0.5) * (m.c[1].x + m.c[0].y),
                  T(0.5) * (m.c[2].x + m.c[0].z), T(0.5) * (m.c[2].y + m.c[1].z));
    }
    V3<T> operator*(const V3<T> &v) const {
        return {xx * v.x + xy * v.y + xz * v.z, xy * v.x + yy * v.y + yz * v.z,
                xz * v.x + yz * v.y + zz * v.z};
    }
    SM operator*(T s) const { return SM(xx * s, yy * s, zz * s, xy * s, xz * s, yz * s); }
    T det() const { return xx * yy * zz + T(2) * xy * xz * yz - xx * yz * yz - yy * xz * xz - zz * xy * xy; }
    SM inverse() const {   // mat3.hpp:246-266
        T d = det();
        if (std::abs(d) < C<T>::EPS) return SM();
        T inv = T(1.0) / d;
        SM r;
        r.xx = (yy * zz - yz * yz) * inv;
        r.yy = (xx * zz - xz * xz) * inv;
        r.zz = (xx * yy - xy * xy) * inv;
        r.xy = (xz * yz - xy * zz) * inv;
        r.xz = (xy * yz - xz * yy) * inv;
        r.yz = (xy * xz - xx * yz) * inv;
        return r;
    }
};
template <class T> inline SM<T> operator*(T s, const SM<T> &m) { return m * s; }
// mat3 * smat3 (mat3.hpp:284-296)
template <class T> inline M3<T> operator*(const M3<T> &l, const SM<T> &r) {
    M3<T> o;
    o.c[0] = l * V3<T>{r.xx, r.xy, r.xz};
    o.c[1] = l * V3<T>{r.xy, r.yy, r.yz};
    o.c[2] = l * V3<T>{r.xz, r.yz, r.zz};
    return o;
}
// World-frame tensor R * I * R^T, symmetrised (Body.cpp:11-16).
template <class T> inline SM<T> world_tensor(const M3<T> &R, const SM<T> &I) {
    return SM<T>::sym((R * I) * R.transpose());
}

template <class T> struct TF {
    V3<T> pos;
    Q<T> rot;
    V3<T> point(const V3<T> &p) const { return pos + rotate(rot, p); }
    V3<T> vec(const V3<T> &d) const { return rotate(rot, d); }
    V3<T> inv_vec(const V3<T> &d) const { return rotate(conj(rot), d); }
};

template <class T> struct AABB {
    V3<T> mn, mx;
};
// AABB.hpp:54-59 (inclusive)
template <class T> inline bool overlap(const AABB<T> &a, const AABB<T> &b) {
    return (a.mn.x <= b.mx.x && a.mx.x >= b.mn.x) && (a.mn.y <= b.mx.y && a.mx.y >= b.mn.y) &&
           (a.mn.z <= b.mx.z && a.mx.z >= b.mn.z);
}

}  // namespace orc
