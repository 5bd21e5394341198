// FP32 vector / quaternion / symmetric-tensor algebra for the XPBD pipeline,
// usable from host and device code.
//
// The operand order of every expression follows the reference math layer
// (include/math3d/vec3.hpp, quat.hpp, mat3.hpp) so that results are reproducible
// against a CPU evaluation of the same float expressions.  The library is built
// with -ffp-contract=off and correctly rounded division / square root.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define RB_HD __host__ __device__ __forceinline__

namespace rb {

constexpr float EPS = 1e-9f;                 // math3d/core.hpp:49
constexpr float QUAT_ULP = 2.107342e-08f;    // math3d/quat.hpp:157

RB_HD float fmin_ref(float a, float b) { return (b < a) ? b : a; }   // std::min
RB_HD float fmax_ref(float a, float b) { return (a < b) ? b : a; }   // std::max
RB_HD float clampf(float v, float lo, float hi) { return (v < lo) ? lo : ((v > hi) ? hi : v); }
RB_HD float absf(float x) { return fabsf(x); }

struct V3 {
    float x, y, z;
};
RB_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
RB_HD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RB_HD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RB_HD V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
RB_HD V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
RB_HD V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
RB_HD V3 vdiv(V3 a, float s) { float inv = 1.0f / s; return {a.x * inv, a.y * inv, a.z * inv}; }  // vec3.hpp:56-60
RB_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RB_HD V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
RB_HD float len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
RB_HD float len(V3 a) { return sqrtf(len2(a)); }
RB_HD V3 normalize(V3 v) {
    float l2 = len2(v);
    if (l2 < EPS) return {0.0f, 0.0f, 0.0f};
    float inv = 1.0f / sqrtf(l2);
    return v * inv;
}
RB_HD float comp(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

struct Q {
    float w, x, y, z;
};
RB_HD Q qmul(Q a, Q r) {
    Q o;
    o.w = a.w * r.w - a.x * r.x - a.y * r.y - a.z * r.z;
    o.x = a.w * r.x + a.x * r.w + a.y * r.z - a.z * r.y;
    o.y = a.w * r.y + a.y * r.w + a.z * r.x - a.x * r.z;
    o.z = a.w * r.z + a.z * r.w + a.x * r.y - a.y * r.x;
    return o;
}
RB_HD Q qadd(Q a, Q b) { return {a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z}; }
RB_HD Q qscale(Q a, float s) { return {a.w * s, a.x * s, a.y * s, a.z * s}; }
RB_HD Q conj(Q q) { return {q.w, -q.x, -q.y, -q.z}; }
RB_HD Q qnormalize(Q q) {
    float m2 = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
    if (absf(1.0f - m2) < QUAT_ULP) return qscale(q, 2.0f / (1.0f + m2));
    if (m2 < EPS) return {1.0f, 0.0f, 0.0f, 0.0f};
    return qscale(q, 1.0f / sqrtf(m2));
}
RB_HD V3 rotate(Q q, V3 v) {
    V3 u{q.x, q.y, q.z};
    V3 uv = cross(u, v);
    V3 uuv = cross(u, uv);
    return v + ((uv * q.w) + uuv) * 2.0f;
}
// q <- normalize(q + (0, d) * q): the orientation update shared by integration and
// positional impulses (Body.cpp:36-37, 66-67).
RB_HD Q qstep(Q q, V3 d) { return qnormalize(qadd(q, qmul(Q{0.0f, d.x, d.y, d.z}, q))); }

struct M3 {
    V3 c0, c1, c2;   // columns
};
RB_HD V3 mv(const M3 &m, V3 v) {
    return {m.c0.x * v.x + m.c1.x * v.y + m.c2.x * v.z, m.c0.y * v.x + m.c1.y * v.y + m.c2.y * v.z,
            m.c0.z * v.x + m.c1.z * v.y + m.c2.z * v.z};
}
RB_HD M3 mm(const M3 &a, const M3 &b) { return {mv(a, b.c0), mv(a, b.c1), mv(a, b.c2)}; }
RB_HD M3 transpose(const M3 &m) {
    return {{m.c0.x, m.c1.x, m.c2.x}, {m.c0.y, m.c1.y, m.c2.y}, {m.c0.z, m.c1.z, m.c2.z}};
}
RB_HD M3 mat3_cast(Q q) {   // mat3.hpp:166-178
    float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
    float yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    M3 m;
    m.c0 = {1.0f - (yy + zz), xy + wz, xz - wy};
    m.c1 = {xy - wz, 1.0f - (xx + zz), yz + wx};
    m.c2 = {xz + wy, yz - wx, 1.0f - (xx + yy)};
    return m;
}

struct SM {
    float xx, yy, zz, xy, xz, yz;
};
RB_HD V3 smv(const SM &m, V3 v) {
    return {m.xx * v.x + m.xy * v.y + m.xz * v.z, m.xy * v.x + m.yy * v.y + m.yz * v.z,
            m.xz * v.x + m.yz * v.y + m.zz * v.z};
}
RB_HD SM sscale(const SM &m, float s) { return {m.xx * s, m.yy * s, m.zz * s, m.xy * s, m.xz * s, m.yz * s}; }
RB_HD SM sinverse(const SM &m) {   // mat3.hpp:246-266
    float d = m.xx * m.yy * m.zz + 2.0f * m.xy * m.xz * m.yz - m.xx * m.yz * m.yz - m.yy * m.xz * m.xz - m.zz * m.xy * m.xy;
    if (absf(d) < EPS) return {0, 0, 0, 0, 0, 0};
    float inv = 1.0f / d;
    return {(m.yy * m.zz - m.yz * m.yz) * inv, (m.xx * m.zz - m.xz * m.xz) * inv, (m.xx * m.yy - m.xy * m.xy) * inv,
            (m.xz * m.yz - m.xy * m.zz) * inv, (m.xy * m.yz - m.xz * m.yy) * inv, (m.xy * m.xz - m.xx * m.yz) * inv};
}
// sym((R * I) * R^T) (Body.cpp:11-16 with mat3.hpp:214-218, 284-296)
RB_HD SM world_tensor(const M3 &R, const SM &I) {
    M3 RI{mv(R, V3{I.xx, I.xy, I.xz}), mv(R, V3{I.xy, I.yy, I.yz}), mv(R, V3{I.xz, I.yz, I.zz})};
    M3 W = mm(RI, transpose(R));
    return {W.c0.x, W.c1.y, W.c2.z, 0.5f * (W.c1.x + W.c0.y), 0.5f * (W.c2.x + W.c0.z), 0.5f * (W.c2.y + W.c1.z)};
}

struct TF {
    V3 pos;
    Q rot;
};
RB_HD V3 tf_point(const TF &t, V3 p) { return t.pos + rotate(t.rot, p); }
RB_HD V3 tf_vec(const TF &t, V3 d) { return rotate(t.rot, d); }
RB_HD V3 tf_inv_vec(const TF &t, V3 d) { return rotate(conj(t.rot), d); }

struct Box3 {
    V3 mn, mx;
};

// ---------------------------------------------------------------- sin, cos, atan2
// Built only from IEEE +, -, *, / and comparisons so that a CPU evaluation of the same
// expressions reproduces them bit for bit.  Argument reduction x = k*(pi/2) + r with
// pi/2 split in three parts (Cody & Waite 1980): 1.5703125 + 4.837512969970703e-4 +
// 7.549790126404332e-8 (k*P1, k*P2 exact for |k| < 2^13); Taylor polynomials of sin
// and cos on |r| <= pi/4 through r^11 / r^12; atan via its Taylor series through t^19
// for t <= tan(pi/8) and atan(t) = pi/4 + atan((t-1)/(t+1)) above, then the octant
// fix-up.  A few ulp from the correctly rounded values.
RB_HD float rb_sin_poly(float r) {
    const float r2 = r * r;
    float p = -1.0f / 39916800.0f;
    p = p * r2 + 1.0f / 362880.0f;
    p = p * r2 + -1.0f / 5040.0f;
    p = p * r2 + 1.0f / 120.0f;
    p = p * r2 + -1.0f / 6.0f;
    return r + (r * r2) * p;
}
RB_HD float rb_cos_poly(float r) {
    const float r2 = r * r;
    float p = 1.0f / 479001600.0f;
    p = p * r2 + -1.0f / 3628800.0f;
    p = p * r2 + 1.0f / 40320.0f;
    p = p * r2 + -1.0f / 720.0f;
    p = p * r2 + 1.0f / 24.0f;
    return (1.0f - 0.5f * r2) + (r2 * r2) * p;
}
RB_HD float rb_reduce(float x, int &k) {
    const float kf = rintf(x * 0.636619772f);
    k = (int)kf;
    return ((x - kf * 1.5703125f) - kf * 4.837512969970703e-4f) - kf * 7.549790126404332e-8f;
}
RB_HD float rb_sin(float x) {
    int k;
    const float r = rb_reduce(x, k);
    switch (k & 3) {
        case 0: return rb_sin_poly(r);
        case 1: return rb_cos_poly(r);
        case 2: return -rb_sin_poly(r);
        default: return -rb_cos_poly(r);
    }
}
RB_HD float rb_cos(float x) {
    int k;
    const float r = rb_reduce(x, k);
    switch (k & 3) {
        case 0: return rb_cos_poly(r);
        case 1: return -rb_sin_poly(r);
        case 2: return -rb_cos_poly(r);
        default: return rb_sin_poly(r);
    }
}
RB_HD float rb_atan_core(float u) {
    const float u2 = u * u;
    float p = 1.0f / 19.0f;
    p = p * u2 + -1.0f / 17.0f;
    p = p * u2 + 1.0f / 15.0f;
    p = p * u2 + -1.0f / 13.0f;
    p = p * u2 + 1.0f / 11.0f;
    p = p * u2 + -1.0f / 9.0f;
    p = p * u2 + 1.0f / 7.0f;
    p = p * u2 + -1.0f / 5.0f;
    p = p * u2 + 1.0f / 3.0f;
    return u - (u * u2) * p;
}
RB_HD float rb_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    if (mx == 0.0f) return 0.0f;
    const float t = mn / mx;
    float a = t > 0.414213562f ? 0.785398163f + rb_atan_core((t - 1.0f) / (t + 1.0f)) : rb_atan_core(t);
    if (ay > ax) a = 1.57079633f - a;
    if (x < 0.0f) a = 3.14159265f - a;
    return y < 0.0f ? -a : a;
}
// quat.hpp:52-58
RB_HD Q from_axis_angle(V3 axis, float angle) {
    const float half = angle * 0.5f;
    const float s = rb_sin(half);
    const V3 n = normalize(axis);
    return Q{rb_cos(half), n.x * s, n.y * s, n.z * s};
}

}  // namespace rb
