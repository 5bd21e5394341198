// Oracle: sine, cosine and atan2 of the float instantiation.
//
// TEST INFRASTRUCTURE ONLY (see orc_math.hpp).
//
// The reference evaluates joint angles with std::sin / std::cos / std::atan2 in
// FP64 (include/math3d/core.hpp:102-108, src/rbps/constraints/Joint.cpp:124-152);
// the double instantiation calls exactly those.  For FP32 the GPU evaluates its
// own functions built only from IEEE +, -, *, / and comparisons, so that this
// CPU restatement reproduces them bit for bit; they are restated here:
//   * argument reduction x = k*(pi/2) + r, |r| <= ~pi/4, with pi/2 split in three
//     parts (Cody & Waite, "Software Manual for the Elementary Functions", 1980):
//     1.5703125 + 4.837512969970703e-4 + 7.549790126404332e-8 (11, 11, 24 bits;
//     k*P1 and k*P2 are exact for |k| < 2^13);
//   * Taylor polynomials of sin and cos on |r| <= pi/4 through r^11 and r^12
//     (truncation below 2e-9);
//   * atan on [0, 1]: t <= tan(pi/8) uses the Taylor series through t^19,
//     larger t uses atan(t) = pi/4 + atan((t - 1)/(t + 1)); octant fix-up by
//     |y| > |x|, sign of x and sign of y.
// Accuracy (checked against std:: in the tests): a few ulp.
#pragma once
#include <cmath>

namespace orc {

inline float rb_sin_poly(float r) {
    const float r2 = r * r;
    float p = -1.0f / 39916800.0f;
    p = p * r2 + 1.0f / 362880.0f;
    p = p * r2 + -1.0f / 5040.0f;
    p = p * r2 + 1.0f / 120.0f;
    p = p * r2 + -1.0f / 6.0f;
    return r + (r * r2) * p;
}
inline float rb_cos_poly(float r) {
    const float r2 = r * r;
    float p = 1.0f / 479001600.0f;
    p = p * r2 + -1.0f / 3628800.0f;
    p = p * r2 + 1.0f / 40320.0f;
    p = p * r2 + -1.0f / 720.0f;
    p = p * r2 + 1.0f / 24.0f;
    return (1.0f - 0.5f * r2) + (r2 * r2) * p;
}
// Quadrant k (as an integer) and the reduced argument r.
inline float rb_reduce(float x, int &k) {
    const float kf = std::nearbyint(x * 0.636619772f);
    k = (int)kf;
    return ((x - kf * 1.5703125f) - kf * 4.837512969970703e-4f) - kf * 7.549790126404332e-8f;
}
inline float rb_sin(float x) {
    int k;
    const float r = rb_reduce(x, k);
    switch (k & 3) {
        case 0: return rb_sin_poly(r);
        case 1: return rb_cos_poly(r);
        case 2: return -rb_sin_poly(r);
        default: return -rb_cos_poly(r);
    }
}
inline float rb_cos(float x) {
    int k;
    const float r = rb_reduce(x, k);
    switch (k & 3) {
        case 0: return rb_cos_poly(r);
        case 1: return -rb_sin_poly(r);
        case 2: return -rb_cos_poly(r);
        default: return rb_sin_poly(r);
    }
}
inline float rb_atan_core(float u) {   // |u| <= tan(pi/8)
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
inline float rb_atan2(float y, float x) {
    const float ax = std::fabs(x), ay = std::fabs(y);
    const float mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    if (mx == 0.0f) return 0.0f;
    const float t = mn / mx;
    float a = t > 0.414213562f ? 0.785398163f + rb_atan_core((t - 1.0f) / (t + 1.0f)) : rb_atan_core(t);
    if (ay > ax) a = 1.57079633f - a;
    if (x < 0.0f) a = 3.14159265f - a;
    return y < 0.0f ? -a : a;
}

template <class T> inline T t_sin(T x) { return std::sin(x); }
template <class T> inline T t_cos(T x) { return std::cos(x); }
template <class T> inline T t_atan2(T y, T x) { return std::atan2(y, x); }
template <> inline float t_sin<float>(float x) { return rb_sin(x); }
template <> inline float t_cos<float>(float x) { return rb_cos(x); }
template <> inline float t_atan2<float>(float y, float x) { return rb_atan2(y, x); }

}  // namespace orc
