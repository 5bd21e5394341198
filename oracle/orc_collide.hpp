// Oracle: shapes, AABBs and the analytic narrowphase (sphere / box / capsule / plane).
//
// TEST INFRASTRUCTURE ONLY (see orc_math.hpp).  Each routine cites the reference
// header it restates; operand order follows the reference so that the double
// instantiation is bit-exact against it.
#pragma once
#include "orc_math.hpp"

namespace orc {

// Shape kinds use the reference's Shape::v variant indices (ShapeTypes.hpp:69).
enum ShapeKind { SH_SPHERE = 0, SH_BOX = 1, SH_CAPSULE = 3, SH_PLANE = 6, SH_HULL = 9 };

template <class T> struct Hull;   // orc_gjk.hpp

template <class T> struct ShapeP {
    int kind = SH_SPHERE;
    // sphere: a=radius | box: half extents (a,b,c) | capsule: a=half_height b=radius
    // plane: normal (a,b,c), d | hull: the hull's tables
    T a = 0, b = 0, c = 0, d = 0;
    const Hull<T> *hull = nullptr;
};

template <class T> inline AABB<T> hull_aabb(const Hull<T> &h, const TF<T> &tf);

template <class T> struct Manifold {
    V3<T> normal;
    V3<T> pos[4];
    T depth[4];
    int n = 0;
};

// ---------------------------------------------------------------- AABBs
// Sphere.hpp:83-87, Box.hpp:101-112, Capsule.hpp:129-139
template <class T> inline AABB<T> shape_aabb(const ShapeP<T> &s, const TF<T> &tf) {
    AABB<T> r;
    if (s.kind == SH_SPHERE) {
        V3<T> e(s.a, s.a, s.a);
        r.mn = tf.pos - e;
        r.mx = tf.pos + e;
    } else if (s.kind == SH_BOX) {
        M3<T> R = mat3_cast(tf.rot);
        V3<T> ext(std::abs(R.c[0].x) * s.a + std::abs(R.c[1].x) * s.b + std::abs(R.c[2].x) * s.c,
                  std::abs(R.c[0].y) * s.a + std::abs(R.c[1].y) * s.b + std::abs(R.c[2].y) * s.c,
                  std::abs(R.c[0].z) * s.a + std::abs(R.c[1].z) * s.b + std::abs(R.c[2].z) * s.c);
        r.mn = tf.pos - ext;
        r.mx = tf.pos + ext;
    } else if (s.kind == SH_CAPSULE) {
        V3<T> ax = tf.vec(V3<T>(T(0), s.a, T(0)));
        V3<T> p1 = tf.pos + ax, p2 = tf.pos - ax;
        V3<T> rr(s.b, s.b, s.b);
        r.mn = V3<T>(smin(p1.x, p2.x), smin(p1.y, p2.y), smin(p1.z, p2.z)) - rr;
        r.mx = V3<T>(smax(p1.x, p2.x), smax(p1.y, p2.y), smax(p1.z, p2.z)) + rr;
    } else if (s.kind == SH_HULL) {
        r = hull_aabb(*s.hull, tf);
    } else {
        T inf = std::numeric_limits<T>::infinity();
        r.mn = V3<T>(-inf);
        r.mx = V3<T>(inf);
    }
    return r;
}

template <class T> inline void capsule_ends(const ShapeP<T> &c, const TF<T> &tf, V3<T> &p1, V3<T> &p2) {
    V3<T> ax = tf.vec(V3<T>(T(0), c.a, T(0)));   // Capsule.hpp:43-49
    p1 = tf.pos + ax;
    p2 = tf.pos - ax;
}

// ---------------------------------------------------------------- pair tests
// SphereSphere.hpp:24-51
template <class T>
inline bool sphere_sphere(const ShapeP<T> &a, const TF<T> &ta, const ShapeP<T> &b, const TF<T> &tb, Manifold<T> &m) {
    V3<T> delta = tb.pos - ta.pos;
    T dist = length(delta);
    T rsum = a.a + b.a;
    if (dist >= rsum) return false;
    m.normal = (dist > C<T>::EPS) ? delta / dist : V3<T>(T(1), T(0), T(0));
    m.n = 1;
    m.depth[0] = rsum - dist;
    m.pos[0] = ta.pos + m.normal * a.a;
    return true;
}

// SphereBox.hpp:31-108
template <class T>
inline bool sphere_box(const ShapeP<T> &sp, const TF<T> &ts, const ShapeP<T> &bx, const TF<T> &tb, Manifold<T> &m) {
    V3<T> lc = tb.inv_vec(ts.pos - tb.pos);
    V3<T> cl(sclamp(lc.x, -bx.a, bx.a), sclamp(lc.y, -bx.b, bx.b), sclamp(lc.z, -bx.c, bx.c));
    V3<T> ld = lc - cl;
    T dist = length(ld);
    if (dist > C<T>::EPS) {
        if (dist >= sp.a) return false;
        m.n = 1;
        m.normal = -tb.vec(ld / dist);
        m.depth[0] = sp.a - dist;
        m.pos[0] = tb.point(cl);
        return true;
    }
    T px = bx.a - std::abs(lc.x);
    T py = bx.b - std::abs(lc.y);
    T pz = bx.c - std::abs(lc.z);
    V3<T> ln;
    if (px <= py && px <= pz)
        ln = V3<T>(lc.x > 0 ? T(1.0) : T(-1.0), T(0), T(0));
    else if (py <= px && py <= pz)
        ln = V3<T>(T(0), lc.y > 0 ? T(1.0) : T(-1.0), T(0));
    else
        ln = V3<T>(T(0), T(0), lc.z > 0 ? T(1.0) : T(-1.0));
    m.n = 1;
    m.normal = -tb.vec(ln);
    m.depth[0] = sp.a + smin(px, smin(py, pz));
    V3<T> fp(ln.x != T(0) ? ln.x * bx.a : sclamp(lc.x, -bx.a, bx.a),
             ln.y != T(0) ? ln.y * bx.b : sclamp(lc.y, -bx.b, bx.b),
             ln.z != T(0) ? ln.z * bx.c : sclamp(lc.z, -bx.c, bx.c));
    m.pos[0] = tb.point(fp);
    return true;
}

// SphereCapsule.hpp:24-67
template <class T>
inline bool sphere_capsule(const ShapeP<T> &sp, const TF<T> &ts, const ShapeP<T> &cp, const TF<T> &tc, Manifold<T> &m) {
    V3<T> p1, p2;
    capsule_ends(cp, tc, p1, p2);
    V3<T> d = p2 - p1;
    T len2 = length_sq(d);
    T t = T(0);
    if (len2 > C<T>::EPS) t = sclamp(dot(ts.pos - p1, d) / len2, T(0), T(1));
    V3<T> closest = p1 + d * t;
    V3<T> delta = ts.pos - closest;
    T dist = length(delta);
    T rsum = sp.a + cp.b;
    if (dist >= rsum) return false;
    m.n = 1;
    m.normal = (dist > C<T>::EPS) ? -delta / dist : V3<T>(T(1), T(0), T(0));
    m.depth[0] = rsum - dist;
    m.pos[0] = closest - m.normal * cp.b;
    return true;
}

// CapsuleCapsule.hpp:24-100
template <class T>
inline bool capsule_capsule(const ShapeP<T> &a, const TF<T> &ta, const ShapeP<T> &b, const TF<T> &tb, Manifold<T> &m) {
    V3<T> p1, p2, q1, q2;
    capsule_ends(a, ta, p1, p2);
    capsule_ends(b, tb, q1, q2);
    V3<T> d1 = p2 - p1, d2 = q2 - q1, r = p1 - q1;
    T A = length_sq(d1), E = length_sq(d2), F = dot(d2, r);
    T s = T(0), t = T(0);
    T B = dot(d1, d2);
    T denom = A * E - B * B;
    bool par = denom < (C<T>::EPS * C<T>::EPS);
    if (par) {
        if (A > C<T>::EPS) s = sclamp(-dot(d1, r) / A, T(0), T(1));
        if (E > C<T>::EPS) t = sclamp(dot(d2, -r) / E, T(0), T(1));
    } else {
        T Cc = dot(d1, r);
        s = sclamp((B * F - Cc * E) / denom, T(0), T(1));
        t = (B * s + F) / E;
        if (t < T(0)) { t = T(0); s = sclamp(-Cc / A, T(0), T(1)); }
        else if (t > T(1)) { t = T(1); s = sclamp((B - Cc) / A, T(0), T(1)); }
    }
    V3<T> ca = p1 + d1 * s, cb = q1 + d2 * t;
    V3<T> delta = ca - cb;
    T dist = length(delta);
    T rsum = a.b + b.b;
    if (dist >= rsum) return false;
    V3<T> n = (dist > C<T>::EPS) ? delta / dist : V3<T>(T(1), T(0), T(0));
    m.normal = n;
    m.depth[0] = rsum - dist;
    m.pos[0] = cb + n * b.b;
    if (!par) {
        m.n = 1;
    } else {
        m.n = 2;   // second point is the reference's placeholder q1 + n*r_b (CapsuleCapsule.hpp:97-99)
        m.pos[1] = q1 + n * b.b;
        m.depth[1] = rsum - dist;
    }
    return true;
}

// BoxBox.hpp:29-37
template <class T> inline T obb_radius(const ShapeP<T> &bx, const TF<T> &tf, const V3<T> &axis) {
    V3<T> ax = tf.vec(V3<T>(1, 0, 0)), ay = tf.vec(V3<T>(0, 1, 0)), az = tf.vec(V3<T>(0, 0, 1));
    return std::abs(dot(ax, axis)) * bx.a + std::abs(dot(ay, axis)) * bx.b + std::abs(dot(az, axis)) * bx.c;
}
// BoxBox.hpp:39-69
template <class T>
inline bool bb_test_axis(const V3<T> &axis, const ShapeP<T> &a, const TF<T> &ta, const ShapeP<T> &b, const TF<T> &tb,
                         const V3<T> &cd, T &best_depth, V3<T> &best_n, int feat, int &best_feat, T bias) {
    T l2 = length_sq(axis);
    if (l2 < C<T>::EPS * C<T>::EPS) return true;
    V3<T> na = axis / std::sqrt(l2);
    T rA = obb_radius(a, ta, na), rB = obb_radius(b, tb, na);
    T dist = std::abs(dot(cd, na));
    T ov = rA + rB - dist;
    if (ov < T(0)) return false;
    if ((ov + bias) < best_depth) {
        best_depth = ov + bias;
        best_n = (dot(cd, na) >= T(0)) ? na : -na;
        best_feat = feat;
    }
    return true;
}
// BoxBox.hpp:71-103
template <class T>
inline void bb_best_face(const ShapeP<T> &bx, const TF<T> &tf, const V3<T> &dir, V3<T> out[4], V3<T> &out_n) {
    V3<T> axes[3] = {tf.vec(V3<T>(1, 0, 0)), tf.vec(V3<T>(0, 1, 0)), tf.vec(V3<T>(0, 0, 1))};
    T half[3] = {bx.a, bx.b, bx.c};
    int best = 0;
    T bd = std::abs(dot(axes[0], dir));
    for (int i = 1; i < 3; ++i) {
        T dd = std::abs(dot(axes[i], dir));
        if (dd > bd) { bd = dd; best = i; }
    }
    T sign = (dot(axes[best], dir) >= T(0)) ? T(1.0) : T(-1.0);
    out_n = axes[best] * sign;
    V3<T> fc = tf.pos + out_n * half[best];
    int u = (best + 1) % 3, v = (best + 2) % 3;
    out[0] = fc + axes[u] * half[u] + axes[v] * half[v];
    out[1] = fc - axes[u] * half[u] + axes[v] * half[v];
    out[2] = fc - axes[u] * half[u] - axes[v] * half[v];
    out[3] = fc + axes[u] * half[u] - axes[v] * half[v];
}
// BoxBox.hpp:105-127
template <class T>
inline int bb_clip(const V3<T> *poly, int n, const V3<T> &pn, const V3<T> &pp, V3<T> *out) {
    int k = 0;
    for (int i = 0; i < n; ++i) {
        const V3<T> &A = poly[i];
        const V3<T> &B = poly[(i + 1) % n];
        T dA = dot(A - pp, pn), dB = dot(B - pp, pn);
        if (dA >= T(0)) out[k++] = A;
        if ((dA > T(0)) != (dB > T(0))) {
            T t = dA / (dA - dB);
            out[k++] = A + (B - A) * t;
        }
    }
    return k;
}
// BoxBox.hpp:130-231 (coplanar area maximisation)
template <class T>
inline int reduce4(const V3<T> *p, const T *dep, int n, V3<T> *op, T *od) {
    if (n <= 4) {
        for (int i = 0; i < n; ++i) { op[i] = p[i]; od[i] = dep[i]; }
        return n;
    }
    int i0 = 0;
    for (int i = 1; i < n; ++i)
        if (dep[i] > dep[i0]) i0 = i;
    int i1 = (i0 == 0) ? 1 : 0;
    T best = T(-1.0);
    for (int i = 0; i < n; ++i) {
        if (i == i0) continue;
        T dd = length_sq(p[i] - p[i0]);
        if (dd > best) { best = dd; i1 = i; }
    }
    int i2 = -1;
    best = T(-1.0);
    V3<T> seg = p[i1] - p[i0];
    for (int i = 0; i < n; ++i) {
        if (i == i0 || i == i1) continue;
        T dd = length_sq(cross(p[i] - p[i0], seg));
        if (dd > best) { best = dd; i2 = i; }
    }
    if (i2 < 0) { i2 = 0; while (i2 == i0 || i2 == i1) ++i2; }
    int i3 = -1;
    T mind = T(1.0);
    V3<T> tn = cross(p[i2] - p[i0], seg);
    for (int i = 0; i < n; ++i) {
        if (i == i0 || i == i1 || i == i2) continue;
        T dd = dot(cross(p[i] - p[i0], seg), tn);
        if (dd < mind) { mind = dd; i3 = i; }
    }
    if (i3 < 0 || mind >= T(0)) {
        best = T(-1.0);
        for (int i = 0; i < n; ++i) {
            if (i == i0 || i == i1 || i == i2) continue;
            T dd = length_sq(p[i] - p[i2]);
            if (dd > best) { best = dd; i3 = i; }
        }
    }
    if (i3 < 0) { i3 = 0; while (i3 == i0 || i3 == i1 || i3 == i2) ++i3; }
    op[0] = p[i0]; od[0] = dep[i0];
    op[1] = p[i1]; od[1] = dep[i1];
    op[2] = p[i2]; od[2] = dep[i2];
    op[3] = p[i3]; od[3] = dep[i3];
    return 4;
}
// BoxBox.hpp:239-399
template <class T>
inline bool box_box(const ShapeP<T> &a, const TF<T> &ta, const ShapeP<T> &b, const TF<T> &tb, Manifold<T> &m) {
    V3<T> cd = tb.pos - ta.pos;
    V3<T> ax[3] = {ta.vec(V3<T>(1, 0, 0)), ta.vec(V3<T>(0, 1, 0)), ta.vec(V3<T>(0, 0, 1))};
    V3<T> bx[3] = {tb.vec(V3<T>(1, 0, 0)), tb.vec(V3<T>(0, 1, 0)), tb.vec(V3<T>(0, 0, 1))};
    T best_depth = std::numeric_limits<T>::max();
    V3<T> best_n(T(1), T(0), T(0));
    int best_feat = 0, feat = 0;
    for (int i = 0; i < 3; ++i, ++feat)
        if (!bb_test_axis(ax[i], a, ta, b, tb, cd, best_depth, best_n, feat, best_feat, T(0))) return false;
    for (int i = 0; i < 3; ++i, ++feat)
        if (!bb_test_axis(bx[i], a, ta, b, tb, cd, best_depth, best_n, feat, best_feat, T(0))) return false;
    const T edge_bias = T(0.001f);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j, ++feat)
            if (!bb_test_axis(cross(ax[i], bx[j]), a, ta, b, tb, cd, best_depth, best_n, feat, best_feat, edge_bias))
                return false;
    if (best_feat >= 6) best_depth -= edge_bias;
    m.normal = best_n;
    m.n = 0;
    if (best_feat < 6) {
        bool refA = best_feat < 3;
        const ShapeP<T> &rb = refA ? a : b;
        const TF<T> &rt = refA ? ta : tb;
        const ShapeP<T> &ib = refA ? b : a;
        const TF<T> &it = refA ? tb : ta;
        V3<T> target = refA ? best_n : -best_n;
        V3<T> rc[4], rn, ic[4], inn;
        bb_best_face(rb, rt, target, rc, rn);
        bb_best_face(ib, it, -target, ic, inn);
        V3<T> b0[16], b1[16];
        int cnt = 4;
        for (int i = 0; i < 4; ++i) b0[i] = ic[i];
        V3<T> rcen = (rc[0] + rc[1] + rc[2] + rc[3]) * T(0.25f);
        for (int i = 0; i < 4; ++i) {
            V3<T> edge = rc[(i + 1) % 4] - rc[i];
            V3<T> sn = normalize(cross(rn, edge));
            if (dot(sn, rcen - rc[i]) < T(0.0f)) sn = -sn;
            cnt = bb_clip(b0, cnt, sn, rc[i], b1);
            if (cnt == 0) break;
            for (int k = 0; k < cnt; ++k) b0[k] = b1[k];
        }
        if (cnt == 0) return false;
        T rd = dot(rc[0], rn);
        V3<T> kp[16];
        T kd[16];
        int kn = 0;
        for (int i = 0; i < cnt; ++i) {
            T sd = dot(b0[i], rn) - rd;
            if (sd <= C<T>::EPS) {
                kp[kn] = b0[i] - rn * (sd * T(0.5f));
                kd[kn] = -sd;
                ++kn;
            }
        }
        if (kn == 0) return false;
        m.n = reduce4(kp, kd, kn, m.pos, m.depth);
    } else {
        int ei = (best_feat - 6) / 3, ej = (best_feat - 6) % 3;
        T eA[3] = {a.a, a.b, a.c}, eB[3] = {b.a, b.b, b.c};
        V3<T> pa = ta.pos + ta.vec(V3<T>((dot(ax[0], best_n) >= 0 ? T(1.0) : T(-1.0)) * (ei == 0 ? T(0.0) : eA[0]),
                                         (dot(ax[1], best_n) >= 0 ? T(1.0) : T(-1.0)) * (ei == 1 ? T(0.0) : eA[1]),
                                         (dot(ax[2], best_n) >= 0 ? T(1.0) : T(-1.0)) * (ei == 2 ? T(0.0) : eA[2])));
        V3<T> nn = -best_n;
        V3<T> pb = tb.pos + tb.vec(V3<T>((dot(bx[0], nn) >= 0 ? T(1.0) : T(-1.0)) * (ej == 0 ? T(0.0) : eB[0]),
                                         (dot(bx[1], nn) >= 0 ? T(1.0) : T(-1.0)) * (ej == 1 ? T(0.0) : eB[1]),
                                         (dot(bx[2], nn) >= 0 ? T(1.0) : T(-1.0)) * (ej == 2 ? T(0.0) : eB[2])));
        V3<T> d1 = ax[ei], d2 = bx[ej];
        V3<T> r = pa - pb;
        T bd = dot(d1, d2), cdt = dot(d1, r), fd = dot(d2, r);
        T den = T(1.0f) - bd * bd;
        T s = T(0.0f), t = T(0.0f);
        if (den > C<T>::EPS) s = (bd * fd - cdt) / den;
        T ea = eA[ei], eb = eB[ej];
        if (s < -ea || s > ea) s = sclamp(s, -ea, ea);
        t = s * bd + fd;
        if (t < -eb || t > eb) {
            t = sclamp(t, -eb, eb);
            s = sclamp(t * bd - cdt, -ea, ea);
        }
        m.n = 1;
        m.pos[0] = (pa + d1 * s + pb + d2 * t) * T(0.5f);
        m.depth[0] = best_depth;
    }
    return true;
}

// CapsuleBox.hpp:28-221
template <class T>
inline bool capsule_box(const ShapeP<T> &cp, const TF<T> &tc, const ShapeP<T> &bx, const TF<T> &tb, Manifold<T> &m) {
    V3<T> p1, p2;
    capsule_ends(cp, tc, p1, p2);
    V3<T> d = p2 - p1;
    V3<T> l = d * T(0.5);
    V3<T> c = p1 + l;
    V3<T> ax[3] = {tb.vec(V3<T>(1, 0, 0)), tb.vec(V3<T>(0, 1, 0)), tb.vec(V3<T>(0, 0, 1))};
    T e[3] = {bx.a, bx.b, bx.c};
    const T rad = cp.b;
    V3<T> diff = c - tb.pos;
    T best_depth = std::numeric_limits<T>::max();
    V3<T> best_n(T(1), T(0), T(0));
    int best_feat = -1;
    for (int i = 0; i < 3; ++i) {
        const V3<T> &axis = ax[i];
        T rbox = e[i];
        T rcap = std::abs(dot(l, axis)) + rad;
        T dist = std::abs(dot(diff, axis));
        T ov = rbox + rcap - dist;
        if (ov < T(0)) return false;
        if (ov < best_depth) {
            best_depth = ov;
            best_n = (dot(diff, axis) >= T(0)) ? axis : -axis;
            best_feat = i;
        }
    }
    for (int i = 0; i < 3; ++i) {
        V3<T> axis = cross(ax[i], l);
        T l2 = length_sq(axis);
        if (l2 < C<T>::EPS * C<T>::EPS) continue;
        axis = axis / std::sqrt(l2);
        T rbox = e[0] * std::abs(dot(ax[0], axis)) + e[1] * std::abs(dot(ax[1], axis)) + e[2] * std::abs(dot(ax[2], axis));
        T rcap = rad;
        T dist = std::abs(dot(diff, axis));
        T ov = rbox + rcap - dist;
        if (ov < T(0)) return false;
        if (ov < best_depth) {
            best_depth = ov;
            best_n = (dot(diff, axis) >= T(0)) ? axis : -axis;
            best_feat = i + 3;
        }
    }
    m.normal = -best_n;
    m.n = 0;
    if (best_feat < 3) {
        T t0 = T(0), t1 = T(1);
        auto clip = [](T p, T q, T &tmin, T &tmax) -> bool {
            if (std::abs(p) < C<T>::EPS) return q >= T(0);
            T t = q / p;
            if (p > T(0)) { if (t < tmax) tmax = t; }
            else { if (t > tmin) tmin = t; }
            return tmin <= tmax;
        };
        int iA = (best_feat + 1) % 3, iB = (best_feat + 2) % 3;
        V3<T> p1r = p1 - tb.pos;
        T pA = dot(d, ax[iA]), vA = dot(p1r, ax[iA]);
        if (!clip(pA, e[iA] - vA, t0, t1)) return true;
        if (!clip(-pA, e[iA] + vA, t0, t1)) return true;
        T pB = dot(d, ax[iB]), vB = dot(p1r, ax[iB]);
        if (!clip(pB, e[iB] - vB, t0, t1)) return true;
        if (!clip(-pB, e[iB] + vB, t0, t1)) return true;
        V3<T> pt0 = p1 + d * t0;
        T d0 = rad + e[best_feat] - dot(pt0 - tb.pos, best_n);
        if (d0 > T(0)) { m.pos[m.n] = pt0 - best_n * rad; m.depth[m.n] = d0; m.n++; }
        if (t1 - t0 > C<T>::EPS) {
            V3<T> pt1 = p1 + d * t1;
            T d1 = rad + e[best_feat] - dot(pt1 - tb.pos, best_n);
            if (d1 > T(0)) { m.pos[m.n] = pt1 - best_n * rad; m.depth[m.n] = d1; m.n++; }
        }
    } else {
        int ea = best_feat - 3;
        V3<T> v = ax[ea];
        V3<T> sb = tb.pos + ax[0] * ((dot(ax[0], best_n) >= T(0)) ? e[0] : -e[0]) +
                   ax[1] * ((dot(ax[1], best_n) >= T(0)) ? e[1] : -e[1]) +
                   ax[2] * ((dot(ax[2], best_n) >= T(0)) ? e[2] : -e[2]);
        V3<T> r = p1 - sb;
        T a = length_sq(d), b = dot(d, v), f = dot(v, r), cdot = dot(d, r);
        T den = a - b * b;
        T t = (den > C<T>::EPS) ? sclamp((b * f - cdot) / den, T(0), T(1)) : T(0.5);
        V3<T> pt = p1 + d * t;
        m.n = 1;
        m.pos[0] = pt - best_n * rad;
        m.depth[0] = best_depth;
    }
    if (m.n == 0) {
        m.n = 1;
        m.pos[0] = c - best_n * rad;
        m.depth[0] = best_depth;
    }
    return true;
}

// Plane world frame (Plane.hpp:47-57)
template <class T> inline void plane_world(const ShapeP<T> &pl, const TF<T> &tp, V3<T> &wn, T &wd) {
    wn = tp.vec(V3<T>(pl.a, pl.b, pl.c));
    wd = pl.d + dot(wn, tp.pos);
}

// convex_vs_plane for a sphere (PlaneCollision.hpp:40-61); Sphere support Sphere.hpp:52-55
template <class T>
inline bool sphere_plane(const ShapeP<T> &sp, const TF<T> &ts, const ShapeP<T> &pl, const TF<T> &tp, Manifold<T> &m) {
    V3<T> wn;
    T wd;
    plane_world(pl, tp, wn, wd);
    V3<T> ldir = ts.inv_vec(-wn);
    V3<T> lsup = normalize(ldir) * sp.a;
    V3<T> wsup = ts.pos + ts.vec(lsup);
    T sd = dot(wsup, wn) - wd;
    if (sd >= T(0)) return false;
    m.normal = -wn;
    m.n = 1;
    m.depth[0] = -sd;
    m.pos[0] = wsup;
    return true;
}

// PlaneCollision.hpp:293-337
template <class T>
inline bool box_plane(const ShapeP<T> &bx, const TF<T> &tb, const ShapeP<T> &pl, const TF<T> &tp, Manifold<T> &m) {
    V3<T> wn;
    T wd;
    plane_world(pl, tp, wn, wd);
    V3<T> u[3] = {tb.vec(V3<T>(1, 0, 0)), tb.vec(V3<T>(0, 1, 0)), tb.vec(V3<T>(0, 0, 1))};
    const T eps = T(1e-4f);
    T r = bx.a * std::abs(dot(u[0], wn)) + bx.b * std::abs(dot(u[1], wn)) + bx.c * std::abs(dot(u[2], wn));
    T cdist = dot(tb.pos, wn) - wd;
    if (cdist - r > eps) return false;
    m.normal = -wn;
    m.n = 0;
    for (int sx = -1; sx <= 1; sx += 2)
        for (int sy = -1; sy <= 1; sy += 2)
            for (int sz = -1; sz <= 1; sz += 2) {
                V3<T> corner = tb.pos + T(sx) * bx.a * u[0] + T(sy) * bx.b * u[1] + T(sz) * bx.c * u[2];
                T dist = dot(corner, wn) - wd;
                if (dist <= eps && m.n < 4) {
                    m.pos[m.n] = corner;
                    m.depth[m.n] = -dist;
                    m.n++;
                }
            }
    return m.n > 0;
}

// PlaneCollision.hpp:344-389
template <class T>
inline bool capsule_plane(const ShapeP<T> &cp, const TF<T> &tc, const ShapeP<T> &pl, const TF<T> &tp, Manifold<T> &m) {
    V3<T> wn;
    T wd;
    plane_world(pl, tp, wn, wd);
    V3<T> p1, p2;
    capsule_ends(cp, tc, p1, p2);
    T d1 = dot(p1, wn) - wd, d2 = dot(p2, wn) - wd;
    m.n = 0;
    m.normal = -wn;
    const T eps = T(1e-4);
    V3<T> roff = wn * cp.b;
    T thr = cp.b + eps;
    if (d1 <= thr) { m.pos[m.n] = p1 - roff; m.depth[m.n] = cp.b - d1; m.n++; }
    if (d2 <= thr) { m.pos[m.n] = p2 - roff; m.depth[m.n] = cp.b - d2; m.n++; }
    return m.n > 0;
}

// Dispatcher.hpp:104-116 and 212-217: analytic table; (B,A) = (A,B) with the normal flipped.
}  // namespace orc
#include "orc_gjk.hpp"
namespace orc {

template <class T> struct EpaGuard;   // orc_gjk.hpp
// guard: the EPA guard of GJK/EPA pairs (gjk_epa_manifold), else unused.
template <class T>
inline bool dispatch(const ShapeP<T> &a, const TF<T> &ta, const ShapeP<T> &b, const TF<T> &tb, Manifold<T> &m,
                     const EpaGuard<T> *guard = nullptr, T *guard_cut = nullptr) {
    const int ka = a.kind, kb = b.kind;
    auto sym = [&](bool hit) { if (hit) m.normal = -m.normal; return hit; };
    if (ka == SH_SPHERE && kb == SH_SPHERE) return sphere_sphere(a, ta, b, tb, m);
    if (ka == SH_SPHERE && kb == SH_BOX) return sphere_box(a, ta, b, tb, m);
    if (ka == SH_BOX && kb == SH_SPHERE) return sym(sphere_box(b, tb, a, ta, m));
    if (ka == SH_BOX && kb == SH_BOX) return box_box(a, ta, b, tb, m);
    if (ka == SH_SPHERE && kb == SH_CAPSULE) return sphere_capsule(a, ta, b, tb, m);
    if (ka == SH_CAPSULE && kb == SH_SPHERE) return sym(sphere_capsule(b, tb, a, ta, m));
    if (ka == SH_CAPSULE && kb == SH_CAPSULE) return capsule_capsule(a, ta, b, tb, m);
    if (ka == SH_CAPSULE && kb == SH_BOX) return capsule_box(a, ta, b, tb, m);
    if (ka == SH_BOX && kb == SH_CAPSULE) return sym(capsule_box(b, tb, a, ta, m));
    // hulls: hull-plane analytic (PlaneCollision.hpp:90-289); every other pair with a hull
    // takes the GJK/EPA primary template, in the given order (Dispatcher.hpp:69-91)
    if (ka == SH_HULL && kb == SH_PLANE) return hull_plane(a, ta, b, tb, m);
    if (ka == SH_PLANE && kb == SH_HULL) return sym(hull_plane(b, tb, a, ta, m));
    if ((ka == SH_HULL || kb == SH_HULL) && ka != SH_PLANE && kb != SH_PLANE)
        return gjk_epa_manifold<T>(a, ta, b, tb, m, nullptr, nullptr, nullptr, nullptr, nullptr, guard, guard_cut);
    if (kb == SH_PLANE) {
        if (ka == SH_SPHERE) return sphere_plane(a, ta, b, tb, m);
        if (ka == SH_BOX) return box_plane(a, ta, b, tb, m);
        if (ka == SH_CAPSULE) return capsule_plane(a, ta, b, tb, m);
        return false;
    }
    if (ka == SH_PLANE) {
        if (kb == SH_SPHERE) return sym(sphere_plane(b, tb, a, ta, m));
        if (kb == SH_BOX) return sym(box_plane(b, tb, a, ta, m));
        if (kb == SH_CAPSULE) return sym(capsule_plane(b, tb, a, ta, m));
        return false;
    }
    return false;
}

}  // namespace orc
