// FP32 shape AABBs and analytic pair tests (sphere / box / capsule / plane),
// usable from host and device code.
//
// Each function cites the reference routine whose contact convention it keeps
// (normal from A to B, up to four points, positive depth when overlapping).
#pragma once
#include "rbgpu_math.hpp"

namespace rb {

enum ShapeKind : int { SH_SPHERE = 0, SH_BOX = 1, SH_CAPSULE = 3, SH_PLANE = 6, SH_HULL = 9 };

struct Shape {
    int kind;
    float a, b, c, d;   // sphere r | box half extents | capsule half_height, radius | plane n, d
};

struct Manifold {
    V3 n;
    V3 p[4];
    float dep[4];
    int np;
};

// Sphere.hpp:83-87, Box.hpp:101-112, Capsule.hpp:129-139
RB_HD Box3 shape_aabb(const Shape &s, const TF &tf) {
    Box3 r;
    if (s.kind == SH_SPHERE) {
        V3 e{s.a, s.a, s.a};
        r.mn = tf.pos - e;
        r.mx = tf.pos + e;
    } else if (s.kind == SH_BOX) {
        M3 R = mat3_cast(tf.rot);
        V3 ext{absf(R.c0.x) * s.a + absf(R.c1.x) * s.b + absf(R.c2.x) * s.c,
               absf(R.c0.y) * s.a + absf(R.c1.y) * s.b + absf(R.c2.y) * s.c,
               absf(R.c0.z) * s.a + absf(R.c1.z) * s.b + absf(R.c2.z) * s.c};
        r.mn = tf.pos - ext;
        r.mx = tf.pos + ext;
    } else if (s.kind == SH_CAPSULE) {
        V3 ax = tf_vec(tf, V3{0.0f, s.a, 0.0f});
        V3 p1 = tf.pos + ax, p2 = tf.pos - ax;
        V3 rr{s.b, s.b, s.b};
        r.mn = V3{fmin_ref(p1.x, p2.x), fmin_ref(p1.y, p2.y), fmin_ref(p1.z, p2.z)} - rr;
        r.mx = V3{fmax_ref(p1.x, p2.x), fmax_ref(p1.y, p2.y), fmax_ref(p1.z, p2.z)} + rr;
    } else {
        const float inf = __builtin_inff();
        r.mn = V3{-inf, -inf, -inf};
        r.mx = V3{inf, inf, inf};
    }
    return r;
}

// ConvexHull.hpp:484-499: centre and half extent of the local AABB, projected like a box.
RB_HD Box3 hull_aabb(V3 mn, V3 mx, const TF &tf) {
    const V3 c = (mn + mx) * 0.5f;
    const V3 e = (mx - mn) * 0.5f;
    const M3 R = mat3_cast(tf.rot);
    const V3 wc = tf.pos + tf_vec(tf, c);
    const V3 ext{absf(R.c0.x) * e.x + absf(R.c1.x) * e.y + absf(R.c2.x) * e.z,
                 absf(R.c0.y) * e.x + absf(R.c1.y) * e.y + absf(R.c2.y) * e.z,
                 absf(R.c0.z) * e.x + absf(R.c1.z) * e.y + absf(R.c2.z) * e.z};
    return Box3{wc - ext, wc + ext};
}

RB_HD void capsule_ends(const Shape &c, const TF &tf, V3 &p1, V3 &p2) {   // Capsule.hpp:43-49
    V3 ax = tf_vec(tf, V3{0.0f, c.a, 0.0f});
    p1 = tf.pos + ax;
    p2 = tf.pos - ax;
}

// SphereSphere.hpp:24-51
RB_HD bool sphere_sphere(const Shape &a, const TF &ta, const Shape &b, const TF &tb, Manifold &m) {
    V3 delta = tb.pos - ta.pos;
    float dist = len(delta);
    float rsum = a.a + b.a;
    if (dist >= rsum) return false;
    m.n = (dist > EPS) ? vdiv(delta, dist) : V3{1.0f, 0.0f, 0.0f};
    m.np = 1;
    m.dep[0] = rsum - dist;
    m.p[0] = ta.pos + m.n * a.a;
    return true;
}

// SphereBox.hpp:31-108
RB_HD bool sphere_box(const Shape &sp, const TF &ts, const Shape &bx, const TF &tb, Manifold &m) {
    V3 lc = tf_inv_vec(tb, ts.pos - tb.pos);
    V3 cl{clampf(lc.x, -bx.a, bx.a), clampf(lc.y, -bx.b, bx.b), clampf(lc.z, -bx.c, bx.c)};
    V3 ld = lc - cl;
    float dist = len(ld);
    if (dist > EPS) {
        if (dist >= sp.a) return false;
        m.np = 1;
        m.n = -tf_vec(tb, vdiv(ld, dist));
        m.dep[0] = sp.a - dist;
        m.p[0] = tf_point(tb, cl);
        return true;
    }
    float px = bx.a - absf(lc.x), py = bx.b - absf(lc.y), pz = bx.c - absf(lc.z);
    V3 ln;
    if (px <= py && px <= pz) ln = V3{lc.x > 0.0f ? 1.0f : -1.0f, 0.0f, 0.0f};
    else if (py <= px && py <= pz) ln = V3{0.0f, lc.y > 0.0f ? 1.0f : -1.0f, 0.0f};
    else ln = V3{0.0f, 0.0f, lc.z > 0.0f ? 1.0f : -1.0f};
    m.np = 1;
    m.n = -tf_vec(tb, ln);
    m.dep[0] = sp.a + fmin_ref(px, fmin_ref(py, pz));
    V3 fp{ln.x != 0.0f ? ln.x * bx.a : clampf(lc.x, -bx.a, bx.a), ln.y != 0.0f ? ln.y * bx.b : clampf(lc.y, -bx.b, bx.b),
          ln.z != 0.0f ? ln.z * bx.c : clampf(lc.z, -bx.c, bx.c)};
    m.p[0] = tf_point(tb, fp);
    return true;
}

// SphereCapsule.hpp:24-67
RB_HD bool sphere_capsule(const Shape &sp, const TF &ts, const Shape &cp, const TF &tc, Manifold &m) {
    V3 p1, p2;
    capsule_ends(cp, tc, p1, p2);
    V3 d = p2 - p1;
    float l2 = len2(d);
    float t = 0.0f;
    if (l2 > EPS) t = clampf(dot(ts.pos - p1, d) / l2, 0.0f, 1.0f);
    V3 closest = p1 + d * t;
    V3 delta = ts.pos - closest;
    float dist = len(delta);
    float rsum = sp.a + cp.b;
    if (dist >= rsum) return false;
    m.np = 1;
    m.n = (dist > EPS) ? vdiv(-delta, dist) : V3{1.0f, 0.0f, 0.0f};
    m.dep[0] = rsum - dist;
    m.p[0] = closest - m.n * cp.b;
    return true;
}

// CapsuleCapsule.hpp:24-100 (the parallel case keeps the reference's second point q1 + n*r_b)
RB_HD bool capsule_capsule(const Shape &a, const TF &ta, const Shape &b, const TF &tb, Manifold &m) {
    V3 p1, p2, q1, q2;
    capsule_ends(a, ta, p1, p2);
    capsule_ends(b, tb, q1, q2);
    V3 d1 = p2 - p1, d2 = q2 - q1, r = p1 - q1;
    float A = len2(d1), E = len2(d2), F = dot(d2, r);
    float s = 0.0f, t = 0.0f;
    float B = dot(d1, d2);
    float den = A * E - B * B;
    bool par = den < (EPS * EPS);
    if (par) {
        if (A > EPS) s = clampf(-dot(d1, r) / A, 0.0f, 1.0f);
        if (E > EPS) t = clampf(dot(d2, -r) / E, 0.0f, 1.0f);
    } else {
        float Cc = dot(d1, r);
        s = clampf((B * F - Cc * E) / den, 0.0f, 1.0f);
        t = (B * s + F) / E;
        if (t < 0.0f) { t = 0.0f; s = clampf(-Cc / A, 0.0f, 1.0f); }
        else if (t > 1.0f) { t = 1.0f; s = clampf((B - Cc) / A, 0.0f, 1.0f); }
    }
    V3 ca = p1 + d1 * s, cb = q1 + d2 * t;
    V3 delta = ca - cb;
    float dist = len(delta);
    float rsum = a.b + b.b;
    if (dist >= rsum) return false;
    V3 n = (dist > EPS) ? vdiv(delta, dist) : V3{1.0f, 0.0f, 0.0f};
    m.n = n;
    m.dep[0] = rsum - dist;
    m.p[0] = cb + n * b.b;
    if (!par) {
        m.np = 1;
    } else {
        m.np = 2;
        m.p[1] = q1 + n * b.b;
        m.dep[1] = rsum - dist;
    }
    return true;
}

// ---------------------------------------------------------------- box-box (BoxBox.hpp:29-399)
RB_HD float obb_radius(const V3 ax[3], const Shape &bx, V3 axis) {
    return absf(dot(ax[0], axis)) * bx.a + absf(dot(ax[1], axis)) * bx.b + absf(dot(ax[2], axis)) * bx.c;
}
RB_HD bool bb_axis(V3 axis, const V3 ax[3], const Shape &a, const V3 bx[3], const Shape &b, V3 cd, float &best, V3 &bn,
                   int feat, int &bf, float bias) {
    float l2 = len2(axis);
    if (l2 < EPS * EPS) return true;
    V3 na = vdiv(axis, sqrtf(l2));
    float rA = obb_radius(ax, a, na), rB = obb_radius(bx, b, na);
    float dist = absf(dot(cd, na));
    float ov = rA + rB - dist;
    if (ov < 0.0f) return false;
    if ((ov + bias) < best) {
        best = ov + bias;
        bn = (dot(cd, na) >= 0.0f) ? na : -na;
        bf = feat;
    }
    return true;
}
RB_HD void bb_face(const V3 axes[3], const Shape &bx, V3 pos, V3 dir, V3 out[4], V3 &on) {
    const float half[3] = {bx.a, bx.b, bx.c};
    int best = 0;
    float bd = absf(dot(axes[0], dir));
    for (int i = 1; i < 3; ++i) {
        float dd = absf(dot(axes[i], dir));
        if (dd > bd) { bd = dd; best = i; }
    }
    float sign = (dot(axes[best], dir) >= 0.0f) ? 1.0f : -1.0f;
    on = axes[best] * sign;
    V3 fc = pos + on * half[best];
    int u = (best + 1) % 3, v = (best + 2) % 3;
    out[0] = fc + axes[u] * half[u] + axes[v] * half[v];
    out[1] = fc - axes[u] * half[u] + axes[v] * half[v];
    out[2] = fc - axes[u] * half[u] - axes[v] * half[v];
    out[3] = fc + axes[u] * half[u] - axes[v] * half[v];
}
// Clip buffer capacity of the reference (BoxBox.hpp:310); writes are bounded by it.
constexpr int CLIP_CAP = 16;
RB_HD int bb_clip(const V3 *poly, int n, V3 pn, V3 pp, V3 *out) {
    int k = 0;
    for (int i = 0; i < n; ++i) {
        V3 A = poly[i];
        V3 B = poly[(i + 1) % n];
        float dA = dot(A - pp, pn), dB = dot(B - pp, pn);
        if (dA >= 0.0f && k < CLIP_CAP) out[k++] = A;
        if ((dA > 0.0f) != (dB > 0.0f) && k < CLIP_CAP) {
            float t = dA / (dA - dB);
            out[k++] = A + (B - A) * t;
        }
    }
    return k;
}
RB_HD int reduce4(const V3 *p, const float *dep, int n, V3 *op, float *od) {
    if (n <= 4) {
        for (int i = 0; i < n; ++i) { op[i] = p[i]; od[i] = dep[i]; }
        return n;
    }
    int i0 = 0;
    for (int i = 1; i < n; ++i)
        if (dep[i] > dep[i0]) i0 = i;
    int i1 = (i0 == 0) ? 1 : 0;
    float best = -1.0f;
    for (int i = 0; i < n; ++i) {
        if (i == i0) continue;
        float dd = len2(p[i] - p[i0]);
        if (dd > best) { best = dd; i1 = i; }
    }
    int i2 = -1;
    best = -1.0f;
    V3 seg = p[i1] - p[i0];
    for (int i = 0; i < n; ++i) {
        if (i == i0 || i == i1) continue;
        float dd = len2(cross(p[i] - p[i0], seg));
        if (dd > best) { best = dd; i2 = i; }
    }
    if (i2 < 0) { i2 = 0; while (i2 == i0 || i2 == i1) ++i2; }
    int i3 = -1;
    float mind = 1.0f;
    V3 tn = cross(p[i2] - p[i0], seg);
    for (int i = 0; i < n; ++i) {
        if (i == i0 || i == i1 || i == i2) continue;
        float dd = dot(cross(p[i] - p[i0], seg), tn);
        if (dd < mind) { mind = dd; i3 = i; }
    }
    if (i3 < 0 || mind >= 0.0f) {
        best = -1.0f;
        for (int i = 0; i < n; ++i) {
            if (i == i0 || i == i1 || i == i2) continue;
            float dd = len2(p[i] - p[i2]);
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
// The separating-axis part of box_box alone (same axes, order and arithmetic): false
// exactly when box_box returns false before building a manifold.
RB_HD bool box_box_overlap(const Shape &a, const TF &ta, const Shape &b, const TF &tb) {
    V3 cd = tb.pos - ta.pos;
    V3 ax[3] = {tf_vec(ta, V3{1, 0, 0}), tf_vec(ta, V3{0, 1, 0}), tf_vec(ta, V3{0, 0, 1})};
    V3 bx[3] = {tf_vec(tb, V3{1, 0, 0}), tf_vec(tb, V3{0, 1, 0}), tf_vec(tb, V3{0, 0, 1})};
    float best = __FLT_MAX__;
    V3 bn{1.0f, 0.0f, 0.0f};
    int bf = 0, feat = 0;
    for (int i = 0; i < 3; ++i, ++feat)
        if (!bb_axis(ax[i], ax, a, bx, b, cd, best, bn, feat, bf, 0.0f)) return false;
    for (int i = 0; i < 3; ++i, ++feat)
        if (!bb_axis(bx[i], ax, a, bx, b, cd, best, bn, feat, bf, 0.0f)) return false;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j, ++feat)
            if (!bb_axis(cross(ax[i], bx[j]), ax, a, bx, b, cd, best, bn, feat, bf, 0.001f)) return false;
    return true;
}
RB_HD bool box_box(const Shape &a, const TF &ta, const Shape &b, const TF &tb, Manifold &m) {
    V3 cd = tb.pos - ta.pos;
    V3 ax[3] = {tf_vec(ta, V3{1, 0, 0}), tf_vec(ta, V3{0, 1, 0}), tf_vec(ta, V3{0, 0, 1})};
    V3 bx[3] = {tf_vec(tb, V3{1, 0, 0}), tf_vec(tb, V3{0, 1, 0}), tf_vec(tb, V3{0, 0, 1})};
    float best = __FLT_MAX__;
    V3 bn{1.0f, 0.0f, 0.0f};
    int bf = 0, feat = 0;
    for (int i = 0; i < 3; ++i, ++feat)
        if (!bb_axis(ax[i], ax, a, bx, b, cd, best, bn, feat, bf, 0.0f)) return false;
    for (int i = 0; i < 3; ++i, ++feat)
        if (!bb_axis(bx[i], ax, a, bx, b, cd, best, bn, feat, bf, 0.0f)) return false;
    const float edge_bias = 0.001f;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j, ++feat)
            if (!bb_axis(cross(ax[i], bx[j]), ax, a, bx, b, cd, best, bn, feat, bf, edge_bias)) return false;
    if (bf >= 6) best -= edge_bias;
    m.n = bn;
    m.np = 0;
    if (bf < 6) {
        bool refA = bf < 3;
        V3 target = refA ? bn : -bn;
        V3 rc[4], rn, ic[4], inn;
        bb_face(refA ? ax : bx, refA ? a : b, refA ? ta.pos : tb.pos, target, rc, rn);
        bb_face(refA ? bx : ax, refA ? b : a, refA ? tb.pos : ta.pos, -target, ic, inn);
        V3 b0[CLIP_CAP], b1[CLIP_CAP];
        int cnt = 4;
        for (int i = 0; i < 4; ++i) b0[i] = ic[i];
        V3 rcen = (rc[0] + rc[1] + rc[2] + rc[3]) * 0.25f;
        for (int i = 0; i < 4; ++i) {
            V3 edge = rc[(i + 1) % 4] - rc[i];
            V3 sn = normalize(cross(rn, edge));
            if (dot(sn, rcen - rc[i]) < 0.0f) sn = -sn;
            cnt = bb_clip(b0, cnt, sn, rc[i], b1);
            if (cnt == 0) break;
            for (int k = 0; k < cnt; ++k) b0[k] = b1[k];
        }
        if (cnt == 0) return false;
        float rd = dot(rc[0], rn);
        V3 kp[CLIP_CAP];
        float kd[CLIP_CAP];
        int kn = 0;
        for (int i = 0; i < cnt; ++i) {
            float sd = dot(b0[i], rn) - rd;
            if (sd <= EPS) {
                kp[kn] = b0[i] - rn * (sd * 0.5f);
                kd[kn] = -sd;
                ++kn;
            }
        }
        if (kn == 0) return false;
        m.np = reduce4(kp, kd, kn, m.p, m.dep);
    } else {
        int ei = (bf - 6) / 3, ej = (bf - 6) % 3;
        const float eA[3] = {a.a, a.b, a.c}, eB[3] = {b.a, b.b, b.c};
        V3 pa = ta.pos + tf_vec(ta, V3{(dot(ax[0], bn) >= 0.0f ? 1.0f : -1.0f) * (ei == 0 ? 0.0f : eA[0]),
                                       (dot(ax[1], bn) >= 0.0f ? 1.0f : -1.0f) * (ei == 1 ? 0.0f : eA[1]),
                                       (dot(ax[2], bn) >= 0.0f ? 1.0f : -1.0f) * (ei == 2 ? 0.0f : eA[2])});
        V3 nn = -bn;
        V3 pb = tb.pos + tf_vec(tb, V3{(dot(bx[0], nn) >= 0.0f ? 1.0f : -1.0f) * (ej == 0 ? 0.0f : eB[0]),
                                       (dot(bx[1], nn) >= 0.0f ? 1.0f : -1.0f) * (ej == 1 ? 0.0f : eB[1]),
                                       (dot(bx[2], nn) >= 0.0f ? 1.0f : -1.0f) * (ej == 2 ? 0.0f : eB[2])});
        V3 d1 = ax[ei], d2 = bx[ej];
        V3 r = pa - pb;
        float bd = dot(d1, d2), cdt = dot(d1, r), fd = dot(d2, r);
        float den = 1.0f - bd * bd;
        float s = 0.0f, t = 0.0f;
        if (den > EPS) s = (bd * fd - cdt) / den;
        float ea = eA[ei], eb = eB[ej];
        if (s < -ea || s > ea) s = clampf(s, -ea, ea);
        t = s * bd + fd;
        if (t < -eb || t > eb) {
            t = clampf(t, -eb, eb);
            s = clampf(t * bd - cdt, -ea, ea);
        }
        m.np = 1;
        m.p[0] = (pa + d1 * s + pb + d2 * t) * 0.5f;
        m.dep[0] = best;
    }
    return true;
}

// CapsuleBox.hpp:28-221 (normal points from the box to the capsule as in the reference)
RB_HD bool cl_clip(float p, float q, float &tmin, float &tmax) {
    if (absf(p) < EPS) return q >= 0.0f;
    float t = q / p;
    if (p > 0.0f) { if (t < tmax) tmax = t; }
    else { if (t > tmin) tmin = t; }
    return tmin <= tmax;
}
RB_HD bool capsule_box(const Shape &cp, const TF &tc, const Shape &bx, const TF &tb, Manifold &m) {
    V3 p1, p2;
    capsule_ends(cp, tc, p1, p2);
    V3 d = p2 - p1;
    V3 l = d * 0.5f;
    V3 c = p1 + l;
    V3 ax[3] = {tf_vec(tb, V3{1, 0, 0}), tf_vec(tb, V3{0, 1, 0}), tf_vec(tb, V3{0, 0, 1})};
    const float e[3] = {bx.a, bx.b, bx.c};
    const float rad = cp.b;
    V3 diff = c - tb.pos;
    float best = __FLT_MAX__;
    V3 bn{1.0f, 0.0f, 0.0f};
    int bf = -1;
    for (int i = 0; i < 3; ++i) {
        V3 axis = ax[i];
        float ov = e[i] + (absf(dot(l, axis)) + rad) - absf(dot(diff, axis));
        if (ov < 0.0f) return false;
        if (ov < best) {
            best = ov;
            bn = (dot(diff, axis) >= 0.0f) ? axis : -axis;
            bf = i;
        }
    }
    for (int i = 0; i < 3; ++i) {
        V3 axis = cross(ax[i], l);
        float l2 = len2(axis);
        if (l2 < EPS * EPS) continue;
        axis = vdiv(axis, sqrtf(l2));
        float rbox = e[0] * absf(dot(ax[0], axis)) + e[1] * absf(dot(ax[1], axis)) + e[2] * absf(dot(ax[2], axis));
        float ov = rbox + rad - absf(dot(diff, axis));
        if (ov < 0.0f) return false;
        if (ov < best) {
            best = ov;
            bn = (dot(diff, axis) >= 0.0f) ? axis : -axis;
            bf = i + 3;
        }
    }
    m.n = -bn;
    m.np = 0;
    if (bf < 3) {
        float t0 = 0.0f, t1 = 1.0f;
        int iA = (bf + 1) % 3, iB = (bf + 2) % 3;
        V3 p1r = p1 - tb.pos;
        float pA = dot(d, ax[iA]), vA = dot(p1r, ax[iA]);
        if (!cl_clip(pA, e[iA] - vA, t0, t1)) return true;
        if (!cl_clip(-pA, e[iA] + vA, t0, t1)) return true;
        float pB = dot(d, ax[iB]), vB = dot(p1r, ax[iB]);
        if (!cl_clip(pB, e[iB] - vB, t0, t1)) return true;
        if (!cl_clip(-pB, e[iB] + vB, t0, t1)) return true;
        V3 pt0 = p1 + d * t0;
        float d0 = rad + e[bf] - dot(pt0 - tb.pos, bn);
        if (d0 > 0.0f) { m.p[m.np] = pt0 - bn * rad; m.dep[m.np] = d0; m.np++; }
        if (t1 - t0 > EPS) {
            V3 pt1 = p1 + d * t1;
            float d1 = rad + e[bf] - dot(pt1 - tb.pos, bn);
            if (d1 > 0.0f) { m.p[m.np] = pt1 - bn * rad; m.dep[m.np] = d1; m.np++; }
        }
    } else {
        V3 v = ax[bf - 3];
        V3 sb = tb.pos + ax[0] * ((dot(ax[0], bn) >= 0.0f) ? e[0] : -e[0]) + ax[1] * ((dot(ax[1], bn) >= 0.0f) ? e[1] : -e[1]) +
                ax[2] * ((dot(ax[2], bn) >= 0.0f) ? e[2] : -e[2]);
        V3 r = p1 - sb;
        float aa = len2(d), bb = dot(d, v), f = dot(v, r), cdot = dot(d, r);
        float den = aa - bb * bb;
        float t = (den > EPS) ? clampf((bb * f - cdot) / den, 0.0f, 1.0f) : 0.5f;
        V3 pt = p1 + d * t;
        m.np = 1;
        m.p[0] = pt - bn * rad;
        m.dep[0] = best;
    }
    if (m.np == 0) {
        m.np = 1;
        m.p[0] = c - bn * rad;
        m.dep[0] = best;
    }
    return true;
}

// ---------------------------------------------------------------- shape vs plane (PlaneCollision.hpp)
RB_HD void plane_world(const Shape &pl, const TF &tp, V3 &wn, float &wd) {   // Plane.hpp:47-57
    wn = tf_vec(tp, V3{pl.a, pl.b, pl.c});
    wd = pl.d + dot(wn, tp.pos);
}
// convex_vs_plane with the sphere support (PlaneCollision.hpp:40-61, Sphere.hpp:52-55)
RB_HD bool sphere_plane(const Shape &sp, const TF &ts, const Shape &pl, const TF &tp, Manifold &m) {
    V3 wn;
    float wd;
    plane_world(pl, tp, wn, wd);
    V3 lsup = normalize(tf_inv_vec(ts, -wn)) * sp.a;
    V3 wsup = ts.pos + tf_vec(ts, lsup);
    float sd = dot(wsup, wn) - wd;
    if (sd >= 0.0f) return false;
    m.n = -wn;
    m.np = 1;
    m.dep[0] = -sd;
    m.p[0] = wsup;
    return true;
}
// PlaneCollision.hpp:293-337 (the first four corners within 1e-4f, in sx,sy,sz loop order)
RB_HD bool box_plane(const Shape &bx, const TF &tb, const Shape &pl, const TF &tp, Manifold &m) {
    V3 wn;
    float wd;
    plane_world(pl, tp, wn, wd);
    V3 u[3] = {tf_vec(tb, V3{1, 0, 0}), tf_vec(tb, V3{0, 1, 0}), tf_vec(tb, V3{0, 0, 1})};
    const float eps = 1e-4f;
    float r = bx.a * absf(dot(u[0], wn)) + bx.b * absf(dot(u[1], wn)) + bx.c * absf(dot(u[2], wn));
    float cdist = dot(tb.pos, wn) - wd;
    if (cdist - r > eps) return false;
    m.n = -wn;
    m.np = 0;
    for (int sx = -1; sx <= 1; sx += 2)
        for (int sy = -1; sy <= 1; sy += 2)
            for (int sz = -1; sz <= 1; sz += 2) {
                V3 corner = tb.pos + (float)sx * bx.a * u[0] + (float)sy * bx.b * u[1] + (float)sz * bx.c * u[2];
                float dist = dot(corner, wn) - wd;
                if (dist <= eps && m.np < 4) {
                    m.p[m.np] = corner;
                    m.dep[m.np] = -dist;
                    m.np++;
                }
            }
    return m.np > 0;
}
// PlaneCollision.hpp:344-389
RB_HD bool capsule_plane(const Shape &cp, const TF &tc, const Shape &pl, const TF &tp, Manifold &m) {
    V3 wn;
    float wd;
    plane_world(pl, tp, wn, wd);
    V3 p1, p2;
    capsule_ends(cp, tc, p1, p2);
    float d1 = dot(p1, wn) - wd, d2 = dot(p2, wn) - wd;
    m.np = 0;
    m.n = -wn;
    const float eps = 1e-4f;
    V3 roff = wn * cp.b;
    float thr = cp.b + eps;
    if (d1 <= thr) { m.p[m.np] = p1 - roff; m.dep[m.np] = cp.b - d1; m.np++; }
    if (d2 <= thr) { m.p[m.np] = p2 - roff; m.dep[m.np] = cp.b - d2; m.np++; }
    return m.np > 0;
}

}  // namespace rb
