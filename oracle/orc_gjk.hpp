// Oracle: convex hulls, GJK, EPA and the hull contact manifolds.
//
// TEST INFRASTRUCTURE ONLY (see orc_math.hpp).  Operand order follows the reference
// so that the double instantiation is bit-exact against it; the float instantiation
// is the arithmetic of the GPU's wave-per-pair kernels.
//
// Restated reference routines:
//   hull tables          include/rbc/shapes/ConvexHull.hpp:149-420 (triangle normals,
//                        coplanar merge, boundary loops, adjacency tables)
//   hull mass            ConvexHull.hpp:575-745 (volume, tetrahedral inertia)
//   supports             Sphere.hpp:48-51, Box.hpp:49-55, Capsule.hpp:72-80,
//                        ConvexHull.hpp:457-473; MinkowskiDiff.hpp:28-65
//   GJK                  src/rbc/gjk/GJK.cpp:42-96, simplex/DoSimplex.cpp:5-27,
//                        Line.cpp:6-48, Triangle.cpp:6-80, Tetrahedron.cpp:6-217
//   EPA                  src/rbc/gjk/EPA.cpp:37-114 (face normal, edge cancellation),
//                        :141-378 (evaluate); pools of EPA.hpp:43-46
//   manifold             src/rbc/gjk/ContactManifoldGenerator.cpp:56-438,
//                        Clipping.cpp:6-128, ContactManifoldGenerator.hpp:88-154
//   hull-plane           include/rbc/analytic/PlaneCollision.hpp:90-285
#pragma once
#include <cstdint>
#include <limits>
#include <vector>

namespace orc {

// ------------------------------------------------------------------ hull tables
// Built once in double (the reference's scalar) and then rounded to T.
struct HullTables {
    std::vector<V3<double>> v;
    std::vector<uint32_t> tri;
    std::vector<V3<double>> fn, pn;
    std::vector<uint32_t> pidx, poff;
    std::vector<uint32_t> v2p_o, v2p_i, v2n_o, v2n_i, p2p_o, p2p_i;
};

inline void flatten_lists(const std::vector<std::vector<uint32_t>> &src, std::vector<uint32_t> &offs,
                          std::vector<uint32_t> &idx) {
    offs.assign(1, 0u);
    idx.clear();
    for (const auto &l : src) {
        for (uint32_t x : l) idx.push_back(x);
        offs.push_back((uint32_t)idx.size());
    }
}

inline HullTables build_hull_tables(const double *verts, uint32_t nv, const uint32_t *faces, uint32_t nf) {
    HullTables h;
    for (uint32_t i = 0; i < nv; ++i) h.v.push_back(V3<double>(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]));
    if (!faces || nf == 0) return h;
    h.tri.assign(faces, faces + 3 * nf);
    for (uint32_t f = 0; f < nf; ++f) {
        const V3<double> &A = h.v[faces[3 * f]], &B = h.v[faces[3 * f + 1]], &Cc = h.v[faces[3 * f + 2]];
        h.fn.push_back(normalize(cross(B - A, Cc - A)));
    }
    // coplanar groups: connected components of "normals within kCoplanarDot"
    const double kCoplanarDot = 0.99995;
    std::vector<int> parent(nf);
    for (uint32_t i = 0; i < nf; ++i) parent[i] = (int)i;
    auto find = [&parent](int x) {
        while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; }
        return x;
    };
    for (uint32_t i = 0; i < nf; ++i)
        for (uint32_t j = i + 1; j < nf; ++j)
            if (dot(h.fn[i], h.fn[j]) > kCoplanarDot) {
                int ri = find((int)i), rj = find((int)j);
                if (ri != rj) parent[ri] = rj;
            }
    std::vector<std::vector<uint32_t>> groups;
    std::vector<int> gid(nf, -1);
    for (uint32_t i = 0; i < nf; ++i) {
        int r = find((int)i);
        if (gid[r] < 0) { gid[r] = (int)groups.size(); groups.emplace_back(); }
        groups[gid[r]].push_back(i);
    }
    h.poff.push_back(0);
    for (const auto &g : groups) {
        std::vector<std::pair<uint32_t, uint32_t>> te, bd;
        for (uint32_t f : g) {
            uint32_t a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
            te.push_back({a, b}); te.push_back({b, c}); te.push_back({c, a});
        }
        for (size_t i = 0; i < te.size(); ++i) {
            bool rev = false;
            for (size_t j = 0; j < te.size() && !rev; ++j)
                rev = (i != j) && te[j].first == te[i].second && te[j].second == te[i].first;
            if (!rev) bd.push_back(te[i]);
        }
        if (bd.empty()) continue;
        std::vector<char> used(bd.size(), 0);
        std::vector<uint32_t> loop{bd[0].first};
        used[0] = 1;
        uint32_t cur = bd[0].second;
        for (size_t step = 1; step < bd.size(); ++step) {
            size_t k = 0;
            while (k < bd.size() && (used[k] || bd[k].first != cur)) ++k;
            if (k == bd.size()) break;
            used[k] = 1;
            loop.push_back(bd[k].first);
            cur = bd[k].second;
        }
        if (loop.size() < 3) continue;
        V3<double> an(0, 0, 0);
        for (uint32_t f : g) an = an + h.fn[f];
        const double al = length(an);
        if (al < C<double>::EPS) continue;
        an = an / al;
        for (uint32_t x : loop) h.pidx.push_back(x);
        h.poff.push_back((uint32_t)h.pidx.size());
        h.pn.push_back(an);
    }
    const uint32_t P = (uint32_t)h.pn.size();
    if (P == 0) { h.poff.clear(); return h; }
    std::vector<std::vector<uint32_t>> v2p(nv), v2n(nv), p2p(P);
    auto add_unique = [](std::vector<uint32_t> &l, uint32_t x) {
        for (uint32_t y : l) if (y == x) return;
        l.push_back(x);
    };
    for (uint32_t p = 0; p < P; ++p) {
        const uint32_t s = h.poff[p], n = h.poff[p + 1] - s;
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t a = h.pidx[s + i], b = h.pidx[s + (i + 1) % n];
            add_unique(v2p[a], p);
            add_unique(v2n[a], b);
            add_unique(v2n[b], a);
        }
    }
    for (uint32_t p = 0; p < P; ++p) {
        const uint32_t sp = h.poff[p], np = h.poff[p + 1] - sp;
        for (uint32_t q = 0; q < P; ++q) {
            if (q == p) continue;
            const uint32_t sq = h.poff[q], nq = h.poff[q + 1] - sq;
            bool shared = false;
            for (uint32_t i = 0; i < np && !shared; ++i) {
                const uint32_t a = h.pidx[sp + i], b = h.pidx[sp + (i + 1) % np];
                for (uint32_t j = 0; j < nq && !shared; ++j)
                    shared = a == h.pidx[sq + (j + 1) % nq] && b == h.pidx[sq + j];
            }
            if (shared) p2p[p].push_back(q);
        }
    }
    flatten_lists(v2p, h.v2p_o, h.v2p_i);
    flatten_lists(v2n, h.v2n_o, h.v2n_i);
    flatten_lists(p2p, h.p2p_o, h.p2p_i);
    return h;
}

// Volume and unit-density inertia about the centroid (ConvexHull.hpp:575-745).  Vertex-only
// hulls use the local AABB (volume and box inertia).  out: volume, xx yy zz xy xz yz.
inline void hull_mass_properties(const HullTables &h, double *out) {
    for (int i = 0; i < 7; ++i) out[i] = 0.0;
    if (h.v.empty()) return;
    if (h.tri.empty()) {
        V3<double> mn(1e30f, 1e30f, 1e30f), mx(-1e30f, -1e30f, -1e30f);
        for (const auto &p : h.v) {
            mn = V3<double>(smin(mn.x, p.x), smin(mn.y, p.y), smin(mn.z, p.z));
            mx = V3<double>(smax(mx.x, p.x), smax(mx.y, p.y), smax(mx.z, p.z));
        }
        const V3<double> e = mx - mn;
        const double m = e.x * e.y * e.z;
        out[0] = m;
        out[1] = (1.0 / 12.0) * m * (e.y * e.y + e.z * e.z);
        out[2] = (1.0 / 12.0) * m * (e.x * e.x + e.z * e.z);
        out[3] = (1.0 / 12.0) * m * (e.x * e.x + e.y * e.y);
        return;
    }
    double v6 = 0.0;
    for (size_t f = 0; f < h.tri.size() / 3; ++f)
        v6 += dot(h.v[h.tri[3 * f]], cross(h.v[h.tri[3 * f + 1]], h.v[h.tri[3 * f + 2]]));
    out[0] = v6 / 6.0;
    double total_v6 = 0.0, Ixx = 0.0, Iyy = 0.0, Izz = 0.0, Ixy = 0.0, Ixz = 0.0, Iyz = 0.0;
    V3<double> fm(0.0, 0.0, 0.0);
    for (size_t f = 0; f < h.tri.size() / 3; ++f) {
        const V3<double> &A = h.v[h.tri[3 * f]], &B = h.v[h.tri[3 * f + 1]], &Cc = h.v[h.tri[3 * f + 2]];
        const double det = dot(A, cross(B, Cc));
        const double dV = det / 6.0;
        total_v6 += det;
        fm = fm + (A + B + Cc) * (dV * 0.25);
        auto sq = [det](double a, double b, double c) { return (det / 60.0) * (a * a + b * b + c * c + a * b + a * c + b * c); };
        auto pr = [det](double ax, double ay, double bx, double by, double cx, double cy) {
            return (det / 120.0) * (2.0 * (ax * ay + bx * by + cx * cy) + ax * by + ay * bx + ax * cy + ay * cx + bx * cy + by * cx);
        };
        const double mxx = sq(A.x, B.x, Cc.x), myy = sq(A.y, B.y, Cc.y), mzz = sq(A.z, B.z, Cc.z);
        const double mxy = pr(A.x, A.y, B.x, B.y, Cc.x, Cc.y), mxz = pr(A.x, A.z, B.x, B.z, Cc.x, Cc.z),
                     myz = pr(A.y, A.z, B.y, B.z, Cc.y, Cc.z);
        Ixx += myy + mzz; Iyy += mxx + mzz; Izz += mxx + myy;
        Ixy -= mxy; Ixz -= mxz; Iyz -= myz;
    }
    const double vol = total_v6 / 6.0;
    if (std::abs(vol) < C<double>::EPS) return;
    const V3<double> c = fm / vol;
    const double m = vol;
    out[1] = Ixx - m * (c.y * c.y + c.z * c.z);
    out[2] = Iyy - m * (c.x * c.x + c.z * c.z);
    out[3] = Izz - m * (c.x * c.x + c.y * c.y);
    out[4] = Ixy + m * c.x * c.y;
    out[5] = Ixz + m * c.x * c.z;
    out[6] = Iyz + m * c.y * c.z;
}

// Hull data in the working precision.
template <class T> struct Hull {
    std::vector<V3<T>> v, pn;
    std::vector<uint32_t> pidx, poff, v2p_o, v2p_i, v2n_o, v2n_i, p2p_o, p2p_i;
    V3<T> amin, amax;   // local AABB
    bool adj = false;   // polygons and adjacency present (ContactManifoldGenerator.cpp:48-54)
    static Hull from(const HullTables &t) {
        Hull h;
        auto cv = [](const V3<double> &a) { return V3<T>(T(a.x), T(a.y), T(a.z)); };
        for (const auto &p : t.v) h.v.push_back(cv(p));
        for (const auto &p : t.pn) h.pn.push_back(cv(p));
        h.pidx = t.pidx; h.poff = t.poff;
        h.v2p_o = t.v2p_o; h.v2p_i = t.v2p_i; h.v2n_o = t.v2n_o; h.v2n_i = t.v2n_i; h.p2p_o = t.p2p_o; h.p2p_i = t.p2p_i;
        h.adj = !t.pn.empty() && !t.v2p_i.empty() && !t.v2n_i.empty() && !t.p2p_i.empty();
        T big = T(1e30f);
        h.amin = V3<T>(big, big, big);
        h.amax = V3<T>(-big, -big, -big);
        for (const auto &p : h.v) {
            h.amin = V3<T>(smin(h.amin.x, p.x), smin(h.amin.y, p.y), smin(h.amin.z, p.z));
            h.amax = V3<T>(smax(h.amax.x, p.x), smax(h.amax.y, p.y), smax(h.amax.z, p.z));
        }
        return h;
    }
};

// ConvexHull.hpp:484-499
template <class T> inline AABB<T> hull_aabb(const Hull<T> &h, const TF<T> &tf) {
    const V3<T> c = (h.amin + h.amax) * T(0.5);
    const V3<T> e = (h.amax - h.amin) * T(0.5);
    const M3<T> R = mat3_cast(tf.rot);
    const V3<T> wc = tf.pos + tf.vec(c);
    const V3<T> ext(std::abs(R.c[0].x) * e.x + std::abs(R.c[1].x) * e.y + std::abs(R.c[2].x) * e.z,
                    std::abs(R.c[0].y) * e.x + std::abs(R.c[1].y) * e.y + std::abs(R.c[2].y) * e.z,
                    std::abs(R.c[0].z) * e.x + std::abs(R.c[1].z) * e.y + std::abs(R.c[2].z) * e.z);
    return AABB<T>{wc - ext, wc + ext};
}

// ------------------------------------------------------------------ supports
template <class T> inline uint32_t hull_support_index(const Hull<T> &h, const V3<T> &d) {
    uint32_t best = 0;
    T bd = dot(h.v[0], d);
    for (uint32_t i = 1; i < h.v.size(); ++i) {
        const T x = dot(h.v[i], d);
        if (x > bd) { bd = x; best = i; }
    }
    return best;
}
template <class T> inline V3<T> shape_support(const ShapeP<T> &s, const V3<T> &d) {
    switch (s.kind) {
        case SH_SPHERE: return normalize(d) * s.a;
        case SH_BOX: return V3<T>(d.x > 0 ? s.a : -s.a, d.y > 0 ? s.b : -s.b, d.z > 0 ? s.c : -s.c);
        case SH_CAPSULE: {
            const V3<T> e(T(0), d.y >= T(0) ? s.a : -s.a, T(0));
            const T l = length(d);
            if (l < C<T>::EPS) return e;
            return e + (d / l) * s.b;
        }
        case SH_HULL: return s.hull->v.empty() ? V3<T>() : s.hull->v[hull_support_index(*s.hull, d)];
        default: return V3<T>();
    }
}

template <class T> struct SVert { V3<T> w0, w1, w; };

template <class T> struct MinkDiff {
    const ShapeP<T> *a, *b;
    TF<T> ta, tb;
    V3<T> sup_a(const V3<T> &d) const { return ta.point(shape_support(*a, ta.inv_vec(d))); }
    V3<T> sup_b(const V3<T> &d) const { return tb.point(shape_support(*b, tb.inv_vec(d))); }
    SVert<T> sample(const V3<T> &d) const {
        SVert<T> s;
        s.w0 = sup_a(d);
        s.w1 = sup_b(-d);
        s.w = s.w0 - s.w1;
        return s;
    }
};

// ------------------------------------------------------------------ GJK
enum { GJK_VALID = 0, GJK_INSIDE = 1, GJK_FAILED = 2 };
template <class T> struct Simplex {
    SVert<T> v[4];   // v[0] = newest
    int rank = 0;
    void push(const SVert<T> &s) {
        for (int i = (rank < 3 ? rank : 3); i > 0; --i) v[i] = v[i - 1];
        v[0] = s;
        if (rank < 4) ++rank;
    }
};
template <class T> inline V3<T> triple(const V3<T> &a, const V3<T> &b, const V3<T> &c) { return cross(cross(a, b), c); }

template <class T> inline void simplex2(Simplex<T> &s, V3<T> &dir) {
    const V3<T> a = s.v[0].w, b = s.v[1].w;
    const V3<T> ao = -a, ab = b - a;
    if (dot(ab, ao) >= T(0)) {
        s.rank = 2;
        dir = triple(ab, ao, ab);
        if (length_sq(dir) < T(1e-6)) {
            const T ax = std::abs(ab.x), ay = std::abs(ab.y), az = std::abs(ab.z);
            const V3<T> axis = (ax <= ay && ax <= az) ? V3<T>(1, 0, 0) : (ay <= az ? V3<T>(0, 1, 0) : V3<T>(0, 0, 1));
            dir = cross(ab, axis);
            if (length_sq(dir) < T(1e-6)) dir = cross(ab, V3<T>(0, 0, 1));
        }
    } else {
        s.rank = 1;
        dir = ao;
    }
}
template <class T> inline void simplex3(Simplex<T> &s, V3<T> &dir) {
    const V3<T> a = s.v[0].w, b = s.v[1].w, c = s.v[2].w;
    const V3<T> ao = -a, ab = b - a, ac = c - a, abc = cross(ab, ac);
    if (dot(cross(abc, ac), ao) >= T(0)) {
        if (dot(ac, ao) >= T(0)) { s.v[1] = s.v[2]; s.rank = 2; dir = triple(ac, ao, ac); }
        else if (dot(ab, ao) >= T(0)) { s.rank = 2; dir = triple(ab, ao, ab); }
        else { s.rank = 1; dir = ao; }
    } else if (dot(cross(ab, abc), ao) >= T(0)) {
        if (dot(ab, ao) >= T(0)) { s.rank = 2; dir = triple(ab, ao, ab); }
        else { s.rank = 1; dir = ao; }
    } else if (dot(abc, ao) >= T(0)) {
        s.rank = 3; dir = abc;
    } else {
        std::swap(s.v[1], s.v[2]); s.rank = 3; dir = -abc;
    }
}
// Reduction of a triangle face (a, b, c) of the tetrahedron with normal n toward the
// origin; the three outcomes of Tetrahedron.cpp's per-face cases.
template <class T>
inline void tet_face(Simplex<T> &s, V3<T> &dir, const SVert<T> &A, const SVert<T> &Bv, const SVert<T> &Cv,
                     const V3<T> &ab, const V3<T> &ac, const V3<T> &n, const V3<T> &ao) {
    if (dot(cross(n, ac), ao) >= T(0)) {
        if (dot(ac, ao) >= T(0)) { s.v[0] = A; s.v[1] = Cv; s.rank = 2; dir = triple(ac, ao, ac); }
        else if (dot(ab, ao) >= T(0)) { s.v[0] = A; s.v[1] = Bv; s.rank = 2; dir = triple(ab, ao, ab); }
        else { s.v[0] = A; s.rank = 1; dir = ao; }
    } else if (dot(cross(ab, n), ao) >= T(0)) {
        if (dot(ab, ao) >= T(0)) { s.v[0] = A; s.v[1] = Bv; s.rank = 2; dir = triple(ab, ao, ab); }
        else { s.v[0] = A; s.rank = 1; dir = ao; }
    } else {
        s.v[0] = A; s.v[1] = Bv; s.v[2] = Cv; s.rank = 3; dir = n;
    }
}
template <class T> inline bool simplex4(Simplex<T> &s, V3<T> &dir) {
    const SVert<T> va = s.v[0], vb = s.v[1], vc = s.v[2], vd = s.v[3];
    const V3<T> a = va.w, b = vb.w, c = vc.w, d = vd.w;
    const V3<T> ao = -a, ab = b - a, ac = c - a, ad = d - a;
    const V3<T> abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
    const T eps = T(1e-10);
    int pi = 0;
    if (dot(abc, ao) > eps) pi |= 1;
    if (dot(acd, ao) > eps) pi |= 2;
    if (dot(adb, ao) > eps) pi |= 4;
    auto line = [&](const SVert<T> &X, const V3<T> &ax) {
        if (dot(ax, ao) >= T(0)) { s.v[0] = va; s.v[1] = X; s.rank = 2; dir = triple(ax, ao, ax); }
        else { s.v[0] = va; s.rank = 1; dir = ao; }
    };
    switch (pi) {
        case 0: return true;
        case 1: tet_face(s, dir, va, vb, vc, ab, ac, abc, ao); break;   // ABC
        case 2: tet_face(s, dir, va, vc, vd, ac, ad, acd, ao); break;   // ACD
        case 3: line(vc, ac); break;                                    // AC
        case 4: tet_face(s, dir, va, vd, vb, ad, ab, adb, ao); break;   // ADB
        case 5: line(vb, ab); break;                                    // AB
        case 6: line(vd, ad); break;                                    // AD
        default: s.v[0] = va; s.rank = 1; dir = ao; break;              // A
    }
    return false;
}

template <class T> struct GjkOut {
    int status = GJK_FAILED, iters = 0;
    Simplex<T> s;
    V3<T> dir;   // search direction at exit (a separating direction when Valid)
};
template <class T> inline GjkOut<T> gjk(const MinkDiff<T> &md, const V3<T> &guess, int max_iter = 128) {
    GjkOut<T> o;
    V3<T> dir = (length_sq(guess) > C<T>::EPS) ? guess : V3<T>(1, 0, 0);
    const SVert<T> v0 = md.sample(dir);
    o.s.push(v0);
    dir = -v0.w;
    for (int it = 0; it < max_iter; ++it) {
        o.iters = it + 1;
        const SVert<T> nx = md.sample(dir);
        if (dot(nx.w, dir) < T(0)) { o.status = GJK_VALID; o.dir = dir; return o; }
        o.s.push(nx);
        bool inside = false;
        if (o.s.rank == 2) simplex2(o.s, dir);
        else if (o.s.rank == 3) simplex3(o.s, dir);
        else if (o.s.rank == 4) inside = simplex4(o.s, dir);
        if (inside) { o.status = GJK_INSIDE; return o; }
    }
    o.status = GJK_FAILED;
    return o;
}

// ------------------------------------------------------------------ EPA
enum { EPA_VALID = 0, EPA_FAILED = 1, EPA_OUT_OF_FACES = 2, EPA_OUT_OF_VERTICES = 3, EPA_NONLINEAR = 4 };
constexpr int EPA_MAX_FACES = 256, EPA_MAX_VERTS = 128, EPA_MAX_ITER = 128;

template <class T> struct EpaFace {
    V3<T> n;
    T d;
    int v[3];
    bool obsolete;
};
template <class T> struct EpaOut {
    int status = EPA_FAILED, iters = 0;
    V3<T> n, p;
    T depth = 0;
};

template <class T> inline bool face_normal(EpaFace<T> &f, const std::vector<SVert<T>> &vs) {
    const T tol = T(1e-12);
    const V3<T> a = vs[f.v[0]].w, b = vs[f.v[1]].w, c = vs[f.v[2]].w;
    V3<T> n = cross(b - a, c - a);
    const T nl = length(n);
    if (nl < C<T>::EPS) return false;
    n = n / nl;
    T d = dot(n, a);
    if (d < -tol) {
        n = -n;
        d = -d;
    } else if (d <= tol) {
        bool fixed = false;
        for (size_t i = 0; i < vs.size(); ++i) {
            const T t = dot(n, vs[i].w);
            if (t > tol) { n = -n; d = T(0); fixed = true; break; }
            if (t < -tol) { d = T(0); fixed = true; break; }
        }
        if (!fixed) return false;
    }
    f.n = n;
    f.d = d;
    f.obsolete = false;
    return true;
}

template <class T> inline EpaOut<T> epa(const Simplex<T> &seed, const MinkDiff<T> &md, const V3<T> *seed_n = nullptr) {
    EpaOut<T> o;
    if (seed.rank != 4) { o.status = EPA_NONLINEAR; return o; }
    std::vector<SVert<T>> vs(seed.v, seed.v + 4);
    {   // coplanar seed: touching contact (EPA.cpp:174-212)
        const V3<T> va = vs[0].w, vb = vs[1].w, vc = vs[2].w, vd = vs[3].w;
        const T e2 = C<T>::EPS * C<T>::EPS;
        V3<T> pn = cross(vb - va, vc - va);
        if (length_sq(pn) < e2) pn = cross(vb - va, vd - va);
        if (length_sq(pn) < e2) pn = cross(vc - va, vd - va);
        if (length_sq(pn) > e2) {
            pn = normalize(pn);
            const T dd = std::abs(dot(vd - va, pn));
            if (dd < T(1e-5)) {
                const V3<T> sp = md.sup_a(pn) - md.sup_b(-pn);
                const V3<T> sn = md.sup_a(-pn) - md.sup_b(pn);
                const T dp = dot(sp - va, pn);
                const T dn = dot(sn - va, -pn);
                if (dp < T(1e-4) || dn < T(1e-4)) {
                    o.n = (dp <= dn) ? pn : -pn;
                    o.depth = T(0);
                    o.p = (vs[0].w0 + vs[1].w0 + vs[2].w0 + vs[3].w0) * T(0.25);
                    o.status = EPA_VALID;
                    return o;
                }
            }
        }
    }
    std::vector<EpaFace<T>> fs;
    auto add_face = [&](int p, int q, int r) {
        EpaFace<T> f;
        f.v[0] = p; f.v[1] = q; f.v[2] = r;
        f.obsolete = false;
        if (!face_normal(f, vs)) return false;
        fs.push_back(f);
        return true;
    };
    if (!add_face(0, 1, 2) || !add_face(0, 2, 3) || !add_face(0, 3, 1) || !add_face(1, 2, 3)) {
        o.status = EPA_NONLINEAR;
        return o;
    }
    std::vector<std::pair<int, int>> edges;
    // Adds w to the polytope: faces visible from w become obsolete, the horizon of the
    // visible region (add_edge toggling) is stitched to w.  false: face pool exhausted.
    auto expand = [&](const SVert<T> &w) {
        vs.push_back(w);
        const int wi = (int)vs.size() - 1;
        edges.clear();
        auto add_edge = [&edges](int a, int b) {
            for (size_t i = 0; i < edges.size(); ++i)
                if ((edges[i].first == a && edges[i].second == b) || (edges[i].first == b && edges[i].second == a)) {
                    edges.erase(edges.begin() + (long)i);
                    return;
                }
            edges.push_back({a, b});
        };
        for (auto &f : fs) {
            if (f.obsolete) continue;
            const V3<T> cen = (vs[f.v[0]].w + vs[f.v[1]].w + vs[f.v[2]].w) * T(1.0 / 3.0);
            if (dot(f.n, w.w - cen) > T(0)) {
                f.obsolete = true;
                add_edge(f.v[0], f.v[1]);
                add_edge(f.v[1], f.v[2]);
                add_edge(f.v[2], f.v[0]);
            }
        }
        for (const auto &e : edges) {
            if ((int)fs.size() >= EPA_MAX_FACES) return false;
            add_face(e.first, e.second, wi);
        }
        return true;
    };
    const T big = std::numeric_limits<T>::max();
    int it0 = 0;
    if (seed_n) {
        // warm start: one expansion along a predicted normal before the first closest-face
        // step, when its support is a new vertex that some face sees
        it0 = 1;
        const SVert<T> w = md.sample(*seed_n);
        bool fresh = true, seen = false;
        for (size_t i = 0; i < vs.size(); ++i) fresh = fresh && !(length_sq(vs[i].w - w.w) < C<T>::EPS * C<T>::EPS);
        for (const auto &f : fs) {
            const V3<T> cen = (vs[f.v[0]].w + vs[f.v[1]].w + vs[f.v[2]].w) * T(1.0 / 3.0);
            seen = seen || (!f.obsolete && dot(f.n, w.w - cen) > T(0));
        }
        if (fresh && seen && !expand(w)) { o.iters = 1; o.status = EPA_OUT_OF_FACES; return o; }
    }
    for (int it = 0; it < EPA_MAX_ITER; ++it) {
        o.iters = it0 + it + 1;
        int ci = -1;
        T md_ = big;
        for (size_t i = 0; i < fs.size(); ++i)
            if (!fs[i].obsolete && fs[i].d < md_) { md_ = fs[i].d; ci = (int)i; }
        if (ci < 0) { o.status = EPA_FAILED; return o; }
        const EpaFace<T> cf = fs[ci];
        const SVert<T> w = md.sample(cf.n);
        const T dproj = dot(w.w, cf.n);
        if (dproj - md_ < T(1e-6)) {
            o.n = cf.n;
            o.depth = md_;
            const V3<T> q = cf.n * md_;
            const V3<T> p0 = vs[cf.v[0]].w, p1 = vs[cf.v[1]].w, p2 = vs[cf.v[2]].w;
            const V3<T> e0 = p1 - p0, e1 = p2 - p0, r = q - p0;
            const T d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1), d20 = dot(r, e0), d21 = dot(r, e1);
            const T den = d00 * d11 - d01 * d01;
            if (std::abs(den) < C<T>::EPS) {
                o.p = vs[cf.v[0]].w0;
            } else {
                const T vw = (d11 * d20 - d01 * d21) / den;
                const T ww = (d00 * d21 - d01 * d20) / den;
                const T uw = T(1.0) - vw - ww;
                o.p = vs[cf.v[0]].w0 * uw + vs[cf.v[1]].w0 * vw + vs[cf.v[2]].w0 * ww;
            }
            o.status = EPA_VALID;
            return o;
        }
        const T e2 = C<T>::EPS * C<T>::EPS;
        for (size_t i = 0; i < vs.size(); ++i)
            if (length_sq(vs[i].w - w.w) < e2) {
                o.n = cf.n;
                o.depth = md_;
                o.p = vs[cf.v[0]].w0;
                o.status = EPA_VALID;
                return o;
            }
        if ((int)vs.size() >= EPA_MAX_VERTS) { o.status = EPA_OUT_OF_VERTICES; return o; }
        if (!expand(w)) { o.status = EPA_OUT_OF_FACES; return o; }
    }
    o.status = EPA_FAILED;
    return o;
}

// ------------------------------------------------------------------ manifolds
template <class T> struct ClipPlane { V3<T> p, n; };

template <class T> inline bool in_plane(const ClipPlane<T> &pl, const V3<T> &p) {
    const T d = -dot(pl.n, pl.p);
    return dot(p, pl.n) + d >= T(0);
}
template <class T> inline bool plane_edge(const ClipPlane<T> &pl, const V3<T> &s, const V3<T> &e, V3<T> &out) {
    const V3<T> ab = e - s;
    const T abp = dot(pl.n, ab);
    if (std::abs(abp) <= T(1e-6)) return false;
    const T dist = -dot(pl.n, pl.p);
    const V3<T> pco = pl.n * (-dist);
    T fac = -dot(pl.n, s - pco) / abp;
    if (fac < T(0)) fac = T(0);
    if (fac > T(1)) fac = T(1);
    out = s + ab * fac;
    return true;
}
// Clipping.cpp:38-102 (output capacity never binds for convex inputs and is not modelled)
template <class T>
inline std::vector<V3<T>> clip_polygon(std::vector<V3<T>> a, const std::vector<ClipPlane<T>> &planes, bool remove_only) {
    std::vector<V3<T>> b;
    for (const auto &pl : planes) {
        if (a.empty()) break;
        b.clear();
        V3<T> sp = a.back();
        for (const auto &ep : a) {
            const bool si = in_plane(pl, sp), ei = in_plane(pl, ep);
            if (remove_only) {
                if (ei) b.push_back(ep);
            } else if (si && ei) {
                b.push_back(ep);
            } else if (si && !ei) {
                V3<T> cut;
                if (plane_edge(pl, sp, ep, cut)) b.push_back(cut);
            } else if (!si && ei) {
                V3<T> cut;
                if (plane_edge(pl, sp, ep, cut)) b.push_back(cut);
                b.push_back(ep);
            }
            sp = ep;
        }
        a.swap(b);
    }
    return a;
}
// Clipping.cpp:104-128
template <class T>
inline bool skew_lines(const V3<T> &p1, const V3<T> &d1, const V3<T> &p2, const V3<T> &d2, V3<T> &l1, V3<T> &l2) {
    const T n1 = d1.x * d2.x + d1.y * d2.y + d1.z * d2.z;
    const T n2 = d2.x * d2.x + d2.y * d2.y + d2.z * d2.z;
    const T m1 = -d1.x * d1.x - d1.y * d1.y - d1.z * d1.z;
    const T m2 = -d2.x * d1.x - d2.y * d1.y - d2.z * d1.z;
    const T r1 = -d1.x * p2.x + d1.x * p1.x - d1.y * p2.y + d1.y * p1.y - d1.z * p2.z + d1.z * p1.z;
    const T r2 = -d2.x * p2.x + d2.x * p1.x - d2.y * p2.y + d2.y * p1.y - d2.z * p2.z + d2.z * p1.z;
    const T den = (n1 * m2) - (n2 * m1);
    if (std::abs(den) < C<T>::EPS) return false;
    const T n = ((r1 * m2) - (r2 * m1)) / den;
    const T m = ((n1 * r2) - (n2 * r1)) / den;
    l1 = p1 + d1 * m;
    l2 = p2 + d2 * n;
    return true;
}
// ContactManifoldGenerator.hpp:88-154
template <class T> inline int reduce4_area(const std::vector<V3<T>> &p, const std::vector<T> &d, V3<T> *op, T *od) {
    const int n = (int)p.size();
    if (n <= 4) {
        for (int i = 0; i < n; ++i) { op[i] = p[i]; od[i] = d[i]; }
        return n;
    }
    int i0 = 0;
    for (int i = 1; i < n; ++i) if (d[i] > d[i0]) i0 = i;
    int i1 = (i0 == 0) ? 1 : 0;
    T best = length_sq(p[i1] - p[i0]);
    for (int i = 0; i < n; ++i) {
        if (i == i0) continue;
        const T x = length_sq(p[i] - p[i0]);
        if (x > best) { best = x; i1 = i; }
    }
    int i2 = -1;
    best = T(0);
    const V3<T> seg = p[i1] - p[i0];
    for (int i = 0; i < n; ++i) {
        if (i == i0 || i == i1) continue;
        const T x = length_sq(cross(p[i] - p[i0], seg));
        if (x > best) { best = x; i2 = i; }
    }
    if (i2 < 0) for (int i = 0; i < n; ++i) if (i != i0 && i != i1) { i2 = i; break; }
    int i3 = -1;
    best = T(0);
    if (i2 >= 0) {
        const V3<T> tn = cross(p[i1] - p[i0], p[i2] - p[i0]);
        for (int i = 0; i < n; ++i) {
            if (i == i0 || i == i1 || i == i2) continue;
            const T x = std::abs(dot(p[i] - p[i0], tn));
            if (x > best) { best = x; i3 = i; }
        }
        if (i3 < 0) for (int i = 0; i < n; ++i) if (i != i0 && i != i1 && i != i2) { i3 = i; break; }
    }
    op[0] = p[i0]; od[0] = d[i0];
    op[1] = p[i1]; od[1] = d[i1];
    if (i2 >= 0) { op[2] = p[i2]; od[2] = d[i2]; }
    if (i3 >= 0) { op[3] = p[i3]; od[3] = d[i3]; return 4; }
    return (i2 >= 0) ? 3 : 2;
}

template <class T> inline void single_point(Manifold<T> &m, const V3<T> &n, const V3<T> &p, T dep) {
    m.normal = n;
    m.n = 1;
    m.pos[0] = p;
    m.depth[0] = dep;
}

// ContactManifoldGenerator.cpp:176-379 (both shapes hulls with adjacency tables)
template <class T>
inline void manifold_hulls(const V3<T> &en, T edep, const V3<T> &ep, const Hull<T> &ha, const TF<T> &ta,
                           const Hull<T> &hb, const TF<T> &tb, Manifold<T> &m) {
    const T lowest = -std::numeric_limits<T>::max();
    m.normal = en;
    m.n = 0;
    const uint32_t sa = hull_support_index(ha, ta.inv_vec(en));
    const uint32_t sb = hull_support_index(hb, tb.inv_vec(-en));
    auto best_poly = [&](const Hull<T> &h, const TF<T> &tf, uint32_t sv, const V3<T> &wd, T &al) {
        const uint32_t s = h.v2p_o[sv], e = h.v2p_o[sv + 1];
        uint32_t best = (s < e) ? h.v2p_i[s] : 0;
        T bd = lowest;
        for (uint32_t i = s; i < e; ++i) {
            const uint32_t p = h.v2p_i[i];
            const T x = dot(tf.vec(h.pn[p]), wd);
            if (x > bd) { bd = x; best = p; }
        }
        al = bd;
        return best;
    };
    T al_a, al_b;
    const uint32_t pa = best_poly(ha, ta, sa, en, al_a);
    const uint32_t pb = best_poly(hb, tb, sb, -en, al_b);
    // best edge pair (deg_a x deg_b candidates, both signs of the cross product)
    bool has_edge = false;
    T al_e = lowest;
    uint32_t ea0 = 0, ea1 = 0, eb0 = 0, eb1 = 0;
    {
        const V3<T> wsa = ta.point(ha.v[sa]), wsb = tb.point(hb.v[sb]);
        const uint32_t as = ha.v2n_o[sa], ae = ha.v2n_o[sa + 1], bs = hb.v2n_o[sb], be = hb.v2n_o[sb + 1];
        if (as < ae && bs < be) {
            for (uint32_t i = as; i < ae; ++i) {
                const uint32_t na = ha.v2n_i[i];
                const V3<T> ea = wsa - ta.point(ha.v[na]);
                for (uint32_t j = bs; j < be; ++j) {
                    const uint32_t nb = hb.v2n_i[j];
                    const V3<T> eb = wsb - tb.point(hb.v[nb]);
                    const V3<T> c = cross(ea, eb);
                    const T cl = length(c);
                    if (cl < C<T>::EPS) continue;
                    const V3<T> cn = c / cl;
                    const T dp = dot(cn, en);
                    if (dp > al_e) { al_e = dp; ea0 = sa; ea1 = na; eb0 = sb; eb1 = nb; has_edge = true; }
                    const T dn = -dp;
                    if (dn > al_e) { al_e = dn; ea0 = sa; ea1 = na; eb0 = sb; eb1 = nb; has_edge = true; }
                }
            }
        }
    }
    const T EPS4 = T(1e-4);
    if (has_edge && al_e > al_a + EPS4 && al_e > al_b + EPS4) {
        const V3<T> p1 = ta.point(ha.v[ea0]);
        const V3<T> d1 = ta.point(ha.v[ea1]) - p1;
        const V3<T> p2 = tb.point(hb.v[eb0]);
        const V3<T> d2 = tb.point(hb.v[eb1]) - p2;
        V3<T> l1, l2;
        if (skew_lines(p1, d1, p2, d2, l1, l2)) single_point(m, en, (l1 + l2) * T(0.5), edep);
        else single_point(m, en, ep, edep);
        return;
    }
    const bool ref_a = al_a > al_b;
    const Hull<T> &hr = ref_a ? ha : hb;
    const Hull<T> &hi = ref_a ? hb : ha;
    const TF<T> &tr = ref_a ? ta : tb;
    const TF<T> &tinc = ref_a ? tb : ta;
    const uint32_t rp = ref_a ? pa : pb, ip = ref_a ? pb : pa;
    const V3<T> rn = tr.vec(hr.pn[rp]);
    const uint32_t is = hi.poff[ip], ie = hi.poff[ip + 1];
    if (ie - is < 3) { single_point(m, en, ep, edep); return; }
    std::vector<V3<T>> inc;
    for (uint32_t i = is; i < ie; ++i) inc.push_back(tinc.point(hi.v[hi.pidx[i]]));
    const V3<T> rpt = tr.point(hr.v[hr.pidx[hr.poff[rp]]]);
    std::vector<ClipPlane<T>> planes;
    for (uint32_t i = hr.p2p_o[rp]; i < hr.p2p_o[rp + 1]; ++i) {
        const uint32_t q = hr.p2p_i[i];
        planes.push_back({tr.point(hr.v[hr.pidx[hr.poff[q]]]), -tr.vec(hr.pn[q])});
    }
    std::vector<V3<T>> clipped = planes.empty() ? inc : clip_polygon(inc, planes, false);
    const std::vector<V3<T>> fin = clip_polygon(clipped, {ClipPlane<T>{rpt, -rn}}, true);
    std::vector<V3<T>> kp;
    std::vector<T> kd;
    for (const auto &p : fin) {
        const T dtp = dot(p - rpt, rn);
        const V3<T> diff = rn * dtp;
        const T pen = ref_a ? dot(diff, en) : -dot(diff, en);
        if (pen <= T(0)) { kp.push_back(p); kd.push_back(-pen); }
    }
    if (kp.empty()) { single_point(m, en, ep, edep); return; }
    m.n = reduce4_area(kp, kd, m.pos, m.depth);
}

// GJK + EPA + manifold (ContactManifoldGenerator.cpp:382-438).  Warm start (no
// counterpart in the reference): guess_local / seed_local, in A's frame, replace the
// GJK seed direction tf_b.pos - tf_a.pos and add one EPA expansion along the seed
// normal; when the warm-started run fails (GJK cap, EPA not valid) the plain run is the
// fallback (also when the warm-started GJK is not done within 16 iterations).  stats,
// when given, receive the statuses and iteration counts of the runs
// made (both runs' iterations when the fallback ran) and the exit directions.
//
// EPA guard (no counterpart in the reference; off unless `guard` is given): any unit
// direction n bounds the penetration depth from above by the overlap of the shapes along n,
// h(n) = support_{A-B}(n) . n.  An FP32 EPA that settles on a face deeper than the smallest
// h over the guard's directions by more than EPA_GUARD_TOL has not found the minimum; the
// result is replaced by (that direction, its h) with A's support point along it as the
// witness.  Directions, in this order (the first smallest wins): the pair's contact normal of
// the previous substep (has_prev), then with `faces` (when the depth in question exceeds
// EPA_GUARD_SCAN) the face normals of A (a hull's merged polygons, a box's +-x +-y +-z) and
// the reversed face normals of B.
constexpr float EPA_GUARD_TOL = 1e-4f;
// The face normals are tried only when the depth still in question exceeds this (m): resting
// contacts are a fraction of a millimetre deep, the faces FP32 EPA settles on wrongly are
// centimetres deep, and the scan costs a support query per face.
constexpr float EPA_GUARD_SCAN = 2e-3f;
template <class T> struct EpaGuard {
    bool has_prev = false, faces = false;
    V3<T> prev;
};
template <class T> inline int face_normal_count(const ShapeP<T> &s) {
    return s.kind == SH_HULL ? (int)s.hull->pn.size() : s.kind == SH_BOX ? 6 : 0;
}
template <class T> inline V3<T> face_normal_local(const ShapeP<T> &s, int k) {
    if (s.kind == SH_HULL) return s.hull->pn[k];
    const T sg = (k & 1) ? T(-1) : T(1);
    return k < 2 ? V3<T>(sg, T(0), T(0)) : k < 4 ? V3<T>(T(0), sg, T(0)) : V3<T>(T(0), T(0), sg);
}
struct GjkEpaStats { int gjk_status, gjk_iters, epa_status, epa_iters, fallback; };
template <class T>
inline bool gjk_epa_manifold(const ShapeP<T> &a, const TF<T> &ta, const ShapeP<T> &b, const TF<T> &tb, Manifold<T> &m,
                             const V3<T> *guess_local = nullptr, GjkEpaStats *st = nullptr, EpaOut<T> *epa_out = nullptr,
                             const V3<T> *seed_local = nullptr, V3<T> *gjk_dir = nullptr,
                             const EpaGuard<T> *guard = nullptr, T *guard_cut = nullptr) {
    MinkDiff<T> md{&a, &b, ta, tb};
    V3<T> plain = tb.pos - ta.pos;
    if (length_sq(plain) < C<T>::EPS) plain = V3<T>(1, 0, 0);
    const bool warm = guess_local != nullptr;
    const V3<T> g = warm ? ta.vec(*guess_local) : plain;
    const V3<T> sn = (warm && seed_local) ? ta.vec(*seed_local) : V3<T>();
    GjkOut<T> gr = gjk(md, g, warm ? 16 : 128);   // a warm start gets 16 iterations before the fallback
    EpaOut<T> e;
    if (gr.status == GJK_INSIDE) e = epa(gr.s, md, (warm && seed_local) ? &sn : nullptr);
    int gi = gr.iters, ei = gr.status == GJK_INSIDE ? e.iters : 0, fb = 0;
    if (warm && (gr.status == GJK_FAILED || (gr.status == GJK_INSIDE && e.status != EPA_VALID))) {
        fb = 1;
        gr = gjk(md, plain);
        e = EpaOut<T>();
        if (gr.status == GJK_INSIDE) e = epa(gr.s, md);
        gi += gr.iters;
        ei += gr.status == GJK_INSIDE ? e.iters : 0;
    }
    if (st) { st->gjk_status = gr.status; st->gjk_iters = gi; st->epa_status = gr.status == GJK_INSIDE ? e.status : -1;
              st->epa_iters = ei; st->fallback = fb; }
    if (guard && gr.status == GJK_INSIDE && e.status == EPA_VALID) {
        bool any = false;
        V3<T> bn, bp;
        T bh = T(0);
        auto consider = [&](const V3<T> &n) {
            const SVert<T> w = md.sample(n);
            const T h = dot(w.w, n);
            if (!any || h < bh) { any = true; bh = h; bn = n; bp = w.w0; }
        };
        if (guard->has_prev) consider(guard->prev);
        if (guard->faces && e.depth > T(EPA_GUARD_SCAN) && (!any || bh > T(EPA_GUARD_SCAN))) {
            for (int k = 0, nf = face_normal_count(a); k < nf; ++k) consider(ta.vec(face_normal_local(a, k)));
            for (int k = 0, nf = face_normal_count(b); k < nf; ++k) consider(-tb.vec(face_normal_local(b, k)));
        }
        if (any && bh < e.depth - T(EPA_GUARD_TOL)) {
            if (guard_cut) *guard_cut = e.depth - bh;
            e.n = bn;
            e.depth = bh;
            e.p = bp;
        }
    }
    if (gjk_dir) *gjk_dir = gr.dir;
    if (epa_out) *epa_out = e;
    if (gr.status != GJK_INSIDE || e.status != EPA_VALID) return false;
    m.normal = e.n;
    m.n = 0;
    if (a.kind == SH_HULL && b.kind == SH_HULL && a.hull->adj && b.hull->adj)
        manifold_hulls(e.n, e.depth, e.p, *a.hull, ta, *b.hull, tb, m);
    else
        single_point(m, e.n, e.p, e.depth);
    return true;
}

// PlaneCollision.hpp:90-285
template <class T>
inline bool hull_plane(const ShapeP<T> &hs, const TF<T> &ta, const ShapeP<T> &pl, const TF<T> &tp, Manifold<T> &m) {
    const Hull<T> &h = *hs.hull;
    if (h.v.empty()) return false;
    V3<T> wn;
    T wd;
    plane_world(pl, tp, wn, wd);
    const V3<T> ln = ta.inv_vec(wn);
    const T ld = wd - dot(ta.pos, wn);
    const T eps = T(1e-4f);
    const V3<T> deep = h.v[hull_support_index(h, -ln)];
    if (dot(deep, ln) - ld > eps) return false;
    V3<T> tp_[32];
    T td[32];
    int tc = 0;
    for (size_t i = 0; i < h.v.size(); ++i) {
        const T dist = dot(h.v[i], ln) - ld;
        if (dist <= eps && tc < 32) { tp_[tc] = h.v[i]; td[tc] = -dist; ++tc; }
    }
    if (tc == 0) return false;
    m.normal = -wn;
    m.n = 0;
    if (tc <= 4) {
        for (int i = 0; i < tc; ++i) { m.pos[i] = ta.point(tp_[i]); m.depth[i] = td[i]; }
        m.n = tc;
        return true;
    }
    const T low = T(-1e18f);
    int idx[4] = {-1, -1, -1, -1};
    T best = low;
    for (int i = 0; i < tc; ++i) if (td[i] > best) { best = td[i]; idx[0] = i; }
    best = low;
    for (int i = 0; i < tc; ++i) {
        if (i == idx[0]) continue;
        const V3<T> d = tp_[i] - tp_[idx[0]];
        const T x = dot(d, d);
        if (x > best) { best = x; idx[1] = i; }
    }
    best = low;
    {
        const V3<T> e = tp_[idx[1]] - tp_[idx[0]];
        for (int i = 0; i < tc; ++i) {
            if (i == idx[0] || i == idx[1]) continue;
            const V3<T> c = cross(e, tp_[i] - tp_[idx[0]]);
            const T x = dot(c, c);
            if (x > best) { best = x; idx[2] = i; }
        }
    }
    best = low;
    {
        const V3<T> p0 = tp_[idx[0]], p1 = tp_[idx[1]], p2 = tp_[idx[2]];
        const V3<T> e01 = p1 - p0, e12 = p2 - p1, e20 = p0 - p2;
        for (int i = 0; i < tc; ++i) {
            if (i == idx[0] || i == idx[1] || i == idx[2]) continue;
            const V3<T> &x = tp_[i];
            const T a0 = dot(cross(e01, x - p0), ln);
            const T a1 = dot(cross(e12, x - p1), ln);
            const T a2 = dot(cross(e20, x - p2), ln);
            T sc = a0;
            if (a1 > sc) sc = a1;
            if (a2 > sc) sc = a2;
            if (sc > best) { best = sc; idx[3] = i; }
        }
    }
    const int emit = (idx[3] >= 0) ? 4 : 3;
    for (int i = 0; i < emit; ++i) { m.pos[i] = ta.point(tp_[idx[i]]); m.depth[i] = td[idx[i]]; }
    m.n = emit;
    return m.n > 0;
}

}  // namespace orc
