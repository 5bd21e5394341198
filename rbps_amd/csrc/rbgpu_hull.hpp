// Host-side convex hull preparation: the tables rbc::convex_hull_data_create builds
// (include/rbc/shapes/ConvexHull.hpp:149-420) and the unit-density mass properties of
// ConvexHull.hpp:575-745, in double precision (the reference's scalar).
//
// Tables (all indices local to the hull):
//   polygons   coplanar triangles merged into one CCW loop each: triangles whose unit
//              normals have a dot product above 0.99995 form a group (transitively);
//              groups in order of their lowest triangle; the loop is the group's boundary
//              (directed edges whose reverse is not in the group) chained from its first
//              boundary edge; the normal is the normalised sum of the group's triangle
//              normals.  Groups with an empty boundary, a loop shorter than 3 or a zero
//              normal sum are dropped.
//   v2p        polygons touching each vertex, in polygon order (first appearance)
//   v2n        polygon-edge neighbours of each vertex (first appearance over the loops)
//   p2p        polygons sharing an edge with each polygon, in polygon order
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace rbh {

struct D3 {
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline D3 scale_inv(D3 a, double s) { const double inv = 1.0 / s; return a * inv; }   // vec3.hpp:56-60
inline D3 unit(D3 v) {                                                                  // vec3.hpp:196-203
    const double l2 = dot(v, v);
    if (l2 < 1e-9) return {0.0, 0.0, 0.0};
    return v * (1.0 / std::sqrt(l2));
}

struct HullTables {
    std::vector<D3> v, pn;
    std::vector<uint32_t> poff, pidx, v2p_o, v2p_i, v2n_o, v2n_i, p2p_o, p2p_i;
    bool adjacency() const { return !pn.empty() && !v2p_i.empty() && !v2n_i.empty() && !p2p_i.empty(); }
};

// CSR form of per-item lists.
inline void to_csr(const std::vector<std::vector<uint32_t>> &lists, std::vector<uint32_t> &off, std::vector<uint32_t> &idx) {
    off.resize(lists.size() + 1);
    off[0] = 0;
    idx.clear();
    for (size_t i = 0; i < lists.size(); ++i) {
        idx.insert(idx.end(), lists[i].begin(), lists[i].end());
        off[i + 1] = (uint32_t)idx.size();
    }
}

inline void push_once(std::vector<uint32_t> &l, uint32_t x) {
    for (uint32_t y : l)
        if (y == x) return;
    l.push_back(x);
}

inline HullTables prepare_hull(const double *verts, uint32_t nv, const uint32_t *tri, uint32_t nt) {
    HullTables h;
    h.v.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) h.v[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    if (!tri || nt == 0) return h;
    std::vector<D3> tn(nt);
    for (uint32_t f = 0; f < nt; ++f) {
        const D3 a = h.v[tri[3 * f]], b = h.v[tri[3 * f + 1]], c = h.v[tri[3 * f + 2]];
        tn[f] = unit(cross(b - a, c - a));
    }
    // connected components of the "nearly parallel normals" relation
    std::vector<uint32_t> comp(nt, UINT32_MAX), stack;
    uint32_t ncomp = 0;
    for (uint32_t s = 0; s < nt; ++s) {
        if (comp[s] != UINT32_MAX) continue;
        comp[s] = ncomp;
        stack.assign(1, s);
        while (!stack.empty()) {
            const uint32_t f = stack.back();
            stack.pop_back();
            for (uint32_t g = 0; g < nt; ++g)
                if (comp[g] == UINT32_MAX && dot(tn[f], tn[g]) > 0.99995) {
                    comp[g] = ncomp;
                    stack.push_back(g);
                }
        }
        ++ncomp;
    }
    // components are numbered by their lowest triangle, so iterating them in number order
    // visits groups in order of first triangle
    std::vector<std::vector<uint32_t>> groups(ncomp);
    for (uint32_t f = 0; f < nt; ++f) groups[comp[f]].push_back(f);
    h.poff.push_back(0);
    for (const auto &g : groups) {
        std::vector<uint32_t> ea, eb;   // directed edges of the group, triangle order
        for (uint32_t f : g)
            for (int k = 0; k < 3; ++k) {
                ea.push_back(tri[3 * f + k]);
                eb.push_back(tri[3 * f + (k + 1) % 3]);
            }
        std::vector<uint32_t> ba, bb;   // boundary: the reverse edge is absent
        for (size_t i = 0; i < ea.size(); ++i) {
            bool interior = false;
            for (size_t j = 0; j < ea.size() && !interior; ++j) interior = j != i && ea[j] == eb[i] && eb[j] == ea[i];
            if (!interior) { ba.push_back(ea[i]); bb.push_back(eb[i]); }
        }
        if (ba.empty()) continue;
        std::vector<char> taken(ba.size(), 0);
        std::vector<uint32_t> loop{ba[0]};
        taken[0] = 1;
        uint32_t at = bb[0];
        for (size_t n = 1; n < ba.size(); ++n) {
            size_t k = 0;
            for (; k < ba.size(); ++k)
                if (!taken[k] && ba[k] == at) break;
            if (k == ba.size()) break;
            taken[k] = 1;
            loop.push_back(ba[k]);
            at = bb[k];
        }
        if (loop.size() < 3) continue;
        D3 s{0.0, 0.0, 0.0};
        for (uint32_t f : g) s = s + tn[f];
        const double sl = std::sqrt(dot(s, s));
        if (sl < 1e-9) continue;
        h.pn.push_back(scale_inv(s, sl));
        h.pidx.insert(h.pidx.end(), loop.begin(), loop.end());
        h.poff.push_back((uint32_t)h.pidx.size());
    }
    const uint32_t P = (uint32_t)h.pn.size();
    if (!P) { h.poff.clear(); return h; }
    std::vector<std::vector<uint32_t>> v2p(nv), v2n(nv), p2p(P);
    for (uint32_t p = 0; p < P; ++p) {
        const uint32_t s = h.poff[p], n = h.poff[p + 1] - s;
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t a = h.pidx[s + i], b = h.pidx[s + (i + 1) % n];
            push_once(v2p[a], p);
            push_once(v2n[a], b);
            push_once(v2n[b], a);
        }
    }
    for (uint32_t p = 0; p < P; ++p) {
        const uint32_t sp = h.poff[p], np = h.poff[p + 1] - sp;
        for (uint32_t q = 0; q < P; ++q) {
            if (q == p) continue;
            const uint32_t sq = h.poff[q], nq = h.poff[q + 1] - sq;
            bool shares = false;
            for (uint32_t i = 0; i < np && !shares; ++i) {
                const uint32_t a = h.pidx[sp + i], b = h.pidx[sp + (i + 1) % np];
                for (uint32_t j = 0; j < nq && !shares; ++j) shares = h.pidx[sq + j] == b && h.pidx[sq + (j + 1) % nq] == a;
            }
            if (shares) p2p[p].push_back(q);
        }
    }
    to_csr(v2p, h.v2p_o, h.v2p_i);
    to_csr(v2n, h.v2n_o, h.v2n_i);
    to_csr(p2p, h.p2p_o, h.p2p_i);
    return h;
}

// Volume and inertia (xx yy zz xy xz yz) at unit density about the centroid.
inline void hull_mass(const HullTables &h, const uint32_t *tri, uint32_t nt, double &vol, double I[6]) {
    vol = 0.0;
    for (int k = 0; k < 6; ++k) I[k] = 0.0;
    if (h.v.empty()) return;
    if (!tri || nt == 0) {   // local AABB as a box
        D3 mn{1e30f, 1e30f, 1e30f}, mx{-1e30f, -1e30f, -1e30f};
        for (const D3 &p : h.v) {
            mn = {p.x < mn.x ? p.x : mn.x, p.y < mn.y ? p.y : mn.y, p.z < mn.z ? p.z : mn.z};
            mx = {mx.x < p.x ? p.x : mx.x, mx.y < p.y ? p.y : mx.y, mx.z < p.z ? p.z : mx.z};
        }
        const D3 e = mx - mn;
        vol = e.x * e.y * e.z;
        I[0] = (1.0 / 12.0) * vol * (e.y * e.y + e.z * e.z);
        I[1] = (1.0 / 12.0) * vol * (e.x * e.x + e.z * e.z);
        I[2] = (1.0 / 12.0) * vol * (e.x * e.x + e.y * e.y);
        return;
    }
    double v6 = 0.0, xx = 0.0, yy = 0.0, zz = 0.0, xy = 0.0, xz = 0.0, yz = 0.0;
    D3 m1{0.0, 0.0, 0.0};
    for (uint32_t f = 0; f < nt; ++f) {
        const D3 A = h.v[tri[3 * f]], B = h.v[tri[3 * f + 1]], C = h.v[tri[3 * f + 2]];
        const double det = dot(A, cross(B, C));
        v6 += det;
        m1 = m1 + (A + B + C) * ((det / 6.0) * 0.25);
        const double s60 = det / 60.0, s120 = det / 120.0;
        const double sxx = s60 * (A.x * A.x + B.x * B.x + C.x * C.x + A.x * B.x + A.x * C.x + B.x * C.x);
        const double syy = s60 * (A.y * A.y + B.y * B.y + C.y * C.y + A.y * B.y + A.y * C.y + B.y * C.y);
        const double szz = s60 * (A.z * A.z + B.z * B.z + C.z * C.z + A.z * B.z + A.z * C.z + B.z * C.z);
        auto mixed = [&](double ax, double ay, double bx, double by, double cx, double cy) {
            return s120 * (2.0 * (ax * ay + bx * by + cx * cy) + ax * by + ay * bx + ax * cy + ay * cx + bx * cy + by * cx);
        };
        xx += syy + szz;
        yy += sxx + szz;
        zz += sxx + syy;
        xy -= mixed(A.x, A.y, B.x, B.y, C.x, C.y);
        xz -= mixed(A.x, A.z, B.x, B.z, C.x, C.z);
        yz -= mixed(A.y, A.z, B.y, B.z, C.y, C.z);
    }
    vol = v6 / 6.0;
    if (std::fabs(vol) < 1e-9) return;   // degenerate: zero tensor
    const D3 c = scale_inv(m1, vol);
    I[0] = xx - vol * (c.y * c.y + c.z * c.z);
    I[1] = yy - vol * (c.x * c.x + c.z * c.z);
    I[2] = zz - vol * (c.x * c.x + c.y * c.y);
    I[3] = xy + vol * c.x * c.y;
    I[4] = xz + vol * c.x * c.z;
    I[5] = yz + vol * c.y * c.z;
}

}  // namespace rbh
