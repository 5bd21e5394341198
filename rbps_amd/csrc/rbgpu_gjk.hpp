// Convex hulls on CDNA4: GJK, EPA and the hull contact manifolds, one 64-lane
// workgroup (one wavefront) per pair.
//
// Work split inside the wave:
//   * support mappings: each lane holds up to four hull vertices in registers (vertex
//     lane + 64k); a support is a lane-local scan plus a butterfly arg-max with the
//     lowest vertex index winning ties, which is the reference's first-maximum scan
//     (ConvexHull.hpp:457-473);
//   * GJK's simplex logic is uniform scalar code every lane executes alike;
//   * EPA keeps its polytope in LDS (vertices: w and the A-side witness; faces: normal,
//     distance, vertex indices, obsolete flag).  The closest face is a parallel arg-min
//     (lowest face index on ties, like the reference's first-minimum scan); the visible
//     faces are tested in parallel and listed in face order; the horizon is computed in
//     parallel from that list: the reference's add_edge (EPA.cpp:102-114) toggles an
//     undirected edge in an ordered list, so the final list holds the edges that occur an
//     odd number of times, each in the direction and at the place of its last occurrence;
//     new faces are built one per lane and packed in horizon order;
//   * the hull-hull manifold (ContactManifoldGenerator.cpp:176-379): supports and the
//     edge-pair search are wave-parallel; the polygon clipping and the 4-point reduction
//     run on lane 0 over LDS buffers.
// Every floating-point expression keeps the reference's operand order, so the FP32
// oracle's restatement reproduces these results bit for bit.
#pragma once
#include "rbgpu_kernels.hpp"

namespace rb {

constexpr int HULL_MAX_VERTS = 256;
constexpr int HULL_VPL = HULL_MAX_VERTS / 64;     // vertices per lane (kernels for hulls of up to 64 use 1)
constexpr int HULL_MAX_POLY = 64;                 // vertices of one polygon, neighbours of one polygon
constexpr int EPA_FACES = 256, EPA_VERTS = 128, EPA_ITERS = 128, GJK_ITERS = 128;
constexpr int GJK_WARM_ITERS = 16;   // a warm-started GJK not done by then falls back to the plain seed
constexpr int GJK_CYCLE_FROM = 8;    // first iteration whose loop state the cycle detection saves
constexpr int EPA_SEQ = 3 * EPA_FACES;            // directed edges of the visible faces
constexpr int HCLIP_CAP = 2 * HULL_MAX_POLY + 8;   // Sutherland-Hodgman output bound (n + planes)
constexpr int HP_MAX_POINTS = 32;                 // hull-plane candidates (PlaneCollision.hpp:118)

enum { GJK_VALID = 0, GJK_INSIDE = 1, GJK_FAILED = 2 };
enum { EPA_VALID = 0, EPA_FAILED = 1, EPA_OUT_OF_FACES = 2, EPA_OUT_OF_VERTICES = 3, EPA_NONLINEAR = 4 };
// Not a reference status: a polytope or clip buffer smaller than the reference's limits ran
// out; the pair is run again with the full ones (never stored).
constexpr int EPA_DEFER = 14;

__device__ __forceinline__ int lane64() { return (int)(threadIdx.x & 63); }

// ------------------------------------------------------------------ wave reductions
// Arg-max over (value, index): larger value, then lower index.
__device__ __forceinline__ void wave_argmax(float &d, int &i) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float od = __shfl_xor(d, o, 64);
        const int oi = __shfl_xor(i, o, 64);
        if (od > d || (od == d && oi < i)) { d = od; i = oi; }
    }
}
// Arg-min over (value, index): smaller value, then lower index.
__device__ __forceinline__ void wave_argmin(float &d, int &i) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float od = __shfl_xor(d, o, 64);
        const int oi = __shfl_xor(i, o, 64);
        if (od < d || (od == d && oi < i)) { d = od; i = oi; }
    }
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ uint32_t dpp_row_max(uint32_t v);
__device__ __forceinline__ uint32_t dpp_row_min(uint32_t v);
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v);

// The lanes that hold one pair: LPP = 64 (a pair per wavefront, one-wave workgroups) or 16
// (a pair per DPP row, four per wavefront).  Ballots come back shifted to the group's first
// lane; reductions stay inside the group.  Every lane of a group must be active at each call
// (the group's control flow is uniform); other groups of the wavefront may be inactive.
template <int LPP> struct Grp {
    static_assert(LPP == 16 || LPP == 64, "a pair spans a DPP row or a wavefront");
    __device__ static __forceinline__ int lane() { return lane64() & (LPP - 1); }
    __device__ static __forceinline__ uint64_t ballot(bool v) {
        const uint64_t m = __ballot(v);
        if constexpr (LPP == 64) return m;
        else return (m >> (lane64() & ~(LPP - 1))) & ((1ull << LPP) - 1ull);
    }
    // set bits of a group ballot below this lane
    __device__ static __forceinline__ uint32_t rank(uint64_t m) {
        if constexpr (LPP == 64) return lane_rank(m);
        else return (uint32_t)__popcll(m & ((1ull << lane()) - 1ull));
    }
    __device__ static __forceinline__ uint32_t min_u32(uint32_t v) {
        if constexpr (LPP == 64) return wave_min_u32(v);
        else return dpp_row_min(v);
    }
    __device__ static __forceinline__ uint32_t sum(uint32_t v) {
#pragma unroll
        for (int o = LPP / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        return v;
    }
    // arg-max over (value, index): larger value, then lower index
    __device__ static __forceinline__ void argmax(float &d, int &i) {
#pragma unroll
        for (int o = LPP / 2; o > 0; o >>= 1) {
            const float od = __shfl_xor(d, o, 64);
            const int oi = __shfl_xor(i, o, 64);
            if (od > d || (od == d && oi < i)) { d = od; i = oi; }
        }
    }
    // Orders the group's LDS accesses: a wavefront runs its LDS instructions in order, so
    // within one wavefront a compiler barrier suffices (a workgroup barrier there would need
    // every group of the workgroup to arrive).
    __device__ static __forceinline__ void sync() {
        if constexpr (LPP == 64) {
            __syncthreads();
        } else {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
};

// ------------------------------------------------------------------ shapes in a wave
template <int VPL, int LPP = 64> struct WaveHull {   // this lane's vertices r + LPP k (r: lane within the pair's LPP lanes)
    float x[VPL], y[VPL], z[VPL];
    int nv;
    uint32_t v0;
};
// Lanes per pair (LPP): 64 = one pair per wavefront; 16 = four pairs per wavefront, one per
// DPP row (hulls of at most 16 * VPL vertices).
template <int LPP> __device__ __forceinline__ int pair_lane() { return lane64() & (LPP - 1); }

template <int VPL, int LPP> __device__ __forceinline__ void load_wave_hull(WaveHull<VPL, LPP> &w, const HullSet &hs, uint32_t h) {
    const HullMeta m = hs.meta[h];
    w.nv = (int)m.nv;
    w.v0 = m.v0;
#pragma unroll
    for (int k = 0; k < VPL; ++k) {
        const int i = pair_lane<LPP>() + LPP * k;
        float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (i < w.nv) p = hs.v[m.v0 + i];
        w.x[k] = p.x; w.y[k] = p.y; w.z[k] = p.z;
    }
}
// Cross-lane max / min: DPP within each 16-lane row (quad swaps, half-row and row
// mirrors), then the four row results through readlane.  Every lane must be active.
__device__ __forceinline__ uint32_t dpp_row_max(uint32_t v) {
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));    // quad_perm [1,0,3,2]
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));    // quad_perm [2,3,0,1]
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));   // row_half_mirror
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false));   // row_mirror
    return v;
}
__device__ __forceinline__ uint32_t dpp_row_min(uint32_t v) {
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false));
    return v;
}
__device__ __forceinline__ uint32_t rl(uint32_t v, int lane) { return (uint32_t)__builtin_amdgcn_readlane((int)v, lane); }
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    v = dpp_row_max(v);
    return max(max(rl(v, 0), rl(v, 16)), max(rl(v, 32), rl(v, 48)));
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    v = dpp_row_min(v);
    return min(min(rl(v, 0), rl(v, 16)), min(rl(v, 32), rl(v, 48)));
}
// Order-preserving map of a non-NaN float to u32 (-0 maps with +0, as the comparisons treat them).
__device__ __forceinline__ uint32_t ord_f(float x) {
    const uint32_t u = __float_as_uint(x + 0.0f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Index of the first vertex with the largest dot(v, d) (uniform result); its local
// coordinates in *vout.  Lane-local scan in index order, wave max of the mapped value,
// then the lowest index among the lanes holding it.  A NaN at vertex 0 is never beaten
// in the sequential scan; NaN elsewhere never wins.
template <int VPL, int LPP> __device__ __forceinline__ int wave_support_index(const WaveHull<VPL, LPP> &w, V3 d, V3 *vout = nullptr) {
    if constexpr (LPP == 16) {
        // one pair per 16-lane row: vertex r + 16 k on lane r of the row; a lane keeps its
        // first maximum, the row takes the lowest index among the lanes holding the row
        // maximum (dpp_row_max / dpp_row_min leave the row result in every lane of the row)
        const int r = pair_lane<16>(), row = lane64() >> 4;
        uint32_t best = 0u;
        int bk = 0;
        bool nan0 = false;
#pragma unroll
        for (int k = 0; k < VPL; ++k) {
            const int i = r + 16 * k;
            if (i < w.nv) {
                float x = w.x[k] * d.x + w.y[k] * d.y + w.z[k] * d.z;
                if (x != x) {
                    nan0 |= i == 0;
                    x = -__builtin_inff();
                }
                const uint32_t o = ord_f(x);
                if (o > best) { best = o; bk = k; }
            }
        }
        const uint32_t m = dpp_row_max(best);
        int idx = (int)dpp_row_min(best == m ? (uint32_t)(r + 16 * bk) : 0xffffffffu);
        if ((__ballot(nan0) >> (16 * row)) & 1ull) idx = 0;
        if (vout) {
            const int src = (row << 4) | (idx & 15), slot = idx >> 4;
            float x = w.x[0], y = w.y[0], z = w.z[0];
#pragma unroll
            for (int k = 1; k < VPL; ++k)
                if (slot == k) { x = w.x[k]; y = w.y[k]; z = w.z[k]; }
            *vout = V3{__shfl(x, src, 64), __shfl(y, src, 64), __shfl(z, src, 64)};
        }
        return idx;
    }
    const int lane = lane64();
    uint32_t best = 0u;
    int bk = 0;
    bool nan0 = false;
#pragma unroll
    for (int k = 0; k < VPL; ++k) {
        const int i = lane + 64 * k;
        if (i < w.nv) {
            float x = w.x[k] * d.x + w.y[k] * d.y + w.z[k] * d.z;
            if (x != x) {
                nan0 |= i == 0;
                x = -__builtin_inff();
            }
            const uint32_t o = ord_f(x);
            if (o > best) { best = o; bk = k; }
        }
    }
    const uint32_t m = wave_max_u32(best);
    const bool cand = best == m;   // lanes without vertices hold 0, below any mapped value
    int idx;
    if (w.nv <= 64) {
        idx = __ffsll((long long)__ballot(cand)) - 1;
    } else {
        idx = (int)wave_min_u32(cand ? (uint32_t)(lane + 64 * bk) : 0xffffffffu);
    }
    if (rl((uint32_t)nan0, 0)) idx = 0;
    if (vout) {
        const int src = idx & 63, slot = idx >> 6;
        float x = w.x[0], y = w.y[0], z = w.z[0];
#pragma unroll
        for (int k = 1; k < VPL; ++k)
            if (slot == k) { x = w.x[k]; y = w.y[k]; z = w.z[k]; }
        *vout = V3{__uint_as_float(rl(__float_as_uint(x), src)), __uint_as_float(rl(__float_as_uint(y), src)),
                   __uint_as_float(rl(__float_as_uint(z), src))};
    }
    return idx;
}

template <int VPL, int LPP = 64> struct Side {
    static constexpr int kLanes = LPP;   // lanes that hold one pair
    int kind;
    float a, b, c;   // sphere r | box half extents | capsule half_height, radius
    TF tf;
    WaveHull<VPL, LPP> wh;
};
template <int VPL, int LPP> __device__ __forceinline__ V3 hull_vertex(const HullSet &hs, const WaveHull<VPL, LPP> &w, int i) {
    const float4 p = hs.v[w.v0 + i];
    return V3{p.x, p.y, p.z};
}
// Local support of a sphere, box or capsule (Sphere.hpp:48-51, Box.hpp:49-55, Capsule.hpp:72-80).
__device__ __forceinline__ V3 prim_support(int kind, float a, float b, float c, V3 d) {
    if (kind == SH_SPHERE) return normalize(d) * a;
    if (kind == SH_BOX) return V3{d.x > 0.0f ? a : -a, d.y > 0.0f ? b : -b, d.z > 0.0f ? c : -c};
    const V3 e{0.0f, d.y >= 0.0f ? a : -a, 0.0f};   // capsule
    const float l = len(d);
    if (l < EPS) return e;
    return e + vdiv(d, l) * b;
}
// Local support (and ConvexHull.hpp:457-473 for hulls, over the wave's lanes).
template <int VPL, int LPP> __device__ __forceinline__ V3 local_support(const Side<VPL, LPP> &s, const HullSet &hs, V3 d) {
    if (s.kind == SH_HULL) {
        V3 v;
        wave_support_index(s.wh, d, &v);
        return v;
    }
    return prim_support(s.kind, s.a, s.b, s.c, d);
}
// World support of one side (MinkowskiDiff.hpp:52-63).
template <int VPL, int LPP> __device__ __forceinline__ V3 side_support(const Side<VPL, LPP> &s, const HullSet &hs, V3 dir) {
    return tf_point(s.tf, local_support(s, hs, tf_inv_vec(s.tf, dir)));
}

// A side of a pair that one lane holds by itself: the shape and, for a hull, where its
// vertices start in the world's vertex array.
struct LaneSide {
    static constexpr int kLanes = 1;
    int kind;
    float a, b, c;
    TF tf;
    uint32_t v0;
    int nv;
};
// ConvexHull.hpp:457-473 on one lane: the first vertex (index order) with the largest
// dot(v, d).  Strict comparisons give the first maximum, treat -0 and +0 alike, never let a
// NaN win and never let anything beat a NaN at vertex 0 -- the reference's scan.  Loads
// run four vertices ahead of the ordered comparisons.
__device__ __forceinline__ V3 lane_hull_support(const HullSet &hs, uint32_t v0, int nv, V3 d) {
    const float4 p = hs.v[v0];
    float best = p.x * d.x + p.y * d.y + p.z * d.z;
    V3 bv{p.x, p.y, p.z};
    for (int i = 1; i < nv; i += 4) {
        float4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = i + u < nv ? hs.v[v0 + i + u] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float x = q[u].x * d.x + q[u].y * d.y + q[u].z * d.z;
            if (i + u < nv && x > best) { best = x; bv = V3{q[u].x, q[u].y, q[u].z}; }
        }
    }
    return bv;
}
__device__ __forceinline__ V3 side_support(const LaneSide &s, const HullSet &hs, V3 dir) {
    const V3 d = tf_inv_vec(s.tf, dir);
    return tf_point(s.tf, s.kind == SH_HULL ? lane_hull_support(hs, s.v0, s.nv, d) : prim_support(s.kind, s.a, s.b, s.c, d));
}
__device__ __forceinline__ void load_lane_side(LaneSide &s, const Colliders &C, const HullSet &hs, uint32_t c, const TF &tf) {
    const int kind = C.meta[c].x;
    const float4 d = C.dims[c];
    s.kind = kind;
    s.a = d.x; s.b = d.y; s.c = d.z;
    s.tf = tf;
    s.v0 = 0;
    s.nv = 0;
    if (kind == SH_HULL) {
        const HullMeta m = hs.meta[(uint32_t)C.hull[c]];
        s.v0 = m.v0;
        s.nv = (int)m.nv;
    }
}

struct SV {
    V3 w0, w;   // A-side support, Minkowski-difference point (the B-side support is not kept)
};
template <class SideT> __device__ __forceinline__ SV md_sample(const SideT &A, const SideT &Bs, const HullSet &hs, V3 d) {
    SV s;
    s.w0 = side_support(A, hs, d);
    s.w = s.w0 - side_support(Bs, hs, -d);
    return s;
}

// ------------------------------------------------------------------ GJK (GJK.cpp:42-96)
struct Simplex {
    SV v[4];   // v[0] newest
    int rank;
    V3 dir;    // search direction at exit (a separating direction when GJK returns valid)
};
// Vertex i of a simplex by value (selects, so that the simplex stays in registers).
__device__ __forceinline__ SV sx_vertex(const Simplex &s, int i) {
    SV r = s.v[0];
    if (i == 1) r = s.v[1];
    if (i == 2) r = s.v[2];
    if (i == 3) r = s.v[3];
    return r;
}
// Field-wise select of two simplex vertices (keeps the simplex in registers: conditional
// whole-struct copies were lowered to stack copies).
__device__ __forceinline__ SV sv_sel(bool c, const SV &a, const SV &b) {
    SV r;
    r.w0 = V3{c ? a.w0.x : b.w0.x, c ? a.w0.y : b.w0.y, c ? a.w0.z : b.w0.z};
    r.w = V3{c ? a.w.x : b.w.x, c ? a.w.y : b.w.y, c ? a.w.z : b.w.z};
    return r;
}
__device__ __forceinline__ void sx_push(Simplex &s, const SV &n) {   // DoSimplex.cpp:5-17
    const int r = s.rank;
    const SV v0 = s.v[0], v1 = s.v[1], v2 = s.v[2], v3 = s.v[3];
    s.v[3] = sv_sel(r >= 3, v2, v3);
    s.v[2] = sv_sel(r >= 2, v1, v2);
    s.v[1] = sv_sel(r >= 1, v0, v1);
    s.v[0] = n;
    if (r < 4) s.rank = r + 1;
}
__device__ __forceinline__ V3 triple(V3 a, V3 b, V3 c) { return cross(cross(a, b), c); }

__device__ __forceinline__ void sx_line(Simplex &s, V3 &dir) {   // Line.cpp:6-48
    const V3 a = s.v[0].w, b = s.v[1].w;
    const V3 ao = -a, ab = b - a;
    if (dot(ab, ao) >= 0.0f) {
        s.rank = 2;
        dir = triple(ab, ao, ab);
        if (len2(dir) < 1e-6f) {
            const float ax = absf(ab.x), ay = absf(ab.y), az = absf(ab.z);
            const V3 axis = (ax <= ay && ax <= az) ? V3{1.0f, 0.0f, 0.0f}
                                                   : (ay <= az ? V3{0.0f, 1.0f, 0.0f} : V3{0.0f, 0.0f, 1.0f});
            dir = cross(ab, axis);
            if (len2(dir) < 1e-6f) dir = cross(ab, V3{0.0f, 0.0f, 1.0f});
        }
    } else {
        s.rank = 1;
        dir = ao;
    }
}
__device__ __forceinline__ void sx_triangle(Simplex &s, V3 &dir) {   // Triangle.cpp:6-80
    const V3 a = s.v[0].w, b = s.v[1].w, c = s.v[2].w;
    const V3 ao = -a, ab = b - a, ac = c - a, abc = cross(ab, ac);
    if (dot(cross(abc, ac), ao) >= 0.0f) {
        if (dot(ac, ao) >= 0.0f) { s.v[1] = s.v[2]; s.rank = 2; dir = triple(ac, ao, ac); }
        else if (dot(ab, ao) >= 0.0f) { s.rank = 2; dir = triple(ab, ao, ab); }
        else { s.rank = 1; dir = ao; }
    } else if (dot(cross(ab, abc), ao) >= 0.0f) {
        if (dot(ab, ao) >= 0.0f) { s.rank = 2; dir = triple(ab, ao, ab); }
        else { s.rank = 1; dir = ao; }
    } else if (dot(abc, ao) >= 0.0f) {
        s.rank = 3;
        dir = abc;
    } else {
        const SV t = s.v[1];
        s.v[1] = s.v[2];
        s.v[2] = t;
        s.rank = 3;
        dir = -abc;
    }
}
// One face (A, B, C) of the tetrahedron with outward normal n (Tetrahedron.cpp:55-165).
__device__ __forceinline__ void sx_tet_face(Simplex &s, V3 &dir, const SV &A, const SV &Bv, const SV &Cv, V3 ab, V3 ac,
                                            V3 n, V3 ao) {
    if (dot(cross(n, ac), ao) >= 0.0f) {
        if (dot(ac, ao) >= 0.0f) { s.v[0] = A; s.v[1] = Cv; s.rank = 2; dir = triple(ac, ao, ac); }
        else if (dot(ab, ao) >= 0.0f) { s.v[0] = A; s.v[1] = Bv; s.rank = 2; dir = triple(ab, ao, ab); }
        else { s.v[0] = A; s.rank = 1; dir = ao; }
    } else if (dot(cross(ab, n), ao) >= 0.0f) {
        if (dot(ab, ao) >= 0.0f) { s.v[0] = A; s.v[1] = Bv; s.rank = 2; dir = triple(ab, ao, ab); }
        else { s.v[0] = A; s.rank = 1; dir = ao; }
    } else {
        s.v[0] = A; s.v[1] = Bv; s.v[2] = Cv; s.rank = 3; dir = n;
    }
}
__device__ __forceinline__ bool sx_tetra(Simplex &s, V3 &dir) {   // Tetrahedron.cpp:6-217
    const SV va = s.v[0], vb = s.v[1], vc = s.v[2], vd = s.v[3];
    const V3 ao = -va.w, ab = vb.w - va.w, ac = vc.w - va.w, ad = vd.w - va.w;
    const V3 abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
    int pi = 0;
    if (dot(abc, ao) > 1e-10f) pi |= 1;
    if (dot(acd, ao) > 1e-10f) pi |= 2;
    if (dot(adb, ao) > 1e-10f) pi |= 4;
    switch (pi) {
        case 0: return true;
        case 1: sx_tet_face(s, dir, va, vb, vc, ab, ac, abc, ao); break;
        case 2: sx_tet_face(s, dir, va, vc, vd, ac, ad, acd, ao); break;
        case 4: sx_tet_face(s, dir, va, vd, vb, ad, ab, adb, ao); break;
        case 3:
            if (dot(ac, ao) >= 0.0f) { s.v[0] = va; s.v[1] = vc; s.rank = 2; dir = triple(ac, ao, ac); }
            else { s.v[0] = va; s.rank = 1; dir = ao; }
            break;
        case 5:
            if (dot(ab, ao) >= 0.0f) { s.v[0] = va; s.v[1] = vb; s.rank = 2; dir = triple(ab, ao, ab); }
            else { s.v[0] = va; s.rank = 1; dir = ao; }
            break;
        case 6:
            if (dot(ad, ao) >= 0.0f) { s.v[0] = va; s.v[1] = vd; s.rank = 2; dir = triple(ad, ao, ad); }
            else { s.v[0] = va; s.rank = 1; dir = ao; }
            break;
        default: s.v[0] = va; s.rank = 1; dir = ao; break;
    }
    return false;
}
// Exact cycle detection for GJK (Brent 1980).  An iteration's outcome depends only on the
// loop state -- rank, the simplex points w below the rank and the search direction (the
// A-side points w0 ride along and decide nothing) -- so a state equal to an earlier one
// repeats forever and the loop can only end at its cap, GJK_FAILED.  One earlier state is
// kept, re-saved at the 1st, 2nd, 4th, 8th, ... call (calls from iteration GJK_CYCLE_FROM on);
// a match ends the loop with the status and
// iteration count the full run reports.  (In FP32, GJK on touching hulls cycles with
// periods of 4-15 iterations from about its tenth: such pairs used the full 128.)
// The state is 16 words: 0 rank, 1-12 the points, 13-15 the direction.  Where a pair spans
// 16 or 64 lanes (its state uniform over them) lane r keeps word r & 15 and the match is a
// ballot; a pair on one lane keeps all 16 words.
// Word r of the state, picked with masks rather than selects or an index: a run-time pick
// among the simplex's fields was lowered to a stack copy of the simplex.
__device__ __forceinline__ uint32_t gjk_state_word(const Simplex &s, V3 dir, int r, bool &used) {
    used = r == 0 || r >= 13 || (r - 1) / 3 < s.rank;
    const V3 a = s.v[0].w, b = s.v[1].w, c = s.v[2].w, d = s.v[3].w;
    uint32_t w = (uint32_t)s.rank & (0u - (uint32_t)(r == 0));
    const float f[15] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, d.x, d.y, d.z, dir.x, dir.y, dir.z};
#pragma unroll
    for (int q = 0; q < 15; ++q) w |= __float_as_uint(f[q]) & (0u - (uint32_t)(r == q + 1));
    return w;
}
template <int LANES> struct GjkCycle {
    uint32_t saved[16];
    int power = 1, lam = 0;
    __device__ __forceinline__ bool repeats(const Simplex &s, V3 dir) {
        bool same = lam > 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            bool used;
            const uint32_t w = gjk_state_word(s, dir, r, used);
            same = same && (!used || w == saved[r]);
            if (power == lam || lam == 0) saved[r] = w;
        }
        if (same) return true;
        if (power == lam) { power *= 2; lam = 0; }
        ++lam;
        return false;
    }
};
template <> struct GjkCycle<16> {   // one pair per 16-lane row: lane r keeps word r
    uint32_t saved = 0u;
    int power = 1, lam = 0;
    __device__ __forceinline__ bool repeats(const Simplex &s, V3 dir) {
        const int r = lane64() & 15;
        bool used;
        const uint32_t w = gjk_state_word(s, dir, r, used);
        const uint32_t diff = (uint32_t)((__ballot(used && w != saved) >> (lane64() & ~15)) & 0xffffull);
        const bool same = lam > 0 && diff == 0u;
        if (power == lam || lam == 0) saved = w;
        if (same) return true;
        if (power == lam) { power *= 2; lam = 0; }
        ++lam;
        return false;
    }
};
template <> struct GjkCycle<64> {   // one pair per wavefront: lanes 0-15 keep the words
    uint32_t saved = 0u;
    int power = 1, lam = 0;
    __device__ __forceinline__ bool repeats(const Simplex &s, V3 dir) {
        const int r = lane64() & 15;
        bool used;
        const uint32_t w = gjk_state_word(s, dir, r, used);
        const bool same = lam > 0 && (__ballot(used && w != saved) & 0xffffull) == 0ull;
        if (power == lam || lam == 0) saved = w;
        if (same) return true;
        if (power == lam) { power *= 2; lam = 0; }
        ++lam;
        return false;
    }
};

template <class SideT> __device__ __forceinline__ int gjk(const SideT &A, const SideT &Bs, const HullSet &hs, V3 guess, Simplex &s,
                                                         int &iters, int max_iter = GJK_ITERS) {
    s.rank = 0;
    V3 dir = (len2(guess) > EPS) ? guess : V3{1.0f, 0.0f, 0.0f};
    const SV v0 = md_sample(A, Bs, hs, dir);
    sx_push(s, v0);
    dir = -v0.w;
    GjkCycle<SideT::kLanes> cyc;
    for (int it = 0; it < max_iter; ++it) {
        iters = it + 1;
        const SV nx = md_sample(A, Bs, hs, dir);
        if (dot(nx.w, dir) < 0.0f) {
            s.dir = dir;
            return GJK_VALID;
        }
        sx_push(s, nx);
        bool inside = false;
        if (s.rank == 2) sx_line(s, dir);
        else if (s.rank == 3) sx_triangle(s, dir);
        else inside = sx_tetra(s, dir);
        if (inside) return GJK_INSIDE;
        // the remaining iterations would only go round the cycle; the detection starts at
        // iteration GJK_CYCLE_FROM (a state that repeats from an earlier iteration keeps
        // repeating from there, and most pairs end before it)
        if (it >= GJK_CYCLE_FROM && cyc.repeats(s, dir)) {
            iters = max_iter;
            return GJK_FAILED;
        }
    }
    return GJK_FAILED;
}

// ------------------------------------------------------------------ EPA (EPA.cpp:141-378)
// Polytope of NV vertices and NF faces.  The reference's limits (EPA_VERTS, EPA_FACES) give
// its OUT_OF_VERTICES / OUT_OF_FACES statuses; a smaller polytope ends with EPA_DEFER where
// the reference's would still grow, and the pair is run again with the full one.  Up to that
// point the two runs are the same instructions on the same data.
template <int NV, int NF> struct EpaLdsT {
    static_assert(NF % 4 == 0, "the horizon scan reads the edge list four entries at a time");
    float wx[NV], wy[NV], wz[NV];     // polytope vertices (w)
    float ax[NV], ay[NV], az[NV];     // their A-side supports (w0)
    float nx[NF], ny[NF], nz[NF], nd[NF];
    uint32_t fv[NF];                  // v0 | v1 << 8 | v2 << 16 | obsolete << 24
    uint16_t vis[NF];                 // visible faces, face order
    alignas(8) uint16_t seq[3 * NF];  // their directed edges (a | b << 8); read four at a time
    uint16_t hor[3 * NF];             // horizon edges, in order
};
using EpaLds = EpaLdsT<EPA_VERTS, EPA_FACES>;
// Manifold buffers, overlaid on the EPA polytope once EPA is done: two clip buffers of HC
// points and the kept points' depths.
template <int HC> struct ClipLdsT {
    float px[2][HC], py[2][HC], pz[2][HC];
    float dep[HC];
};
using ClipLds = ClipLdsT<HCLIP_CAP>;
template <int NV, int NF, int HC> union PairLdsT {
    EpaLdsT<NV, NF> e;
    ClipLdsT<HC> c;
};
using PairLds = PairLdsT<EPA_VERTS, EPA_FACES, HCLIP_CAP>;

template <class L> __device__ __forceinline__ V3 lds_w(const L &l, int i) { return V3{l.wx[i], l.wy[i], l.wz[i]}; }
template <class L> __device__ __forceinline__ V3 lds_a(const L &l, int i) { return V3{l.ax[i], l.ay[i], l.az[i]}; }

// compute_face_normal (EPA.cpp:37-87) over the first nall polytope vertices.
template <class L> __device__ __forceinline__ bool epa_face_normal(const L &l, int i0, int i1, int i2, int nall, V3 &n_out,
                                                                    float &d_out) {
    const V3 a = lds_w(l, i0), b = lds_w(l, i1), c = lds_w(l, i2);
    V3 n = cross(b - a, c - a);
    const float nl = len(n);
    if (nl < EPS) return false;
    n = vdiv(n, nl);
    float d = dot(n, a);
    if (d < -1e-12f) {
        n = -n;
        d = -d;
    } else if (d <= 1e-12f) {
        // the first vertex off the plane decides (reads four vertices ahead, in order)
        bool fixed = false;
        for (int i0 = 0; i0 < nall && !fixed; i0 += 4) {
            float t[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) t[u] = i0 + u < nall ? dot(n, lds_w(l, i0 + u)) : 0.0f;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (fixed) break;
                if (t[u] > 1e-12f) { n = -n; d = 0.0f; fixed = true; }
                else if (t[u] < -1e-12f) { d = 0.0f; fixed = true; }
            }
        }
        if (!fixed) return false;
    }
    n_out = n;
    d_out = d;
    return true;
}

// Adds w to the polytope (EPA.cpp:337-374): faces visible from w become obsolete and
// the horizon of the visible region is stitched to w.  Returns EPA_OUT_OF_FACES when
// the face pool runs out (EPA_DEFER for a pool smaller than the reference's), else 0.
template <int LPP, int NV, int NF>
__device__ __forceinline__ int epa_expand(EpaLdsT<NV, NF> &L, const SV &w, int &nvtx, int &nf, int &nf_tot) {
    using G = Grp<LPP>;
    const int lane = G::lane();
    const int wi = nvtx;
    if (lane == 0) {
        L.wx[wi] = w.w.x; L.wy[wi] = w.w.y; L.wz[wi] = w.w.z;
        L.ax[wi] = w.w0.x; L.ay[wi] = w.w0.y; L.az[wi] = w.w0.z;
    }
    ++nvtx;
    // 4. visible faces, listed in face order: every face of this lane tested first (the reads
    // need not wait for the list's writes), then listed
    constexpr int FPL = (NF + LPP - 1) / LPP;   // faces per lane
    bool vf[FPL];
    uint32_t fvr[FPL];
#pragma unroll
    for (int u = 0; u < FPL; ++u) {
        const int f = u * LPP + lane;
        vf[u] = false;
        fvr[u] = 0u;
        if (f < nf) {
            const uint32_t fv = L.fv[f];
            fvr[u] = fv;
            if (!(fv >> 24)) {
                const V3 cen = ((lds_w(L, (int)(fv & 255u)) + lds_w(L, (int)((fv >> 8) & 255u))) +
                                lds_w(L, (int)((fv >> 16) & 255u))) * (1.0f / 3.0f);
                const V3 fn{L.nx[f], L.ny[f], L.nz[f]};
                vf[u] = dot(fn, w.w - cen) > 0.0f;
            }
        }
    }
    int nvis = 0;
#pragma unroll
    for (int u = 0; u < FPL; ++u) {
        if (u * LPP >= nf) break;   // uniform over the group
        const int f = u * LPP + lane;
        const bool v = vf[u];
        const uint64_t m = G::ballot(v);
        if (v) {
            const uint32_t fv = fvr[u];
            L.fv[f] = fv | (1u << 24);
            const int pos = nvis + (int)G::rank(m);
            L.vis[pos] = (uint16_t)f;
            // its directed edges (v0,v1) (v1,v2) (v2,v0), in visible-face order
            const uint32_t a = fv & 255u, b = (fv >> 8) & 255u, c = (fv >> 16) & 255u;
            L.seq[3 * pos] = (uint16_t)(a | (b << 8));
            L.seq[3 * pos + 1] = (uint16_t)(b | (c << 8));
            L.seq[3 * pos + 2] = (uint16_t)(c | (a << 8));
        }
        nvis += __popcll(m);
    }
    G::sync();
    const int ns = 3 * nvis;
    // horizon: odd occurrence count, kept at the last occurrence.  Every lane scans the whole
    // sequence (no early exit: the reads do not wait for one another)
    auto ukey = [](uint32_t e) {
        const uint32_t a = e & 255u, b = e >> 8;
        return a < b ? (a | (b << 8)) : (b | (a << 8));
    };
    int nh = 0;
    for (int base = 0; base < ns; base += LPP) {
        const int i = base + lane;
        bool keep = false;
        if (i < ns) {
            const uint32_t key = ukey(L.seq[i]);
            int later = 0, before = 0;
            const uint64_t *sq = reinterpret_cast<const uint64_t *>(L.seq);   // 3 NF entries: a multiple of four
#pragma unroll 4
            for (int j4 = 0; j4 < ns; j4 += 4) {
                const uint64_t q = sq[j4 >> 2];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int j = j4 + t;
                    const bool same = j < ns && ukey((uint32_t)(q >> (16 * t)) & 0xffffu) == key;
                    later += (int)(same & (j > i));
                    before += (int)(same & (j < i));
                }
            }
            keep = later == 0 && (before & 1) == 0;
        }
        const uint64_t m = G::ballot(keep);
        if (keep) L.hor[nh + (int)G::rank(m)] = L.seq[i];
        nh += __popcll(m);
    }
    G::sync();
    // A pool smaller than the reference's makes room when the new faces might not fit: the live
    // faces move down, in order (the closest-face ties and the visible-face lists depend only on
    // the live faces' order); the reference's own pool, which keeps obsolete faces, is counted
    // in nf_tot.
    if constexpr (NF < EPA_FACES) {
        if (nf + nh > NF) {   // uniform over the group
            float rx[FPL], ry[FPL], rz[FPL], rd[FPL];
            uint32_t rf[FPL];
            int rp[FPL];
            int live = 0;
#pragma unroll
            for (int u = 0; u < FPL; ++u) {
                const int f = u * LPP + lane;
                const bool lv = f < nf && !(L.fv[f] >> 24);
                const uint64_t m = G::ballot(lv);
                rp[u] = lv ? live + (int)G::rank(m) : -1;
                live += __popcll(m);
                rx[u] = lv ? L.nx[f] : 0.0f; ry[u] = lv ? L.ny[f] : 0.0f; rz[u] = lv ? L.nz[f] : 0.0f;
                rd[u] = lv ? L.nd[f] : 0.0f; rf[u] = lv ? L.fv[f] : 0u;
            }
            G::sync();
#pragma unroll
            for (int u = 0; u < FPL; ++u)
                if (rp[u] >= 0) {
                    L.nx[rp[u]] = rx[u]; L.ny[rp[u]] = ry[u]; L.nz[rp[u]] = rz[u]; L.nd[rp[u]] = rd[u];
                    L.fv[rp[u]] = rf[u];
                }
            nf = live;
            G::sync();
        }
    }
    // stitch: one candidate face per horizon edge, packed in horizon order
    int made_before_last = 0, made = 0;
    for (int base = 0; base < nh; base += LPP) {
        const int k = base + lane;
        bool ok = false;
        V3 n{0.0f, 0.0f, 0.0f};
        float d = 0.0f;
        uint32_t e = 0;
        if (k < nh) {
            e = L.hor[k];
            ok = epa_face_normal(L, (int)(e & 255u), (int)(e >> 8), wi, nvtx, n, d);
        }
        const uint64_t m = G::ballot(ok);
        const int slot = nf + made + (int)G::rank(m);
        if (ok && k != nh - 1) ++made_before_last;   // counted per lane, summed below
        if (ok && slot < NF) {
            L.nx[slot] = n.x; L.ny[slot] = n.y; L.nz[slot] = n.z; L.nd[slot] = d;
            L.fv[slot] = e | ((uint32_t)wi << 16);
        }
        made += __popcll(m);
    }
    // the reference checks its pool (every face made) before each new face, failing ones
    // included; a smaller pool hands the pair over where its live faces would not fit
    const int mb = (int)G::sum((uint32_t)made_before_last);
    if (nh > 0 && nf_tot + mb >= EPA_FACES) return EPA_OUT_OF_FACES;
    if (NF < EPA_FACES && nh > 0 && nf + mb >= NF) return EPA_DEFER;
    nf += made;
    nf_tot += made;
    G::sync();
    return 0;
}

constexpr float EPA_GUARD_TOL = 1e-4f;   // depth excess (m) over the guard direction's overlap that counts
// Face normals are tried only when the depth in question exceeds this (m): resting contacts
// are a fraction of a millimetre deep, the faces FP32 EPA settles on wrongly centimetres.
constexpr float EPA_GUARD_SCAN = 2e-3f;
// Face normals the EPA guard tries (local frame): a hull's merged polygons, a box's
// +x -x +y -y +z -z; other shapes none.
template <class SideT> __device__ __forceinline__ int face_count(const SideT &S, const HullSet &hs, int hull) {
    return S.kind == SH_HULL ? (int)hs.meta[hull].np : S.kind == SH_BOX ? 6 : 0;
}
template <class SideT> __device__ __forceinline__ V3 face_normal_local(const SideT &S, const HullSet &hs, int hull, int k) {
    if (S.kind == SH_HULL) {
        const float4 p = hs.pn[hs.meta[hull].pn0 + k];
        return V3{p.x, p.y, p.z};
    }
    const float sg = (k & 1) ? -1.0f : 1.0f;
    return k < 2 ? V3{sg, 0.0f, 0.0f} : k < 4 ? V3{0.0f, sg, 0.0f} : V3{0.0f, 0.0f, sg};
}
struct EpaResult {
    int status, iters;
    V3 n, p;
    float depth;
};

template <class SideT, int NV, int NF>
__device__ __forceinline__ EpaResult epa(const Simplex &seed, const SideT &A, const SideT &Bs, const HullSet &hs,
                                         EpaLdsT<NV, NF> &L, bool seeded, V3 seed_n) {
    constexpr int LPP = SideT::kLanes;
    using G = Grp<LPP>;
    EpaResult r{EPA_FAILED, 0, V3{0.0f, 0.0f, 0.0f}, V3{0.0f, 0.0f, 0.0f}, 0.0f};
    const int lane = G::lane();
    if (seed.rank != 4) { r.status = EPA_NONLINEAR; return r; }
    {   // coplanar seed: a touching contact (EPA.cpp:174-212)
        const V3 va = seed.v[0].w, vb = seed.v[1].w, vc = seed.v[2].w, vd = seed.v[3].w;
        const float e2 = EPS * EPS;
        V3 pn = cross(vb - va, vc - va);
        if (len2(pn) < e2) pn = cross(vb - va, vd - va);
        if (len2(pn) < e2) pn = cross(vc - va, vd - va);
        if (len2(pn) > e2) {
            pn = normalize(pn);
            const float dd = absf(dot(vd - va, pn));
            if (dd < 1e-5f) {
                const V3 sp = side_support(A, hs, pn) - side_support(Bs, hs, -pn);
                const V3 sn = side_support(A, hs, -pn) - side_support(Bs, hs, pn);
                const float dp = dot(sp - va, pn);
                const float dn = dot(sn - va, -pn);
                if (dp < 1e-4f || dn < 1e-4f) {
                    r.n = (dp <= dn) ? pn : -pn;
                    r.depth = 0.0f;
                    r.p = (((seed.v[0].w0 + seed.v[1].w0) + seed.v[2].w0) + seed.v[3].w0) * 0.25f;
                    r.status = EPA_VALID;
                    return r;
                }
            }
        }
    }
    if (lane == 0) {   // constant indices (a lane-dependent one put the simplex on the stack)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const SV &s = seed.v[k];
            L.wx[k] = s.w.x; L.wy[k] = s.w.y; L.wz[k] = s.w.z;
            L.ax[k] = s.w0.x; L.ay[k] = s.w0.y; L.az[k] = s.w0.z;
        }
    }
    G::sync();
    int nvtx = 4, nf = 0, nf_tot = 0;
    {   // seed faces ABC, ACD, ADB, BCD (every lane evaluates them alike)
        const int fa[4] = {0, 0, 0, 1}, fb[4] = {1, 2, 3, 2}, fc[4] = {2, 3, 1, 3};
        for (int f = 0; f < 4; ++f) {
            V3 n;
            float d;
            if (!epa_face_normal(L, fa[f], fb[f], fc[f], 4, n, d)) { r.status = EPA_NONLINEAR; return r; }
            if (lane == f) {
                L.nx[f] = n.x; L.ny[f] = n.y; L.nz[f] = n.z; L.nd[f] = d;
                L.fv[f] = (uint32_t)fa[f] | ((uint32_t)fb[f] << 8) | ((uint32_t)fc[f] << 16);
            }
        }
        nf = nf_tot = 4;
    }
    G::sync();
    int it0 = 0;
    if (seeded) {
        // warm start: one expansion along a predicted normal, when its support is a new
        // vertex that a face sees
        it0 = 1;
        const SV w = md_sample(A, Bs, hs, seed_n);
        bool dup = false, seen = false;
        if (lane < nvtx) dup = len2(lds_w(L, lane) - w.w) < EPS * EPS;
        if (lane < nf) {
            const uint32_t fv = L.fv[lane];
            const V3 cen = ((lds_w(L, (int)(fv & 255u)) + lds_w(L, (int)((fv >> 8) & 255u))) +
                            lds_w(L, (int)((fv >> 16) & 255u))) * (1.0f / 3.0f);
            seen = dot(V3{L.nx[lane], L.ny[lane], L.nz[lane]}, w.w - cen) > 0.0f;
        }
        if (!G::ballot(dup) && G::ballot(seen)) {
            const int rc = epa_expand<LPP>(L, w, nvtx, nf, nf_tot);
            if (rc) {
                r.iters = 1;
                r.status = rc;
                return r;
            }
        }
    }
    for (int it = 0; it < EPA_ITERS; ++it) {
        r.iters = it0 + it + 1;
        // 1. closest live face: lane-local first minimum, wave minimum of the mapped
        // distance, lowest face index among the lanes holding it; faces at or above
        // FLT_MAX (and NaN) never qualify, as in the reference's scan from DBL_MAX
        uint32_t best = 0xffffffffu;
        int bi = 0x7fffffff;
        for (int f = lane; f < nf; f += LPP) {
            const float d = L.nd[f];
            if (!(L.fv[f] >> 24) && d < 3.402823466e+38f) {
                const uint32_t o = ord_f(d);
                if (o < best) { best = o; bi = f; }
            }
        }
        const uint32_t mo = G::min_u32(best);
        if (mo == 0xffffffffu) { r.status = EPA_FAILED; return r; }
        const int ci = (int)G::min_u32(best == mo ? (uint32_t)bi : 0xffffffffu);
        const float mind = L.nd[ci];
        const V3 cn{L.nx[ci], L.ny[ci], L.nz[ci]};
        const uint32_t cv = L.fv[ci];
        const int c0 = (int)(cv & 255u), c1 = (int)((cv >> 8) & 255u), c2 = (int)((cv >> 16) & 255u);
        // 2-3. support along its normal; convergence
        const SV w = md_sample(A, Bs, hs, cn);
        const float dproj = dot(w.w, cn);
        if (dproj - mind < 1e-6f) {
            r.n = cn;
            r.depth = mind;
            const V3 q = cn * mind;
            const V3 p0 = lds_w(L, c0), p1 = lds_w(L, c1), p2 = lds_w(L, c2);
            const V3 e0 = p1 - p0, e1 = p2 - p0, rr = q - p0;
            const float d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1), d20 = dot(rr, e0), d21 = dot(rr, e1);
            const float den = d00 * d11 - d01 * d01;
            if (absf(den) < EPS) {
                r.p = lds_a(L, c0);
            } else {
                const float vw = (d11 * d20 - d01 * d21) / den;
                const float ww = (d00 * d21 - d01 * d20) / den;
                const float uw = 1.0f - vw - ww;
                r.p = (lds_a(L, c0) * uw + lds_a(L, c1) * vw) + lds_a(L, c2) * ww;
            }
            r.status = EPA_VALID;
            return r;
        }
        // duplicate support: stop on the closest face
        bool dup = false;
        for (int i = lane; i < nvtx; i += LPP) dup |= len2(lds_w(L, i) - w.w) < EPS * EPS;
        if (G::ballot(dup)) {
            r.n = cn;
            r.depth = mind;
            r.p = lds_a(L, c0);
            r.status = EPA_VALID;
            return r;
        }
        if (nvtx >= NV) { r.status = NV < EPA_VERTS ? EPA_DEFER : EPA_OUT_OF_VERTICES; return r; }
        if (const int rc = epa_expand<LPP>(L, w, nvtx, nf, nf_tot)) { r.status = rc; return r; }
    }
    r.status = EPA_FAILED;
    return r;
}

// ------------------------------------------------------------------ manifolds
// One hull's tables as offsets into the world's arrays (vertices, polygon normals, the u32
// table): 32-bit offsets rather than pointers keep the per-pair copies (two hulls, and the
// reference/incident pick) at half the registers.
struct HullTabs {
    uint32_t v, pn, poff, pidx, v2p_o, v2p_i, v2n_o, v2n_i, p2p_o, p2p_i;
};
__device__ __forceinline__ HullTabs hull_tabs(const HullSet &hs, uint32_t h) {
    const HullMeta m = hs.meta[h];
    return HullTabs{m.v0, m.pn0, m.poff, m.pidx, m.v2p_o, m.v2p_i, m.v2n_o, m.v2n_i, m.p2p_o, m.p2p_i};
}
__device__ __forceinline__ V3 tv(const float4 *p, uint32_t i) {
    const float4 q = p[i];
    return V3{q.x, q.y, q.z};
}

// Polygon at a support vertex most aligned with wd (ContactManifoldGenerator.cpp:73-95).
__device__ __forceinline__ uint32_t best_polygon(const HullSet &hs, const HullTabs &h, const TF &tf, uint32_t sv, V3 wd,
                                                 float &al) {
    const uint32_t s = hs.tab[h.v2p_o + sv], e = hs.tab[h.v2p_o + sv + 1];
    uint32_t best = (s < e) ? hs.tab[h.v2p_i + s] : 0u;
    float bd = -3.402823466e+38f;
    for (uint32_t i = s; i < e; ++i) {
        const uint32_t p = hs.tab[h.v2p_i + i];
        const float x = dot(tf_vec(tf, tv(hs.pn + h.pn, p)), wd);
        if (x > bd) { bd = x; best = p; }
    }
    al = bd;
    return best;
}

// Clipping.cpp:6-36
__device__ __forceinline__ bool clip_in(V3 pp, V3 pn, V3 p) { return dot(p, pn) + (-dot(pn, pp)) >= 0.0f; }
__device__ __forceinline__ bool clip_edge(V3 pp, V3 pn, V3 s, V3 e, V3 &out) {
    const V3 ab = e - s;
    const float abp = dot(pn, ab);
    if (absf(abp) <= 1e-6f) return false;
    const float dist = -dot(pn, pp);
    const V3 pco = pn * (-dist);
    float fac = -dot(pn, s - pco) / abp;
    if (fac < 0.0f) fac = 0.0f;
    if (fac > 1.0f) fac = 1.0f;
    out = s + ab * fac;
    return true;
}
// Sutherland-Hodgman of buffer `src` (n points) against one plane into buffer 1 - src.
// A convex n-gon gains at most one point per plane, so HCLIP_CAP never binds; the bound
// only keeps a numerically non-convex input inside the buffer.  A smaller buffer (HC below
// HCLIP_CAP) sets `full` where the reference-sized one would take the point.
template <int HC>
__device__ __forceinline__ int clip_pass(ClipLdsT<HC> &C, int src, int n, V3 pp, V3 pn, bool remove_only, bool &full) {
    const int dst = 1 - src;
    int m = 0;
    auto put = [&](V3 p) {
        if (m < HC) { C.px[dst][m] = p.x; C.py[dst][m] = p.y; C.pz[dst][m] = p.z; ++m; }
        else if (HC < HCLIP_CAP) full = true;
    };
    V3 sp{C.px[src][n - 1], C.py[src][n - 1], C.pz[src][n - 1]};
    for (int j = 0; j < n; ++j) {
        const V3 ep{C.px[src][j], C.py[src][j], C.pz[src][j]};
        const bool si = clip_in(pp, pn, sp), ei = clip_in(pp, pn, ep);
        V3 cut;
        if (remove_only) {
            if (ei) put(ep);
        } else if (si && ei) {
            put(ep);
        } else if (si && !ei) {
            if (clip_edge(pp, pn, sp, ep, cut)) put(cut);
        } else if (!si && ei) {
            if (clip_edge(pp, pn, sp, ep, cut)) put(cut);
            put(ep);
        }
        sp = ep;
    }
    return m;
}
// Clipping.cpp:104-128
__device__ __forceinline__ bool skew_lines(V3 p1, V3 d1, V3 p2, V3 d2, V3 &l1, V3 &l2) {
    const float n1 = d1.x * d2.x + d1.y * d2.y + d1.z * d2.z;
    const float n2 = d2.x * d2.x + d2.y * d2.y + d2.z * d2.z;
    const float m1 = -d1.x * d1.x - d1.y * d1.y - d1.z * d1.z;
    const float m2 = -d2.x * d1.x - d2.y * d1.y - d2.z * d1.z;
    const float r1 = -d1.x * p2.x + d1.x * p1.x - d1.y * p2.y + d1.y * p1.y - d1.z * p2.z + d1.z * p1.z;
    const float r2 = -d2.x * p2.x + d2.x * p1.x - d2.y * p2.y + d2.y * p1.y - d2.z * p2.z + d2.z * p1.z;
    const float den = (n1 * m2) - (n2 * m1);
    if (absf(den) < EPS) return false;
    const float n = ((r1 * m2) - (r2 * m1)) / den;
    const float m = ((n1 * r2) - (n2 * r1)) / den;
    l1 = p1 + d1 * m;
    l2 = p2 + d2 * n;
    return true;
}
// ContactManifoldGenerator.hpp:88-154 over the kept points in C.px[buf]/C.dep (lane 0).
template <int HC> __device__ __forceinline__ int reduce4_area(const ClipLdsT<HC> &C, int buf, int n, Manifold &m) {
    auto P = [&](int i) { return V3{C.px[buf][i], C.py[buf][i], C.pz[buf][i]}; };
    if (n <= 4) {
#pragma unroll
        for (int i = 0; i < 4; ++i)   // constant indices keep the manifold in registers
            if (i < n) { m.p[i] = P(i); m.dep[i] = C.dep[i]; }
        return n;
    }
    int i0 = 0;
    for (int i = 1; i < n; ++i) if (C.dep[i] > C.dep[i0]) i0 = i;
    int i1 = (i0 == 0) ? 1 : 0;
    float best = len2(P(i1) - P(i0));
    for (int i = 0; i < n; ++i) {
        if (i == i0) continue;
        const float x = len2(P(i) - P(i0));
        if (x > best) { best = x; i1 = i; }
    }
    int i2 = -1;
    best = 0.0f;
    const V3 seg = P(i1) - P(i0);
    for (int i = 0; i < n; ++i) {
        if (i == i0 || i == i1) continue;
        const float x = len2(cross(P(i) - P(i0), seg));
        if (x > best) { best = x; i2 = i; }
    }
    if (i2 < 0) for (int i = 0; i < n; ++i) if (i != i0 && i != i1) { i2 = i; break; }
    int i3 = -1;
    best = 0.0f;
    if (i2 >= 0) {
        const V3 tn = cross(P(i1) - P(i0), P(i2) - P(i0));
        for (int i = 0; i < n; ++i) {
            if (i == i0 || i == i1 || i == i2) continue;
            const float x = absf(dot(P(i) - P(i0), tn));
            if (x > best) { best = x; i3 = i; }
        }
        if (i3 < 0) for (int i = 0; i < n; ++i) if (i != i0 && i != i1 && i != i2) { i3 = i; break; }
    }
    m.p[0] = P(i0); m.dep[0] = C.dep[i0];
    m.p[1] = P(i1); m.dep[1] = C.dep[i1];
    if (i2 >= 0) { m.p[2] = P(i2); m.dep[2] = C.dep[i2]; }
    if (i3 >= 0) { m.p[3] = P(i3); m.dep[3] = C.dep[i3]; return 4; }
    return (i2 >= 0) ? 3 : 2;
}

// Hull-hull manifold from the EPA result (ContactManifoldGenerator.cpp:176-379).  The
// result is valid on the group's lane 0, which returns false when a clip buffer smaller than
// the reference's ran out (the pair is then run again with the full buffers).
template <class SideT, int HC>
__device__ __forceinline__ bool manifold_hulls(const EpaResult &e, const SideT &A, const SideT &Bs, const HullSet &hs,
                                               uint32_t ha_id, uint32_t hb_id, ClipLdsT<HC> &C, Manifold &m) {
    constexpr int LPP = SideT::kLanes;
    using G = Grp<LPP>;
    const int lane = G::lane();
    const float lowest = -3.402823466e+38f;
    m.n = e.n;
    m.np = 0;
    const HullTabs ha = hull_tabs(hs, ha_id), hb = hull_tabs(hs, hb_id);
    const uint32_t sa = (uint32_t)wave_support_index(A.wh, tf_inv_vec(A.tf, e.n));
    const uint32_t sb = (uint32_t)wave_support_index(Bs.wh, tf_inv_vec(Bs.tf, -e.n));
    float al_a, al_b;
    const uint32_t pa = best_polygon(hs, ha, A.tf, sa, e.n, al_a);
    const uint32_t pb = best_polygon(hs, hb, Bs.tf, sb, -e.n, al_b);
    // edge pairs at the two support vertices, in (i, j, sign) order
    const float4 *va = hs.v + ha.v, *vb = hs.v + hb.v;
    const V3 wsa = tf_point(A.tf, tv(va, sa)), wsb = tf_point(Bs.tf, tv(vb, sb));
    const uint32_t as = hs.tab[ha.v2n_o + sa], ae = hs.tab[ha.v2n_o + sa + 1];
    const uint32_t bs = hs.tab[hb.v2n_o + sb], be = hs.tab[hb.v2n_o + sb + 1];
    const int dega = (int)(ae - as), degb = (int)(be - bs);
    float al_e = lowest;
    int key = 0x7fffffff;
    if (dega > 0 && degb > 0) {
        for (int c = lane; c < dega * degb; c += LPP) {
            const uint32_t na = hs.tab[ha.v2n_i + as + c / degb], nb = hs.tab[hb.v2n_i + bs + c % degb];
            const V3 ea = wsa - tf_point(A.tf, tv(va, na));
            const V3 eb = wsb - tf_point(Bs.tf, tv(vb, nb));
            const V3 cc = cross(ea, eb);
            const float cl = len(cc);
            if (cl < EPS) continue;
            const V3 cn = vdiv(cc, cl);
            const float dp = dot(cn, e.n);
            if (dp > al_e) { al_e = dp; key = 2 * c; }
            const float dn = -dp;
            if (dn > al_e) { al_e = dn; key = 2 * c + 1; }
        }
    }
    G::argmax(al_e, key);
    const bool has_edge = key != 0x7fffffff;
    if (lane != 0) return true;
    if (has_edge && al_e > al_a + 1e-4f && al_e > al_b + 1e-4f) {
        const int c = key / 2;
        const uint32_t na = hs.tab[ha.v2n_i + as + c / degb], nb = hs.tab[hb.v2n_i + bs + c % degb];
        const V3 p1 = tf_point(A.tf, tv(va, sa));
        const V3 d1 = tf_point(A.tf, tv(va, na)) - p1;
        const V3 p2 = tf_point(Bs.tf, tv(vb, sb));
        const V3 d2 = tf_point(Bs.tf, tv(vb, nb)) - p2;
        V3 l1, l2;
        m.np = 1;
        m.dep[0] = e.depth;
        m.p[0] = skew_lines(p1, d1, p2, d2, l1, l2) ? (l1 + l2) * 0.5f : e.p;
        return true;
    }
    const bool ref_a = al_a > al_b;
    // by value: a reference to one of two locals would put both on the stack
    // the reference / incident hull's tables read again (copies of ha / hb by value held four
    // table sets in registers at once)
    const HullTabs hr = hull_tabs(hs, ref_a ? ha_id : hb_id);
    const HullTabs hi = hull_tabs(hs, ref_a ? hb_id : ha_id);
    const TF tr = ref_a ? A.tf : Bs.tf;
    const TF ti = ref_a ? Bs.tf : A.tf;
    const uint32_t rp = ref_a ? pa : pb, ip = ref_a ? pb : pa;
    const float4 *vr = hs.v + hr.v, *vi = hs.v + hi.v;
    const V3 rn = tf_vec(tr, tv(hs.pn + hr.pn, rp));
    const uint32_t is = hs.tab[hi.poff + ip], ie = hs.tab[hi.poff + ip + 1];
    const int inc_n = (int)(ie - is);
    if (inc_n < 3) {
        m.np = 1; m.p[0] = e.p; m.dep[0] = e.depth;
        return true;
    }
    if (HC < HCLIP_CAP && inc_n > HC) return false;
    for (int i = 0; i < inc_n; ++i) {
        const V3 p = tf_point(ti, tv(vi, hs.tab[hi.pidx + is + i]));
        C.px[0][i] = p.x; C.py[0][i] = p.y; C.pz[0][i] = p.z;
    }
    const V3 rpt = tf_point(tr, tv(vr, hs.tab[hr.pidx + hs.tab[hr.poff + rp]]));
    int n = inc_n, buf = 0;
    bool full = false;
    const uint32_t k1 = hs.tab[hr.p2p_o + rp + 1];
    for (uint32_t k = hs.tab[hr.p2p_o + rp]; k < k1 && n > 0; ++k) {
        const uint32_t q = hs.tab[hr.p2p_i + k];
        const V3 pp = tf_point(tr, tv(vr, hs.tab[hr.pidx + hs.tab[hr.poff + q]]));
        const V3 pn = -tf_vec(tr, tv(hs.pn + hr.pn, q));
        n = clip_pass(C, buf, n, pp, pn, false, full);
        buf = 1 - buf;
    }
    if (n > 0) {
        n = clip_pass(C, buf, n, rpt, -rn, true, full);
        buf = 1 - buf;
    }
    if (full) return false;
    int kept = 0;   // compacted in place
    for (int i = 0; i < n; ++i) {
        const V3 p{C.px[buf][i], C.py[buf][i], C.pz[buf][i]};
        const float dtp = dot(p - rpt, rn);
        const V3 diff = rn * dtp;
        const float pen = ref_a ? dot(diff, e.n) : -dot(diff, e.n);
        if (pen <= 0.0f) {
            C.px[buf][kept] = p.x; C.py[buf][kept] = p.y; C.pz[buf][kept] = p.z;
            C.dep[kept] = -pen;
            ++kept;
        }
    }
    if (!kept) {
        m.np = 1; m.p[0] = e.p; m.dep[0] = e.depth;
        return true;
    }
    m.np = reduce4_area(C, buf, kept, m);
    return true;
}

// ------------------------------------------------------------------ kernels
template <int VPL, int LPP>
__device__ __forceinline__ void load_side(Side<VPL, LPP> &s, const Colliders &C, const HullSet &hs, uint32_t c, const TF &tf) {
    const int kind = C.meta[c].x;
    const float4 d = C.dims[c];
    s.kind = kind;
    s.a = d.x; s.b = d.y; s.c = d.z;
    s.tf = tf;
    s.wh.nv = 0;
    s.wh.v0 = 0;
    if (kind == SH_HULL) load_wave_hull(s.wh, hs, (uint32_t)C.hull[c]);
}

// GJK + EPA + manifold of the hull-hull / hull-primitive bins (ContactManifoldGenerator.cpp:415-438)
// in two launches.  Pass 1 (k_gjk_pass): GJK only, one wavefront per pair, four per
// workgroup, no LDS; a separated pair is finished there (no contact), an overlapping
// pair's final simplex is stored and its bin position listed in its workgroup's
// four-entry segment.  Pass 2 (k_epa_pass): one wavefront per pass-1 workgroup runs EPA
// (polytope in LDS) and the manifold for the listed pairs.  Most broadphase pairs are
// separated, so the heavy EPA/manifold code and its LDS only occupy the GPU for the few
// that overlap.
// gstat: per slot, GJK iterations | EPA iterations << 8 | GJK status << 16 | EPA status
// << 20 (15: not run) | fallback << 24; iterations of a fallback run are added.
// guess (optional, per bin position two float4 in A's frame): warm start, the GJK
// direction and (seed_epa) an EPA seed normal; a zero direction means the plain seed.
// A warm-started GJK gets GJK_WARM_ITERS iterations; when it fails, or EPA from its
// simplex fails, the plain run follows.
// tlog (optional, per candidate two float4 in A's frame, after *tlog_base rows): training targets, the EPA
// normal (or, for a separated pair, the reverse of the GJK exit direction) with a valid
// flag, and the EPA normal with a valid flag.
constexpr int GJK_PASS_WAVES = 4;
struct GjkSeeds {
    V3 plain, g, sn;
    bool warm, seeded;
};
__device__ __forceinline__ GjkSeeds gjk_seeds(const TF &ta, const TF &tb, const float4 *guess, uint32_t t, int seed_epa) {
    GjkSeeds q;
    q.plain = tb.pos - ta.pos;
    if (len2(q.plain) < EPS) q.plain = V3{1.0f, 0.0f, 0.0f};
    q.g = q.plain;
    q.sn = V3{0.0f, 0.0f, 0.0f};
    q.warm = q.seeded = false;
    if (guess) {
        const V3 gl = xyz(guess[2 * t]);
        if (len2(gl) > 0.5f) { q.g = tf_vec(ta, gl); q.warm = true; }
        const V3 sl = xyz(guess[2 * t + 1]);
        if (q.warm && seed_epa && len2(sl) > 0.5f) { q.sn = tf_vec(ta, sl); q.seeded = true; }
    }
    return q;
}
__device__ __forceinline__ uint32_t pack_gstat(int gsum, int esum, int gs, int es, int fb) {
    return (uint32_t)(gsum < 255 ? gsum : 255) | ((uint32_t)(esum < 255 ? esum : 255) << 8) | ((uint32_t)gs << 16) |
           ((uint32_t)es << 20) | ((uint32_t)fb << 24);
}

#ifndef GJK_PASS_OCCUPANCY
#define GJK_PASS_OCCUPANCY 4
#endif
// Candidate lists (k_gjk_cull): with `cand`, the passes run over the *ncand candidates,
// candidate tc being bin position cand[tc]; the warm starts are per candidate, the training
// targets per bin position.  Both passes stride over "virtual blocks" of gjk_pass_pairs
// candidates (block b takes b, b + gridDim.x, ...), so the grid does not depend on the count.
__device__ __forceinline__ uint32_t gjk_count(uint32_t n, const uint32_t *ncand) { return ncand ? min(*ncand, n) : n; }

// Pairs per workgroup of the GJK pass: four wavefronts of 64 / LPP pairs each; the EPA pass
// reads the overlapping pairs in segments of this size.
template <int LPP> constexpr int gjk_pass_pairs() { return GJK_PASS_WAVES * (64 / LPP); }

template <int VPL, int LPP, int OCC = GJK_PASS_OCCUPANCY>
__global__ void __launch_bounds__(64 * GJK_PASS_WAVES) __attribute__((amdgpu_waves_per_eu(OCC)))
k_gjk_pass(const uint32_t *bin_slot, uint32_t off, uint32_t n, const uint32_t *cand, const uint32_t *ncand,
                                                                  Colliders C, Bodies B, Contacts K, HullSet hs,
                                                                  uint32_t *gstat, const float4 *guess, float4 *tlog,
                                                                  const uint32_t *tlog_base, uint32_t *hits,
                                                                  uint32_t *blk_hits, float4 *sxs, float4 *axis) {
    constexpr int PPB = gjk_pass_pairs<LPP>();
    __shared__ uint32_t inside_g[PPB];
    const int g = (int)threadIdx.x / LPP, pl = pair_lane<LPP>();   // pair group in the block, lane within it
    const uint32_t nc = gjk_count(n, ncand);
    if (tlog && tlog_base) tlog += 2 * (size_t)*tlog_base;
    for (uint32_t vb = blockIdx.x; vb * PPB < nc; vb += gridDim.x) {
        const uint32_t tc = vb * PPB + g;   // candidate
        bool inside = false;
        Simplex sx;
        if (tc < nc) {
            const uint32_t t = cand ? cand[tc] : tc;   // bin position
            const uint32_t s = bin_slot[off + t];
            const uint2 cc = K.cc[s], bb = K.bb[s];
            const TF ta = collider_tf(C, B, cc.x, bb.x), tb = collider_tf(C, B, cc.y, bb.y);
            Side<VPL, LPP> A, Bs;
            load_side(A, C, hs, cc.x, ta);
            load_side(Bs, C, hs, cc.y, tb);
            const GjkSeeds q = gjk_seeds(ta, tb, guess, tc, 0);
            int it = 0;
            int gs = gjk(A, Bs, hs, q.g, sx, it, q.warm ? GJK_WARM_ITERS : GJK_ITERS);
            int gsum = it, fb = 0;
            if (q.warm && gs == GJK_FAILED) {
                fb = 1;
                gs = gjk(A, Bs, hs, q.plain, sx, it);
                gsum += it;
            }
            inside = gs == GJK_INSIDE;
            if (pl == 0) gstat[s] = pack_gstat(gsum, 0, gs, 15, fb);   // pass 2 completes an overlapping pair
            if (pl == 0 && axis) {
                // a separated pair's last direction: the support of A - B along it is
                // negative, i.e. B lies beyond A along +dir (a hint for the next rejection test)
                const float ln = len(sx.dir);
                axis[s] = gs == GJK_VALID && ln > 1e-12f ? f4(vdiv(sx.dir, ln), 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            if (!inside && pl == 0) {
                K.cnt[s] = 0;
                if (tlog) {
                    V3 tg{0.0f, 0.0f, 0.0f};
                    float vg = 0.0f;
                    const float ln = len(sx.dir);
                    // a seed against the exit direction puts the first support on the far side of
                    // the Minkowski difference, so that -v0 already points across it at the origin
                    if (gs == GJK_VALID && ln > 1e-12f) { tg = tf_inv_vec(ta, vdiv(-sx.dir, ln)); vg = 1.0f; }
                    tlog[2 * tc] = f4(tg, vg);
                    tlog[2 * tc + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
            }
        }
        if (pl == 0) inside_g[g] = inside ? 1u : 0u;
        __syncthreads();
        int pos = 0;
        for (int k = 0; k < g; ++k) pos += (int)inside_g[k];
        if (inside && pl == 0) {
            const uint32_t h = vb * PPB + pos;
            hits[h] = tc;
            // the four simplex vertices: (w.xyz, w0.x) (w0.y, w0.z); constant indices keep the
            // simplex in registers (a lane-dependent index put it on the stack)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const SV &v = sx.v[k];
                sxs[(size_t)h * 8 + 2 * k] = make_float4(v.w.x, v.w.y, v.w.z, v.w0.x);
                sxs[(size_t)h * 8 + 2 * k + 1] = make_float4(v.w0.y, v.w0.z, 0.0f, 0.0f);
            }
        }
        if (threadIdx.x == 0) {
            uint32_t c = 0;
            for (int k = 0; k < PPB; ++k) c += inside_g[k];
            blk_hits[vb] = c;
        }
        __syncthreads();   // inside_g is rewritten by the next virtual block
    }
}

// Pass 1 with one pair per lane: each thread runs GJK for one candidate, scanning its hulls'
// vertices itself (lane_hull_support), so a wavefront serves 64 pairs with one instruction
// stream and no cross-lane reductions.  Same per-slot results as k_gjk_pass and the same
// lists of overlapping pairs: candidates [16 j, 16 j + 16) form segment j (hits at 16 j + r,
// blk_hits[j] entries, simplices beside them), which the EPA pass reads with seg = 16.
constexpr uint32_t GJK_LANE_SEG = 16;
__global__ void __launch_bounds__(64) k_gjk_lanes(const uint32_t *bin_slot, uint32_t off, uint32_t n, const uint32_t *cand,
                                                   const uint32_t *ncand, Colliders C, Bodies B, Contacts K, HullSet hs,
                                                   uint32_t *gstat, const float4 *guess, float4 *tlog,
                                                   const uint32_t *tlog_base, uint32_t *hits, uint32_t *blk_hits,
                                                   float4 *sxs, float4 *axis) {
    const uint32_t nc = gjk_count(n, ncand);
    if (tlog && tlog_base) tlog += 2 * (size_t)*tlog_base;
    const int lane = lane64();
    for (uint32_t base = blockIdx.x * blockDim.x; base < nc; base += gridDim.x * blockDim.x) {   // uniform per block
        const uint32_t tc = base + threadIdx.x;   // candidate
        bool inside = false;
        Simplex sx;
        if (tc < nc) {
            const uint32_t t = cand ? cand[tc] : tc;   // bin position
            const uint32_t s = bin_slot[off + t];
            const uint2 cc = K.cc[s], bb = K.bb[s];
            const TF ta = collider_tf(C, B, cc.x, bb.x), tb = collider_tf(C, B, cc.y, bb.y);
            LaneSide A, Bs;
            load_lane_side(A, C, hs, cc.x, ta);
            load_lane_side(Bs, C, hs, cc.y, tb);
            const GjkSeeds q = gjk_seeds(ta, tb, guess, tc, 0);
            int it = 0;
            int gs = gjk(A, Bs, hs, q.g, sx, it, q.warm ? GJK_WARM_ITERS : GJK_ITERS);
            int gsum = it, fb = 0;
            if (q.warm && gs == GJK_FAILED) {
                fb = 1;
                gs = gjk(A, Bs, hs, q.plain, sx, it);
                gsum += it;
            }
            inside = gs == GJK_INSIDE;
            gstat[s] = pack_gstat(gsum, 0, gs, 15, fb);   // pass 2 completes an overlapping pair
            if (axis) {   // as in k_gjk_pass: a separated pair's last direction (rejection hint)
                const float ln = len(sx.dir);
                axis[s] = gs == GJK_VALID && ln > 1e-12f ? f4(vdiv(sx.dir, ln), 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            if (!inside) {
                K.cnt[s] = 0;
                if (tlog) {
                    V3 tg{0.0f, 0.0f, 0.0f};
                    float vg = 0.0f;
                    const float ln = len(sx.dir);
                    if (gs == GJK_VALID && ln > 1e-12f) { tg = tf_inv_vec(ta, vdiv(-sx.dir, ln)); vg = 1.0f; }
                    tlog[2 * tc] = f4(tg, vg);
                    tlog[2 * tc + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
            }
        }
        // segments: this lane's group of 16 lanes holds the candidates of one segment
        const uint32_t gm = (uint32_t)((__ballot(inside) >> (lane & ~15)) & 0xffffull);
        const uint32_t seg = tc / GJK_LANE_SEG;
        if (inside) {
            const uint32_t h = seg * GJK_LANE_SEG + (uint32_t)__popc(gm & ((1u << (lane & 15)) - 1u));
            hits[h] = tc;
#pragma unroll
            for (int k = 0; k < 4; ++k) {   // the simplex: (w.xyz, w0.x) (w0.y, w0.z)
                const SV &v = sx.v[k];
                sxs[(size_t)h * 8 + 2 * k] = make_float4(v.w.x, v.w.y, v.w.z, v.w0.x);
                sxs[(size_t)h * 8 + 2 * k + 1] = make_float4(v.w0.y, v.w0.z, 0.0f, 0.0f);
            }
        }
        if ((lane & 15) == 0 && tc < nc) blk_hits[seg] = (uint32_t)__popc(gm);
    }
}

// Pass 1 with one pair per 16-lane DPP row and rows refilled as they finish: a wavefront
// owns a range of GJK_REFILL_SEG candidates and runs GJK one iteration at a time on its four
// rows; a row whose pair is done takes the range's next candidate at once, so a wavefront
// does not wait for the slowest of four pairs (k_gjk_pass does, every four).  Each pair runs
// exactly the iterations gjk() runs (same seeds, warm-start budget, fallback, cycle test),
// so statuses, iteration counts and simplices are those of k_gjk_pass.  Overlapping pairs of
// range j are listed at GJK_REFILL_SEG * j + r (in completion order), blk_hits[j] of them,
// the simplices beside them: the EPA pass reads them with seg = GJK_REFILL_SEG.
constexpr uint32_t GJK_REFILL_SEG = 16;
template <int OCC>
__global__ void __launch_bounds__(64 * GJK_PASS_WAVES) __attribute__((amdgpu_waves_per_eu(OCC)))
k_gjk_refill(const uint32_t *bin_slot, uint32_t off, uint32_t n, const uint32_t *cand, const uint32_t *ncand, Colliders C,
             Bodies B, Contacts K, HullSet hs, uint32_t *gstat, const float4 *guess, float4 *tlog, const uint32_t *tlog_base,
             uint32_t *hits, uint32_t *blk_hits, float4 *sxs, float4 *axis) {
    constexpr int LPP = 16;
    const uint32_t nc = gjk_count(n, ncand);
    if (tlog && tlog_base) tlog += 2 * (size_t)*tlog_base;
    const int lane = lane64(), row = lane >> 4, pl = lane & 15;
    const uint32_t wave = blockIdx.x * GJK_PASS_WAVES + threadIdx.x / 64, nwaves = gridDim.x * GJK_PASS_WAVES;
    for (uint32_t rg = wave; rg * GJK_REFILL_SEG < nc; rg += nwaves) {   // uniform per wavefront
        const uint32_t end = min(nc, (rg + 1) * GJK_REFILL_SEG);
        uint32_t next = rg * GJK_REFILL_SEG, nhit = 0;   // wavefront-uniform
        // this row's pair
        bool active = false;
        uint32_t tc = 0, s = 0;
        Side<4, LPP> A, Bs;   // hulls of at most 64 vertices: four per lane
        Simplex sx;
        V3 dir{0.0f, 0.0f, 0.0f}, plain{0.0f, 0.0f, 0.0f};
        GjkCycle<LPP> cyc;
        int it = 0, max_it = GJK_ITERS, gsum = 0, fb = 0;
        bool warm = false;
        for (;;) {
            // rows without a pair take the next candidates, in row order
            const uint32_t need = (uint32_t)__popcll(__ballot(!active && pl == 0) & 0x0001000100010001ull);
            const uint32_t before = (uint32_t)__popcll(__ballot(!active && pl == 0) & ((1ull << (row * 16)) - 1ull));
            if (!active && next + before < end) {
                tc = next + before;
                const uint32_t t = cand ? cand[tc] : tc;   // bin position
                s = bin_slot[off + t];
                const uint2 cc = K.cc[s], bb = K.bb[s];
                const TF ta = collider_tf(C, B, cc.x, bb.x), tb = collider_tf(C, B, cc.y, bb.y);
                load_side(A, C, hs, cc.x, ta);
                load_side(Bs, C, hs, cc.y, tb);
                const GjkSeeds q = gjk_seeds(ta, tb, guess, tc, 0);
                warm = q.warm;
                plain = q.plain;
                max_it = warm ? GJK_WARM_ITERS : GJK_ITERS;
                gsum = 0;
                fb = 0;
                // the start of gjk(): the first support along the seed
                sx.rank = 0;
                dir = (len2(q.g) > EPS) ? q.g : V3{1.0f, 0.0f, 0.0f};
                const SV v0 = md_sample(A, Bs, hs, dir);
                sx_push(sx, v0);
                dir = -v0.w;
                cyc = GjkCycle<LPP>();
                it = 0;
                active = true;
            }
            next = min(end, next + need);
            if (__ballot(active) == 0ull) break;   // the range is done
            int gs = -1;   // this row's pair ended with status gs (>= 0)
            if (active) {
                // one iteration of gjk()'s loop
                const SV nx = md_sample(A, Bs, hs, dir);
                if (dot(nx.w, dir) < 0.0f) {
                    sx.dir = dir;
                    gs = GJK_VALID;
                } else {
                    sx_push(sx, nx);
                    bool inside = false;
                    if (sx.rank == 2) sx_line(sx, dir);
                    else if (sx.rank == 3) sx_triangle(sx, dir);
                    else inside = sx_tetra(sx, dir);
                    if (inside) gs = GJK_INSIDE;
                    else if (cyc.repeats(sx, dir)) { it = max_it - 1; gs = GJK_FAILED; }
                    else if (it + 1 == max_it) gs = GJK_FAILED;
                }
                ++it;
                if (gs >= 0) {
                    gsum += it;
                    if (warm && gs == GJK_FAILED) {   // the warm-started run failed: the plain run follows
                        fb = 1;
                        warm = false;
                        max_it = GJK_ITERS;
                        sx.rank = 0;
                        dir = (len2(plain) > EPS) ? plain : V3{1.0f, 0.0f, 0.0f};
                        const SV v0 = md_sample(A, Bs, hs, dir);
                        sx_push(sx, v0);
                        dir = -v0.w;
                        cyc = GjkCycle<LPP>();
                        it = 0;
                        gs = -1;
                    }
                }
                if (gs >= 0 && pl == 0) {
                    gstat[s] = pack_gstat(gsum, 0, gs, 15, fb);   // pass 2 completes an overlapping pair
                    if (axis) {
                        const float ln = len(sx.dir);
                        axis[s] = gs == GJK_VALID && ln > 1e-12f ? f4(vdiv(sx.dir, ln), 1.0f)
                                                                 : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    }
                    if (gs != GJK_INSIDE) {
                        K.cnt[s] = 0;
                        if (tlog) {
                            V3 tg{0.0f, 0.0f, 0.0f};
                            float vg = 0.0f;
                            const float ln = len(sx.dir);
                            if (gs == GJK_VALID && ln > 1e-12f) { tg = tf_inv_vec(A.tf, vdiv(-sx.dir, ln)); vg = 1.0f; }
                            tlog[2 * tc] = f4(tg, vg);
                            tlog[2 * tc + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        }
                    }
                }
            }
            // list this iteration's overlapping pairs, in row order
            const uint64_t ins = __ballot(gs == GJK_INSIDE && pl == 0);
            if (gs == GJK_INSIDE && pl == 0) {
                const uint32_t h = rg * GJK_REFILL_SEG + nhit + (uint32_t)__popcll(ins & ((1ull << lane) - 1ull));
                hits[h] = tc;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const SV &v = sx.v[k];
                    sxs[(size_t)h * 8 + 2 * k] = make_float4(v.w.x, v.w.y, v.w.z, v.w0.x);
                    sxs[(size_t)h * 8 + 2 * k + 1] = make_float4(v.w0.y, v.w0.z, 0.0f, 0.0f);
                }
            }
            nhit += (uint32_t)__popcll(ins);
            if (gs >= 0) active = false;
        }
        if (lane == 0) blk_hits[rg] = nhit;
    }
}

// seg: pairs per GJK-pass workgroup (gjk_pass_pairs of the pass that listed the pairs); a
// workgroup takes `sub` of a segment's entries at a time (sub divides seg).
// OCC: waves per EU the compiler budgets registers for (0: no budget).
template <int VPL, int OCC = 0>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(OCC ? OCC : 1))) k_epa_pass(const uint32_t *bin_slot, uint32_t off, uint32_t n, const uint32_t *cand,
                                                 const uint32_t *ncand, uint32_t seg, uint32_t sub, Colliders C,
                                                 Bodies B, Contacts K,
                                                 HullSet hs, uint32_t *gstat, const float4 *guess, int seed_epa,
                                                 float4 *tlog, const uint32_t *tlog_base, const uint32_t *hits,
                                                 const uint32_t *blk_hits, const float4 *sxs, int guard,
                                                 const uint32_t *nflat = nullptr) {
    __shared__ PairLds L;
    const int lane = lane64();
    const uint32_t nc = gjk_count(n, ncand);
    if (tlog && tlog_base) tlog += 2 * (size_t)*tlog_base;
    // work items of `sub` listed pairs: a GJK segment of seg entries splits into seg / sub;
    // with nflat, items are the *nflat entries of one flat list whose simplices are stored
    // per candidate (the pairs the fused pass handed over)
    const uint32_t per = seg / sub, nitems = nflat ? *nflat : (nc + seg - 1) / seg * per;
    for (uint32_t ve = blockIdx.x; ve < nitems; ve += gridDim.x)
    for (uint32_t vb = nflat ? 0u : ve / per, r = nflat ? 0u : (ve % per) * sub,
                  re = nflat ? 1u : min(blk_hits[vb], r + sub); r < re; ++r) {
        const uint32_t h = nflat ? ve : vb * seg + r;
        const uint32_t tc = hits[h];
        const size_t hx = nflat ? (size_t)tc : (size_t)h;   // simplex row
        const uint32_t t = cand ? cand[tc] : tc;
        const uint32_t s = bin_slot[off + t];
        const uint2 cc = K.cc[s], bb = K.bb[s];
        const TF ta = collider_tf(C, B, cc.x, bb.x), tb = collider_tf(C, B, cc.y, bb.y);
        Side<VPL> A, Bs;
        load_side(A, C, hs, cc.x, ta);
        load_side(Bs, C, hs, cc.y, tb);
        Simplex sx;
        sx.rank = 4;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 a = sxs[hx * 8 + 2 * k], b = sxs[hx * 8 + 2 * k + 1];
            sx.v[k].w = V3{a.x, a.y, a.z};
            sx.v[k].w0 = V3{a.w, b.x, b.y};
        }
        const uint32_t g0 = gstat[s];
        int gsum = (int)(g0 & 255u), fb = (int)((g0 >> 24) & 1u);
        int gs = GJK_INSIDE;
        const GjkSeeds q = gjk_seeds(ta, tb, guess, tc, seed_epa);
        const bool warm = q.warm && !fb;   // a pass-1 fallback already ran the plain GJK
        __syncthreads();
        // the seed by value: a pointer into q would keep q in scratch memory
        EpaResult e = epa(sx, A, Bs, hs, L.e, warm && q.seeded, q.sn);
        int esum = e.iters;
        if (warm && e.status != EPA_VALID) {
            fb = 1;
            __syncthreads();
            int it = 0;
            gs = gjk(A, Bs, hs, q.plain, sx, it);
            gsum += it;
            e = EpaResult{EPA_FAILED, 0, V3{0.0f, 0.0f, 0.0f}, V3{0.0f, 0.0f, 0.0f}, 0.0f};
            if (gs == GJK_INSIDE) {
                e = epa(sx, A, Bs, hs, L.e, false, V3{0.0f, 0.0f, 0.0f});
                esum += e.iters;
            }
        }
        const bool ok = gs == GJK_INSIDE && e.status == EPA_VALID;
        // EPA guard: the overlap along any direction bounds the depth from above; an EPA
        // face deeper than the smallest overlap over the guard's directions by more than the
        // tolerance is not the minimum and gives way to that direction.  Directions, first
        // smallest wins: the pair's previous-substep normal (K still holds the previous
        // substep's contacts of this slot; none in a frame's first substep, the pair gather
        // clears them), then with guard 2, while the depth in question exceeds
        // EPA_GUARD_SCAN, A's face normals and B's reversed.  The loops are uniform over the
        // wavefront (one pair).
        if (ok && guard) {
            bool any = false;
            V3 bn{0.0f, 0.0f, 0.0f}, bp{0.0f, 0.0f, 0.0f};
            float bh = 0.0f;
            auto consider = [&](V3 gn) {
                const SV w = md_sample(A, Bs, hs, gn);
                const float hov = dot(w.w, gn);
                if (!any || hov < bh) { any = true; bh = hov; bn = gn; bp = w.w0; }
            };
            if (K.cnt[s] > 0) {
                const float4 hp = hd_row(K, s)[0];
                consider(V3{hp.x, hp.y, hp.z});
            }
            if (guard == 2 && e.depth > EPA_GUARD_SCAN && (!any || bh > EPA_GUARD_SCAN)) {
                const int nfa = face_count(A, hs, C.hull[cc.x]), nfb = face_count(Bs, hs, C.hull[cc.y]);
                for (int k = 0; k < nfa; ++k) consider(tf_vec(A.tf, face_normal_local(A, hs, C.hull[cc.x], k)));
                for (int k = 0; k < nfb; ++k) consider(-tf_vec(Bs.tf, face_normal_local(Bs, hs, C.hull[cc.y], k)));
            }
            if (any && bh < e.depth - EPA_GUARD_TOL) {
                e.n = bn;
                e.depth = bh;
                e.p = bp;
            }
        }
        if (lane == 0) {
            gstat[s] = pack_gstat(gsum, esum, gs, gs == GJK_INSIDE ? e.status : 15, fb);
            if (tlog) {
                V3 tg{0.0f, 0.0f, 0.0f};
                float vg = 0.0f;
                if (ok) { tg = tf_inv_vec(ta, e.n); vg = 1.0f; }
                else if (gs == GJK_VALID) {
                    const float ln = len(sx.dir);
                    if (ln > 1e-12f) { tg = tf_inv_vec(ta, vdiv(-sx.dir, ln)); vg = 1.0f; }
                }
                tlog[2 * tc] = f4(tg, vg);
                tlog[2 * tc + 1] = ok ? f4(tf_inv_vec(ta, e.n), 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        }
        Manifold m;
        m.np = 0;
        if (ok) {
            const bool adj = A.kind == SH_HULL && Bs.kind == SH_HULL && hs.meta[C.hull[cc.x]].adj && hs.meta[C.hull[cc.y]].adj;
            if (adj) {
                __syncthreads();   // the manifold buffers overlay the polytope
                manifold_hulls(e, A, Bs, hs, (uint32_t)C.hull[cc.x], (uint32_t)C.hull[cc.y], L.c, m);
            } else {
                m.n = e.n;
                m.np = 1;
                m.p[0] = e.p;
                m.dep[0] = e.depth;
            }
        }
        if (lane == 0) {
            if (m.np > 0) emit(m, B, K, s, bb);
            else K.cnt[s] = 0;
        }
        __syncthreads();
    }
}

// GJK, EPA and the manifold in one pass, one pair per 16-lane DPP row (four per one-wave
// workgroup, hulls of at most 64 vertices): an overlapping pair goes on to EPA in the row
// that ran its GJK, with the hull vertices, transforms and simplex still in registers and a
// polytope of NV vertices / NF faces and clip buffers of HC points in LDS per row.  The
// steps and their order are those of k_gjk_pass followed by k_epa_pass, so every per-slot
// result is the same.  A pair whose polytope or clip buffer runs out (EPA_DEFER) keeps its
// GJK outcome (status, the simplex stored per candidate) and is listed in dhits (*dcount
// entries, any order) for k_epa_pass in flat-list mode, which runs EPA and the manifold again
// with the reference-sized buffers; nothing else of the pair is written here.
// Row polytope and clip buffer sizes of the fused pass; workgroups of the full-size EPA pass
// that takes the pairs they could not hold.
constexpr int FUSED_NV = 36, FUSED_NF = 72, FUSED_HC = 40;
constexpr uint32_t FUSED_DEFER_GRID = 256;
template <int NV, int NF, int HC, int OCC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(OCC)))
k_gjk_epa_rows(const uint32_t *bin_slot, uint32_t off, uint32_t n, const uint32_t *cand, const uint32_t *ncand, Colliders C,
               Bodies B, Contacts K, HullSet hs, uint32_t *gstat, const float4 *guess, int seed_epa, float4 *tlog,
               const uint32_t *tlog_base, uint32_t *dcount, uint32_t *dhits, float4 *sxs, float4 *axis, int guard) {
    using G = Grp<16>;
    __shared__ PairLdsT<NV, NF, HC> Ls[4];
    const int row = lane64() >> 4, pl = G::lane();
    PairLdsT<NV, NF, HC> &L = Ls[row];
    const uint32_t nc = gjk_count(n, ncand);
    if (tlog && tlog_base) tlog += 2 * (size_t)*tlog_base;
    for (uint32_t vb = blockIdx.x; vb * 4 < nc; vb += gridDim.x) {
        const uint32_t tc = vb * 4 + (uint32_t)row;   // candidate
        if (tc >= nc) continue;
        const uint32_t t = cand ? cand[tc] : tc;   // bin position
        const uint32_t s = bin_slot[off + t];
        const uint2 cc = K.cc[s], bb = K.bb[s];
        const TF ta = collider_tf(C, B, cc.x, bb.x), tb = collider_tf(C, B, cc.y, bb.y);
        Side<4, 16> A, Bs;
        load_side(A, C, hs, cc.x, ta);
        load_side(Bs, C, hs, cc.y, tb);
        const GjkSeeds q = gjk_seeds(ta, tb, guess, tc, seed_epa);
        // GJK (k_gjk_pass)
        Simplex sx;
        int it = 0;
        int gs = gjk(A, Bs, hs, q.g, sx, it, q.warm ? GJK_WARM_ITERS : GJK_ITERS);
        int gsum = it, fb = 0;
        if (q.warm && gs == GJK_FAILED) {
            fb = 1;
            gs = gjk(A, Bs, hs, q.plain, sx, it);
            gsum += it;
        }
        if (pl == 0 && axis) {
            const float ln = len(sx.dir);
            axis[s] = gs == GJK_VALID && ln > 1e-12f ? f4(vdiv(sx.dir, ln), 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        if (gs != GJK_INSIDE) {
            if (pl == 0) {
                gstat[s] = pack_gstat(gsum, 0, gs, 15, fb);
                K.cnt[s] = 0;
                if (tlog) {
                    V3 tg{0.0f, 0.0f, 0.0f};
                    float vg = 0.0f;
                    const float ln = len(sx.dir);
                    if (gs == GJK_VALID && ln > 1e-12f) { tg = tf_inv_vec(ta, vdiv(-sx.dir, ln)); vg = 1.0f; }
                    tlog[2 * tc] = f4(tg, vg);
                    tlog[2 * tc + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
            }
            continue;
        }
        if (pl == 0) {   // the simplex, for a pair handed over (constant indices keep it in registers)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const SV &v = sx.v[k];
                sxs[(size_t)tc * 8 + 2 * k] = make_float4(v.w.x, v.w.y, v.w.z, v.w0.x);
                sxs[(size_t)tc * 8 + 2 * k + 1] = make_float4(v.w0.y, v.w0.z, 0.0f, 0.0f);
            }
        }
        // EPA and the manifold (k_epa_pass)
        const uint32_t gsum1 = (uint32_t)gsum, fb1 = (uint32_t)fb;
        const bool warm = q.warm && !fb;   // a fallback already ran the plain GJK
        bool defer = false;
        G::sync();
        EpaResult e = epa(sx, A, Bs, hs, L.e, warm && q.seeded, q.sn);
        int esum = e.iters;
        defer = e.status == EPA_DEFER;
        if (!defer && warm && e.status != EPA_VALID) {
            fb = 1;
            G::sync();
            gs = gjk(A, Bs, hs, q.plain, sx, it);
            gsum += it;
            e = EpaResult{EPA_FAILED, 0, V3{0.0f, 0.0f, 0.0f}, V3{0.0f, 0.0f, 0.0f}, 0.0f};
            if (gs == GJK_INSIDE) {
                e = epa(sx, A, Bs, hs, L.e, false, V3{0.0f, 0.0f, 0.0f});
                esum += e.iters;
                defer = e.status == EPA_DEFER;
            }
        }
        const bool ok = !defer && gs == GJK_INSIDE && e.status == EPA_VALID;
        if (ok && guard) {   // the EPA guard of k_epa_pass
            bool any = false;
            V3 bn{0.0f, 0.0f, 0.0f}, bp{0.0f, 0.0f, 0.0f};
            float bh = 0.0f;
            auto consider = [&](V3 gn) {
                const SV w = md_sample(A, Bs, hs, gn);
                const float hov = dot(w.w, gn);
                if (!any || hov < bh) { any = true; bh = hov; bn = gn; bp = w.w0; }
            };
            if (K.cnt[s] > 0) {
                const float4 hp = hd_row(K, s)[0];
                consider(V3{hp.x, hp.y, hp.z});
            }
            if (guard == 2 && e.depth > EPA_GUARD_SCAN && (!any || bh > EPA_GUARD_SCAN)) {
                const int nfa = face_count(A, hs, C.hull[cc.x]), nfb = face_count(Bs, hs, C.hull[cc.y]);
                for (int k = 0; k < nfa; ++k) consider(tf_vec(A.tf, face_normal_local(A, hs, C.hull[cc.x], k)));
                for (int k = 0; k < nfb; ++k) consider(-tf_vec(Bs.tf, face_normal_local(Bs, hs, C.hull[cc.y], k)));
            }
            if (any && bh < e.depth - EPA_GUARD_TOL) {
                e.n = bn;
                e.depth = bh;
                e.p = bp;
            }
        }
        Manifold m;
        m.np = 0;
        if (ok) {
            const bool adj = A.kind == SH_HULL && Bs.kind == SH_HULL && hs.meta[C.hull[cc.x]].adj && hs.meta[C.hull[cc.y]].adj;
            if (adj) {
                G::sync();   // the manifold buffers overlay the polytope
                defer = !manifold_hulls(e, A, Bs, hs, (uint32_t)C.hull[cc.x], (uint32_t)C.hull[cc.y], L.c, m);
                defer = G::ballot(defer) != 0ull;   // lane 0 decides
            } else {
                m.n = e.n;
                m.np = 1;
                m.p[0] = e.p;
                m.dep[0] = e.depth;
            }
        }
        if (pl == 0) {
            if (defer) {
                gstat[s] = pack_gstat((int)gsum1, 0, GJK_INSIDE, 15, (int)fb1);
                dhits[atomicAdd(dcount, 1u)] = tc;
            } else {
                gstat[s] = pack_gstat(gsum, esum, gs, gs == GJK_INSIDE ? e.status : 15, fb);
                if (tlog) {
                    V3 tg{0.0f, 0.0f, 0.0f};
                    float vg = 0.0f;
                    if (ok) { tg = tf_inv_vec(ta, e.n); vg = 1.0f; }
                    else if (gs == GJK_VALID) {
                        const float ln = len(sx.dir);
                        if (ln > 1e-12f) { tg = tf_inv_vec(ta, vdiv(-sx.dir, ln)); vg = 1.0f; }
                    }
                    tlog[2 * tc] = f4(tg, vg);
                    tlog[2 * tc + 1] = ok ? f4(tf_inv_vec(ta, e.n), 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
                if (m.np > 0) emit(m, B, K, s, bb);
                else K.cnt[s] = 0;
            }
        }
        G::sync();   // the next pair of this row rewrites its LDS
    }
}

// Totals over the GJK pairs bin_slot[off .. off + n) from gstat: pairs, GJK iterations,
// EPA iterations, overlapping pairs, EPA converged, fallback runs.
// Extent of a collider beyond its origin along the unit direction u (world frame): the
// support distance max over the shape of (p - origin) . u; a hull takes it over its vertices,
// or (bound_only) over the corners of its local AABB, an upper bound.
__device__ __forceinline__ float support_extent(const Colliders &C, const HullSet &hs, uint32_t c, const TF &tf, V3 u,
                                                bool bound_only = false) {
    const int4 m = C.meta[c];
    const float4 dm = C.dims[c];
    const V3 l = tf_inv_vec(tf, u);
    if (m.x == SH_SPHERE) return dm.x;
    if (m.x == SH_BOX) return fabsf(l.x) * dm.x + fabsf(l.y) * dm.y + fabsf(l.z) * dm.z;
    if (m.x == SH_CAPSULE) return fabsf(l.y) * dm.x + dm.y;
    const uint32_t h = (uint32_t)C.hull[c];
    if (bound_only) {   // the smaller of the local AABB's and the bounding ellipsoid's extents
        const float4 lo = hs.amin[h], hi = hs.amax[h];
        const float box = (l.x >= 0.0f ? hi.x * l.x : lo.x * l.x) + (l.y >= 0.0f ? hi.y * l.y : lo.y * l.y) +
                          (l.z >= 0.0f ? hi.z * l.z : lo.z * l.z);
        if (!hs.ell) return box;
        const float4 c = hs.ell[2 * h], a = hs.ell[2 * h + 1];
        const float ax = a.x * l.x, ay = a.y * l.y, az = a.z * l.z;
        return fminf(box, (c.x * l.x + c.y * l.y + c.z * l.z) + sqrtf(ax * ax + ay * ay + az * az));
    }
    const HullMeta hm = hs.meta[h];
    float e = -3.4e38f;
    for (uint32_t k = 0; k < hm.nv; ++k) {
        const float4 v = hs.v[hm.v0 + k];
        e = fmaxf(e, v.x * l.x + v.y * l.y + v.z * l.z);
    }
    return e;
}

// Separating-axis hints across a frame's new pair list: the old list's hints of its GJK
// slots go into an open-addressing table keyed by collider pair (linear probing, at most
// half full), and the new list's GJK slots look their pair up (a slot whose pair is absent
// starts without a hint; a hint proves nothing by itself: k_gjk_cull tests it exactly).
// The slot of a pair changes with the sweep order, so without this the first substep of a
// frame would mostly test other pairs' axes.
constexpr unsigned long long AXIS_EMPTY = ~0ull;
__device__ __forceinline__ uint32_t axis_hash(unsigned long long k, uint32_t mask) {
    // MurmurHash3's 64-bit finaliser
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return (uint32_t)k & mask;
}
__device__ __forceinline__ unsigned long long axis_key(const Contacts &K, uint32_t s) {
    const uint2 cc = K.cc[s];
    return ((unsigned long long)cc.x << 32) | cc.y;
}
__global__ void k_axis_clear(unsigned long long *keys, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = AXIS_EMPTY;
}
// slots: the old list's GJK slots (each < cap, the pair buffers' capacity; others are skipped)
__global__ void k_axis_save(const uint32_t *slots, uint32_t n, uint32_t cap, Contacts K, const float4 *axis,
                            unsigned long long *keys, float4 *vals, uint32_t mask) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = slots[i];
    if (s >= cap) return;
    const float4 a = axis[s];
    if (a.w == 0.0f) return;
    const unsigned long long key = axis_key(K, s);
    uint32_t h = axis_hash(key, mask);
    for (uint32_t p = 0; p <= mask; ++p, h = (h + 1u) & mask) {
        const unsigned long long prev = atomicCAS(&keys[h], AXIS_EMPTY, key);
        if (prev == AXIS_EMPTY || prev == key) {
            vals[h] = a;
            return;
        }
    }
}
// Every slot of the new list gets its pair's saved hint, or none (w = 0): a slot's hint then
// depends only on the table, never on which pair held the slot before.  keys == nullptr: no
// table (nothing carried), every slot starts without a hint.
__global__ void k_axis_load(const uint32_t *slots, uint32_t n, Contacts K, const unsigned long long *keys,
                            const float4 *vals, uint32_t mask, float4 *axis) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = slots[i];
    float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (keys) {
        const unsigned long long key = axis_key(K, s);
        uint32_t h = axis_hash(key, mask);
        for (uint32_t p = 0; p <= mask; ++p, h = (h + 1u) & mask) {
            const unsigned long long k = keys[h];
            if (k == key) {
                a = vals[h];
                break;
            }
            if (k == AXIS_EMPTY) break;
        }
    }
    axis[s] = a;
}

// Exact rejection in front of GJK.  Two tests on the line between the collider origins:
// bounding spheres (radius about each origin in lpos.w, rounded up) apart, or the two
// shapes' support extents along that line apart (a separating plane).  Either way the
// shapes cannot overlap, so plain GJK would end separated and emit nothing; the pair is
// finished here (no contact, status separated after 0 iterations, not logged for
// training) and the others are flagged for k_gjk_compact.  The margin (1e-4 relative +
// 1e-3 m + 1e-6 of the largest coordinate) covers the FP32 rounding of the centres and of
// the tests themselves.
// A third test re-checks the axis that separated the pair the last time GJK ran on its
// slot (`axis`, world frame, written by k_gjk_pass): bodies move little between substeps, so
// that axis usually still separates.  Any axis that separates the current poses proves the
// pair apart, however old or whichever pair it came from, so the cache needs no clearing.
//
// Work split for the wavefront: every lane first runs the cheap tests of its own pair (the
// bounding spheres, and each axis against the hulls' local-AABB bound of their extent).  The
// exact extents of hulls (a maximum over their vertices) then run for the lanes still open,
// four pairs at a time, one 16-lane row each (each lane a quarter of the vertices, a DPP row
// maximum): a per-lane vertex loop left ~87 % of the lanes idle.  The tests and their
// outcome are the same either way (a maximum does not depend on the order it is taken in).
struct CullPair {
    uint32_t s;
    uint2 cc;
    TF ta, tb;
    V3 d;
    float d2, slack;
};
__device__ __forceinline__ CullPair cull_pair(const uint32_t *bin_slot, uint32_t off, uint32_t i, const Colliders &C,
                                              const Bodies &B, const Contacts &K, const uint4 *rec, const uint32_t *recb) {
    CullPair p;
    uint2 bb;
    if (rec) {   // the frame's record: slot, colliders and bodies in one round of loads
        const uint4 r = rec[i];
        p.s = r.x;
        p.cc = make_uint2(r.y, r.z);
        bb = make_uint2(r.w, recb[i]);
    } else {
        p.s = bin_slot[off + i];
        p.cc = K.cc[p.s];
        bb = K.bb[p.s];
    }
    p.ta = collider_tf(C, B, p.cc.x, bb.x);
    p.tb = collider_tf(C, B, p.cc.y, bb.y);
    p.d = p.tb.pos - p.ta.pos;
    const float big = fmaxf(fmaxf(fmaxf(fabsf(p.ta.pos.x), fabsf(p.ta.pos.y)), fmaxf(fabsf(p.ta.pos.z), fabsf(p.tb.pos.x))),
                            fmaxf(fabsf(p.tb.pos.y), fabsf(p.tb.pos.z)));
    p.slack = 1e-3f + 1e-6f * big;
    p.d2 = len2(p.d);
    return p;
}
__device__ __forceinline__ float dpp_row_fmax(float v) {
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, false)));
    return v;
}
// support_extent of one side, a hull's maximum taken by the 16 lanes of a row (all active).
__device__ __forceinline__ float row_support_extent(const Colliders &C, const HullSet &hs, uint32_t c, const TF &tf, V3 u) {
    const int4 m = C.meta[c];
    if (m.x != SH_HULL) return support_extent(C, hs, c, tf, u);
    const V3 l = tf_inv_vec(tf, u);
    const HullMeta hm = hs.meta[(uint32_t)C.hull[c]];
    float e = -3.4e38f;
    for (uint32_t k = (uint32_t)(lane64() & 15); k < hm.nv; k += 16) {
        const float4 v = hs.v[hm.v0 + k];
        e = fmaxf(e, v.x * l.x + v.y * l.y + v.z * l.z);
    }
    return dpp_row_fmax(e);
}
// The frame's GJK-bin records (per bin position: slot, colliders, bodies), written once per
// frame after the pair list is binned, so that each substep's rejection pass reaches the poses
// after one round of loads instead of two.
__global__ void k_gjk_records(const uint32_t *bin_slot, uint32_t n, Contacts K, uint4 *rec, uint32_t *recb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = bin_slot[i];
    const uint2 cc = K.cc[s], bb = K.bb[s];
    rec[i] = make_uint4(s, cc.x, cc.y, bb.x);
    recb[i] = bb.y;
}
// keep[i]: the pair stays a GJK candidate; bsum (optional): kept pairs per 256-pair workgroup.
__global__ void __launch_bounds__(256) k_gjk_cull(const uint32_t *bin_slot, uint32_t off, uint32_t n, Colliders C, Bodies B,
                                                  Contacts K, HullSet hs, const float4 *axis, uint32_t *gstat, uint32_t *keep,
                                                  uint32_t *bsum, const uint4 *rec = nullptr, const uint32_t *recb = nullptr) {
    __shared__ uint32_t bkept;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = lane64(), row = lane >> 4, pl = lane & 15;
    if (threadIdx.x == 0) bkept = 0u;
    __syncthreads();
    bool need = false, kept = false;
    CullPair p{};   // this lane's pair (the exact stage takes the open pairs' from their lanes)
    if (i < n) {
        p = cull_pair(bin_slot, off, i, C, B, K, rec, recb);
        const float ra = C.lpos[p.cc.x].w, rb = C.lpos[p.cc.y].w;
        const float rs = (ra + rb) * 1.0001f + p.slack;
        bool apart = p.d2 > rs * rs;
        auto bound_apart = [&](V3 u, float gap) {
            const float ext = support_extent(C, hs, p.cc.x, p.ta, u, true) + support_extent(C, hs, p.cc.y, p.tb, -u, true);
            return ext * 1.0001f + p.slack < gap;
        };
        bool axes = false;
        if (!apart && axis) {
            const float4 a = axis[p.s];
            if (a.w != 0.0f) {
                const V3 u{a.x, a.y, a.z};
                apart = bound_apart(u, dot(p.d, u));
                axes = true;
            }
        }
        if (!apart && p.d2 > 1e-8f) {
            const float dl = sqrtf(p.d2);
            apart = bound_apart(vdiv(p.d, dl), dl);
            axes = true;
        }
        // the exact extents can only differ from the bounds for hulls
        need = !apart && axes && (C.meta[p.cc.x].x == SH_HULL || C.meta[p.cc.y].x == SH_HULL);
        if (!need) {
            keep[i] = apart ? 0u : 1u;
            kept = !apart;
            if (apart) {
                K.cnt[p.s] = 0;
                gstat[p.s] = pack_gstat(0, 0, GJK_VALID, 15, 0);
            }
        }
    }
    {
        const uint64_t km = __ballot(kept);
        if (bsum && lane == 0 && km) atomicAdd(&bkept, (uint32_t)__popcll(km));
    }
    // exact extents of the open pairs, a row per pair
    uint64_t todo = __ballot(need);
    while (todo) {   // uniform over the wavefront
        uint64_t m = todo;
        for (int k = 0; k < row; ++k) m &= m - 1ull;
        for (int k = 0; k < 4; ++k) todo &= todo - 1ull;
        // the open pair as its lane computed it, shuffled (not loaded again) while every lane of
        // the wavefront is active; a row without a pair this round takes lane 0's and waits
        const int src = m ? __ffsll((long long)m) - 1 : 0;
        auto shf = [&](float x) { return __shfl(x, src, 64); };
        auto shu = [&](uint32_t x) { return (uint32_t)__shfl((int)x, src, 64); };
        CullPair q;
        q.s = shu(p.s);
        q.cc = make_uint2(shu(p.cc.x), shu(p.cc.y));
        q.ta.pos = V3{shf(p.ta.pos.x), shf(p.ta.pos.y), shf(p.ta.pos.z)};
        q.ta.rot = Q{shf(p.ta.rot.w), shf(p.ta.rot.x), shf(p.ta.rot.y), shf(p.ta.rot.z)};
        q.tb.pos = V3{shf(p.tb.pos.x), shf(p.tb.pos.y), shf(p.tb.pos.z)};
        q.tb.rot = Q{shf(p.tb.rot.w), shf(p.tb.rot.x), shf(p.tb.rot.y), shf(p.tb.rot.z)};
        q.d = V3{shf(p.d.x), shf(p.d.y), shf(p.d.z)};
        q.d2 = shf(p.d2);
        q.slack = shf(p.slack);
        if (m == 0ull) continue;   // fewer than four pairs left: this row waits (whole row inactive)
        const uint32_t j = blockIdx.x * blockDim.x + (threadIdx.x & ~63u) + (uint32_t)src;   // the open pair's index
        auto exact_apart = [&](V3 u, float gap) {
            const float ext = row_support_extent(C, hs, q.cc.x, q.ta, u) + row_support_extent(C, hs, q.cc.y, q.tb, -u);
            return ext * 1.0001f + q.slack < gap;
        };
        bool apart = false;
        if (axis) {
            const float4 a = axis[q.s];
            if (a.w != 0.0f) {
                const V3 u{a.x, a.y, a.z};
                apart = exact_apart(u, dot(q.d, u));
            }
        }
        if (!apart && q.d2 > 1e-8f) {
            const float dl = sqrtf(q.d2);
            apart = exact_apart(vdiv(q.d, dl), dl);
        }
        if (pl == 0) {
            keep[j] = apart ? 0u : 1u;
            if (apart) {
                K.cnt[q.s] = 0;
                gstat[q.s] = pack_gstat(0, 0, GJK_VALID, 15, 0);
            } else if (bsum) {
                atomicAdd(&bkept, 1u);
            }
        }
    }
    if (bsum) {
        __syncthreads();
        if (threadIdx.x == 0) bsum[blockIdx.x] = bkept;
    }
}

// Algorithmic bytes of the rejection test per pair: slot, colliders and bodies 20 B, per side
// collider offset/rotation 32 B and body pose 32 B, the flag 4 B (a rejected pair also
// writes its count and status, 8 B).
constexpr double GJK_CULL_BYTES = 20.0 + 2.0 * 64.0 + 4.0 + 8.0;

__global__ void k_add_count(const uint32_t *n, unsigned long long *acc) { atomicAdd(acc, (unsigned long long)*n); }
// *cnt += the candidate count (or n_host without a list): rows appended to the training log.
__global__ void k_log_advance(uint32_t *cnt, const uint32_t *n_dev, uint32_t n_host) {
    *cnt += n_dev ? min(*n_dev, n_host) : n_host;
}

// Stable compaction of the kept bin positions (exclusive scan `pos` of `keep`).
__global__ void k_gjk_compact(const uint32_t *keep, const uint32_t *pos, uint32_t n, uint32_t *cand) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && keep[i]) cand[pos[i]] = i;
}
// The same from the cull's per-workgroup counts (boff: their exclusive scan; 256-pair
// workgroups as in k_gjk_cull): the rank inside the workgroup from the wavefronts' ballots.
__global__ void __launch_bounds__(256) k_gjk_compact_blk(const uint32_t *keep, const uint32_t *boff, uint32_t n,
                                                         uint32_t *cand) {
    __shared__ uint32_t ws[4];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x, w = threadIdx.x >> 6;
    const bool k = i < n && keep[i];
    const uint64_t m = __ballot(k);
    if ((threadIdx.x & 63u) == 0u) ws[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = boff[blockIdx.x];
    for (uint32_t u = 0; u < w; ++u) base += ws[u];
    if (k) cand[base + lane_rank(m)] = i;
}
// The same compaction with a sort key per candidate from its slot's previous substep (a
// frame's first substep reads the slots' old contents): mode 1 the GJK iterations (capped at
// 15), so the four pairs of a wavefront need similar iteration counts; mode 2 whether GJK
// found the pair overlapping, so the rows of a wavefront mostly all run EPA or none does;
// mode 3 both.  The order of the candidates changes no result.
__global__ void k_gjk_compact_keyed(const uint32_t *keep, const uint32_t *pos, uint32_t n, const uint32_t *bin_slot,
                                    uint32_t off, const uint32_t *gstat, uint32_t *cand, uint32_t *key, int mode) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && keep[i]) {
        const uint32_t p = pos[i];
        const uint32_t g = gstat[bin_slot[off + i]];
        const uint32_t in = ((g >> 16) & 15u) == (uint32_t)GJK_INSIDE ? 1u : 0u;
        cand[p] = i;
        key[p] = mode == 1 ? min(g & 255u, 15u) : mode == 2 ? in : (in << 4) | min(g & 255u, 15u);
    }
}

// A grid of at most GJK_TOTALS_BLOCKS workgroups of 256 threads strides over the pairs; one
// atomic per counter per workgroup (per-wavefront atomics on six addresses serialised into
// milliseconds per frame).
constexpr uint32_t GJK_TOTALS_BLOCKS = 256;
__global__ void __launch_bounds__(256) k_gjk_totals(const uint32_t *bin_slot, uint32_t off, uint32_t n, const uint32_t *gstat,
                                                    unsigned long long *tot) {
    __shared__ uint32_t part[4][6];
    uint32_t v[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t g = gstat[bin_slot[off + i]];
        const uint32_t gs = (g >> 16) & 15u, es = (g >> 20) & 15u;
        v[0] += 1u; v[1] += g & 255u; v[2] += (g >> 8) & 255u; v[3] += gs == GJK_INSIDE;
        v[4] += gs == GJK_INSIDE && es == EPA_VALID; v[5] += (g >> 24) & 1u;
    }
    for (int k = 0; k < 6; ++k) {
        const uint32_t x = wave_sum(v[k]);
        if (lane64() == 0) part[threadIdx.x >> 6][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const unsigned long long x = (unsigned long long)part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                                     part[3][threadIdx.x];
        if (x) atomicAdd(&tot[threadIdx.x], x);
    }
}

// Hull against a plane (PlaneCollision.hpp:90-285): deepest support, then up to 32
// vertices within 1e-4 of the plane in vertex order, reduced to 4 by area.
template <int VPL>
__global__ void __launch_bounds__(64) k_narrow_hull_plane(const uint32_t *bin_slot, uint32_t off, Colliders C, Bodies B,
                                                          Contacts K, HullSet hs) {
    __shared__ float tx[HP_MAX_POINTS], ty[HP_MAX_POINTS], tz[HP_MAX_POINTS], td[HP_MAX_POINTS];
    const int lane = lane64();
    const uint32_t s = bin_slot[off + blockIdx.x];
    const uint2 cc = K.cc[s], bb = K.bb[s];
    const TF ta = collider_tf(C, B, cc.x, bb.x), tp = collider_tf(C, B, cc.y, bb.y);
    WaveHull<VPL> wh;
    load_wave_hull(wh, hs, (uint32_t)C.hull[cc.x]);
    const float4 pd = C.dims[cc.y];
    const V3 wn = tf_vec(tp, V3{pd.x, pd.y, pd.z});   // Plane.hpp:47-57
    const float wd = pd.w + dot(wn, tp.pos);
    const V3 ln = tf_inv_vec(ta, wn);
    const float ld = wd - dot(ta.pos, wn);
    const V3 deep = hull_vertex(hs, wh, wave_support_index(wh, -ln));
    if (dot(deep, ln) - ld > 1e-4f) {
        if (lane == 0) K.cnt[s] = 0;
        return;
    }
    int tc = 0;
#pragma unroll
    for (int k = 0; k < VPL; ++k) {
        const int i = lane + 64 * k;
        float dist = 0.0f;
        bool pen = false;
        if (i < wh.nv) {
            dist = wh.x[k] * ln.x + wh.y[k] * ln.y + wh.z[k] * ln.z - ld;
            pen = dist <= 1e-4f;
        }
        const uint64_t m = __ballot(pen);
        const int slot = tc + (int)lane_rank(m);
        if (pen && slot < HP_MAX_POINTS) {
            tx[slot] = wh.x[k]; ty[slot] = wh.y[k]; tz[slot] = wh.z[k]; td[slot] = -dist;
        }
        tc += __popcll(m);
    }
    if (tc > HP_MAX_POINTS) tc = HP_MAX_POINTS;
    __syncthreads();
    if (lane != 0) return;
    if (tc == 0) { K.cnt[s] = 0; return; }
    auto P = [&](int i) { return V3{tx[i], ty[i], tz[i]}; };
    Manifold m;
    m.n = -wn;
    if (tc <= 4) {
        for (int i = 0; i < tc; ++i) { m.p[i] = tf_point(ta, P(i)); m.dep[i] = td[i]; }
        m.np = tc;
        emit(m, B, K, s, bb);
        return;
    }
    const float low = -1e18f;
    int idx[4] = {-1, -1, -1, -1};
    float best = low;
    for (int i = 0; i < tc; ++i) if (td[i] > best) { best = td[i]; idx[0] = i; }
    best = low;
    for (int i = 0; i < tc; ++i) {
        if (i == idx[0]) continue;
        const V3 d = P(i) - P(idx[0]);
        const float x = dot(d, d);
        if (x > best) { best = x; idx[1] = i; }
    }
    best = low;
    {
        const V3 e = P(idx[1]) - P(idx[0]);
        for (int i = 0; i < tc; ++i) {
            if (i == idx[0] || i == idx[1]) continue;
            const V3 c = cross(e, P(i) - P(idx[0]));
            const float x = dot(c, c);
            if (x > best) { best = x; idx[2] = i; }
        }
    }
    best = low;
    {
        const V3 p0 = P(idx[0]), p1 = P(idx[1]), p2 = P(idx[2]);
        const V3 e01 = p1 - p0, e12 = p2 - p1, e20 = p0 - p2;
        for (int i = 0; i < tc; ++i) {
            if (i == idx[0] || i == idx[1] || i == idx[2]) continue;
            const V3 x = P(i);
            const float a0 = dot(cross(e01, x - p0), ln);
            const float a1 = dot(cross(e12, x - p1), ln);
            const float a2 = dot(cross(e20, x - p2), ln);
            float sc = a0;
            if (a1 > sc) sc = a1;
            if (a2 > sc) sc = a2;
            if (sc > best) { best = sc; idx[3] = i; }
        }
    }
    const int ne = (idx[3] >= 0) ? 4 : 3;
    for (int i = 0; i < ne; ++i) { m.p[i] = tf_point(ta, P(idx[i])); m.dep[i] = td[idx[i]]; }
    m.np = ne;
    emit(m, B, K, s, bb);
}

}  // namespace rb
