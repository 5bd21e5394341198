// Joints on the GPU: error evaluation, XPBD row solve and velocity damping.
//
// Reference: src/rbps/constraints/Joint.cpp (errors 18-182, damping 184-220),
// src/rbps/constraints/Constraint.cpp (rows 5-46), include/rbps/API/JointAPI.hpp
// (rows per joint type 160-265), World.hpp:166-178 (order inside a substep).
//
// A joint is one unit of work: one thread walks its 1-4 rows in order, exactly as the
// reference walks a joint's consecutive constraint rows.  Joints sharing a dynamic body
// are separated by a colouring computed once when the joints are uploaded; one launch
// per colour.  All errors of a substep are evaluated before any row is solved (the
// reference's compute_joint_errors pass precedes solve_constraints).
//
// Layout (one row per joint, 16-byte rows):
//   jm   int4   (type | actuation << 8 | limited << 16 | rows << 24, body 1, body 2, -)
//   jr1  float4 (r_1 local, damping)      jr2  float4 (r_2 local, compliance)
//   jax  float4 (axis, lower limit)        jlx  float4 (limit axis, upper limit)
//   jtg  float4 (target position, target speed, current position, -)
//   jrow float4 x 4 per joint (direction, magnitude), jlam float x 4 per joint (lambda,
//   never reset, as in the reference).
// Row kinds per type (JointAPI.hpp): prismatic R P P(compliance), revolute R P R R(c),
// fixed R P; spherical (this build's extension): P(c) and, when limited, a swing-limit R.
// A constraint row made by World::create_constraint (ConstraintAPI.hpp:37-64) is a
// one-row unit of type ROW (actuation = its kind, 1 ROTATIONAL): nothing evaluates its
// errors, it keeps the direction and magnitude its caller writes.
#pragma once
#include "rbgpu_math.hpp"

namespace rb {

enum JointType { JT_PRISMATIC = 0, JT_REVOLUTE = 1, JT_FIXED = 2, JT_SPHERICAL = 3, JT_ROW = 4 };
enum JointAct { JA_FREE = 0, JA_POSITION = 1, JA_SPEED = 2 };

struct Joints {
    int4 *jm;
    float4 *jr1, *jr2, *jax, *jlx, *jtg;
    float4 *jrow;
    float *jlam;
};

__device__ __forceinline__ int jt_type(int4 m) { return m.x & 0xff; }
__device__ __forceinline__ int jt_act(int4 m) { return (m.x >> 8) & 0xff; }
__device__ __forceinline__ bool jt_limited(int4 m) { return ((m.x >> 16) & 0xff) != 0; }
__device__ __forceinline__ int jt_rows(int4 m) { return (m.x >> 24) & 0xff; }
// ROTATIONAL row?  (POSITIONAL otherwise)
__device__ __forceinline__ bool jt_rot(int4 m, int k) {
    switch (jt_type(m)) {
        case JT_PRISMATIC: return k == 0;
        case JT_REVOLUTE: return k != 1;
        case JT_FIXED: return k == 0;
        case JT_ROW: return jt_act(m) == 1;
        default: return k == 1;
    }
}
// The row carrying the joint's compliance (the others have 0).
__device__ __forceinline__ int jt_comp_row(int type) {
    return type == JT_PRISMATIC ? 2 : (type == JT_REVOLUTE ? 3 : (type == JT_SPHERICAL || type == JT_ROW ? 0 : -1));
}

// Joint solver modes.
//   0 (the reference): errors of every joint evaluated once per substep before any row is
//     solved (World.hpp:168-170); lambda never reset; row direction normalize(v)
//     (Constraint.cpp:5-12), which is zero for errors in (1e-9, ~3.2e-5) although the row is
//     kept, so a ROTATIONAL row with compliance 0 then divides 0 by 0; revolute alignment
//     a1 x a2 (Joint.cpp:101), whose correction turns the axes apart.
//   1 (XPBD, Mueller et al. 2020): each row's error re-evaluated from the current state just
//     before it is solved, lambda reset every substep, direction v/|v| whenever |v| > EPS,
//     revolute alignment a2 x a1.
__device__ __forceinline__ float4 row_value(V3 v, int mode) {
    const float m = len(v);
    V3 d{1.0f, 0.0f, 0.0f};
    if (m > EPS) d = mode == 0 ? normalize(v) : v * (1.0f / sqrtf(len2(v)));
    return make_float4(d.x, d.y, d.z, m);
}

// row[k] for a run-time k by selects: constant indices keep `row` in registers, where a
// run-time index would put it in scratch memory.
__device__ __forceinline__ float4 row_at(const float4 row[4], int k) {
    return k == 0 ? row[0] : k == 1 ? row[1] : k == 2 ? row[2] : row[3];
}

// Rows of one joint from the two bodies' poses (Joint.cpp:18-182; spherical: attachment row
// and a swing limit built like the revolute angle limit, Joint.cpp:73-90).  `advance`
// moves a speed-driven target by speed * h.  tg.z receives the joint's current position.
__device__ __forceinline__ void joint_rows(int4 m, float4 r1, float4 r2, float4 ax, float4 lx, float4 &tg, V3 x1, Q q1,
                                           V3 x2, Q q2, float h, bool advance, int mode, float4 row[4]) {
    // the rows are set in locals and stored once at the end: stores to different rows in
    // different branches could otherwise be merged into one store at a run-time index,
    // which puts `row` in scratch memory
    const int type = jt_type(m), act = jt_act(m);
    if (type == JT_ROW) return;   // a constraint row keeps what its caller wrote
    float4 o0 = row[0], o1 = row[1], o2 = row[2], o3 = row[3];
    const V3 p1 = x1 + rotate(q1, xyz(r1)), p2 = x2 + rotate(q2, xyz(r2));
    if (type == JT_PRISMATIC || type == JT_FIXED) {   // Joint.cpp:27-29, 170-172
        const Q dq = qmul(q1, conj(q2));
        o0 = row_value(V3{dq.x, dq.y, dq.z} * 2.0f, mode);
    }
    if (type == JT_FIXED) {
        o1 = row_value(p1 - p2, mode);
    } else if (type == JT_PRISMATIC) {   // Joint.cpp:32-70
        const V3 dp = p1 - p2;
        const V3 axis = rotate(q1, xyz(ax));
        const float ps = dot(dp, axis);
        float over = 0.0f;
        if (jt_limited(m)) over = ps - clampf(ps, ax.w, lx.w);
        o1 = row_value((dp - axis * ps) + axis * over, mode);
        V3 drive{0.0f, 0.0f, 0.0f};
        if (act == JA_SPEED && advance) tg.x = tg.x + tg.y * h;
        if (act == JA_SPEED || act == JA_POSITION) {
            const float curr = dot(p2 - p1, axis);
            drive = axis * (tg.x - curr);
        }
        o2 = row_value(drive, mode);
        tg.z = ps;
    } else if (type == JT_REVOLUTE) {   // Joint.cpp:92-159
        const V3 a1 = rotate(q1, xyz(ax)), a2 = rotate(q2, xyz(ax));
        o0 = row_value(mode == 0 ? cross(a1, a2) : cross(a2, a1), mode);
        o1 = row_value(p1 - p2, mode);
        const V3 l1 = rotate(q1, xyz(lx)), l2 = rotate(q2, xyz(lx));
        float phi = rb_atan2(dot(cross(l1, l2), a1), dot(l1, l2));
        V3 lim{0.0f, 0.0f, 0.0f};
        if (jt_limited(m) && (phi < ax.w || phi > lx.w)) {
            phi = clampf(phi, ax.w, lx.w);
            const V3 n1 = rotate(from_axis_angle(a1, phi), l1);
            lim = cross(n1, l2);
        }
        o2 = row_value(-lim, mode);
        V3 drive{0.0f, 0.0f, 0.0f};
        if (act == JA_SPEED && advance) tg.x = tg.x + tg.y * h;
        if (act == JA_SPEED || act == JA_POSITION) {
            const float theta = tg.x + 3.14159265358979323846f;
            const V3 bt = (l1 * rb_cos(theta) + cross(a1, l1) * rb_sin(theta)) + a1 * dot(a1, l1) * (1.0f - rb_cos(theta));
            drive = cross(bt, l2);
        }
        o3 = row_value(drive, mode);
        tg.z = phi;
    } else {   // spherical
        o0 = row_value(p1 - p2, mode);
        if (jt_limited(m)) {
            const V3 a1 = rotate(q1, xyz(ax)), a2 = rotate(q2, xyz(ax));
            const V3 c = cross(a1, a2);
            const float th = rb_atan2(len(c), dot(a1, a2));
            V3 lim{0.0f, 0.0f, 0.0f};
            if (th > lx.w) {
                const V3 n1 = rotate(from_axis_angle(c, lx.w), a1);
                lim = cross(n1, a2);
            }
            o1 = row_value(-lim, mode);
            tg.z = th;
        }
    }
    row[0] = o0;
    row[1] = o1;
    row[2] = o2;
    row[3] = o3;
}

// Row k of joint_rows alone, from the same expressions in the same order (the XPBD mode
// re-evaluates one row just before solving it; evaluating all four there repeated the angle
// functions four times per row).  With `cur`, tg.z also receives the joint's current
// position as joint_rows leaves it.  A constraint row (type ROW) keeps `old`.
__device__ __forceinline__ float4 joint_row(int4 m, float4 r1, float4 r2, float4 ax, float4 lx, float4 &tg, V3 x1, Q q1,
                                            V3 x2, Q q2, int mode, int k, bool cur, float4 old) {
    const int type = jt_type(m), act = jt_act(m);
    if (type == JT_ROW) return old;
    if ((type == JT_PRISMATIC || type == JT_FIXED) && k == 0) {
        const Q dq = qmul(q1, conj(q2));
        return row_value(V3{dq.x, dq.y, dq.z} * 2.0f, mode);
    }
    const V3 p1 = x1 + rotate(q1, xyz(r1)), p2 = x2 + rotate(q2, xyz(r2));
    if (type == JT_FIXED) return k == 1 ? row_value(p1 - p2, mode) : old;
    if (type == JT_PRISMATIC) {
        const V3 dp = p1 - p2;
        const V3 axis = rotate(q1, xyz(ax));
        const float ps = dot(dp, axis);
        if (cur) tg.z = ps;
        if (k == 1) {
            float over = 0.0f;
            if (jt_limited(m)) over = ps - clampf(ps, ax.w, lx.w);
            return row_value((dp - axis * ps) + axis * over, mode);
        }
        if (k == 2) {
            V3 drive{0.0f, 0.0f, 0.0f};
            if (act == JA_SPEED || act == JA_POSITION) {
                const float curr = dot(p2 - p1, axis);
                drive = axis * (tg.x - curr);
            }
            return row_value(drive, mode);
        }
        return old;
    }
    if (type == JT_REVOLUTE) {
        if (k == 1) {
            if (!cur) return row_value(p1 - p2, mode);
        }
        const V3 a1 = rotate(q1, xyz(ax)), a2 = rotate(q2, xyz(ax));
        if (k == 0 && !cur) return row_value(mode == 0 ? cross(a1, a2) : cross(a2, a1), mode);
        const V3 l1 = rotate(q1, xyz(lx)), l2 = rotate(q2, xyz(lx));
        float4 out = old;
        if (k == 0) out = row_value(mode == 0 ? cross(a1, a2) : cross(a2, a1), mode);
        if (k == 1) out = row_value(p1 - p2, mode);
        if (k == 2 || cur) {
            float phi = rb_atan2(dot(cross(l1, l2), a1), dot(l1, l2));
            V3 lim{0.0f, 0.0f, 0.0f};
            if (jt_limited(m) && (phi < ax.w || phi > lx.w)) {
                phi = clampf(phi, ax.w, lx.w);
                if (k == 2) {
                    const V3 n1 = rotate(from_axis_angle(a1, phi), l1);
                    lim = cross(n1, l2);
                }
            }
            if (k == 2) out = row_value(-lim, mode);
            if (cur) tg.z = phi;
        }
        if (k == 3) {
            V3 drive{0.0f, 0.0f, 0.0f};
            if (act == JA_SPEED || act == JA_POSITION) {
                const float theta = tg.x + 3.14159265358979323846f;
                const V3 bt = (l1 * rb_cos(theta) + cross(a1, l1) * rb_sin(theta)) + a1 * dot(a1, l1) * (1.0f - rb_cos(theta));
                drive = cross(bt, l2);
            }
            out = row_value(drive, mode);
        }
        return out;
    }
    // spherical
    if (k == 0) return row_value(p1 - p2, mode);
    if (k == 1 && jt_limited(m)) {
        const V3 a1 = rotate(q1, xyz(ax)), a2 = rotate(q2, xyz(ax));
        const V3 c = cross(a1, a2);
        const float th = rb_atan2(len(c), dot(a1, a2));
        V3 lim{0.0f, 0.0f, 0.0f};
        if (th > lx.w) {
            const V3 n1 = rotate(from_axis_angle(c, lx.w), a1);
            lim = cross(n1, a2);
        }
        if (cur) tg.z = th;
        return row_value(-lim, mode);
    }
    return old;
}

// compute_joint_errors (Joint.cpp:222-244) for every joint; mode 1 also resets lambda.
// KNOWN (here and below): the joint's meta row comes with the call (staged by the caller).
template <bool KNOWN = false>
__device__ __forceinline__ void joint_errors(Bodies &B, Joints &J, uint32_t i, float h, int mode, int4 mk = {}) {
    const int4 m = KNOWN ? mk : J.jm[i];
    const uint32_t b1 = (uint32_t)m.y, b2 = (uint32_t)m.z;
    float4 tg = J.jtg[i];
    float4 row[4];
    for (int k = 0; k < 4; ++k) row[k] = J.jrow[(size_t)i * 4 + k];
    joint_rows(m, J.jr1[i], J.jr2[i], J.jax[i], J.jlx[i], tg, xyz(B.r[b1].x), quat(B.r[b1].q), xyz(B.r[b2].x),
               quat(B.r[b2].q), h, true, mode, row);
    const int nrows = jt_rows(m);
    for (int k = 0; k < nrows; ++k) {
        J.jrow[(size_t)i * 4 + k] = row_at(row, k);
        if (mode == 1) J.jlam[(size_t)i * 4 + k] = 0.0f;
    }
    J.jtg[i] = tg;
}
__global__ void k_joint_errors(Bodies B, Joints J, uint32_t nj, float h, int mode) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nj) joint_errors(B, J, i, h, mode);
}

// Body.cpp:103-109 (refreshes the world tensor)
__device__ __forceinline__ float w_rot(BodyReg &r, V3 n) {
    if (!r.dyn) return 0.0f;
    r.Iwi = world_tensor(mat3_cast(r.q), r.Ii);
    return dot(n, smv(r.Iwi, n));
}
__device__ __forceinline__ void apply_rot(BodyReg &r, V3 imp) {   // Body.cpp:70-77
    if (!r.dyn) return;
    const V3 rot = smv(r.Iwi, imp) * 0.5f;
    r.q = qstep(r.q, rot);
}

// solve_constraints over one joint's rows (Constraint.cpp:14-46, 89-106); mode 1 first
// re-evaluates the rows from the bodies' current poses.
template <bool KNOWN = false>
__device__ __forceinline__ void joint_solve(Bodies &B, Joints &J, uint32_t i, float inv_h, float h, int mode,
                                            int4 mk = {}) {
    const int4 m = KNOWN ? mk : J.jm[i];
    const int type = jt_type(m), nrows = jt_rows(m), crow = jt_comp_row(type);
    const uint32_t b1 = (uint32_t)m.y, b2 = (uint32_t)m.z;
    const float4 r1f = J.jr1[i], r2f = J.jr2[i], axf = J.jax[i], lxf = J.jlx[i];
    const V3 r1 = xyz(r1f), r2 = xyz(r2f);
    float4 tg = J.jtg[i];
    BodyReg A, Bd;
    load_body(B, b1, A);
    load_body(B, b2, Bd);
    // the rows live in four named values, selected by value: an array indexed at run time
    // would be placed in scratch memory
    const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 o0 = nrows > 0 ? J.jrow[(size_t)i * 4] : z4, o1 = nrows > 1 ? J.jrow[(size_t)i * 4 + 1] : z4;
    float4 o2 = nrows > 2 ? J.jrow[(size_t)i * 4 + 2] : z4, o3 = nrows > 3 ? J.jrow[(size_t)i * 4 + 3] : z4;
    for (int k = 0; k < nrows; ++k) {
        float4 rv = k == 0 ? o0 : k == 1 ? o1 : k == 2 ? o2 : o3;
        if (mode == 1)   // this row from the current poses; the joint's position at its last row
            rv = joint_row(m, r1f, r2f, axf, lxf, tg, A.x, A.q, Bd.x, Bd.q, mode, k, k == nrows - 1, rv);
        const V3 n = xyz(rv);
        const float comp = (k == crow) ? r2f.w : 0.0f;
        float lam = J.jlam[(size_t)i * 4 + k];
        if (jt_rot(m, k)) {
            const float w1 = w_rot(A, n), w2 = w_rot(Bd, n);
            const float ap = comp * inv_h * inv_h;
            const float dl = (-rv.w - ap * lam) / (w1 + w2 + ap);
            lam = lam + dl;
            const V3 imp = dl * n;
            apply_rot(A, imp);
            apply_rot(Bd, -imp);
        } else {
            const V3 r1w = rotate(A.q, r1), r2w = rotate(Bd.q, r2);
            const float w1 = w_pos(A, r1w, n), w2 = w_pos(Bd, r2w, n);
            const float ap = comp * inv_h * inv_h;
            const float dl = (-rv.w - ap * lam) / (w1 + w2 + ap);
            lam = lam + dl;
            const V3 imp = dl * n;
            apply_pos(A, imp, r1w);
            apply_pos(Bd, -imp, r2w);
        }
        J.jlam[(size_t)i * 4 + k] = lam;
    }
    if (mode == 1) J.jtg[i] = tg;
    if (A.dyn) { B.r[b1].x = f4(A.x, A.im); B.r[b1].q = f4q(A.q); }
    if (Bd.dyn) { B.r[b2].x = f4(Bd.x, Bd.im); B.r[b2].q = f4q(Bd.q); }
}

__global__ void __launch_bounds__(256) k_joint_solve(Bodies B, Joints J, const uint32_t *order, uint32_t off,
                                                     uint32_t count, float inv_h, float h, int mode) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) joint_solve(B, J, order[off + t], inv_h, h, mode);
}

// Joint.cpp:184-220: prismatic (linear) and revolute (angular) damping; others none.
__device__ __forceinline__ float w_rot_vel(VelReg &r, V3 n) {
    if (!r.dyn) return 0.0f;
    r.Iwi = world_tensor(mat3_cast(r.q), r.Ii);
    return dot(n, smv(r.Iwi, n));
}
template <bool KNOWN = false>
__device__ __forceinline__ void joint_damp(Bodies &B, Joints &J, uint32_t i, float h, int4 mk = {}) {
    const int4 m = KNOWN ? mk : J.jm[i];
    const int type = jt_type(m);
    if (type != JT_PRISMATIC && type != JT_REVOLUTE) return;
    const uint32_t b1 = (uint32_t)m.y, b2 = (uint32_t)m.z;
    const float4 r1f = J.jr1[i];
    VelReg A, Bd;
    load_vel(B, b1, A);
    load_vel(B, b2, Bd);
    const float f = fmin_ref(r1f.w * h, 1.0f);
    if (type == JT_PRISMATIC) {
        const V3 dv = (Bd.v - A.v) * f;
        if (len2(dv) < EPS) return;
        const V3 r1w = rotate(A.q, xyz(r1f)), r2w = rotate(Bd.q, xyz(J.jr2[i]));
        const V3 n = normalize(dv);
        const float w1 = w_vel(A, r1w, n), w2 = w_vel(Bd, r2w, n);
        const V3 imp = vdiv(dv, w1 + w2);
        apply_vel(A, imp, r1w);
        apply_vel(Bd, -imp, r2w);
    } else {
        const V3 dw = (Bd.w - A.w) * f;
        if (len2(dw) < EPS) return;
        const V3 n = normalize(dw);
        const float w1 = w_rot_vel(A, n), w2 = w_rot_vel(Bd, n);
        const V3 imp = vdiv(dw, w1 + w2);
        if (A.dyn) A.w = A.w + smv(A.Iwi, imp);
        if (Bd.dyn) Bd.w = Bd.w + smv(Bd.Iwi, -imp);
    }
    if (A.dyn) { B.r[b1].v = f4(A.v, 0.0f); B.r[b1].w = f4(A.w, 0.0f); }
    if (Bd.dyn) { B.r[b2].v = f4(Bd.v, 0.0f); B.r[b2].w = f4(Bd.w, 0.0f); }
}

__global__ void __launch_bounds__(256) k_joint_damp(Bodies B, Joints J, const uint32_t *order, uint32_t off,
                                                    uint32_t count, float h) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) joint_damp(B, J, order[off + t], h);
}

// Colour-ordered solve of every scene, one workgroup per scene (POS: XPBD position solve;
// otherwise the velocity solve).  Lanes iterate alike, so the point counter may reduce per
// wave.  With per-scene joint tables the scene's joints follow in the same workgroup: after
// the position solve their errors (Joint.cpp:222-244) and rows colour by colour
// (Constraint.cpp:89-106), after the velocity solve their damping (Joint.cpp:249-267) — the
// order of the global joint launches restricted to the scene, whose joints touch no other
// scene's bodies.
// The scene's colour offsets, its pair list and joint order are read into LDS once (up to
// SOLVE_LDS_PAIRS pairs, SOLVE_LDS_JOINTS joints; larger scenes read them from the global
// lists), so that each colour step starts with the contact and body loads.
constexpr uint32_t SOLVE_LDS_PAIRS = 1024, SOLVE_LDS_JOINTS = 512, SOLVE_LDS_COLOURS = 256, SOLVE_LDS_JCOLS = 64;
template <bool POS, int OCC = 1, bool ALL = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC)))
k_scene_solve(Bodies B, Contacts K, SceneTabs T, float h, unsigned long long *pts, Joints J, float inv_h, int jmode) {
    static_assert(POS || !ALL, "the whole substep solve starts with the positions");
    __shared__ uint32_t s_co[SOLVE_LDS_COLOURS + 1], s_lst[SOLVE_LDS_PAIRS];
    __shared__ uint2 s_bb[SOLVE_LDS_PAIRS];       // the listed pairs' bodies
    __shared__ uint8_t s_np[SOLVE_LDS_PAIRS];     // and point counts
    __shared__ uint32_t s_jc[SOLVE_LDS_JCOLS + 1], s_jo[SOLVE_LDS_JOINTS];
    __shared__ int4 s_jm[SOLVE_LDS_JOINTS];       // the scene's joints' meta rows (type, bodies)
    const uint32_t s = blockIdx.x;
    const uint32_t nc = T.ncol[s];
    const uint32_t *lst = T.list + T.base[s];
    const uint32_t *co = T.coff + (size_t)s * (MAX_COLOURS + 1);
    const bool jn = T.jorder != nullptr;
    const uint32_t j0 = jn ? T.jbase[s] : 0u, j1 = jn ? T.jbase[s + 1] : 0u;
    const uint32_t *jc = jn ? T.jcoff + (size_t)s * (T.njc + 1) : nullptr;
    const bool co_l = nc <= SOLVE_LDS_COLOURS, jc_l = jn && T.njc <= SOLVE_LDS_JCOLS;
    for (uint32_t c = threadIdx.x; co_l && c <= nc; c += blockDim.x) s_co[c] = co[c];
    for (uint32_t c = threadIdx.x; jc_l && c <= T.njc; c += blockDim.x) s_jc[c] = jc[c];
    const uint32_t nj = j1 - j0;
    const bool jo_l = nj <= SOLVE_LDS_JOINTS;
    for (uint32_t i = threadIdx.x; jo_l && i < nj; i += blockDim.x) {
        const uint32_t q = T.jorder[j0 + i];
        s_jo[i] = q;
        s_jm[i] = J.jm[q];
    }
    __syncthreads();
    const uint32_t np = co_l ? s_co[nc] : co[nc];
    const bool ls_l = np <= SOLVE_LDS_PAIRS;
    for (uint32_t i = threadIdx.x; ls_l && i < np; i += blockDim.x) {
        const uint32_t p = lst[i];
        s_lst[i] = p;
        s_bb[i] = K.bb[p];
        s_np[i] = (uint8_t)K.cnt[p];
    }
    __syncthreads();
    // one solve: the contacts colour by colour, then the scene's joints (errors and rows after
    // the position solve, damping after the velocity solve)
    auto solve = [&](bool pos) {
        for (uint32_t c = 0; c < nc; ++c) {
            const uint32_t o = co_l ? s_co[c] : co[c], e = co_l ? s_co[c + 1] : co[c + 1];
            for (uint32_t i0 = o; i0 < e; i0 += blockDim.x) {
                const uint32_t i = i0 + threadIdx.x;
                const bool v = i < e;
                const uint32_t p = v ? (ls_l ? s_lst[i] : lst[i]) : 0u;
                if (pts && pos == POS) count_points(K, p, v, pts);
                if (v) {   // bodies and point count staged in LDS (or read here)
                    const int npp = ls_l ? (int)s_np[i] : K.cnt[p];
                    const uint2 bbp = ls_l ? s_bb[i] : K.bb[p];
                    if (pos) solve_pos_pair<true>(B, K, p, npp, bbp);
                    else solve_vel_pair<true>(B, K, p, h, npp, bbp);
                }
            }
            __syncthreads();
        }
        if (!jn || j0 == j1) return;
        if (pos) {
            for (uint32_t i = j0 + threadIdx.x; i < j1; i += blockDim.x) {
                const uint32_t q = jo_l ? s_jo[i - j0] : T.jorder[i];
                joint_errors<true>(B, J, q, h, jmode, jo_l ? s_jm[i - j0] : J.jm[q]);
            }
            __syncthreads();
        }
        for (uint32_t c = 0; c < T.njc; ++c) {
            const uint32_t e = j0 + (jc_l ? s_jc[c + 1] : jc[c + 1]);
            for (uint32_t i = j0 + (jc_l ? s_jc[c] : jc[c]) + threadIdx.x; i < e; i += blockDim.x) {
                const uint32_t q = jo_l ? s_jo[i - j0] : T.jorder[i];
                const int4 mq = jo_l ? s_jm[i - j0] : J.jm[q];   // the meta row (staged in LDS)
                if (pos) joint_solve<true>(B, J, q, inv_h, h, jmode, mq);
                else joint_damp<true>(B, J, q, h, mq);
            }
            __syncthreads();
        }
    };
    solve(POS);
    if (ALL) {
        // the rest of the substep for this scene (World.hpp:203-211): velocities from the
        // solved positions (k_update_vel on the scene's bodies), then the velocity solve
        const uint32_t b1 = T.body_off[s + 1];
        for (uint32_t i = T.body_off[s] + threadIdx.x; i < b1; i += blockDim.x) update_vel_body(B, T.body_idx[i], inv_h);
        __syncthreads();
        solve(false);
    }
}

}  // namespace rb
