// C-ABI driver around the reference engine, compiled together with the
// reference's own sources (see oracle/Makefile, target `ref`) into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: it exists to pin the oracle restatement against the
// real reference and to time the reference CPU step.  The descriptor layouts are
// the same as oracle/orc_capi.cpp.
#include <cstdint>
#include <cstring>
#include <deque>
#include <unordered_map>
#include <vector>
#include "rbps/API/World.hpp"
#include "rbc/Dispatcher.hpp"

extern "C" {

typedef struct {
    int32_t type, _pad;
    double mass;
    double inertia[6];
    double position[3];
    double orientation[4];
    double linear_velocity[3];
    double angular_velocity[3];
} ref_body_desc;

typedef struct {
    int32_t shape;
    uint32_t body_id;
    int32_t is_static;
    int32_t hull_index;
    double dims[4];
    double local_pos[3];
    double local_rot[4];
    double restitution, static_friction, dynamic_friction;
} ref_collider_desc;

typedef struct {
    int32_t type, actuation;
    uint32_t body_1, body_2;
    double r_1[3], r_2[3];
    double axis[3], limit_axis[3];
    double compliance, damping;
    int32_t limited, _pad;
    double lower_limit, upper_limit;
    double target_position, target_speed;
} ref_joint_desc;

}  // extern "C"

namespace {

m3d::vec3 v3(const double *a) { return m3d::vec3(a[0], a[1], a[2]); }

// The driver's joint list: joints (types 0-2, their JointCollection id) and free constraint
// rows (type 4, made by World::create_constraint; their ConstraintCollection id), in
// creation order, per world.
struct Entry {
    int type;
    uint32_t id;
};
std::unordered_map<const void *, std::vector<Entry>> g_entries;

// Creates one joint through the reference's World API; target fields are the public
// joint data a caller sets after creation.  Spherical (3) has no reference counterpart.
// Type 4: a constraint row (World::create_constraint, ConstraintAPI.hpp:37-64) with
// actuation = row type (0 POSITIONAL, 1 ROTATIONAL), r_1, r_2 and compliance.
int add_joint(rbps::World &w, const ref_joint_desc &d) {
    const auto act = static_cast<rbps::JointActuationType>(d.actuation);
    uint32_t id;
    if (d.type == 4) {
        rbps::ConstraintParams p;
        p.body_1 = d.body_1; p.body_2 = d.body_2; p.r_1 = v3(d.r_1); p.r_2 = v3(d.r_2); p.compliance = d.compliance;
        p.type = d.actuation == 1 ? rbps::ConstraintType::ROTATIONAL : rbps::ConstraintType::POSITIONAL;
        g_entries[&w].push_back({4, w.create_constraint(p)});
        return 0;
    }
    if (d.type == 0) {
        rbps::PrismaticJointParams p;
        p.body_1 = d.body_1; p.body_2 = d.body_2; p.moving_axis = v3(d.axis); p.r_1 = v3(d.r_1); p.r_2 = v3(d.r_2);
        p.actuation_type = act; p.compliance = d.compliance; p.damping = d.damping; p.limited = d.limited != 0;
        p.lower_limit = d.lower_limit; p.upper_limit = d.upper_limit;
        id = w.create_prismatic_joint(p);
    } else if (d.type == 1) {
        rbps::RevoluteJointParams p;
        p.body_1 = d.body_1; p.body_2 = d.body_2; p.aligned_axis = v3(d.axis); p.limit_axis = v3(d.limit_axis);
        p.r_1 = v3(d.r_1); p.r_2 = v3(d.r_2); p.actuation_type = act; p.compliance = d.compliance;
        p.damping = d.damping; p.limited = d.limited != 0; p.lower_limit = d.lower_limit; p.upper_limit = d.upper_limit;
        id = w.create_revolute_joint(p);
    } else if (d.type == 2) {
        rbps::FixedJointParams p;
        p.body_1 = d.body_1; p.body_2 = d.body_2; p.r_1 = v3(d.r_1); p.r_2 = v3(d.r_2);
        id = w.create_fixed_joint(p);
    } else {
        return -1;
    }
    const uint32_t i = w.joints.index_of(id);
    w.joints.target_position[i] = d.target_position;
    w.joints.target_speed[i] = d.target_speed;
    g_entries[&w].push_back({d.type, id});
    return 0;
}

// Registered hulls: the caller-owned arrays ConvexHullData points at, kept for the
// process lifetime, and the data made by convex_hull_data_create.
struct HullStore {
    std::vector<m3d::vec3> v;
    std::vector<uint32_t> f;
    rbc::ConvexHullData *data = nullptr;
};
std::deque<HullStore> g_hulls;

rbc::Shape make_shape(const ref_collider_desc &d) {
    switch (d.shape) {
        case 9: return rbc::Shape(rbc::ConvexHull(g_hulls.at(d.hull_index).data));
        case 0: return rbc::Shape(rbc::Sphere(d.dims[0]));
        case 1: return rbc::Shape(rbc::Box(m3d::vec3(d.dims[0], d.dims[1], d.dims[2])));
        case 3: return rbc::Shape(rbc::Capsule(d.dims[0], d.dims[1]));
        case 6: return rbc::Shape(rbc::Plane(m3d::vec3(d.dims[0], d.dims[1], d.dims[2]), d.dims[3]));
        default: return rbc::Shape(rbc::Sphere(d.dims[0]));
    }
}

m3d::tf make_tf(const double *p7) {
    return m3d::tf(m3d::vec3(p7[0], p7[1], p7[2]), m3d::quat(p7[3], p7[4], p7[5], p7[6]));
}

}  // namespace

extern "C" {

void *ref_create(double timestep, int substeps) {
    auto *w = new rbps::World(timestep, substeps);
    w->broad_phase_state.dirty = false;
    return w;
}
void ref_destroy(void *h) {
    g_entries.erase(h);
    delete static_cast<rbps::World *>(h);
}

int ref_add_bodies(void *h, int n, const ref_body_desc *d) {
    auto *w = static_cast<rbps::World *>(h);
    for (int i = 0; i < n; ++i) {
        rbps::BodyParams p;
        p.type = d[i].type == 0 ? rbps::STATIC : rbps::DYNAMIC;
        p.mass = d[i].mass;
        p.inertia_tensor = m3d::smat3(d[i].inertia[0], d[i].inertia[1], d[i].inertia[2], d[i].inertia[3],
                                      d[i].inertia[4], d[i].inertia[5]);
        p.position = m3d::vec3(d[i].position[0], d[i].position[1], d[i].position[2]);
        p.orientation = m3d::quat(d[i].orientation[0], d[i].orientation[1], d[i].orientation[2], d[i].orientation[3]);
        p.linear_velocity = m3d::vec3(d[i].linear_velocity[0], d[i].linear_velocity[1], d[i].linear_velocity[2]);
        p.angular_velocity = m3d::vec3(d[i].angular_velocity[0], d[i].angular_velocity[1], d[i].angular_velocity[2]);
        w->create_body(p);
    }
    return 0;
}

int ref_add_colliders(void *h, int n, const ref_collider_desc *d) {
    auto *w = static_cast<rbps::World *>(h);
    for (int i = 0; i < n; ++i) {
        rbps::ColliderParams p;
        p.shape = make_shape(d[i]);
        p.body_id = d[i].body_id;
        p.is_static = d[i].is_static != 0;
        p.local_pos = m3d::vec3(d[i].local_pos[0], d[i].local_pos[1], d[i].local_pos[2]);
        p.local_rot = m3d::quat(d[i].local_rot[0], d[i].local_rot[1], d[i].local_rot[2], d[i].local_rot[3]);
        p.restitution = d[i].restitution;
        p.static_friction = d[i].static_friction;
        p.dynamic_friction = d[i].dynamic_friction;
        w->create_collider(p);
    }
    return 0;
}

int ref_step(void *h, int frames) {
    auto *w = static_cast<rbps::World *>(h);
    for (int f = 0; f < frames; ++f) w->step();
    return 0;
}

int ref_num_bodies(void *h) { return (int)static_cast<rbps::World *>(h)->bodies.count(); }

void ref_get_bodies(void *h, double *p, double *q, double *v, double *av) {
    auto *w = static_cast<rbps::World *>(h);
    const auto &b = w->bodies;
    for (uint32_t i = 0; i < b.count(); ++i) {
        if (p) { p[3 * i] = b.position[i].x; p[3 * i + 1] = b.position[i].y; p[3 * i + 2] = b.position[i].z; }
        if (q) { q[4 * i] = b.orientation[i].w; q[4 * i + 1] = b.orientation[i].x; q[4 * i + 2] = b.orientation[i].y; q[4 * i + 3] = b.orientation[i].z; }
        if (v) { v[3 * i] = b.linear_velocity[i].x; v[3 * i + 1] = b.linear_velocity[i].y; v[3 * i + 2] = b.linear_velocity[i].z; }
        if (av) { av[3 * i] = b.angular_velocity[i].x; av[3 * i + 1] = b.angular_velocity[i].y; av[3 * i + 2] = b.angular_velocity[i].z; }
    }
}

void ref_set_bodies(void *h, const double *p, const double *q, const double *v, const double *av) {
    auto *w = static_cast<rbps::World *>(h);
    auto &b = w->bodies;
    for (uint32_t i = 0; i < b.count(); ++i) {
        if (p) b.position[i] = m3d::vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        if (q) b.orientation[i] = m3d::quat(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]);
        if (v) b.linear_velocity[i] = m3d::vec3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        if (av) b.angular_velocity[i] = m3d::vec3(av[3 * i], av[3 * i + 1], av[3 * i + 2]);
    }
}

int ref_num_pairs(void *h) { return (int)static_cast<rbps::World *>(h)->broad_phase_state.pairs.size(); }
void ref_get_pairs(void *h, uint32_t *out) {
    const auto &pr = static_cast<rbps::World *>(h)->broad_phase_state.pairs;
    for (size_t i = 0; i < pr.size(); ++i) { out[2 * i] = pr[i].id_a; out[2 * i + 1] = pr[i].id_b; }
}

int ref_num_contacts(void *h) { return (int)static_cast<rbps::World *>(h)->contacts.n_contacts; }
void ref_get_contacts(void *h, uint32_t *ids4, double *n3, double *ra3, double *rb3, double *dep, double *lam2) {
    const auto &c = static_cast<rbps::World *>(h)->contacts;
    for (uint32_t i = 0; i < c.n_contacts; ++i) {
        if (ids4) { ids4[4 * i] = c.body_a[i]; ids4[4 * i + 1] = c.body_b[i]; ids4[4 * i + 2] = c.collider_a[i]; ids4[4 * i + 3] = c.collider_b[i]; }
        if (n3) { n3[3 * i] = c.normal[i].x; n3[3 * i + 1] = c.normal[i].y; n3[3 * i + 2] = c.normal[i].z; }
        if (ra3) { ra3[3 * i] = c.r_a_local[i].x; ra3[3 * i + 1] = c.r_a_local[i].y; ra3[3 * i + 2] = c.r_a_local[i].z; }
        if (rb3) { rb3[3 * i] = c.r_b_local[i].x; rb3[3 * i + 1] = c.r_b_local[i].y; rb3[3 * i + 2] = c.r_b_local[i].z; }
        if (dep) dep[i] = c.penetration_depth[i];
        if (lam2) { lam2[2 * i] = c.normal_lambda[i]; lam2[2 * i + 1] = c.tangent_lambda[i]; }
    }
}

// One narrowphase call through the reference dispatcher.  tf = (px,py,pz,qw,qx,qy,qz).
// out: [n, nx, ny, nz, (px, py, pz, depth) x4]; returns hit.
int ref_collide(const ref_collider_desc *a, const double *tfa, const ref_collider_desc *b, const double *tfb, double *out) {
    rbc::ContactManifold m;
    bool hit;
    try {
        hit = rbc::dispatch(make_shape(*a), make_tf(tfa), make_shape(*b), make_tf(tfb), m);
    } catch (const std::exception &) {
        return -1;
    }
    out[0] = hit ? m.num_points : 0;
    out[1] = m.normal.x; out[2] = m.normal.y; out[3] = m.normal.z;
    for (int i = 0; i < 4; ++i) {
        out[4 + 4 * i] = m.points[i].position.x;
        out[5 + 4 * i] = m.points[i].position.y;
        out[6 + 4 * i] = m.points[i].position.z;
        out[7 + 4 * i] = m.points[i].penetration_depth;
    }
    return hit ? 1 : 0;
}

// The reference's greedy colouring (CollisionPipeline.cpp:223-293) on the current contact list.
// out_colour[i] = group index of contact i (groups sorted by size as the reference returns them).
int ref_add_joints(void *h, int n, const ref_joint_desc *d) {
    auto *w = static_cast<rbps::World *>(h);
    for (int i = 0; i < n; ++i)
        if (add_joint(*w, d[i])) return -1;
    return 0;
}
int ref_num_joints(void *h) { return (int)g_entries[h].size(); }
void ref_get_joints(void *h, double *cur, double *tgt, double *lam4, int32_t *colour) {
    auto *w = static_cast<rbps::World *>(h);
    auto &jc = w->joints;
    auto &cc = w->constraints;
    const auto &E = g_entries[h];
    for (uint32_t e = 0; e < E.size(); ++e) {
        if (E[e].type == 4) {   // a constraint row: its lambda
            if (cur) cur[e] = 0.0;
            if (tgt) tgt[e] = 0.0;
            if (lam4) for (uint32_t k = 0; k < 4; ++k) lam4[4 * e + k] = k == 0 ? cc.lambda[cc.index_of(E[e].id)] : 0.0;
            if (colour) colour[e] = -1;
            continue;
        }
        const uint32_t i = jc.index_of(E[e].id);
        if (cur) cur[e] = jc.current_position[i];
        if (tgt) tgt[e] = jc.target_position[i];
        if (lam4)
            for (uint32_t k = 0; k < 4; ++k)
                lam4[4 * e + k] = k < jc.constraint_count[i] ? cc.lambda[cc.index_of(jc.constraints[i].ids[k])] : 0.0;
        if (colour) colour[e] = -1;
    }
}
void ref_set_joint_targets(void *h, int n, const uint32_t *ids, const double *tgt, const double *speed) {
    auto &jc = static_cast<rbps::World *>(h)->joints;
    const auto &E = g_entries[h];
    for (int k = 0; k < n; ++k) {
        if (E.at(ids[k]).type == 4) continue;
        const uint32_t i = jc.index_of(E[ids[k]].id);
        if (tgt) jc.target_position[i] = tgt[k];
        if (speed) jc.target_speed[i] = speed[k];
    }
}
// The public ConstraintCollection::direction / magnitude of constraint rows (entries of
// type 4), which the reference's callers write directly.
void ref_set_rows(void *h, int n, const uint32_t *ids, const double *dir3, const double *mag) {
    auto &cc = static_cast<rbps::World *>(h)->constraints;
    const auto &E = g_entries[h];
    for (int k = 0; k < n; ++k) {
        if (E.at(ids[k]).type != 4) continue;
        const uint32_t i = cc.index_of(E[ids[k]].id);
        if (dir3) cc.direction[i] = m3d::vec3(dir3[3 * k], dir3[3 * k + 1], dir3[3 * k + 2]);
        if (mag) cc.magnitude[i] = mag[k];
    }
}
void ref_set_forces(void *h, const double *f3, const double *t3) {
    auto &b = static_cast<rbps::World *>(h)->bodies;
    for (uint32_t i = 0; i < b.count(); ++i) {
        if (f3) b.force[i] = m3d::vec3(f3[3 * i], f3[3 * i + 1], f3[3 * i + 2]);
        if (t3) b.torque[i] = m3d::vec3(t3[3 * i], t3[3 * i + 1], t3[3 * i + 2]);
    }
}
// compute_*_joint_errors (Joint.cpp:18-182) for one joint between two bodies at given poses.
int ref_joint_errors(const ref_joint_desc *d, const double *pose1, const double *pose2, double h, double *out) {
    rbps::World w(1.0 / 60.0, 1);
    for (const double *p : {pose1, pose2}) {
        rbps::BodyParams bp;
        bp.type = rbps::DYNAMIC;
        bp.mass = 1.0;
        bp.inertia_tensor = m3d::smat3(1, 1, 1, 0, 0, 0);
        bp.position = v3(p);
        bp.orientation = m3d::quat(p[3], p[4], p[5], p[6]);
        w.create_body(bp);
    }
    ref_joint_desc dd = *d;
    dd.body_1 = 0; dd.body_2 = 1;
    const int rc = add_joint(w, dd);
    g_entries.erase(&w);
    if (rc) return -1;
    rbps::compute_joint_errors(w.joints, w.bodies, w.constraints, h);
    auto &jc = w.joints;
    auto &cc = w.constraints;
    for (uint32_t k = 0; k < 4; ++k) {
        if (k < jc.constraint_count[0]) {
            const uint32_t r = cc.index_of(jc.constraints[0].ids[k]);
            out[4 * k] = cc.direction[r].x; out[4 * k + 1] = cc.direction[r].y; out[4 * k + 2] = cc.direction[r].z;
            out[4 * k + 3] = cc.magnitude[r];
        } else {
            out[4 * k] = 1.0; out[4 * k + 1] = 0.0; out[4 * k + 2] = 0.0; out[4 * k + 3] = 0.0;
        }
    }
    out[16] = jc.current_position[0];
    out[17] = jc.target_position[0];
    return 0;
}

// convex_hull_data_create (ConvexHull.hpp:149-420); returns the index for hull_index.
int ref_hull_register(const double *verts, int nv, const uint32_t *faces, int nf) {
    g_hulls.emplace_back();
    HullStore &h = g_hulls.back();
    for (int i = 0; i < nv; ++i) h.v.push_back(m3d::vec3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]));
    if (faces && nf > 0) h.f.assign(faces, faces + 3 * nf);
    h.data = rbc::convex_hull_data_create(h.v.data(), (uint32_t)nv, h.f.empty() ? nullptr : h.f.data(),
                                          h.f.empty() ? 0u : (uint32_t)nf);
    return (int)g_hulls.size() - 1;
}
int ref_hull_sizes(int i, int32_t *out6) {
    if (i < 0 || (size_t)i >= g_hulls.size()) return -1;
    const rbc::ConvexHullData *d = g_hulls[i].data;
    const uint32_t P = d->poly_count, nv = d->vert_count;
    out6[0] = (int32_t)P;
    out6[1] = P ? (int32_t)d->poly_offsets[P] : 0;
    out6[2] = P ? (int32_t)d->vertex_to_polys_offsets[nv] : 0;
    out6[3] = P ? (int32_t)d->vertex_to_neighbors_offsets[nv] : 0;
    out6[4] = P ? (int32_t)d->poly_to_polys_offsets[P] : 0;
    out6[5] = -1;
    return 0;
}
int ref_hull_tables(int i, double *pn, uint32_t *poff, uint32_t *pidx, uint32_t *v2p_o, uint32_t *v2p_i,
                    uint32_t *v2n_o, uint32_t *v2n_i, uint32_t *p2p_o, uint32_t *p2p_i) {
    if (i < 0 || (size_t)i >= g_hulls.size()) return -1;
    const rbc::ConvexHullData *d = g_hulls[i].data;
    const uint32_t P = d->poly_count, nv = d->vert_count;
    if (!P) return 0;
    for (uint32_t k = 0; k < P; ++k) { pn[3 * k] = d->poly_normals[k].x; pn[3 * k + 1] = d->poly_normals[k].y; pn[3 * k + 2] = d->poly_normals[k].z; }
    auto cp = [](const uint32_t *src, uint32_t n, uint32_t *o) { for (uint32_t k = 0; k < n; ++k) o[k] = src[k]; };
    cp(d->poly_offsets, P + 1, poff);
    cp(d->poly_indices, d->poly_offsets[P], pidx);
    cp(d->vertex_to_polys_offsets, nv + 1, v2p_o);
    cp(d->vertex_to_polys_indices, d->vertex_to_polys_offsets[nv], v2p_i);
    cp(d->vertex_to_neighbors_offsets, nv + 1, v2n_o);
    cp(d->vertex_to_neighbors_indices, d->vertex_to_neighbors_offsets[nv], v2n_i);
    cp(d->poly_to_polys_offsets, P + 1, p2p_o);
    cp(d->poly_to_polys_indices, d->poly_to_polys_offsets[P], p2p_i);
    return 0;
}
// compute_volume / compute_inertia_tensor (ConvexHull.hpp:575-745): volume, xx yy zz xy xz yz.
int ref_hull_mass(int i, double *out7) {
    if (i < 0 || (size_t)i >= g_hulls.size()) return -1;
    rbc::ConvexHull h(g_hulls[i].data);
    out7[0] = rbc::compute_volume(h);
    const m3d::smat3 I = rbc::compute_inertia_tensor(h);
    out7[1] = I.xx; out7[2] = I.yy; out7[3] = I.zz; out7[4] = I.xy; out7[5] = I.xz; out7[6] = I.yz;
    return 0;
}
// GJK + EPA of one pair as gjk_epa_manifold runs them (ContactManifoldGenerator.cpp:415-438):
// out = gjk status, -1, epa status (-1 if not run), -1, normal xyz, depth, witness xyz.
int ref_gjk_epa(const ref_collider_desc *a, const double *tfa, const ref_collider_desc *b, const double *tfb,
                double *out) {
    try {
        const rbc::Shape sa = make_shape(*a), sb = make_shape(*b);
        const m3d::tf ta = make_tf(tfa), tb = make_tf(tfb);
        rbc::MinkowskiDiff md(&sa, &sb, ta, tb);
        m3d::vec3 guess = tb.pos - ta.pos;
        if (m3d::length_sq(guess) < m3d::EPSILON) guess = m3d::vec3(1.0, 0.0, 0.0);
        rbc::GJK gjk;
        const int gs = gjk.evaluate(md, guess);
        for (int k = 0; k < 11; ++k) out[k] = 0.0;
        out[0] = gs; out[1] = -1; out[2] = -1; out[3] = -1;
        if (gs != rbc::GJK::Inside) return 0;
        rbc::EPA epa;
        const int es = epa.evaluate(gjk, md);
        out[2] = es;
        out[4] = epa.normal.x; out[5] = epa.normal.y; out[6] = epa.normal.z; out[7] = epa.depth;
        out[8] = epa.contact_point.x; out[9] = epa.contact_point.y; out[10] = epa.contact_point.z;
        return es == rbc::EPA::Valid ? 1 : 0;
    } catch (const std::exception &) {
        return -1;
    }
}

int ref_collision_groups(void *h, int32_t *out_colour) {
    auto *w = static_cast<rbps::World *>(h);
    auto groups = rbps::get_collision_groups(w->contacts, w->bodies);
    for (size_t g = 0; g < groups.size(); ++g)
        for (auto ci : groups[g]) out_colour[ci] = (int32_t)g;
    return (int)groups.size();
}

}  // extern "C"
