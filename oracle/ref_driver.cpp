// C-ABI driver around the reference engine, compiled together with the
// reference's own sources (see oracle/Makefile, target `ref`) into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: it exists to pin the oracle restatement against the
// real reference and to time the reference CPU step.  The descriptor layouts are
// the same as oracle/orc_capi.cpp.
#include <cstdint>
#include <cstring>
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

}  // extern "C"

namespace {

rbc::Shape make_shape(const ref_collider_desc &d) {
    switch (d.shape) {
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
void ref_destroy(void *h) { delete static_cast<rbps::World *>(h); }

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
    bool hit = rbc::dispatch(make_shape(*a), make_tf(tfa), make_shape(*b), make_tf(tfb), m);
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
int ref_collision_groups(void *h, int32_t *out_colour) {
    auto *w = static_cast<rbps::World *>(h);
    auto groups = rbps::get_collision_groups(w->contacts, w->bodies);
    for (size_t g = 0; g < groups.size(); ++g)
        for (auto ci : groups[g]) out_colour[ci] = (int32_t)g;
    return (int)groups.size();
}

}  // extern "C"
