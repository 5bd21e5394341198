// The reference's World tests (tests/rbps/API/test_world.cpp, every case, same names and
// assertions) followed by cases for the rest of the public surface (constraints, joint
// targets, contacts, broad phase, forces, hulls, errors), run against the MI355X rbps::World
// (include/rbps/API/World.hpp -> include/rbps_amd/World.hpp).  Built and run by
// tests/test_gpu_facade.py.
#include <cmath>
#include <vector>

#include "rbps/API/World.hpp"
#include "harness.hpp"
#include "tests/rbps/pipeline_helpers.hpp"

using namespace rbps;
using namespace m3d;

using test::add_dynamic_sphere;
using test::add_static_box;

// ---------------------------------------------------------------- test_world.cpp cases

TEST(world_default_construction) {   // commented out in the reference; holds here
    World w;
    ASSERT_NEAR(w.timestep, 1.0 / 60.0, 1e-9);
    ASSERT_TRUE(w.substeps == 20);
    ASSERT_TRUE(w.bodies.count() == 0);
}

TEST(world_parameterized_construction) {
    World w(1.0 / 120.0, 10);
    ASSERT_NEAR(w.timestep, 1.0 / 120.0, 1e-9);
    ASSERT_TRUE(w.substeps == 10);
}

TEST(world_create_body_returns_valid_id) {
    World w;
    BodyParams bp;
    bp.position = vec3{0, 0, 0};
    bp.mass = 1.0;
    bp.type = BodyType::DYNAMIC;
    uint32_t id = w.create_body(bp);
    int32_t slot = w.bodies.index_of(id);
    ASSERT_TRUE(slot >= 0);
    ASSERT_NEAR(w.bodies.position[slot].y, 0.0, 1e-6);
}

TEST(world_create_multiple_bodies) {
    World w;
    uint32_t a = add_dynamic_sphere(w, vec3{0, 2, 0});
    uint32_t b = add_dynamic_sphere(w, vec3{0, -2, 0});
    ASSERT_TRUE(a != b);
    ASSERT_TRUE(w.bodies.count() == 2);
    ASSERT_TRUE(w.colliders.count() == 2);
}

TEST(world_create_collider_registers_in_broad_phase) {
    World w;
    add_dynamic_sphere(w, vec3{0, 0, 0});
    ASSERT_TRUE(w.broad_phase_state.objects.size() >= 1);
}

TEST(world_step_empty_world) {
    World w;
    w.step();
    w.step();
}

TEST(world_step_single_body_no_collider) {
    World w;
    BodyParams bp;
    bp.position = vec3{0, 10, 0};
    bp.mass = 1.0;
    bp.type = BodyType::DYNAMIC;
    w.create_body(bp);
    w.step();
    ASSERT_TRUE(w.contacts.n_contacts == 0);
}

TEST(world_free_fall_moves_body_downward) {
    World w(1.0 / 60.0, 20);
    uint32_t bid = add_dynamic_sphere(w, vec3{0, 10, 0});
    int32_t slot = w.bodies.index_of(bid);
    scalar y_before = w.bodies.position[slot].y;
    w.step();
    ASSERT_TRUE(w.bodies.position[slot].y < y_before);
}

TEST(world_free_fall_velocity_increases) {
    World w(1.0 / 60.0, 20);
    int32_t slot = w.bodies.index_of(add_dynamic_sphere(w, vec3{0, 10, 0}));
    w.step();
    scalar vy_1 = w.bodies.linear_velocity[slot].y;
    w.step();
    ASSERT_TRUE(w.bodies.linear_velocity[slot].y < vy_1);
}

TEST(world_static_body_does_not_move) {
    World w;
    int32_t slot = w.bodies.index_of(add_static_box(w, vec3{0, 0, 0}));
    vec3 before = w.bodies.position[slot];
    for (int i = 0; i < 10; ++i) w.step();
    vec3 after = w.bodies.position[slot];
    ASSERT_NEAR(after.x, before.x, 1e-6);
    ASSERT_NEAR(after.y, before.y, 1e-6);
    ASSERT_NEAR(after.z, before.z, 1e-6);
}

TEST(world_sphere_lands_on_ground) {
    World w(1.0 / 60.0, 20);
    add_static_box(w, vec3{0, -3, 0});
    int32_t slot = w.bodies.index_of(add_dynamic_sphere(w, vec3{0, 1, 0}, 0.5));
    for (int i = 0; i < 180; ++i) w.step();
    ASSERT_TRUE(m3d::magnitude(w.bodies.linear_velocity[slot]) < 2.0);
    ASSERT_TRUE(w.bodies.position[slot].y > -4.0);
}

TEST(world_sphere_does_not_tunnel_through_ground) {
    World w(1.0 / 60.0, 25);
    add_static_box(w, vec3{0, -3, 0});
    int32_t slot = w.bodies.index_of(add_dynamic_sphere(w, vec3{0, 2, 0}, 0.5));
    w.auto_sync(false);   // 300 frames, one read-back
    for (int i = 0; i < 300; ++i) w.step();
    w.sync();
    ASSERT_TRUE(w.bodies.position[slot].y > -5.0);
}

TEST(world_two_spheres_collide_and_separate) {
    World w(1.0 / 60.0, 20);
    int32_t sa = w.bodies.index_of(add_dynamic_sphere(w, vec3{-2, 0, 0}, 0.5, 1.0, vec3{5, 0, 0}));
    int32_t sb = w.bodies.index_of(add_dynamic_sphere(w, vec3{2, 0, 0}, 0.5, 1.0, vec3{-5, 0, 0}));
    for (int i = 0; i < 30; ++i) w.step();
    ASSERT_TRUE(w.bodies.linear_velocity[sa].x < 0.0 || w.bodies.linear_velocity[sb].x > 0.0);
}

TEST(world_collision_conserves_momentum_approximately) {
    World w(1.0 / 60.0, 20);
    int32_t sa = w.bodies.index_of(add_dynamic_sphere(w, vec3{-2, 0, 0}, 0.5, 1.0, vec3{3, 0, 0}));
    int32_t sb = w.bodies.index_of(add_dynamic_sphere(w, vec3{2, 0, 0}, 0.5, 1.0, vec3{0, 0, 0}));
    scalar px_before = w.bodies.linear_velocity[sa].x * w.bodies.mass[sa] + w.bodies.linear_velocity[sb].x * w.bodies.mass[sb];
    for (int i = 0; i < 60; ++i) w.step();
    scalar px_after = w.bodies.linear_velocity[sa].x * w.bodies.mass[sa] + w.bodies.linear_velocity[sb].x * w.bodies.mass[sb];
    ASSERT_NEAR(px_after, px_before, m3d::abs(px_before) * 0.2 + 0.1);
}

TEST(world_contacts_cleared_each_frame) {
    World w(1.0 / 60.0, 20);
    add_static_box(w, vec3{0, -1, 0});
    add_dynamic_sphere(w, vec3{0, 0, 0}, 0.5);
    for (int i = 0; i < 10; ++i) w.step();
    ASSERT_TRUE(w.contacts.n_contacts <= 16);
}

TEST(world_no_contacts_when_bodies_separated) {
    World w;
    add_dynamic_sphere(w, vec3{0, 100, 0});
    add_dynamic_sphere(w, vec3{100, 0, 0});
    w.step();
    ASSERT_TRUE(w.contacts.n_contacts == 0);
}

// ---------------------------------------------------------------- the rest of the surface

// Contact fields of a sphere resting on a slab: one point, unit normal along y, lever arms
// reaching the touching surfaces, material mixing of emit_contact (min restitution, mean
// frictions), points on both surfaces, a positive normal multiplier.
TEST(world_contact_fields) {
    World w(1.0 / 60.0, 20);
    uint32_t g = add_static_box(w, vec3{0, -2, 0});   // top face at y = 0
    BodyParams bp;
    bp.position = vec3{0, 0.5, 0};
    uint32_t s = w.create_body(bp);
    ColliderParams cp;
    cp.shape = rbc::Sphere(0.5);
    cp.body_id = s;
    cp.restitution = 0.2;
    cp.static_friction = 0.9;
    cp.dynamic_friction = 0.7;
    w.create_collider(cp);
    for (int i = 0; i < 30; ++i) w.step();
    const ContactList &c = w.contacts;
    ASSERT_TRUE(c.n_contacts == 1);
    ASSERT_TRUE((c.body_a[0] == g && c.body_b[0] == s) || (c.body_a[0] == s && c.body_b[0] == g));
    ASSERT_NEAR(m3d::magnitude(c.normal[0]), 1.0, 1e-5);
    ASSERT_NEAR(m3d::abs(c.normal[0].y), 1.0, 1e-4);
    ASSERT_NEAR(c.restitution[0], 0.2, 1e-12);
    ASSERT_NEAR(c.static_friction[0], (0.9 + 0.5) * 0.5, 1e-12);
    ASSERT_NEAR(c.dynamic_friction[0], (0.7 + 0.5) * 0.5, 1e-12);
    ASSERT_TRUE(c.normal_lambda[0] != 0.0);
    ASSERT_TRUE(m3d::abs(c.penetration_depth[0]) < 0.01);
    const vec3 pa = c.point_on_a[0], pb = c.point_on_b[0];
    ASSERT_NEAR(pa.y, 0.0, 0.01);
    ASSERT_NEAR(pb.y, 0.0, 0.01);
    const vec3 sp = w.bodies.position[w.bodies.index_of(s)];
    const vec3 r_sphere = c.body_a[0] == s ? c.r_a_local[0] : c.r_b_local[0];
    ASSERT_NEAR(m3d::magnitude(r_sphere), 0.5, 1e-4);
    ASSERT_NEAR(sp.y, 0.5, 0.01);
}

// broad_phase_state: objects carry fat boxes around their colliders and the collider ids;
// pairs list the overlapping pair, and nothing for a far-away body.
TEST(world_broad_phase_state_mirror) {
    World w;
    add_dynamic_sphere(w, vec3{0, 0.9, 0}, 0.5);
    add_dynamic_sphere(w, vec3{0, 0, 0}, 0.5);
    add_dynamic_sphere(w, vec3{50, 0, 0}, 0.5);
    ASSERT_TRUE(w.broad_phase_state.objects.size() == 3);
    for (uint32_t i = 0; i < 3; ++i) {
        const rbc::BPObject &o = w.broad_phase_state.objects[i];
        ASSERT_TRUE(o.user_id == i);
        const vec3 p = w.bodies.position[i];
        ASSERT_TRUE(o.fat_aabb.min.x < p.x - 0.5 && o.fat_aabb.max.x > p.x + 0.5);
        ASSERT_TRUE(o.fat_aabb.min.y < p.y - 0.5 && o.fat_aabb.max.y > p.y + 0.5);
    }
    w.step();
    ASSERT_TRUE(w.broad_phase_state.pairs.size() == 1);
    const rbc::BroadPhasePair pr = w.broad_phase_state.pairs[0];
    ASSERT_TRUE(std::min(pr.id_a, pr.id_b) == 0 && std::max(pr.id_a, pr.id_b) == 1);
}

// create_constraint: rows start at direction (1, 0, 0), magnitude 0.  A positional row whose
// caller writes direction (0, 1, 0) and magnitude c moves the dynamic body up by c in every
// substep (Constraint.cpp:14-46, compliance 0), and the velocity update keeps each push:
// after N substeps the body is c * N (N + 1) / 2 above its free-falling twin.  Magnitude 0
// stops the pushes.  lambda is mirrored back.
TEST(world_create_constraint_row) {
    World w(1.0 / 60.0, 20);
    BodyParams anchor;
    anchor.type = STATIC;
    anchor.mass = 0.0;
    uint32_t a = w.create_body(anchor);
    BodyParams bp;
    bp.position = vec3{0, -1, 0};
    uint32_t b = w.create_body(bp);
    BodyParams tp;   // twin in free fall
    tp.position = vec3{3, -1, 0};
    uint32_t t = w.create_body(tp);
    ConstraintParams cp;
    cp.body_1 = a;
    cp.body_2 = b;
    cp.type = ConstraintType::POSITIONAL;
    uint32_t c = w.constraints.index_of(w.create_constraint(cp));
    ASSERT_TRUE(w.constraints.count() == 1);
    ASSERT_TRUE(w.constraints.direction[c] == vec3(1.0, 0.0, 0.0) && w.constraints.magnitude[c] == 0.0);
    set_value(w.constraints, c, vec3{0, 0.002, 0});
    ASSERT_NEAR(w.constraints.magnitude[c], 0.002, 1e-15);
    ASSERT_NEAR(w.constraints.direction[c].y, 1.0, 1e-15);
    w.step();
    ASSERT_NEAR(w.bodies.position[b].y - w.bodies.position[t].y, 0.002 * 20 * 21 / 2, 2e-3);
    ASSERT_NEAR(w.bodies.linear_velocity[b].y - w.bodies.linear_velocity[t].y, 0.002 * 20 * 1200, 0.5);
    ASSERT_TRUE(get_lagrange_multiplier(w.constraints, c) < 0.0);
    ASSERT_NEAR(w.constraints.magnitude[c], 0.002, 1e-9);   // the caller's value, kept
    w.constraints.magnitude[c] = 0.0;
    const scalar gap = w.bodies.position[b].y - w.bodies.position[t].y;
    const scalar dv = w.bodies.linear_velocity[b].y - w.bodies.linear_velocity[t].y;
    w.step();
    ASSERT_NEAR(w.bodies.position[b].y - w.bodies.position[t].y, gap + dv / 60.0, 0.02);
    ASSERT_NEAR(w.bodies.linear_velocity[b].y - w.bodies.linear_velocity[t].y, dv, 0.5);
}

// Joint rows in the reference's numbering (JointAPI.hpp:160-265): prismatic R P P,
// revolute R P R R, fixed R P, each joint's rows consecutive in creation order.
TEST(world_joint_rows_numbering) {
    World w;
    BodyParams s;
    s.type = STATIC;
    s.mass = 0.0;
    uint32_t g = w.create_body(s);
    BodyParams d;
    uint32_t b1 = w.create_body(d), b2 = w.create_body(d), b3 = w.create_body(d);
    PrismaticJointParams pj;
    pj.body_1 = g;
    pj.body_2 = b1;
    RevoluteJointParams rj;
    rj.body_1 = g;
    rj.body_2 = b2;
    FixedJointParams fj;
    fj.body_1 = g;
    fj.body_2 = b3;
    uint32_t jp = w.create_prismatic_joint(pj), jr = w.create_revolute_joint(rj), jf = w.create_fixed_joint(fj);
    ASSERT_TRUE(jp == 0 && jr == 1 && jf == 2 && w.joints.count() == 3);
    ASSERT_TRUE(w.joints.type[jp] == PRISMATIC && w.joints.type[jr] == REVOLUTE && w.joints.type[jf] == FIXED);
    const ConstraintType R = ConstraintType::ROTATIONAL, P = ConstraintType::POSITIONAL;
    const std::vector<ConstraintType> want = {R, P, P, R, P, R, R, R, P};
    ASSERT_TRUE(w.constraints.count() == want.size());
    for (size_t i = 0; i < want.size(); ++i) ASSERT_TRUE(w.constraints.type[i] == want[i]);
    ASSERT_TRUE(w.constraints.body_2[0] == b1 && w.constraints.body_2[3] == b2 && w.constraints.body_2[7] == b3);
    w.step();
    // the fixed joint holds b3 at the anchor: its positional row carries the weight
    ASSERT_TRUE(m3d::abs(w.constraints.lambda[8]) > 0.0);
    ASSERT_NEAR(w.bodies.position[b3].y, 0.0, 1e-3);
}

// joints.target_position written through the mirror drives a position-actuated prismatic
// joint there; current_position follows.
TEST(world_joint_target_write) {
    World w(1.0 / 60.0, 20);
    w.set_option(RBGPU_OPT_JOINT_MODE, 1);   // multipliers reset per substep: the drive converges
    BodyParams s;
    s.type = STATIC;
    s.mass = 0.0;
    uint32_t rail = w.create_body(s);
    BodyParams d;
    uint32_t car = w.create_body(d);
    PrismaticJointParams pj;
    pj.body_1 = rail;
    pj.body_2 = car;
    pj.moving_axis = vec3(1, 0, 0);
    pj.actuation_type = POSITION;
    pj.compliance = 1e-5;
    pj.damping = 0.5;
    uint32_t j = w.create_prismatic_joint(pj);
    for (int f = 0; f < 30; ++f) w.step();
    ASSERT_NEAR(w.joints.current_position[j], 0.0, 0.02);
    w.joints.target_position[j] = 0.5;
    for (int f = 0; f < 120; ++f) w.step();
    ASSERT_NEAR(w.joints.target_position[j], 0.5, 1e-7);
    // the reference measures current_position as (p1 - p2) . axis and drives (p2 - p1) . axis
    // to the target (Joint.cpp:36-70): at rest the two have opposite signs
    ASSERT_NEAR(w.joints.current_position[j], -0.5, 0.05);
    ASSERT_NEAR(m3d::abs(w.bodies.position[car].x), 0.5, 0.05);
    ASSERT_NEAR(w.bodies.position[car].y, 0.0, 0.01);
}

// bodies.force / torque: create_body's load is the weight, and the force is the whole
// external load (Body.cpp:28): writing zero holds a body still, writing the weight upwards
// lifts it at g, a torque spins it up about its axis.
TEST(world_body_force_and_torque) {
    World w(1.0 / 60.0, 20);
    BodyParams bp;
    bp.mass = 2.0;
    bp.position = vec3{0, 0, 0};   // near the origin: FP32 keeps the per-substep increments
    uint32_t b = w.create_body(bp);
    ASSERT_NEAR(w.bodies.force[b].y, -9.8f * 2.0, 1e-12);
    ASSERT_NEAR(w.bodies.inverse_mass[b], 0.5, 1e-15);
    ASSERT_NEAR(w.bodies.inverse_inertia_tensor[b].xx, 1.0, 1e-15);
    w.bodies.force[b] = vec3(0.0);
    w.bodies.torque[b] = vec3(0.0, 1.0, 0.0);
    for (int f = 0; f < 30; ++f) w.step();
    ASSERT_TRUE(w.bodies.position[b].y == 0.0 && w.bodies.linear_velocity[b].y == 0.0);
    ASSERT_NEAR(w.bodies.angular_velocity[b].y, 0.5, 2e-3);   // 1 N m / 1 kg m^2 for 0.5 s
    w.bodies.force[b] = vec3(0.0, 9.8f * 2.0, 0.0);
    for (int f = 0; f < 30; ++f) w.step();
    // 600 substeps of 1/1200 s at +g: h^2 g 600 * 601 / 2 = 1.2270 m (FP32: 1.2265)
    ASSERT_NEAR(w.bodies.position[b].y, 1.2270, 0.002);
}

// Writes to the body mirror reach the GPU on the next step.
TEST(world_body_write_through) {
    World w(1.0 / 60.0, 20);
    uint32_t b = add_dynamic_sphere(w, vec3{0, 10, 0});
    w.step();
    w.bodies.position[b] = vec3{5, 20, 0};
    w.bodies.linear_velocity[b] = vec3{1, 0, 0};
    w.step();
    ASSERT_NEAR(w.bodies.position[b].x, 5.0 + 1.0 / 60.0, 1e-4);
    ASSERT_TRUE(w.bodies.position[b].y < 20.0 && w.bodies.position[b].y > 19.9);
}

TEST(world_timestep_change_between_steps) {
    World w(1.0 / 60.0, 20);
    uint32_t s = add_dynamic_sphere(w, vec3{0, 10, 0});
    w.step();
    const scalar v1 = w.bodies.linear_velocity[s].y;
    w.substeps = 10;
    w.timestep = 1.0 / 120.0;
    w.step();
    // FP32 positions near 10 m differenced over a 1/1200 s substep: about 1e-3 m/s of noise
    ASSERT_NEAR(w.bodies.linear_velocity[s].y - v1, -9.8 / 120.0, 3e-3);
}

TEST(world_error_is_an_exception) {
    World w;
    bool thrown = false;
    try {
        ColliderParams cp;
        cp.body_id = 42;   // no such body
        w.create_collider(cp);
    } catch (const Error &) {
        thrown = true;
    }
    ASSERT_TRUE(thrown);
}

// The reference demo's pendulum (src/visr_demo/main.cpp:168-185) on the XPBD joint mode.
TEST(revolute_pendulum_demo) {
    World w;
    w.set_option(RBGPU_OPT_JOINT_MODE, 1);
    BodyParams pp;
    pp.type = STATIC;
    pp.mass = 0.0;
    pp.position = vec3(0.0, 4.0, 0.0);
    uint32_t pivot = w.create_body(pp);
    uint32_t bob = add_dynamic_sphere(w, vec3(1.8, 4.0, 0.0), 0.3, 20.0);
    RevoluteJointParams rj;
    rj.body_1 = pivot;
    rj.body_2 = bob;
    rj.aligned_axis = vec3(0, 0, 1);
    rj.limit_axis = vec3(1, 0, 0);
    rj.r_2 = vec3(-1.8, 0, 0);
    w.create_revolute_joint(rj);
    scalar lowest = 10.0, worst = 0.0;
    for (int f = 0; f < 60; ++f) {
        w.step();
        vec3 p = w.bodies.position[w.bodies.index_of(bob)];
        worst = std::fmax(worst, m3d::abs(m3d::magnitude(p - vec3(0.0, 4.0, 0.0)) - 1.8));
        lowest = std::fmin(lowest, p.y);
    }
    ASSERT_TRUE(worst < 0.02);
    ASSERT_TRUE(lowest < 2.5);
}

TEST(spherical_chain) {
    World w;
    w.set_option(RBGPU_OPT_JOINT_MODE, 1);
    BodyParams ap;
    ap.type = STATIC;
    ap.mass = 0.0;
    ap.position = vec3(0.0, 5.0, 0.0);
    uint32_t prev = w.create_body(ap);
    vec3 tip(0.0, 0.0, 0.0);
    for (int k = 0; k < 4; ++k) {
        uint32_t b = add_dynamic_sphere(w, vec3(0.6 * (k + 1), 5.0, 0.0), 0.1);
        SphericalJointParams sj;
        sj.body_1 = prev;
        sj.body_2 = b;
        sj.r_1 = tip;
        sj.r_2 = vec3(-0.3, 0.0, 0.0);
        w.create_spherical_joint(sj);
        prev = b;
        tip = vec3(0.3, 0.0, 0.0);
    }
    for (int f = 0; f < 60; ++f) w.step();
    vec3 end = w.bodies.position[w.bodies.index_of(prev)];
    ASSERT_TRUE(end.y < 4.0);
    ASSERT_TRUE(m3d::magnitude(end - vec3(0.0, 5.0, 0.0)) < 2.45);
}

// A box-shaped convex hull (12 CCW triangles) dropped on the ground plane comes to rest on
// a face: centre 0.3 above the plane (its half height).
TEST(convex_hull_rests_on_plane) {
    static const vec3 verts[8] = {vec3(-0.4, -0.3, -0.5), vec3(0.4, -0.3, -0.5), vec3(0.4, 0.3, -0.5),
                                  vec3(-0.4, 0.3, -0.5),  vec3(-0.4, -0.3, 0.5), vec3(0.4, -0.3, 0.5),
                                  vec3(0.4, 0.3, 0.5),    vec3(-0.4, 0.3, 0.5)};
    static const uint32_t tris[36] = {0, 3, 2, 0, 2, 1, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4,
                                      3, 7, 6, 3, 6, 2, 0, 4, 7, 0, 7, 3, 1, 2, 6, 1, 6, 5};
    rbc::ConvexHullData *hd = rbc::convex_hull_data_create(verts, 8, tris, 12);
    rbc::ConvexHull hull(hd);
    ASSERT_NEAR(rbc::compute_volume(hull), 0.48, 1e-12);
    World w;
    BodyParams g;
    g.type = STATIC;
    g.mass = 0.0;
    uint32_t gid = w.create_body(g);
    ColliderParams gc;
    gc.shape = rbc::Plane(vec3(0, 1, 0), 0.0);
    gc.body_id = gid;
    gc.is_static = true;
    w.create_collider(gc);
    ASSERT_TRUE(w.broad_phase_state.objects.empty());   // planes stay out of the broad phase
    BodyParams bp;
    bp.mass = 1000.0 * rbc::compute_volume(hull);
    smat3 I = rbc::compute_inertia_tensor(hull);
    bp.inertia_tensor = smat3(I.xx * 1000.0, I.yy * 1000.0, I.zz * 1000.0, 0, 0, 0);
    bp.position = vec3(0.0, 1.0, 0.0);
    uint32_t b = w.create_body(bp);
    ColliderParams hc;
    hc.shape = rbc::ConvexHull(hd);
    hc.body_id = b;
    hc.restitution = 0.0;
    w.create_collider(hc);
    for (int f = 0; f < 120; ++f) w.step();
    vec3 p = w.bodies.position[w.bodies.index_of(b)];
    ASSERT_NEAR(p.y, 0.3, 0.01);
    ASSERT_TRUE(m3d::magnitude(w.bodies.linear_velocity[w.bodies.index_of(b)]) < 0.05);
    rbc::convex_hull_data_destroy(hd);
}

// The autotuner steps real frames; a twin stepped as many frames untuned ends in the same
// state.
TEST(autotune_keeps_the_trajectory) {
    World a(1.0 / 60.0, 10), b(1.0 / 60.0, 10);
    for (World *w : {&a, &b}) {
        add_static_box(*w, vec3(0.0, -2.0, 0.0));
        for (int i = 0; i < 6; ++i) add_dynamic_sphere(*w, vec3(0.3 * i, 1.0 + 1.1 * i, 0.0));
    }
    a.step();
    b.step();
    const rbgpu_tune_report r = b.autotune(1);
    ASSERT_TRUE(r.trials > 0 && r.frames == r.trials + 1);
    for (int f = 0; f < r.frames; ++f) a.step();
    for (uint32_t i = 1; i < 7; ++i) ASSERT_TRUE(a.bodies.position[i] == b.bodies.position[i]);
}

TEST_SUITE(
    RUN_TEST(world_default_construction),
    RUN_TEST(world_parameterized_construction),
    RUN_TEST(world_create_body_returns_valid_id),
    RUN_TEST(world_create_multiple_bodies),
    RUN_TEST(world_create_collider_registers_in_broad_phase),
    RUN_TEST(world_step_empty_world),
    RUN_TEST(world_step_single_body_no_collider),
    RUN_TEST(world_free_fall_moves_body_downward),
    RUN_TEST(world_free_fall_velocity_increases),
    RUN_TEST(world_static_body_does_not_move),
    RUN_TEST(world_sphere_lands_on_ground),
    RUN_TEST(world_sphere_does_not_tunnel_through_ground),
    RUN_TEST(world_two_spheres_collide_and_separate),
    RUN_TEST(world_collision_conserves_momentum_approximately),
    RUN_TEST(world_contacts_cleared_each_frame),
    RUN_TEST(world_no_contacts_when_bodies_separated),
    RUN_TEST(world_contact_fields),
    RUN_TEST(world_broad_phase_state_mirror),
    RUN_TEST(world_create_constraint_row),
    RUN_TEST(world_joint_rows_numbering),
    RUN_TEST(world_joint_target_write),
    RUN_TEST(world_body_force_and_torque),
    RUN_TEST(world_body_write_through),
    RUN_TEST(world_timestep_change_between_steps),
    RUN_TEST(world_error_is_an_exception),
    RUN_TEST(revolute_pendulum_demo),
    RUN_TEST(spherical_chain),
    RUN_TEST(convex_hull_rests_on_plane),
    RUN_TEST(autotune_keeps_the_trajectory)
)
