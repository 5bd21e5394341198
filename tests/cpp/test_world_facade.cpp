// The reference's World tests (tests/rbps/API/test_world.cpp) written against
// rbps_amd::World, the drop-in facade over the rbgpu C-ABI.  Built and run by
// tests/test_gpu_facade.py; prints one line per case and exits non-zero on failure.
#include <cmath>
#include <cstdio>
#include "rbps_amd/World.hpp"

using namespace rbps_amd;

static int failures = 0;
#define CHECK(cond)                                                        \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::printf("  FAILED %s:%d %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                    \
        }                                                                  \
    } while (0)

static uint32_t add_dynamic_sphere(World &w, vec3 pos, scalar r = 0.5, scalar m = 1.0, vec3 v = vec3(0.0)) {
    BodyParams bp;
    bp.position = pos;
    bp.linear_velocity = v;
    bp.mass = m;
    bp.type = DYNAMIC;
    uint32_t id = w.create_body(bp);
    ColliderParams cp;
    cp.body_id = id;
    cp.shape = Sphere(r);
    w.create_collider(cp);
    return id;
}
static uint32_t add_static_box(World &w, vec3 pos, vec3 half = vec3(10, 2, 10)) {
    BodyParams bp;
    bp.position = pos;
    bp.mass = 0.0;
    bp.type = STATIC;
    uint32_t id = w.create_body(bp);
    ColliderParams cp;
    cp.body_id = id;
    cp.shape = Box(half);
    w.create_collider(cp);
    return id;
}
static scalar mag(vec3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

int main() {
    {
        std::printf("world_parameterized_construction\n");
        World w(1.0 / 120.0, 10);
        CHECK(std::fabs(w.timestep - 1.0 / 120.0) < 1e-9 && w.substeps == 10);
    }
    {
        std::printf("world_step_empty_world\n");
        World w;
        w.step();
        w.step();
    }
    {
        std::printf("world_free_fall\n");
        World w(1.0 / 60.0, 20);
        uint32_t s = w.bodies.index_of(add_dynamic_sphere(w, vec3(0, 10, 0)));
        scalar y0 = w.bodies.position[s].y;
        w.step();
        scalar vy1 = w.bodies.linear_velocity[s].y;
        CHECK(w.bodies.position[s].y < y0);
        w.step();
        CHECK(w.bodies.linear_velocity[s].y < vy1);
    }
    {
        std::printf("world_static_body_does_not_move\n");
        World w;
        uint32_t s = w.bodies.index_of(add_static_box(w, vec3(0, 0, 0)));
        for (int i = 0; i < 10; ++i) w.step();
        CHECK(std::fabs(w.bodies.position[s].y) < 1e-6);
    }
    {
        std::printf("world_sphere_lands_on_ground\n");
        World w(1.0 / 60.0, 20);
        add_static_box(w, vec3(0, -3, 0));
        uint32_t s = w.bodies.index_of(add_dynamic_sphere(w, vec3(0, 1, 0), 0.5));
        w.auto_sync(false);
        for (int i = 0; i < 180; ++i) w.step();
        w.sync_bodies();
        CHECK(mag(w.bodies.linear_velocity[s]) < 2.0);
        CHECK(w.bodies.position[s].y > -4.0);
    }
    {
        std::printf("world_two_spheres_collide_and_separate\n");
        World w(1.0 / 60.0, 20);
        uint32_t a = add_dynamic_sphere(w, vec3(-2, 0, 0), 0.5, 1.0, vec3(5, 0, 0));
        uint32_t b = add_dynamic_sphere(w, vec3(2, 0, 0), 0.5, 1.0, vec3(-5, 0, 0));
        for (int i = 0; i < 30; ++i) w.step();
        CHECK(w.bodies.linear_velocity[a].x < 0.0 || w.bodies.linear_velocity[b].x > 0.0);
    }
    {
        std::printf("world_timestep_change_between_steps\n");
        World w(1.0 / 60.0, 20);
        uint32_t s = add_dynamic_sphere(w, vec3(0, 10, 0));
        w.step();
        w.substeps = 10;
        w.timestep = 1.0 / 120.0;
        w.step();
        CHECK(w.bodies.position[s].y < 10.0);
    }
    {
        std::printf("world_no_contacts_when_separated\n");
        World w;
        add_dynamic_sphere(w, vec3(0, 100, 0));
        add_dynamic_sphere(w, vec3(100, 0, 0));
        w.step();
        CHECK(w.contact_count() == 0);
    }
    {
        std::printf("world_error_is_an_exception\n");
        World w;
        bool thrown = false;
        try {
            ColliderParams cp;
            cp.body_id = 42;   // no such body
            w.create_collider(cp);
        } catch (const Error &e) {
            thrown = true;
        }
        CHECK(thrown);
    }
    {
        // the reference demo's pendulum (src/visr_demo/main.cpp:168-185) on the XPBD joint mode
        std::printf("revolute_pendulum_demo\n");
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
            vec3 d(p.x, p.y - 4.0, p.z);
            worst = std::fmax(worst, std::fabs(mag(d) - 1.8));
            lowest = std::fmin(lowest, p.y);
        }
        CHECK(worst < 0.02);
        CHECK(lowest < 2.5);
        CHECK(w.joint_positions().size() == 1);
    }
    {
        std::printf("spherical_chain\n");
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
        CHECK(end.y < 4.0);
        CHECK(std::sqrt(end.x * end.x + (end.y - 5.0) * (end.y - 5.0) + end.z * end.z) < 2.45);
    }
    {
        // a box-shaped convex hull (12 CCW triangles) dropped on the ground plane comes to
        // rest on a face: centre 0.3 above the plane (its half height), upright
        std::printf("convex_hull_rests_on_plane\n");
        static const vec3 verts[8] = {vec3(-0.4, -0.3, -0.5), vec3(0.4, -0.3, -0.5), vec3(0.4, 0.3, -0.5),
                                      vec3(-0.4, 0.3, -0.5), vec3(-0.4, -0.3, 0.5), vec3(0.4, -0.3, 0.5),
                                      vec3(0.4, 0.3, 0.5), vec3(-0.4, 0.3, 0.5)};
        static const uint32_t tris[36] = {0, 3, 2, 0, 2, 1, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4,
                                          3, 7, 6, 3, 6, 2, 0, 4, 7, 0, 7, 3, 1, 2, 6, 1, 6, 5};
        ConvexHullData *hd = convex_hull_data_create(verts, 8, tris, 12);
        ConvexHull hull(hd);
        CHECK(std::fabs(compute_volume(hull) - 0.48) < 1e-12);
        World w;
        BodyParams g;
        g.type = STATIC;
        g.mass = 0.0;
        uint32_t gid = w.create_body(g);
        ColliderParams gc;
        gc.shape = Plane(vec3(0, 1, 0), 0.0);
        gc.body_id = gid;
        gc.is_static = true;
        w.create_collider(gc);
        BodyParams bp;
        bp.mass = 1000.0 * compute_volume(hull);
        smat3 I = compute_inertia_tensor(hull);
        bp.inertia_tensor = smat3(I.xx * 1000.0, I.yy * 1000.0, I.zz * 1000.0, 0, 0, 0);
        bp.position = vec3(0.0, 1.0, 0.0);
        uint32_t b = w.create_body(bp);
        ColliderParams hc;
        hc.shape = ConvexHull(hd);
        hc.body_id = b;
        hc.restitution = 0.0;
        w.create_collider(hc);
        for (int f = 0; f < 120; ++f) w.step();
        vec3 p = w.bodies.position[w.bodies.index_of(b)];
        CHECK(std::fabs(p.y - 0.3) < 0.01);
        CHECK(mag(w.bodies.linear_velocity[w.bodies.index_of(b)]) < 0.05);
        convex_hull_data_destroy(hd);
    }
    {
        // extension: the autotuner steps real frames; a twin stepped as many frames untuned
        // ends in the same state
        std::printf("autotune_keeps_the_trajectory\n");
        World a(1.0 / 60.0, 10), b(1.0 / 60.0, 10);
        for (World *w : {&a, &b}) {
            add_static_box(*w, vec3(0.0, -2.0, 0.0));
            for (int i = 0; i < 6; ++i) add_dynamic_sphere(*w, vec3(0.3 * i, 1.0 + 1.1 * i, 0.0));
        }
        a.step();
        b.step();
        const rbgpu_tune_report r = b.autotune(1);
        CHECK(r.trials > 0 && r.frames == r.trials + 1);
        for (int f = 0; f < r.frames; ++f) a.step();
        for (uint32_t i = 1; i < 7; ++i) {
            const vec3 pa = a.bodies.position[a.bodies.index_of(i)], pb = b.bodies.position[b.bodies.index_of(i)];
            CHECK(pa.x == pb.x && pa.y == pb.y && pa.z == pb.z);
        }
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "ALL PASSED", failures);
    return failures ? 1 : 0;
}
