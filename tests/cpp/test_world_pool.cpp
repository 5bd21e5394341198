// WorldPool over the MI355X batch, in the cases of the reference's multi-world design
// (docs/design/multi-world/multi-world.md:63-66): (a) N = 8 worlds of different scenes
// stepped together equal the same worlds stepped alone, bit for bit (bodies, joints,
// contacts, broadphase pairs); mirror writes made from parallel_each threads reach the GPU;
// (c) the pool's step against the worlds' own steps (timed, printed).  Built and run by
// tests/test_gpu_facade.py.
#include <chrono>
#include <cstdio>
#include <vector>

#include "rbps/api/WorldPool.hpp"
#include "harness.hpp"
#include "tests/rbps/pipeline_helpers.hpp"

using namespace rbps;
using namespace m3d;

static const vec3 kHullVerts[8] = {vec3(-0.4, -0.3, -0.5), vec3(0.4, -0.3, -0.5), vec3(0.4, 0.3, -0.5),
                                   vec3(-0.4, 0.3, -0.5),  vec3(-0.4, -0.3, 0.5), vec3(0.4, -0.3, 0.5),
                                   vec3(0.4, 0.3, 0.5),    vec3(-0.4, 0.3, 0.5)};
static const uint32_t kHullTris[36] = {0, 3, 2, 0, 2, 1, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4,
                                       3, 7, 6, 3, 6, 2, 0, 4, 7, 0, 7, 3, 1, 2, 6, 1, 6, 5};
static rbc::ConvexHullData *hull_data() {
    static rbc::ConvexHullData *d = rbc::convex_hull_data_create(kHullVerts, 8, kHullTris, 12);
    return d;
}

// World i of the pool: a slab, a column of 2 + i spheres, two boxes, a pendulum on a
// revolute joint in odd worlds, a convex hull on a plane in worlds divisible by 3.
static void build(World &w, int i) {
    test::add_static_box(w, vec3{0, -2, 0});
    for (int k = 0; k < 2 + i; ++k) test::add_dynamic_sphere(w, vec3{0.05 * k, 0.6 + 1.05 * k, 0.02 * i}, 0.5);
    for (int k = 0; k < 2; ++k) {
        BodyParams bp;
        bp.position = vec3{2.0 + 1.2 * k, 0.5 + 0.3 * i, 0.0};
        bp.orientation = quat{0.9950041652780258, 0.0, 0.09983341664682815, 0.0};
        uint32_t b = w.create_body(bp);
        ColliderParams cp;
        cp.shape = rbc::Box(vec3{0.5, 0.5, 0.5});
        cp.body_id = b;
        w.create_collider(cp);
    }
    if (i % 2) {
        BodyParams pp;
        pp.type = STATIC;
        pp.mass = 0.0;
        pp.position = vec3{-3.0, 4.0, 0.0};
        uint32_t pivot = w.create_body(pp);
        uint32_t bob = test::add_dynamic_sphere(w, vec3{-1.5, 4.0, 0.0}, 0.3, 5.0);
        RevoluteJointParams rj;
        rj.body_1 = pivot;
        rj.body_2 = bob;
        rj.aligned_axis = vec3(0, 0, 1);
        rj.limit_axis = vec3(1, 0, 0);
        rj.r_2 = vec3(-1.5, 0, 0);
        rj.damping = 0.1 * i;
        w.create_revolute_joint(rj);
    }
    if (i % 3 == 0) {
        BodyParams g;
        g.type = STATIC;
        g.mass = 0.0;
        g.position = vec3{6.0, 0.0, 0.0};
        uint32_t gid = w.create_body(g);
        ColliderParams gc;
        gc.shape = rbc::Plane(vec3(0, 1, 0), 0.0);
        gc.body_id = gid;
        gc.is_static = true;
        w.create_collider(gc);
        BodyParams hb;
        hb.mass = 480.0;
        hb.position = vec3{6.0, 1.0 + 0.1 * i, 0.0};
        uint32_t h = w.create_body(hb);
        ColliderParams hc;
        hc.shape = rbc::ConvexHull(hull_data());
        hc.body_id = h;
        w.create_collider(hc);
    }
}

static void same(const World &a, const World &b, const char *tag) {
    ASSERT_TRUE(a.bodies.count() == b.bodies.count());
    for (uint32_t i = 0; i < a.bodies.count(); ++i) {
        if (!(a.bodies.position[i] == b.bodies.position[i]) || !(a.bodies.orientation[i] == b.bodies.orientation[i]) ||
            !(a.bodies.linear_velocity[i] == b.bodies.linear_velocity[i]) ||
            !(a.bodies.angular_velocity[i] == b.bodies.angular_velocity[i]))
            harness::fail(__FILE__, __LINE__, std::string(tag) + ": body " + std::to_string(i) + " differs");
    }
    ASSERT_TRUE(a.contacts.n_contacts == b.contacts.n_contacts);
    for (uint32_t i = 0; i < a.contacts.n_contacts; ++i) {
        ASSERT_TRUE(a.contacts.body_a[i] == b.contacts.body_a[i] && a.contacts.collider_b[i] == b.contacts.collider_b[i]);
        ASSERT_TRUE(a.contacts.normal[i] == b.contacts.normal[i] && a.contacts.normal_lambda[i] == b.contacts.normal_lambda[i]);
    }
    ASSERT_TRUE(a.broad_phase_state.pairs.size() == b.broad_phase_state.pairs.size());
    for (size_t i = 0; i < a.broad_phase_state.pairs.size(); ++i)
        ASSERT_TRUE(a.broad_phase_state.pairs[i].id_a == b.broad_phase_state.pairs[i].id_a &&
                    a.broad_phase_state.pairs[i].id_b == b.broad_phase_state.pairs[i].id_b);
    ASSERT_TRUE(a.joints.count() == b.joints.count());
    for (uint32_t j = 0; j < a.joints.count(); ++j) ASSERT_TRUE(a.joints.current_position[j] == b.joints.current_position[j]);
    for (uint32_t c = 0; c < a.constraints.count(); ++c) ASSERT_TRUE(a.constraints.lambda[c] == b.constraints.lambda[c]);
}

TEST(pool_worlds_equal_worlds_alone) {
    const int N = 8;
    WorldPool pool(N);
    ASSERT_TRUE(pool.size() == N);
    pool.parallel_each([](World &w, size_t i) { build(w, (int)i); }, 1);   // creation in world order
    std::vector<World> alone;
    for (int i = 0; i < N; ++i) {
        alone.emplace_back(1.0 / 60.0, 20);
        build(alone.back(), i);
    }
    for (int i = 0; i < N; ++i) ASSERT_TRUE(pool.worlds[i].bodies.count() == alone[i].bodies.count());
    ASSERT_TRUE(pool.worlds[3].broad_phase_state.objects.size() == alone[3].broad_phase_state.objects.size());
    for (int f = 0; f < 60; ++f) {
        pool.parallel_step(1.0 / 60.0);
        for (World &w : alone) w.step();
    }
    for (int i = 0; i < N; ++i) same(pool.worlds[i], alone[i], "after 60 frames");
    ASSERT_TRUE(pool.worlds[7].contacts.n_contacts > 0);
    // writes into the mirrors from parallel_each threads: a kick and a joint target per world
    pool.parallel_each([](World &w, size_t i) {
        w.bodies.linear_velocity[1] = vec3(0.5 + 0.1 * i, 2.0, 0.0);
        if (w.joints.count()) w.joints.target_speed[0] = 0.0;
    });
    for (int i = 0; i < N; ++i) {
        alone[i].bodies.linear_velocity[1] = vec3(0.5 + 0.1 * i, 2.0, 0.0);
        if (alone[i].joints.count()) alone[i].joints.target_speed[0] = 0.0;
    }
    for (int f = 0; f < 30; ++f) {
        pool.parallel_step(1.0 / 60.0);
        for (World &w : alone) w.step();
    }
    for (int i = 0; i < N; ++i) same(pool.worlds[i], alone[i], "after the writes");
}

TEST(pool_worlds_step_together) {
    WorldPool pool(2);
    test::add_dynamic_sphere(pool.worlds[0], vec3{0, 1, 0});
    bool thrown = false;
    try {
        pool.worlds[0].step();
    } catch (const Error &) {
        thrown = true;
    }
    ASSERT_TRUE(thrown);
    pool.worlds[1].substeps = 10;   // the pool's worlds share one substep count
    thrown = false;
    try {
        pool.parallel_step(1.0 / 60.0);
    } catch (const Error &) {
        thrown = true;
    }
    ASSERT_TRUE(thrown);
}

TEST(pool_step_time_against_worlds_alone) {
    const int N = 64, F = 20;
    WorldPool pool(N);
    pool.parallel_each([](World &w, size_t i) { build(w, (int)(i % 8)); }, 1);
    std::vector<World> alone;
    for (int i = 0; i < N; ++i) {
        alone.emplace_back(1.0 / 60.0, 20);
        build(alone.back(), i % 8);
        alone.back().auto_sync(false);
    }
    pool.auto_sync(false);
    pool.parallel_step(1.0 / 60.0);
    for (World &w : alone) w.step();
    auto t0 = std::chrono::steady_clock::now();
    for (int f = 0; f < F; ++f) pool.parallel_step(1.0 / 60.0);
    pool.sync();
    auto t1 = std::chrono::steady_clock::now();
    for (int f = 0; f < F; ++f)
        for (World &w : alone) w.step();
    for (World &w : alone) w.sync();
    auto t2 = std::chrono::steady_clock::now();
    const double a = std::chrono::duration<double, std::milli>(t1 - t0).count() / F;
    const double b = std::chrono::duration<double, std::milli>(t2 - t1).count() / F;
    std::printf("\n  %d worlds: pool %.2f ms per frame, worlds alone %.2f ms per frame (%.1fx)\n", N, a, b, b / a);
    for (int i = 0; i < N; ++i) same(pool.worlds[i], alone[i], "timed frames");
}

TEST_SUITE(
    RUN_TEST(pool_worlds_equal_worlds_alone),
    RUN_TEST(pool_worlds_step_together),
    RUN_TEST(pool_step_time_against_worlds_alone)
)
