// World-level scene helpers of the reference's tests (include/tests/rbps/pipeline_helpers.hpp:
// add_dynamic_sphere :70-90, add_static_box :92-110) over the MI355X rbps::World.  Found
// first on the include path, so the reference's tests/rbps/API/test_world.cpp compiles
// against rbps/API/World.hpp of this repository unchanged.
#pragma once
#include "rbps/API/World.hpp"

namespace test {

// A dynamic sphere: body + collider; returns the body id.
inline uint32_t add_dynamic_sphere(rbps::World &w, const m3d::vec3 &pos, m3d::scalar radius = 0.5,
                                   m3d::scalar mass = 1.0, const m3d::vec3 &vel = m3d::vec3(0, 0, 0)) {
    rbps::BodyParams body;
    body.type = rbps::BodyType::DYNAMIC;
    body.mass = mass;
    body.position = pos;
    body.linear_velocity = vel;
    const uint32_t id = w.create_body(body);
    rbps::ColliderParams col;
    col.shape = rbc::Sphere(radius);
    col.body_id = id;
    w.create_collider(col);
    return id;
}

// A static box (zero mass), by default the wide (10, 2, 10) ground slab; returns the body id.
inline uint32_t add_static_box(rbps::World &w, const m3d::vec3 &pos,
                               const m3d::vec3 &half_extents = m3d::vec3(10, 2, 10)) {
    rbps::BodyParams body;
    body.type = rbps::BodyType::STATIC;
    body.mass = 0.0;
    body.position = pos;
    const uint32_t id = w.create_body(body);
    rbps::ColliderParams col;
    col.shape = rbc::Box(half_extents);
    col.body_id = id;
    w.create_collider(col);
    return id;
}

}  // namespace test
