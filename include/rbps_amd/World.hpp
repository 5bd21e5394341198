// rbps::World over the MI355X pipeline: a source-compatible stand-in for the reference's
// World (include/rbps/API/World.hpp:87-212) whose step() runs the rbgpu C-ABI
// (include/rbgpu.h).  Include it as "rbps/API/World.hpp" (include/rbps/API/World.hpp
// forwards here) or directly.
//
// The reference's names, namespaces and argument meaning:
//   m3d   scalar, vec3, quat, smat3, dot, cross, magnitude, normalize, abs
//   rbc   Sphere, Box, Capsule, Plane, ConvexHull(Data), convex_hull_data_create/destroy,
//         compute_volume, compute_inertia_tensor, Shape, AABB, BPObject, BroadPhasePair,
//         BroadPhaseState
//   rbps  World{timestep, substeps, bodies, constraints, joints, contacts, colliders,
//         broad_phase_state}, BodyParams, ColliderParams, ConstraintParams, the joint params,
//         create_body / create_collider / create_constraint / create_*_joint, step()
// The public collections are host mirrors of the state on the GPU:
//   * step() first pushes what the caller changed in the writable fields (the body state,
//     joints' target_position / target_speed, constraint rows' direction / magnitude), then
//     runs the frame, then (auto_sync, default on) refreshes every mirror; auto_sync(false)
//     and sync() batch the read-backs of large worlds;
//   * contacts, colliders and broad_phase_state are read-only mirrors; contacts holds the
//     last substep's contacts; broad_phase_state.pairs the frame's pairs after the
//     narrowphase driver's filters;
//   * the reference's silent failure modes become exceptions (rbps::Error) carrying the
//     library's message.
// Header-only; link with -lrbgpu.
#pragma once
#include <algorithm>
#include <cmath>
#include <initializer_list>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../rbgpu.h"

// ------------------------------------------------------------------ math (include/math3d)
namespace m3d {

using scalar = double;
constexpr scalar EPSILON = 1e-9;   // include/math3d/core.hpp:49

struct vec3 {
    scalar x = 0, y = 0, z = 0;
    vec3() = default;
    vec3(scalar a, scalar b, scalar c) : x(a), y(b), z(c) {}
    explicit vec3(scalar d) : x(d), y(d), z(d) {}
};
inline vec3 operator+(const vec3 &a, const vec3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(const vec3 &a, const vec3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator-(const vec3 &a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator*(const vec3 &a, scalar s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator*(scalar s, const vec3 &a) { return {a.x * s, a.y * s, a.z * s}; }
inline bool operator==(const vec3 &a, const vec3 &b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool operator!=(const vec3 &a, const vec3 &b) { return !(a == b); }
inline scalar dot(const vec3 &a, const vec3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(const vec3 &a, const vec3 &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline scalar magnitude(const vec3 &a) { return std::sqrt(dot(a, a)); }
inline scalar length_sq(const vec3 &a) { return dot(a, a); }
inline vec3 normalize(const vec3 &a) {   // zero below EPSILON squared length (vec3.hpp:197-203)
    const scalar l2 = length_sq(a);
    return l2 < EPSILON ? vec3(0.0) : a * (1.0 / std::sqrt(l2));
}
inline scalar abs(scalar a) { return std::fabs(a); }

struct quat {
    scalar w = 1, x = 0, y = 0, z = 0;
    quat() = default;
    quat(scalar a, scalar b, scalar c, scalar d) : w(a), x(b), y(c), z(d) {}
};
inline bool operator==(const quat &a, const quat &b) { return a.w == b.w && a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool operator!=(const quat &a, const quat &b) { return !(a == b); }

// Symmetric 3x3: xx yy zz xy xz yz.
struct smat3 {
    scalar xx = 0, yy = 0, zz = 0, xy = 0, xz = 0, yz = 0;
    smat3() = default;
    smat3(scalar a, scalar b, scalar c, scalar d, scalar e, scalar f) : xx(a), yy(b), zz(c), xy(d), xz(e), yz(f) {}
    // adjugate / determinant
    smat3 inverse() const {
        const scalar a = yy * zz - yz * yz, b = xz * yz - xy * zz, c = xy * yz - xz * yy;
        const scalar det = xx * a + xy * b + xz * c;
        if (det == 0) return smat3{};
        const scalar k = 1.0 / det;
        return smat3(a * k, (xx * zz - xz * xz) * k, (xx * yy - xy * xy) * k, b * k, c * k, (xy * xz - xx * yz) * k);
    }
};

}  // namespace m3d

namespace rbps {
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
inline void check(int rc) {
    if (rc != RBGPU_OK) throw Error(rc, rbgpu_last_error());
}
}  // namespace rbps

// ------------------------------------------------------------------ shapes and broadphase (include/rbc)
namespace rbc {
using m3d::scalar;
using m3d::vec3;

struct Sphere {
    scalar radius = 0.5;
    Sphere() = default;
    Sphere(scalar r) : radius(r) {}
};
struct Box {
    vec3 half_extents = vec3(0.5, 0.5, 0.5);
    Box() = default;
    Box(const vec3 &h) : half_extents(h) {}
};
struct Capsule {
    scalar half_height = 0.5, radius = 0.25;
    Capsule() = default;
    Capsule(scalar hh, scalar r) : half_height(hh), radius(r) {}
};
struct Plane {
    vec3 normal = vec3(0.0, 1.0, 0.0);
    scalar d = 0.0;
    Plane() = default;
    Plane(const vec3 &n, scalar dd) : normal(n), d(dd) {}
};
// ConvexHullData (include/rbc/shapes/ConvexHull.hpp:79-98): non-owning over the caller's
// vertices and CCW triangles (triangles optional).  A world copies the data (rbgpu_add_hull)
// the first time a collider uses it; the arrays must stay valid until then.
struct ConvexHullData {
    const vec3 *vertices = nullptr;
    const uint32_t *face_indices = nullptr;
    uint32_t vert_count = 0, face_count = 0;
};
// convex_hull_data_create / convex_hull_data_destroy (ConvexHull.hpp:149-448)
inline ConvexHullData *convex_hull_data_create(const vec3 *vertices, uint32_t vert_count,
                                               const uint32_t *face_indices = nullptr, uint32_t face_count = 0) {
    return new ConvexHullData{vertices, face_indices, vert_count, face_indices ? face_count : 0u};
}
inline void convex_hull_data_destroy(ConvexHullData *d) { delete d; }
// ConvexHull (ConvexHull.hpp:101-121)
struct ConvexHull {
    const ConvexHullData *data = nullptr;
    ConvexHull() = default;
    explicit ConvexHull(const ConvexHullData *d) : data(d) {}
};
// compute_volume / compute_inertia_tensor (ConvexHull.hpp:575-745), unit density
inline scalar compute_volume(const ConvexHull &h) {
    double vol = 0.0;
    if (h.data)
        rbps::check(rbgpu_hull_mass_properties(&h.data->vertices[0].x, h.data->vert_count, h.data->face_indices,
                                               h.data->face_count, &vol, nullptr));
    return vol;
}
inline m3d::smat3 compute_inertia_tensor(const ConvexHull &h) {
    double I[6] = {0, 0, 0, 0, 0, 0};
    if (h.data)
        rbps::check(rbgpu_hull_mass_properties(&h.data->vertices[0].x, h.data->vert_count, h.data->face_indices,
                                               h.data->face_count, nullptr, I));
    return m3d::smat3(I[0], I[1], I[2], I[3], I[4], I[5]);
}

// Shape (ShapeTypes.hpp:69): the kinds this pipeline runs.
struct Shape {
    int kind = RBGPU_SPHERE;
    scalar dims[4] = {0.5, 0, 0, 0};
    const ConvexHullData *hull = nullptr;
    Shape() = default;
    Shape(const ConvexHull &h) : kind(RBGPU_HULL), dims{0, 0, 0, 0}, hull(h.data) {}
    Shape(const Sphere &s) : kind(RBGPU_SPHERE), dims{s.radius, 0, 0, 0} {}
    Shape(const Box &b) : kind(RBGPU_BOX), dims{b.half_extents.x, b.half_extents.y, b.half_extents.z, 0} {}
    Shape(const Capsule &c) : kind(RBGPU_CAPSULE), dims{c.half_height, c.radius, 0, 0} {}
    Shape(const Plane &p) : kind(RBGPU_PLANE), dims{p.normal.x, p.normal.y, p.normal.z, p.d} {}
};

// AABB, BroadPhasePair, BPObject, BroadPhaseState (AABB.hpp, BroadPhase.hpp:64-117)
struct AABB {
    vec3 min, max;
};
struct BroadPhasePair {
    uint32_t id_a, id_b;
};
struct BPObject {
    AABB fat_aabb;
    AABB tight_aabb;   // not kept by this pipeline (the fat box is what the sweep uses)
    uint32_t user_id = 0;
    bool alive = true;
};
struct BroadPhaseState {
    std::vector<BPObject> objects;
    std::vector<BroadPhasePair> pairs;
};

}  // namespace rbc

// ------------------------------------------------------------------ the world (include/rbps)
namespace rbps {
using m3d::quat;
using m3d::scalar;
using m3d::smat3;
using m3d::vec3;

enum BodyType { STATIC = 0, DYNAMIC = 1 };   // Body.hpp:11-15

// BodyParams (BodyAPI.hpp:22-31)
struct BodyParams {
    BodyType type = DYNAMIC;
    scalar mass = 1.0;
    smat3 inertia_tensor = smat3{1.0, 1.0, 1.0, 0, 0, 0};
    vec3 position = vec3(0.0);
    quat orientation = quat{1, 0, 0, 0};
    vec3 linear_velocity = vec3(0.0);
    vec3 angular_velocity = vec3(0.0);
};

// ColliderParams (ColliderAPI.hpp:17-27)
struct ColliderParams {
    rbc::Shape shape;
    uint32_t body_id = 0;
    bool is_static = false;
    vec3 local_pos = vec3(0, 0, 0);
    quat local_rot = quat(1, 0, 0, 0);
    scalar restitution = 0.5;
    scalar static_friction = 0.5;
    scalar dynamic_friction = 0.5;
};

// ConstraintType, ConstraintParams (Constraint.hpp, ConstraintAPI.hpp:18-26)
enum class ConstraintType { POSITIONAL = 0, ROTATIONAL = 1 };
struct ConstraintParams {
    size_t body_1 = 0, body_2 = 0;
    ConstraintType type = ConstraintType::POSITIONAL;
    vec3 r_1 = {0.0, 0.0, 0.0};
    vec3 r_2 = {0.0, 0.0, 0.0};
    scalar compliance = 0.0;
};

// Joints (Joint.hpp:25-47, JointAPI.hpp:22-68)
enum JointType { PRISMATIC = 0, REVOLUTE = 1, FIXED = 2, SPHERICAL = 3 };   // SPHERICAL: extension
enum JointActuationType { FREE = 0, POSITION = 1, SPEED = 2 };

struct PrismaticJointParams {
    size_t body_1 = 0, body_2 = 0;
    vec3 moving_axis = vec3(1, 0, 0);
    vec3 r_1 = vec3(0, 0, 0), r_2 = vec3(0, 0, 0);
    JointActuationType actuation_type = FREE;
    scalar compliance = 0.0, damping = 0.0;
    bool limited = false;
    scalar lower_limit = 0.0, upper_limit = 0.0;
};
struct RevoluteJointParams {
    size_t body_1 = 0, body_2 = 0;
    vec3 aligned_axis = vec3(1, 0, 0);
    vec3 limit_axis = vec3(0, 1, 0);
    vec3 r_1 = vec3(0, 0, 0), r_2 = vec3(0, 0, 0);
    JointActuationType actuation_type = FREE;
    scalar compliance = 0.0, damping = 0.0;
    bool limited = false;
    scalar lower_limit = 0.0, upper_limit = 0.0;
};
struct FixedJointParams {
    size_t body_1 = 0, body_2 = 0;
    vec3 r_1 = vec3(0, 0, 0), r_2 = vec3(0, 0, 0);
};
// Ball joint (no reference counterpart): attachment row with `compliance`; with `limited`,
// the angle between the bodies' `swing_axis` stays below `swing_limit`.
struct SphericalJointParams {
    size_t body_1 = 0, body_2 = 0;
    vec3 r_1 = vec3(0, 0, 0), r_2 = vec3(0, 0, 0);
    vec3 swing_axis = vec3(0, 1, 0);
    scalar compliance = 0.0;
    bool limited = false;
    scalar swing_limit = 0.0;
};

// Without removals an ID is its dense index (include/storage/Soa.hpp:405-416).
inline uint32_t dense_index(uint32_t id) { return id & 0x00ffffffu; }

// BodyCollection (Body.hpp:52-69): mirror of the public body fields.  position,
// orientation, linear_velocity, angular_velocity, force and torque are writable; type,
// mass, inverse_mass and the inertia tensors are fixed at creation.
struct BodyCollection {
    std::vector<BodyType> type;
    std::vector<scalar> mass, inverse_mass;
    std::vector<vec3> position, linear_velocity, angular_velocity, force, torque;
    std::vector<quat> orientation;
    std::vector<smat3> inertia_tensor, inverse_inertia_tensor;
    uint32_t count() const { return (uint32_t)position.size(); }
    uint32_t index_of(uint32_t id) const { return dense_index(id); }
};
// ColliderCollection (Collider.hpp): read-only mirror of the creation parameters.
struct ColliderCollection {
    std::vector<rbc::Shape> shape;
    std::vector<uint32_t> body_id;
    std::vector<bool> is_static;
    std::vector<vec3> local_pos;
    std::vector<quat> local_rot;
    std::vector<scalar> restitution, static_friction, dynamic_friction;
    uint32_t count() const { return (uint32_t)body_id.size(); }
    uint32_t index_of(uint32_t id) const { return dense_index(id); }
};
// JointCollection (Joint.hpp:82-106): target_position / target_speed are writable,
// current_position is refreshed after each step.
struct JointCollection {
    std::vector<JointType> type;
    std::vector<size_t> body_1, body_2;
    std::vector<scalar> target_position, target_speed, current_position;
    uint32_t count() const { return (uint32_t)type.size(); }
    uint32_t index_of(uint32_t id) const { return dense_index(id); }
};
// ConstraintCollection (Constraint.hpp:31-53): every joint's rows and the rows made by
// create_constraint, in creation order.  direction / magnitude of a create_constraint row
// are writable (a joint's rows follow its errors); lambda is refreshed after each step.
struct ConstraintCollection {
    std::vector<size_t> body_1, body_2;
    std::vector<ConstraintType> type;
    std::vector<vec3> r_1, r_2, direction;
    std::vector<scalar> magnitude, lambda, compliance;
    uint32_t count() const { return (uint32_t)type.size(); }
    uint32_t index_of(uint32_t id) const { return dense_index(id); }
};
// set_value / get_lagrange_multiplier (Constraint.cpp:5-12, :63-66) on the mirror: a row's
// direction and magnitude from one error vector (direction (1, 0, 0) below EPSILON); step()
// pushes the change.
inline void set_value(ConstraintCollection &cc, uint32_t i, vec3 value) {
    cc.magnitude[i] = m3d::magnitude(value);
    cc.direction[i] = cc.magnitude[i] > m3d::EPSILON ? m3d::normalize(value) : vec3(1.0, 0.0, 0.0);
}
inline scalar get_lagrange_multiplier(const ConstraintCollection &cc, uint32_t i) { return cc.lambda[i]; }

// ContactList (CollisionPipeline.hpp:46-139): the last substep's contacts, read-only.
struct ContactList {
    uint32_t n_contacts = 0;
    std::vector<uint32_t> body_a, body_b, collider_a, collider_b;
    std::vector<vec3> r_a_local, r_b_local, normal, point_on_a, point_on_b;
    std::vector<scalar> restitution, static_friction, dynamic_friction;
    std::vector<scalar> normal_lambda, tangent_lambda, penetration_depth, relative_velocity;
    std::vector<bool> collision;
};

struct World {
    scalar timestep = 1.0 / 60.0;
    int substeps = 20;
    BodyCollection bodies;
    ConstraintCollection constraints;
    JointCollection joints;
    ContactList contacts;
    ColliderCollection colliders;
    rbc::BroadPhaseState broad_phase_state;

    World() { create(); }
    World(scalar timestep_, int substeps_) : timestep(timestep_), substeps(substeps_) { create(); }
    ~World() {
        if (h_) rbgpu_world_destroy(h_);
    }
    World(const World &) = delete;
    World &operator=(const World &) = delete;

    // World::create_body (World.hpp:113-116)
    uint32_t create_body(BodyParams p) {
        push_changes();
        rbgpu_body_desc d{};
        d.type = p.type == STATIC ? RBGPU_STATIC : RBGPU_DYNAMIC;
        d.mass = p.mass;
        const scalar I[6] = {p.inertia_tensor.xx, p.inertia_tensor.yy, p.inertia_tensor.zz,
                             p.inertia_tensor.xy, p.inertia_tensor.xz, p.inertia_tensor.yz};
        for (int k = 0; k < 6; ++k) d.inertia[k] = I[k];
        set3(d.position, p.position);
        d.orientation[0] = p.orientation.w; d.orientation[1] = p.orientation.x;
        d.orientation[2] = p.orientation.y; d.orientation[3] = p.orientation.z;
        set3(d.linear_velocity, p.linear_velocity);
        set3(d.angular_velocity, p.angular_velocity);
        uint32_t id = 0;
        check(rbgpu_add_bodies(h_, 1, &d, &id));
        bodies.type.push_back(p.type);
        bodies.mass.push_back(p.mass);
        bodies.inverse_mass.push_back(p.type == STATIC ? 0.0 : 1.0 / p.mass);   // BodyAPI.hpp:58-59
        bodies.inertia_tensor.push_back(p.inertia_tensor);
        bodies.inverse_inertia_tensor.push_back(p.type == STATIC ? smat3{} : p.inertia_tensor.inverse());
        bodies.force.push_back(vec3(0.0, -9.8f * p.mass, 0.0));   // BodyAPI.hpp:61-62
        bodies.torque.push_back(vec3(0.0));
        shadow_.force.push_back(bodies.force.back());
        shadow_.torque.push_back(bodies.torque.back());
        bodies.position.push_back(p.position);
        bodies.orientation.push_back(p.orientation);
        bodies.linear_velocity.push_back(p.linear_velocity);
        bodies.angular_velocity.push_back(p.angular_velocity);
        shadow_bodies();
        return id;
    }

    // World::create_collider (World.hpp:124-132); inserts a non-plane collider into the
    // broad phase.
    uint32_t create_collider(ColliderParams p) {
        push_changes();
        rbgpu_collider_desc d{};
        d.shape = p.shape.kind;
        d.body_id = p.body_id;
        d.is_static = p.is_static ? 1 : 0;
        d.hull_index = p.shape.kind == RBGPU_HULL ? hull_id(p.shape.hull) : -1;
        for (int k = 0; k < 4; ++k) d.dims[k] = p.shape.dims[k];
        set3(d.local_pos, p.local_pos);
        d.local_rot[0] = p.local_rot.w; d.local_rot[1] = p.local_rot.x; d.local_rot[2] = p.local_rot.y;
        d.local_rot[3] = p.local_rot.z;
        d.restitution = p.restitution;
        d.static_friction = p.static_friction;
        d.dynamic_friction = p.dynamic_friction;
        uint32_t id = 0;
        check(rbgpu_add_colliders(h_, 1, &d, &id));
        colliders.shape.push_back(p.shape);
        colliders.body_id.push_back(p.body_id);
        colliders.is_static.push_back(p.is_static);
        colliders.local_pos.push_back(p.local_pos);
        colliders.local_rot.push_back(p.local_rot);
        colliders.restitution.push_back(p.restitution);
        colliders.static_friction.push_back(p.static_friction);
        colliders.dynamic_friction.push_back(p.dynamic_friction);
        if (p.shape.kind != RBGPU_PLANE) pull_broad_phase();
        return id;
    }

    // World::create_constraint (World.hpp:135-138, ConstraintAPI.hpp:37-64): a row with
    // direction (1, 0, 0) and magnitude 0 until the caller writes them.
    uint32_t create_constraint(ConstraintParams p) {
        push_changes();
        rbgpu_constraint_desc d{};
        d.body_1 = (uint32_t)p.body_1;
        d.body_2 = (uint32_t)p.body_2;
        d.type = p.type == ConstraintType::ROTATIONAL ? 1 : 0;
        set3(d.r_1, p.r_1);
        set3(d.r_2, p.r_2);
        d.compliance = p.compliance;
        uint32_t unit = 0;
        check(rbgpu_add_constraints(h_, 1, &d, &unit));
        const uint32_t id = add_row(unit, 0, p.body_1, p.body_2, p.type, p.r_1, p.r_2, p.compliance);
        row_unit_.push_back(unit);
        row_of_unit_.push_back(id);
        return id;
    }

    // World::create_prismatic_joint / create_revolute_joint / create_fixed_joint
    // (World.hpp:141-156, JointAPI.hpp:160-265) and the spherical extension.
    uint32_t create_prismatic_joint(PrismaticJointParams p) {
        rbgpu_joint_desc d = joint_base(RBGPU_JOINT_PRISMATIC, p.body_1, p.body_2, p.r_1, p.r_2);
        set3(d.axis, p.moving_axis);
        d.actuation = p.actuation_type; d.compliance = p.compliance; d.damping = p.damping;
        d.limited = p.limited ? 1 : 0; d.lower_limit = p.lower_limit; d.upper_limit = p.upper_limit;
        return add_joint(d, PRISMATIC, {ConstraintType::ROTATIONAL, ConstraintType::POSITIONAL, ConstraintType::POSITIONAL},
                         {0.0, 0.0, p.compliance});
    }
    uint32_t create_revolute_joint(RevoluteJointParams p) {
        rbgpu_joint_desc d = joint_base(RBGPU_JOINT_REVOLUTE, p.body_1, p.body_2, p.r_1, p.r_2);
        set3(d.axis, p.aligned_axis);
        set3(d.limit_axis, p.limit_axis);
        d.actuation = p.actuation_type; d.compliance = p.compliance; d.damping = p.damping;
        d.limited = p.limited ? 1 : 0; d.lower_limit = p.lower_limit; d.upper_limit = p.upper_limit;
        return add_joint(d, REVOLUTE,
                         {ConstraintType::ROTATIONAL, ConstraintType::POSITIONAL, ConstraintType::ROTATIONAL,
                          ConstraintType::ROTATIONAL},
                         {0.0, 0.0, 0.0, p.compliance});
    }
    uint32_t create_fixed_joint(FixedJointParams p) {
        return add_joint(joint_base(RBGPU_JOINT_FIXED, p.body_1, p.body_2, p.r_1, p.r_2), FIXED,
                         {ConstraintType::ROTATIONAL, ConstraintType::POSITIONAL}, {0.0, 0.0});
    }
    uint32_t create_spherical_joint(SphericalJointParams p) {
        rbgpu_joint_desc d = joint_base(RBGPU_JOINT_SPHERICAL, p.body_1, p.body_2, p.r_1, p.r_2);
        set3(d.axis, p.swing_axis);
        d.compliance = p.compliance;
        d.limited = p.limited ? 1 : 0;
        d.upper_limit = p.swing_limit;
        if (p.limited)
            return add_joint(d, SPHERICAL, {ConstraintType::POSITIONAL, ConstraintType::ROTATIONAL}, {p.compliance, 0.0});
        return add_joint(d, SPHERICAL, {ConstraintType::POSITIONAL}, {p.compliance});
    }

    // Pipeline options (rbgpu_set_option): RBGPU_OPT_* of rbgpu.h.
    void set_option(int option, int value) { check(rbgpu_set_option(h_, option, value)); }
    // Times the result-neutral launch options over real frames and keeps the fastest
    // (rbgpu_autotune; no counterpart in the reference): the world advances report.frames
    // frames exactly as it would untuned.  Refreshes the mirrors like step().
    rbgpu_tune_report autotune(int frames_per_trial = 3) {
        prepare_step();
        rbgpu_tune_report r{};
        check(rbgpu_autotune(h_, frames_per_trial, &r));
        if (auto_sync_) sync();
        return r;
    }

    // World::step (World.hpp:192-212): one frame of `substeps` substeps of timestep/substeps.
    void step() {
        prepare_step();
        check(rbgpu_step(h_, 1));
        if (auto_sync_) sync();
    }

    // Mirrors refreshed after every step (default), or only by sync().
    void auto_sync(bool on) { auto_sync_ = on; }
    // GPU -> every mirror (waits for the world).
    void sync() {
        sync_bodies();
        pull_joints();
        pull_contacts();
        pull_broad_phase();
    }
    // GPU -> the body mirror.
    void sync_bodies() {
        const size_t n = bodies.position.size();
        if (!n) return;
        tmp_.resize(13 * n);
        double *p = tmp_.data(), *q = p + 3 * n, *v = q + 4 * n, *w = v + 3 * n;
        check(rbgpu_get_bodies(h_, p, q, v, w));
        for (size_t i = 0; i < n; ++i) {
            bodies.position[i] = vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
            bodies.orientation[i] = quat(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]);
            bodies.linear_velocity[i] = vec3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
            bodies.angular_velocity[i] = vec3(w[3 * i], w[3 * i + 1], w[3 * i + 2]);
        }
        shadow_bodies();
    }
    // The body mirror -> GPU (step() does this by itself when the mirror changed).
    void push_bodies() {
        const size_t n = bodies.position.size();
        if (!n) return;
        tmp_.resize(13 * n);
        double *p = tmp_.data(), *q = p + 3 * n, *v = q + 4 * n, *w = v + 3 * n;
        for (size_t i = 0; i < n; ++i) {
            p[3 * i] = bodies.position[i].x; p[3 * i + 1] = bodies.position[i].y; p[3 * i + 2] = bodies.position[i].z;
            q[4 * i] = bodies.orientation[i].w; q[4 * i + 1] = bodies.orientation[i].x;
            q[4 * i + 2] = bodies.orientation[i].y; q[4 * i + 3] = bodies.orientation[i].z;
            v[3 * i] = bodies.linear_velocity[i].x; v[3 * i + 1] = bodies.linear_velocity[i].y;
            v[3 * i + 2] = bodies.linear_velocity[i].z;
            w[3 * i] = bodies.angular_velocity[i].x; w[3 * i + 1] = bodies.angular_velocity[i].y;
            w[3 * i + 2] = bodies.angular_velocity[i].z;
        }
        check(rbgpu_set_bodies(h_, p, q, v, w));
        shadow_bodies();
    }

    rbgpu_world *handle() { return h_; }

private:
    // ids: a joint's JointCollection id, a row's ConstraintCollection id; units: the C-ABI's
    // joint/row ids (one per joint or create_constraint row, creation order)
    std::vector<uint32_t> joint_unit_, unit_first_row_, unit_rows_, row_unit_, row_of_unit_;
    std::vector<const rbc::ConvexHullData *> hull_keys_;
    BodyCollection shadow_;
    std::vector<scalar> shadow_tp_, shadow_ts_;
    std::vector<vec3> shadow_dir_;
    std::vector<scalar> shadow_mag_;

    static void set3(double *o, const vec3 &v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }
    static rbgpu_joint_desc joint_base(int type, size_t b1, size_t b2, const vec3 &r1, const vec3 &r2) {
        rbgpu_joint_desc d{};
        d.type = type;
        d.actuation = RBGPU_ACT_FREE;
        d.body_1 = (uint32_t)b1;
        d.body_2 = (uint32_t)b2;
        set3(d.r_1, r1);
        set3(d.r_2, r2);
        return d;
    }
    uint32_t add_row(uint32_t unit, uint32_t k, size_t b1, size_t b2, ConstraintType t, const vec3 &r1, const vec3 &r2,
                     scalar comp) {
        const uint32_t id = constraints.count();
        if (k == 0) {
            if (unit_first_row_.size() <= unit) { unit_first_row_.resize(unit + 1, 0); unit_rows_.resize(unit + 1, 0); }
            unit_first_row_[unit] = id;
        }
        ++unit_rows_[unit];
        constraints.body_1.push_back(b1);
        constraints.body_2.push_back(b2);
        constraints.type.push_back(t);
        constraints.r_1.push_back(r1);
        constraints.r_2.push_back(r2);
        constraints.direction.push_back(vec3(1.0, 0.0, 0.0));   // ConstraintAPI.hpp:52-54
        constraints.magnitude.push_back(0.0);
        constraints.lambda.push_back(0.0);
        constraints.compliance.push_back(comp);
        shadow_dir_.push_back(vec3(1.0, 0.0, 0.0));
        shadow_mag_.push_back(0.0);
        return id;
    }
    uint32_t add_joint(rbgpu_joint_desc d, JointType t, std::initializer_list<ConstraintType> rows,
                       std::initializer_list<scalar> comp) {
        push_changes();
        uint32_t unit = 0;
        check(rbgpu_add_joints(h_, 1, &d, &unit));
        const uint32_t id = joints.count();
        joints.type.push_back(t);
        joints.body_1.push_back(d.body_1);
        joints.body_2.push_back(d.body_2);
        joints.target_position.push_back(d.target_position);
        joints.target_speed.push_back(d.target_speed);
        joints.current_position.push_back(0.0);
        shadow_tp_.push_back(d.target_position);
        shadow_ts_.push_back(d.target_speed);
        joint_unit_.push_back(unit);
        auto c = comp.begin();
        uint32_t k = 0;
        for (ConstraintType r : rows)
            add_row(unit, k++, d.body_1, d.body_2, r, vec3(d.r_1[0], d.r_1[1], d.r_1[2]), vec3(d.r_2[0], d.r_2[1], d.r_2[2]),
                    *c++);
        return id;
    }
    int32_t hull_id(const rbc::ConvexHullData *h) {
        if (!h || !h->vertices || !h->vert_count) throw Error(RBGPU_ERR_INVALID, "convex hull collider without hull data");
        for (size_t i = 0; i < hull_keys_.size(); ++i)
            if (hull_keys_[i] == h) return (int32_t)i;
        uint32_t id = 0;
        check(rbgpu_add_hull(h_, &h->vertices[0].x, h->vert_count, h->face_indices, h->face_count, &id));
        hull_keys_.push_back(h);
        return (int32_t)id;
    }
    void shadow_bodies() {
        shadow_.position = bodies.position;
        shadow_.orientation = bodies.orientation;
        shadow_.linear_velocity = bodies.linear_velocity;
        shadow_.angular_velocity = bodies.angular_velocity;
    }
    bool bodies_changed() const {
        return bodies.position != shadow_.position || bodies.orientation != shadow_.orientation ||
               bodies.linear_velocity != shadow_.linear_velocity || bodies.angular_velocity != shadow_.angular_velocity;
    }
    // The caller's writes into the mirrors -> GPU.
    void push_changes() {
        if (bodies_changed()) push_bodies();
        if (bodies.force != shadow_.force || bodies.torque != shadow_.torque) {
            const size_t n = bodies.force.size();
            std::vector<double> ft(6 * n);
            for (size_t i = 0; i < n; ++i) {
                ft[3 * i] = bodies.force[i].x; ft[3 * i + 1] = bodies.force[i].y; ft[3 * i + 2] = bodies.force[i].z;
                ft[3 * n + 3 * i] = bodies.torque[i].x; ft[3 * n + 3 * i + 1] = bodies.torque[i].y;
                ft[3 * n + 3 * i + 2] = bodies.torque[i].z;
            }
            check(rbgpu_set_body_forces(h_, ft.data(), ft.data() + 3 * n));
            shadow_.force = bodies.force;
            shadow_.torque = bodies.torque;
        }
        for (uint32_t j = 0; j < joints.count(); ++j)
            if (joints.target_position[j] != shadow_tp_[j] || joints.target_speed[j] != shadow_ts_[j]) {
                check(rbgpu_set_joint_targets(h_, 1, &joint_unit_[j], &joints.target_position[j], &joints.target_speed[j]));
                shadow_tp_[j] = joints.target_position[j];
                shadow_ts_[j] = joints.target_speed[j];
            }
        for (size_t r = 0; r < row_unit_.size(); ++r) {
            const uint32_t c = row_of_unit_[r];
            if (constraints.direction[c] != shadow_dir_[c] || constraints.magnitude[c] != shadow_mag_[c]) {
                const double d3[3] = {constraints.direction[c].x, constraints.direction[c].y, constraints.direction[c].z};
                check(rbgpu_set_constraints(h_, 1, &row_unit_[r], d3, &constraints.magnitude[c]));
                shadow_dir_[c] = constraints.direction[c];
                shadow_mag_[c] = constraints.magnitude[c];
            }
        }
    }
    void prepare_step() {
        push_changes();
        if (timestep != ts_ || substeps != sub_) {
            check(rbgpu_set_step(h_, timestep, substeps));
            ts_ = timestep;
            sub_ = substeps;
        }
    }
    // GPU -> joints (current / target position) and every constraint row (direction,
    // magnitude, lambda).
    void pull_joints() {
        const size_t nu = unit_rows_.size();
        if (!nu) return;
        std::vector<double> cur(nu), tgt(nu), lam(4 * nu), dir(12 * nu), mag(4 * nu);
        check(rbgpu_get_joints(h_, cur.data(), tgt.data(), lam.data(), nullptr));
        check(rbgpu_get_constraint_rows(h_, dir.data(), mag.data()));
        for (uint32_t j = 0; j < joints.count(); ++j) {
            const uint32_t u = joint_unit_[j];
            joints.current_position[j] = cur[u];
            joints.target_position[j] = shadow_tp_[j] = tgt[u];   // a speed drive moves its target
        }
        for (uint32_t u = 0; u < nu; ++u)
            for (uint32_t k = 0; k < unit_rows_[u]; ++k) {
                const uint32_t c = unit_first_row_[u] + k;
                constraints.lambda[c] = lam[4 * u + k];
                constraints.direction[c] = shadow_dir_[c] = vec3(dir[12 * u + 3 * k], dir[12 * u + 3 * k + 1], dir[12 * u + 3 * k + 2]);
                constraints.magnitude[c] = shadow_mag_[c] = mag[4 * u + k];
            }
    }
    // GPU -> contacts (the last substep's).
    void pull_contacts() {
        uint32_t n = 0;
        check(rbgpu_get_contacts(h_, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, &n));
        ContactList &c = contacts;
        c = ContactList{};
        c.n_contacts = n;
        if (!n) return;
        std::vector<uint32_t> ids(4 * n);
        std::vector<double> nrm(3 * n), ra(3 * n), rb(3 * n), dep(n), lam(2 * n), pa(3 * n), pb(3 * n), rv(n);
        std::vector<int32_t> col(n);
        check(rbgpu_get_contacts(h_, ids.data(), nrm.data(), ra.data(), rb.data(), dep.data(), lam.data(), n, &n));
        uint32_t n2 = 0;
        check(rbgpu_get_contact_points(h_, pa.data(), pb.data(), rv.data(), col.data(), n, &n2));
        auto v = [](const std::vector<double> &a, uint32_t i) { return vec3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); };
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t ca = ids[4 * i + 2], cb = ids[4 * i + 3];
            c.body_a.push_back(ids[4 * i]);
            c.body_b.push_back(ids[4 * i + 1]);
            c.collider_a.push_back(ca);
            c.collider_b.push_back(cb);
            c.normal.push_back(v(nrm, i));
            c.r_a_local.push_back(v(ra, i));
            c.r_b_local.push_back(v(rb, i));
            c.point_on_a.push_back(v(pa, i));
            c.point_on_b.push_back(v(pb, i));
            c.penetration_depth.push_back(dep[i]);
            c.normal_lambda.push_back(lam[2 * i]);
            c.tangent_lambda.push_back(lam[2 * i + 1]);
            c.relative_velocity.push_back(rv[i]);
            c.collision.push_back(col[i] != 0);
            // material mixing of emit_contact (CollisionPipeline.cpp:75-77): minimum, means
            c.restitution.push_back(std::min(colliders.restitution[ca], colliders.restitution[cb]));
            c.static_friction.push_back((colliders.static_friction[ca] + colliders.static_friction[cb]) * 0.5);
            c.dynamic_friction.push_back((colliders.dynamic_friction[ca] + colliders.dynamic_friction[cb]) * 0.5);
        }
    }
    // GPU -> broad_phase_state (objects: fat boxes and user ids; pairs: the frame's filtered pairs).
    void pull_broad_phase() {
        uint32_t n = 0;
        check(rbgpu_get_broadphase_objects(h_, nullptr, nullptr, nullptr, 0, &n));
        std::vector<double> mn(3 * n), mx(3 * n);
        std::vector<uint32_t> uid(n);
        if (n) check(rbgpu_get_broadphase_objects(h_, mn.data(), mx.data(), uid.data(), n, &n));
        auto &o = broad_phase_state.objects;
        o.assign(n, rbc::BPObject{});
        for (uint32_t i = 0; i < n; ++i) {
            o[i].fat_aabb.min = vec3(mn[3 * i], mn[3 * i + 1], mn[3 * i + 2]);
            o[i].fat_aabb.max = vec3(mx[3 * i], mx[3 * i + 1], mx[3 * i + 2]);
            o[i].user_id = uid[i];
        }
        uint32_t np = 0;
        check(rbgpu_get_sap_pairs(h_, nullptr, 0, &np));
        std::vector<uint32_t> pr(2 * np);
        if (np) check(rbgpu_get_sap_pairs(h_, pr.data(), np, &np));
        broad_phase_state.pairs.resize(np);
        for (uint32_t i = 0; i < np; ++i) broad_phase_state.pairs[i] = rbc::BroadPhasePair{pr[2 * i], pr[2 * i + 1]};
    }
    void create() {
        check(rbgpu_world_create(timestep, substeps, &h_));
        ts_ = timestep;
        sub_ = substeps;
    }
    rbgpu_world *h_ = nullptr;
    scalar ts_ = 0;
    int sub_ = 0;
    bool auto_sync_ = true;
    std::vector<double> tmp_;
};

}  // namespace rbps

// The names this header used before it took the reference's namespaces.
namespace rbps_amd {
using namespace ::m3d;
using namespace ::rbc;
using namespace ::rbps;
}  // namespace rbps_amd
