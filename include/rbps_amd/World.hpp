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
#include <atomic>
#include <cmath>
#include <cstdint>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
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

class WorldPool;

namespace detail {

// The GPU world behind one World, or behind every World of a WorldPool (each pool world a
// scene of the batch).  Holds the C-ABI handle, which world owns which body / collider, and
// the frame's read-backs, fetched once per step for all the worlds that share it.
struct Batch {
    rbgpu_world *h = nullptr;
    std::mutex mu;                       // C-ABI calls of worlds that share the batch
    uint64_t epoch = 1;                  // bumped by every step and every write
    std::vector<uint32_t> body_world, body_local, col_world, col_local;
    uint32_t units = 0;                  // joints + constraint rows (the C-ABI's joint ids)
    std::vector<const rbc::ConvexHullData *> hull_keys;   // hull data -> rbgpu_add_hull id
    std::vector<uint32_t> hull_ids;
    // read-backs of the current epoch
    uint64_t e_bodies = 0, e_joints = 0, e_contacts = 0, e_bp = 0;
    std::vector<double> pos, ori, lv, av;
    std::vector<double> jcur, jtgt, jlam, rdir, rmag;
    uint32_t n_contacts = 0;
    std::vector<uint32_t> cids;
    std::vector<double> cnrm, cra, crb, cdep, clam, cpa, cpb, crv;
    std::vector<int32_t> ccol;
    std::vector<double> fmn, fmx;
    std::vector<uint32_t> fuid, sap;
    double ts = 0;
    int sub = 0;

    Batch(scalar timestep, int substeps) {
        check(rbgpu_world_create(timestep, substeps, &h));
        ts = timestep;
        sub = substeps;
    }
    ~Batch() {
        if (h) rbgpu_world_destroy(h);
    }
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    void touch() { ++epoch; }
    void set_step(scalar timestep, int substeps) {
        if (timestep != ts || substeps != sub) {
            check(rbgpu_set_step(h, timestep, substeps));
            ts = timestep;
            sub = substeps;
        }
    }
    void fetch_bodies() {
        if (e_bodies == epoch) return;
        const size_t n = body_world.size();
        pos.resize(3 * n); ori.resize(4 * n); lv.resize(3 * n); av.resize(3 * n);
        if (n) check(rbgpu_get_bodies(h, pos.data(), ori.data(), lv.data(), av.data()));
        e_bodies = epoch;
    }
    void fetch_joints() {
        if (e_joints == epoch && jcur.size() == units) return;
        jcur.resize(units); jtgt.resize(units); jlam.resize(4 * units); rdir.resize(12 * units); rmag.resize(4 * units);
        if (units) {
            check(rbgpu_get_joints(h, jcur.data(), jtgt.data(), jlam.data(), nullptr));
            check(rbgpu_get_constraint_rows(h, rdir.data(), rmag.data()));
        }
        e_joints = epoch;
    }
    void fetch_contacts() {
        if (e_contacts == epoch) return;
        uint32_t n = 0;
        check(rbgpu_get_contacts(h, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, &n));
        cids.resize(4 * n); cnrm.resize(3 * n); cra.resize(3 * n); crb.resize(3 * n); cdep.resize(n); clam.resize(2 * n);
        cpa.resize(3 * n); cpb.resize(3 * n); crv.resize(n); ccol.resize(n);
        if (n) {
            check(rbgpu_get_contacts(h, cids.data(), cnrm.data(), cra.data(), crb.data(), cdep.data(), clam.data(), n, &n));
            uint32_t n2 = 0;
            check(rbgpu_get_contact_points(h, cpa.data(), cpb.data(), crv.data(), ccol.data(), n, &n2));
        }
        n_contacts = n;
        e_contacts = epoch;
    }
    void fetch_broad_phase() {
        if (e_bp == epoch) return;
        uint32_t n = 0;
        check(rbgpu_get_broadphase_objects(h, nullptr, nullptr, nullptr, 0, &n));
        fmn.resize(3 * n); fmx.resize(3 * n); fuid.resize(n);
        if (n) check(rbgpu_get_broadphase_objects(h, fmn.data(), fmx.data(), fuid.data(), n, &n));
        uint32_t np = 0;
        check(rbgpu_get_sap_pairs(h, nullptr, 0, &np));
        sap.resize(2 * np);
        if (np) check(rbgpu_get_sap_pairs(h, sap.data(), np, &np));
        e_bp = epoch;
    }
};

}  // namespace detail

struct World {
    scalar timestep = 1.0 / 60.0;
    int substeps = 20;
    BodyCollection bodies;
    ConstraintCollection constraints;
    JointCollection joints;
    ContactList contacts;
    ColliderCollection colliders;
    rbc::BroadPhaseState broad_phase_state;

    World() : b_(std::make_shared<detail::Batch>(timestep, substeps)) {}
    World(scalar timestep_, int substeps_)
        : timestep(timestep_), substeps(substeps_), b_(std::make_shared<detail::Batch>(timestep_, substeps_)) {}
    World(World &&) = default;
    World &operator=(World &&) = default;
    World(const World &) = delete;
    World &operator=(const World &) = delete;

    // World::create_body (World.hpp:113-116)
    uint32_t create_body(BodyParams p) {
        std::lock_guard<std::mutex> lk(b_->mu);
        push_changes();
        rbgpu_body_desc d{};
        d.type = p.type == STATIC ? RBGPU_STATIC : RBGPU_DYNAMIC;
        d.scene = scene_;
        d.mass = p.mass;
        const scalar I[6] = {p.inertia_tensor.xx, p.inertia_tensor.yy, p.inertia_tensor.zz,
                             p.inertia_tensor.xy, p.inertia_tensor.xz, p.inertia_tensor.yz};
        for (int k = 0; k < 6; ++k) d.inertia[k] = I[k];
        set3(d.position, p.position);
        d.orientation[0] = p.orientation.w; d.orientation[1] = p.orientation.x;
        d.orientation[2] = p.orientation.y; d.orientation[3] = p.orientation.z;
        set3(d.linear_velocity, p.linear_velocity);
        set3(d.angular_velocity, p.angular_velocity);
        uint32_t g = 0;
        check(rbgpu_add_bodies(b_->h, 1, &d, &g));
        b_->touch();
        const uint32_t id = bodies.count();
        gbody_.push_back(g);
        b_->body_world.push_back(scene_);
        b_->body_local.push_back(id);
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
        std::lock_guard<std::mutex> lk(b_->mu);
        push_changes();
        rbgpu_collider_desc d{};
        d.shape = p.shape.kind;
        d.body_id = global_body(p.body_id);
        d.is_static = p.is_static ? 1 : 0;
        d.hull_index = p.shape.kind == RBGPU_HULL ? hull_id(p.shape.hull) : -1;
        for (int k = 0; k < 4; ++k) d.dims[k] = p.shape.dims[k];
        set3(d.local_pos, p.local_pos);
        d.local_rot[0] = p.local_rot.w; d.local_rot[1] = p.local_rot.x; d.local_rot[2] = p.local_rot.y;
        d.local_rot[3] = p.local_rot.z;
        d.restitution = p.restitution;
        d.static_friction = p.static_friction;
        d.dynamic_friction = p.dynamic_friction;
        uint32_t g = 0;
        check(rbgpu_add_colliders(b_->h, 1, &d, &g));
        b_->touch();
        const uint32_t id = colliders.count();
        b_->col_world.push_back(scene_);
        b_->col_local.push_back(id);
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
        std::lock_guard<std::mutex> lk(b_->mu);
        push_changes();
        rbgpu_constraint_desc d{};
        d.body_1 = global_body(p.body_1);
        d.body_2 = global_body(p.body_2);
        d.type = p.type == ConstraintType::ROTATIONAL ? 1 : 0;
        set3(d.r_1, p.r_1);
        set3(d.r_2, p.r_2);
        d.compliance = p.compliance;
        uint32_t unit = 0;
        check(rbgpu_add_constraints(b_->h, 1, &d, &unit));
        b_->units = std::max(b_->units, unit + 1);
        b_->touch();
        const uint32_t id = add_row(unit, 0, p.body_1, p.body_2, p.type, p.r_1, p.r_2, p.compliance);
        row_unit_.push_back(unit);
        row_of_unit_.push_back(id);
        return id;
    }

    // World::create_prismatic_joint / create_revolute_joint / create_fixed_joint
    // (World.hpp:141-156, JointAPI.hpp:160-265) and the spherical extension.
    uint32_t create_prismatic_joint(PrismaticJointParams p) {
        rbgpu_joint_desc d = joint_base(RBGPU_JOINT_PRISMATIC, p.r_1, p.r_2);
        set3(d.axis, p.moving_axis);
        d.actuation = p.actuation_type; d.compliance = p.compliance; d.damping = p.damping;
        d.limited = p.limited ? 1 : 0; d.lower_limit = p.lower_limit; d.upper_limit = p.upper_limit;
        return add_joint(d, p.body_1, p.body_2, PRISMATIC,
                         {ConstraintType::ROTATIONAL, ConstraintType::POSITIONAL, ConstraintType::POSITIONAL},
                         {0.0, 0.0, p.compliance});
    }
    uint32_t create_revolute_joint(RevoluteJointParams p) {
        rbgpu_joint_desc d = joint_base(RBGPU_JOINT_REVOLUTE, p.r_1, p.r_2);
        set3(d.axis, p.aligned_axis);
        set3(d.limit_axis, p.limit_axis);
        d.actuation = p.actuation_type; d.compliance = p.compliance; d.damping = p.damping;
        d.limited = p.limited ? 1 : 0; d.lower_limit = p.lower_limit; d.upper_limit = p.upper_limit;
        return add_joint(d, p.body_1, p.body_2, REVOLUTE,
                         {ConstraintType::ROTATIONAL, ConstraintType::POSITIONAL, ConstraintType::ROTATIONAL,
                          ConstraintType::ROTATIONAL},
                         {0.0, 0.0, 0.0, p.compliance});
    }
    uint32_t create_fixed_joint(FixedJointParams p) {
        return add_joint(joint_base(RBGPU_JOINT_FIXED, p.r_1, p.r_2), p.body_1, p.body_2, FIXED,
                         {ConstraintType::ROTATIONAL, ConstraintType::POSITIONAL}, {0.0, 0.0});
    }
    uint32_t create_spherical_joint(SphericalJointParams p) {
        rbgpu_joint_desc d = joint_base(RBGPU_JOINT_SPHERICAL, p.r_1, p.r_2);
        set3(d.axis, p.swing_axis);
        d.compliance = p.compliance;
        d.limited = p.limited ? 1 : 0;
        d.upper_limit = p.swing_limit;
        if (p.limited)
            return add_joint(d, p.body_1, p.body_2, SPHERICAL, {ConstraintType::POSITIONAL, ConstraintType::ROTATIONAL},
                             {p.compliance, 0.0});
        return add_joint(d, p.body_1, p.body_2, SPHERICAL, {ConstraintType::POSITIONAL}, {p.compliance});
    }

    // Pipeline options (rbgpu_set_option, RBGPU_OPT_* of rbgpu.h); a WorldPool's worlds
    // share them.
    void set_option(int option, int value) {
        std::lock_guard<std::mutex> lk(b_->mu);
        check(rbgpu_set_option(b_->h, option, value));
    }
    // Times the result-neutral launch options over real frames and keeps the fastest
    // (rbgpu_autotune; no counterpart in the reference): the world advances report.frames
    // frames exactly as it would untuned.  Refreshes the mirrors like step().
    rbgpu_tune_report autotune(int frames_per_trial = 3) {
        solo("autotune");
        rbgpu_tune_report r{};
        {
            std::lock_guard<std::mutex> lk(b_->mu);
            prepare_step();
            check(rbgpu_autotune(b_->h, frames_per_trial, &r));
            b_->touch();
        }
        if (auto_sync_) sync();
        return r;
    }

    // World::step (World.hpp:192-212): one frame of `substeps` substeps of timestep/substeps.
    // A WorldPool's worlds step together (WorldPool::parallel_step).
    void step() {
        solo("step");
        {
            std::lock_guard<std::mutex> lk(b_->mu);
            prepare_step();
            check(rbgpu_step(b_->h, 1));
            b_->touch();
        }
        if (auto_sync_) sync();
    }

    // Mirrors refreshed after every step (default), or only by sync().
    void auto_sync(bool on) { auto_sync_ = on; }
    // GPU -> every mirror (waits for the world).
    void sync() {
        std::lock_guard<std::mutex> lk(b_->mu);
        pull_bodies();
        pull_joints();
        pull_contacts();
        pull_broad_phase();
    }
    // GPU -> the body mirror.
    void sync_bodies() {
        std::lock_guard<std::mutex> lk(b_->mu);
        pull_bodies();
    }
    // The body mirror -> GPU (step() does this by itself when the mirror changed).
    void push_bodies() {
        std::lock_guard<std::mutex> lk(b_->mu);
        write_bodies();
    }

    rbgpu_world *handle() { return b_->h; }

private:
    friend class WorldPool;
    std::shared_ptr<detail::Batch> b_;
    uint32_t scene_ = 0;
    bool pooled_ = false;
    std::vector<uint32_t> gbody_;   // this world's body ids -> the batch's
    // ids: a joint's JointCollection id, a row's ConstraintCollection id; units: the C-ABI's
    // joint/row ids (one per joint or create_constraint row, in the batch)
    std::vector<uint32_t> joint_unit_, row_unit_, row_of_unit_;
    std::vector<uint32_t> unit_list_, unit_first_row_, unit_rows_;   // this world's units, in creation order
    BodyCollection shadow_;
    std::vector<scalar> shadow_tp_, shadow_ts_;
    std::vector<vec3> shadow_dir_;
    std::vector<scalar> shadow_mag_;
    bool auto_sync_ = true;

    World(std::shared_ptr<detail::Batch> b, uint32_t scene, scalar timestep_, int substeps_)
        : timestep(timestep_), substeps(substeps_), b_(std::move(b)), scene_(scene), pooled_(true) {}

    void solo(const char *what) const {
        if (pooled_)
            throw Error(RBGPU_ERR_INVALID, std::string("World::") + what + " on a WorldPool world: the pool's worlds "
                                           "step together (WorldPool::parallel_step)");
    }
    uint32_t global_body(size_t local) const {
        if (local >= gbody_.size()) throw Error(RBGPU_ERR_INVALID, "body id " + std::to_string(local) + " out of range");
        return gbody_[local];
    }
    static void set3(double *o, const vec3 &v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }
    static rbgpu_joint_desc joint_base(int type, const vec3 &r1, const vec3 &r2) {
        rbgpu_joint_desc d{};
        d.type = type;
        d.actuation = RBGPU_ACT_FREE;
        set3(d.r_1, r1);
        set3(d.r_2, r2);
        return d;
    }
    uint32_t add_row(uint32_t unit, uint32_t k, size_t b1, size_t b2, ConstraintType t, const vec3 &r1, const vec3 &r2,
                     scalar comp) {
        const uint32_t id = constraints.count();
        if (k == 0) {
            unit_list_.push_back(unit);
            unit_first_row_.push_back(id);
            unit_rows_.push_back(0);
        }
        ++unit_rows_.back();
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
    uint32_t add_joint(rbgpu_joint_desc d, size_t b1, size_t b2, JointType t, std::initializer_list<ConstraintType> rows,
                       std::initializer_list<scalar> comp) {
        std::lock_guard<std::mutex> lk(b_->mu);
        push_changes();
        d.body_1 = global_body(b1);
        d.body_2 = global_body(b2);
        uint32_t unit = 0;
        check(rbgpu_add_joints(b_->h, 1, &d, &unit));
        b_->units = std::max(b_->units, unit + 1);
        b_->touch();
        const uint32_t id = joints.count();
        joints.type.push_back(t);
        joints.body_1.push_back(b1);
        joints.body_2.push_back(b2);
        joints.target_position.push_back(d.target_position);
        joints.target_speed.push_back(d.target_speed);
        joints.current_position.push_back(0.0);
        shadow_tp_.push_back(d.target_position);
        shadow_ts_.push_back(d.target_speed);
        joint_unit_.push_back(unit);
        auto c = comp.begin();
        uint32_t k = 0;
        for (ConstraintType r : rows)
            add_row(unit, k++, b1, b2, r, vec3(d.r_1[0], d.r_1[1], d.r_1[2]), vec3(d.r_2[0], d.r_2[1], d.r_2[2]), *c++);
        return id;
    }
    // the batch's id of the hull data (added once, shared by every world of a pool)
    int32_t hull_id(const rbc::ConvexHullData *h) {
        if (!h || !h->vertices || !h->vert_count) throw Error(RBGPU_ERR_INVALID, "convex hull collider without hull data");
        detail::Batch &B = *b_;
        for (size_t i = 0; i < B.hull_keys.size(); ++i)
            if (B.hull_keys[i] == h) return (int32_t)B.hull_ids[i];
        uint32_t id = 0;
        check(rbgpu_add_hull(B.h, &h->vertices[0].x, h->vert_count, h->face_indices, h->face_count, &id));
        B.hull_keys.push_back(h);
        B.hull_ids.push_back(id);
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
    // The body mirror -> GPU: the batch's whole state with this world's bodies replaced.
    void write_bodies() {
        const size_t n = bodies.count();
        if (!n) return;
        detail::Batch &B = *b_;
        B.fetch_bodies();
        std::vector<double> p(B.pos), q(B.ori), v(B.lv), w(B.av);
        for (size_t i = 0; i < n; ++i) {
            const size_t g = gbody_[i];
            const vec3 &x = bodies.position[i], &lv = bodies.linear_velocity[i], &av = bodies.angular_velocity[i];
            const quat &o = bodies.orientation[i];
            p[3 * g] = x.x; p[3 * g + 1] = x.y; p[3 * g + 2] = x.z;
            q[4 * g] = o.w; q[4 * g + 1] = o.x; q[4 * g + 2] = o.y; q[4 * g + 3] = o.z;
            v[3 * g] = lv.x; v[3 * g + 1] = lv.y; v[3 * g + 2] = lv.z;
            w[3 * g] = av.x; w[3 * g + 1] = av.y; w[3 * g + 2] = av.z;
        }
        check(rbgpu_set_bodies(B.h, p.data(), q.data(), v.data(), w.data()));
        B.touch();
        B.pos.swap(p); B.ori.swap(q); B.lv.swap(v); B.av.swap(w);
        B.e_bodies = B.epoch;   // the batch holds exactly what was written
        shadow_bodies();
    }
    // The caller's writes into the mirrors -> GPU (batch lock held).
    void push_changes() {
        if (bodies_changed()) write_bodies();
        if (bodies.force != shadow_.force || bodies.torque != shadow_.torque) {
            const size_t nb = b_->body_world.size();
            std::vector<double> f(3 * nb), t(3 * nb);
            check(rbgpu_get_body_forces(b_->h, f.data(), t.data()));
            for (size_t i = 0; i < bodies.force.size(); ++i) {
                const size_t g = gbody_[i];
                f[3 * g] = bodies.force[i].x; f[3 * g + 1] = bodies.force[i].y; f[3 * g + 2] = bodies.force[i].z;
                t[3 * g] = bodies.torque[i].x; t[3 * g + 1] = bodies.torque[i].y; t[3 * g + 2] = bodies.torque[i].z;
            }
            check(rbgpu_set_body_forces(b_->h, f.data(), t.data()));
            b_->touch();
            shadow_.force = bodies.force;
            shadow_.torque = bodies.torque;
        }
        for (uint32_t j = 0; j < joints.count(); ++j)
            if (joints.target_position[j] != shadow_tp_[j] || joints.target_speed[j] != shadow_ts_[j]) {
                check(rbgpu_set_joint_targets(b_->h, 1, &joint_unit_[j], &joints.target_position[j],
                                              &joints.target_speed[j]));
                b_->touch();
                shadow_tp_[j] = joints.target_position[j];
                shadow_ts_[j] = joints.target_speed[j];
            }
        for (size_t r = 0; r < row_unit_.size(); ++r) {
            const uint32_t c = row_of_unit_[r];
            if (constraints.direction[c] != shadow_dir_[c] || constraints.magnitude[c] != shadow_mag_[c]) {
                const double d3[3] = {constraints.direction[c].x, constraints.direction[c].y, constraints.direction[c].z};
                check(rbgpu_set_constraints(b_->h, 1, &row_unit_[r], d3, &constraints.magnitude[c]));
                b_->touch();
                shadow_dir_[c] = constraints.direction[c];
                shadow_mag_[c] = constraints.magnitude[c];
            }
        }
    }
    void prepare_step() {
        push_changes();
        b_->set_step(timestep, substeps);
    }
    // GPU -> the body mirror (batch lock held).
    void pull_bodies() {
        const size_t n = bodies.count();
        if (!n) return;
        detail::Batch &B = *b_;
        B.fetch_bodies();
        for (size_t i = 0; i < n; ++i) {
            const size_t g = gbody_[i];
            bodies.position[i] = vec3(B.pos[3 * g], B.pos[3 * g + 1], B.pos[3 * g + 2]);
            bodies.orientation[i] = quat(B.ori[4 * g], B.ori[4 * g + 1], B.ori[4 * g + 2], B.ori[4 * g + 3]);
            bodies.linear_velocity[i] = vec3(B.lv[3 * g], B.lv[3 * g + 1], B.lv[3 * g + 2]);
            bodies.angular_velocity[i] = vec3(B.av[3 * g], B.av[3 * g + 1], B.av[3 * g + 2]);
        }
        shadow_bodies();
    }
    // GPU -> joints (current / target position) and every constraint row (direction,
    // magnitude, lambda).
    void pull_joints() {
        if (unit_list_.empty()) return;
        detail::Batch &B = *b_;
        B.fetch_joints();
        for (uint32_t j = 0; j < joints.count(); ++j) {
            const uint32_t u = joint_unit_[j];
            joints.current_position[j] = B.jcur[u];
            joints.target_position[j] = shadow_tp_[j] = B.jtgt[u];   // a speed drive moves its target
        }
        for (size_t k = 0; k < unit_list_.size(); ++k) {
            const uint32_t u = unit_list_[k];
            for (uint32_t r = 0; r < unit_rows_[k]; ++r) {
                const uint32_t c = unit_first_row_[k] + r;
                constraints.lambda[c] = B.jlam[4 * u + r];
                constraints.direction[c] = shadow_dir_[c] =
                    vec3(B.rdir[12 * u + 3 * r], B.rdir[12 * u + 3 * r + 1], B.rdir[12 * u + 3 * r + 2]);
                constraints.magnitude[c] = shadow_mag_[c] = B.rmag[4 * u + r];
            }
        }
    }
    // GPU -> contacts (the last substep's, this world's).
    void pull_contacts() {
        detail::Batch &B = *b_;
        B.fetch_contacts();
        ContactList &c = contacts;
        c = ContactList{};
        auto v = [](const std::vector<double> &a, uint32_t i) { return vec3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); };
        for (uint32_t i = 0; i < B.n_contacts; ++i) {
            const uint32_t gca = B.cids[4 * i + 2], gcb = B.cids[4 * i + 3];
            if (B.col_world[gca] != scene_) continue;
            const uint32_t ca = B.col_local[gca], cb = B.col_local[gcb];
            c.body_a.push_back(B.body_local[B.cids[4 * i]]);
            c.body_b.push_back(B.body_local[B.cids[4 * i + 1]]);
            c.collider_a.push_back(ca);
            c.collider_b.push_back(cb);
            c.normal.push_back(v(B.cnrm, i));
            c.r_a_local.push_back(v(B.cra, i));
            c.r_b_local.push_back(v(B.crb, i));
            c.point_on_a.push_back(v(B.cpa, i));
            c.point_on_b.push_back(v(B.cpb, i));
            c.penetration_depth.push_back(B.cdep[i]);
            c.normal_lambda.push_back(B.clam[2 * i]);
            c.tangent_lambda.push_back(B.clam[2 * i + 1]);
            c.relative_velocity.push_back(B.crv[i]);
            c.collision.push_back(B.ccol[i] != 0);
            // material mixing of emit_contact (CollisionPipeline.cpp:75-77): minimum, means
            c.restitution.push_back(std::min(colliders.restitution[ca], colliders.restitution[cb]));
            c.static_friction.push_back((colliders.static_friction[ca] + colliders.static_friction[cb]) * 0.5);
            c.dynamic_friction.push_back((colliders.dynamic_friction[ca] + colliders.dynamic_friction[cb]) * 0.5);
        }
        c.n_contacts = (uint32_t)c.body_a.size();
    }
    // GPU -> broad_phase_state (objects: fat boxes and collider ids; pairs: the frame's
    // filtered pairs), this world's.
    void pull_broad_phase() {
        detail::Batch &B = *b_;
        B.fetch_broad_phase();
        auto &o = broad_phase_state.objects;
        o.clear();
        for (size_t i = 0; i < B.fuid.size(); ++i) {
            const uint32_t g = B.fuid[i];
            if (B.col_world[g] != scene_) continue;
            rbc::BPObject ob;
            ob.fat_aabb.min = vec3(B.fmn[3 * i], B.fmn[3 * i + 1], B.fmn[3 * i + 2]);
            ob.fat_aabb.max = vec3(B.fmx[3 * i], B.fmx[3 * i + 1], B.fmx[3 * i + 2]);
            ob.user_id = B.col_local[g];
            o.push_back(ob);
        }
        auto &pr = broad_phase_state.pairs;
        pr.clear();
        for (size_t i = 0; 2 * i < B.sap.size(); ++i) {
            const uint32_t a = B.sap[2 * i], b = B.sap[2 * i + 1];
            if (B.col_world[a] != scene_) continue;
            pr.push_back(rbc::BroadPhasePair{B.col_local[a], B.col_local[b]});
        }
    }
};

// WorldPool (the reference's multi-world design, docs/design/multi-world/multi-world.md:40-50):
// n independent worlds, each a scene of one batched GPU world.  Build each world with its own
// create_* calls (its ids start at 0, as in a World of its own); parallel_step advances every
// world one frame in one batched step (every kernel runs once over all worlds; a world evolves
// exactly as it would alone); parallel_each runs a function per world on host threads.  The
// worlds share the timestep / substeps given to parallel_step and the pipeline options; a
// pool world's own step() / autotune() throw (the worlds step together).
class WorldPool {
public:
    std::vector<World> worlds;

    explicit WorldPool(size_t n, scalar timestep = 1.0 / 60.0, int substeps = 20)
        : b_(std::make_shared<detail::Batch>(timestep, substeps)) {
        worlds.reserve(n);
        for (size_t i = 0; i < n; ++i) worlds.push_back(World(b_, (uint32_t)i, timestep, substeps));
    }
    WorldPool(const WorldPool &) = delete;
    WorldPool &operator=(const WorldPool &) = delete;

    size_t size() const { return worlds.size(); }

    // One frame of dt for every world (each world's `substeps`, which must agree).
    void parallel_step(scalar dt) {
        {
            std::lock_guard<std::mutex> lk(b_->mu);
            const int sub = worlds.empty() ? b_->sub : worlds[0].substeps;
            for (World &w : worlds) {
                if (w.substeps != sub) throw Error(RBGPU_ERR_INVALID, "the worlds of a WorldPool step with one substep count");
                w.push_changes();
                w.timestep = dt;
            }
            b_->set_step(dt, sub);
            check(rbgpu_step(b_->h, 1));
            b_->touch();
        }
        if (auto_sync_) sync();
    }

    // fn(World &) or fn(World &, size_t index) for every world, on up to `threads` host
    // threads (0: the hardware's).  The first exception thrown is rethrown after all finish.
    template <typename Fn> void parallel_each(Fn &&fn, unsigned threads = 0) {
        const size_t n = worlds.size();
        if (!threads) threads = std::max(1u, std::thread::hardware_concurrency());
        threads = (unsigned)std::min<size_t>(threads, n);
        std::atomic<size_t> next{0};
        std::exception_ptr err;
        std::mutex err_mu;
        auto work = [&]() {
            for (size_t i = next++; i < n; i = next++) {
                try {
                    call(fn, worlds[i], i);
                } catch (...) {
                    std::lock_guard<std::mutex> lk(err_mu);
                    if (!err) err = std::current_exception();
                }
            }
        };
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < threads; ++t) pool.emplace_back(work);
        work();
        for (std::thread &t : pool) t.join();
        if (err) std::rethrow_exception(err);
    }

    // Every world's mirrors from the GPU: the batch is read back once, then distributed.
    void sync() {
        {
            std::lock_guard<std::mutex> lk(b_->mu);
            b_->fetch_bodies();
            b_->fetch_joints();
            b_->fetch_contacts();
            b_->fetch_broad_phase();
        }
        parallel_each([](World &w) {
            w.pull_bodies();
            w.pull_joints();
            w.pull_contacts();
            w.pull_broad_phase();
        });
    }
    void auto_sync(bool on) { auto_sync_ = on; }
    void set_option(int option, int value) {
        std::lock_guard<std::mutex> lk(b_->mu);
        check(rbgpu_set_option(b_->h, option, value));
    }
    // rbgpu_autotune over the pool's batch (every world advances report.frames frames).
    rbgpu_tune_report autotune(scalar dt, int frames_per_trial = 3) {
        rbgpu_tune_report r{};
        {
            std::lock_guard<std::mutex> lk(b_->mu);
            const int sub = worlds.empty() ? b_->sub : worlds[0].substeps;
            for (World &w : worlds) {
                w.push_changes();
                w.timestep = dt;
            }
            b_->set_step(dt, sub);
            check(rbgpu_autotune(b_->h, frames_per_trial, &r));
            b_->touch();
        }
        if (auto_sync_) sync();
        return r;
    }
    rbgpu_world *handle() { return b_->h; }

private:
    std::shared_ptr<detail::Batch> b_;
    bool auto_sync_ = true;

    template <typename Fn> static auto call(Fn &fn, World &w, size_t i) -> decltype(fn(w, i), void()) { fn(w, i); }
    template <typename Fn> static auto call(Fn &fn, World &w, size_t) -> decltype(fn(w), void()) { fn(w); }
};

}  // namespace rbps

// The names this header used before it took the reference's namespaces.
namespace rbps_amd {
using namespace ::m3d;
using namespace ::rbc;
using namespace ::rbps;
}  // namespace rbps_amd
