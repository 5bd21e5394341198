// rbps_amd::World — a source-compatible stand-in for rbps::World
// (include/rbps/API/World.hpp:87-212 of the reference) whose step() runs the
// MI355X pipeline through the rbgpu C-ABI (include/rbgpu.h).
//
// Same names and argument meaning as the reference API:
//   World(timestep, substeps); create_body(BodyParams); create_collider(ColliderParams);
//   step(); bodies.position[bodies.index_of(id)] etc.
// Differences a caller can see:
//   * after step() the body mirror is refreshed from the GPU (set auto_sync(false)
//     and call sync_bodies() to batch read-backs);
//   * writes into the mirror are pushed with push_bodies();
//   * the reference's silent failure modes become exceptions (rbps_amd::Error)
//     carrying the library's message.
// Header-only; link with -lrbgpu.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../rbgpu.h"

namespace rbps_amd {

using scalar = double;

struct vec3 {
    scalar x = 0, y = 0, z = 0;
    vec3() = default;
    vec3(scalar a, scalar b, scalar c) : x(a), y(b), z(c) {}
    explicit vec3(scalar d) : x(d), y(d), z(d) {}
};
struct quat {
    scalar w = 1, x = 0, y = 0, z = 0;
    quat() = default;
    quat(scalar a, scalar b, scalar c, scalar d) : w(a), x(b), y(c), z(d) {}
};
struct smat3 {
    scalar xx = 0, yy = 0, zz = 0, xy = 0, xz = 0, yz = 0;
    smat3() = default;
    smat3(scalar a, scalar b, scalar c, scalar d, scalar e, scalar f) : xx(a), yy(b), zz(c), xy(d), xz(e), yz(f) {}
};

enum BodyType { STATIC = 0, DYNAMIC = 1 };   // include/rbps/Body.hpp:11-15

// rbps::BodyParams (include/rbps/API/BodyAPI.hpp:22-31)
struct BodyParams {
    BodyType type = DYNAMIC;
    scalar mass = 1.0;
    smat3 inertia_tensor = smat3{1.0, 1.0, 1.0, 0, 0, 0};
    vec3 position = vec3(0.0);
    quat orientation = quat{1, 0, 0, 0};
    vec3 linear_velocity = vec3(0.0);
    vec3 angular_velocity = vec3(0.0);
};

// Shapes of rbc::Shape used on the GPU path (include/rbc/shapes/*.hpp defaults kept)
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
// rbc::ConvexHullData (include/rbc/shapes/ConvexHull.hpp:79-98): non-owning over the
// caller's vertices and CCW triangles (triangles optional).  The world copies the data
// (rbgpu_add_hull) the first time a collider uses it; the arrays must stay valid until then.
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
// rbc::ConvexHull (ConvexHull.hpp:101-121)
struct ConvexHull {
    const ConvexHullData *data = nullptr;
    ConvexHull() = default;
    explicit ConvexHull(const ConvexHullData *d) : data(d) {}
};
// compute_volume / compute_inertia_tensor (ConvexHull.hpp:575-745), unit density
inline scalar compute_volume(const ConvexHull &h) {
    double vol = 0.0;
    if (h.data && rbgpu_hull_mass_properties(&h.data->vertices[0].x, h.data->vert_count, h.data->face_indices,
                                             h.data->face_count, &vol, nullptr) != RBGPU_OK)
        throw std::runtime_error(std::string("rbgpu: ") + rbgpu_last_error());
    return vol;
}
inline smat3 compute_inertia_tensor(const ConvexHull &h) {
    double I[6] = {0, 0, 0, 0, 0, 0};
    if (h.data && rbgpu_hull_mass_properties(&h.data->vertices[0].x, h.data->vert_count, h.data->face_indices,
                                             h.data->face_count, nullptr, I) != RBGPU_OK)
        throw std::runtime_error(std::string("rbgpu: ") + rbgpu_last_error());
    return smat3(I[0], I[1], I[2], I[3], I[4], I[5]);
}

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

// rbps::ColliderParams (include/rbps/API/ColliderAPI.hpp:17-27)
struct ColliderParams {
    Shape shape;
    uint32_t body_id = 0;
    bool is_static = false;
    vec3 local_pos = vec3(0, 0, 0);
    quat local_rot = quat(1, 0, 0, 0);
    scalar restitution = 0.5;
    scalar static_friction = 0.5;
    scalar dynamic_friction = 0.5;
};

// Joints (include/rbps/constraints/Joint.hpp:25-47, include/rbps/API/JointAPI.hpp:22-68)
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

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
inline void check(int rc) {
    if (rc != RBGPU_OK) throw Error(rc, rbgpu_last_error());
}

// Host mirror of the public BodyCollection fields (include/rbps/Body.hpp:52-69).
struct BodyMirror {
    std::vector<BodyType> type;
    std::vector<scalar> mass;
    std::vector<vec3> position, linear_velocity, angular_velocity;
    std::vector<quat> orientation;
    uint32_t count() const { return (uint32_t)position.size(); }
    // Without removals an ID is its dense index (include/storage/Soa.hpp:405-416).
    uint32_t index_of(uint32_t id) const { return id & 0x00ffffffu; }
};

struct World {
    scalar timestep = 1.0 / 60.0;
    int substeps = 20;
    BodyMirror bodies;

    World() { create(); }
    World(scalar timestep_, int substeps_) : timestep(timestep_), substeps(substeps_) { create(); }
    ~World() {
        if (h_) rbgpu_world_destroy(h_);
    }
    World(const World &) = delete;
    World &operator=(const World &) = delete;

    // World::create_body (World.hpp:113-116)
    uint32_t create_body(const BodyParams &p) {
        rbgpu_body_desc d{};
        d.type = p.type == STATIC ? RBGPU_STATIC : RBGPU_DYNAMIC;
        d.mass = p.mass;
        const scalar I[6] = {p.inertia_tensor.xx, p.inertia_tensor.yy, p.inertia_tensor.zz,
                             p.inertia_tensor.xy, p.inertia_tensor.xz, p.inertia_tensor.yz};
        for (int k = 0; k < 6; ++k) d.inertia[k] = I[k];
        d.position[0] = p.position.x; d.position[1] = p.position.y; d.position[2] = p.position.z;
        d.orientation[0] = p.orientation.w; d.orientation[1] = p.orientation.x;
        d.orientation[2] = p.orientation.y; d.orientation[3] = p.orientation.z;
        d.linear_velocity[0] = p.linear_velocity.x; d.linear_velocity[1] = p.linear_velocity.y;
        d.linear_velocity[2] = p.linear_velocity.z;
        d.angular_velocity[0] = p.angular_velocity.x; d.angular_velocity[1] = p.angular_velocity.y;
        d.angular_velocity[2] = p.angular_velocity.z;
        uint32_t id = 0;
        check(rbgpu_add_bodies(h_, 1, &d, &id));
        bodies.type.push_back(p.type);
        bodies.mass.push_back(p.mass);
        bodies.position.push_back(p.position);
        bodies.orientation.push_back(p.orientation);
        bodies.linear_velocity.push_back(p.linear_velocity);
        bodies.angular_velocity.push_back(p.angular_velocity);
        return id;
    }

    // The world's copy of a caller's hull data (added on first use).
    int32_t hull_id(const ConvexHullData *h) {
        if (!h || !h->vertices || !h->vert_count) throw Error(RBGPU_ERR_INVALID, "convex hull collider without hull data");
        for (size_t i = 0; i < hull_keys_.size(); ++i)
            if (hull_keys_[i] == h) return (int32_t)i;
        uint32_t id = 0;
        check(rbgpu_add_hull(h_, &h->vertices[0].x, h->vert_count, h->face_indices, h->face_count, &id));
        hull_keys_.push_back(h);
        return (int32_t)id;
    }
    std::vector<const ConvexHullData *> hull_keys_;

    // World::create_collider (World.hpp:124-132)
    uint32_t create_collider(const ColliderParams &p) {
        rbgpu_collider_desc d{};
        d.shape = p.shape.kind;
        d.body_id = p.body_id;
        d.is_static = p.is_static ? 1 : 0;
        d.hull_index = p.shape.kind == RBGPU_HULL ? hull_id(p.shape.hull) : -1;
        for (int k = 0; k < 4; ++k) d.dims[k] = p.shape.dims[k];
        d.local_pos[0] = p.local_pos.x; d.local_pos[1] = p.local_pos.y; d.local_pos[2] = p.local_pos.z;
        d.local_rot[0] = p.local_rot.w; d.local_rot[1] = p.local_rot.x; d.local_rot[2] = p.local_rot.y;
        d.local_rot[3] = p.local_rot.z;
        d.restitution = p.restitution;
        d.static_friction = p.static_friction;
        d.dynamic_friction = p.dynamic_friction;
        uint32_t id = 0;
        check(rbgpu_add_colliders(h_, 1, &d, &id));
        return id;
    }

    // World::create_prismatic_joint / create_revolute_joint / create_fixed_joint
    // (include/rbps/API/World.hpp:141-156) and the spherical extension.
    uint32_t create_prismatic_joint(const PrismaticJointParams &p) {
        rbgpu_joint_desc d = joint_base(RBGPU_JOINT_PRISMATIC, p.body_1, p.body_2, p.r_1, p.r_2);
        set3(d.axis, p.moving_axis);
        d.actuation = p.actuation_type; d.compliance = p.compliance; d.damping = p.damping;
        d.limited = p.limited ? 1 : 0; d.lower_limit = p.lower_limit; d.upper_limit = p.upper_limit;
        return add_joint(d);
    }
    uint32_t create_revolute_joint(const RevoluteJointParams &p) {
        rbgpu_joint_desc d = joint_base(RBGPU_JOINT_REVOLUTE, p.body_1, p.body_2, p.r_1, p.r_2);
        set3(d.axis, p.aligned_axis);
        set3(d.limit_axis, p.limit_axis);
        d.actuation = p.actuation_type; d.compliance = p.compliance; d.damping = p.damping;
        d.limited = p.limited ? 1 : 0; d.lower_limit = p.lower_limit; d.upper_limit = p.upper_limit;
        return add_joint(d);
    }
    uint32_t create_fixed_joint(const FixedJointParams &p) {
        return add_joint(joint_base(RBGPU_JOINT_FIXED, p.body_1, p.body_2, p.r_1, p.r_2));
    }
    uint32_t create_spherical_joint(const SphericalJointParams &p) {
        rbgpu_joint_desc d = joint_base(RBGPU_JOINT_SPHERICAL, p.body_1, p.body_2, p.r_1, p.r_2);
        set3(d.axis, p.swing_axis);
        d.compliance = p.compliance;
        d.limited = p.limited ? 1 : 0;
        d.upper_limit = p.swing_limit;
        return add_joint(d);
    }
    // Pipeline options (rbgpu_set_option): RBGPU_OPT_* of rbgpu.h.
    void set_option(int option, int value) { check(rbgpu_set_option(h_, option, value)); }
    // Times the result-neutral launch options over real frames and keeps the fastest
    // (rbgpu_autotune; no counterpart in the reference): the world advances report.frames
    // frames exactly as it would untuned.  Refreshes the body mirror like step().
    rbgpu_tune_report autotune(int frames_per_trial = 3) {
        if (timestep != ts_ || substeps != sub_) {
            check(rbgpu_set_step(h_, timestep, substeps));
            ts_ = timestep;
            sub_ = substeps;
        }
        rbgpu_tune_report r{};
        check(rbgpu_autotune(h_, frames_per_trial, &r));
        if (auto_sync_) sync_bodies();
        return r;
    }

    // JointCollection::target_position / target_speed are public fields in the reference.
    void set_joint_target(uint32_t joint_id, scalar target_position, scalar target_speed) {
        check(rbgpu_set_joint_targets(h_, 1, &joint_id, &target_position, &target_speed));
    }
    // JointCollection::current_position of every joint after the last step.
    std::vector<scalar> joint_positions() {
        std::vector<scalar> cur(n_joints_);
        if (n_joints_) check(rbgpu_get_joints(h_, cur.data(), nullptr, nullptr, nullptr));
        return cur;
    }

    // World::step (World.hpp:192-212): one frame of `substeps` substeps of timestep/substeps.
    void step() {
        if (timestep != ts_ || substeps != sub_) {
            check(rbgpu_set_step(h_, timestep, substeps));
            ts_ = timestep;
            sub_ = substeps;
        }
        check(rbgpu_step(h_, 1));
        if (auto_sync_) sync_bodies();
    }

    void auto_sync(bool on) { auto_sync_ = on; }

    // GPU -> mirror
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
    }
    // mirror -> GPU
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
    }

    uint32_t contact_count() {
        rbgpu_stats s{};
        check(rbgpu_get_stats(h_, &s));
        return s.contacts;
    }
    rbgpu_world *handle() { return h_; }

private:
    uint32_t n_joints_ = 0;
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
    uint32_t add_joint(const rbgpu_joint_desc &d) {
        uint32_t id = 0;
        check(rbgpu_add_joints(h_, 1, &d, &id));
        ++n_joints_;
        return id;
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

}  // namespace rbps_amd
