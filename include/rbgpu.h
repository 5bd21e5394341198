/*
 * rbgpu — C-ABI of the MI355X XPBD substep pipeline (broadphase, narrowphase,
 * contact generation, constraint colouring, position/velocity solve).
 *
 * This is the boundary that replaces the reference's `rbps::World` step path.
 * Plain pointers and sizes only; all state lives on one GPU in FP32 and the
 * host sees it through the get/set calls.  Every call returns RBGPU_OK (0) or a
 * negative status; rbgpu_last_error() describes the last failure.
 *
 * Body and collider IDs follow the reference: a 24-bit slot plus 8 generation
 * bits (include/storage/Soa.hpp:405-416); without removals an ID equals its
 * dense index, which is how the reference pipeline uses them
 * (src/rbps/CollisionPipeline.cpp:34-35, 124-125).
 */
#ifndef RBGPU_H
#define RBGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RBGPU_OK = 0,
    RBGPU_ERR_INVALID = -1,   /* bad argument or handle */
    RBGPU_ERR_HIP = -2,       /* a HIP runtime call failed */
    RBGPU_ERR_CAPACITY = -3,  /* a fixed capacity was exceeded (e.g. > 1024 colours) */
    RBGPU_ERR_OOM = -4,       /* device allocation failed */
    RBGPU_ERR_UNSUPPORTED = -5
};

/* Body types (include/rbps/Body.hpp:11-15). */
enum { RBGPU_STATIC = 0, RBGPU_DYNAMIC = 1 };

/* Shape codes = the reference's rbc::Shape variant indices (include/rbc/shapes/ShapeTypes.hpp:69). */
enum { RBGPU_SPHERE = 0, RBGPU_BOX = 1, RBGPU_CAPSULE = 3, RBGPU_PLANE = 6, RBGPU_HULL = 9 };

/* Fields of rbps::BodyParams (include/rbps/API/BodyAPI.hpp:22-31). 168 bytes. */
typedef struct {
    int32_t type;                /* RBGPU_STATIC / RBGPU_DYNAMIC */
    uint32_t scene;              /* independent scene of a batch (0 = single world); see below */
    double mass;
    double inertia[6];           /* local inertia tensor: xx yy zz xy xz yz */
    double position[3];
    double orientation[4];       /* quaternion w x y z */
    double linear_velocity[3];
    double angular_velocity[3];
} rbgpu_body_desc;

/* Fields of rbps::ColliderParams (include/rbps/API/ColliderAPI.hpp:17-27). 128 bytes.
 * dims: sphere {radius} | box {hx, hy, hz} | capsule {half_height, radius} | plane {nx, ny, nz, d}. */
typedef struct {
    int32_t shape;
    uint32_t body_id;
    int32_t is_static;
    int32_t hull_index;          /* RBGPU_HULL: id from rbgpu_add_hull (else -1) */
    double dims[4];
    double local_pos[3];
    double local_rot[4];         /* w x y z */
    double restitution, static_friction, dynamic_friction;
} rbgpu_collider_desc;

/* Joint types and actuation: the reference's JointType and JointActuationType order
 * (include/rbps/constraints/Joint.hpp:25-47); SPHERICAL is this build's extension
 * (a ball joint: one positional attachment row, plus a swing limit on `axis` when
 * limited, upper_limit = cone half-angle). */
enum {
    RBGPU_JOINT_PRISMATIC = 0,
    RBGPU_JOINT_REVOLUTE = 1,
    RBGPU_JOINT_FIXED = 2,
    RBGPU_JOINT_SPHERICAL = 3,
    RBGPU_JOINT_ROW = 4   /* a constraint row (rbgpu_add_constraints): actuation = its kind, 0 positional, 1 rotational */
};
enum { RBGPU_ACT_FREE = 0, RBGPU_ACT_POSITION = 1, RBGPU_ACT_SPEED = 2 };

/* Fields of rbps::PrismaticJointParams / RevoluteJointParams / FixedJointParams
 * (include/rbps/API/JointAPI.hpp:22-68) plus the joint's public target fields
 * (JointCollection::target_position / target_speed, Joint.hpp:82-106).  168 bytes.
 * axis: prismatic moving_axis | revolute aligned_axis | spherical swing axis;
 * limit_axis: revolute only.  compliance lands on the row the reference gives it
 * (prismatic drive, revolute drive, spherical attachment). */
typedef struct {
    int32_t type;
    int32_t actuation;
    uint32_t body_1, body_2;
    double r_1[3], r_2[3];       /* body-local attachment points */
    double axis[3];
    double limit_axis[3];
    double compliance, damping;
    int32_t limited, _pad;
    double lower_limit, upper_limit;
    double target_position, target_speed;
} rbgpu_joint_desc;

typedef struct {
    uint32_t frames;             /* World::step calls executed */
    uint32_t sap_pairs;          /* broadphase pairs of the last frame (after the pipeline's filters) */
    uint32_t pairs;              /* + plane pairs = the frame's narrowphase pair list */
    uint32_t colours;            /* colours of the last substep's colouring */
    uint32_t colour_rounds;      /* colouring rounds of the last substep */
    uint32_t contacts;           /* contact points of the last substep */
    uint32_t bins[16];           /* pairs per narrowphase bin of the last frame (SS SB BB SC CC CB SP BP CP HH HX HP other) */
    uint32_t active_pairs;       /* pairs with contact points in the last substep */
    uint32_t graph_frames;       /* frames whose substeps ran as a captured hipGraph */
    uint32_t scene_kernels;      /* 1: the last frame coloured and solved per scene */
} rbgpu_stats;

typedef struct rbgpu_world rbgpu_world;

/* World(timestep, substeps) — include/rbps/API/World.hpp:87-106. */
int rbgpu_world_create(double timestep, int substeps, rbgpu_world **out);
int rbgpu_world_destroy(rbgpu_world *w);

/* Batches: bodies carry a scene index (rbgpu_body_desc.scene).  Scenes never interact:
 * the broadphase sorts endpoints by (scene, value), a plane pairs only with its own
 * scene's colliders, joints must join bodies of one scene, and the colouring hashes a
 * pair's index within its own scene, so every scene evolves exactly as it would alone
 * in a world of its own, while every kernel runs once over all scenes. */

/* World::create_body / rbps::create_body — include/rbps/API/World.hpp:113-116,
 * include/rbps/API/BodyAPI.hpp:53-80.  ids_out may be NULL. */
int rbgpu_add_bodies(rbgpu_world *w, int n, const rbgpu_body_desc *desc, uint32_t *ids_out);

/* World::create_collider / rbps::create_collider — include/rbps/API/World.hpp:124-132,
 * include/rbps/API/ColliderAPI.hpp:61-91 (inserts into the broadphase). */
int rbgpu_add_colliders(rbgpu_world *w, int n, const rbgpu_collider_desc *desc, uint32_t *ids_out);

/* World::create_prismatic_joint / create_revolute_joint / create_fixed_joint —
 * include/rbps/API/World.hpp:141-156, include/rbps/API/JointAPI.hpp:160-265 — and the
 * spherical extension.  Joint ids are creation indices.  ids_out may be NULL. */
int rbgpu_add_joints(rbgpu_world *w, int n, const rbgpu_joint_desc *desc, uint32_t *ids_out);

/* World::create_constraint / rbps::create_constraint — include/rbps/API/World.hpp:135-138,
 * include/rbps/API/ConstraintAPI.hpp:37-64: a constraint row between two bodies (fields of
 * rbps::ConstraintParams).  A row starts with direction (1, 0, 0) and magnitude 0, keeps the
 * direction and magnitude its caller writes (rbgpu_set_constraints) and is solved with the
 * joints' rows, in creation order (Constraint.cpp:89-106; here colour by colour with the
 * joints).  It shares the joints' id space: ids_out are the ids rbgpu_get_joints reports it
 * under (its lambda in slot 0). */
typedef struct {
    uint32_t body_1, body_2;
    int32_t type, _pad;          /* 0 POSITIONAL, 1 ROTATIONAL (Constraint.hpp ConstraintType) */
    double r_1[3], r_2[3];       /* body-local attachment points of a positional row */
    double compliance;
} rbgpu_constraint_desc;
int rbgpu_add_constraints(rbgpu_world *w, int n, const rbgpu_constraint_desc *desc, uint32_t *ids_out);
/* Writes ConstraintCollection::direction / magnitude of constraint rows (public fields the
 * reference's callers assign directly); either array may be NULL. */
int rbgpu_set_constraints(rbgpu_world *w, int n, const uint32_t *ids, const double *direction3, const double *magnitude);
/* Every joint's / row's constraint rows after the last step (ConstraintCollection::direction,
 * magnitude): 4 slots per id (unused slots (1, 0, 0), 0); either pointer may be NULL. */
int rbgpu_get_constraint_rows(rbgpu_world *w, double *direction3, double *magnitude);

/* Public joint state after the last step (JointCollection::current_position,
 * target_position; ConstraintCollection::lambda of the joint's rows, 4 per joint, unused
 * rows 0) and the joint's solve colour.  Any pointer may be NULL. */
int rbgpu_get_joints(rbgpu_world *w, double *current_position, double *target_position, double *lambda4,
                     int32_t *colour);

/* Writes JointCollection::target_position / target_speed of the given joints (the
 * reference's callers assign these public fields directly).  Either array may be NULL. */
int rbgpu_set_joint_targets(rbgpu_world *w, int n, const uint32_t *ids, const double *target_position,
                            const double *target_speed);

/* World::step — include/rbps/API/World.hpp:192-212; runs n_frames frames, asynchronously
 * on the world's stream (the call returns once the work is queued and the per-frame
 * bookkeeping read-backs are done). */
int rbgpu_step(rbgpu_world *w, int n_frames);

/* Changes World::timestep / World::substeps (public members of the reference World,
 * include/rbps/API/World.hpp:89-90) for the following steps. */
int rbgpu_set_step(rbgpu_world *w, double timestep, int substeps);

/* Waits for all queued work of the world. */
int rbgpu_synchronize(rbgpu_world *w);

/* Reads / writes the public body state (BodyCollection fields position, orientation,
 * linear_velocity, angular_velocity; include/rbps/Body.hpp:52-69) as doubles,
 * n*3 / n*4 values; any pointer may be NULL. */
int rbgpu_get_bodies(rbgpu_world *w, double *position, double *orientation, double *linear_velocity,
                     double *angular_velocity);
int rbgpu_set_bodies(rbgpu_world *w, const double *position, const double *orientation,
                     const double *linear_velocity, const double *angular_velocity);

/* BodyCollection::force / torque (Body.hpp:52-69; create_body sets force = (0, -9.8 m, 0)
 * and torque = 0, BodyAPI.hpp:61-62): the constant external load the integrator applies
 * every substep (Body.cpp:28-35).  Writes / reads n*3 values each; either pointer may be
 * NULL.  Held until overwritten; bodies added later start with the defaults. */
int rbgpu_set_body_forces(rbgpu_world *w, const double *force3, const double *torque3);
int rbgpu_get_body_forces(rbgpu_world *w, double *force3, double *torque3);

/* Broadphase output of the last frame (rbc::BroadPhaseState::pairs, include/rbc/BroadPhase.hpp:117)
 * with the narrowphase driver's filters applied, in canonical sweep order: (id_a, id_b)
 * with id_a < id_b.  Writes min(cap, n) pairs, *n = total. */
int rbgpu_get_sap_pairs(rbgpu_world *w, uint32_t *pairs2, uint32_t cap, uint32_t *n);

/* The last frame's pair list (SAP pairs then plane pairs) with each pair's colour. */
int rbgpu_get_pair_colours(rbgpu_world *w, uint32_t *pairs2, int32_t *colour, uint32_t cap, uint32_t *n);

/* Contacts of the last substep (rbps::ContactList, include/rbps/CollisionPipeline.hpp:46-139),
 * listed per pair in pair-list order, points in manifold order.  ids4 = body_a, body_b,
 * collider_a, collider_b; lam2 = normal_lambda, tangent_lambda.  Any pointer may be NULL. */
int rbgpu_get_contacts(rbgpu_world *w, uint32_t *ids4, double *normal3, double *r_a_local3, double *r_b_local3,
                       double *depth, double *lam2, uint32_t cap, uint32_t *n);

/* More of the last substep's ContactList (CollisionPipeline.hpp:46-139), in the order of
 * rbgpu_get_contacts: point_on_a / point_on_b (world points after the position solve),
 * relative_velocity (pre-solve normal velocity) and collision (0 once the bodies separated
 * in the position solve).  Any pointer may be NULL. */
int rbgpu_get_contact_points(rbgpu_world *w, double *point_on_a3, double *point_on_b3, double *relative_velocity,
                             int32_t *collision, uint32_t cap, uint32_t *n);

/* rbc::BroadPhaseState::objects (BroadPhase.hpp:91-97, 117): per broadphase object (every
 * collider but planes, in creation order) its fat AABB and user id (the collider id);
 * writes min(cap, n), *n = total.  Any array may be NULL. */
int rbgpu_get_broadphase_objects(rbgpu_world *w, double *fat_min3, double *fat_max3, uint32_t *user_id, uint32_t cap,
                                 uint32_t *n);

int rbgpu_get_stats(rbgpu_world *w, rbgpu_stats *out);

/* Kernel timing with HIP events around every launch of one kernel family on the world's
 * stream.  kernel: 0 off, 1 position solve (per-colour launches), 2 velocity solve
 * (per-colour launches), 3 narrowphase (with the colouring state's reset, which runs
 * beside it), 4 integrate, 5 velocity update, 6 broadphase
 * (refit, endpoint sort, sweep, pair list; one span per frame), 7 colouring (one span per
 * substep), 8 position solve runs of small colours (single workgroup), 9 velocity solve
 * runs, 10 joints (empty when the per-scene solves run the joints), 11 the GJK pass
 * (k_gjk_pass, GJK on the candidates of the hull bins).  rbgpu_get_timing
 * returns the summed milliseconds and launches since the last reset (synchronises);
 * bytes = algorithmic bytes of the timed launches for 1, 2 and 11. */
/* Pipeline options (no counterpart in the reference, whose step is sequential).
 *   RBGPU_OPT_PALETTE_MAX  colouring: a pair draws its tentative colour from at most this
 *                          many free colours (default 8).  Changes the colouring, hence the
 *                          solve order and the trajectory; the other options do not change
 *                          any result.
 *   RBGPU_OPT_HEAD_ROUNDS  colouring: most grid-wide rounds per substep (default 96, <= 256)
 *   RBGPU_OPT_ROUND_BATCH  colouring: grid-wide rounds launched between host checks (12)
 *   RBGPU_OPT_TAIL_MIN     colouring: at most this many uncoloured pairs go to the
 *                          single-workgroup rounds (2048)
 *   RBGPU_OPT_RUN_MIN      solve: colours with fewer pairs run in single-workgroup runs (256)
 *   RBGPU_OPT_TIMING_STRIDE  kernel timing (rbgpu_set_timing) covers every n-th frame (1)
 *   RBGPU_OPT_JOINT_MODE     joint solver (0 = the reference's, default).  The reference
 *                          evaluates every joint's errors once per substep before any row
 *                          is solved, never resets lambda, gives rows with errors in
 *                          (1e-9, ~3.2e-5) a zero direction (a ROTATIONAL row of compliance 0
 *                          then evaluates 0/0: NaN poses) and builds the revolute alignment
 *                          row as a1 x a2, which turns the axes apart; chains come apart and
 *                          hinges diverge.  1 = XPBD rows (Mueller et al. 2020): each row's
 *                          error re-evaluated just before it is solved, lambda reset every
 *                          substep, direction v/|v| whenever |v| > 1e-9, alignment a2 x a1.
 *   RBGPU_OPT_SCENE_KERNELS  batches: colour and solve each scene in one workgroup (one launch
 *                          each per substep, no host round trip) instead of grid-wide rounds
 *                          and per-colour launches; -1 auto (default: on for 16 or more
 *                          scenes of at most 4096 bodies), 0 off, 1 on.  Same colours, same
 *                          states.  A colouring error (more than 1024 colours, or no
 *                          convergence) is then read back after the frame that raised it: by
 *                          the next frame of the same rbgpu_step call, or, for the call's last
 *                          frames, by the next call that waits for the world (rbgpu_synchronize,
 *                          rbgpu_get_bodies/_sap_pairs/_pair_colours/_contacts/_stats, or the
 *                          next rbgpu_step), which returns RBGPU_ERR_CAPACITY.  Past 1024
 *                          colours the frame's colouring is not valid.
 *   RBGPU_OPT_SCENE_THREADS  threads per workgroup of the per-scene solve (64, 128, 256; 64).  The
 *                          per-scene solves also run the scene's joints (timing classes 1 and 2
 *                          then include them; class 10 stays empty).
 *   RBGPU_OPT_GJK_CULL     exact bounding-sphere rejection in front of GJK (1, default; 0 runs
 *                          GJK on every pair of the hull bins)
 *   RBGPU_OPT_GJK_GRID, RBGPU_OPT_EPA_GRID  workgroups of the two GJK/EPA passes (4096, 8192)
 *   RBGPU_OPT_GJK_ROWS     layout of the GJK pass when every hull has at most 64 vertices: 4
 *                          (default) GJK, EPA and the manifold in one pass, four pairs per
 *                          wavefront, one per 16-lane DPP row, polytope and clip buffers per row in
 *                          LDS (a pair that outgrows them is finished by the full-size EPA pass);
 *                          5 the same with row buffers most pairs outgrow (tests); 1 GJK alone per
 *                          row, EPA in a second pass (one pair per wavefront); 0 one pair per
 *                          wavefront; 2 one pair per lane (each lane scans its hulls' vertices); 3
 *                          rows refilled one candidate at a time.  Larger hulls: one pair per
 *                          wavefront.  Every layout gives the same results
 *   RBGPU_OPT_EPA_SUB      listed pairs per EPA work item (1, 2, 4, 8, 16; default 16)
 *   RBGPU_OPT_GJK_OCCUPANCY  waves per EU the row GJK pass budgets its registers for (2, 3, 4)
 *   RBGPU_OPT_EPA_OCCUPANCY, RBGPU_OPT_SOLVE_OCCUPANCY  the same for the EPA pass and the
 *                          per-scene position solve (0 = no budget, 4)
 *   RBGPU_OPT_SIDE_STREAM  run the narrowphase bins other than GJK/EPA on a second stream of
 *                          the world, beside the GJK chain, joined before the colouring; the
 *                          value is where it forks: 1 before the GJK chain, 2 after its
 *                          rejection tests (default), 3 after the GJK pass; 0: all on the
 *                          world's stream.  Results do not change.
 *   RBGPU_OPT_AXIS_CARRY   carry the GJK rejection hints (the axis that last separated a pair)
 *                          across the frame's new pair list, keyed by collider pair (1,
 *                          default; 0: every pair starts a frame without a hint).  Results do
 *                          not change.
 *   RBGPU_OPT_GJK_SORT     order of the GJK candidates, from their slots' previous substep:
 *                          0 (default) bin order; 1 by GJK iterations, so the four pairs of a
 *                          GJK wavefront take similar numbers of iterations; 2 overlapping pairs
 *                          after separated ones (stable), so the rows of a wavefront mostly all
 *                          run EPA or none does; 3 both.  Results do not change.
 *   RBGPU_OPT_EPA_GUARD    0 (default): EPA depths and normals as the reference computes them.
 *                          Any direction n bounds the penetration depth from above by the
 *                          shapes' overlap along n (the Minkowski difference's support along
 *                          n); when EPA's face is deeper than the smallest overlap over a set of
 *                          directions by more than 1e-4 m, the FP32 EPA has settled on a face
 *                          that is not the minimum, and the contact takes that direction and
 *                          overlap instead.  1: the set is the pair's contact normal of the
 *                          previous substep (same frame); 2: that and, while the depth in
 *                          question exceeds 2 mm, the face normals of both shapes (hull
 *                          polygons, box faces).  Changes results (of pairs where FP32 EPA goes
 *                          wrong, typically touching hulls).
 *   RBGPU_OPT_GRAPH        with the per-scene kernels: capture one substep per frame in a
 *                          hipGraph and launch it for every substep (1, default; frames with
 *                          kernel timing or an open training log run without it)
 *   RBGPU_OPT_CONTACT_MODE  0 (default): contacts exactly as the reference makes them.  The
 *                          reference's capsule-capsule routine returns a normal from B to A
 *                          (CapsuleCapsule.hpp:68) where every other routine returns one from
 *                          A to B, so overlapping capsules are pushed into each other and
 *                          capsule piles gain energy without bound.  1: that normal is turned
 *                          around (points and depths unchanged). */
enum {
    RBGPU_OPT_PALETTE_MAX = 0,
    RBGPU_OPT_HEAD_ROUNDS = 1,
    RBGPU_OPT_ROUND_BATCH = 2,
    RBGPU_OPT_TAIL_MIN = 3,
    RBGPU_OPT_RUN_MIN = 4,
    RBGPU_OPT_TIMING_STRIDE = 5,
    RBGPU_OPT_JOINT_MODE = 6,
    RBGPU_OPT_SCENE_KERNELS = 7,
    RBGPU_OPT_SCENE_THREADS = 8,
    RBGPU_OPT_GJK_CULL = 9,
    RBGPU_OPT_GJK_GRID = 10,
    RBGPU_OPT_EPA_GRID = 11,
    RBGPU_OPT_GRAPH = 12,
    RBGPU_OPT_CONTACT_MODE = 13,
    RBGPU_OPT_GJK_ROWS = 14,
    RBGPU_OPT_EPA_SUB = 15,
    RBGPU_OPT_GJK_OCCUPANCY = 16,
    RBGPU_OPT_EPA_OCCUPANCY = 17,
    RBGPU_OPT_SOLVE_OCCUPANCY = 18,
    RBGPU_OPT_SIDE_STREAM = 19,
    RBGPU_OPT_AXIS_CARRY = 20,
    RBGPU_OPT_EPA_GUARD = 21,
    RBGPU_OPT_GJK_SORT = 22
};
int rbgpu_set_option(rbgpu_world *w, int option, int value);
int rbgpu_get_option(rbgpu_world *w, int option, int *value);

/* Autotuner (no counterpart in the reference).  Picks this world's launch options that
 * cannot change any result (per-scene solve workgroup size and the substep graph, or the
 * grid path's run minimum and round batch) by coordinate descent: each candidate value is timed twice over `frames_per_trial` real
 * frames (wall clock between two stream synchronisations; two passes over the values, the second in reverse order, since the scene
 * moves between trials) and the value with the smallest total is kept.  The trial
 * frames advance the world exactly as untuned frames would, since every candidate gives
 * bit-identical states; callers typically tune during warm-up.  Requires timing off. */
typedef struct {
    int32_t scene_kernels, scene_threads, graph, run_min, round_batch, gjk_grid, epa_grid;
    int32_t trials, frames;      /* timings taken (two per candidate), frames stepped (trials + one settling frame) */
    double best_ms_per_frame;    /* of the last option tuned */
} rbgpu_tune_report;
int rbgpu_autotune(rbgpu_world *w, int frames_per_trial, rbgpu_tune_report *out);

int rbgpu_set_timing(rbgpu_world *w, int kernel);
int rbgpu_get_timing(rbgpu_world *w, double *total_ms, uint64_t *launches, double *bytes);

/* Narrowphase of n independent shape pairs, no world (rbc::test_narrow_phase,
 * src/rbc/narrow_phase.cpp:8-13, dispatching as include/rbc/Dispatcher.hpp:212-217).
 * kind: shape codes; dims: 4 per shape; tf: 7 per shape (px py pz qw qx qy qz).
 * out: 20 per pair = n_points, normal xyz, then (point xyz, depth) x 4 (zeros when no hit). */
int rbgpu_collide_pairs(int n, const int32_t *kind_a, const double *dims_a, const double *tf_a, const int32_t *kind_b,
                        const double *dims_b, const double *tf_b, double *out);

/* The same narrowphase as rbgpu_collide_pairs, same arguments and output, but it allocates
 * nothing: no hipMalloc/hipFree or other allocating call.  No counterpart in the reference;
 * for callers that cannot allocate device memory (tight serving loops, harnesses that forbid
 * it).  Inputs and results are staged in one fixed device array per GPU (8.5 MB, set up by
 * the first call on that GPU), so one call takes at most 16384 pairs (RBGPU_ERR_CAPACITY
 * beyond that).  Every call on a GPU shares that GPU's array and writes every element it
 * reads, so nothing carries over between calls; calls that overlap on one GPU (from several
 * threads) race on it and must be serialised by the caller.  Runs on the null stream and
 * returns when the result is in `out`. */
int rbgpu_collide_pairs_noalloc(int n, const int32_t *kind_a, const double *dims_a, const double *tf_a,
                                const int32_t *kind_b, const double *dims_b, const double *tf_b, double *out);

/* ---- convex hulls ------------------------------------------------------------------
 * rbc::convex_hull_data_create (include/rbc/shapes/ConvexHull.hpp:149-420): vertices
 * (3 doubles each, local frame) and optional CCW outward triangles; the world copies them,
 * merges coplanar triangles into polygons and builds the adjacency tables the hull-hull
 * manifold uses.  A hull collider (shape RBGPU_HULL) names the returned id in hull_index.
 * Limits: 1..256 vertices, merged polygons of at most 64 vertices and 64 neighbours
 * (RBGPU_ERR_CAPACITY otherwise).  Pairs: hull-plane is analytic (PlaneCollision.hpp:90-285);
 * any other pair with a hull runs GJK + EPA + manifold (ContactManifoldGenerator.cpp:415-438),
 * with a face/edge manifold when both hulls have triangles, else one point. */
int rbgpu_add_hull(rbgpu_world *w, const double *vertices, uint32_t n_vertices, const uint32_t *triangles,
                   uint32_t n_triangles, uint32_t *hull_id_out);
/* compute_volume / compute_inertia_tensor (ConvexHull.hpp:575-745): unit-density volume
 * and inertia about the centroid (xx yy zz xy xz yz); vertex-only hulls use the local AABB. */
int rbgpu_hull_mass_properties(const double *vertices, uint32_t n_vertices, const uint32_t *triangles,
                               uint32_t n_triangles, double *volume, double *inertia6);
/* The tables rbgpu_add_hull builds (ConvexHullData's poly_* and adjacency arrays,
 * ConvexHull.hpp:79-98).  sizes6 = polygons, polygon indices, v2p, v2n and p2p entries,
 * adjacency flag; call with null arrays first to size them.  Offsets arrays hold
 * polygons + 1 (poly, p2p) or vertices + 1 (v2p, v2n) entries. */
int rbgpu_hull_tables(const double *vertices, uint32_t n_vertices, const uint32_t *triangles, uint32_t n_triangles,
                      uint32_t *sizes6, double *poly_normals, uint32_t *poly_offsets, uint32_t *poly_indices,
                      uint32_t *v2p_offsets, uint32_t *v2p_indices, uint32_t *v2n_offsets, uint32_t *v2n_indices,
                      uint32_t *p2p_offsets, uint32_t *p2p_indices);
/* The narrowphase (rbc::dispatch, Dispatcher.hpp:212-217) of given collider pairs at the
 * world's current state, through the pipeline's own binned kernels; a plane must be the
 * second collider (as in the pipeline's plane loop, CollisionPipeline.cpp:156-197).
 * out32 per pair: points, normal xyz, then per point the two surface points p_a, p_b
 * (CollisionPipeline.cpp:59-104) and the depth (7 values) x 4.  gjk_stats (optional) per
 * pair: GJK iterations | EPA iterations << 8 | GJK status << 16 | EPA status << 20 (EPA
 * status 15 = not run; 0xffffffff = not a GJK pair; bit 24: the plain fallback ran).
 * guess6 (optional, per pair, A's frame): GJK warm-start direction xyz and EPA seed normal
 * xyz (used when seed_epa) in place of the network's predictions (zero: plain seed).
 * Replaces the last substep's contacts. */
int rbgpu_collide_world_pairs(rbgpu_world *w, uint32_t n, const uint32_t *pairs2, double *out32, uint32_t *gjk_stats,
                              const float *guess6, int seed_epa);

/* ---- learned warm start of GJK/EPA (no counterpart in the reference) -----------------
 * A small MLP (32 -> 128 -> 128 -> 128 -> 16, ReLU; BF16 weights on MFMA, FP32
 * accumulation) maps a GJK pair's relative pose and shape descriptors to a GJK initial
 * direction and an EPA seed normal.  The reference seeds GJK with tf_b.pos - tf_a.pos
 * (ContactManifoldGenerator.cpp:423-425) and EPA with nothing; here plain GJK/EPA still
 * run to the same tolerances and are the fallback when a warm-started run fails, so
 * predictions change iteration counts (and, within the EPA tolerance, which face EPA
 * stops on), not the convergence criteria.
 *   rbgpu_mlp_init        allocate and initialise the network (He-uniform, seeded)
 *   rbgpu_mlp_set/get_weights  the 39,312 FP32 parameters: per layer W[out][in], bias
 *   rbgpu_mlp_set_mode    0 plain (default), 1 predicted GJK direction, 2 + EPA seed
 *   rbgpu_mlp_log         log training samples (features, targets) of the GJK pairs of
 *                         every substep until `capacity` samples (0: drop the log)
 *   rbgpu_mlp_logged      samples logged (stop != 0 ends logging); rbgpu_mlp_get_log reads them
 *   rbgpu_mlp_train       Adam steps on batches drawn from the log; mean batch loss
 *   rbgpu_mlp_eval        inference on given features (n x 32 float) -> n x 6
 *   rbgpu_mlp_gradients   gradient of the mean loss on given features and targets
 *                         (n x 8: GJK target xyz, valid, EPA target xyz, valid)
 *   rbgpu_get_gjk_totals  since the last reset: GJK pairs, GJK iterations, EPA
 *                         iterations, overlapping pairs, EPA converged, fallbacks
 *   rbgpu_get_gjk_counts  the same n <= RBGPU_GJK_COUNTS counters, the seventh the pairs the
 *                         GJK + EPA row pass (RBGPU_OPT_GJK_ROWS 4, 5) handed to the full-size
 *                         EPA pass (no reference counterpart: instrumentation) */
int rbgpu_mlp_init(rbgpu_world *w, uint64_t seed);
int rbgpu_mlp_set_weights(rbgpu_world *w, const float *params, uint32_t n);
int rbgpu_mlp_get_weights(rbgpu_world *w, float *params, uint32_t n);
int rbgpu_mlp_set_mode(rbgpu_world *w, int mode);
int rbgpu_mlp_log(rbgpu_world *w, uint64_t capacity);
int rbgpu_mlp_logged(rbgpu_world *w, uint64_t *n, int stop);
/* Logged samples [start, start + n): features (n x 32) and targets (n x 8), as float. */
int rbgpu_mlp_get_log(rbgpu_world *w, uint64_t start, uint32_t n, float *features32, float *targets8);
int rbgpu_mlp_train(rbgpu_world *w, int steps, uint32_t batch, double lr, double *mean_loss);
int rbgpu_mlp_eval(rbgpu_world *w, uint32_t n, const float *features32, float *out6);
int rbgpu_mlp_gradients(rbgpu_world *w, uint32_t n, const float *features32, const float *targets8, float *grad,
                        double *loss);
int rbgpu_get_gjk_totals(rbgpu_world *w, uint64_t *out6, int reset);
#define RBGPU_GJK_COUNTS 7
int rbgpu_get_gjk_counts(rbgpu_world *w, uint64_t *out, uint32_t n, int reset);

/* Device-resident snapshot and restore of a world's simulation state (the reference's
 * WorldSnapshot design, docs/design/world-snapshot-restore/world-snapshot-restore.md:29-45,
 * held in device memory).  A snapshot holds what carries from one frame to the next: body
 * poses, velocities and previous poses, forces and torques, the broadphase's fat AABBs and
 * endpoint order, joint targets, current positions and lambdas, constraint rows, the GJK
 * rejection hints, the frame count, timestep / substeps and every option.  Restoring it and
 * stepping repeats the original run bit for bit.  Contacts, pair lists and colours are
 * rebuilt by the next frame (their getters describe the last frame stepped until then);
 * the learned warm start's weights are parameters, not state (rbgpu_mlp_get_weights).
 *   create   allocates the snapshot's device block and captures the world (builds it first
 *            if needed); capture re-captures into an existing snapshot, reusing its block
 *            when it is large enough.
 *   restore  needs a world with the same bodies, colliders, joints, scenes and hulls (the
 *            world it came from, or one loaded from the same scene), waits for it, and
 *            copies device to device.
 *   size / save / load  the serialised form for a checkpoint file: header + the section
 *            block, with a checksum that load verifies. */
typedef struct rbgpu_snapshot rbgpu_snapshot;
int rbgpu_snapshot_create(rbgpu_world *w, rbgpu_snapshot **out);
int rbgpu_snapshot_capture(rbgpu_world *w, rbgpu_snapshot *s);
int rbgpu_snapshot_restore(rbgpu_world *w, const rbgpu_snapshot *s);
int rbgpu_snapshot_size(const rbgpu_snapshot *s, uint64_t *bytes);
int rbgpu_snapshot_save(const rbgpu_snapshot *s, void *host, uint64_t cap);
int rbgpu_snapshot_load(const void *host, uint64_t n, rbgpu_snapshot **out);
int rbgpu_snapshot_destroy(rbgpu_snapshot *s);

const char *rbgpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* RBGPU_H */
