"""Scene descriptors and seeded scene generators.

The descriptor dtypes mirror the C structs ``rbgpu_body_desc``,
``rbgpu_collider_desc`` and ``rbgpu_joint_desc`` declared in ``include/rbgpu.h``
(the fields of the reference's ``rbps::BodyParams``, include/rbps/API/BodyAPI.hpp:22-31,
``rbps::ColliderParams``, include/rbps/API/ColliderAPI.hpp:17-27, and the
``Prismatic/Revolute/FixedJointParams`` of include/rbps/API/JointAPI.hpp:22-68 plus the
joint's public target fields).

Shape codes are the reference's ``rbc::Shape`` variant indices
(include/rbc/shapes/ShapeTypes.hpp:69): sphere 0, box 1, capsule 3, plane 6, convex
hull 9.  A hull collider's ``hull_index`` indexes the scene's ``hulls`` list of
(vertices (n,3), CCW triangles (m,3) or None), the arrays a caller hands to
``rbc::convex_hull_data_create`` (include/rbc/shapes/ConvexHull.hpp:149).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

STATIC, DYNAMIC = 0, 1
PRISMATIC, REVOLUTE, FIXED, SPHERICAL = 0, 1, 2, 3      # Joint.hpp JointType (+ spherical)
ROW = 4   # a constraint row (World::create_constraint): actuation = its type, 0 POSITIONAL, 1 ROTATIONAL
ROW_POSITIONAL, ROW_ROTATIONAL = 0, 1                    # Constraint.hpp ConstraintType
ACT_FREE, ACT_POSITION, ACT_SPEED = 0, 1, 2            # Joint.hpp JointActuationType
SPHERE, BOX, CAPSULE, PLANE, HULL = 0, 1, 3, 6, 9

BODY_DTYPE = np.dtype([
    ("type", "<i4"), ("scene", "<u4"), ("mass", "<f8"), ("inertia", "<f8", (6,)),
    ("position", "<f8", (3,)), ("orientation", "<f8", (4,)),
    ("linear_velocity", "<f8", (3,)), ("angular_velocity", "<f8", (3,)),
], align=True)
assert BODY_DTYPE.itemsize == 168

COLLIDER_DTYPE = np.dtype([
    ("shape", "<i4"), ("body_id", "<u4"), ("is_static", "<i4"), ("hull_index", "<i4"),
    ("dims", "<f8", (4,)), ("local_pos", "<f8", (3,)), ("local_rot", "<f8", (4,)),
    ("restitution", "<f8"), ("static_friction", "<f8"), ("dynamic_friction", "<f8"),
], align=True)
assert COLLIDER_DTYPE.itemsize == 128

JOINT_DTYPE = np.dtype([
    ("type", "<i4"), ("actuation", "<i4"), ("body_1", "<u4"), ("body_2", "<u4"),
    ("r_1", "<f8", (3,)), ("r_2", "<f8", (3,)), ("axis", "<f8", (3,)), ("limit_axis", "<f8", (3,)),
    ("compliance", "<f8"), ("damping", "<f8"), ("limited", "<i4"), ("_pad", "<i4"),
    ("lower_limit", "<f8"), ("upper_limit", "<f8"), ("target_position", "<f8"), ("target_speed", "<f8"),
], align=True)
assert JOINT_DTYPE.itemsize == 168

CONSTRAINT_DTYPE = np.dtype([   # rbgpu_constraint_desc: the fields of rbps::ConstraintParams
    ("body_1", "<u4"), ("body_2", "<u4"), ("type", "<i4"), ("_pad", "<i4"),
    ("r_1", "<f8", (3,)), ("r_2", "<f8", (3,)), ("compliance", "<f8"),
], align=True)
assert CONSTRAINT_DTYPE.itemsize == 72


# ---------------------------------------------------------------- unit-density mass properties
# Sphere.hpp:54-78, Box.hpp:58-86, Capsule.hpp:83-122 (density 1).
def sphere_volume(r):
    return (4.0 / 3.0) * math.pi * r * r * r


def sphere_inertia(r):
    i = (2.0 / 5.0) * sphere_volume(r) * r * r
    return (i, i, i, 0.0, 0.0, 0.0)


def box_volume(h):
    return 8.0 * h[0] * h[1] * h[2]


def box_inertia(h):
    x2, y2, z2 = 4.0 * h[0] * h[0], 4.0 * h[1] * h[1], 4.0 * h[2] * h[2]
    m = box_volume(h)
    return ((1.0 / 12.0) * m * (y2 + z2), (1.0 / 12.0) * m * (x2 + z2), (1.0 / 12.0) * m * (x2 + y2), 0.0, 0.0, 0.0)


def capsule_volume(hh, r):
    return math.pi * r * r * (2.0 * hh + (4.0 / 3.0) * r)


def capsule_inertia(hh, r):
    r2, h2 = r * r, hh * hh
    mass = capsule_volume(hh, r)
    m_cyl = math.pi * r2 * 2.0 * hh
    m_hemi = mass - m_cyl
    iyy = (m_cyl * r2 / 2.0) + (m_hemi * 2.0 * r2 / 5.0)
    ixx_cyl = m_cyl * (r2 / 4.0 + h2 / 3.0)
    d_hemi = hh + 3.0 * r / 8.0
    ixx = ixx_cyl + m_hemi * (2.0 * r2 / 5.0 + d_hemi * d_hemi)
    return (ixx, iyy, ixx, 0.0, 0.0, 0.0)


def hull_mass_properties(verts, faces=None):
    """Volume and unit-density inertia (xx yy zz xy xz yz) about the centroid of a hull
    (ConvexHull.hpp:575-745): tetrahedra from the origin per CCW triangle, moved to the
    centroid; a vertex-only hull uses its local AABB as a box."""
    v = np.asarray(verts, np.float64)
    if faces is None or len(faces) == 0:
        e = v.max(0) - v.min(0)
        m = e[0] * e[1] * e[2]
        x2, y2, z2 = e * e
        return m, (m * (y2 + z2) / 12.0, m * (x2 + z2) / 12.0, m * (x2 + y2) / 12.0, 0.0, 0.0, 0.0)
    f = np.asarray(faces, np.int64)
    A, B, Cc = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    det = np.einsum("ij,ij->i", A, np.cross(B, Cc))
    vol = det.sum() / 6.0
    if abs(vol) < 1e-9:
        return 0.0, (0.0,) * 6
    cen = ((A + B + Cc) * (det / 24.0)[:, None]).sum(0) / vol

    def sq(a, b, c):
        return (det / 60.0) * (a * a + b * b + c * c + a * b + a * c + b * c)

    def pr(i, j):
        return (det / 120.0) * (2.0 * (A[:, i] * A[:, j] + B[:, i] * B[:, j] + Cc[:, i] * Cc[:, j])
                                + A[:, i] * B[:, j] + A[:, j] * B[:, i] + A[:, i] * Cc[:, j] + A[:, j] * Cc[:, i]
                                + B[:, i] * Cc[:, j] + B[:, j] * Cc[:, i])
    mxx, myy, mzz = (sq(A[:, k], B[:, k], Cc[:, k]).sum() for k in range(3))
    mxy, mxz, myz = pr(0, 1).sum(), pr(0, 2).sum(), pr(1, 2).sum()
    cx, cy, cz = cen
    return vol, (myy + mzz - vol * (cy * cy + cz * cz), mxx + mzz - vol * (cx * cx + cz * cz),
                 mxx + myy - vol * (cx * cx + cy * cy), -mxy + vol * cx * cy, -mxz + vol * cx * cz,
                 -myz + vol * cy * cz)


def convex_hull_faces(points):
    """Vertices and CCW-outward triangles of the convex hull of a point set (Qhull via
    scipy; vertices not on the hull are dropped)."""
    from scipy.spatial import ConvexHull
    p = np.asarray(points, np.float64)
    h = ConvexHull(p)
    keep = np.unique(h.simplices)
    remap = np.full(len(p), -1, np.int64)
    remap[keep] = np.arange(len(keep))
    v = p[keep]
    f = remap[h.simplices]
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    flip = np.einsum("ij,ij->i", n, h.equations[:, :3]) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f.astype(np.uint32)


def ellipsoid_hull(rng, n_min=16, n_max=64, axis_min=0.2, axis_max=0.8, faces=True):
    """A random convex hull: n ~ U{n_min..n_max} points on an ellipsoid with semi-axes
    U(axis_min, axis_max), with its triangles (or vertex-only)."""
    n = int(rng.integers(n_min, n_max + 1))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pts = d * rng.uniform(axis_min, axis_max, 3)
    if not faces:
        return pts, None
    return convex_hull_faces(pts)


@dataclass
class Scene:
    """One world: bodies, colliders and the World(timestep, substeps) settings."""
    name: str
    bodies: np.ndarray
    colliders: np.ndarray
    timestep: float = 1.0 / 60.0
    substeps: int = 20
    frames: int = 1
    meta: dict = field(default_factory=dict)
    joints: np.ndarray = field(default_factory=lambda: np.zeros(0, JOINT_DTYPE))
    hulls: list = field(default_factory=list)

    @property
    def n_dynamic(self) -> int:
        return int((self.bodies["type"] == DYNAMIC).sum())


class SceneBuilder:
    def __init__(self):
        self.b = []
        self.c = []
        self.j = []
        self.h = []

    def hull(self, verts, faces=None) -> int:
        """Adds hull data (vertices, optional CCW triangles); returns its hull_index."""
        v = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
        f = None if faces is None else np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        self.h.append((v, f))
        return len(self.h) - 1

    def shared_hull(self, data) -> int:
        """Adds hull data from a library (a tuple made by hull_library); `batch` uploads
        data shared by several scenes once."""
        self.h.append(data)
        return len(self.h) - 1

    def body(self, type_=DYNAMIC, mass=1.0, inertia=(1.0, 1.0, 1.0, 0.0, 0.0, 0.0), pos=(0, 0, 0),
             quat=(1, 0, 0, 0), vel=(0, 0, 0), angvel=(0, 0, 0)) -> int:
        r = np.zeros((), BODY_DTYPE)
        r["type"] = type_
        r["mass"] = mass
        r["inertia"] = inertia
        r["position"] = pos
        r["orientation"] = quat
        r["linear_velocity"] = vel
        r["angular_velocity"] = angvel
        self.b.append(r)
        return len(self.b) - 1

    def collider(self, shape, body, dims, is_static=False, local_pos=(0, 0, 0), local_rot=(1, 0, 0, 0),
                 restitution=0.5, static_friction=0.5, dynamic_friction=0.5, hull_index=-1) -> int:
        r = np.zeros((), COLLIDER_DTYPE)
        r["shape"] = shape
        r["body_id"] = body
        r["is_static"] = 1 if is_static else 0
        r["hull_index"] = hull_index
        d = list(dims) + [0.0] * (4 - len(dims))
        r["dims"] = d
        r["local_pos"] = local_pos
        r["local_rot"] = local_rot
        r["restitution"] = restitution
        r["static_friction"] = static_friction
        r["dynamic_friction"] = dynamic_friction
        self.c.append(r)
        return len(self.c) - 1

    def joint(self, type_, body_1, body_2, r_1=(0, 0, 0), r_2=(0, 0, 0), axis=(1, 0, 0), limit_axis=(0, 1, 0),
              actuation=ACT_FREE, compliance=0.0, damping=0.0, limited=False, lower=0.0, upper=0.0,
              target_position=0.0, target_speed=0.0) -> int:
        r = np.zeros((), JOINT_DTYPE)
        r["type"], r["actuation"], r["body_1"], r["body_2"] = type_, actuation, body_1, body_2
        r["r_1"], r["r_2"], r["axis"], r["limit_axis"] = r_1, r_2, axis, limit_axis
        r["compliance"], r["damping"], r["limited"] = compliance, damping, 1 if limited else 0
        r["lower_limit"], r["upper_limit"] = lower, upper
        r["target_position"], r["target_speed"] = target_position, target_speed
        self.j.append(r)
        return len(self.j) - 1

    def build(self, name, timestep=1.0 / 60.0, substeps=20, frames=1, **meta) -> Scene:
        bodies = np.array(self.b, dtype=BODY_DTYPE) if self.b else np.zeros(0, BODY_DTYPE)
        cols = np.array(self.c, dtype=COLLIDER_DTYPE) if self.c else np.zeros(0, COLLIDER_DTYPE)
        joints = np.array(self.j, dtype=JOINT_DTYPE) if self.j else np.zeros(0, JOINT_DTYPE)
        return Scene(name, bodies, cols, timestep, substeps, frames, dict(meta), joints, list(self.h))


def _rand_quat(rng) -> tuple:
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    return tuple(float(x) for x in q)


def add_ground_plane(sb: SceneBuilder, friction=0.5, restitution=0.0) -> int:
    g = sb.body(STATIC, mass=0.0, pos=(0, 0, 0))
    sb.collider(PLANE, g, (0.0, 1.0, 0.0, 0.0), is_static=True, restitution=restitution,
                static_friction=friction, dynamic_friction=friction)
    return g


def pile_small(seed=12345, nx=10, ny=10, nz=10, spacing=1.2, y0=1.0, jitter=0.05, density=1000.0,
               friction=0.5, restitution=0.0, substeps=20, frames=300) -> Scene:
    """1,000 unit boxes (half 0.5) and spheres (r 0.5) on a jittered grid over a plane.

    Index parity selects box or sphere; orientations are normalised N(0,1)^4.
    """
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    add_ground_plane(sb, friction, restitution)
    idx = 0
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                p = np.array([(i - (nx - 1) / 2) * spacing, y0 + j * spacing, (k - (nz - 1) / 2) * spacing])
                p = p + rng.uniform(-jitter, jitter, 3)
                q = _rand_quat(rng)
                if idx % 2 == 0:
                    h = (0.5, 0.5, 0.5)
                    m = density * box_volume(h)
                    inert = tuple(density * x for x in box_inertia(h))
                    b = sb.body(DYNAMIC, m, inert, tuple(p), q)
                    sb.collider(BOX, b, h, restitution=restitution, static_friction=friction, dynamic_friction=friction)
                else:
                    r = 0.5
                    m = density * sphere_volume(r)
                    inert = tuple(density * x for x in sphere_inertia(r))
                    b = sb.body(DYNAMIC, m, inert, tuple(p), q)
                    sb.collider(SPHERE, b, (r,), restitution=restitution, static_friction=friction,
                                dynamic_friction=friction)
                idx += 1
    return sb.build(f"pile_small_{nx}x{ny}x{nz}", 1.0 / 60.0, substeps, frames, seed=seed)


def resting_layer(seed=300, n=1000, cols=32, spacing=1.6, drop=(0.02, 0.25), speed=0.2, density=1000.0,
                  friction=0.5, restitution=0.0, substeps=20, frames=300) -> Scene:
    """The small plain configuration's bodies (unit boxes, half 0.5, and spheres, r 0.5,
    by index parity; density 1000; mu 0.5; e 0; dt 1/60; 20 substeps; 300 frames) in a
    well-posed arrangement: one layer on a `cols`-wide grid (spacing 1.6 m, so no two
    bodies can touch).  Boxes start flat on the plane (turned about the vertical axis
    only) and slide with a horizontal velocity of at most `speed` m/s until friction
    stops them within millimetres; spheres drop from 0.02-0.25 m above their resting
    height with no horizontal velocity.  A box dropped flat would be ill-posed (rounding
    decides which corner lands first) and a sphere rolling sideways would never stop, so
    neither is used.  No contact pair shares a dynamic body, so any solve order gives the
    same result."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    add_ground_plane(sb, friction, restitution)
    rows = (n + cols - 1) // cols
    for k in range(n):
        i, j = k % cols, k // cols
        x0, z0 = (i - (cols - 1) / 2) * spacing, (j - (rows - 1) / 2) * spacing
        drop_k = float(rng.uniform(*drop))
        yaw = float(rng.uniform(-math.pi, math.pi))
        vx, vz = (float(x) for x in rng.uniform(-speed, speed, 2))
        if k % 2 == 0:
            h = (0.5, 0.5, 0.5)
            q = (math.cos(yaw / 2), 0.0, math.sin(yaw / 2), 0.0)
            b = sb.body(DYNAMIC, density * box_volume(h), tuple(density * x for x in box_inertia(h)), (x0, 0.5, z0), q,
                        (vx, 0.0, vz))
            sb.collider(BOX, b, h, restitution=restitution, static_friction=friction, dynamic_friction=friction)
        else:
            r = 0.5
            b = sb.body(DYNAMIC, density * sphere_volume(r), tuple(density * x for x in sphere_inertia(r)),
                        (x0, r + drop_k, z0))
            sb.collider(SPHERE, b, (r,), restitution=restitution, static_friction=friction, dynamic_friction=friction)
    return sb.build(f"resting_layer_{n}", 1.0 / 60.0, substeps, frames, seed=seed)


def accuracy_batch(n_scenes=50, per_scene=20, cols=5, seed=400) -> Scene:
    """The well-posed accuracy workload: 1,000 bodies of the small plain configuration as
    50 independent resting layers of 20 bodies (5 x 4 grid each, so every coordinate stays
    within a few metres of its scene's origin), one batched world."""
    return batch([resting_layer(seed=seed + s, n=per_scene, cols=cols) for s in range(n_scenes)],
                 f"resting_layers_{n_scenes}x{per_scene}")


def pile_mixed(seed=2024, nx=64, ny=16, nz=64, spacing=2.2, y0=1.5, jitter=0.05, density=1000.0,
               friction=0.5, restitution=0.0, substeps=10, frames=500) -> Scene:
    """Mixed convex pile: boxes (half extents U(0.2,0.6)), spheres (r U(0.2,0.5)) and
    capsules (r U(0.1,0.3), half-length U(0.2,0.6)), type = index mod 3, on a seeded
    grid (no initial overlap at spacing 2.2) over a static plane."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    add_ground_plane(sb, friction, restitution)
    idx = 0
    for j in range(ny):
        for i in range(nx):
            for k in range(nz):
                p = np.array([(i - (nx - 1) / 2) * spacing, y0 + j * spacing, (k - (nz - 1) / 2) * spacing])
                p = p + rng.uniform(-jitter, jitter, 3)
                q = _rand_quat(rng)
                t = idx % 3
                if t == 0:
                    h = tuple(float(x) for x in rng.uniform(0.2, 0.6, 3))
                    m = density * box_volume(h)
                    inert = tuple(density * x for x in box_inertia(h))
                    b = sb.body(DYNAMIC, m, inert, tuple(p), q)
                    sb.collider(BOX, b, h, restitution=restitution, static_friction=friction, dynamic_friction=friction)
                elif t == 1:
                    r = float(rng.uniform(0.2, 0.5))
                    m = density * sphere_volume(r)
                    inert = tuple(density * x for x in sphere_inertia(r))
                    b = sb.body(DYNAMIC, m, inert, tuple(p), q)
                    sb.collider(SPHERE, b, (r,), restitution=restitution, static_friction=friction,
                                dynamic_friction=friction)
                else:
                    r = float(rng.uniform(0.1, 0.3))
                    hh = float(rng.uniform(0.2, 0.6))
                    m = density * capsule_volume(hh, r)
                    inert = tuple(density * x for x in capsule_inertia(hh, r))
                    b = sb.body(DYNAMIC, m, inert, tuple(p), q)
                    sb.collider(CAPSULE, b, (hh, r), restitution=restitution, static_friction=friction,
                                dynamic_friction=friction)
                idx += 1
    return sb.build(f"pile_mixed_{nx}x{ny}x{nz}", 1.0 / 60.0, substeps, frames, seed=seed)


def pile_mixed_dense(seed=7, nx=8, ny=6, nz=8, spacing=1.05, y0=0.7, **kw) -> Scene:
    """A small, initially interpenetrating mixed pile: every pair type and many
    contacts from the first frame (parity stress case)."""
    return pile_mixed(seed=seed, nx=nx, ny=ny, nz=nz, spacing=spacing, y0=y0, **kw)


# ---------------------------------------------------------------- jointed scenes
def _box(sb, pos, half, density=1000.0, quat=(1, 0, 0, 0), vel=(0, 0, 0), type_=DYNAMIC, friction=0.5):
    m = density * box_volume(half)
    inertia = tuple(density * x for x in box_inertia(half))
    b = sb.body(type_, m if type_ == DYNAMIC else 0.0, inertia, pos, quat, vel)
    sb.collider(BOX, b, half, is_static=type_ == STATIC, restitution=0.0, static_friction=friction,
                dynamic_friction=friction)
    return b


def _anchor(sb, pos):
    """A static body without a collider (a joint anchor nothing collides with)."""
    return sb.body(STATIC, 0.0, (1.0, 1.0, 1.0, 0.0, 0.0, 0.0), pos)


def _capsule(sb, pos, hh, r, density=1000.0, quat=(1, 0, 0, 0), friction=0.5):
    m = density * capsule_volume(hh, r)
    inertia = tuple(density * x for x in capsule_inertia(hh, r))
    b = sb.body(DYNAMIC, m, inertia, pos, quat)
    sb.collider(CAPSULE, b, (hh, r), restitution=0.0, static_friction=friction, dynamic_friction=friction)
    return b


def _chain(sb, anchor, anchor_pos, n, hh, r, compliance, limited_every=0, swing=math.radians(45)):
    """n capsules hanging sideways from `anchor` (capsule axis = local y, laid along +x),
    linked end to end by spherical joints half a radius beyond the capsule tips (the
    reference does not filter collisions between jointed bodies, so links must not touch
    when straight)."""
    qz = (math.cos(math.pi / 4), 0.0, 0.0, -math.sin(math.pi / 4))   # local y -> world +x
    tip = hh + 1.5 * r
    link = 2.0 * tip
    prev, prev_tip = anchor, (0.0, 0.0, 0.0)
    x0, y0, z0 = anchor_pos
    for k in range(n):
        c = _capsule(sb, (x0 + link * (k + 0.5), y0, z0), hh, r, quat=qz)
        lim = limited_every and (k % limited_every == limited_every - 1)
        sb.joint(SPHERICAL, prev, c, r_1=prev_tip, r_2=(0.0, -tip, 0.0), axis=(0, 1, 0),
                 compliance=compliance, limited=bool(lim), upper=swing)
        prev, prev_tip = c, (0.0, tip, 0.0)
    return prev


def joint_zoo(seed=31, spherical=True) -> Scene:
    """Every joint type and feature next to contacts: limited and driven revolutes,
    a limited and driven prismatic, a fixed pair falling on the plane, damping, and
    (with `spherical`) a capsule chain on spherical joints with one swing limit."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    add_ground_plane(sb)
    anchor = _box(sb, (0.0, 4.0, 0.0), (0.2, 0.2, 0.2), type_=STATIC)
    # pendulum on a limited, damped revolute (hinge about z), released sideways
    p = _box(sb, (1.0, 4.0, 0.0), (0.4, 0.1, 0.1))
    sb.joint(REVOLUTE, anchor, p, r_1=(0.0, 0.0, 0.0), r_2=(-1.0, 0.0, 0.0), axis=(0, 0, 1), limit_axis=(1, 0, 0),
             damping=0.2, limited=True, lower=-math.radians(60), upper=math.radians(60))
    # motor: speed-driven revolute about y with a compliant drive
    a2 = _box(sb, (3.0, 3.0, 0.0), (0.2, 0.2, 0.2), type_=STATIC)
    m = _box(sb, (3.0, 2.4, 0.0), (0.5, 0.05, 0.1))
    sb.joint(REVOLUTE, a2, m, r_1=(0.0, -0.3, 0.0), r_2=(0.0, 0.3, 0.0), axis=(0, 1, 0), limit_axis=(1, 0, 0),
             actuation=ACT_SPEED, target_speed=2.0, compliance=1e-4)
    # position-driven revolute with a target the limit cuts off
    a3 = _box(sb, (5.0, 3.0, 0.0), (0.2, 0.2, 0.2), type_=STATIC)
    m3 = _box(sb, (5.8, 3.0, 0.0), (0.5, 0.1, 0.1))
    sb.joint(REVOLUTE, a3, m3, r_2=(-0.8, 0.0, 0.0), axis=(0, 0, 1), limit_axis=(1, 0, 0), actuation=ACT_POSITION,
             target_position=1.2, limited=True, lower=-0.5, upper=0.8, damping=0.5)
    # slider: limited prismatic, position-driven, damped, between two dynamic boxes
    s1 = _box(sb, (-3.0, 1.0, 0.0), (0.3, 0.3, 0.3))
    s2 = _box(sb, (-2.0, 1.0, 0.0), (0.2, 0.2, 0.2))
    sb.joint(PRISMATIC, s1, s2, axis=(1, 0, 0), actuation=ACT_POSITION, target_position=0.4, compliance=1e-5,
             damping=0.5, limited=True, lower=-1.5, upper=-0.3)
    # speed-driven prismatic on a static rail
    rail = _box(sb, (-6.0, 2.0, 0.0), (0.1, 0.1, 0.1), type_=STATIC)
    car = _box(sb, (-6.0, 1.5, 0.0), (0.2, 0.2, 0.2))
    sb.joint(PRISMATIC, rail, car, axis=(0, 1, 0), actuation=ACT_SPEED, target_speed=-0.5, damping=0.1)
    # fixed pair tumbling onto the plane
    f1 = _box(sb, (0.0, 1.5, 3.0), (0.3, 0.3, 0.3), quat=_rand_quat(rng))
    f2 = _box(sb, (0.5, 1.9, 3.0), (0.2, 0.2, 0.2))
    sb.joint(FIXED, f1, f2, r_1=(0.25, 0.2, 0.0), r_2=(-0.25, -0.2, 0.0))
    # loose spheres for contacts with the jointed bodies
    for k in range(6):
        rr = 0.2 + 0.1 * rng.random()
        pos = (-1.0 + 0.7 * k, 2.0 + rng.random(), 3.0 + 0.2 * rng.standard_normal())
        b = sb.body(DYNAMIC, 1000.0 * sphere_volume(rr), tuple(1000.0 * x for x in sphere_inertia(rr)), pos)
        sb.collider(SPHERE, b, (rr,), restitution=0.0)
    if spherical:
        a4 = _anchor(sb, (0.0, 3.0, -3.0))
        _chain(sb, a4, (0.0, 3.0, -3.0), 6, 0.2, 0.08, 1e-6, limited_every=3)
    return sb.build(f"joint_zoo{'_sph' if spherical else ''}", 1.0 / 60.0, 20, 120, seed=seed)


def stack_chain(n_bodies=64, seed=0, origin=(0.0, 0.0, 0.0), chain_links=16, density=1000.0) -> Scene:
    """One batched-workload scene: a seeded box stack of (n_bodies - chain_links) boxes
    plus a chain of `chain_links` capsules on spherical joints (compliance 1e-6) hanging
    from a static anchor, over a static plane."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    add_ground_plane(sb)
    ox, oy, oz = origin
    nb = max(n_bodies - chain_links, 0)
    cols = max(1, int(math.ceil(math.sqrt(nb / 6.0))))
    for k in range(nb):
        c, lvl = k % (cols * cols), k // (cols * cols)
        h = 0.25 + 0.05 * rng.random()
        x = ox + (c % cols) * 0.7 + rng.uniform(-0.02, 0.02)
        z = oz + (c // cols) * 0.7 + rng.uniform(-0.02, 0.02)
        _box(sb, (x, oy + 0.3 + lvl * 0.62, z), (h, h, h), density)
    a = _anchor(sb, (ox - 1.5, oy + 6.0, oz))
    _chain(sb, a, (ox - 1.5, oy + 6.0, oz), chain_links, 0.15, 0.06, 1e-6)
    return sb.build(f"stack_chain_{n_bodies}", 1.0 / 60.0, 10, 100, seed=seed)


def planar_joints(seed=5) -> Scene:
    """Joints whose motion stays in the xy plane (hinges about z, sliders along x or y with
    attachments at the body centres): the reference's rotational rows then keep exactly
    zero out-of-plane errors and its step stays finite (see joint_zoo for the general case)."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    add_ground_plane(sb)
    piv = [_box(sb, (4.0 * k, 5.0, 0.0), (0.1, 0.1, 0.1), type_=STATIC) for k in range(4)]
    # free pendulum, limited + damped pendulum, speed motor, position servo (hinges about z)
    bobs = [_box(sb, (4.0 * k + 1.2, 5.0 - 0.3 * k, 0.0), (0.35, 0.1, 0.1)) for k in range(4)]
    for k in range(4):
        sb.joint(REVOLUTE, piv[k], bobs[k], r_2=(-1.2 if k != 1 else -1.2, 0.3 * k, 0.0), axis=(0, 0, 1),
                 limit_axis=(1, 0, 0), damping=(0.3 if k == 1 else 0.0), limited=(k == 1),
                 lower=-math.radians(50), upper=math.radians(50),
                 actuation=(ACT_SPEED if k == 2 else ACT_POSITION if k == 3 else ACT_FREE),
                 target_speed=(1.5 if k == 2 else 0.0), target_position=(0.6 if k == 3 else 0.0),
                 compliance=(1e-3 if k >= 2 else 0.0))
    # a double pendulum in the plane
    d1 = _box(sb, (17.0, 5.0, 0.0), (0.5, 0.08, 0.08))
    d2 = _box(sb, (18.0, 5.0, 0.0), (0.5, 0.08, 0.08))
    p5 = _box(sb, (16.5, 5.0, 0.0), (0.1, 0.1, 0.1), type_=STATIC)
    sb.joint(REVOLUTE, p5, d1, r_2=(-0.5, 0.0, 0.0), axis=(0, 0, 1), limit_axis=(1, 0, 0))
    sb.joint(REVOLUTE, d1, d2, r_1=(0.5, 0.0, 0.0), r_2=(-0.5, 0.0, 0.0), axis=(0, 0, 1), limit_axis=(1, 0, 0))
    # sliders: a limited, position-driven, damped one between two free boxes on the plane, and
    # a speed-driven vertical one on a static rail
    s1 = _box(sb, (-3.0, 0.3, 0.0), (0.3, 0.3, 0.3))
    s2 = _box(sb, (-2.0, 0.2, 0.0), (0.2, 0.2, 0.2))
    sb.joint(PRISMATIC, s1, s2, axis=(1, 0, 0), actuation=ACT_POSITION, target_position=0.5, compliance=1e-5,
             damping=0.5, limited=True, lower=-1.5, upper=-0.3)
    rail = _box(sb, (-6.0, 3.0, 0.0), (0.1, 0.1, 0.1), type_=STATIC)
    car = _box(sb, (-6.0, 2.0, 0.0), (0.2, 0.2, 0.2))
    sb.joint(PRISMATIC, rail, car, axis=(0, 1, 0), actuation=ACT_SPEED, target_speed=-0.5, damping=0.1,
             limited=True, lower=-2.5, upper=2.5)
    # spheres dropped on the slider pair (contacts with jointed bodies, still planar)
    for k in range(3):
        rr = 0.2 + 0.05 * k
        b = sb.body(DYNAMIC, 1000.0 * sphere_volume(rr), tuple(1000.0 * x for x in sphere_inertia(rr)),
                    (-3.2 + 0.6 * k, 1.5 + 0.5 * k + 0.1 * rng.random(), 0.0))
        sb.collider(SPHERE, b, (rr,), restitution=0.0)
    return sb.build("planar_joints", 1.0 / 60.0, 20, 120, seed=seed)


def constraint_rows(seed=11) -> Scene:
    """Constraint rows made by World::create_constraint (ConstraintAPI.hpp:37-64) between
    jointed bodies, created between the joints (the reference solves all rows in creation
    order), planar like planar_joints: two hinged pendulums about z, a POSITIONAL row along
    x between the bobs, a ROTATIONAL row about z between them and a zero row from a box on
    the plane to a static anchor; spheres land on that box.  meta["rows"]: (index in
    `joints`, direction, magnitude) that the caller writes after loading, as reference
    callers write ConstraintCollection::direction / magnitude."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    add_ground_plane(sb)
    p1 = _box(sb, (0.0, 5.0, 0.0), (0.1, 0.1, 0.1), type_=STATIC)
    b1 = _box(sb, (1.0, 5.0, 0.0), (0.4, 0.1, 0.1))
    p2 = _box(sb, (3.0, 5.0, 0.0), (0.1, 0.1, 0.1), type_=STATIC)
    b2 = _box(sb, (4.0, 5.0, 0.0), (0.4, 0.1, 0.1))
    sb.joint(REVOLUTE, p1, b1, r_2=(-1.0, 0.0, 0.0), axis=(0, 0, 1), limit_axis=(1, 0, 0))
    rows = []
    rows.append((len(sb.j), (1.0, 0.0, 0.0), 0.01))
    sb.joint(ROW, b1, b2, r_1=(0.4, 0.0, 0.0), r_2=(-0.4, 0.0, 0.0), actuation=ROW_POSITIONAL, compliance=1e-4)
    sb.joint(REVOLUTE, p2, b2, r_2=(-1.0, 0.0, 0.0), axis=(0, 0, 1), limit_axis=(1, 0, 0), compliance=0.0)
    rows.append((len(sb.j), (0.0, 0.0, 1.0), 0.002))
    sb.joint(ROW, b1, b2, actuation=ROW_ROTATIONAL, compliance=1e-3)
    base = _box(sb, (-3.0, 0.3, 0.0), (0.6, 0.3, 0.3))
    anchor = _box(sb, (-3.0, 3.0, 0.0), (0.1, 0.1, 0.1), type_=STATIC)
    sb.joint(ROW, base, anchor, r_1=(0.0, 0.3, 0.0), actuation=ROW_POSITIONAL)   # direction (1,0,0), magnitude 0
    for k in range(3):
        rr = 0.15 + 0.05 * k
        b = sb.body(DYNAMIC, 1000.0 * sphere_volume(rr), tuple(1000.0 * x for x in sphere_inertia(rr)),
                    (-3.3 + 0.3 * k, 1.2 + 0.5 * k + 0.1 * rng.random(), 0.0))
        sb.collider(SPHERE, b, (rr,), restitution=0.0)
    return sb.build("constraint_rows", 1.0 / 60.0, 20, 60, seed=seed, rows=rows)


def batch(parts, name="batch") -> Scene:
    """Independent scenes as one batched world: every body carries its scene index;
    colliders and joints are re-indexed.  Scenes may overlap in space, they never
    interact.  Timestep and substeps come from the first scene."""
    bodies, cols, joints, hulls = [], [], [], []
    hull_ids = {}   # hull data shared between scenes (the same arrays) is uploaded once
    nb = 0
    for s, sc in enumerate(parts):
        b = sc.bodies.copy()
        b["scene"] = s
        c = sc.colliders.copy()
        c["body_id"] += nb
        remap = np.zeros(len(sc.hulls), np.int32)
        for k, h in enumerate(sc.hulls):
            key = (id(h[0]), id(h[1]))
            if key not in hull_ids:
                hull_ids[key] = len(hulls)
                hulls.append(h)
            remap[k] = hull_ids[key]
        sel = c["shape"] == HULL
        c["hull_index"][sel] = remap[c["hull_index"][sel]]
        j = sc.joints.copy()
        j["body_1"] += nb
        j["body_2"] += nb
        bodies.append(b)
        cols.append(c)
        joints.append(j)
        nb += len(b)
    first = parts[0]
    return Scene(name, np.concatenate(bodies), np.concatenate(cols), first.timestep, first.substeps, first.frames,
                 {"scenes": len(parts), "bodies_per_scene": [len(p.bodies) for p in parts]},
                 np.concatenate(joints) if joints else np.zeros(0, JOINT_DTYPE), hulls)


def split(sc: Scene, only=None) -> list:
    """The scenes of a batch (see `batch`) as worlds of their own, in scene order (or the
    scene indices in `only`): bodies, colliders and joints re-indexed, the batch's hull
    list kept whole."""
    out = []
    scene = sc.bodies["scene"]
    body_of = sc.colliders["body_id"].astype(np.int64)
    for s in (range(int(scene.max()) + 1 if len(scene) else 0) if only is None else only):
        sel = np.nonzero(scene == s)[0]
        if not len(sel):
            continue
        b0, b1 = int(sel[0]), int(sel[-1]) + 1   # a batch keeps each scene's bodies contiguous
        b = sc.bodies[b0:b1].copy()
        b["scene"] = 0
        c = sc.colliders[(body_of >= b0) & (body_of < b1)].copy()
        c["body_id"] -= np.uint32(b0)
        j = sc.joints[(sc.joints["body_1"] >= b0) & (sc.joints["body_1"] < b1)].copy() if len(sc.joints) else sc.joints
        if len(j):
            j["body_1"] -= np.uint32(b0)
            j["body_2"] -= np.uint32(b0)
        out.append(Scene(f"{sc.name}_scene{s}", b, c, sc.timestep, sc.substeps, sc.frames, {"scene": s}, j,
                         list(sc.hulls)))
    return out


def batched_stack_chains(n_scenes=4096, seed=77, n_min=32, n_max=512) -> Scene:
    """The batched workload: n_scenes independent scenes with body counts uniform in
    [n_min, n_max], each a seeded box stack plus a 16-link capsule chain on spherical
    joints (compliance 1e-6) over its own ground plane."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(n_min, n_max + 1, n_scenes)
    parts = [stack_chain(int(n), seed=seed + 1 + s) for s, n in enumerate(counts)]
    return batch(parts, f"batched_stack_chains_{n_scenes}")


# ---------------------------------------------------------------- hulls
def _hull_body(sb, pos, hidx, density=1000.0, quat=(1, 0, 0, 0), friction=0.5, restitution=0.0, vel=(0, 0, 0)):
    v, f = sb.h[hidx]
    vol, inert = hull_mass_properties(v, f)
    b = sb.body(DYNAMIC, density * vol, tuple(density * x for x in inert), tuple(pos), quat, vel)
    sb.collider(HULL, b, (), hull_index=hidx, restitution=restitution, static_friction=friction,
                dynamic_friction=friction)
    return b


def tetrahedron_hull():
    """The demo tetrahedron (src/visr_demo/main.cpp:12-30): 4 vertices, 4 CCW triangles."""
    v = np.array([[0.6, 0.6, 0.6], [-0.6, -0.6, 0.6], [-0.6, 0.6, -0.6], [0.6, -0.6, -0.6]])
    f = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.uint32)
    return v, f


def hex_prism_hull(radius=0.8, half_height=0.5):
    """A hexagonal prism whose caps are fans of four triangles (merged into hexagons by
    the coplanar merge) and whose sides are quads of two triangles."""
    ang = np.arange(6) * (2.0 * math.pi / 6.0)
    top = np.stack([np.cos(ang) * radius, np.full(6, half_height), np.sin(ang) * radius], 1)
    bot = top * np.array([1.0, -1.0, 1.0])
    v = np.concatenate([top, bot])
    f = []
    for i in range(1, 5):
        f.append((0, i + 1, i))
        f.append((6, i + 6, i + 7))
    for i in range(6):
        j = (i + 1) % 6
        f.append((i, j, i + 6))
        f.append((j, j + 6, i + 6))
    return v, np.array(f, np.uint32)


def box_hull(half=(0.5, 0.5, 0.5)):
    """A box as a hull of 12 triangles (merged into 6 quads)."""
    pts = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float64) * np.asarray(half)
    return convex_hull_faces(pts)


def hull_pile(n=32768, seed=4, faces=True, spacing=1.7, y0=1.0, jitter=0.05, density=1000.0, friction=0.5,
              substeps=10, frames=500) -> Scene:
    """The GJK/EPA-heavy workload: n random hulls (16-64 vertices on ellipsoids with
    semi-axes U(0.2, 0.8)) on a seeded, jittered, square-based grid pile over a plane,
    random orientations, each hull its own data."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    add_ground_plane(sb, friction, 0.0)
    side = max(1, int(round(math.sqrt(n / 8.0))))
    layers = (n + side * side - 1) // (side * side)
    k = 0
    for j in range(layers):
        for i in range(side):
            for l in range(side):
                if k == n:
                    break
                v, f = ellipsoid_hull(rng, faces=faces)
                h = sb.hull(v, f)
                p = np.array([(i - (side - 1) / 2) * spacing, y0 + j * spacing, (l - (side - 1) / 2) * spacing])
                p = p + rng.uniform(-jitter, jitter, 3)
                _hull_body(sb, p, h, density, _rand_quat(rng), friction)
                k += 1
    return sb.build(f"hull_pile_{n}{'' if faces else '_vertex_only'}", 1.0 / 60.0, substeps, frames, seed=seed)


def hull_zoo(seed=9) -> Scene:
    """Hulls (tetrahedra, hexagonal prisms, box hulls, random ellipsoid hulls with and
    without faces) dropped among boxes, spheres and capsules on a plane: every hull pair
    kind (hull-hull manifold, hull-hull single point, hull vs primitive, hull-plane)."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    add_ground_plane(sb, 0.5, 0.0)
    hs = [sb.hull(*tetrahedron_hull()), sb.hull(*hex_prism_hull()), sb.hull(*box_hull((0.4, 0.3, 0.5)))]
    hs += [sb.hull(*ellipsoid_hull(rng, faces=True)) for _ in range(3)]
    hs += [sb.hull(*ellipsoid_hull(rng, faces=False)) for _ in range(2)]
    k = 0
    for j in range(4):
        for i in range(4):
            for l in range(3):
                p = np.array([(i - 1.5) * 1.3, 0.9 + j * 1.25, (l - 1.0) * 1.3]) + rng.uniform(-0.05, 0.05, 3)
                q = _rand_quat(rng)
                kind = k % 4
                if kind == 3:
                    which = int(rng.integers(0, 3))
                    if which == 0:
                        _box(sb, tuple(p), (0.35, 0.3, 0.4), quat=q)
                    elif which == 1:
                        r = 0.4
                        b = sb.body(DYNAMIC, 1000.0 * sphere_volume(r), tuple(1000.0 * x for x in sphere_inertia(r)),
                                    tuple(p), q)
                        sb.collider(SPHERE, b, (r,), restitution=0.0)
                    else:
                        _capsule(sb, tuple(p), 0.35, 0.2, quat=q)
                else:
                    _hull_body(sb, p, hs[int(rng.integers(0, len(hs)))], quat=q)
                k += 1
    return sb.build("hull_zoo", 1.0 / 60.0, 10, 1, seed=seed)


# ---------------------------------------------------------------- combined (ragdolls + hulls + boxes)
def hull_library(n=16, seed=101, axis_min=0.15, axis_max=0.45):
    """n ellipsoid hulls (16-64 vertices, Qhull triangles) shared by the scenes of a batch."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        v, f = ellipsoid_hull(rng, axis_min=axis_min, axis_max=axis_max, faces=True)
        out.append((np.ascontiguousarray(v, np.float64), np.ascontiguousarray(f, np.uint32)))
    return out


RAGDOLL_LIMIT = math.radians(60.0)


def ragdoll(sb, origin, yaw=0.0, density=1000.0, compliance=1e-6, friction=0.5) -> list:
    """An 8-segment ragdoll standing at `origin` (pelvis centre), turned by `yaw` about y:
    pelvis, abdomen and chest boxes, a sphere head, capsule arms held out sideways and
    capsule legs.  Joints: pelvis-abdomen, abdomen-chest and the hips are revolute about the
    body's x axis with limits of +-60 deg; neck and shoulders are spherical with a 60 deg
    swing limit.  Segments are spaced so that no two of one ragdoll touch anywhere inside
    the limits (stacked boxes of half-depth d stay clear while the gap exceeds 2 d tan 30
    deg): the reference does not filter contacts between jointed bodies, and a contact
    fighting a joint limit pumps energy into the pair."""
    cy, sy = math.cos(0.5 * yaw), math.sin(0.5 * yaw)
    q = (cy, 0.0, sy, 0.0)
    ox, oy, oz = origin

    def at(x, y, z):   # ragdoll-frame offset -> world
        c, s = math.cos(yaw), math.sin(yaw)
        return (ox + c * x + s * z, oy + y, oz - s * x + c * z)

    pelvis = _box(sb, at(0.0, 0.0, 0.0), (0.15, 0.08, 0.1), density, quat=q, friction=friction)
    belly = _box(sb, at(0.0, 0.30, 0.0), (0.13, 0.08, 0.09), density, quat=q, friction=friction)
    chest = _box(sb, at(0.0, 0.62, 0.0), (0.17, 0.1, 0.1), density, quat=q, friction=friction)
    rh = 0.1
    head = sb.body(DYNAMIC, density * sphere_volume(rh), tuple(density * x for x in sphere_inertia(rh)), at(0.0, 0.92, 0.0), q)
    sb.collider(SPHERE, head, (rh,), restitution=0.0, static_friction=friction, dynamic_friction=friction)
    # arms: bodies aligned with the chest, capsule colliders turned from y onto x
    ahh, ar = 0.18, 0.05
    ixx, iyy, _, _, _, _ = capsule_inertia(ahh, ar)
    along_x = (math.cos(math.pi / 4), 0.0, 0.0, -math.sin(math.pi / 4))
    arms = []
    for sx in (-1.0, 1.0):
        a = sb.body(DYNAMIC, density * capsule_volume(ahh, ar), (density * iyy, density * ixx, density * ixx, 0.0, 0.0, 0.0),
                    at(sx * 0.48, 0.66, 0.0), q)
        sb.collider(CAPSULE, a, (ahh, ar), local_rot=along_x, restitution=0.0, static_friction=friction,
                    dynamic_friction=friction)
        arms.append(a)
    legs = [_capsule(sb, at(sx * 0.09, -0.42, 0.0), 0.22, 0.06, density, quat=q, friction=friction) for sx in (-1.0, 1.0)]
    lim = RAGDOLL_LIMIT
    hinge = dict(axis=(1, 0, 0), limit_axis=(0, 1, 0), compliance=compliance, limited=True, lower=-lim, upper=lim)
    sb.joint(REVOLUTE, pelvis, belly, r_1=(0.0, 0.15, 0.0), r_2=(0.0, -0.15, 0.0), **hinge)
    sb.joint(REVOLUTE, belly, chest, r_1=(0.0, 0.15, 0.0), r_2=(0.0, -0.17, 0.0), **hinge)
    sb.joint(SPHERICAL, chest, head, r_1=(0.0, 0.13, 0.0), r_2=(0.0, -0.17, 0.0), axis=(0, 1, 0),
             compliance=compliance, limited=True, upper=lim)
    for sx, a in zip((-1.0, 1.0), arms):
        sb.joint(SPHERICAL, chest, a, r_1=(sx * 0.20, 0.04, 0.0), r_2=(-sx * 0.28, 0.0, 0.0), axis=(sx, 0, 0),
                 compliance=compliance, limited=True, upper=lim)
    for sx, l in zip((-1.0, 1.0), legs):
        sb.joint(REVOLUTE, pelvis, l, r_1=(sx * 0.09, -0.11, 0.0), r_2=(0.0, 0.31, 0.0), **hinge)
    return [pelvis, belly, chest, head] + arms + legs


def combined_scene(n_bodies=256, seed=0, library=None, ragdoll_share=0.375, spacing=1.35, y0=2.8, density=1000.0,
                   frames=1000, ragdoll_spacing=1.6) -> Scene:
    """One scene of the combined workload: a ground plane, ragdolls standing on a grid
    (about `ragdoll_share` of the bodies, 8 per ragdoll), and above them a pile of convex
    hulls (from the shared `library`) and boxes (half-extents U(0.15, 0.35)), half each,
    that falls onto the ragdolls.  dt = 1/120 s, 8 substeps."""
    rng = np.random.default_rng(seed)
    lib = library if library is not None else hull_library()
    sb = SceneBuilder()
    add_ground_plane(sb)
    hidx = [sb.shared_hull(h) for h in lib]
    n_rag = int(round(n_bodies * ragdoll_share / 8.0))
    n_rest = n_bodies - 8 * n_rag
    side = max(1, int(math.ceil(math.sqrt(max(n_rag, 1)))))
    for k in range(n_rag):
        x = (k % side - (side - 1) / 2) * ragdoll_spacing + rng.uniform(-0.05, 0.05)
        z = (k // side - (side - 1) / 2) * ragdoll_spacing + rng.uniform(-0.05, 0.05)
        ragdoll(sb, (x, 0.95, z), yaw=float(rng.uniform(-math.pi, math.pi)), density=density)
    cols = max(side, int(math.ceil(math.sqrt(n_rest / 6.0))))
    for k in range(n_rest):
        c, lvl = k % (cols * cols), k // (cols * cols)
        p = ((c % cols - (cols - 1) / 2) * spacing + rng.uniform(-0.05, 0.05), y0 + lvl * spacing,
             (c // cols - (cols - 1) / 2) * spacing + rng.uniform(-0.05, 0.05))
        if k % 2 == 0:
            _hull_body(sb, p, hidx[int(rng.integers(0, len(hidx)))], density, _rand_quat(rng))
        else:
            _box(sb, p, tuple(rng.uniform(0.15, 0.35, 3)), density, quat=_rand_quat(rng))
    return sb.build(f"combined_{n_bodies}", 1.0 / 120.0, 8, frames, seed=seed)


def combined_counts(n_scenes=1024, total=262144, n_min=64, n_max=512, seed=2026):
    """Uneven body counts: uniform in [n_min, n_max], then scaled about n_min and
    adjusted by single bodies so that they sum to `total` (every count stays in range)."""
    rng = np.random.default_rng(seed)
    u = rng.integers(n_min, n_max + 1, n_scenes).astype(np.float64)
    f = (total - n_min * n_scenes) / max((u - n_min).sum(), 1.0)
    c = np.clip(np.round(n_min + (u - n_min) * f), n_min, n_max).astype(np.int64)
    while c.sum() != total:
        i = int(rng.integers(0, n_scenes))
        d = 1 if c.sum() < total else -1
        if n_min <= c[i] + d <= n_max:
            c[i] += d
    return c


def combined_batch(n_scenes=1024, total=262144, seed=2026, n_min=64, n_max=512, frames=1000) -> Scene:
    """The combined workload: n_scenes independent scenes of uneven size (summing to
    `total` dynamic bodies) mixing convex hulls, boxes and 8-segment ragdolls, one world."""
    lib = hull_library()
    counts = combined_counts(n_scenes, total, n_min, n_max, seed)
    parts = [combined_scene(int(n), seed=seed + 1 + s, library=lib, frames=frames) for s, n in enumerate(counts)]
    return batch(parts, f"combined_{n_scenes}_scenes_{total}")
