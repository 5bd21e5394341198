"""Scene descriptors and seeded scene generators.

The descriptor dtypes mirror the C structs ``rbgpu_body_desc`` and
``rbgpu_collider_desc`` declared in ``include/rbgpu.h`` (the fields of the
reference's ``rbps::BodyParams``, include/rbps/API/BodyAPI.hpp:22-31, and
``rbps::ColliderParams``, include/rbps/API/ColliderAPI.hpp:17-27).

Shape codes are the reference's ``rbc::Shape`` variant indices
(include/rbc/shapes/ShapeTypes.hpp:69): sphere 0, box 1, capsule 3, plane 6.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

STATIC, DYNAMIC = 0, 1
SPHERE, BOX, CAPSULE, PLANE, HULL = 0, 1, 3, 6, 9

BODY_DTYPE = np.dtype([
    ("type", "<i4"), ("_pad", "<i4"), ("mass", "<f8"), ("inertia", "<f8", (6,)),
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

    @property
    def n_dynamic(self) -> int:
        return int((self.bodies["type"] == DYNAMIC).sum())


class SceneBuilder:
    def __init__(self):
        self.b = []
        self.c = []

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
                 restitution=0.5, static_friction=0.5, dynamic_friction=0.5) -> int:
        r = np.zeros((), COLLIDER_DTYPE)
        r["shape"] = shape
        r["body_id"] = body
        r["is_static"] = 1 if is_static else 0
        r["hull_index"] = -1
        d = list(dims) + [0.0] * (4 - len(dims))
        r["dims"] = d
        r["local_pos"] = local_pos
        r["local_rot"] = local_rot
        r["restitution"] = restitution
        r["static_friction"] = static_friction
        r["dynamic_friction"] = dynamic_friction
        self.c.append(r)
        return len(self.c) - 1

    def build(self, name, timestep=1.0 / 60.0, substeps=20, frames=1, **meta) -> Scene:
        bodies = np.array(self.b, dtype=BODY_DTYPE) if self.b else np.zeros(0, BODY_DTYPE)
        cols = np.array(self.c, dtype=COLLIDER_DTYPE) if self.c else np.zeros(0, COLLIDER_DTYPE)
        return Scene(name, bodies, cols, timestep, substeps, frames, dict(meta))


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
