"""Python mirror of ``rbps::World`` (include/rbps/API/World.hpp:87-212) over the
rbgpu C-ABI (include/rbgpu.h).

Same names and argument meaning as the reference API: ``World(timestep,
substeps)``, ``create_body``, ``create_collider``, ``step``; bodies are read back
as the reference's ``BodyCollection`` fields.  All computation runs in the HIP
library; if ``librbgpu.so`` is missing this module raises instead of falling
back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import scenes
from .scenes import BODY_DTYPE, COLLIDER_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("RBGPU_LIBRARY") or os.path.join(HERE, "librbgpu.so")   # override: another build of the library

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_U = C.POINTER(C.c_uint32)
_I = C.POINTER(C.c_int32)

EXPORTS = [
    "rbgpu_world_create", "rbgpu_world_destroy", "rbgpu_add_bodies", "rbgpu_add_colliders", "rbgpu_step",
    "rbgpu_set_step", "rbgpu_synchronize", "rbgpu_get_bodies", "rbgpu_set_bodies", "rbgpu_get_sap_pairs", "rbgpu_get_pair_colours",
    "rbgpu_get_contacts", "rbgpu_get_stats", "rbgpu_set_timing", "rbgpu_get_timing", "rbgpu_collide_pairs",
    "rbgpu_collide_pairs_noalloc", "rbgpu_add_constraints", "rbgpu_set_constraints", "rbgpu_get_constraint_rows",
    "rbgpu_get_contact_points", "rbgpu_get_broadphase_objects", "rbgpu_set_body_forces", "rbgpu_get_body_forces",
    "rbgpu_autotune",
    "rbgpu_set_option", "rbgpu_get_option", "rbgpu_add_joints", "rbgpu_get_joints", "rbgpu_set_joint_targets",
    "rbgpu_add_hull", "rbgpu_hull_mass_properties", "rbgpu_hull_tables", "rbgpu_collide_world_pairs",
    "rbgpu_mlp_init", "rbgpu_mlp_set_weights", "rbgpu_mlp_get_weights", "rbgpu_mlp_set_mode", "rbgpu_mlp_log",
    "rbgpu_mlp_logged", "rbgpu_mlp_get_log", "rbgpu_mlp_train", "rbgpu_mlp_eval", "rbgpu_mlp_gradients", "rbgpu_get_gjk_totals", "rbgpu_get_gjk_counts",
    "rbgpu_snapshot_create", "rbgpu_snapshot_capture", "rbgpu_snapshot_restore", "rbgpu_snapshot_size",
    "rbgpu_snapshot_save", "rbgpu_snapshot_load", "rbgpu_snapshot_destroy",
    "rbgpu_last_error",
]


class RbgpuError(RuntimeError):
    pass


class TuneReport(C.Structure):
    _fields_ = [("scene_kernels", C.c_int32), ("scene_threads", C.c_int32), ("graph", C.c_int32), ("run_min", C.c_int32),
                ("round_batch", C.c_int32), ("gjk_grid", C.c_int32), ("epa_grid", C.c_int32), ("trials", C.c_int32),
                ("frames", C.c_int32), ("best_ms_per_frame", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("frames", C.c_uint32), ("sap_pairs", C.c_uint32), ("pairs", C.c_uint32), ("colours", C.c_uint32),
                ("colour_rounds", C.c_uint32), ("contacts", C.c_uint32), ("bins", C.c_uint32 * 16),
                ("active_pairs", C.c_uint32), ("graph_frames", C.c_uint32), ("scene_kernels", C.c_uint32)]


_lib = None


def lib():
    """Loads librbgpu.so (built by ``python rbps_amd/build.py``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RbgpuError(f"{LIB_PATH} is not built; run `python rbps_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        L.rbgpu_world_create.argtypes = [C.c_double, C.c_int, C.POINTER(_P)]
        L.rbgpu_world_destroy.argtypes = [_P]
        L.rbgpu_add_bodies.argtypes = [_P, C.c_int, _P, _U]
        L.rbgpu_add_colliders.argtypes = [_P, C.c_int, _P, _U]
        L.rbgpu_step.argtypes = [_P, C.c_int]
        L.rbgpu_set_step.argtypes = [_P, C.c_double, C.c_int]
        L.rbgpu_synchronize.argtypes = [_P]
        L.rbgpu_get_bodies.argtypes = [_P, _D, _D, _D, _D]
        L.rbgpu_set_bodies.argtypes = [_P, _D, _D, _D, _D]
        L.rbgpu_set_body_forces.argtypes = [_P, _D, _D]
        L.rbgpu_get_body_forces.argtypes = [_P, _D, _D]
        L.rbgpu_get_sap_pairs.argtypes = [_P, _U, C.c_uint32, _U]
        L.rbgpu_get_pair_colours.argtypes = [_P, _U, _I, C.c_uint32, _U]
        L.rbgpu_get_contacts.argtypes = [_P, _U, _D, _D, _D, _D, _D, C.c_uint32, _U]
        L.rbgpu_get_stats.argtypes = [_P, C.POINTER(Stats)]
        L.rbgpu_set_timing.argtypes = [_P, C.c_int]
        L.rbgpu_add_joints.argtypes = [_P, C.c_int, _P, _U]
        L.rbgpu_get_joints.argtypes = [_P, _D, _D, _D, _I]
        L.rbgpu_set_joint_targets.argtypes = [_P, C.c_int, _U, _D, _D]
        L.rbgpu_add_constraints.argtypes = [_P, C.c_int, _P, _U]
        L.rbgpu_set_constraints.argtypes = [_P, C.c_int, _U, _D, _D]
        L.rbgpu_get_constraint_rows.argtypes = [_P, _D, _D]
        L.rbgpu_get_contact_points.argtypes = [_P, _D, _D, _D, _I, C.c_uint32, _U]
        L.rbgpu_get_broadphase_objects.argtypes = [_P, _D, _D, _U, C.c_uint32, _U]
        L.rbgpu_set_option.argtypes = [_P, C.c_int, C.c_int]
        L.rbgpu_autotune.argtypes = [_P, C.c_int, C.POINTER(TuneReport)]
        L.rbgpu_get_option.argtypes = [_P, C.c_int, _I]
        L.rbgpu_get_timing.argtypes = [_P, _D, C.POINTER(C.c_uint64), _D]
        L.rbgpu_collide_pairs.argtypes = [C.c_int, _I, _D, _D, _I, _D, _D, _D]
        L.rbgpu_collide_pairs_noalloc.argtypes = [C.c_int, _I, _D, _D, _I, _D, _D, _D]
        L.rbgpu_add_hull.argtypes = [_P, _D, C.c_uint32, _U, C.c_uint32, _U]
        L.rbgpu_hull_mass_properties.argtypes = [_D, C.c_uint32, _U, C.c_uint32, _D, _D]
        L.rbgpu_hull_tables.argtypes = [_D, C.c_uint32, _U, C.c_uint32, _U, _D] + [_U] * 8
        _F = C.POINTER(C.c_float)
        L.rbgpu_collide_world_pairs.argtypes = [_P, C.c_uint32, _U, _D, _U, _F, C.c_int]
        L.rbgpu_mlp_init.argtypes = [_P, C.c_uint64]
        L.rbgpu_mlp_set_weights.argtypes = [_P, _F, C.c_uint32]
        L.rbgpu_mlp_get_weights.argtypes = [_P, _F, C.c_uint32]
        L.rbgpu_mlp_set_mode.argtypes = [_P, C.c_int]
        L.rbgpu_mlp_log.argtypes = [_P, C.c_uint64]
        L.rbgpu_mlp_logged.argtypes = [_P, C.POINTER(C.c_uint64), C.c_int]
        L.rbgpu_mlp_train.argtypes = [_P, C.c_int, C.c_uint32, C.c_double, _D]
        L.rbgpu_mlp_get_log.argtypes = [_P, C.c_uint64, C.c_uint32, _F, _F]
        L.rbgpu_mlp_eval.argtypes = [_P, C.c_uint32, _F, _F]
        L.rbgpu_mlp_gradients.argtypes = [_P, C.c_uint32, _F, _F, _F, _D]
        L.rbgpu_get_gjk_totals.argtypes = [_P, C.POINTER(C.c_uint64), C.c_int]
        L.rbgpu_get_gjk_counts.argtypes = [_P, C.POINTER(C.c_uint64), C.c_uint32, C.c_int]
        L.rbgpu_snapshot_create.argtypes = [_P, C.POINTER(_P)]
        L.rbgpu_snapshot_capture.argtypes = [_P, _P]
        L.rbgpu_snapshot_restore.argtypes = [_P, _P]
        L.rbgpu_snapshot_size.argtypes = [_P, C.POINTER(C.c_uint64)]
        L.rbgpu_snapshot_save.argtypes = [_P, _P, C.c_uint64]
        L.rbgpu_snapshot_load.argtypes = [_P, C.c_uint64, C.POINTER(_P)]
        L.rbgpu_snapshot_destroy.argtypes = [_P]
        L.rbgpu_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise RbgpuError(f"rbgpu error {rc}: {lib().rbgpu_last_error().decode()}")


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


NOALLOC_MAX_PAIRS = 16384


def collide_pairs(kind_a, dims_a, tf_a, kind_b, dims_b, tf_b, noalloc=False) -> np.ndarray:
    """Single-pair narrowphase of a batch on the GPU; rows = n, normal, (point, depth) x 4.
    noalloc: through rbgpu_collide_pairs_noalloc (fixed device staging, at most
    NOALLOC_MAX_PAIRS pairs per call)."""
    ka = np.ascontiguousarray(kind_a, np.int32)
    kb = np.ascontiguousarray(kind_b, np.int32)
    da, db = np.ascontiguousarray(dims_a, np.float64), np.ascontiguousarray(dims_b, np.float64)
    ta, tb = np.ascontiguousarray(tf_a, np.float64), np.ascontiguousarray(tf_b, np.float64)
    out = np.zeros((len(ka), 20))
    fn = lib().rbgpu_collide_pairs_noalloc if noalloc else lib().rbgpu_collide_pairs
    _check(fn(len(ka), _ptr(ka, _I), _ptr(da, _D), _ptr(ta, _D), _ptr(kb, _I), _ptr(db, _D), _ptr(tb, _D), _ptr(out, _D)))
    return out


def _hull_arrays(verts, faces):
    v = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
    f = None if faces is None else np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
    return v, f


def hull_mass_properties(verts, faces=None):
    """(volume, inertia xx yy zz xy xz yz) at unit density, ConvexHull.hpp:575-745."""
    v, f = _hull_arrays(verts, faces)
    vol, inert = np.zeros(1), np.zeros(6)
    _check(lib().rbgpu_hull_mass_properties(_ptr(v, _D), len(v), _ptr(f, _U), 0 if f is None else len(f),
                                            _ptr(vol, _D), _ptr(inert, _D)))
    return float(vol[0]), inert


def hull_tables(verts, faces=None) -> dict:
    """Merged polygons and adjacency tables as rbgpu_add_hull builds them."""
    v, f = _hull_arrays(verts, faces)
    nf = 0 if f is None else len(f)
    sz = np.zeros(6, np.uint32)
    _check(lib().rbgpu_hull_tables(_ptr(v, _D), len(v), _ptr(f, _U), nf, _ptr(sz, _U), *([None] * 9)))
    P, NI, N2P, N2N, NP2P = (int(x) for x in sz[:5])
    nv = len(v)
    out = {"pn": np.zeros((P, 3)), "poff": np.zeros(P + 1 if P else 0, np.uint32), "pidx": np.zeros(NI, np.uint32),
           "v2p_o": np.zeros(nv + 1 if P else 0, np.uint32), "v2p_i": np.zeros(N2P, np.uint32),
           "v2n_o": np.zeros(nv + 1 if P else 0, np.uint32), "v2n_i": np.zeros(N2N, np.uint32),
           "p2p_o": np.zeros(P + 1 if P else 0, np.uint32), "p2p_i": np.zeros(NP2P, np.uint32)}
    _check(lib().rbgpu_hull_tables(_ptr(v, _D), len(v), _ptr(f, _U), nf, _ptr(sz, _U), _ptr(out["pn"], _D),
                                   *(_ptr(out[k], _U) for k in ("poff", "pidx", "v2p_o", "v2p_i", "v2n_o", "v2n_i",
                                                                "p2p_o", "p2p_i"))))
    return out


OPTIONS = {"palette_max": 0, "head_rounds": 1, "round_batch": 2, "tail_min": 3, "run_min": 4, "timing_stride": 5,
           "joint_mode": 6, "scene_kernels": 7, "scene_threads": 8, "gjk_cull": 9, "gjk_grid": 10, "epa_grid": 11,
           "graph": 12, "contact_mode": 13, "gjk_rows": 14,
           "epa_sub": 15, "gjk_occupancy": 16,
           "epa_occupancy": 17, "solve_occupancy": 18,
           "side_stream": 19, "axis_carry": 20, "epa_guard": 21, "gjk_sort": 22}

TIMING = {"off": 0, "solve_pos": 1, "solve_vel": 2, "narrowphase": 3, "integrate": 4, "update_vel": 5, "broadphase": 6,
          "colouring": 7, "solve_pos_runs": 8, "solve_vel_runs": 9, "joints": 10, "gjk": 11}


class Snapshot:
    """A device-resident snapshot of a world's state (World.snapshot(); the reference's
    WorldSnapshot design, docs/design/world-snapshot-restore/world-snapshot-restore.md:29-66).
    ``bytes(snap)`` / ``Snapshot.load(data)`` are the checkpoint form."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def load(cls, data: bytes) -> "Snapshot":
        buf = C.create_string_buffer(bytes(data), len(data))
        h = _P()
        _check(lib().rbgpu_snapshot_load(C.cast(buf, _P), len(data), C.byref(h)))
        return cls(h)

    def __len__(self):
        n = C.c_uint64()
        _check(lib().rbgpu_snapshot_size(self.h, C.byref(n)))
        return n.value

    def __bytes__(self):
        n = len(self)
        buf = C.create_string_buffer(n)
        _check(lib().rbgpu_snapshot_save(self.h, C.cast(buf, _P), n))
        return buf.raw

    def close(self):
        if getattr(self, "h", None):
            lib().rbgpu_snapshot_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class World:
    """A world on the GPU: ``World(timestep, substeps)`` as in World.hpp:89-106."""

    def __init__(self, timestep: float = 1.0 / 60.0, substeps: int = 20):
        self.timestep, self.substeps = timestep, substeps
        h = _P()
        _check(lib().rbgpu_world_create(timestep, substeps, C.byref(h)))
        self.h = h
        self.n_bodies = 0
        self.n_colliders = 0
        self.n_joints = 0
        self.n_hulls = 0

    def close(self):
        if getattr(self, "h", None):
            lib().rbgpu_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- creation (BodyAPI / ColliderAPI)
    def add_bodies(self, desc: np.ndarray) -> np.ndarray:
        desc = np.ascontiguousarray(desc, dtype=BODY_DTYPE)
        ids = np.zeros(len(desc), np.uint32)
        _check(lib().rbgpu_add_bodies(self.h, len(desc), desc.ctypes.data_as(_P), _ptr(ids, _U)))
        self.n_bodies += len(desc)
        return ids

    def add_colliders(self, desc: np.ndarray) -> np.ndarray:
        desc = np.ascontiguousarray(desc, dtype=COLLIDER_DTYPE)
        ids = np.zeros(len(desc), np.uint32)
        _check(lib().rbgpu_add_colliders(self.h, len(desc), desc.ctypes.data_as(_P), _ptr(ids, _U)))
        self.n_colliders += len(desc)
        return ids

    def create_body(self, type=scenes.DYNAMIC, mass=1.0, inertia_tensor=(1.0, 1.0, 1.0, 0.0, 0.0, 0.0),
                    position=(0, 0, 0), orientation=(1, 0, 0, 0), linear_velocity=(0, 0, 0),
                    angular_velocity=(0, 0, 0)) -> int:
        """rbps::create_body (BodyAPI.hpp:53-80); defaults of BodyParams (BodyAPI.hpp:22-31)."""
        sb = scenes.SceneBuilder()
        sb.body(type, mass, inertia_tensor, position, orientation, linear_velocity, angular_velocity)
        return int(self.add_bodies(sb.build("b").bodies)[0])

    def create_collider(self, shape, dims, body_id=0, is_static=False, local_pos=(0, 0, 0),
                        local_rot=(1, 0, 0, 0), restitution=0.5, static_friction=0.5, dynamic_friction=0.5) -> int:
        """rbps::create_collider (ColliderAPI.hpp:61-91); defaults of ColliderParams (ColliderAPI.hpp:17-27)."""
        sb = scenes.SceneBuilder()
        sb.collider(shape, body_id, dims, is_static, local_pos, local_rot, restitution, static_friction,
                    dynamic_friction)
        return int(self.add_colliders(sb.build("c").colliders)[0])

    def add_joints(self, desc) -> np.ndarray:
        """World::create_{prismatic,revolute,fixed}_joint (JointAPI.hpp:160-265) and the
        spherical extension, from JOINT_DTYPE rows; returns the joint ids."""
        desc = np.ascontiguousarray(desc, dtype=scenes.JOINT_DTYPE)
        ids = np.zeros(len(desc), np.uint32)
        _check(lib().rbgpu_add_joints(self.h, len(desc), desc.ctypes.data_as(_P), _ptr(ids, _U)))
        self.n_joints += len(desc)
        return ids

    def joints(self) -> dict:
        n = self.n_joints
        cur, tgt, lam, col = np.zeros(n), np.zeros(n), np.zeros((n, 4)), np.zeros(n, np.int32)
        if n:
            _check(lib().rbgpu_get_joints(self.h, _ptr(cur, _D), _ptr(tgt, _D), _ptr(lam, _D), _ptr(col, _I)))
        return {"current_position": cur, "target_position": tgt, "lambda": lam, "colour": col}

    def set_joint_targets(self, ids, target_position=None, target_speed=None):
        ids = np.ascontiguousarray(ids, np.uint32)
        tp = None if target_position is None else np.ascontiguousarray(target_position, np.float64)
        ts = None if target_speed is None else np.ascontiguousarray(target_speed, np.float64)
        _check(lib().rbgpu_set_joint_targets(self.h, len(ids), _ptr(ids, _U), _ptr(tp, _D), _ptr(ts, _D)))

    def add_constraints(self, desc) -> np.ndarray:
        """World::create_constraint (ConstraintAPI.hpp:37-64) from CONSTRAINT_DTYPE rows;
        returns their ids (shared with the joints)."""
        desc = np.ascontiguousarray(desc, dtype=scenes.CONSTRAINT_DTYPE)
        ids = np.zeros(len(desc), np.uint32)
        _check(lib().rbgpu_add_constraints(self.h, len(desc), desc.ctypes.data_as(_P), _ptr(ids, _U)))
        self.n_joints += len(desc)
        return ids

    def set_rows(self, ids, direction=None, magnitude=None):
        """ConstraintCollection::direction / magnitude of constraint rows, as the reference's
        callers write them."""
        ids = np.ascontiguousarray(ids, np.uint32)
        d = None if direction is None else np.ascontiguousarray(direction, np.float64).reshape(-1)
        m = None if magnitude is None else np.ascontiguousarray(magnitude, np.float64)
        _check(lib().rbgpu_set_constraints(self.h, len(ids), _ptr(ids, _U), _ptr(d, _D), _ptr(m, _D)))

    def constraint_rows(self) -> dict:
        n = self.n_joints
        d, m = np.zeros((n, 4, 3)), np.zeros((n, 4))
        if n:
            _check(lib().rbgpu_get_constraint_rows(self.h, _ptr(d, _D), _ptr(m, _D)))
        return {"direction": d, "magnitude": m}

    def add_hull(self, verts, faces=None) -> int:
        """rbc::convex_hull_data_create (ConvexHull.hpp:149-420): the world copies the hull;
        returns the id a hull collider names in hull_index."""
        v, f = _hull_arrays(verts, faces)
        hid = np.zeros(1, np.uint32)
        _check(lib().rbgpu_add_hull(self.h, _ptr(v, _D), len(v), _ptr(f, _U), 0 if f is None else len(f), _ptr(hid, _U)))
        self.n_hulls += 1
        return int(hid[0])

    def load(self, scene) -> "World":
        self.add_bodies(scene.bodies)
        cols = scene.colliders
        hulls = getattr(scene, "hulls", None) or []
        if hulls:
            ids = np.array([self.add_hull(v, f) for v, f in hulls], np.int32)
            cols = np.array(cols, copy=True)
            sel = cols["shape"] == scenes.HULL
            cols["hull_index"][sel] = ids[cols["hull_index"][sel]]
        self.add_colliders(cols)
        js = getattr(scene, "joints", None)
        if js is not None and len(js):
            self.add_joints(js)
        return self

    def collide_pairs(self, pairs, guess=None, seed_epa=False) -> dict:
        """Narrowphase of collider pairs (a, b) at the current state through the pipeline's
        binned kernels: points, normal, surface points p_a / p_b and depths per pair, and the
        GJK/EPA statistics (iterations, statuses, fallback; -1 where no GJK ran).  guess
        (n x 6, A's frame): explicit warm starts (GJK direction, EPA seed normal)."""
        pr = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        n = len(pr)
        out = np.zeros((n, 32))
        gs = np.zeros(n, np.uint32)
        g = None if guess is None else np.ascontiguousarray(guess, np.float32).reshape(n, 6)
        _check(lib().rbgpu_collide_world_pairs(self.h, n, _ptr(pr, _U), _ptr(out, _D), _ptr(gs, _U),
                                               _ptr(g, C.POINTER(C.c_float)), 1 if seed_epa else 0))
        cnt = out[:, 0].astype(int)
        pts = out[:, 4:].reshape(n, 4, 7)
        nogjk = gs == 0xFFFFFFFF
        g = gs.astype(np.int64)
        return {"n": cnt, "normal": out[:, 1:4], "p_a": pts[:, :, 0:3], "p_b": pts[:, :, 3:6], "depth": pts[:, :, 6],
                "gjk_iters": np.where(nogjk, -1, g & 255), "epa_iters": np.where(nogjk, -1, (g >> 8) & 255),
                "gjk_status": np.where(nogjk, -1, (g >> 16) & 15), "epa_status": np.where(nogjk, -1, (g >> 20) & 15),
                "fallback": np.where(nogjk, -1, (g >> 24) & 1)}

    # ---------------------------------------------------------------- learned warm start
    MLP_NPARAM = 39312
    MLP_MODES = {"plain": 0, "gjk": 1, "gjk+epa": 2}

    def mlp_init(self, seed: int = 1):
        _check(lib().rbgpu_mlp_init(self.h, seed))

    def mlp_weights(self) -> np.ndarray:
        p = np.zeros(self.MLP_NPARAM, np.float32)
        _check(lib().rbgpu_mlp_get_weights(self.h, _ptr(p, C.POINTER(C.c_float)), len(p)))
        return p

    def mlp_set_weights(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().rbgpu_mlp_set_weights(self.h, _ptr(p, C.POINTER(C.c_float)), len(p)))

    def mlp_mode(self, mode: str):
        _check(lib().rbgpu_mlp_set_mode(self.h, self.MLP_MODES[mode]))

    def mlp_log(self, capacity: int):
        _check(lib().rbgpu_mlp_log(self.h, int(capacity)))

    def mlp_logged(self, stop: bool = False) -> int:
        n = C.c_uint64()
        _check(lib().rbgpu_mlp_logged(self.h, C.byref(n), 1 if stop else 0))
        return int(n.value)

    def mlp_log_samples(self, start: int, n: int):
        """Logged features (n x 32) and targets (n x 8: GJK target xyz, valid, EPA target xyz, valid)."""
        f = np.zeros((n, 32), np.float32)
        t = np.zeros((n, 8), np.float32)
        _F = C.POINTER(C.c_float)
        _check(lib().rbgpu_mlp_get_log(self.h, start, n, _ptr(f, _F), _ptr(t, _F)))
        return f, t

    def mlp_train(self, steps: int, batch: int = 4096, lr: float = 1e-3) -> float:
        loss = np.zeros(1)
        _check(lib().rbgpu_mlp_train(self.h, steps, batch, lr, _ptr(loss, _D)))
        return float(loss[0])

    def mlp_eval(self, features) -> np.ndarray:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, 32)
        out = np.zeros((len(f), 6), np.float32)
        _check(lib().rbgpu_mlp_eval(self.h, len(f), _ptr(f, C.POINTER(C.c_float)), _ptr(out, C.POINTER(C.c_float))))
        return out

    def mlp_gradients(self, features, targets):
        f = np.ascontiguousarray(features, np.float32).reshape(-1, 32)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 8)
        g = np.zeros(self.MLP_NPARAM, np.float32)
        loss = np.zeros(1)
        _F = C.POINTER(C.c_float)
        _check(lib().rbgpu_mlp_gradients(self.h, len(f), _ptr(f, _F), _ptr(t, _F), _ptr(g, _F), _ptr(loss, _D)))
        return g, float(loss[0])

    def gjk_totals(self, reset: bool = True) -> dict:
        out = (C.c_uint64 * 7)()
        _check(lib().rbgpu_get_gjk_counts(self.h, out, 7, 1 if reset else 0))
        k = ("pairs", "gjk_iters", "epa_iters", "overlapping", "epa_valid", "fallbacks", "handed_over")
        return {a: int(b) for a, b in zip(k, out)}

    # ---------------------------------------------------------------- step
    def step(self, frames: int = 1):
        _check(lib().rbgpu_step(self.h, frames))

    def set_step(self, timestep: float, substeps: int):
        """Changes World::timestep / World::substeps for the following steps."""
        _check(lib().rbgpu_set_step(self.h, timestep, substeps))
        self.timestep, self.substeps = timestep, substeps

    def synchronize(self):
        _check(lib().rbgpu_synchronize(self.h))

    # ---------------------------------------------------------------- state
    def state(self) -> dict:
        n = self.n_bodies
        p, q, v, w = np.zeros((n, 3)), np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3))
        _check(lib().rbgpu_get_bodies(self.h, _ptr(p, _D), _ptr(q, _D), _ptr(v, _D), _ptr(w, _D)))
        return {"position": p, "orientation": q, "linear_velocity": v, "angular_velocity": w}

    def set_state(self, st: dict):
        a = {k: np.ascontiguousarray(st[k], dtype=np.float64) for k in st}
        _check(lib().rbgpu_set_bodies(self.h, _ptr(a.get("position"), _D), _ptr(a.get("orientation"), _D),
                                      _ptr(a.get("linear_velocity"), _D), _ptr(a.get("angular_velocity"), _D)))

    def set_forces(self, force=None, torque=None):
        """BodyCollection::force / torque (n x 3 each; Body.hpp:52-69): the constant load the
        integrator applies each substep, in place of gravity (BodyAPI.hpp:61-62)."""
        f = None if force is None else np.ascontiguousarray(force, np.float64).reshape(-1)
        t = None if torque is None else np.ascontiguousarray(torque, np.float64).reshape(-1)
        for a in (f, t):
            if a is not None and a.size != 3 * self.n_bodies:
                raise ValueError(f"expected {self.n_bodies} x 3 values, got {a.size}")
        _check(lib().rbgpu_set_body_forces(self.h, _ptr(f, _D), _ptr(t, _D)))

    def forces(self) -> dict:
        n = self.n_bodies
        f, t = np.zeros((n, 3)), np.zeros((n, 3))
        _check(lib().rbgpu_get_body_forces(self.h, _ptr(f, _D), _ptr(t, _D)))
        return {"force": f, "torque": t}

    def snapshot(self, into: "Snapshot" = None) -> "Snapshot":
        """Captures the state that carries between frames into device memory (a new
        snapshot, or `into`, reusing its block)."""
        if into is not None:
            _check(lib().rbgpu_snapshot_capture(self.h, into.h))
            return into
        h = _P()
        _check(lib().rbgpu_snapshot_create(self.h, C.byref(h)))
        return Snapshot(h)

    def restore(self, snap: "Snapshot"):
        """Puts the world back into the snapshot's state (same bodies, colliders, joints,
        scenes and hulls); the following steps repeat the original run bit for bit."""
        _check(lib().rbgpu_snapshot_restore(self.h, snap.h))

    def sap_pairs(self) -> np.ndarray:
        n = C.c_uint32()
        _check(lib().rbgpu_get_sap_pairs(self.h, None, 0, C.byref(n)))
        out = np.zeros((n.value, 2), np.uint32)
        if n.value:
            _check(lib().rbgpu_get_sap_pairs(self.h, _ptr(out, _U), n.value, C.byref(n)))
        return out

    def pair_colours(self):
        n = C.c_uint32()
        _check(lib().rbgpu_get_pair_colours(self.h, None, None, 0, C.byref(n)))
        pr, col = np.zeros((n.value, 2), np.uint32), np.zeros(n.value, np.int32)
        if n.value:
            _check(lib().rbgpu_get_pair_colours(self.h, _ptr(pr, _U), _ptr(col, _I), n.value, C.byref(n)))
        return pr, col

    def contacts(self) -> dict:
        n = C.c_uint32()
        _check(lib().rbgpu_get_contacts(self.h, None, None, None, None, None, None, 0, C.byref(n)))
        k = n.value
        ids = np.zeros((k, 4), np.uint32)
        nrm, ra, rb = np.zeros((k, 3)), np.zeros((k, 3)), np.zeros((k, 3))
        dep, lam = np.zeros(k), np.zeros((k, 2))
        if k:
            _check(lib().rbgpu_get_contacts(self.h, _ptr(ids, _U), _ptr(nrm, _D), _ptr(ra, _D), _ptr(rb, _D),
                                            _ptr(dep, _D), _ptr(lam, _D), k, C.byref(n)))
        return {"ids": ids, "normal": nrm, "r_a_local": ra, "r_b_local": rb, "depth": dep, "lambda": lam}

    def stats(self) -> dict:
        s = Stats()
        _check(lib().rbgpu_get_stats(self.h, C.byref(s)))
        return {"frames": s.frames, "sap_pairs": s.sap_pairs, "pairs": s.pairs, "colours": s.colours,
                "colour_rounds": s.colour_rounds, "active_pairs": s.active_pairs, "contacts": s.contacts,
                "bins": list(s.bins), "graph_frames": s.graph_frames, "scene_kernels": s.scene_kernels}

    # ---------------------------------------------------------------- options
    def set_option(self, name: str, value: int):
        """Pipeline options (OPTIONS); palette_max changes the colouring, the others only speed."""
        _check(lib().rbgpu_set_option(self.h, OPTIONS[name], int(value)))

    def option(self, name: str) -> int:
        v = C.c_int32()
        _check(lib().rbgpu_get_option(self.h, OPTIONS[name], C.byref(v)))
        return v.value

    # ---------------------------------------------------------------- timing
    def autotune(self, frames_per_trial: int = 2) -> dict:
        """Times the result-neutral launch options over real frames and keeps the fastest
        (rbgpu_autotune); the world advances as it would untuned."""
        r = TuneReport()
        _check(lib().rbgpu_autotune(self.h, int(frames_per_trial), C.byref(r)))
        return {k: getattr(r, k) for k, _ in TuneReport._fields_}

    def set_timing(self, kernel: str):
        _check(lib().rbgpu_set_timing(self.h, TIMING[kernel]))

    def timing(self) -> dict:
        ms, n, b = C.c_double(), C.c_uint64(), C.c_double()
        _check(lib().rbgpu_get_timing(self.h, C.byref(ms), C.byref(n), C.byref(b)))
        return {"ms": ms.value, "launches": n.value, "bytes": b.value}


class WorldPool:
    """The reference's multi-world design (docs/design/multi-world/multi-world.md:40-50) in
    Python: independent worlds, one per scene, stepped together as the scenes of one batched
    GPU world (every kernel runs once over all worlds; each world evolves exactly as it would
    alone).  ``worlds`` are the member scenes; states come back per world, or stacked
    ``(n_worlds, ...)`` when every world has the same body count."""

    def __init__(self, worlds, timestep: float = None, substeps: int = None):
        self.worlds = list(worlds)
        if not self.worlds:
            raise ValueError("a WorldPool needs at least one world")
        sc = scenes.batch(self.worlds, "pool")
        self.timestep = sc.timestep if timestep is None else timestep
        self.substeps = sc.substeps if substeps is None else substeps
        self.batch = World(self.timestep, self.substeps).load(sc)
        self._sizes = [len(w.bodies) for w in self.worlds]
        self._off = np.concatenate([[0], np.cumsum(self._sizes)]).astype(np.int64)

    def __len__(self):
        return len(self.worlds)

    def parallel_step(self, dt: float = None, frames: int = 1):
        """`frames` frames of dt (default: the pool's timestep) for every world."""
        if dt is not None and dt != self.timestep:
            self.batch.set_step(dt, self.substeps)
            self.timestep = dt
        self.batch.step(frames)

    def parallel_each(self, fn):
        """fn(index, state) for every world's state (host side), in world order; returns
        the list of results."""
        return [fn(i, st) for i, st in enumerate(self.states())]

    def states(self) -> list:
        st = self.batch.state()
        return [{k: v[self._off[i]:self._off[i + 1]] for k, v in st.items()} for i in range(len(self.worlds))]

    def stacked(self) -> dict:
        """Every world's state as (n_worlds, bodies, ...) arrays (equal body counts only)."""
        if len(set(self._sizes)) != 1:
            raise ValueError("stacked states need worlds of equal body count")
        st = self.batch.state()
        return {k: v.reshape(len(self.worlds), self._sizes[0], *v.shape[1:]) for k, v in st.items()}

    def set_states(self, states):
        """Writes every world's state (a list of per-world dicts, as states() returns)."""
        st = {k: np.concatenate([s[k] for s in states]) for k in states[0]}
        self.batch.set_state(st)

    def snapshot(self, into: "Snapshot" = None) -> "Snapshot":
        return self.batch.snapshot(into)

    def restore(self, snap: "Snapshot"):
        self.batch.restore(snap)

    def set_option(self, name: str, value: int):
        self.batch.set_option(name, value)
