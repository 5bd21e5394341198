"""ctypes bindings for the oracle (liboracle.so) and the compiled reference
(_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

ORDER_REFERENCE, ORDER_COLOUR = 0, 1

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_U = C.POINTER(C.c_uint32)
_I = C.POINTER(C.c_int32)


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _bind(lib, prefix):
    f = lambda n: getattr(lib, prefix + n)
    f("destroy").argtypes = [_P]
    f("add_bodies").argtypes = [_P, C.c_int, _P]
    f("add_colliders").argtypes = [_P, C.c_int, _P]
    f("step").argtypes = [_P, C.c_int]
    f("num_bodies").argtypes = [_P]
    f("get_bodies").argtypes = [_P, _D, _D, _D, _D]
    f("set_bodies").argtypes = [_P, _D, _D, _D, _D]
    f("num_pairs").argtypes = [_P]
    f("get_pairs").argtypes = [_P, _U]
    f("num_contacts").argtypes = [_P]
    f("get_contacts").argtypes = [_P, _U, _D, _D, _D, _D, _D]
    f("collide").restype = C.c_int
    f("add_joints").argtypes = [_P, C.c_int, _P]
    f("num_joints").argtypes = [_P]
    f("get_joints").argtypes = [_P, _D, _D, _D, _I]
    f("set_joint_targets").argtypes = [_P, C.c_int, _U, _D, _D]
    f("set_rows").argtypes = [_P, C.c_int, _U, _D, _D]
    f("set_forces").argtypes = [_P, _D, _D]
    f("hull_register").restype = C.c_int
    f("hull_register").argtypes = [_D, C.c_int, _U, C.c_int]
    f("hull_sizes").argtypes = [C.c_int, _I]
    f("hull_tables").argtypes = [C.c_int, _D, _U, _U, _U, _U, _U, _U, _U, _U]
    f("hull_mass").argtypes = [C.c_int, _D]


_libs = {}
_hull_ids = {}   # (lib prefix, content key) -> registered index


def _hull_key(verts, faces):
    v = np.ascontiguousarray(verts, np.float64)
    f = None if faces is None else np.ascontiguousarray(faces, np.uint32)
    return v, f, (v.tobytes(), None if f is None else f.tobytes())


def register_hull(prefix, verts, faces=None):
    """Registers a hull (vertices (n,3), CCW triangles (m,3) or None) in the oracle ("orc_")
    or the reference build ("ref_"); the same content maps to the same index."""
    lib = oracle_lib() if prefix == "orc_" else ref_lib()
    v, f, key = _hull_key(verts, faces)
    k = (prefix, key)
    if k not in _hull_ids:
        i = getattr(lib, prefix + "hull_register")(_ptr(v, _D), len(v), _ptr(f, _U), 0 if f is None else len(f))
        if i < 0:
            raise ValueError("invalid hull")
        _hull_ids[k] = i
    return _hull_ids[k]


def hull_tables(prefix, verts, faces=None):
    """Merged polygons and adjacency tables of a hull as the library built them."""
    lib = oracle_lib() if prefix == "orc_" else ref_lib()
    i = register_hull(prefix, verts, faces)
    sz = np.zeros(6, np.int32)
    getattr(lib, prefix + "hull_sizes")(i, _ptr(sz, _I))
    P, NI, N2P, N2N, NP2P = (int(x) for x in sz[:5])
    nv = len(verts)
    out = {"pn": np.zeros((P, 3)), "poff": np.zeros(P + 1 if P else 0, np.uint32), "pidx": np.zeros(NI, np.uint32),
           "v2p_o": np.zeros(nv + 1 if P else 0, np.uint32), "v2p_i": np.zeros(N2P, np.uint32),
           "v2n_o": np.zeros(nv + 1 if P else 0, np.uint32), "v2n_i": np.zeros(N2N, np.uint32),
           "p2p_o": np.zeros(P + 1 if P else 0, np.uint32), "p2p_i": np.zeros(NP2P, np.uint32)}
    getattr(lib, prefix + "hull_tables")(i, *(_ptr(out[k], _D if k == "pn" else _U) for k in
                                              ("pn", "poff", "pidx", "v2p_o", "v2p_i", "v2n_o", "v2n_i", "p2p_o", "p2p_i")))
    return out


def hull_mass(prefix, verts, faces=None):
    """(volume, inertia xx yy zz xy xz yz) at unit density."""
    lib = oracle_lib() if prefix == "orc_" else ref_lib()
    out = np.zeros(7)
    getattr(lib, prefix + "hull_mass")(register_hull(prefix, verts, faces), _ptr(out, _D))
    return out[0], out[1:]


def map_hulls(prefix, colliders, hulls):
    """Copy of collider descriptors whose hull_index (into the scene's hull list) is
    replaced by the library's registered index."""
    c = np.array(colliders, copy=True)
    if hulls is None or not len(c):
        return c
    sel = c["shape"] == 9
    for k in np.nonzero(sel)[0]:
        hv, hf = hulls[int(c["hull_index"][k])]
        c["hull_index"][k] = register_hull(prefix, hv, hf)
    return c


def oracle_lib():
    if "orc" not in _libs:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run `make -C oracle`)")
        lib = C.CDLL(ORACLE_SO)
        lib.orc_create.restype = _P
        lib.orc_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int]
        _bind(lib, "orc_")
        lib.orc_collide.argtypes = [C.c_int, _P, _D, _P, _D, _D]
        lib.orc_collide_warm.restype = C.c_int
        lib.orc_collide_warm.argtypes = [C.c_int, _P, _D, _P, _D, _D, _D]
        lib.orc_gjk_epa.restype = C.c_int
        lib.orc_gjk_epa.argtypes = [C.c_int, _P, _D, _P, _D, _D, _D]
        lib.orc_num_solve_pairs.argtypes = [_P]
        lib.orc_get_solve_pairs.argtypes = [_P, _U, _I, _I, _I]
        lib.orc_num_colours.argtypes = [_P]
        lib.orc_colour_rounds.argtypes = [_P]
        lib.orc_set_palette_max.argtypes = [_P, C.c_int]
        lib.orc_set_joint_mode.argtypes = [_P, C.c_int]
        lib.orc_set_contact_mode.argtypes = [_P, C.c_int]
        lib.orc_set_epa_guard.argtypes = [_P, C.c_int]
        lib.orc_guard_stats.argtypes = [_P, _D]
        lib.orc_gjk_epa_guarded.restype = C.c_int
        lib.orc_gjk_epa_guarded.argtypes = [C.c_int, _P, _D, _P, _D, _D, _D]
        lib.orc_set_planted_fault.argtypes = [_P, C.c_int]
        lib.orc_get_solve_order.argtypes = [_P, _U]
        lib.orc_gjk_trace.argtypes = [C.c_int, _P, _D, _P, _D, _I]
        lib.orc_joint_errors.argtypes = [C.c_int, _P, _D, _D, C.c_double, _D]
        for f in ("orc_sinf", "orc_cosf"):
            getattr(lib, f).restype = C.c_float
            getattr(lib, f).argtypes = [C.c_float]
        lib.orc_atan2f.restype = C.c_float
        lib.orc_atan2f.argtypes = [C.c_float, C.c_float]
        lib.orc_mix32.restype = C.c_uint32
        lib.orc_mix32.argtypes = [C.c_uint32]
        _libs["orc"] = lib
    return _libs["orc"]


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    if "ref" not in _libs:
        if not ref_available():
            raise RuntimeError(f"reference build missing: {REF_SO} (run `make -C oracle ref` where the reference is present)")
        lib = C.CDLL(REF_SO)
        lib.ref_create.restype = _P
        lib.ref_create.argtypes = [C.c_double, C.c_int]
        _bind(lib, "ref_")
        lib.ref_collide.argtypes = [_P, _D, _P, _D, _D]
        lib.ref_gjk_epa.restype = C.c_int
        lib.ref_gjk_epa.argtypes = [_P, _D, _P, _D, _D]
        lib.ref_collision_groups.argtypes = [_P, _I]
        lib.ref_add_joints.restype = C.c_int
        lib.ref_joint_errors.restype = C.c_int
        lib.ref_joint_errors.argtypes = [_P, _D, _D, C.c_double, _D]
        _libs["ref"] = lib
    return _libs["ref"]


class _WorldBase:
    prefix = ""

    def __init__(self, lib, handle):
        self.lib = lib
        self.h = handle
        self._fn = lambda n: getattr(lib, self.prefix + n)

    def close(self):
        if self.h:
            self._fn("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, scene):
        b = np.ascontiguousarray(scene.bodies)
        c = np.ascontiguousarray(map_hulls(self.prefix, scene.colliders, getattr(scene, "hulls", None)))
        self._fn("add_bodies")(self.h, len(b), b.ctypes.data_as(_P))
        if self._fn("add_colliders")(self.h, len(c), c.ctypes.data_as(_P)):
            raise RuntimeError(f"{self.prefix}add_colliders: invalid collider")
        j = np.ascontiguousarray(getattr(scene, "joints", np.zeros(0)))
        if len(j):
            if self._fn("add_joints")(self.h, len(j), j.ctypes.data_as(_P)):
                raise RuntimeError(f"{self.prefix}add_joints: unsupported joint type")
        return self

    def joints(self):
        n = self._fn("num_joints")(self.h)
        cur, tgt, lam, col = np.zeros(n), np.zeros(n), np.zeros((n, 4)), np.zeros(n, np.int32)
        if n:
            self._fn("get_joints")(self.h, _ptr(cur, _D), _ptr(tgt, _D), _ptr(lam, _D), _ptr(col, _I))
        return {"current_position": cur, "target_position": tgt, "lambda": lam, "colour": col}

    def set_joint_targets(self, ids, target_position=None, target_speed=None):
        ids = np.ascontiguousarray(ids, np.uint32)
        tp = None if target_position is None else np.ascontiguousarray(target_position, np.float64)
        ts = None if target_speed is None else np.ascontiguousarray(target_speed, np.float64)
        self._fn("set_joint_targets")(self.h, len(ids), _ptr(ids, _U), _ptr(tp, _D), _ptr(ts, _D))

    def set_rows(self, ids, direction=None, magnitude=None):
        """Direction / magnitude of constraint rows (joints of type 4), as a caller of the
        reference writes ConstraintCollection::direction / magnitude."""
        ids = np.ascontiguousarray(ids, np.uint32)
        d = None if direction is None else np.ascontiguousarray(direction, np.float64).reshape(-1)
        m = None if magnitude is None else np.ascontiguousarray(magnitude, np.float64)
        self._fn("set_rows")(self.h, len(ids), _ptr(ids, _U), _ptr(d, _D), _ptr(m, _D))

    def set_forces(self, force=None, torque=None):
        """BodyCollection::force / torque of every body (n x 3 each), as a caller of the
        reference writes them."""
        f = None if force is None else np.ascontiguousarray(force, np.float64).reshape(-1)
        t = None if torque is None else np.ascontiguousarray(torque, np.float64).reshape(-1)
        self._fn("set_forces")(self.h, _ptr(f, _D), _ptr(t, _D))

    def step(self, frames=1):
        if self._fn("step")(self.h, frames) != 0:
            raise RuntimeError(f"{self.prefix}step failed (colour capacity exceeded)")

    def state(self):
        n = self._fn("num_bodies")(self.h)
        p, q, v, w = (np.zeros((n, 3)), np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3)))
        self._fn("get_bodies")(self.h, _ptr(p, _D), _ptr(q, _D), _ptr(v, _D), _ptr(w, _D))
        return {"position": p, "orientation": q, "linear_velocity": v, "angular_velocity": w}

    def set_state(self, st):
        a = {k: np.ascontiguousarray(st[k], dtype=np.float64) for k in st}
        self._fn("set_bodies")(self.h, _ptr(a.get("position"), _D), _ptr(a.get("orientation"), _D),
                               _ptr(a.get("linear_velocity"), _D), _ptr(a.get("angular_velocity"), _D))

    def pairs(self):
        n = self._fn("num_pairs")(self.h)
        out = np.zeros((n, 2), np.uint32)
        if n:
            self._fn("get_pairs")(self.h, _ptr(out, _U))
        return out

    def contacts(self):
        n = self._fn("num_contacts")(self.h)
        ids = np.zeros((n, 4), np.uint32)
        nrm, ra, rb = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        dep, lam = np.zeros(n), np.zeros((n, 2))
        if n:
            self._fn("get_contacts")(self.h, _ptr(ids, _U), _ptr(nrm, _D), _ptr(ra, _D), _ptr(rb, _D),
                                     _ptr(dep, _D), _ptr(lam, _D))
        return {"ids": ids, "normal": nrm, "r_a_local": ra, "r_b_local": rb, "depth": dep, "lambda": lam}


class OracleWorld(_WorldBase):
    """The oracle restatement: fp32 selects float arithmetic, order selects
    the reference's sequential order or the GPU's colour order."""
    prefix = "orc_"

    def __init__(self, timestep=1.0 / 60.0, substeps=20, fp32=False, order=ORDER_REFERENCE):
        lib = oracle_lib()
        super().__init__(lib, lib.orc_create(1 if fp32 else 0, order, timestep, substeps))
        self.fp32, self.order = fp32, order

    def set_joint_mode(self, mode):
        """0: the reference's joint solver; 1: XPBD rows (see orc_world.hpp)."""
        self.lib.orc_set_joint_mode(self.h, int(mode))

    def set_contact_mode(self, mode):
        """0: contacts as the reference makes them; 1: capsule-capsule normal from A to B."""
        self.lib.orc_set_contact_mode(self.h, int(mode))

    def set_epa_guard(self, on):
        """1: bound EPA depths by the overlap along the pair's previous-substep normal
        (colour order; see orc_gjk.hpp gjk_epa_manifold)."""
        self.lib.orc_set_epa_guard(self.h, int(on))

    def guard_stats(self):
        """(EPA results the guard replaced, largest depth it cut in m)."""
        out = np.zeros(2)
        self.lib.orc_guard_stats(self.h, _ptr(out, _D))
        return int(out[0]), float(out[1])

    def plant_fault(self, kind):
        """Plants a colour-schedule mistake (1: the last colour's pairs skipped, 2: the
        first pair solved twice) so tests can show their pins catch it."""
        self.lib.orc_set_planted_fault(self.h, int(kind))

    def set_palette_max(self, m):
        if self.lib.orc_set_palette_max(self.h, m) != 0:
            raise ValueError("palette maximum must be >= 1")

    def solve_pairs(self):
        n = self.lib.orc_num_solve_pairs(self.h)
        cc, col, cnt = np.zeros((n, 2), np.uint32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        kind = np.zeros(n, np.int32)
        if n:
            self.lib.orc_get_solve_pairs(self.h, _ptr(cc, _U), _ptr(col, _I), _ptr(cnt, _I), _ptr(kind, _I))
        order = np.zeros(self.lib.orc_get_solve_order(self.h, None), np.uint32)
        if len(order):
            self.lib.orc_get_solve_order(self.h, _ptr(order, _U))
        return {"colliders": cc, "colour": col, "count": cnt, "kind": kind, "ncolours": self.lib.orc_num_colours(self.h),
                "rounds": self.lib.orc_colour_rounds(self.h), "order": order.astype(np.int64)}


class RefWorld(_WorldBase):
    """The reference engine compiled from its own sources (FP64, sequential)."""
    prefix = "ref_"

    def __init__(self, timestep=1.0 / 60.0, substeps=20):
        lib = ref_lib()
        super().__init__(lib, lib.ref_create(timestep, substeps))

    def collision_groups(self):
        n = self.lib.ref_num_contacts(self.h)
        out = np.full(n, -1, np.int32)
        g = self.lib.ref_collision_groups(self.h, _ptr(out, _I)) if n else 0
        return out, g


def _collide(fn, pre, a, tfa, b, tfb, hulls=None, prefix="orc_"):
    out = np.zeros(20)
    a = np.ascontiguousarray(map_hulls(prefix, np.atleast_1d(a), hulls))
    b = np.ascontiguousarray(map_hulls(prefix, np.atleast_1d(b), hulls))
    tfa = np.ascontiguousarray(tfa, dtype=np.float64)
    tfb = np.ascontiguousarray(tfb, dtype=np.float64)
    hit = fn(*pre, a.ctypes.data_as(_P), _ptr(tfa, _D), b.ctypes.data_as(_P), _ptr(tfb, _D), _ptr(out, _D))
    if hit < 0:
        raise RuntimeError("collide: invalid shape")
    n = int(out[0])
    return {"hit": bool(hit), "n": n, "normal": out[1:4].copy(),
            "points": out[4:4 + 4 * n].reshape(n, 4)[:, :3].copy(), "depth": out[4:4 + 4 * n].reshape(n, 4)[:, 3].copy()}


def oracle_collide(a, tfa, b, tfb, fp32=False, hulls=None):
    return _collide(oracle_lib().orc_collide, (1 if fp32 else 0,), a, tfa, b, tfb, hulls, "orc_")


def oracle_collide_warm(a, tfa, b, tfb, guess6, fp32=True, hulls=None):
    """Narrowphase of one pair with a warm start for GJK pairs (guess6 in A's frame: GJK
    direction, EPA seed normal; zeros: plain)."""
    g = np.ascontiguousarray(guess6, np.float64)
    fn = lambda f32, pa, ta, pb, tb, out: oracle_lib().orc_collide_warm(f32, pa, ta, pb, tb, _ptr(g, _D), out)  # noqa: E731
    return _collide(fn, (1 if fp32 else 0,), a, tfa, b, tfb, hulls, "orc_")


def ref_collide(a, tfa, b, tfb, hulls=None):
    return _collide(ref_lib().ref_collide, (), a, tfa, b, tfb, hulls, "ref_")


GJK_VALID, GJK_INSIDE, GJK_FAILED = 0, 1, 2
EPA_VALID, EPA_FAILED, EPA_OUT_OF_FACES, EPA_OUT_OF_VERTICES, EPA_NONLINEAR = 0, 1, 2, 3, 4


def _gjk_epa_out(out, hit):
    return {"hit": bool(hit), "gjk_status": int(out[0]), "gjk_iters": int(out[1]), "epa_status": int(out[2]),
            "epa_iters": int(out[3]), "normal": out[4:7].copy(), "depth": float(out[7]), "point": out[8:11].copy(),
            "gjk_dir": out[11:14].copy(), "fallback": int(out[14])}


def oracle_gjk_epa(a, tfa, b, tfb, fp32=False, hulls=None, guess=None, guard=None, guard_faces=False):
    """GJK + EPA of one pair with statuses, iteration counts and the GJK exit direction.
    guess (optional, A's frame): GJK direction (3) and EPA seed normal (3, zero = none).
    guard (optional, world unit vector) / guard_faces: the EPA guard's directions (no warm
    start then)."""
    out = np.zeros(15)
    if guard is not None or guard_faces:
        guard = np.concatenate([np.zeros(3) if guard is None else np.asarray(guard, np.float64),
                                [1.0 if guard_faces else 0.0]])
        a = np.ascontiguousarray(map_hulls("orc_", np.atleast_1d(a), hulls))
        b = np.ascontiguousarray(map_hulls("orc_", np.atleast_1d(b), hulls))
        gd = np.ascontiguousarray(guard, np.float64)
        hit = oracle_lib().orc_gjk_epa_guarded(1 if fp32 else 0, a.ctypes.data_as(_P),
                                               _ptr(np.ascontiguousarray(tfa, np.float64), _D), b.ctypes.data_as(_P),
                                               _ptr(np.ascontiguousarray(tfb, np.float64), _D), _ptr(gd, _D),
                                               _ptr(out, _D))
        if hit < 0:
            raise RuntimeError("gjk_epa: invalid shape")
        return _gjk_epa_out(out, hit)
    if guess is not None and len(guess) == 3:
        guess = np.concatenate([guess, np.zeros(3)])
    a = np.ascontiguousarray(map_hulls("orc_", np.atleast_1d(a), hulls))
    b = np.ascontiguousarray(map_hulls("orc_", np.atleast_1d(b), hulls))
    g = None if guess is None else np.ascontiguousarray(guess, np.float64)
    hit = oracle_lib().orc_gjk_epa(1 if fp32 else 0, a.ctypes.data_as(_P), _ptr(np.ascontiguousarray(tfa, np.float64), _D),
                                   b.ctypes.data_as(_P), _ptr(np.ascontiguousarray(tfb, np.float64), _D), _ptr(g, _D),
                                   _ptr(out, _D))
    if hit < 0:
        raise RuntimeError("gjk_epa: invalid shape")
    return _gjk_epa_out(out, hit)


def oracle_gjk_trace(a, tfa, b, tfb, fp32=True, hulls=None):
    """Plain-seed GJK of one pair with its loop states recorded: status, iterations, the
    first iteration whose state (rank, simplex points, direction) equals the state `period`
    iterations earlier (-1: none within 16), and that period."""
    out = np.zeros(4, np.int32)
    a = np.ascontiguousarray(map_hulls("orc_", np.atleast_1d(a), hulls))
    b = np.ascontiguousarray(map_hulls("orc_", np.atleast_1d(b), hulls))
    oracle_lib().orc_gjk_trace(1 if fp32 else 0, a.ctypes.data_as(_P), _ptr(np.ascontiguousarray(tfa, np.float64), _D),
                               b.ctypes.data_as(_P), _ptr(np.ascontiguousarray(tfb, np.float64), _D), _ptr(out, _I))
    return {"gjk_status": int(out[0]), "gjk_iters": int(out[1]), "first_repeat": int(out[2]), "period": int(out[3])}


def ref_gjk_epa(a, tfa, b, tfb, hulls=None):
    out = np.zeros(15)
    a = np.ascontiguousarray(map_hulls("ref_", np.atleast_1d(a), hulls))
    b = np.ascontiguousarray(map_hulls("ref_", np.atleast_1d(b), hulls))
    hit = ref_lib().ref_gjk_epa(a.ctypes.data_as(_P), _ptr(np.ascontiguousarray(tfa, np.float64), _D),
                                b.ctypes.data_as(_P), _ptr(np.ascontiguousarray(tfb, np.float64), _D), _ptr(out, _D))
    if hit < 0:
        raise RuntimeError("gjk_epa: invalid shape")
    return _gjk_epa_out(out, hit)


def _joint_err(fn, pre, desc, pose1, pose2, h):
    out = np.zeros(18)
    d = np.ascontiguousarray(desc)
    p1 = np.ascontiguousarray(pose1, np.float64)
    p2 = np.ascontiguousarray(pose2, np.float64)
    rc = fn(*pre, d.ctypes.data_as(_P), _ptr(p1, _D), _ptr(p2, _D), h, _ptr(out, _D))
    if rc not in (None, 0):
        raise RuntimeError("joint type not supported")
    return {"direction": out[:16].reshape(4, 4)[:, :3].copy(), "magnitude": out[:16].reshape(4, 4)[:, 3].copy(),
            "current_position": out[16], "target_position": out[17]}


def oracle_joint_errors(desc, pose1, pose2, h, fp32=False):
    """Rows of one joint between two bodies at poses (x y z qw qx qy qz)."""
    return _joint_err(oracle_lib().orc_joint_errors, (1 if fp32 else 0,), desc, pose1, pose2, h)


def ref_joint_errors(desc, pose1, pose2, h):
    return _joint_err(ref_lib().ref_joint_errors, (), desc, pose1, pose2, h)


def sinf(x):
    return oracle_lib().orc_sinf(x)


def cosf(x):
    return oracle_lib().orc_cosf(x)


def atan2f(y, x):
    return oracle_lib().orc_atan2f(y, x)


def mix32(x: int) -> int:
    return int(oracle_lib().orc_mix32(x))
