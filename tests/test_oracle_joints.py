"""Joints in the oracle: pinned against the reference (goldens written from the
reference compiled from its own sources), its unit-test known answers
(tests/rbps/test_joint.cpp), the mathematics of the FP32 trig, and physical
invariants for the spherical extension, which has no reference counterpart.
"""
import math

import numpy as np
import pytest

from oracle import orc
from rbps_amd import scenes as S
from tests.conftest import golden

I7 = (0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0)


def _desc(t, **kw):
    sb = S.SceneBuilder()
    sb.joint(t, 0, 1, **kw)
    return sb.j[0]


def _rpy(roll=0.0, pitch=0.0, yaw=0.0):
    """quat::from_rpy (quat.hpp:67-80) for the single-axis cases of the reference tests."""
    cr, sr = math.cos(roll / 2), math.sin(roll / 2)
    cp, sp = math.cos(pitch / 2), math.sin(pitch / 2)
    cy, sy = math.cos(yaw / 2), math.sin(yaw / 2)
    return (cr * cp * cy + sr * sp * sy, sr * cp * cy - cr * sp * sy, cr * sp * cy + sr * cp * sy,
            cr * cp * sy - sr * sp * cy)


def _pose(pos=(0, 0, 0), q=(1, 0, 0, 0)):
    return tuple(pos) + tuple(q)


@pytest.mark.parametrize("fp32", [False, True])
def test_joint_error_known_answers(fp32):
    """The assertions of tests/rbps/test_joint.cpp:20-581 (tolerances as there)."""
    E = lambda d, p2, h=0.01, p1=I7: orc.oracle_joint_errors(d, p1, p2, h, fp32=fp32)
    # prismatic
    r = E(_desc(S.PRISMATIC, axis=(1, 0, 0), damping=0.1), I7)
    assert np.allclose(r["magnitude"][:3], 0.0, atol=1e-3)
    assert E(_desc(S.PRISMATIC, axis=(1, 0, 0)), _pose(q=_rpy(yaw=math.pi / 2)))["magnitude"][0] > 0.01
    r = E(_desc(S.PRISMATIC, axis=(1, 0, 0)), _pose((2, 0, 0)))
    assert abs(r["current_position"] + 2.0) < 1e-3 and abs(r["magnitude"][1]) < 1e-3
    r = E(_desc(S.PRISMATIC, axis=(1, 0, 0), limited=True, lower=-1, upper=1), _pose((0.5, 0, 0)))
    assert abs(r["magnitude"][1]) < 1e-3
    r = E(_desc(S.PRISMATIC, axis=(1, 0, 0), limited=True, lower=-1, upper=1), _pose((2, 0, 0)))
    assert r["magnitude"][1] > 0.5
    r = E(_desc(S.PRISMATIC, axis=(1, 0, 0), actuation=S.ACT_POSITION, target_position=1.0), _pose((0.5, 0, 0)))
    assert r["magnitude"][2] > 0.1
    r = E(_desc(S.PRISMATIC, axis=(1, 0, 0), actuation=S.ACT_SPEED, target_speed=2.0), I7, h=0.1)
    assert abs(r["target_position"] - 0.2) < 1e-3 and r["magnitude"][2] > 0.01
    r = E(_desc(S.PRISMATIC, axis=(1, 0, 0)), _pose((0, 1, 0)))
    assert abs(r["magnitude"][1] - 1.0) < 1e-3
    # revolute
    rd = dict(axis=(0, 0, 1), limit_axis=(1, 0, 0))
    assert np.allclose(E(_desc(S.REVOLUTE, **rd), I7)["magnitude"], 0.0, atol=1e-3)
    assert E(_desc(S.REVOLUTE, **rd), _pose(q=_rpy(roll=math.pi / 4)))["magnitude"][0] > 0.01
    assert abs(E(_desc(S.REVOLUTE, **rd), _pose((1, 0, 0)))["magnitude"][1] - 1.0) < 1e-3
    assert abs(E(_desc(S.REVOLUTE, **rd), _pose(q=_rpy(yaw=math.pi / 4)))["current_position"] - math.pi / 4) < 1e-2
    lim = dict(rd, limited=True, lower=-math.pi / 2, upper=math.pi / 2)
    assert abs(E(_desc(S.REVOLUTE, **lim), _pose(q=_rpy(yaw=0.5)))["magnitude"][2]) < 1e-2
    lim = dict(rd, limited=True, lower=-math.pi / 4, upper=math.pi / 4)
    assert E(_desc(S.REVOLUTE, **lim), _pose(q=_rpy(yaw=math.pi)))["magnitude"][2] > 0.1
    assert E(_desc(S.REVOLUTE, actuation=S.ACT_POSITION, target_position=math.pi / 4, **rd), I7)["magnitude"][3] > 0.1
    r = E(_desc(S.REVOLUTE, actuation=S.ACT_SPEED, target_speed=1.0, **rd), I7, h=0.1)
    assert abs(r["target_position"] - 0.1) < 1e-3 and r["magnitude"][3] > 0.01
    # fixed
    assert np.allclose(E(_desc(S.FIXED), I7)["magnitude"][:2], 0.0, atol=1e-3)
    assert E(_desc(S.FIXED), _pose(q=_rpy(yaw=math.pi / 2)))["magnitude"][0] > 0.5
    assert abs(E(_desc(S.FIXED), _pose((1, 2, 3)))["magnitude"][1] - math.sqrt(14.0)) < 1e-3
    assert abs(E(_desc(S.FIXED, r_1=(1, 0, 0), r_2=(-1, 0, 0)), I7)["magnitude"][1] - 2.0) < 1e-3


def test_joint_errors_bit_exact_vs_reference_golden():
    """900 random joints/poses through the reference's compute_*_joint_errors
    (tests/golden/joint_errors.npz): the FP64 oracle reproduces every bit."""
    g = golden("joint_errors")
    for i in range(len(g["desc"])):
        r = orc.oracle_joint_errors(g["desc"][i], g["pose1"][i], g["pose2"][i], float(g["h"][i]))
        got = np.concatenate([np.concatenate([r["direction"], r["magnitude"][:, None]], axis=1).ravel(),
                              [r["current_position"], r["target_position"]]])
        nrows = {0: 3, 1: 4, 2: 2}[int(g["desc"][i]["type"])]
        exp = g["result"][i]
        assert np.array_equal(got[:4 * nrows], exp[:4 * nrows]) and np.array_equal(got[16:], exp[16:]), i


@pytest.mark.parametrize("name", ["world_joint_zoo", "world_planar_joints"])
def test_joint_world_trajectory_bit_exact(name):
    """Whole steps with joints and contacts against the reference's trajectory.  The
    reference's own rotational rows produce NaN poses in these scenes (a zero row
    direction for errors in (1e-9, 3.2e-5), see oracle/orc_world.hpp); the oracle must
    reproduce those bit for bit as well (NaN where the reference has NaN)."""
    g = golden(name)
    sc = S.Scene("g", g["bodies"], g["colliders"], float(g["timestep"]), int(g["substeps"]), joints=g["joints"])
    w = orc.OracleWorld(sc.timestep, sc.substeps).load(sc)
    done = 0
    for f in g["frames"]:
        w.step(int(f) - done)
        done = int(f)
        st = w.state()
        for k in ("position", "orientation", "linear_velocity", "angular_velocity"):
            assert np.array_equal(st[k], g[f"{k}_{f}"], equal_nan=True), (name, f, k)
        j = w.joints()
        for k in ("current_position", "target_position", "lambda"):
            assert np.array_equal(j[k], g[f"joint_{k}_{f}"], equal_nan=True), (name, f, k)
        assert np.array_equal(w.pairs(), g[f"pairs_{f}"])


def test_reference_dead_zone_gives_nan_and_robust_mode_does_not():
    """The finding behind the robust option: a hinge whose attachment starts exactly
    satisfied turns NaN within one frame in the reference order (rounding noise feeds a
    rotational row into the zero-direction band); the robust direction keeps it finite."""
    sb = S.SceneBuilder()
    a = S._box(sb, (3.0, 3.0, 0.0), (0.2, 0.2, 0.2), type_=S.STATIC)
    m = S._box(sb, (3.0, 2.4, 0.0), (0.5, 0.05, 0.1))
    sb.joint(S.REVOLUTE, a, m, r_1=(0, -0.3, 0), r_2=(0, 0.3, 0), axis=(0, 1, 0), limit_axis=(1, 0, 0))
    sc = sb.build("hinge", 1 / 60, 20)
    w = orc.OracleWorld(sc.timestep, sc.substeps).load(sc)
    w.step(1)
    assert np.isnan(w.state()["position"][m]).all()
    w = orc.OracleWorld(sc.timestep, sc.substeps).load(sc)
    w.set_joint_mode(1)
    w.step(120)
    p = w.state()["position"][m]
    assert np.isfinite(p).all() and abs(p[1] - 2.4) < 1e-3


def test_fp32_trig_accuracy():
    """The FP32 sin/cos/atan2 the GPU evaluates (and the FP32 oracle restates) against
    float64 numpy: within 2 ulp of 1 for sin/cos on |x| <= 40, 3 ulp of pi for atan2, and
    exact at the special points."""
    rng = np.random.default_rng(3)
    xs = np.float32(rng.uniform(-40, 40, 4000))
    es = max(abs(orc.sinf(float(x)) - math.sin(float(x))) for x in xs)
    ec = max(abs(orc.cosf(float(x)) - math.cos(float(x))) for x in xs)
    assert es < 2.4e-7 and ec < 2.4e-7, (es, ec)
    ys = np.float32(rng.standard_normal((4000, 2)))
    ea = max(abs(orc.atan2f(float(a), float(b)) - math.atan2(float(a), float(b))) for a, b in ys)
    assert ea < 7.2e-7, ea
    assert orc.sinf(0.0) == 0.0 and orc.cosf(0.0) == 1.0 and orc.atan2f(0.0, 1.0) == 0.0
    assert abs(orc.atan2f(1.0, 0.0) - math.pi / 2) < 2e-7 and abs(orc.atan2f(0.0, -1.0) - math.pi) < 3e-7


def _chain_scene(n=6, limited_every=0, compliance=0.0):
    sb = S.SceneBuilder()
    a = S._anchor(sb, (0.0, 5.0, 0.0))
    S._chain(sb, a, (0.0, 5.0, 0.0), n, 0.2, 0.08, compliance, limited_every=limited_every)
    return sb.build("chain", 1 / 60, 20)


def _rot(qq, v):
    qw, u = qq[0], np.asarray(qq[1:])
    return v + 2.0 * np.cross(u, np.cross(u, v) + qw * v)


def _attachment_errors(sc, st):
    p, q = st["position"], st["orientation"]
    out = []
    for j in sc.joints:
        b1, b2 = int(j["body_1"]), int(j["body_2"])
        out.append(np.linalg.norm((p[b1] + _rot(q[b1], j["r_1"])) - (p[b2] + _rot(q[b2], j["r_2"]))))
    return out


def test_spherical_chain_holds_together_and_swings():
    """Spherical extension (no reference), XPBD joint mode: a 6-link chain released
    horizontally keeps every attachment within 5 cm through the swing (one XPBD
    iteration per substep), swings its free end below y = 2.5, and leaves the static
    anchor untouched.  The reference's joint scheme (mode 0) lets the same chain come
    apart, which is checked too."""
    sc = _chain_scene(6)
    w = orc.OracleWorld(sc.timestep, sc.substeps).load(sc)
    w.set_joint_mode(1)
    worst, lowest = 0.0, 10.0
    for _ in range(120):
        w.step(1)
        st = w.state()
        worst = max(worst, max(_attachment_errors(sc, st)))
        lowest = min(lowest, st["position"][-1][1])
    assert np.array_equal(st["position"][0], [0.0, 5.0, 0.0])
    assert worst < 0.05, worst
    assert lowest < 2.5, lowest
    w0 = orc.OracleWorld(sc.timestep, sc.substeps).load(sc)
    w0.step(30)
    err0 = _attachment_errors(sc, w0.state())
    assert not (max(err0) < 0.05), err0


def test_xpbd_hinge_stays_in_its_plane():
    """XPBD mode: a pendulum on a revolute hinge (axis z), released with a small
    out-of-plane tilt, keeps its hinge axis aligned (alignment row a2 x a1); the
    reference's a1 x a2 row pushes the axes apart (checked in mode 0)."""
    tilt = (math.cos(0.01), math.sin(0.01), 0.0, 0.0)
    res = {}
    for mode in (0, 1):
        sb = S.SceneBuilder()
        a = S._anchor(sb, (0.0, 5.0, 0.0))
        b = S._box(sb, (1.0, 5.0, 0.0), (0.4, 0.1, 0.1), quat=tilt)
        sb.joint(S.REVOLUTE, a, b, r_2=(-1.0, 0.0, 0.0), axis=(0, 0, 1), limit_axis=(1, 0, 0))
        sc = sb.build("hinge", 1 / 60, 20)
        w = orc.OracleWorld(sc.timestep, sc.substeps).load(sc)
        w.set_joint_mode(mode)
        worst = 0.0
        for _ in range(90):
            w.step(1)
            qq = w.state()["orientation"][b]
            az = _rot(qq, np.array([0.0, 0.0, 1.0]))
            worst = max(worst, float(np.nan_to_num(math.degrees(math.acos(np.clip(az[2], -1, 1))), nan=180.0)))
        res[mode] = worst
    assert res[1] < 1.5, res        # starts at 1.15 deg and is pulled back
    assert res[0] > 5.0, res


def test_spherical_swing_limit_holds():
    """A limited spherical joint (45 deg cone on the capsule axis) keeps the two link
    axes within the cone (plus the XPBD residual), where the unlimited chain bends far more."""
    bend = {}
    for lim in (0, 1):
        sc = _chain_scene(2, limited_every=lim)
        w = orc.OracleWorld(sc.timestep, sc.substeps).load(sc)
        w.set_joint_mode(1)
        w.step(60)
        worst = 0.0
        for _ in range(60):
            w.step(1)
            j = w.joints()
            worst = max(worst, float(j["current_position"][1]) if lim else 0.0)
            q = w.state()["orientation"]
            a = [np.array([2 * (x * y - w_ * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w_ * x)]) for w_, x, y, z in q[1:3]]
            bend[lim] = max(bend.get(lim, 0.0), math.degrees(math.acos(np.clip(np.dot(a[0], a[1]), -1, 1))))
    assert bend[1] < 45.0 + 3.0, bend
    assert bend[0] > bend[1] + 10.0, bend


def test_joint_colouring_valid():
    sc = S.stack_chain(64)
    w = orc.OracleWorld(sc.timestep, sc.substeps, order=orc.ORDER_COLOUR).load(sc)
    w.step(1)
    col = w.joints()["colour"]
    dyn = sc.bodies["type"] == S.DYNAMIC
    for c in range(col.max() + 1):
        bodies = [int(b) for j in sc.joints[col == c] for b in (j["body_1"], j["body_2"]) if dyn[int(b)]]
        assert len(bodies) == len(set(bodies)), c
    assert col.max() + 1 == 2   # a chain needs two colours


def test_constraint_rows_fp64_equal_reference_golden():
    """Constraint rows made by World::create_constraint (ConstraintAPI.hpp:37-64), created
    between the joints and written by the caller after loading (direction, magnitude):
    the FP64 restatement in the reference's order equals the reference engine
    (tests/golden/world_constraint_rows.npz) bit for bit -- states and every joint's and
    row's lambda.  A row keeps the direction and magnitude its caller wrote; the zero row
    stays at lambda 0."""
    g = golden("world_constraint_rows")
    sc = S.Scene("rows", g["bodies"], g["colliders"], float(g["timestep"]), int(g["substeps"]), joints=g["joints"])
    w = orc.OracleWorld(sc.timestep, sc.substeps).load(sc)
    w.set_rows(g["row_ids"], g["row_dir"], g["row_mag"])
    done = 0
    for f in (int(x) for x in g["frames"]):
        w.step(f - done)
        done = f
        st = w.state()
        for k in st:
            assert np.array_equal(st[k], g[f"{k}_{f}"]), (f, k)
        assert np.array_equal(w.joints()["lambda"], g[f"joint_lambda_{f}"]), f
    lam = g["joint_lambda_60"]
    assert lam[1, 0] != 0.0 and lam[3, 0] != 0.0 and lam[4, 0] == 0.0
