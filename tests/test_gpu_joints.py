"""Joints on the GPU (through the rbgpu C-ABI) against the FP32 colour-order oracle.

Bar: bit-identical states, joint lambdas and joint positions in practice (the same
float expressions, including this build's own sin/cos/atan2, no FMA contraction);
asserted within 1e-5 relative.  NaN appears where the oracle has NaN: in joint mode 0
the reference's own scheme turns several of these joints NaN (the oracle reproduces
the reference bit for bit there, tests/test_oracle_joints.py).
"""
import numpy as np
import pytest

from oracle import orc
from rbps_amd import scenes
from tests.helpers import state_close

pytestmark = pytest.mark.gpu
RTOL = 1e-5


def _pair(sc, mode):
    from rbps_amd import world
    g = world.World(sc.timestep, sc.substeps).load(sc)
    g.set_option("joint_mode", mode)
    o = orc.OracleWorld(sc.timestep, sc.substeps, fp32=True, order=orc.ORDER_COLOUR).load(sc)
    o.set_joint_mode(mode)
    return g, o


def _nan_safe_close(a, b):
    for k in a:
        if not np.array_equal(np.isnan(a[k]), np.isnan(b[k])):
            return False, f"NaN pattern of {k}"
    fa = {k: np.nan_to_num(v) for k, v in a.items()}
    fb = {k: np.nan_to_num(v) for k, v in b.items()}
    return state_close(fa, fb, RTOL)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("make,frames", [
    (lambda: scenes.joint_zoo(), 40),
    (lambda: scenes.planar_joints(), 40),
    (lambda: scenes.stack_chain(64, seed=3), 20),
])
def test_gpu_joints_match_fp32_colour_oracle(make, frames, mode):
    sc = make()
    print("case", sc.name, "mode", mode, flush=True)
    g, o = _pair(sc, mode)
    for f in range(frames):
        g.step(1)
        o.step(1)
        ok, worst = _nan_safe_close(g.state(), o.state())
        assert ok, f"{sc.name} mode {mode} frame {f + 1}: state {worst}"
        jg, jo = g.joints(), o.joints()
        ok, worst = _nan_safe_close({k: jg[k] for k in ("current_position", "target_position", "lambda")},
                                    {k: jo[k] for k in ("current_position", "target_position", "lambda")})
        assert ok, f"{sc.name} mode {mode} frame {f + 1}: joints {worst}"
        assert np.array_equal(jg["colour"], jo["colour"])


def test_gpu_joint_targets_and_api_errors():
    """Driving a joint through rbgpu_set_joint_targets matches the oracle's field write;
    invalid joint descriptors are rejected."""
    from rbps_amd import world
    sc = scenes.planar_joints()
    g, o = _pair(sc, 1)
    g.step(5)
    o.step(5)
    g.set_joint_targets([3], target_position=[-0.4])
    o.set_joint_targets([3], target_position=[-0.4])
    g.step(10)
    o.step(10)
    ok, worst = state_close(g.state(), o.state(), RTOL)
    assert ok, worst
    bad = scenes.SceneBuilder()
    b0 = bad.body()
    bad.joint(scenes.REVOLUTE, b0, b0)
    with pytest.raises(world.RbgpuError):   # body out of range
        world.World().add_joints(np.array(bad.j, dtype=scenes.JOINT_DTYPE))
    w = world.World()
    w.add_bodies(bad.build("x").bodies)
    with pytest.raises(world.RbgpuError):   # a joint needs two different bodies
        w.add_joints(np.array(bad.j, dtype=scenes.JOINT_DTYPE))


@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_constraint_rows_match_fp32_colour_oracle(mode):
    """Constraint rows (World::create_constraint) created between joints and written after
    loading: GPU = FP32 colour-order oracle (states, lambdas of joints and rows); a row
    keeps the direction and magnitude written to it; writing a joint's row is refused."""
    from rbps_amd import world
    sc = scenes.constraint_rows()
    print("case constraint_rows mode", mode, flush=True)
    g, o = _pair(sc, mode)
    ids = np.array([r[0] for r in sc.meta["rows"]], np.uint32)
    dirs, mags = np.array([r[1] for r in sc.meta["rows"]]), np.array([r[2] for r in sc.meta["rows"]])
    g.set_rows(ids, dirs, mags)
    o.set_rows(ids, dirs, mags)
    for f in range(40):
        g.step(1)
        o.step(1)
        ok, worst = _nan_safe_close(g.state(), o.state())
        assert ok, f"mode {mode} frame {f + 1}: state {worst}"
        ok, worst = _nan_safe_close({"lambda": g.joints()["lambda"]}, {"lambda": o.joints()["lambda"]})
        assert ok, f"mode {mode} frame {f + 1}: lambdas {worst}"
    rows = g.constraint_rows()
    assert np.allclose(rows["direction"][ids, 0], dirs) and np.allclose(rows["magnitude"][ids, 0], mags)
    assert (g.joints()["lambda"][ids, 0] != 0.0).all()
    with pytest.raises(world.RbgpuError, match="not a constraint row"):
        g.set_rows([0], [(0.0, 1.0, 0.0)], [1.0])
