"""The reference's World tests (tests/rbps/API/test_world.cpp) run on the GPU
through the Python mirror of the World API, plus the FP64-reference golden of
two of those scenes (FP32 GPU vs FP64 reference, stated tolerance)."""
import numpy as np
import pytest

from rbps_amd import scenes
from tests.conftest import golden

pytestmark = pytest.mark.gpu


def W(dt=1 / 60, sub=20):
    from rbps_amd import world
    return world.World(dt, sub)


def add_dynamic_sphere(w, pos, radius=0.5, mass=1.0, vel=(0, 0, 0)):
    b = w.create_body(scenes.DYNAMIC, mass, position=pos, linear_velocity=vel)
    w.create_collider(scenes.SPHERE, (radius,), body_id=b)
    return b


def add_static_box(w, pos, half=(10, 2, 10)):
    b = w.create_body(scenes.STATIC, 0.0, position=pos)
    w.create_collider(scenes.BOX, half, body_id=b)
    return b


def test_free_fall_moves_down_and_speeds_up():          # test_world.cpp:69-95
    w = W()
    b = add_dynamic_sphere(w, (0, 10, 0))
    w.step(1)
    y1, v1 = w.state()["position"][b, 1], w.state()["linear_velocity"][b, 1]
    w.step(1)
    assert y1 < 10 and w.state()["linear_velocity"][b, 1] < v1


def test_static_body_does_not_move():                   # test_world.cpp:97-107
    w = W()
    g = add_static_box(w, (0, 0, 0))
    w.step(10)
    assert np.abs(w.state()["position"][g]).max() < 1e-6


def test_sphere_lands_on_ground():                      # test_world.cpp:109-124
    w = W()
    add_static_box(w, (0, -3, 0))
    b = add_dynamic_sphere(w, (0, 1, 0), 0.5)
    w.step(180)
    st = w.state()
    assert np.linalg.norm(st["linear_velocity"][b]) < 2.0 and st["position"][b, 1] > -4.0
    ref = golden("world_api_scenes")
    assert np.abs(st["position"] - ref["land_position"]).max() < 1e-3   # FP32 GPU vs FP64 reference


def test_sphere_does_not_tunnel():                      # test_world.cpp:126-137
    w = W(1 / 60, 25)
    add_static_box(w, (0, -3, 0))
    b = add_dynamic_sphere(w, (0, 2, 0), 0.5)
    w.step(300)
    assert w.state()["position"][b, 1] > -5.0


def test_two_spheres_collide_and_separate():            # test_world.cpp:139-156
    w = W()
    a = add_dynamic_sphere(w, (-2, 0, 0), 0.5, 1.0, (5, 0, 0))
    b = add_dynamic_sphere(w, (2, 0, 0), 0.5, 1.0, (-5, 0, 0))
    w.step(30)
    v = w.state()["linear_velocity"]
    assert v[a, 0] < 0.0 or v[b, 0] > 0.0
    ref = golden("world_api_scenes")
    # FP32 vs the FP64 reference after a 10 m/s impact: the FP32 oracle shows the same 2.09e-3 m gap
    assert np.abs(w.state()["position"] - ref["bounce_position"]).max() < 5e-3


def test_momentum_conserved_approximately():            # test_world.cpp:158-176
    w = W()
    a = add_dynamic_sphere(w, (-2, 0, 0), 0.5, 1.0, (3, 0, 0))
    b = add_dynamic_sphere(w, (2, 0, 0), 0.5, 1.0, (0, 0, 0))
    before = 3.0
    w.step(60)
    v = w.state()["linear_velocity"]
    after = v[a, 0] + v[b, 0]
    assert abs(after - before) <= abs(before) * 0.2 + 0.1


def test_contacts_bounded():                            # test_world.cpp:165-176
    w = W()
    add_static_box(w, (0, -1, 0))
    add_dynamic_sphere(w, (0, 0, 0), 0.5)
    w.step(10)
    assert w.stats()["contacts"] <= 16


def test_caller_forces_match_fp32_colour_oracle():
    """Caller-written force / torque (rbgpu_set_body_forces; Body.hpp:52-69) on the pile of
    tests/golden/world_forces.npz: the GPU equals the FP32 colour-order oracle given the same
    loads (states within 1e-5 relative, bit-identical in practice), before the first step
    and again rewritten after 10 frames (the built world's device copy)."""
    from oracle import orc
    from tests.helpers import state_close
    g = golden("world_forces")
    sc = scenes.Scene("forces", g["bodies"], g["colliders"], float(g["timestep"]), int(g["substeps"]))
    w = W(sc.timestep, sc.substeps).load(sc)
    o = orc.OracleWorld(sc.timestep, sc.substeps, fp32=True, order=orc.ORDER_COLOUR).load(sc)
    w.set_forces(g["force"], g["torque"])
    o.set_forces(g["force"], g["torque"])
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    assert np.array_equal(w.forces()["force"], f32(g["force"]))
    for f in range(30):
        if f == 10:
            w.set_forces(-0.5 * g["force"], None)
            o.set_forces(-0.5 * g["force"], None)
        w.step(1)
        o.step(1)
        ok, worst = state_close(w.state(), o.state(), 1e-5)
        assert ok, f"frame {f + 1}: {worst}"
    assert np.array_equal(w.forces()["torque"], f32(g["torque"]))
