"""A frame that needs more than 1024 colours returns RBGPU_ERR_CAPACITY, on the grid path
from the same rbgpu_step call, and on the per-scene path (no host read-back inside a
frame) from the next call that waits for the world.  The condition comes from the input:
one dynamic slab carries 1,156 small spheres, so all 1,156 sphere-slab pairs share one
dynamic body.  Every kernel runs to its normal end (colours past the capacity are clamped
to the last one and flagged)."""
import numpy as np
import pytest

from rbps_amd import scenes

pytestmark = pytest.mark.gpu


def _hub_scene(seed=0):
    sb = scenes.SceneBuilder()
    scenes.add_ground_plane(sb)
    h = (12.0, 0.5, 12.0)
    slab = sb.body(scenes.DYNAMIC, 1000.0 * scenes.box_volume(h), tuple(1000.0 * x for x in scenes.box_inertia(h)),
                   (0.0, 0.5, 0.0))
    sb.collider(scenes.BOX, slab, h, restitution=0.0)
    r = 0.25
    for i in range(34):
        for k in range(34):
            b = sb.body(scenes.DYNAMIC, 1000.0 * scenes.sphere_volume(r),
                        tuple(1000.0 * x for x in scenes.sphere_inertia(r)),
                        ((i - 16.5) * 0.6, 1.0 + r - 0.002, (k - 16.5) * 0.6))
            sb.collider(scenes.SPHERE, b, (r,), restitution=0.0)
    return sb.build(f"hub_{seed}", 1.0 / 60.0, 4)


def _single(r=0.3, seed=0):
    sb = scenes.SceneBuilder()
    scenes.add_ground_plane(sb)
    b = sb.body(scenes.DYNAMIC, 1.0, pos=(0.0, 0.3 + 0.1 * seed, 0.0))
    sb.collider(scenes.SPHERE, b, (r,))
    return sb.build("one", 1.0 / 60.0, 4)


def test_grid_path_reports_colour_capacity_in_the_same_step():
    from rbps_amd import world
    sc = _hub_scene()
    print("case hub scene, grid path", flush=True)
    w = world.World(sc.timestep, sc.substeps).load(sc)
    with pytest.raises(world.RbgpuError, match="1024 colours"):
        w.step(1)


def test_scene_path_reports_colour_capacity_at_the_next_wait():
    from rbps_amd import world
    sc = scenes.batch([_hub_scene()] + [_single(seed=k) for k in range(15)])
    print("case hub scene in a 16-scene batch, per-scene path", flush=True)
    w = world.World(sc.timestep, sc.substeps).load(sc)
    w.set_option("scene_kernels", 1)
    w.step(1)   # queued; the error is read back after the frame
    with pytest.raises(world.RbgpuError, match="1024 colours"):
        w.synchronize()
    w.synchronize()   # reported once
    assert np.isfinite(w.state()["position"]).all()
