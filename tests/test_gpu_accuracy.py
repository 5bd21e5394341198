"""GPU position error against the reference engine over 300 frames on a well-posed scene.

The scene (scenes.accuracy_batch) is the small plain configuration's 1,000 bodies (unit
boxes and spheres, density 1000, mu 0.5, e 0, dt 1/60, 20 substeps) as 50 independent
layers of 20 bodies, each body resting on the plane alone: boxes start flat and slide to a
stop, spheres drop.  No two contact pairs share a dynamic body, so the GPU's colour order
cannot differ from the reference's sequential order (tests/test_oracle_colour_order.py);
what remains is FP32 against the reference's FP64.  The reference trajectory is the
fixture tests/golden/accuracy_resting_layers.npz, written by the reference engine compiled
from its sources (one World per layer).

Bars (stated here, measured on the FP32 oracle, which the GPU equals bit for bit):
  * frames 1..20: the GPU equals the FP32 colour-order oracle bit for bit (batched path,
    one workgroup per scene);
  * spheres: within 2e-5 m of the reference at every checked frame (1, 10, 30, 100, 300);
  * median body: within 1e-4 m at frame 300;
  * every body: within 5e-3 m at frame 300.  A flat box sliding to a stop ends a few
    FP32 rounding steps of friction apart from FP64 (up to ~2e-3 m here); the error is
    reported per frame rather than hidden.
"""
import numpy as np
import pytest

from oracle import orc
from rbps_amd import scenes
from tests.conftest import golden

pytestmark = pytest.mark.gpu


def test_gpu_accuracy_vs_reference_300_frames():
    from rbps_amd import world
    fx = golden("accuracy_resting_layers")
    sc = scenes.accuracy_batch()
    assert np.array_equal(sc.bodies, fx["bodies"]) and np.array_equal(sc.colliders, fx["colliders"])
    print("case accuracy batch 50 x 20 bodies, 300 frames", flush=True)
    g = world.World(sc.timestep, sc.substeps).load(sc)
    o = orc.OracleWorld(sc.timestep, sc.substeps, fp32=True, order=orc.ORDER_COLOUR).load(sc)
    for f in range(20):
        g.step(1)
        o.step(1)
        sg, so = g.state(), o.state()
        for k in sg:
            assert np.array_equal(sg[k], so[k]), f"frame {f + 1} {k}"
    assert g.stats()["scene_kernels"] == 1
    dyn = sc.bodies["type"] == scenes.DYNAMIC
    box = np.zeros(len(sc.bodies), bool)
    box[sc.colliders["body_id"][sc.colliders["shape"] == scenes.BOX]] = True
    done = 20
    curve = {}
    e = None
    for f in (int(x) for x in fx["frames"]):
        if f > done:
            g.step(f - done)
            done = f
        if f < 20:
            continue
        e = np.linalg.norm(g.state()["position"] - fx[f"position_{f}"], axis=1)
        curve[f] = (float(e[dyn].max()), float(np.median(e[dyn])), float(e[dyn & ~box].max()))
        assert e[dyn & ~box].max() <= 2e-5, f"frame {f}: sphere error {e[dyn & ~box].max()}"
    print("frame: (max, median, spheres max) m", curve)
    assert np.median(e[dyn]) <= 1e-4
    assert e[dyn].max() <= 5e-3
