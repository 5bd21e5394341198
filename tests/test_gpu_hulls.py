"""Convex hulls on the GPU (wave-per-pair GJK / EPA / manifold and hull-plane kernels,
through the rbgpu C-ABI) against the FP32 oracle.

Bar: the kernels evaluate the oracle's float expressions in the same order (no FMA
contraction, IEEE division and square root), so single-pair manifolds, GJK/EPA statuses
and iteration counts are compared bit for bit; world states within 1e-5 relative (the
other parity tests' bar; they are bit-identical in practice).  The FP32 oracle itself is
pinned to the reference through its FP64 instantiation (tests/test_oracle_hulls.py).
"""
import numpy as np
import pytest

from oracle import orc
from rbps_amd import scenes
from tests.conftest import golden
from tests.helpers import state_close, unpack_hulls

pytestmark = pytest.mark.gpu
RTOL = 1e-5


def _pair_world(a, tfa, b, tfb, hulls):
    """One world holding both colliders of every pair at its poses (no stepping)."""
    from rbps_amd import world
    sb = scenes.SceneBuilder()
    for v, f in hulls:
        sb.hull(v, f)
    for i in range(len(a)):
        for rec, tf in ((a[i], tfa[i]), (b[i], tfb[i])):
            st = rec["shape"] == scenes.PLANE
            body = sb.body(scenes.STATIC if st else scenes.DYNAMIC, 0.0 if st else 1.0, pos=tuple(tf[:3]),
                           quat=tuple(tf[3:]))
            sb.collider(int(rec["shape"]), body, tuple(rec["dims"]), is_static=st, hull_index=int(rec["hull_index"]))
    sc = sb.build("pairs")
    w = world.World().load(sc)
    return w, np.arange(2 * len(a), dtype=np.uint32).reshape(-1, 2)


def _expected_surface_points(r):
    """p_a = p + n * (d * 0.5), p_b = p - n * (d * 0.5) in float (CollisionPipeline.cpp:59-104)."""
    n = np.float32(r["normal"]).astype(np.float32)
    half = (r["depth"].astype(np.float32) * np.float32(0.5)).astype(np.float32)
    p = r["points"].astype(np.float32)
    return p + n[None, :] * half[:, None], p - n[None, :] * half[:, None]


# GJK pass layouts (option gjk_rows): a pair per wavefront, per 16-lane row, per lane, per refilled row;
# GJK + EPA per 16-lane row (4), the same with row buffers most pairs outgrow, so that the full-size
# EPA pass takes them over (5)
LAYOUTS = [0, 1, 2, 3, 4, 5]


@pytest.mark.parametrize("layout", LAYOUTS)
def test_gpu_hull_pairs_bit_exact_vs_fp32_oracle(layout):
    g = golden("hull_pairs")
    hulls = unpack_hulls(g)
    a, b, tfa, tfb = g["a"], g["b"], g["tfa"], g["tfb"]
    print("case hull_pairs", len(a), "GJK layout", layout, flush=True)
    w, pairs = _pair_world(a, tfa, b, tfb, hulls)
    w.set_option("gjk_rows", layout)
    out = w.collide_pairs(pairs)
    n_gjk = 0
    for i in range(len(a)):
        r = orc.oracle_collide(a[i], tfa[i], b[i], tfb[i], fp32=True, hulls=hulls)
        assert out["n"][i] == r["n"], (i, out["n"][i], r["n"])
        if r["n"]:
            assert np.array_equal(out["normal"][i], np.float32(r["normal"]).astype(np.float64)), i
            pa, pb = _expected_surface_points(r)
            k = r["n"]
            assert np.array_equal(out["p_a"][i, :k], pa.astype(np.float64)), i
            assert np.array_equal(out["p_b"][i, :k], pb.astype(np.float64)), i
            assert np.array_equal(out["depth"][i, :k], r["depth"].astype(np.float32).astype(np.float64)), i
        if b[i]["shape"] != scenes.PLANE:
            s = orc.oracle_gjk_epa(a[i], tfa[i], b[i], tfb[i], fp32=True, hulls=hulls)
            assert out["gjk_status"][i] == s["gjk_status"] and out["gjk_iters"][i] == s["gjk_iters"], i
            if s["gjk_status"] == orc.GJK_INSIDE:
                assert out["epa_status"][i] == s["epa_status"] and out["epa_iters"][i] == s["epa_iters"], i
            n_gjk += 1
    assert n_gjk > 1000


@pytest.mark.parametrize("layout", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("make,frames", [
    (lambda: scenes.hull_zoo(), 40),
    (lambda: scenes.hull_pile(96, seed=3), 30),
    (lambda: scenes.hull_pile(96, seed=5, faces=False), 30),
])
def test_gpu_hull_worlds_match_fp32_colour_oracle(make, frames, layout):
    from rbps_amd import world
    sc = make()
    print("case", sc.name, "GJK layout", layout, flush=True)
    g = world.World(sc.timestep, sc.substeps).load(sc)
    g.set_option("gjk_rows", layout)
    o = orc.OracleWorld(sc.timestep, sc.substeps, fp32=True, order=orc.ORDER_COLOUR).load(sc)
    for f in range(frames):
        g.step(1)
        o.step(1)
        ok, worst = state_close(g.state(), o.state(), RTOL)
        assert ok, f"{sc.name} frame {f + 1}: state rel err {worst}"
        assert np.array_equal(g.sap_pairs(), o.pairs()), f"{sc.name} frame {f + 1}: SAP pairs"
        cg, co = g.contacts(), o.contacts()
        assert np.array_equal(cg["ids"], co["ids"]), f"{sc.name} frame {f + 1}: contact ids"
        assert np.allclose(cg["normal"], co["normal"], rtol=RTOL, atol=RTOL)
    st = g.stats()
    assert st["bins"][9] + st["bins"][10] > 0 and st["bins"][11] > 0   # GJK pairs and hull-plane pairs ran


@pytest.mark.parametrize("layout", LAYOUTS)
def test_gpu_gjk_cycling_pairs_match_oracle(layout):
    """Pairs on which FP32 GJK cycles (tests/golden/gjk_cycling_pairs.npz): the GPU ends
    the loop at the repeated state and reports what the capped run reports -- GJK_FAILED
    after 128 iterations, no contact -- as the FP32 oracle runs it to the cap."""
    g = golden("gjk_cycling_pairs")
    hulls = unpack_hulls(g)
    a, b, tfa, tfb = g["a"], g["b"], g["tfa"], g["tfb"]
    print("case gjk_cycling_pairs", len(a), "GJK layout", layout, flush=True)
    w, pairs = _pair_world(a, tfa, b, tfb, hulls)
    w.set_option("gjk_rows", layout)
    out = w.collide_pairs(pairs)
    for i in range(len(a)):
        s = orc.oracle_gjk_epa(a[i], tfa[i], b[i], tfb[i], fp32=True, hulls=hulls)
        assert out["gjk_status"][i] == s["gjk_status"] == orc.GJK_FAILED, i
        assert out["gjk_iters"][i] == s["gjk_iters"] == 128, i
        assert out["n"][i] == 0, i


@pytest.mark.parametrize("layout", [1, 4])
@pytest.mark.parametrize("mode", [1, 2])
def test_gpu_epa_guard_matches_fp32_oracle(mode, layout):
    """The EPA guard (RBGPU_OPT_EPA_GUARD; 1: the previous substep's normal, 2: and both
    shapes' face normals) on a settling pile of 256 hulls: the GPU equals the FP32
    colour-order oracle with the same guard every frame, through the frames where the guard
    replaces FP32 EPA results (the oracle counts them: at least one)."""
    from rbps_amd import world
    sc = scenes.hull_pile(256, seed=3)
    print("case hull pile 256, EPA guard", mode, "GJK layout", layout, flush=True)
    g = world.World(sc.timestep, sc.substeps).load(sc)
    g.set_option("epa_guard", mode)
    g.set_option("gjk_rows", layout)
    o = orc.OracleWorld(sc.timestep, sc.substeps, fp32=True, order=orc.ORDER_COLOUR).load(sc)
    o.set_epa_guard(mode)
    for f in range(60):
        g.step(1)
        o.step(1)
        ok, worst = state_close(g.state(), o.state(), 1e-5)
        assert ok, f"frame {f + 1}: {worst}"
    cg, co = g.contacts(), o.contacts()
    assert np.array_equal(cg["ids"], co["ids"])
    assert np.allclose(cg["normal"], co["normal"], rtol=1e-5, atol=1e-5)
    assert o.guard_stats()[0] > 0
