"""Batched scenes on the GPU: one world holding many independent scenes.  Each scene
must evolve bit for bit as in a world of its own — checked against the batched FP32
oracle, against single-scene GPU worlds, and (at the full 4,096-scene size) for
finiteness and the absence of cross-scene pairs."""
import numpy as np
import pytest

from oracle import orc
from rbps_amd import scenes as S
from tests.helpers import state_close

pytestmark = pytest.mark.gpu


def _gpu(sc, mode=1, contact_mode=0):
    from rbps_amd import world
    w = world.World(sc.timestep, sc.substeps).load(sc)
    w.set_option("joint_mode", mode)
    w.set_option("contact_mode", contact_mode)
    return w


def _parts():
    parts = [S.stack_chain(int(n), seed=5 + k) for k, n in enumerate((40, 70, 33, 52))]
    parts.append(S.pile_mixed_dense(seed=3, nx=4, ny=3, nz=4))
    return parts


def test_gpu_batch_matches_batched_oracle_and_single_worlds():
    parts = _parts()
    bt = S.batch(parts)
    print("case batch of", len(parts), flush=True)
    g = _gpu(bt)
    o = orc.OracleWorld(bt.timestep, bt.substeps, fp32=True, order=orc.ORDER_COLOUR).load(bt)
    o.set_joint_mode(1)
    singles = [_gpu(p) for p in parts]
    for f in range(20):
        g.step(1)
        o.step(1)
        for s in singles:
            s.step(1)
        ok, worst = state_close(g.state(), o.state(), 1e-5)
        assert ok, f"frame {f + 1}: batch vs oracle {worst}"
    sb, off = g.state(), 0
    for p, s in zip(parts, singles):
        ss, n = s.state(), len(p.bodies)
        for k in ss:
            assert np.array_equal(sb[k][off:off + n], ss[k]), (p.name, k)
        off += n


def test_gpu_full_batch_runs_without_cross_scene_pairs():
    bt = S.batched_stack_chains(4096)
    print("case batched_stack_chains_4096 bodies", len(bt.bodies), flush=True)
    g = _gpu(bt)
    g.step(3)
    st = g.state()
    assert all(np.isfinite(v).all() for v in st.values())
    pr = g.sap_pairs()
    b = bt.colliders["body_id"][pr]
    assert (bt.bodies["scene"][b[:, 0]] == bt.bodies["scene"][b[:, 1]]).all()
    assert len(pr) > 0


@pytest.mark.parametrize("layout", [1, 2, 3, 4, 5])
def test_gpu_combined_batch_matches_fp32_oracle_and_single_worlds(layout):
    """Hulls from a shared library, boxes and ragdolls (revolute + spherical joints with
    60 deg limits) in three uneven scenes; 80 frames at dt = 1/120 s, 8 substeps, so the
    pile has landed on the ragdolls (hull-box, hull-capsule, hull-sphere contacts), in the
    contact mode that orients the capsule-capsule normal like the other pairs; GJK pass
    with a pair per 16-lane row (1), per lane (2), per refilled row (3), or GJK and EPA per
    16-lane row (4; 5 with row buffers that hand most overlapping pairs to the full EPA pass)."""
    lib = S.hull_library()
    parts = [S.combined_scene(n, seed=40 + k, library=lib) for k, n in enumerate((64, 90, 130))]
    bt = S.batch(parts)
    assert len(bt.hulls) == len(lib)
    print("case combined batch bodies", len(bt.bodies), "joints", len(bt.joints), "GJK layout", layout, flush=True)
    g = _gpu(bt, contact_mode=1)
    g.set_option("gjk_rows", layout)
    g.gjk_totals()   # counts from here on
    o = orc.OracleWorld(bt.timestep, bt.substeps, fp32=True, order=orc.ORDER_COLOUR).load(bt)
    o.set_joint_mode(1)
    o.set_contact_mode(1)
    singles = [_gpu(p, contact_mode=1) for p in parts]
    hull_pairs = 0
    for f in range(80):
        g.step(1)
        o.step(1)
        for s in singles:
            s.step(1)
        ok, worst = state_close(g.state(), o.state(), 1e-5)
        assert ok, f"frame {f + 1}: combined batch vs oracle {worst}"
        bins = g.stats()["bins"]
        hull_pairs = max(hull_pairs, bins[9] + bins[10])
    assert hull_pairs > 0
    tot = g.gjk_totals()
    print("GJK totals", tot)
    if layout == 5:   # the small row buffers hand pairs to the full-size EPA pass
        assert tot["handed_over"] > 0
    sb, off = g.state(), 0
    for p, s in zip(parts, singles):
        ss, n = s.state(), len(p.bodies)
        for k in ss:
            assert np.array_equal(sb[k][off:off + n], ss[k]), (p.name, k)
        off += n


@pytest.mark.parametrize("threads", [64, 256])
def test_gpu_scene_kernels_equal_grid_path(threads):
    """Per-scene colouring and solve (one workgroup per scene, no host round trip) against
    the grid-wide rounds and per-colour launches on the same batch: the same colour for
    every pair, the same round and colour counts, bit-identical states, every frame."""
    lib = S.hull_library()
    parts = _parts() + [S.combined_scene(n, seed=80 + k, library=lib) for k, n in enumerate((64, 120))]
    bt = S.batch(parts)
    print("case scene kernels threads", threads, "scenes", len(parts), flush=True)
    a, b = _gpu(bt, contact_mode=1), _gpu(bt, contact_mode=1)
    a.set_option("scene_kernels", 0)
    b.set_option("scene_kernels", 1)
    b.set_option("scene_threads", threads)
    for f in range(70):
        a.step(1)
        b.step(1)
        sa, sb = a.state(), b.state()
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), (f + 1, k)
        pa, ca = a.pair_colours()
        pb, cb = b.pair_colours()
        assert np.array_equal(pa, pb) and np.array_equal(ca, cb), f + 1
        ta, tb = a.stats(), b.stats()
        assert (ta["colours"], ta["colour_rounds"], ta["active_pairs"]) == \
               (tb["colours"], tb["colour_rounds"], tb["active_pairs"]), (f + 1, ta, tb)
    assert tb["active_pairs"] > 0


def test_gpu_substep_graph_equals_direct_launches():
    """The per-scene substep captured once per frame in a hipGraph and replayed for every
    substep against the same launches issued directly: bit-identical states each frame."""
    lib = S.hull_library()
    parts = [S.combined_scene(n, seed=90 + k, library=lib) for k, n in enumerate((64, 80, 100, 72))] * 5
    bt = S.batch(parts)
    print("case substep graph scenes", len(parts), flush=True)
    a, b = _gpu(bt), _gpu(bt)
    for w, g in ((a, 0), (b, 1)):
        w.set_option("scene_kernels", 1)
        w.set_option("graph", g)
    for f in range(60):
        a.step(1)
        b.step(1)
        sa, sb = a.state(), b.state()
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), (f + 1, k)
    assert a.stats()["graph_frames"] == 0 and b.stats()["graph_frames"] == 60
    assert b.stats()["scene_kernels"] == 1


@pytest.mark.parametrize("scene_kernels,graph,side", [(0, 0, 2), (1, 0, 1), (1, 1, 3), (1, 1, 1)])
def test_gpu_side_stream_equals_one_stream(scene_kernels, graph, side):
    """The narrowphase bins other than GJK/EPA on a second stream beside the GJK chain
    (forked before it, after its rejection tests or after the GJK pass), against every bin on
    the world's stream (grid path, per-scene path, per-scene path in a hipGraph):
    bit-identical states and the same contact counts every frame."""
    lib = S.hull_library()
    parts = _parts() + [S.combined_scene(n, seed=70 + k, library=lib) for k, n in enumerate((64, 96, 128))] * 3
    bt = S.batch(parts)
    print("case side stream", side, "scene_kernels", scene_kernels, "graph", graph, flush=True)
    a, b = _gpu(bt, contact_mode=1), _gpu(bt, contact_mode=1)
    for w, v in ((a, 0), (b, side)):
        w.set_option("scene_kernels", scene_kernels)
        w.set_option("graph", graph)
        w.set_option("side_stream", v)
    assert b.option("side_stream") == side and a.option("side_stream") == 0
    forked = 0   # frames with pairs both in the GJK bins and in the others
    for f in range(60):
        a.step(1)
        b.step(1)
        sa, sb = a.state(), b.state()
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), (f + 1, k)
        ta, tb = a.stats(), b.stats()
        assert (ta["contacts"], ta["active_pairs"]) == (tb["contacts"], tb["active_pairs"]), (f + 1, ta, tb)
        gjk = tb["bins"][9] + tb["bins"][10]
        forked += gjk > 0 and sum(tb["bins"]) > gjk
    assert tb["contacts"] > 0 and forked > 30


def test_gpu_gjk_rejection_and_carried_hints_change_nothing():
    """The exact rejection tests in front of GJK, with and without the separating-axis hints
    carried across each frame's new pair list, against every hull pair through GJK:
    bit-identical states every frame; the hints only take pairs away from GJK (fewer GJK
    iterations with the cull, fewer again with the carried hints)."""
    lib = S.hull_library()
    parts = [S.combined_scene(n, seed=60 + k, library=lib) for k, n in enumerate((64, 96, 128, 80))] * 2
    bt = S.batch(parts)
    print("case gjk rejection scenes", len(parts), flush=True)
    ws = [_gpu(bt, contact_mode=1) for _ in range(3)]
    for w, cull, carry in zip(ws, (0, 1, 1), (0, 0, 1)):
        w.set_option("gjk_cull", cull)
        w.set_option("axis_carry", carry)
        w.gjk_totals(reset=True)
    for f in range(50):
        for w in ws:
            w.step(1)
        s0 = ws[0].state()
        for w in ws[1:]:
            sw = w.state()
            for k in s0:
                assert np.array_equal(s0[k], sw[k]), (f + 1, k)
    it = [w.gjk_totals()["gjk_iters"] for w in ws]
    print("GJK iterations: all pairs, cull, cull + carried hints", it, flush=True)
    assert it[0] > 0 and it[2] <= it[1] <= it[0]


def test_gpu_autotune_keeps_the_trajectory():
    """The autotuner times result-neutral options over real frames: the tuned world's
    states equal an untuned world's after the same number of frames, and the chosen values
    are among the candidates."""
    lib = S.hull_library()
    parts = [S.combined_scene(n, seed=95 + k, library=lib) for k, n in enumerate((64, 96, 128, 70))] * 5
    bt = S.batch(parts)
    print("case autotune scenes", len(parts), flush=True)
    a, b = _gpu(bt), _gpu(bt)
    a.step(20)
    b.step(20)
    rep = b.autotune(1)
    assert rep["scene_kernels"] == 1 and rep["scene_threads"] in (64, 128, 256) and rep["graph"] in (0, 1)
    assert rep["trials"] == 10 and rep["gjk_grid"] == 4096 and rep["epa_grid"] == 8192   # grids are not tuned
    a.step(rep["frames"])
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k


def test_gpu_world_pool_equals_worlds_alone():
    """rbps_amd.world.WorldPool (multi-world.md:40-50): 8 piles of equal size stepped
    together equal the same piles stepped as worlds of their own, bit for bit; stacked
    states come back as (worlds, bodies, ...) arrays; a state written per world reaches
    only that world."""
    from rbps_amd import world
    parts = [S.pile_small(nx=3, ny=3, nz=2, seed=100 + k) for k in range(8)]
    pool = world.WorldPool(parts)
    alone = [_gpu(p) for p in parts]
    pool.parallel_step(frames=20)
    for w in alone:
        w.step(20)
    for st, w in zip(pool.states(), alone):
        ws = w.state()
        for k in ws:
            assert np.array_equal(st[k], ws[k]), k
    stk = pool.stacked()
    assert stk["position"].shape == (8, len(parts[0].bodies), 3)
    sts = pool.states()
    sts[5]["linear_velocity"] = sts[5]["linear_velocity"] + np.array([0.0, 3.0, 0.0])
    pool.set_states(sts)
    ws = alone[5].state()
    ws["linear_velocity"] = ws["linear_velocity"] + np.array([0.0, 3.0, 0.0])
    alone[5].set_state(ws)
    pool.parallel_step(frames=10)
    for w in alone:
        w.step(10)
    for st, w in zip(pool.states(), alone):
        ws = w.state()
        for k in ws:
            assert np.array_equal(st[k], ws[k]), k


def test_gpu_gjk_candidate_order_keeps_results():
    """RBGPU_OPT_GJK_SORT (GJK candidates ordered by their previous iteration count) and the
    GJK layouts change only which wavefront runs which pair: the combined batch's states are
    bit-identical with the sort on and off, for the row layout, the refilled rows and GJK + EPA
    per row (with row buffers of both sizes)."""
    lib = S.hull_library()
    parts = [S.combined_scene(n, seed=90 + k, library=lib) for k, n in enumerate((80, 120))]
    bt = S.batch(parts)
    print("case GJK candidate order", len(bt.bodies), "bodies", flush=True)
    runs = []
    for sort, layout in ((0, 1), (1, 1), (1, 3), (0, 4), (1, 5)):
        g = _gpu(bt, contact_mode=1)
        g.set_option("gjk_sort", sort)
        g.set_option("gjk_rows", layout)
        g.step(60)
        runs.append(g.state())
    for st in runs[1:]:
        for k in st:
            assert np.array_equal(st[k], runs[0][k]), k
