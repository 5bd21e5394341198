"""Batches (many independent scenes in one world) in the oracle: every scene of a
batch evolves bit for bit as it does in a world of its own (colour order, FP32 and
FP64), because the broadphase sorts by (scene, value), planes pair within their
scene, and the colouring hashes scene-local pair ids."""
import numpy as np
import pytest

from oracle import orc
from rbps_amd import scenes as S


@pytest.mark.parametrize("fp32", [True, False])
def test_batch_scenes_equal_single_worlds(fp32):
    parts = [S.stack_chain(int(n), seed=5 + k) for k, n in enumerate((40, 70, 33, 52))]
    parts.append(S.pile_mixed_dense(seed=3, nx=4, ny=3, nz=4))
    bt = S.batch(parts)
    ob = orc.OracleWorld(bt.timestep, bt.substeps, fp32=fp32, order=orc.ORDER_COLOUR).load(bt)
    ob.set_joint_mode(1)
    singles = []
    for p in parts:
        o = orc.OracleWorld(bt.timestep, bt.substeps, fp32=fp32, order=orc.ORDER_COLOUR).load(p)
        o.set_joint_mode(1)
        singles.append(o)
    for f in range(25):
        ob.step(1)
        for o in singles:
            o.step(1)
    sb, off = ob.state(), 0
    for p, o in zip(parts, singles):
        so, n = o.state(), len(p.bodies)
        for k in so:
            assert np.array_equal(sb[k][off:off + n], so[k]), (p.name, k)
        off += n
    # the batch holds every scene's pairs and nothing across scenes
    bodies = bt.colliders["body_id"][ob.solve_pairs()["colliders"]]
    assert (bt.bodies["scene"][bodies[:, 0]] == bt.bodies["scene"][bodies[:, 1]]).all()


def test_combined_scenes_generator_and_batch_equal_single_worlds():
    """The combined workload's generator (uneven counts summing to the total, a shared hull
    library uploaded once, 8-body ragdolls with 7 joints) and its batch in the FP32 colour
    oracle: each scene bit for bit as alone, through the landing of the pile (70 frames)."""
    c = S.combined_counts()
    assert c.sum() == 262144 and c.min() >= 64 and c.max() <= 512 and len(c) == 1024
    lib = S.hull_library()
    parts = [S.combined_scene(n, seed=60 + k, library=lib) for k, n in enumerate((64, 100))]
    bt = S.batch(parts)
    assert len(bt.hulls) == len(lib)
    for p in parts:
        n_rag = int(round(p.n_dynamic * 0.375 / 8.0))
        assert len(p.joints) == 7 * n_rag and p.timestep == 1.0 / 120.0 and p.substeps == 8
        assert (p.joints["limited"] == 1).all()
    ob = orc.OracleWorld(bt.timestep, bt.substeps, fp32=True, order=orc.ORDER_COLOUR).load(bt)
    ob.set_joint_mode(1)
    ob.set_contact_mode(1)
    singles = []
    for p in parts:
        o = orc.OracleWorld(bt.timestep, bt.substeps, fp32=True, order=orc.ORDER_COLOUR).load(p)
        o.set_joint_mode(1)
        o.set_contact_mode(1)
        singles.append(o)
    for f in range(70):
        ob.step(1)
        for o in singles:
            o.step(1)
    sb, off = ob.state(), 0
    for p, o in zip(parts, singles):
        so, n = o.state(), len(p.bodies)
        for k in so:
            assert np.array_equal(sb[k][off:off + n], so[k]), (p.name, k)
        off += n
    assert all(np.isfinite(v).all() for v in sb.values())
