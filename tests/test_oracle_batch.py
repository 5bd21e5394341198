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
