"""Contact mode on the GPU (capsule-capsule normal turned to point from A to B) against
the FP32 colour-order oracle in the same mode, and the reference mode beside it."""
import numpy as np
import pytest

from oracle import orc
from tests.helpers import state_close
from tests.test_oracle_contact_mode import _capsule_pile

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_capsule_pile_matches_oracle_in_contact_mode(mode):
    from rbps_amd import world
    sc = _capsule_pile()
    print("case capsule pile contact mode", mode, flush=True)
    g = world.World(sc.timestep, sc.substeps).load(sc)
    g.set_option("contact_mode", mode)
    o = orc.OracleWorld(sc.timestep, sc.substeps, fp32=True, order=orc.ORDER_COLOUR).load(sc)
    o.set_contact_mode(mode)
    for f in range(40):
        g.step(1)
        o.step(1)
        ok, worst = state_close(g.state(), o.state(), 1e-5)
        assert ok, f"mode {mode} frame {f + 1}: {worst}"
    assert g.stats()["bins"][4] > 0   # capsule-capsule pairs took part
