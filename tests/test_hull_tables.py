"""The product's hull preparation (rbgpu_hull_tables / rbgpu_hull_mass_properties, host
code of librbgpu.so; no GPU needed) against the reference's tables and mass properties
held in tests/golden/hull_pairs.npz, and its input checks."""
import numpy as np
import pytest

from rbps_amd import scenes
from tests.conftest import golden
from tests.helpers import unpack_hulls

TABLES = ("pn", "poff", "pidx", "v2p_o", "v2p_i", "v2n_o", "v2n_i", "p2p_o", "p2p_i")


def test_product_hull_tables_match_reference():
    from rbps_amd import world
    g = golden("hull_pairs")
    for i, (v, f) in enumerate(unpack_hulls(g)):
        t = world.hull_tables(v, f)
        for k in TABLES:
            assert np.array_equal(t[k], g[f"t{i}_{k}"]), (i, k)
        vol, inert = world.hull_mass_properties(v, f)
        assert vol == g[f"t{i}_mass"][0] and np.array_equal(inert, g[f"t{i}_mass"][1:]), i


def test_product_hull_limits():
    from rbps_amd import world
    rng = np.random.default_rng(0)
    with pytest.raises(world.RbgpuError):   # more than 256 vertices
        world.hull_tables(rng.standard_normal((300, 3)))
    with pytest.raises(world.RbgpuError):   # triangle index out of range
        world.hull_tables(np.eye(3), np.array([[0, 1, 3]]))
    v, f = scenes.hex_prism_hull()
    assert len(world.hull_tables(v, f)["pn"]) == 8
