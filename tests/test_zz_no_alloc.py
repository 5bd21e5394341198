"""rbgpu_collide_pairs_noalloc (fixed device staging, no allocation) against
rbgpu_collide_pairs on the same inputs, bit for bit: two calls in a row, the second with
fewer pairs, so stale staging from the first call would show; then a call beyond the
16384-pair limit, which the host rejects before any GPU work."""
import numpy as np
import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def _inputs(g, sel):
    return (g["kind_a"][sel], g["dims_a"][sel], g["tfa"][sel], g["kind_b"][sel], g["dims_b"][sel], g["tfb"][sel])


def test_noalloc_equals_allocating_entry_on_two_calls():
    from rbps_amd import world
    g = golden("narrowphase_pairs")
    n = len(g["kind_a"])
    big = np.concatenate([np.arange(n)] * 8)            # 12,000 pairs
    small = np.arange(n)[::-1][: n // 2 + 7]             # 757 pairs, reversed order
    for sel in (big, small):
        print(f"case noalloc vs allocating, {len(sel)} pairs", flush=True)
        args = _inputs(g, sel)
        ref = world.collide_pairs(*args)
        got = world.collide_pairs(*args, noalloc=True)
        assert got.shape == ref.shape == (len(sel), 20)
        assert np.array_equal(got, ref)
        assert (got[:, 0] > 0).sum() > len(sel) // 10   # the sample has contacts


def test_noalloc_rejects_more_than_the_staging_holds():
    from rbps_amd import world
    g = golden("narrowphase_pairs")
    sel = np.resize(np.arange(len(g["kind_a"])), world.NOALLOC_MAX_PAIRS + 1)
    with pytest.raises(world.RbgpuError, match="16384"):
        world.collide_pairs(*_inputs(g, sel), noalloc=True)
