"""The C-ABI library loads and exports every entry point include/rbgpu.h declares
(no GPU work), and the scene generators' mass properties match closed forms."""
import math
import os
import re
import subprocess
import sys

import numpy as np

from rbps_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "rbgpu.h")).read()
    return sorted(set(re.findall(r"\b(rbgpu_\w+)\s*\(", src)))


def test_library_exports_all_declared_symbols():
    from rbps_amd import world
    if not os.path.exists(world.LIB_PATH):
        subprocess.run([sys.executable, os.path.join(ROOT, "rbps_amd", "build.py")], check=True, capture_output=True)
    lib = world.lib()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(world.EXPORTS) == set(names)


def test_invalid_arguments_fail_loudly():
    """Errors come back as status codes with a message, no GPU call needed."""
    import ctypes as C
    from rbps_amd import world
    lib = world.lib()
    h = C.c_void_p()
    assert lib.rbgpu_world_create(1 / 60, 0, C.byref(h)) != 0
    assert b"substeps" in lib.rbgpu_last_error()


def test_noalloc_narrowphase_rejects_oversized_calls_on_the_host():
    """The allocation-free narrowphase refuses more pairs than its fixed staging holds
    before any GPU work (so this runs without a GPU)."""
    import ctypes as C
    from rbps_amd import world
    lib = world.lib()
    n = world.NOALLOC_MAX_PAIRS + 1
    k = np.zeros(n, np.int32)
    d = np.zeros(4 * n)
    t = np.zeros(7 * n)
    out = np.zeros(20 * n)
    P = lambda a, ty: a.ctypes.data_as(C.POINTER(ty))   # noqa: E731
    rc = lib.rbgpu_collide_pairs_noalloc(n, P(k, C.c_int32), P(d, C.c_double), P(t, C.c_double), P(k, C.c_int32),
                                         P(d, C.c_double), P(t, C.c_double), P(out, C.c_double))
    assert rc == -3 and b"16384" in lib.rbgpu_last_error()


def test_mass_properties_closed_forms():
    r = 0.5
    assert math.isclose(scenes.sphere_volume(r), 4 / 3 * math.pi * r ** 3)
    assert math.isclose(scenes.sphere_inertia(r)[0], 0.4 * scenes.sphere_volume(r) * r * r)
    h = (0.5, 0.25, 0.1)
    m = scenes.box_volume(h)
    assert math.isclose(m, 8 * 0.5 * 0.25 * 0.1)
    assert math.isclose(scenes.box_inertia(h)[0], m / 12 * ((0.5) ** 2 + (0.2) ** 2))
    # capsule -> sphere as the half height -> 0
    assert math.isclose(scenes.capsule_volume(0.0, r), scenes.sphere_volume(r))
    assert math.isclose(scenes.capsule_inertia(0.0, r)[1], scenes.sphere_inertia(r)[1])


def test_scene_generators_shapes():
    sc = scenes.pile_mixed()
    assert sc.n_dynamic == 65536 and sc.substeps == 10
    kinds = sc.colliders["shape"]
    assert (kinds == scenes.PLANE).sum() == 1
    assert abs((kinds == scenes.BOX).sum() - 65536 / 3) < 2
    small = scenes.pile_small()
    assert small.n_dynamic == 1000 and small.substeps == 20
