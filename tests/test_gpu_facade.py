"""The source-compatible rbps::World (include/rbps/API/World.hpp -> include/rbps_amd/World.hpp)
compiled against the C-ABI library: the reference's World test cases ported with the
rest of the public surface (tests/cpp/test_world_facade.cpp), the WorldPool over the batch
(tests/cpp/test_world_pool.cpp), and the reference's own tests/rbps/API/test_world.cpp
compiled unchanged (oracle/_ref/test_world, built where the reference sources are present)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "rbps_amd")
REF_TEST = os.path.join(ROOT, "oracle", "_ref", "test_world")


def _build(tmp_path, name="test_world_facade"):
    exe = str(tmp_path / name)
    src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-pthread", "-I",
                    os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "tests", "cpp", "compat"), src, "-L", LIBDIR,
                    "-lrbgpu", f"-Wl,-rpath,{LIBDIR}", "-o", exe], check=True)
    return exe


def _run(exe):
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout
    return r.stdout


def test_facade_compiles(tmp_path):
    """The facade and pool headers compile warning-free as plain C++ against the C header."""
    assert os.path.exists(_build(tmp_path))
    assert os.path.exists(_build(tmp_path, "test_world_pool"))


@pytest.mark.gpu
def test_facade_world_cases(tmp_path):
    out = _run(_build(tmp_path))
    assert out.count("PASS") == 29


@pytest.mark.gpu
def test_world_pool_cases(tmp_path):
    out = _run(_build(tmp_path, "test_world_pool"))
    assert out.count("PASS") == 3


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(REF_TEST), reason="reference sources were absent when oracle/ was built")
def test_reference_world_tests_unchanged():
    """tests/rbps/API/test_world.cpp of the reference, every case, against the MI355X World."""
    out = _run(REF_TEST)
    assert out.count("PASS") == 15
