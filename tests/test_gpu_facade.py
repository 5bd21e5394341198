"""The C++ drop-in facade (include/rbps_amd/World.hpp) compiled against the C-ABI
library and run with the reference's World test cases (tests/cpp/test_world_facade.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp_path):
    exe = str(tmp_path / "test_world_facade")
    src = os.path.join(ROOT, "tests", "cpp", "test_world_facade.cpp")
    libdir = os.path.join(ROOT, "rbps_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src, "-L", libdir, "-lrbgpu",
                    f"-Wl,-rpath,{libdir}", "-o", exe], check=True)
    return exe


def test_facade_compiles(tmp_path):
    """The facade header compiles as plain C++ against the C header (no GPU needed)."""
    assert os.path.exists(_build(tmp_path))


@pytest.mark.gpu
def test_facade_world_cases(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
