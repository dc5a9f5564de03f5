"""The C++ drop-in layer on the device: reference-style use of Bbox,
MortonKeys, the staged tree/octree functions, computeBarnesHutAcceleration,
the leapfrog functions and Simulation through the library, run as a compiled
C++ program (amsc-nbody_amd/cpp/tests/shim_test.cpp)."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "amsc-nbody_amd", "cpp", "tests", "bin", "shim_test")


def test_cpp_drop_in_layer():
    assert os.path.exists(BIN), "build the C++ tests first (make -C amsc-nbody_amd cpp-tests)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") == 9
