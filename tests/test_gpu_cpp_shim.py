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


ARITH_BIN = os.path.join(ROOT, "amsc-nbody_amd", "cpp", "tests", "bin", "fp64_arith_test")


def test_fp64_walk_arithmetic_equals_ieee_operations():
    """The FP64 traversal's reduced square root and reciprocal equal the
    device's IEEE sqrt() and 1.0 / x bit for bit in their range
    (cpp/tests/fp64_arith_test.hip: random operands and near-ties)."""
    assert os.path.exists(ARITH_BIN), "build the C++ tests first (make -C amsc-nbody_amd cpp-tests)"
    r = subprocess.run([ARITH_BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout
