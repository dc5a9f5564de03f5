"""The FP64 walk's exact opening threshold on the host: the library's
header compiled with g++ and checked against sqrt(d2) > crit over the ulp
neighbourhoods of the threshold and of crit^2, across the double range
(amsc-nbody_amd/cpp/tests/mac_threshold_test.cpp)."""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "amsc-nbody_amd", "cpp", "tests", "mac_threshold_test.cpp")


def test_mac_threshold_equals_the_square_root_test(tmp_path):
    exe = tmp_path / "mac_threshold_test"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-o", str(exe),
                    SRC], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "checks passed" in r.stdout
