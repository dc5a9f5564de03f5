"""The particle file reader of the C++ drop-in layer (host code): format
detection and validation cases of the reference's loader."""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "amsc-nbody_amd")
BIN = os.path.join(PKG, "cpp", "tests", "bin", "particle_file_test")


def test_particle_file_reader_cases():
    # host-only program: built here with g++ when the build step has not
    subprocess.run(["make", "-s", "-C", PKG, "cpp/tests/bin/particle_file_test"],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") == 17
