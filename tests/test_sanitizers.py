"""The CPU sanitizer builds, run from their make targets:
- oracle/Makefile sanitize-check: the oracle's CPU tests on an
  AddressSanitizer + UBSan build of the oracle, and the oracle's parallel
  paths (race_check.cpp) under ASan + UBSan and under ThreadSanitizer with
  LLVM's OpenMP runtime and its Archer tool;
- amsc-nbody_amd/Makefile sanitize-check: the particle-file and frame tests
  of the drop-in headers and the host C ABI test (generators, particle
  files, frames, Hilbert tables) under ASan + UBSan with leak checking."""

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("where", ["oracle", "amsc-nbody_amd"])
def test_sanitize_check(where):
    env = dict(os.environ)
    env.pop("ORACLE_LIB", None)
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, where), "sanitize-check"],
                       capture_output=True, text=True, timeout=900, env=env)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "FAIL" not in r.stdout
