"""Register budgets of the persistent traversal kernels (CPU: compiles only).

The persistent traversals size their grids to the wavefronts that stay
resident at once (walk_common.hpp: 8 per SIMD for the mixed whole walk, 7
or 6 for the FP64 walk).  On gfx950 a wavefront of at most 64 VGPRs and 80 SGPRs
(VCC included) gets 8 slots per SIMD; 81 to 89 SGPRs were measured at 7
resident wavefronts per SIMD and 97 SGPRs at 6, although the compiler's own
occupancy remark reports 8 for all of them.  A grid sized for more
wavefronts than fit leaves the last ones to start after the queue is empty,
which lengthened the 1M-particle traversal by a fifth.  These tests read the
compiler's resource remarks (make kernel-resources) and check the counts
that the launch code assumes.
"""

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "amsc-nbody_amd")

pytestmark = pytest.mark.skipif(shutil.which("make") is None or
                                not os.path.exists("/opt/rocm/bin/hipcc"),
                                reason="needs make and hipcc")


def resources(src):
    r = subprocess.run(["make", "-s", "-C", PKG, "kernel-resources", f"SRC={src}"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    out, cur = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"remark:\s+Function Name: (\S+)", line)
        if m:
            cur = m.group(1)
            out[cur] = {}
            continue
        m = re.search(r"remark:\s+([A-Za-z ]+?)(?: \[[^\]]*\])?: (\d+)\s", line + " ")
        if m and cur:
            out[cur][m.group(1).strip()] = int(m.group(2))
    return out


def template_bools(name, kernel):
    m = re.search(kernel + r"I((?:Lb[01]E)+)E", name)
    return [int(b) for b in re.findall(r"Lb([01])E", m.group(1))] if m else None


def test_mixed_whole_walk_fits_eight_wavefronts():
    # walk_mixed_kernel<CONTAINS, SOFT, COUNT, WIDE, SPLIT>: every whole-walk
    # instantiation the step launches (COUNT = SPLIT = false)
    res = resources("csrc/walk_mixed.hip")
    checked = 0
    for name, r in res.items():
        b = template_bools(name, "walk_mixed_kernel")
        if b is None or b[2] or b[4]:
            continue
        checked += 1
        assert r["TotalSGPRs"] <= 80, (name, r)
        assert r["VGPRs"] + r.get("AGPRs", 0) <= 64, (name, r)
        assert r["ScratchSize"] == 0, (name, r)
    assert checked == 8


def test_fp64_walk_fits_its_residency():
    # walk_fp64_kernel<CONTAINS, COUNT, WIDE, FAST>, COUNT = false: the
    # launch assumes 7 resident wavefronts per SIMD for FAST without
    # CONTAINS (at most 96 SGPRs and 72 VGPRs) and 6 for the others (at
    # most 112 SGPRs and 80 VGPRs)
    res = resources("csrc/walk_fp64.hip")
    checked = 0
    for name, r in res.items():
        b = template_bools(name, "walk_fp64_kernel")
        if b is None or b[1]:
            continue
        checked += 1
        seven = b[3] and not b[0]
        assert r["TotalSGPRs"] <= (96 if seven else 112), (name, r)
        assert r["VGPRs"] + r.get("AGPRs", 0) <= (72 if seven else 80), (name, r)
        assert r["ScratchSize"] == 0, (name, r)
    assert checked == 6  # WIDE and FAST are not instantiated together
