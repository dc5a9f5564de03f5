"""Graphical frames on the host: the C++ frame cases, the frame oracle pinned
by exact arithmetic and the reference's frame vectors, and the spool and
readers through the C ABI (no device needed).

Golden vectors: tests/unit/cpu/streaming_frame_test.cpp (header fields,
quantized endpoints, decoded positions) and
tests/unit/javascript/browser_frame_parser_test.mjs (byte offsets of every
header field, a version-2 float frame and a version-3 quantized frame).
"""

import math
import os
import struct
import subprocess
from fractions import Fraction

import numpy as np
import pytest

import nbody_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "amsc-nbody_amd")
BIN = os.path.join(PKG, "cpp", "tests", "bin", "frame_test")


@pytest.fixture
def nb():
    # host-only calls: no device needed
    return nbody_amd


# ------------------------------------------------------------ C++ cases
def test_cpp_frame_cases():
    subprocess.run(["make", "-s", "-C", PKG, "cpp/tests/bin/frame_test"], check=True,
                   capture_output=True, text=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") == 16


# --------------------------------------------------------- frame oracle
def test_oracle_sample_indices(orc):
    # j * n // s by hand for n = 20, s = 7
    assert orc.frame_sample_indices(20, 7).tolist() == [0, 2, 5, 8, 11, 14, 17]
    assert orc.frame_sample_indices(9, 9).tolist() == list(range(9))
    assert orc.frame_sample_indices(1000, 1).tolist() == [0]
    rng = np.random.default_rng(5)
    for _ in range(20):
        n = int(rng.integers(1, 5000))
        s = int(rng.integers(1, n + 1))
        idx = orc.frame_sample_indices(n, s)
        assert idx[0] == 0 and idx[-1] < n
        assert np.all(np.diff(idx.astype(np.int64)) >= 1)  # distinct particles
        # the capture samples exactly those particles: quantize x = index
        x = np.arange(n, dtype=np.float64)
        b, q, _ = orc.frame_capture(x, x, x, s)
        assert b[0] == idx[0] and b[3] == idx[-1]


def _exact_quantize(p, lo, hi):
    """quantizePosition with the fma evaluated exactly and rounded once."""
    if hi == lo:
        return 0
    t = (p - lo) / (hi - lo)          # IEEE double ops, as on the device
    t = min(1.0, max(0.0, t))
    return int(float(Fraction(t) * 65535 + Fraction(1, 2)))


def _ordered_min_max(v):
    key = lambda a: (a, 0 if math.copysign(1.0, a) < 0 else 1)  # -0 below +0
    return min(v, key=key), max(v, key=key)


def test_oracle_quantization_exact(orc):
    rng = np.random.default_rng(11)
    n = 3000
    x = rng.normal(0.0, 1e3, n)
    y = rng.uniform(-1e-9, 1e-9, n)
    z = np.full(n, 7.25)                           # collapsed axis
    x[:4] = [-0.0, 0.0, 1e-300, -1e-300]
    y[10] = -0.0
    s = 1234
    b, q, _ = orc.frame_capture(x, y, z, s)
    idx = orc.frame_sample_indices(n, s)
    for c, arr in enumerate((x, y, z)):
        lo, hi = _ordered_min_max([float(arr[i]) for i in idx])
        assert b[c] == lo and b[3 + c] == hi
        assert math.copysign(1.0, b[c]) == math.copysign(1.0, lo)
        want = [_exact_quantize(float(arr[i]), lo, hi) for i in idx]
        assert q[:, c].tolist() == want, c
    assert np.all(q[:, 2] == 0)
    # endpoints and the half-step bound of the decode
    for c, arr in enumerate((x, y)):
        lo, hi = b[c], b[3 + c]
        assert q[:, c].min() == 0 and q[:, c].max() == 65535
        dec = lo + (hi - lo) * q[:, c].astype(np.float64) / 65535.0
        assert np.max(np.abs(dec - arr[idx])) <= (hi - lo) / 65535.0 * 0.5 * (1 + 1e-9)


def test_oracle_reference_frame_vector(orc):
    # streaming_frame_test.cpp: bounds {-2.5, 3.5, -5}..{1.25, 4.75, 6.125}
    # and quantized {65535, 0, 0, 0, 65535, 65535} for the two corners.
    x, y, z = [1.25, -2.5], [3.5, 4.75], [-5.0, 6.125]
    b, q, t = orc.frame_capture(x, y, z, types=[1, 4])
    assert b.tolist() == [-2.5, 3.5, -5.0, 1.25, 4.75, 6.125]
    assert q.ravel().tolist() == [65535, 0, 0, 0, 65535, 65535]
    assert t.tolist() == [1, 4]


def test_oracle_rejections(orc):
    x = np.arange(10, dtype=np.float64)
    bad = x.copy()
    bad[3] = np.nan
    with pytest.raises(orc.OracleError, match="non-finite positions"):
        orc.frame_capture(bad, x, x)
    # particle 3 is not sampled when 2 of 10 are taken (indices 0 and 5)
    orc.frame_capture(bad, x, x, 2)
    with pytest.raises(orc.OracleError, match="invalid particle types"):
        orc.frame_capture(x, x, x, types=np.array([0, 1, 2, 5, 4, 3, 0, 0, 0, 0]))
    _, _, t = orc.frame_capture(x, x, x, 2, types=np.array([0, 1, 2, 5, 4, 3, 0, 0, 0, 0]))
    assert t.tolist() == [0, 3]


# ------------------------------------------ the reference's frame bytes
def _common(version, header_bytes, sequence, step, time, source, particles, scalar, payload):
    return (b"NBSNAP01" + struct.pack("<IIQQdQQIIQ", version, header_bytes, sequence, step,
                                      time, source, particles, scalar, 3, payload))


def archived_typed_frame():
    """makeArchivedTypedFrame of browser_frame_parser_test.mjs."""
    pos = [1, 2, 3, -4, 5, -6]
    types = bytes([1, 4])
    head = _common(2, 80, 3, 7, 1.25, 20, 2, 4, len(pos) * 4 + 2) + struct.pack("<Q", 40)
    return head + struct.pack("<6f", *pos) + types


def current_typed_frame():
    """makeCurrentTypedFrame of browser_frame_parser_test.mjs."""
    q = [65535, 0, 0, 0, 65535, 65535]
    head = (_common(3, 128, 8, 42, 0.5, 20, 2, 2, len(q) * 2 + 2)
            + struct.pack("<3d", -2.5, 3.5, -5) + struct.pack("<3d", 1.25, 4.75, 6.125)
            + struct.pack("<Q", 2))
    return head + struct.pack("<6H", *q) + bytes([1, 4])


def test_reads_reference_frame_bytes(nb, tmp_path):
    f = tmp_path / "archived.nbsnap"
    f.write_bytes(archived_typed_frame())
    h = nb.read_frame_header(f)
    assert (h["version"], h["header_bytes"], h["simulation_step"], h["total_steps"],
            h["particle_count"]) == (2, 80, 7, 40, 2)
    pos, types = nb.read_frame(f)
    assert pos.ravel().tolist() == [1, 2, 3, -4, 5, -6]
    assert types.tolist() == [1, 4]

    f = tmp_path / "current.nbsnap"
    f.write_bytes(current_typed_frame())
    h = nb.read_frame_header(f)
    assert (h["version"], h["header_bytes"], h["particle_type_bytes"]) == (3, 128, 2)
    assert h["bounds_min"] == [-2.5, 3.5, -5] and h["bounds_max"] == [1.25, 4.75, 6.125]
    pos, types = nb.read_frame(f)
    assert pos.ravel().tolist() == [1.25, 3.5, -5, -2.5, 4.75, 6.125]
    assert types.tolist() == [1, 4]


def test_writes_reference_frame_bytes(nb, tmp_path):
    spool = nb.FrameSpool(tmp_path / "spool", "bytes")
    path = spool.write_quantized_positions(
        8, 42, 0.5, np.array([[65535, 0, 0], [0, 65535, 65535]]),
        [-2.5, 3.5, -5, 1.25, 4.75, 6.125], source_count=20, types=[1, 4])
    assert os.path.basename(path) == "run-bytes-frame-00000000000000000008.nbsnap"
    assert open(path, "rb").read() == current_typed_frame()
    path = spool.write_typed_positions(3, 7, 1.25, 40, [[1, 2, 3], [-4, 5, -6]], [1, 4],
                                       source_count=20)
    assert open(path, "rb").read() == archived_typed_frame()
    assert not [p for p in os.listdir(spool.directory) if p.endswith(".part")]


def test_spool_layouts_and_failures(nb, tmp_path):
    spool = nb.FrameSpool(tmp_path, "layouts")
    p = spool.write_quantized_positions(1, 0, 0.0, [[1, 2, 3]], [0, 0, 0, 1, 1, 1])
    h = nb.read_frame_header(p)
    assert (h["version"], h["header_bytes"], h["payload_bytes"]) == (2, 120, 6)
    assert os.path.getsize(p) == 126
    p = spool.write_positions(2, 0, 0.0, [[0.5, 1.5, 2.5]], source_count=9)
    h = nb.read_frame_header(p)
    assert (h["version"], h["header_bytes"], h["source_particle_count"]) == (1, 72, 9)
    assert nb.read_frame(p)[1].tolist() == [0]
    with pytest.raises(nb.DeviceError, match="Refusing to overwrite queued frame"):
        spool.write_positions(2, 1, 0.1, [[0, 0, 0]])
    with pytest.raises(nb.InvalidArgument, match="finite and ordered"):
        spool.write_quantized_positions(3, 0, 0.0, [[0, 0, 0]], [1, 0, 0, 0, 0, 0])
    with pytest.raises(nb.InvalidArgument, match="cannot be smaller than its sample"):
        spool.write_quantized_positions(3, 0, 0.0, [[0, 0, 0]] * 2, [0] * 6, source_count=1)
    with pytest.raises(nb.DeviceError, match="unsupported value"):
        spool.write_quantized_positions(3, 0, 0.0, [[0, 0, 0]], [0] * 6, types=[7])
    assert sorted(os.listdir(tmp_path)) == [
        "run-layouts-frame-00000000000000000001.nbsnap",
        "run-layouts-frame-00000000000000000002.nbsnap"]
    marker = spool.mark_complete()
    assert os.path.basename(marker) == "run-layouts.complete"
    assert open(marker).read() == "layouts\n"
    assert spool.mark_complete() == marker
    bad = tmp_path / "bad.nbsnap"
    bad.write_bytes(current_typed_frame()[:-1])
    with pytest.raises(nb.DeviceError, match="size does not match its header"):
        nb.read_frame(bad)
    bad.write_bytes(b"NBSNAP02" + current_typed_frame()[8:])
    with pytest.raises(nb.DeviceError, match="Invalid N-body frame magic"):
        nb.read_frame_header(bad)


def test_session_ids(nb, tmp_path):
    s = nb.FrameSpool(tmp_path)
    assert s.session_id.replace("-", "").isdigit()
    for bad in ("", "a b", "../x", "é"):
        with pytest.raises(nb.InvalidArgument):
            nb.FrameSpool(tmp_path, bad)
