"""Particle files through the C ABI and the Python binding: both header
widths, the optional type block, and the loader's failures (host only)."""

import numpy as np
import pytest

import nbody_amd


@pytest.fixture
def nb():
    # host-only calls: no device needed
    return nbody_amd


def sample(nb, n=257):
    p = nb.generate("uniform", n, 3)
    p["type"] = (np.arange(n) % 5).astype(np.uint8)
    return p


@pytest.mark.parametrize("wide", [False, True])
@pytest.mark.parametrize("types", [False, True])
def test_round_trip(nb, tmp_path, wide, types):
    p = sample(nb)
    f = tmp_path / "p.bin"
    nb.save_particles(f, p, wide_header=wide, with_types=types)
    q = nb.load_particles(f)
    for k in ("m", "x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(q[k], p[k]), k
    assert np.array_equal(q["type"], p["type"] if types else np.zeros(len(p["m"]), np.uint8))


def test_rejections(nb, tmp_path):
    p = sample(nb, 16)
    f = tmp_path / "p.bin"
    p["m"][3] = -1.0
    nb.save_particles(f, p)
    with pytest.raises(nb.DeviceError, match="Negative mass value at particle 3"):
        nb.load_particles(f)
    p["m"][3] = 1.0
    p["vy"][5] = np.nan
    nb.save_particles(f, p)
    with pytest.raises(nb.DeviceError, match="Non-finite y velocity value at particle 5"):
        nb.load_particles(f)
    p["vy"][5] = 0.0
    nb.save_particles(f, p)
    with open(f, "ab") as h:
        h.write(b"\\0")
    with pytest.raises(nb.DeviceError, match="Unsupported particle file layout"):
        nb.load_particles(f)
    with pytest.raises(nb.DeviceError, match="Could not open particle input file"):
        nb.load_particles(tmp_path / "missing.bin")


def test_mismatched_array_lengths_rejected_before_the_library(nb, tmp_path):
    # A short array would otherwise let the C side read past a host buffer.
    p = sample(nb)
    p["vy"] = p["vy"][:-1]
    with pytest.raises(nb.InvalidArgument, match="'vy' has 256"):
        nb.save_particles(tmp_path / "p.bin", p)
    with pytest.raises(nb.InvalidArgument, match="'vy' has 256"):
        nb.Simulation(p, 1e-3, 0.5, 0.01)
    q = sample(nb)
    q["x"] = np.stack([q["x"], q["x"]])
    with pytest.raises(nb.InvalidArgument, match="one-dimensional"):
        nb.Simulation(q, 1e-3, 0.5, 0.01)
    r = sample(nb)
    del r["vz"]
    with pytest.raises(nb.InvalidArgument, match="no 'vz'"):
        nb.save_particles(tmp_path / "p.bin", r)
    s = sample(nb)
    s["type"] = s["type"][:10]
    with pytest.raises(nb.InvalidArgument, match="'type' has 10"):
        nb.save_particles(tmp_path / "p.bin", s, with_types=True)
