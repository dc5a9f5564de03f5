"""The C ABI library loads and exports every declared symbol; host-side
workload generators (CPU only)."""

import ctypes
import os

import numpy as np
import pytest


def test_library_exports_every_declared_symbol():
    import nbody_amd
    lib = ctypes.CDLL(nbody_amd.LIB_PATH)
    names = nbody_amd.declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_library_is_built_for_gfx950():
    import nbody_amd
    data = open(nbody_amd.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_status_mapping():
    import nbody_amd as nb
    assert nb.STATUS_EXCEPTIONS[-1] is nb.InvalidArgument
    assert issubclass(nb.DomainError, ValueError)
    with pytest.raises(nb.InvalidArgument):
        nb.check(-1)


def test_phase_names():
    import nbody_amd as nb
    names = nb.phase_names()
    assert "traversal" in names and "sort" in names and len(names) == len(set(names))


@pytest.mark.parametrize("kind", ["plummer", "uniform", "disk", "clumps", "blob"])
def test_generators_deterministic(kind):
    import nbody_amd as nb
    a = nb.generate(kind, 5000, 42)
    b = nb.generate(kind, 5000, 42)
    c = nb.generate(kind, 5000, 43)
    for k in a:
        assert np.array_equal(a[k], b[k])
        assert np.all(np.isfinite(a[k]))
    assert not np.array_equal(a["x"], c["x"])
    assert abs(a["m"].sum() - 1.0) < 1e-12


def test_plummer_statistics():
    # half-mass radius of a Plummer sphere (a = 1): 1 / sqrt(2^(2/3) - 1) = 1.305;
    # truncation at 10 a removes ~1.5% of the mass at most.
    import nbody_amd as nb
    p = nb.generate("plummer", 200_000, 1)
    r = np.sqrt(p["x"] ** 2 + p["y"] ** 2 + p["z"] ** 2)
    assert r.max() <= 10.0
    assert abs(np.median(r) - 1.0 / np.sqrt(2 ** (2 / 3) - 1)) < 0.03
    # virial: 2K ~ |W| = 3 pi / 32 (G = M = a = 1)
    v2 = p["vx"] ** 2 + p["vy"] ** 2 + p["vz"] ** 2
    K = 0.5 * (p["m"] * v2).sum()
    assert abs(2 * K - 3 * np.pi / 32) < 0.02
    assert abs(p["x"].mean()) < 0.02


def test_uniform_statistics():
    import nbody_amd as nb
    p = nb.generate("uniform", 100_000, 2)
    assert p["x"].min() >= -1 and p["x"].max() <= 1
    assert abs(p["vx"].std() - 0.1) < 0.002
    lm = np.log(p["m"] * len(p["m"]))
    assert abs(lm.std() - 0.5) < 0.01


def test_clumps_statistics():
    import nbody_amd as nb
    p = nb.generate("clumps", 65_536, 3)
    assert np.all(p["vx"] == 0.0)
    assert p["x"].min() > -13 and p["x"].max() < 13
