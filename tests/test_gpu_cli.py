"""The nbody command-line driver running simulations on the device: the
summary lines the benchmark scripts parse, for a one-particle file (the
reference's CLI test) and a Plummer file in both precisions."""

import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "amsc-nbody_amd", "cpp", "apps", "bin", "nbody")


def write_particles(path, p, wide=False):
    n = len(p["m"])
    with open(path, "wb") as f:
        f.write(np.array([n], dtype=np.uint64 if wide else np.uint32).tobytes())
        for k in ("m", "x", "y", "z", "vx", "vy", "vz"):
            f.write(np.ascontiguousarray(p[k], dtype=np.float64).tobytes())


def test_one_particle_file(tmp_path):
    f = tmp_path / "one.bin"
    write_particles(f, {k: np.array([v]) for k, v in dict(m=1.0, x=0.0, y=0.0, z=0.0, vx=1.0,
                                                             vy=0.0, vz=0.0).items()})
    r = subprocess.run([CLI, "--input", str(f), "--time-step", "0.25", "--steps", "2",
                        "--theta", "0.6", "--softening", "0.03"],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    for text in ("Loaded 1 particles", "Simulation complete", "steps: 2", "simulated time: 0.5"):
        assert text in r.stdout


@pytest.mark.parametrize("precision,bits", [("fp64", "30"), ("mixed", "63")])
def test_plummer_file(nb, tmp_path, precision, bits):
    f = tmp_path / "plummer.bin"
    write_particles(f, nb.generate("plummer", 20000, 5), wide=True)
    r = subprocess.run([CLI, "-i", str(f), "-t", "1e-3", "-n", "5", "--softening", "0.01",
                        "--precision", precision, "--key-bits", bits,
                        "--gravitational-constant", "1"],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "Loaded 20000 particles" in r.stdout and "steps: 5" in r.stdout
    for key in ("initialization wall time", "evolution wall time", "average step time"):
        m = re.search(key + r": ([0-9.eE+-]+)", r.stdout)
        assert m and float(m.group(1)) > 0.0


def test_stream_dir_writes_frames(nb, tmp_path):
    # frames at the initial state, when due after steps, and at the final step
    f = tmp_path / "plummer.bin"
    p = nb.generate("plummer", 5000, 8)
    n = len(p["m"])
    with open(f, "wb") as out:
        out.write(np.array([n], dtype=np.uint64).tobytes())
        for k in ("m", "x", "y", "z", "vx", "vy", "vz"):
            out.write(p[k].tobytes())
        out.write((np.arange(n) % 5).astype(np.uint8).tobytes())
    spool = tmp_path / "frames"
    r = subprocess.run([CLI, "-i", str(f), "-t", "1e-3", "-n", "20", "--softening", "0.01",
                        "--gravitational-constant", "1", "--stream-dir", str(spool),
                        "--stream-max-particles", "700", "--sample-rate", "1000"],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    m = re.search(r"frames written: (\d+)", r.stdout)
    assert m and "particles per frame: 700 of 5000" in r.stdout
    assert f'frame spool: "{spool}"' in r.stdout
    frames = sorted(x for x in os.listdir(spool) if x.endswith(".nbsnap"))
    markers = [x for x in os.listdir(spool) if x.endswith(".complete")]
    assert len(frames) == int(m.group(1)) >= 2 and len(markers) == 1
    session = markers[0][len("run-"):-len(".complete")]
    steps = []
    for k, name in enumerate(frames):
        assert name == f"run-{session}-frame-{k:020d}.nbsnap"
        h = nb.read_frame_header(spool / name)
        assert (h["version"], h["particle_count"], h["source_particle_count"]) == (3, 700, 5000)
        steps.append(h["simulation_step"])
        pos, types = nb.read_frame(spool / name)
        assert np.array_equal(types, (np.arange(700) * 5000 // 700) % 5)
    assert steps[0] == 0 and steps[-1] == 20 and steps == sorted(set(steps))
