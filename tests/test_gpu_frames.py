"""Frame capture on the device (nbh_frame_capture, nbh_frame_sample_types,
DeviceFrameWriter) against the frame oracle: bounds and quantized positions
bit for bit, for every sample shape, the narrow-store path, collapsed axes,
signed zeros, the rejections, and 4M-particle Plummer sets."""

import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _xyz(nb, rng, n):
    x = rng.normal(0.0, 3.0, n)
    y = rng.uniform(-1e-6, 1e-6, n) + 5.0
    z = rng.standard_cauchy(n) * 1e3
    return x, y, z


def _capture(nb, x, y, z, s):
    q, b = nb.capture_frame(nb.to_device(x), nb.to_device(y), nb.to_device(z), s)
    return b, q.cpu().numpy()


@pytest.mark.parametrize("n,s", [(1, 1), (2, 2), (3, 2), (7, 3), (1000, 1000), (1001, 999),
                                 (4097, 1), (100000, 333), (65536, 65535), (300000, 100000)])
def test_capture_matches_oracle(nb, orc, n, s):
    print(f"capture n={n} s={s}", flush=True)
    rng = np.random.default_rng(n * 7 + s)
    x, y, z = _xyz(nb, rng, n)
    b, q = _capture(nb, x, y, z, s)
    ob, oq, _ = orc.frame_capture(x, y, z, s)
    assert b.tobytes() == ob.tobytes()
    assert np.array_equal(q, oq)


def test_capture_unaligned_output(nb, orc):
    # A 2-byte-aligned output pointer takes the 16-bit store path.
    import torch
    rng = np.random.default_rng(3)
    n, s = 5000, 1237
    x, y, z = _xyz(nb, rng, n)
    buf = torch.zeros(3 * s + 2, dtype=torch.int16, device=nb.device())
    b = (C.c_double * 6)()
    d = [nb.to_device(a) for a in (x, y, z)]
    nb.check(nb.lib.nbh_frame_capture(nb.default_context().handle, *[t.data_ptr() for t in d],
                                      n, s, buf.data_ptr() + 2, b))
    got = buf.cpu().numpy().view(np.uint16)
    ob, oq, _ = orc.frame_capture(x, y, z, s)
    assert np.array(b[:]).tobytes() == ob.tobytes()
    assert np.array_equal(got[1:1 + 3 * s], oq.ravel())
    assert got[0] == 0 and got[-1] == 0  # nothing written outside the sample


def test_collapsed_axes_and_signed_zeros(nb, orc):
    n = 4096
    x = np.full(n, 2.5)
    y = np.zeros(n)
    y[::3] = -0.0
    z = np.linspace(-1.0, 1.0, n)
    z[n // 2] = -0.0
    for s in (n, 1000):
        b, q = _capture(nb, x, y, z, s)
        ob, oq, _ = orc.frame_capture(x, y, z, s)
        assert b.tobytes() == ob.tobytes()
        assert np.array_equal(q, oq)
        assert b[0] == b[3] == 2.5 and np.all(q[:, 0] == 0)
    # -0 orders below +0: the y bounds are (-0, +0) when both are sampled
    b, _ = _capture(nb, x, y, z, n)
    assert np.signbit(b[1]) and not np.signbit(b[4])


def test_rejections(nb, orc):
    n = 100
    x = np.arange(n, dtype=np.float64)
    bad = x.copy()
    bad[7] = np.inf
    with pytest.raises(nb.DeviceError, match="non-finite positions"):
        _capture(nb, x, bad, x, n)
    # 10 of 100 samples particles 0, 10, 20, ...: particle 7 is not read
    b, q = _capture(nb, x, bad, x, 10)
    ob, oq, _ = orc.frame_capture(x, bad, x, 10)
    assert b.tobytes() == ob.tobytes() and np.array_equal(q, oq)
    with pytest.raises(nb.InvalidArgument):
        _capture(nb, x, x, x, 0)
    with pytest.raises(nb.InvalidArgument):
        _capture(nb, x, x, x, n + 1)
    t = (np.arange(n) % 5).astype(np.uint8)
    got = nb.sample_frame_types(nb.to_device(t), 37).cpu().numpy()
    assert np.array_equal(got, orc.frame_capture(x, x, x, 37, types=t)[2])
    t[7] = 6
    with pytest.raises(nb.DeviceError, match="invalid particle types"):
        nb.sample_frame_types(nb.to_device(t))
    assert nb.sample_frame_types(nb.to_device(t), 10).cpu().numpy().tolist() == [0] * 10


@pytest.mark.parametrize("s", [4_194_304, 100_000])
def test_plummer_4m(nb, orc, s):
    n = 4_194_304
    print(f"plummer 4M s={s}", flush=True)
    p = nb.generate("plummer", n, 42)
    b, q = _capture(nb, p["x"], p["y"], p["z"], s)
    ob, oq, _ = orc.frame_capture(p["x"], p["y"], p["z"], s)
    assert b.tobytes() == ob.tobytes()
    assert np.array_equal(q, oq)


def test_device_frame_writer_on_a_simulation(nb, orc, tmp_path):
    n = 8192
    p = nb.generate("plummer", n, 9)
    types = (np.arange(n) % 5).astype(np.uint8)
    sim = nb.Simulation(p, 1e-3, 0.5, 1e-2, gravitational_constant=1.0)
    sim.initialize()
    sim.run(3)
    writer = nb.DeviceFrameWriter.from_simulation(sim, nb.to_device(types), 1000)
    assert (writer.source_particle_count, writer.sample_particle_count) == (n, 1000)
    spool = nb.FrameSpool(tmp_path, "sim")
    path = writer.write(spool, 0, sim.step_number(), sim.time())
    h = nb.read_frame_header(path)
    assert (h["version"], h["particle_count"], h["source_particle_count"],
            h["simulation_step"]) == (3, 1000, n, 3)
    assert h["simulation_time"] == sim.time()
    st = sim.state()
    ob, oq, ot = orc.frame_capture(st["x"], st["y"], st["z"], 1000, types=types)
    assert h["bounds_min"] + h["bounds_max"] == ob.tolist()
    raw = np.fromfile(path, dtype=np.uint8)
    assert np.array_equal(raw[128:128 + 6000].view(np.uint16).reshape(-1, 3), oq)
    pos, t = nb.read_frame(path)
    assert np.array_equal(t, ot)
    idx = orc.frame_sample_indices(n, 1000)
    for c, k in enumerate("xyz"):
        half = (ob[3 + c] - ob[c]) / 65535 * 0.5
        assert np.max(np.abs(pos[:, c] - st[k][idx])) <= half + 1e-6 * max(1, abs(ob[c]))
    # the writer follows the simulation: a later frame sees the new state
    sim.run(2)
    path2 = writer.write(spool, 1, sim.step_number(), sim.time())
    assert nb.read_frame_header(path2)["simulation_step"] == 5
    st = sim.state()
    ob2, oq2, _ = orc.frame_capture(st["x"], st["y"], st["z"], 1000)
    raw = np.fromfile(path2, dtype=np.uint8)
    assert np.array_equal(raw[128:128 + 6000].view(np.uint16).reshape(-1, 3), oq2)
    assert sorted(os.listdir(tmp_path)) == [os.path.basename(path), os.path.basename(path2)]
