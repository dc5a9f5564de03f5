"""Energy, force vector-Jacobian product and the K-step adjoint (GPU), for
positions, velocities and masses.

Pinned by: the numpy direct sum at theta = 0 (potential, VJP, adjoint: the
Barnes-Hut walk with theta = 0 opens every node, so it is the direct sum);
central finite differences of the Barnes-Hut forces and of the K-step
evolution at theta = 0.5; the truncated-Plummer potential energy and the
virial ratio; a two-body Kepler orbit (energy, closure, momentum, order).
"""

import numpy as np
import pytest
import torch

import helpers as H

pytestmark = pytest.mark.gpu

# adjoint vs central finite differences (relative, norm over the probed set)
FD_TOL = 1e-4


def planar(p):
    return np.stack([p["x"], p["y"], p["z"]]), np.stack([p["vx"], p["vy"], p["vz"]])


def sim_of(nb, p, theta, eps, dt=1e-3, precision=None):
    prec = nb.PRECISION_FP64 if precision is None else precision
    s = nb.Simulation(p, dt, theta, eps, gravitational_constant=1.0, precision=prec,
                      bits_per_axis=21)
    s.initialize()
    return s


def with_positions(p, pos):
    q = dict(p)
    q["x"], q["y"], q["z"] = pos[0].copy(), pos[1].copy(), pos[2].copy()
    return q


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def central_fd(f, h):
    """Central difference of f(delta) at 0 with steps h and 2h.  A probe
    whose two estimates disagree has a kink within 2h (an opening-test
    decision or a Morton cell changed): it is reported as invalid."""
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(2 * h) - f(-2 * h)) / (4 * h)
    ok = abs(d1 - d2) <= 3e-5 * max(1.0, abs(d1))
    return d1, ok


def compare_probes(ad, fd, ok, what):
    ad, fd, ok = np.asarray(ad), np.asarray(fd), np.asarray(ok)
    err = rel(ad[ok], fd[ok])
    print(f"{what}: {ok.sum()}/{len(ok)} probes smooth, rel error {err:.2e}")
    assert ok.sum() >= 56
    assert err < FD_TOL


# ------------------------------------------------------------------ energy
def test_potential_theta0_equals_direct_sum(nb):
    p = nb.generate("uniform", 1500, 3)
    s = sim_of(nb, p, 0.0, 0.01)
    phi = s.potential().cpu().numpy()
    ref = H.direct_potential_numpy(p["m"], p["x"], p["y"], p["z"], 0.01)
    assert np.max(np.abs(phi - ref) / np.abs(ref)) < 1e-12


def test_potential_barnes_hut_close_to_direct(nb):
    p = nb.generate("plummer", 3000, 4)
    s = sim_of(nb, p, 0.5, 0.01)
    phi = s.potential().cpu().numpy()
    ref = H.direct_potential_numpy(p["m"], p["x"], p["y"], p["z"], 0.01)
    r = np.abs(phi - ref) / np.abs(ref)
    assert np.sqrt((r ** 2).mean()) < 1e-3


def test_energy_plummer_closed_form_and_virial(nb):
    n = 200_000
    p = nb.generate("plummer", n, 5)
    s = sim_of(nb, p, 0.5, 1e-3)
    k, u = s.energy()
    w = H.truncated_plummer_energy(10.0)
    print(f"Plummer N={n}: U={u:.6f} closed form {w:.6f}  2K/|U|={2 * k / -u:.4f}")
    assert abs(u - w) < 0.015 * abs(w)
    assert abs(2.0 * k / -u - 1.0) < 0.03
    kin = 0.5 * (p["m"] * (p["vx"] ** 2 + p["vy"] ** 2 + p["vz"] ** 2)).sum()
    assert abs(k - kin) < 1e-12 * kin


def test_energy_needs_fp64_walk(nb):
    p = nb.generate("plummer", 500, 1)
    s = sim_of(nb, p, 0.5, 0.01, precision=nb.PRECISION_MIXED)
    with pytest.raises(nb.LogicError):
        s.energy()


# ------------------------------------------------------------- force VJP
def test_force_vjp_theta0_equals_direct_vjp(nb):
    p = nb.generate("uniform", 400, 6)
    pos, _ = planar(p)
    rng = np.random.default_rng(0)
    ab = rng.normal(size=pos.shape)
    s = sim_of(nb, p, 0.0, 0.02)
    g = np.stack([t.cpu().numpy() for t in s.force_vjp(*[torch.from_numpy(a) for a in ab])])
    ref = H.direct_vjp_numpy(p["m"], pos, ab, 0.02)
    assert rel(g, ref) < 1e-11


def test_force_vjp_matches_finite_differences(nb):
    p = nb.generate("plummer", 3000, 7)
    eps = 0.01
    pos, _ = planar(p)
    rng = np.random.default_rng(1)
    ab = rng.normal(size=pos.shape)
    s = sim_of(nb, p, 0.5, eps)
    g = np.stack([t.cpu().numpy() for t in s.force_vjp(*[torch.from_numpy(a) for a in ab])])
    a0 = np.stack(s.accelerations())
    probes = [(int(c), int(i)) for c, i in zip(rng.integers(0, 3, 64), rng.choice(3000, 64, False))]
    fd, ok, gg = [], [], []
    for c, i in probes:
        def f(delta, c=c, i=i):
            q = pos.copy()
            q[c, i] += delta
            a = np.stack(sim_of(nb, with_positions(p, q), 0.5, eps).accelerations())
            # difference first: the sum then carries no rounding of the base
            return ((a - a0) * ab).sum()
        d, good = central_fd(f, 1e-8)
        fd.append(d)
        ok.append(good)
        gg.append(g[c, i])
    compare_probes(gg, fd, ok, "force VJP vs FD on 64 coordinates")


def test_force_vjp_mass_theta0_equals_direct(nb):
    p = nb.generate("uniform", 400, 12)
    p["m"] = np.random.default_rng(7).uniform(0.5, 1.5, 400) / 400
    pos, _ = planar(p)
    ab = np.random.default_rng(8).normal(size=pos.shape)
    s = sim_of(nb, p, 0.0, 0.02)
    out = s.force_vjp(*[torch.from_numpy(a) for a in ab], with_mass=True)
    g = np.stack([t.cpu().numpy() for t in out[:3]])
    gm = out[3].cpu().numpy()
    assert rel(g, H.direct_vjp_numpy(p["m"], pos, ab, 0.02)) < 1e-11
    assert rel(gm, H.direct_mass_vjp_numpy(p["m"], pos, ab, 0.02)) < 1e-11


def test_force_vjp_mass_matches_finite_differences(nb):
    # At theta = 0.5 most sources are node monopoles: the mass cotangent goes
    # through each node's mass and centre of mass to its particles.
    n = 3000
    p = nb.generate("plummer", n, 13)
    p["m"] = p["m"] * np.random.default_rng(9).uniform(0.5, 1.5, n)
    eps = 0.01
    pos, _ = planar(p)
    rng = np.random.default_rng(10)
    ab = rng.normal(size=pos.shape)
    s = sim_of(nb, p, 0.5, eps)
    gm = s.force_vjp(*[torch.from_numpy(a) for a in ab], with_mass=True)[3].cpu().numpy()
    a0 = np.stack(s.accelerations())
    fd, ok, ad = [], [], []
    for i in rng.choice(n, 64, False):
        def f(delta, i=int(i)):
            q = dict(p)
            q["m"] = p["m"].copy()
            q["m"][i] += delta
            a = np.stack(sim_of(nb, q, 0.5, eps).accelerations())
            return ((a - a0) * ab).sum()
        d, good = central_fd(f, 1e-4 * p["m"][i])
        fd.append(d)
        ok.append(good)
        ad.append(gm[i])
    compare_probes(ad, fd, ok, "force mass VJP vs FD on 64 masses")


# --------------------------------------------------------------- adjoint
def test_adjoint_theta0_equals_direct_adjoint(nb):
    n, K, dt, eps = 64, 6, 1e-2, 0.05
    p = nb.generate("blob", n, 8)
    pos, vel = planar(p)
    rng = np.random.default_rng(2)
    xbar, vbar = rng.normal(size=pos.shape), rng.normal(size=pos.shape)
    adj = nb.Adjoint(n, dt, 0.0, eps, K, 4, gravitational_constant=1.0)
    out = adj.forward(*[torch.from_numpy(a) for a in (p["m"], *pos, *vel)])
    xs, vk = H.leapfrog_numpy(p["m"], pos, vel, dt, K, eps)
    assert rel(np.stack([t.cpu().numpy() for t in out[:3]]), xs[-1]) < 1e-13
    assert rel(np.stack([t.cpu().numpy() for t in out[3:]]), vk) < 1e-12
    g = adj.backward(*[torch.from_numpy(a) for a in (*xbar, *vbar)])
    xb, vb = H.leapfrog_adjoint_numpy(p["m"], pos, vel, dt, K, eps, xbar, vbar)
    assert rel(np.stack([t.cpu().numpy() for t in g[:3]]), xb) < 1e-10
    assert rel(np.stack([t.cpu().numpy() for t in g[3:]]), vb) < 1e-10


def test_adjoint_mass_theta0_equals_direct_adjoint(nb):
    n, K, dt, eps = 64, 6, 1e-2, 0.05
    p = nb.generate("blob", n, 14)
    p["m"] = np.random.default_rng(11).uniform(0.5, 1.5, n) / n
    pos, vel = planar(p)
    rng = np.random.default_rng(12)
    xbar, vbar = rng.normal(size=pos.shape), rng.normal(size=pos.shape)
    adj = nb.Adjoint(n, dt, 0.0, eps, K, 4, gravitational_constant=1.0)
    adj.forward(*[torch.from_numpy(a) for a in (p["m"], *pos, *vel)])
    g = adj.backward(*[torch.from_numpy(a) for a in (*xbar, *vbar)], with_mass=True)
    xb, vb, mb = H.leapfrog_adjoint_numpy(p["m"], pos, vel, dt, K, eps, xbar, vbar,
                                          with_mass=True)
    assert rel(np.stack([t.cpu().numpy() for t in g[:3]]), xb) < 1e-10
    assert rel(np.stack([t.cpu().numpy() for t in g[3:6]]), vb) < 1e-10
    assert rel(g[6].cpu().numpy(), mb) < 1e-10


def _disk_problem(nb, n=2048, K=20, every=5):
    p = nb.generate("disk", n, 9)
    pos, vel = planar(p)
    adj = nb.Adjoint(n, 1e-3, 0.5, 0.01, K, every, gravitational_constant=1.0)
    return p, pos, vel, adj


def test_adjoint_matches_finite_differences(nb):
    p, pos, vel, adj = _disk_problem(nb)
    n = pos.shape[1]
    rng = np.random.default_rng(3)
    w, u = rng.normal(size=pos.shape), rng.normal(size=pos.shape)
    m = torch.from_numpy(p["m"])

    def final(q, v):
        out = adj.forward(m, *[torch.from_numpy(a) for a in (*q, *v)])
        return np.stack([t.cpu().numpy() for t in out])

    base = final(pos, vel)

    def loss(q, v):
        # difference to the base run first, then the weighted sum
        d = final(q, v) - base
        return (d[:3] * w).sum() + (d[3:] * u).sum()

    final(pos, vel)
    g = adj.backward(*[torch.from_numpy(a) for a in (*w, *u)])
    gx = np.stack([t.cpu().numpy() for t in g[:3]])
    gv = np.stack([t.cpu().numpy() for t in g[3:]])
    fd, ok, ad = [], [], []
    for k, (c, i) in enumerate(zip(rng.integers(0, 3, 64), rng.choice(n, 64, False))):
        on_pos = k < 32

        def f(delta, c=c, i=i, on_pos=on_pos):
            q, v = pos.copy(), vel.copy()
            (q if on_pos else v)[c, i] += delta
            return loss(q, v)
        d, good = central_fd(f, 1e-8)
        fd.append(d)
        ok.append(good)
        ad.append((gx if on_pos else gv)[c, i])
    compare_probes(ad, fd, ok, f"adjoint (disk N={n}, K=20, checkpoints every 5) vs FD")


def test_adjoint_mass_gradient_matches_finite_differences(nb):
    # the differentiable workload's check: gradient of the final state with
    # respect to the masses, central differences on 64 of them
    p, pos, vel, adj = _disk_problem(nb)
    n = pos.shape[1]
    rng = np.random.default_rng(13)
    w, u = rng.normal(size=pos.shape), rng.normal(size=pos.shape)
    state = [torch.from_numpy(a) for a in (*pos, *vel)]

    def final(m):
        out = adj.forward(torch.from_numpy(m), *state)
        return np.stack([t.cpu().numpy() for t in out])

    base = final(p["m"])

    def loss(m):
        d = final(m) - base
        return (d[:3] * w).sum() + (d[3:] * u).sum()

    final(p["m"])
    gm = adj.backward(*[torch.from_numpy(a) for a in (*w, *u)], with_mass=True)[6]
    gm = gm.cpu().numpy()
    fd, ok, ad = [], [], []
    for i in rng.choice(n, 64, False):
        def f(delta, i=int(i)):
            m = p["m"].copy()
            m[i] += delta
            return loss(m)
        # a mass step moves the centres of mass of the particle's nodes by
        # the same relative amount: 1e-5 keeps opening decisions stable
        # over the 20 steps while the rounding of the loss stays ~1e-6
        d, good = central_fd(f, 1e-5 * p["m"][i])
        fd.append(d)
        ok.append(good)
        ad.append(gm[i])
    compare_probes(ad, fd, ok, f"adjoint mass gradient (disk N={n}, K=20) vs FD")


def test_adjoint_independent_of_checkpoint_interval(nb):
    n, K = 1024, 9
    p = nb.generate("disk", n, 10)
    pos, vel = planar(p)
    rng = np.random.default_rng(4)
    w, u = rng.normal(size=pos.shape), rng.normal(size=pos.shape)
    res = []
    for every in (1, 4, K):
        adj = nb.Adjoint(n, 1e-3, 0.5, 0.01, K, every, gravitational_constant=1.0)
        adj.forward(*[torch.from_numpy(a) for a in (p["m"], *pos, *vel)])
        g = adj.backward(*[torch.from_numpy(a) for a in (*w, *u)], with_mass=True)
        res.append(np.concatenate([t.cpu().numpy() for t in g]))
        adj.close()
    # node cotangents are summed with atomics: equal up to summation order
    assert rel(res[1], res[0]) < 1e-12
    assert rel(res[2], res[0]) < 1e-12


def test_adjoint_backward_requires_forward(nb):
    adj = nb.Adjoint(16, 1e-3, 0.5, 0.01, 3, 1, gravitational_constant=1.0)
    z = torch.zeros(16, dtype=torch.float64)
    with pytest.raises(nb.LogicError):
        adj.backward(z, z, z, z, z, z)


def test_adjoint_rejects_bad_arguments(nb):
    with pytest.raises(nb.InvalidArgument):
        nb.Adjoint(16, 1e-3, 0.5, 0.01, 0, 1)
    with pytest.raises(nb.InvalidArgument):
        nb.Adjoint(16, 1e-3, 0.5, 0.01, 3, 0)


def test_autograd_function_matches_adjoint(nb):
    n, K = 256, 5
    p = nb.generate("plummer", n, 11)
    pos, vel = planar(p)
    adj = nb.Adjoint(n, 1e-3, 0.5, 0.01, K, 2, gravitational_constant=1.0)
    fn = nb.leapfrog_function()
    dev = torch.device("cuda", 0)
    x0 = torch.tensor(pos, device=dev, requires_grad=True)
    v0 = torch.tensor(vel, device=dev, requires_grad=True)
    w = torch.randn(3, n, dtype=torch.float64, device=dev, generator=None)
    m = torch.tensor(p["m"], device=dev, requires_grad=True)
    xk, vk = fn.apply(adj, m, x0, v0)
    (xk * w).sum().backward()
    ref = adj.backward(*w, *torch.zeros(3, n, dtype=torch.float64, device=dev),
                       with_mass=True)
    assert rel(x0.grad.cpu().numpy(), np.stack([t.cpu().numpy() for t in ref[:3]])) < 1e-12
    assert rel(v0.grad.cpu().numpy(), np.stack([t.cpu().numpy() for t in ref[3:6]])) < 1e-12
    assert rel(m.grad.cpu().numpy(), ref[6].cpu().numpy()) < 1e-12


def test_autograd_backward_of_a_superseded_forward_raises(nb):
    # The adjoint keeps the checkpoints of its latest forward only: the
    # backward of an earlier output must refuse rather than differentiate
    # the later trajectory.
    n, K = 128, 3
    p = nb.generate("plummer", n, 12)
    pos, vel = planar(p)
    adj = nb.Adjoint(n, 1e-3, 0.5, 0.01, K, 1, gravitational_constant=1.0)
    fn = nb.leapfrog_function()
    dev = torch.device("cuda", 0)
    m = torch.tensor(p["m"], device=dev)
    x0 = torch.tensor(pos, device=dev, requires_grad=True)
    v0 = torch.tensor(vel, device=dev, requires_grad=True)
    g0 = adj.generation()
    xk1, _ = fn.apply(adj, m, x0, v0)
    xk2, _ = fn.apply(adj, m, x0 * 1.001, v0)
    assert adj.generation() == g0 + 2
    with pytest.raises(RuntimeError, match="another forward"):
        xk1.sum().backward()
    # the latest output still differentiates
    xk2.sum().backward()
    assert torch.isfinite(x0.grad).all()


# ------------------------------------------------------------- two bodies
def _two_body(nb, steps_per_period, periods=1.0):
    # m1 = m2 = 1/2 at separation 1, G = 1: circular, relative speed 1,
    # period 2 pi (the closed-form Kepler orbit)
    p = {"m": np.array([0.5, 0.5]), "x": np.array([-0.5, 0.5]), "y": np.zeros(2),
         "z": np.zeros(2), "vx": np.zeros(2), "vy": np.array([-0.5, 0.5]), "vz": np.zeros(2)}
    period = 2.0 * np.pi
    dt = period / steps_per_period
    s = nb.Simulation(p, dt, 0.5, 0.0, gravitational_constant=1.0,
                      precision=nb.PRECISION_FP64, bits_per_axis=21)
    s.initialize()
    e0 = sum(s.energy())
    s.run(int(round(periods * steps_per_period)))
    return p, s, e0


def test_two_body_orbit_energy_closure_momentum(nb):
    p, s, e0 = _two_body(nb, 1000)
    st = s.state()
    e1 = sum(s.energy())
    assert abs(e0 - (-0.25 + 0.125)) < 1e-15
    assert abs((e1 - e0) / e0) < 5e-6
    sep = np.hypot(st["x"][1] - st["x"][0], st["y"][1] - st["y"][0])
    assert abs(sep - 1.0) < 2e-3
    closure = np.hypot(st["x"] - p["x"], st["y"] - p["y"]).max()
    assert closure < 2e-2
    com = (st["m"] * st["x"]).sum(), (st["m"] * st["y"]).sum()
    assert abs(com[0]) < 1e-12 and abs(com[1]) < 1e-12


def test_two_body_leapfrog_second_order(nb):
    errs = []
    for spp in (500, 1000):
        p, s, _ = _two_body(nb, spp)
        st = s.state()
        errs.append(np.hypot(st["x"] - p["x"], st["y"] - p["y"]).max())
    order = np.log2(errs[0] / errs[1])
    print(f"leapfrog closure errors {errs}, order {order:.3f}")
    assert 1.8 < order < 2.2
