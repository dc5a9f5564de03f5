"""The mixed-precision force VJP and K-step adjoint (GPU).

The mixed VJP differentiates the mixed forces along the mixed walk (FP32
interactions on fixed-point separations, compensated target sums, FP64
source, node and push-down accumulation).  Pinned by: the numpy FP64
direct-sum VJP at theta = 0 (every node opened, so the walk is the direct
sum and only the FP32 interaction arithmetic differs); the FP64 pipeline's
VJP and adjoint at theta 0.5 and 1.2 and without softening (the same
interaction lists up to opening-test decisions that FP32 and FP64 round
differently); central finite differences of the FP64 K-step evolution.

Tolerances (relative L2 over the compared arrays): 2e-6 against the exact
direct sum at theta = 0; 1e-4 against the FP64 pipeline; 1e-3 against
finite differences of the FP64 evolution (whose gradient the mixed adjoint
approximates to the mixed forces' own accuracy).
"""

import numpy as np
import pytest
import torch

import helpers as H

pytestmark = pytest.mark.gpu

TOL_DIRECT = 2e-6
TOL_FP64 = 1e-4
TOL_FD = 1e-3


def planar(p):
    return np.stack([p["x"], p["y"], p["z"]]), np.stack([p["vx"], p["vy"], p["vz"]])


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def sim_of(nb, p, theta, eps, precision, dt=1e-3):
    s = nb.Simulation(p, dt, theta, eps, gravitational_constant=1.0, precision=precision,
                      bits_per_axis=21)
    s.initialize()
    return s


def vjp(s, ab, with_mass=True):
    out = s.force_vjp(*[torch.from_numpy(a) for a in ab], with_mass=with_mass)
    g = np.stack([t.cpu().numpy() for t in out[:3]])
    return (g, out[3].cpu().numpy()) if with_mass else g


def test_mixed_force_vjp_theta0_close_to_direct(nb):
    p = nb.generate("uniform", 400, 6)
    p["m"] = np.random.default_rng(7).uniform(0.5, 1.5, 400) / 400
    pos, _ = planar(p)
    ab = np.random.default_rng(0).normal(size=pos.shape)
    g, gm = vjp(sim_of(nb, p, 0.0, 0.02, nb.PRECISION_MIXED), ab)
    eg = rel(g, H.direct_vjp_numpy(p["m"], pos, ab, 0.02))
    em = rel(gm, H.direct_mass_vjp_numpy(p["m"], pos, ab, 0.02))
    print(f"mixed VJP at theta 0 vs FP64 direct: positions {eg:.2e}, masses {em:.2e}")
    assert eg < TOL_DIRECT
    assert em < TOL_DIRECT


@pytest.mark.parametrize("theta,eps,kind,n", [
    (0.5, 0.01, "plummer", 20000),    # split walk: marked subtree records
    (0.5, 0.01, "plummer", 300000),   # split walk, 8 wavefronts per SIMD
    (0.5, 0.01, "plummer", 1100000),  # whole walks
    (1.2, 0.01, "clumps", 20000),     # cell containment test
    (0.5, 0.0, "plummer", 20000),     # no softening: leaves walked exactly
])
def test_mixed_force_vjp_close_to_fp64(nb, theta, eps, kind, n):
    p = nb.generate(kind, n, 13)
    p["m"] = p["m"] * np.random.default_rng(9).uniform(0.5, 1.5, n)
    pos, _ = planar(p)
    ab = np.random.default_rng(10).normal(size=pos.shape)
    g64, gm64 = vjp(sim_of(nb, p, theta, eps, nb.PRECISION_FP64), ab)
    sm = sim_of(nb, p, theta, eps, nb.PRECISION_MIXED)
    g, gm = vjp(sm, ab)
    eg, em = rel(g, g64), rel(gm, gm64)
    print(f"mixed VJP vs FP64 ({kind} N={n} theta {theta} eps {eps}): positions {eg:.2e}, "
          f"masses {em:.2e}")
    assert eg < TOL_FP64
    assert em < TOL_FP64
    # a second product on the same (now walked) layout gives the same result
    g2, gm2 = vjp(sm, ab)
    assert rel(g2, g) < 1e-12
    assert rel(gm2, gm) < 1e-12


def _disk(nb, n=2048):
    p = nb.generate("disk", n, 9)
    pos, vel = planar(p)
    return p, pos, vel


def _adjoint(nb, n, K, every, precision):
    return nb.Adjoint(n, 1e-3, 0.5, 0.01, K, every, gravitational_constant=1.0,
                      precision=precision)


def test_mixed_adjoint_close_to_fp64_adjoint(nb):
    n, K, every = 2048, 20, 5
    p, pos, vel = _disk(nb, n)
    rng = np.random.default_rng(3)
    w, u = rng.normal(size=pos.shape), rng.normal(size=pos.shape)
    res = {}
    for prec in (nb.PRECISION_FP64, nb.PRECISION_MIXED):
        adj = _adjoint(nb, n, K, every, prec)
        out = adj.forward(*[torch.from_numpy(a) for a in (p["m"], *pos, *vel)])
        g = adj.backward(*[torch.from_numpy(a) for a in (*w, *u)], with_mass=True)
        res[prec] = ([t.cpu().numpy() for t in out], [t.cpu().numpy() for t in g])
        adj.close()
    (o64, g64), (om, gmx) = res[nb.PRECISION_FP64], res[nb.PRECISION_MIXED]
    ex = rel(np.stack(om[:3]), np.stack(o64[:3]))
    eg = rel(np.stack(gmx[:3]), np.stack(g64[:3]))
    ev = rel(np.stack(gmx[3:6]), np.stack(g64[3:6]))
    em = rel(gmx[6], g64[6])
    print(f"mixed vs FP64 adjoint (disk N={n}, K={K}): final positions {ex:.2e}, "
          f"x0 cotangent {eg:.2e}, v0 cotangent {ev:.2e}, mass cotangent {em:.2e}")
    assert ex < 1e-5
    assert eg < TOL_FP64 and ev < TOL_FP64 and em < TOL_FP64


def test_mixed_adjoint_matches_fp64_finite_differences(nb):
    n, K, every = 2048, 20, 5
    p, pos, vel = _disk(nb, n)
    rng = np.random.default_rng(3)
    w, u = rng.normal(size=pos.shape), rng.normal(size=pos.shape)
    m = torch.from_numpy(p["m"])
    ref = _adjoint(nb, n, K, every, nb.PRECISION_FP64)

    def final(q, v):
        out = ref.forward(m, *[torch.from_numpy(a) for a in (*q, *v)])
        return np.stack([t.cpu().numpy() for t in out])

    base = final(pos, vel)

    def loss(q, v):
        d = final(q, v) - base
        return (d[:3] * w).sum() + (d[3:] * u).sum()

    adj = _adjoint(nb, n, K, every, nb.PRECISION_MIXED)
    adj.forward(m, *[torch.from_numpy(a) for a in (*pos, *vel)])
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
        h = 1e-8
        d1 = (f(h) - f(-h)) / (2 * h)
        d2 = (f(2 * h) - f(-2 * h)) / (4 * h)
        fd.append(d1)
        ok.append(abs(d1 - d2) <= 3e-5 * max(1.0, abs(d1)))
        ad.append((gx if on_pos else gv)[c, i])
    ad, fd, ok = np.asarray(ad), np.asarray(fd), np.asarray(ok)
    err = rel(ad[ok], fd[ok])
    print(f"mixed adjoint vs FP64 FD (disk N={n}, K={K}): {ok.sum()}/64 probes smooth, "
          f"rel error {err:.2e}")
    assert ok.sum() >= 56
    assert err < TOL_FD


def test_mixed_adjoint_full_size_velocity_probes(nb):
    # the differentiable workload at its own size: disk + bulge, N = 262,144,
    # K = 100, checkpoints every 10; the mixed gradient of
    # L = sum (r(T) - r*)^2 with respect to 8 initial velocity components
    # against central differences of the FP64 evolution
    n, K, every = 262_144, 100, 10
    p = nb.generate("disk", n, 42)
    state = [torch.from_numpy(p[k]).cuda() for k in ("x", "y", "z", "vx", "vy", "vz")]
    m = torch.from_numpy(p["m"]).cuda()
    ref = _adjoint(nb, n, K, every, nb.PRECISION_FP64)

    def final(st):
        return np.stack([t.cpu().numpy() for t in ref.forward(m, *st)[:3]])

    r_end = final(state)
    target = r_end + 1e-3 * np.random.default_rng(1).normal(size=r_end.shape)

    def loss(st):
        d = final(st) - r_end
        return float((d * (d + 2.0 * (r_end - target))).sum())

    adj = _adjoint(nb, n, K, every, nb.PRECISION_MIXED)
    out = adj.forward(m, *state)
    cot = [torch.from_numpy(2.0 * (o.cpu().numpy() - target[k])).cuda()
           for k, o in enumerate(out[:3])]
    zeros = [torch.zeros_like(state[0]) for _ in range(3)]
    g = adj.backward(*cot, *zeros)
    gv = np.stack([t.cpu().numpy() for t in g[3:6]])
    rng = np.random.default_rng(2)
    errs, smooth = [], 0
    for _ in range(8):
        i, c, h = int(rng.integers(n)), int(rng.integers(3)), 1e-7

        def f(d):
            st = [t.clone() for t in state]
            st[3 + c][i] += d
            return loss(st)
        d1 = (f(h) - f(-h)) / (2 * h)
        d2 = (f(2 * h) - f(-2 * h)) / (4 * h)
        if abs(d1 - d2) <= 1e-4 * abs(d1):
            smooth += 1
            errs.append(abs(gv[c, i] - d1) / abs(d1))
    print(f"mixed adjoint (disk N={n}, K={K}) vs FP64 FD on 8 velocity components: "
          f"{smooth} smooth, max rel error {max(errs) if errs else float('nan'):.2e}")
    assert smooth >= 4
    assert max(errs) < TOL_FD
