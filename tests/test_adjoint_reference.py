"""The numpy direct-sum adjoint that the GPU adjoint is checked against,
pinned on the CPU by central finite differences, and the truncated-Plummer
energy integral pinned by its closed-form limit -3 pi / 32."""

import numpy as np

import helpers as H


def _system(n=16, seed=0):
    rng = np.random.default_rng(seed)
    m = rng.uniform(0.5, 1.5, n) / n
    return m, rng.normal(size=(3, n)), 0.1 * rng.normal(size=(3, n)), rng


def test_direct_vjp_matches_finite_differences():
    m, pos, _, rng = _system()
    eps = 0.05
    ab = rng.normal(size=pos.shape)
    g = H.direct_vjp_numpy(m, pos, ab, eps)
    h = 1e-6
    fd = np.zeros_like(pos)
    for c in range(3):
        for i in range(pos.shape[1]):
            p1, p2 = pos.copy(), pos.copy()
            p1[c, i] += h
            p2[c, i] -= h
            a1 = np.stack(H.direct_numpy(m, *p1, eps))
            a2 = np.stack(H.direct_numpy(m, *p2, eps))
            fd[c, i] = ((a1 - a2) * ab).sum() / (2 * h)
    assert np.linalg.norm(g - fd) / np.linalg.norm(fd) < 1e-8


def test_direct_mass_vjp_matches_finite_differences():
    m, pos, _, rng = _system(14, 3)
    eps = 0.05
    ab = rng.normal(size=pos.shape)
    g = H.direct_mass_vjp_numpy(m, pos, ab, eps)
    h = 1e-7
    fd = np.zeros_like(m)
    for j in range(len(m)):
        m1, m2 = m.copy(), m.copy()
        m1[j] += h
        m2[j] -= h
        a1 = np.stack(H.direct_numpy(m1, *pos, eps))
        a2 = np.stack(H.direct_numpy(m2, *pos, eps))
        fd[j] = ((a1 - a2) * ab).sum() / (2 * h)
    # a is linear in m: the central difference is exact up to rounding
    assert np.linalg.norm(g - fd) / np.linalg.norm(fd) < 1e-7


def test_direct_adjoint_mass_gradient_matches_finite_differences():
    m, pos, vel, rng = _system(10, 4)
    eps, dt, K = 0.05, 1e-2, 5
    xbar, vbar = rng.normal(size=pos.shape), rng.normal(size=pos.shape)
    _, _, mb = H.leapfrog_adjoint_numpy(m, pos, vel, dt, K, eps, xbar, vbar,
                                        with_mass=True)

    def loss(mm):
        xs, vk = H.leapfrog_numpy(mm, pos, vel, dt, K, eps)
        return (xs[-1] * xbar).sum() + (vk * vbar).sum()

    h = 1e-7
    fd = np.zeros_like(m)
    for j in range(len(m)):
        m1, m2 = m.copy(), m.copy()
        m1[j] += h
        m2[j] -= h
        fd[j] = (loss(m1) - loss(m2)) / (2 * h)
    assert np.linalg.norm(mb - fd) / np.linalg.norm(fd) < 1e-6


def test_direct_adjoint_matches_finite_differences():
    m, pos, vel, rng = _system(12, 1)
    eps, dt, K = 0.05, 1e-2, 6
    xbar, vbar = rng.normal(size=pos.shape), rng.normal(size=pos.shape)
    xb, vb = H.leapfrog_adjoint_numpy(m, pos, vel, dt, K, eps, xbar, vbar)

    def loss(p, v):
        xs, vk = H.leapfrog_numpy(m, p, v, dt, K, eps)
        return (xs[-1] * xbar).sum() + (vk * vbar).sum()

    h = 1e-6
    fdx, fdv = np.zeros_like(pos), np.zeros_like(vel)
    for c in range(3):
        for i in range(pos.shape[1]):
            for arr, fd, is_pos in ((pos, fdx, True), (vel, fdv, False)):
                a1, a2 = arr.copy(), arr.copy()
                a1[c, i] += h
                a2[c, i] -= h
                l1 = loss(a1, vel) if is_pos else loss(pos, a1)
                l2 = loss(a2, vel) if is_pos else loss(pos, a2)
                fd[c, i] = (l1 - l2) / (2 * h)
    assert np.linalg.norm(xb - fdx) / np.linalg.norm(fdx) < 1e-7
    assert np.linalg.norm(vb - fdv) / np.linalg.norm(fdv) < 1e-7


def test_direct_potential_matches_pair_sum():
    m, pos, _, _ = _system(9, 2)
    eps = 0.1
    phi = H.direct_potential_numpy(m, *pos, eps)
    for i in range(9):
        ref = -sum(m[j] / np.sqrt(((pos[:, i] - pos[:, j]) ** 2).sum() + eps * eps)
                   for j in range(9) if j != i)
        assert abs(phi[i] - ref) < 1e-14 * abs(ref)


def test_truncated_plummer_energy_limit():
    assert abs(H.truncated_plummer_energy(1e4) - (-3 * np.pi / 32)) < 1e-7
    # truncation at 10 a removes the outer 1.5% of the mass
    w10 = H.truncated_plummer_energy(10.0)
    assert -0.31 < w10 < -0.295
