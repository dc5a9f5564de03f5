"""Independent checks shared by the oracle and GPU parity tests.

Each check is an independent statement of what the reference's tests assert
(tests/unit/cuda/tree_builder_test.cu:243-386, tests/integration/cuda/
sparse_octree_test.cu:207-324) or of what the mathematics fixes (brute-force
octree from key prefixes, subtree sums, direct sums), written without the
oracle's code.
"""

import numpy as np


def spread_bits(q, bits):
    """Bit k of q moves to bit 3k, via string manipulation."""
    s = format(int(q), f"0{bits}b")
    return int("".join("00" + c for c in s), 2)


def morton_key(qx, qy, qz, bits):
    return (spread_bits(qx, bits) << 2) | (spread_bits(qy, bits) << 1) | spread_bits(qz, bits)


def lcp(keys, i, j, width):
    n = len(keys)
    i, j = int(i), int(j)
    if j < 0 or j >= n:
        return -1
    if i == j:
        return 2 * width
    a, b = int(keys[i]), int(keys[j])
    if a != b:
        return width - (a ^ b).bit_length()
    return width + 32 - (i ^ j).bit_length()


def validate_karras(t, keys, width):
    """Topology validator of tree_builder_test.cu:243-386."""
    n = len(keys)
    total = 2 * n - 1
    parent = np.asarray(t["parent"])
    assert len(parent) == total
    if n == 1:
        assert parent[0] == -1
        return
    left, right = np.asarray(t["left"]), np.asarray(t["right"])
    incoming = np.zeros(total, dtype=int)
    for i in range(n - 1):
        node = n + i
        for c in (left[i], right[i]):
            assert 0 <= c < total
            assert parent[c] == node
            incoming[c] += 1
        assert left[i] != right[i]
    roots = [k for k in range(n, total) if parent[k] == -1]
    assert roots == [n]
    assert incoming[n] == 0 and all(incoming[k] == 1 for k in range(total) if k != n)
    seen = np.zeros(total, dtype=bool)

    def visit(node):
        assert not seen[node]
        seen[node] = True
        if node < n:
            return node, node
        i = node - n
        l0, l1 = visit(left[i])
        r0, r1 = visit(right[i])
        assert l1 + 1 == r0
        assert t["range_first"][i] == l0 and t["range_last"][i] == r1
        assert t["prefix_length"][i] == lcp(keys, l0, r1, width)
        return l0, r1

    import sys
    sys.setrecursionlimit(100000)
    assert visit(n) == (0, n - 1)
    assert seen.all()


def brute_octree(unique, bits):
    """Sparse octree by definition: every prefix of every key down to the
    shallowest level at which the key's prefix is unique; returns
    {(level, prefix): group or None} and the parent map."""
    spatial = 3 * bits
    keys = [int(k) for k in unique]
    g = len(keys)

    def pre(k, level):
        if level == 0:
            return 0
        shift = spatial - 3 * level
        return (k >> shift) << shift

    nodes = {(0, 0): None}
    parent = {}
    for gi, k in enumerate(keys):
        # shallowest level separating k from its sorted neighbours
        shared = 0
        for nb in (gi - 1, gi + 1):
            if 0 <= nb < g:
                common = spatial - (k ^ keys[nb]).bit_length()
                shared = max(shared, common // 3)
        leaf_level = shared + 1
        for level in range(1, leaf_level + 1):
            node = (level, pre(k, level))
            if level == leaf_level:
                nodes[node] = gi
            else:
                nodes.setdefault(node, None)
            parent[node] = (level - 1, pre(k, level - 1))
    return nodes, parent


def octree_matches_brute(o, unique, gfirst, gcount, bits):
    """Structural equality of an octree (arrays) with the brute-force one."""
    nodes, bparent = brute_octree(unique, bits)
    M = len(o["parent"])
    assert M == len(nodes)
    ids = {}
    for v in range(M):
        key = (int(o["level"][v]), int(o["prefix"][v]))
        assert key in nodes, key
        assert key not in ids
        ids[key] = v
    for key, v in ids.items():
        if key == (0, 0):
            assert o["parent"][v] == -1
            continue
        assert o["parent"][v] == ids[bparent[key]]
        g = nodes[key]
        if g is None:
            assert o["count"][v] == 0 and o["first"][v] == -1
        else:
            assert o["count"][v] == gcount[g] and o["first"][v] == gfirst[g]
    spatial = 3 * bits
    for v in range(M):
        for oc in range(8):
            c = o["children"][8 * v + oc]
            if c == -1:
                continue
            assert o["parent"][c] == v
            assert o["level"][c] == o["level"][v] + 1
            shift = spatial - 3 * int(o["level"][c])
            assert (int(o["prefix"][c]) >> shift) & 7 == oc


def subtree_sums(o, idx, m, x, y, z):
    """Mass and centre of mass of every node by summing its particles."""
    M = len(o["parent"])
    mass = np.zeros(M)
    wx, wy, wz = np.zeros(M), np.zeros(M), np.zeros(M)
    for v in range(M):
        c = o["count"][v]
        if c <= 0:
            continue
        f = o["first"][v]
        p = idx[f:f + c]
        pm = m[p].sum()
        px, py, pz = (m[p] * x[p]).sum(), (m[p] * y[p]).sum(), (m[p] * z[p]).sum()
        node = v
        while node != -1:
            mass[node] += pm
            wx[node] += px
            wy[node] += py
            wz[node] += pz
            node = o["parent"][node]
    with np.errstate(invalid="ignore", divide="ignore"):
        return mass, np.where(mass > 0, wx / mass, 0), np.where(mass > 0, wy / mass, 0), \
            np.where(mass > 0, wz / mass, 0)


def direct_numpy(m, x, y, z, eps, G=1.0):
    """Softened direct sum, vectorised."""
    dx = x[None, :] - x[:, None]
    dy = y[None, :] - y[:, None]
    dz = z[None, :] - z[:, None]
    r2 = dx * dx + dy * dy + dz * dz + eps * eps
    np.fill_diagonal(r2, np.inf)
    inv3 = r2 ** -1.5
    f = G * m[None, :] * inv3
    return (f * dx).sum(1), (f * dy).sum(1), (f * dz).sum(1)


def rel_l2(a, b):
    a = np.stack(a)
    b = np.stack(b)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def far_cluster_tree(center2_x=8.0):
    """Hand-built 5-node tree of barnes_hut_test.cu:568-800 / 784-1025."""
    children = np.full(40, -1, dtype=np.int32)
    children[0 * 8 + 0] = 1
    children[0 * 8 + 7] = 2
    children[2 * 8 + 0] = 3
    children[2 * 8 + 7] = 4
    return dict(
        children=children,
        parent=np.array([-1, 0, 0, 2, 2], dtype=np.int32),
        first=np.array([-1, 0, -1, 1, 2], dtype=np.int32),
        count=np.array([0, 1, 0, 1, 1], dtype=np.int32),
        mass=np.array([6.0, 1.0, 5.0, 2.0, 3.0]),
        com_x=np.array([46.0 / 6.0, -10.0, 11.2, 10.0, 12.0]),
        com_y=np.zeros(5), com_z=np.zeros(5),
        center_x=np.array([0.0, -8.0, center2_x, 10.0, 12.0]),
        center_y=np.zeros(5), center_z=np.zeros(5),
        half=np.array([16.0, 8.0, 8.0, 1.0, 1.0]),
    )


def single_particle_tree():
    """barnes_hut_test.cu:167-292: one node that is an occupied leaf."""
    return dict(
        children=np.full(8, -1, dtype=np.int32), parent=np.array([-1], dtype=np.int32),
        first=np.array([0], dtype=np.int32), count=np.array([1], dtype=np.int32),
        mass=np.array([2.0]), com_x=np.array([3.0]), com_y=np.array([4.0]),
        com_z=np.array([5.0]), center_x=np.array([3.0]), center_y=np.array([4.0]),
        center_z=np.array([5.0]), half=np.array([1.0]))


# ------------------------------------------------------ direct-sum adjoint
def direct_potential_numpy(m, x, y, z, eps, G=1.0):
    """Softened potential phi_i = -sum_j G m_j (r^2 + eps^2)^{-1/2}, j != i."""
    dx = x[None, :] - x[:, None]
    dy = y[None, :] - y[:, None]
    dz = z[None, :] - z[:, None]
    r2 = dx * dx + dy * dy + dz * dz + eps * eps
    np.fill_diagonal(r2, np.inf)
    return -(G * m[None, :] / np.sqrt(r2)).sum(1)


def direct_vjp_numpy(m, pos, abar, eps, G=1.0):
    """(da/dx)^T abar of the softened direct sum.  pos, abar: (3, N).

    a_i = sum_j G m_j d_ij r_ij^-3 with d_ij = x_j - x_i; the Jacobian block
    J_ij = G m_j (r^-3 I - 3 r^-5 d d^T) is symmetric, so xbar_j += J_ij abar_i
    and xbar_i -= J_ij abar_i.  Written from the closed form, independently of
    the tree code."""
    d = pos[:, None, :] - pos[:, :, None]          # (3, i, j) = x_j - x_i
    r2 = (d * d).sum(0) + eps * eps
    np.fill_diagonal(r2, np.inf)
    inv3 = r2 ** -1.5
    inv5 = inv3 / r2
    gm = G * m[None, :]
    dot = (d * abar[:, :, None]).sum(0)            # d_ij . abar_i
    g = gm[None] * (inv3[None] * abar[:, :, None] - 3.0 * inv5[None] * dot[None] * d)
    return g.sum(1) - g.sum(2)                     # sources gain, targets lose


def direct_mass_vjp_numpy(m, pos, abar, eps, G=1.0):
    """(da/dm)^T abar of the softened direct sum.  a_i = sum_j G m_j d_ij
    r_ij^-3 is linear in each source mass, so mbar_j = sum_{i != j} G
    abar_i . d_ij r_ij^-3 (d_ij = x_j - x_i).  Written from the closed form."""
    d = pos[:, None, :] - pos[:, :, None]          # (3, i, j) = x_j - x_i
    r2 = (d * d).sum(0) + eps * eps
    np.fill_diagonal(r2, np.inf)
    dot = (d * abar[:, :, None]).sum(0)            # d_ij . abar_i
    return G * (dot * r2 ** -1.5).sum(0)


def leapfrog_numpy(m, pos, vel, dt, steps, eps, G=1.0):
    """KDK with the direct sum; returns the positions x_0..x_K and the final
    velocities."""
    def acc(p):
        return np.stack(direct_numpy(m, p[0], p[1], p[2], eps, G))
    xs = [pos.copy()]
    a = acc(pos)
    x, v = pos.copy(), vel.copy()
    for _ in range(steps):
        v = v + 0.5 * dt * a
        x = x + dt * v
        a = acc(x)
        v = v + 0.5 * dt * a
        xs.append(x.copy())
    return xs, v


def leapfrog_adjoint_numpy(m, pos, vel, dt, steps, eps, xbar, vbar, G=1.0,
                           with_mass=False):
    """Reverse sweep of leapfrog_numpy: cotangents of (x_0, v_0), and of the
    masses with with_mass, from those of (x_K, v_K).  Every force
    evaluation a(x_j) adds (da/dm)^T of its cotangent to mbar."""
    xs, _ = leapfrog_numpy(m, pos, vel, dt, steps, eps, G)
    xb, vb = xbar.copy(), vbar.copy()
    mb = np.zeros_like(m)
    for j in range(steps, 0, -1):
        ab = (0.5 * dt if j == steps else dt) * vb
        xb = xb + direct_vjp_numpy(m, xs[j], ab, eps, G)
        mb = mb + direct_mass_vjp_numpy(m, xs[j], ab, eps, G)
        vb = vb + dt * xb
    ab = 0.5 * dt * vb
    xb = xb + direct_vjp_numpy(m, xs[0], ab, eps, G)
    mb = mb + direct_mass_vjp_numpy(m, xs[0], ab, eps, G)
    return (xb, vb, mb) if with_mass else (xb, vb)


def truncated_plummer_energy(rmax=10.0, G=1.0):
    """Potential energy of a unit-mass Plummer sphere (a = 1) truncated at
    rmax and renormalised: W = -G int_0^R M(r) M'(r) / r dr with
    M(r) = r^3 (1 + r^2)^{-3/2} / M(R).  For R -> inf this is -3 pi / 32."""
    from scipy.integrate import quad
    mr = rmax ** 3 / (1.0 + rmax ** 2) ** 1.5

    def integrand(r):
        mass = r ** 3 / (1.0 + r * r) ** 1.5 / mr
        dm = 3.0 * r * r / (1.0 + r * r) ** 2.5 / mr
        return mass * dm / r

    val, _ = quad(integrand, 0.0, rmax, limit=200)
    return -G * val


def hilbert_keys(qx, qy, qz, bits=21):
    """3-D Hilbert curve indices of integer grid coordinates in [0, 2^bits):
    J. Skilling's axes-to-transpose transform ("Programming the Hilbert
    curve", AIP Conf. Proc. 707, 2004), then the transposed bits read from
    the most significant level down, x, y, z.  Vectorised over arrays."""
    X = [np.asarray(q, dtype=np.uint64).copy() for q in (qx, qy, qz)]
    M = np.uint64(1 << (bits - 1))
    Q = int(M)
    while Q > 1:
        q = np.uint64(Q)
        P = np.uint64(Q - 1)
        for i in range(3):
            hi = (X[i] & q) != 0
            # invert where the bit is set, else exchange the low bits of X0, Xi
            t = (X[0] ^ X[i]) & P
            x0_inv = X[0] ^ P
            x0_ex = X[0] ^ t
            xi_ex = X[i] ^ t
            if i == 0:
                X[0] = np.where(hi, x0_inv, X[0])
            else:
                X[0] = np.where(hi, x0_inv, x0_ex)
                X[i] = np.where(hi, X[i], xi_ex)
        Q >>= 1
    X[1] ^= X[0]
    X[2] ^= X[1]
    t = np.zeros_like(X[0])
    Q = int(M)
    while Q > 1:
        t ^= np.where((X[2] & np.uint64(Q)) != 0, np.uint64(Q - 1), np.uint64(0))
        Q >>= 1
    for i in range(3):
        X[i] ^= t
    key = np.zeros_like(X[0])
    one = np.uint64(1)
    for b in range(bits - 1, -1, -1):
        s = np.uint64(b)
        key = (key << np.uint64(3)) | (((X[0] >> s) & one) << np.uint64(2)) | \
            (((X[1] >> s) & one) << one) | ((X[2] >> s) & one)
    return key


def morton_decode(keys, bits=21):
    """Grid coordinates of Morton keys (x in the high bit of each triple)."""
    k = np.asarray(keys, dtype=np.uint64)
    q = [np.zeros_like(k) for _ in range(3)]
    one = np.uint64(1)
    for b in range(bits):
        for a in range(3):
            q[a] |= ((k >> np.uint64(3 * b + 2 - a)) & one) << np.uint64(b)
    return q
