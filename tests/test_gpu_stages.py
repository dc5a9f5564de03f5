"""Stage-by-stage parity of the HIP library with the oracle (GPU).

Integer and index results (box, keys, permutation, groups, radix tree, plan,
octree topology) must match bit for bit; FP64 geometry, mass/COM and
accelerations must match bit for bit too, because both sides state every
fused multiply-add explicitly.
"""

import numpy as np
import pytest
import torch

import helpers as H

pytestmark = pytest.mark.gpu


def dev(a, dtype=None):
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a.copy()).cuda()


def host(t, dtype=None):
    a = t.cpu().numpy()
    return a if dtype is None else a.view(dtype)


def xyz(seed, n, kind="normal"):
    rng = np.random.default_rng(seed)
    if kind == "normal":
        return rng.normal(size=(3, n))
    if kind == "cubed":
        return rng.normal(size=(3, n)) ** 3
    return rng.uniform(-1, 1, size=(3, n))


# ----------------------------------------------------------------- bbox
@pytest.mark.parametrize("coords,expected", [
    (([-2.0, 4.0, 1.0], [10.0, 14.0, -2.0], [-1.0, 0.0, 5.0]), (1.0, 6.0, 2.0, 17.6)),
    (([-2.0, 4.0], [10.0, 14.0], [-1.0, 5.0]), (1.0, 12.0, 2.0, 6.6)),
    (([3.5, 3.5], [-4.0, -4.0], [8.25, 8.25]), (3.5, -4.0, 8.25, 1.0)),
    (([10.0, 14.0], [20.0, 22.0], [30.0, 31.0]), (12.0, 21.0, 30.5, 4.4)),
])
def test_bbox_known_answers(nb, orc, coords, expected):
    x, y, z = (dev(c, np.float64) for c in coords)
    b = nb.Bbox(x, y, z)
    v = b.values()
    assert (v.center_x, v.center_y, v.center_z, v.side) == tuple(orc.bbox(*coords))
    np.testing.assert_allclose([v.center_x, v.center_y, v.center_z, v.side], expected,
                               rtol=4e-16)


def test_bbox_default_sentinel(nb):
    v = nb.Bbox().values()
    assert (v.center_x, v.center_y, v.center_z, v.side) == (0.0, 0.0, 0.0, -1.0)


@pytest.mark.parametrize("n", [1, 2, 255, 256, 1000, 300_001])
def test_bbox_matches_oracle(nb, orc, n):
    x, y, z = xyz(n, n, "cubed")
    b = nb.Bbox(dev(x), dev(y), dev(z))
    assert np.array_equal(host(b.data), orc.bbox(x, y, z))


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_bbox_non_finite(nb, bad):
    x, y, z = xyz(1, 5000)
    x[4321] = bad
    with pytest.raises(nb.DomainError):
        nb.Bbox(dev(x), dev(y), dev(z))


def test_bbox_empty(nb):
    with pytest.raises(nb.InvalidArgument):
        nb.Bbox(dev([1.0]), dev([1.0]), dev([1.0]), count=0)


# --------------------------------------------------------------- Morton
@pytest.mark.parametrize("bits", [10, 21])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 1000, 65_537, 400_003])
def test_morton_matches_oracle(nb, orc, bits, n):
    x, y, z = xyz(100 + n, n, "cubed")
    if n > 10:
        x[::13] = x[0]  # duplicate positions -> equal keys, stability matters
        y[::13] = y[0]
        z[::13] = z[0]
    mk = nb.MortonKeys(dev(x), dev(y), dev(z), bits_per_axis=bits)
    keys, idx = orc.morton_keys(x, y, z, orc.bbox(x, y, z), bits)
    assert np.array_equal(mk.host_keys(), keys)
    assert np.array_equal(mk.host_indices(), idx)


@pytest.mark.parametrize("bits", [10, 21])
def test_morton_known_permutations(nb, bits):
    # morton_test.cu:252-366
    x = y = z = np.array([1.0, 1.0, 0.0])
    mk = nb.MortonKeys(dev(x), dev(y), dev(z), bits_per_axis=bits)
    assert list(mk.host_indices()) == [2, 0, 1]
    x = np.full(4, 3.5)
    mk = nb.MortonKeys(dev(x), dev(np.full(4, -4.0)), dev(np.full(4, 8.25)),
                       bits_per_axis=bits)
    assert list(mk.host_indices()) == [0, 1, 2, 3]
    x = y = z = np.array([0.0, 1.0])
    mk = nb.MortonKeys(dev(x), dev(y), dev(z), bits_per_axis=bits)
    assert list(mk.host_indices()) == [0, 1]


def test_morton_reused_box_identical(nb):
    x, y, z = xyz(3, 777)
    X, Y, Z = dev(x), dev(y), dev(z)
    a = nb.MortonKeys(X, Y, Z)
    b = nb.MortonKeys(X, Y, Z, bbox=nb.Bbox(X, Y, Z))
    assert np.array_equal(a.host_keys(), b.host_keys())


def test_morton_rejects(nb):
    with pytest.raises(nb.InvalidArgument):
        nb.MortonKeys(dev([1.0]), dev([1.0]), dev([1.0]), count=0)
    x, y, z = xyz(2, 10)
    x[3] = np.nan
    with pytest.raises(nb.DomainError):
        nb.MortonKeys(dev(x), dev(y), dev(z))


@pytest.mark.parametrize("width,bits,n", [(4, 30, 1), (4, 32, 70_001), (8, 63, 500_000),
                                          (8, 64, 3), (4, 8, 100_000)])
def test_radix_sort_pairs_stable(nb, width, bits, n):
    rng = np.random.default_rng(n)
    dt = np.uint32 if width == 4 else np.uint64
    hi = (1 << bits) - 1
    keys = rng.integers(0, hi, size=n, dtype=np.uint64, endpoint=True).astype(dt)
    keys[::5] = keys[0]
    vals = np.arange(n, dtype=np.uint32)
    K, V = dev(keys), dev(vals)
    nb.radix_sort_pairs(K, V, bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(host(K, dt), keys[order])
    assert np.array_equal(host(V, np.uint32), order.astype(np.uint32))


# ---------------------------------------------------------- leaf groups
@pytest.mark.parametrize("keys", [
    [42], [2, 5, 8, 10], [7, 7, 7, 7], [1, 1, 2, 4, 4, 4, 9],
    [0, 0, 1, 0xFFFFFFFF, 0xFFFFFFFF]])
def test_leaf_groups_known_answers(nb, orc, keys):
    k = np.array(keys, dtype=np.uint32)
    g = nb.MortonLeafGroups()
    nb.allocate_morton_leaf_groups(g, len(k), 4)
    nb.build_morton_leaf_groups(g, dev(k), len(k))
    u, f, c = orc.leaf_groups(k)
    assert g.n_groups == len(u)
    assert np.array_equal(host(g.unique_keys, np.uint32)[: g.n_groups], u)
    assert np.array_equal(host(g.first_particle)[: g.n_groups], f)
    assert np.array_equal(host(g.particle_count)[: g.n_groups], c)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [2047, 2048, 2049, 300_000])
def test_leaf_groups_random(nb, orc, width, n):
    rng = np.random.default_rng(n + width)
    dt = np.uint32 if width == 4 else np.uint64
    k = np.sort(rng.integers(0, n // 3 + 1, size=n).astype(dt))
    g = nb.MortonLeafGroups()
    nb.allocate_morton_leaf_groups(g, n, width)
    nb.build_morton_leaf_groups(g, dev(k), n)
    u, f, c = orc.leaf_groups(k)
    assert g.n_groups == len(u)
    assert np.array_equal(host(g.unique_keys, dt)[: len(u)], u)
    assert np.array_equal(host(g.first_particle)[: len(u)], f)
    assert np.array_equal(host(g.particle_count)[: len(u)], c)


def test_leaf_groups_capacity(nb):
    g = nb.MortonLeafGroups()
    nb.allocate_morton_leaf_groups(g, 2, 4)
    with pytest.raises(nb.InvalidArgument):
        nb.build_morton_leaf_groups(g, dev(np.array([1, 2, 3], dtype=np.uint32)), 3)


# ----------------------------------------------------------- Karras
def gpu_tree(nb, keys):
    t = nb.Tree()
    nb.allocate_tree_topology(t, len(keys))
    nb.build_tree(t, dev(keys), len(keys))
    out = {"parent": host(t.parent)}
    for f in ("left", "right", "range_first", "range_last", "prefix_length"):
        out[f] = host(getattr(t, f)) if len(keys) > 1 else np.zeros(0, np.int32)
    return out


@pytest.mark.parametrize("keys", [
    np.array([42], dtype=np.uint32),
    np.array([0, 0x80000000], dtype=np.uint32),
    np.array([0, 0x40000000, 0x80000000, 0xC0000000], dtype=np.uint32),
])
def test_karras_known_topologies(nb, orc, keys):
    g = gpu_tree(nb, keys)
    o = orc.tree_build(keys)
    for f in o:
        assert np.array_equal(g[f], o[f]), f


@pytest.mark.parametrize("n", [3, 5, 7, 16, 31, 64, 257, 100_000])
def test_karras_random_matches_oracle(nb, orc, n):
    keys = orc.random_keys(n)
    g = gpu_tree(nb, keys)
    o = orc.tree_build(keys)
    for f in o:
        assert np.array_equal(g[f], o[f]), f
    if n <= 257:
        H.validate_karras(g, keys, 32)


def test_karras_64bit(nb, orc):
    rng = np.random.default_rng(9)
    keys = np.sort(rng.integers(0, 2**63, size=50_000, dtype=np.uint64))
    keys[100:105] = keys[99]
    g = gpu_tree(nb, keys)
    o = orc.tree_build(keys)
    for f in o:
        assert np.array_equal(g[f], o[f]), f


def test_legacy_center_of_mass(nb, orc):
    n = 3001
    keys = orc.random_keys(n)
    rng = np.random.default_rng(4)
    idx = rng.permutation(n).astype(np.uint32)
    m = rng.uniform(0.5, 2, n)
    x, y, z = rng.normal(size=(3, n))
    t = nb.Tree()
    nb.allocate_tree(t, n)
    nb.build_tree(t, dev(keys), n)
    nb.compute_center_of_mass(t, dev(idx), dev(m), dev(x), dev(y), dev(z), n)
    o = orc.tree_com(orc.tree_build(keys), n, idx, m, x, y, z)
    for f in ("mass", "com_x", "com_y", "com_z"):
        assert np.array_equal(host(getattr(t, f)), o[f]), f
    assert np.array_equal(host(t.visit_count), o["visit_count"])


# ------------------------------------------------ plan, octree, geometry
def gpu_build(nb, x, y, z, m, bits):
    X, Y, Z, Mm = dev(x), dev(y), dev(z), dev(m)
    box = nb.Bbox(X, Y, Z)
    mk = nb.MortonKeys(X, Y, Z, bbox=box, bits_per_axis=bits)
    n = len(x)
    g = nb.MortonLeafGroups()
    nb.allocate_morton_leaf_groups(g, n, 4 if bits == 10 else 8)
    nb.build_morton_leaf_groups(g, mk.keys, n)
    t = nb.Tree()
    nb.allocate_tree_topology(t, g.n_groups)
    nb.build_tree_from_morton_groups(t, g)
    p = nb.RadixToOctreePlan()
    nb.allocate_radix_to_octree_plan(p, 2 * g.n_groups - 1)
    nb.build_radix_to_octree_plan(p, t, g, bits)
    o = nb.Octree(bits)
    nb.allocate_octree_topology(o, p.n_octree_nodes)
    nb.build_sparse_octree_topology(o, t, g, p)
    nb.allocate_octree_geometry(o)
    nb.compute_octree_geometry(o, box)
    nb.allocate_octree_physical_data(o)
    nb.compute_octree_mass_and_center_of_mass(o, mk.indices, Mm, X, Y, Z)
    return dict(X=X, Y=Y, Z=Z, M=Mm, box=box, mk=mk, g=g, t=t, p=p, o=o)


def oracle_build(orc, x, y, z, m, bits):
    box = orc.bbox(x, y, z)
    keys, idx = orc.morton_keys(x, y, z, box, bits)
    u, f, c = orc.leaf_groups(keys)
    t = orc.tree_build(u)
    lv, cnt, off, M = orc.plan(len(u), t["parent"], t["prefix_length"], u, bits)
    o = orc.octree(M, len(u), t["parent"], t["range_first"], u, f, c, lv, cnt, off, bits)
    cx, cy, cz, h = orc.geometry(o["level"], o["prefix"], box, bits)
    r = orc.mass_com(o["children"], o["parent"], o["first"], o["count"], len(x), idx, m, x, y, z)
    o.update(center_x=cx, center_y=cy, center_z=cz, half=h, **r)
    return dict(box=box, keys=keys, idx=idx, u=u, f=f, c=c, t=t, lv=lv, cnt=cnt, off=off,
                M=M, o=o)


@pytest.mark.parametrize("bits,n,kind", [(10, 1, "normal"), (10, 2, "normal"),
                                         (10, 5000, "cubed"), (21, 5000, "cubed"),
                                         (21, 200_000, "normal"), (10, 200_000, "uniform")])
def test_octree_pipeline_stages_match_oracle(nb, orc, bits, n, kind):
    x, y, z = xyz(7 * n + bits, n, kind)
    m = np.random.default_rng(n).uniform(0.5, 2.0, n)
    if n > 10:
        x[1::17], y[1::17], z[1::17] = x[0], y[0], z[0]
    G = gpu_build(nb, x, y, z, m, bits)
    O = oracle_build(orc, x, y, z, m, bits)
    kd = np.uint32 if bits == 10 else np.uint64
    g, o = G["g"], G["o"]
    assert g.n_groups == len(O["u"])
    assert np.array_equal(host(G["t"].parent), O["t"]["parent"])
    assert G["p"].n_octree_nodes == O["M"]
    assert np.array_equal(host(G["p"].radix_level), O["lv"])
    assert np.array_equal(host(G["p"].edge_node_count), O["cnt"])
    assert np.array_equal(host(G["p"].edge_node_offset), O["off"])
    for f, of in (("children", "children"), ("parent", "parent"), ("level", "level"),
                  ("first_particle", "first"), ("particle_count", "count")):
        assert np.array_equal(host(getattr(o, f)), O["o"][of]), f
    assert np.array_equal(host(o.prefix, kd), O["o"]["prefix"])
    for f, of in (("center_x", "center_x"), ("center_y", "center_y"),
                  ("center_z", "center_z"), ("half_size", "half"), ("mass", "mass"),
                  ("com_x", "com_x"), ("com_y", "com_y"), ("com_z", "com_z")):
        assert np.array_equal(host(getattr(o, f)), O["o"][of]), f


@pytest.mark.parametrize("keys,levels,counts,offsets,total", [
    ([0], [1], [1], [0], 2),
    ([0, 1 << 29], [1, 1, 0], [1, 1, 0], [0, 1, 2], 3),
    ([0, 1 << 26], [2, 2, 1], [1, 1, 1], [0, 1, 2], 4),
    ([7, 7, 7, 7], [1], [1], [0], 2),
])
def test_plan_known_answers(nb, keys, levels, counts, offsets, total):
    k = np.array(keys, dtype=np.uint32)
    g = nb.MortonLeafGroups()
    nb.allocate_morton_leaf_groups(g, len(k), 4)
    nb.build_morton_leaf_groups(g, dev(k), len(k))
    t = nb.Tree()
    nb.allocate_tree_topology(t, g.n_groups)
    nb.build_tree_from_morton_groups(t, g)
    p = nb.RadixToOctreePlan()
    nb.allocate_radix_to_octree_plan(p, 2 * g.n_groups - 1)
    nb.build_radix_to_octree_plan(p, t, g, 10)
    assert list(host(p.radix_level)) == levels
    assert list(host(p.edge_node_count)) == counts
    assert list(host(p.edge_node_offset)) == offsets
    assert p.n_octree_nodes == total


def test_plan_rejects_keys_outside_layout(nb):
    k = np.array([0, 0x80000000], dtype=np.uint32)
    g = nb.MortonLeafGroups()
    nb.allocate_morton_leaf_groups(g, 2, 4)
    nb.build_morton_leaf_groups(g, dev(k), 2)
    t = nb.Tree()
    nb.allocate_tree(t, 2)
    nb.build_tree_from_morton_groups(t, g)
    p = nb.RadixToOctreePlan()
    nb.allocate_radix_to_octree_plan(p, 3)
    with pytest.raises(nb.DeviceError) as e:
        nb.build_radix_to_octree_plan(p, t, g, 10)
    assert e.value.device_code == 1


def test_geometry_known_answers(nb):
    X = dev([-10.0, 10.0])
    box = nb.Bbox(X, X, X)
    o = nb.Octree(10)
    nb.allocate_octree_topology(o, 4)
    o.level.copy_(dev([0, 1, 1, 3], np.int32))
    o.prefix.copy_(dev(np.array([0, 0, 7 << 27, (5 << 27) | (2 << 24) | (7 << 21)],
                                dtype=np.uint32)))
    o.root = 0
    nb.allocate_octree_geometry(o)
    nb.compute_octree_geometry(o, box)
    got = np.stack([host(o.center_x), host(o.center_y), host(o.center_z), host(o.half_size)])
    np.testing.assert_array_equal(got[:, 0], [0, 0, 0, 11.0])
    np.testing.assert_array_equal(got[:, 1], [-5.5, -5.5, -5.5, 5.5])
    np.testing.assert_array_equal(got[:, 2], [5.5, 5.5, 5.5, 5.5])
    np.testing.assert_array_equal(got[:, 3], [4.125, -1.375, 4.125, 1.375])


@pytest.mark.parametrize("level,prefix,code", [(11, 0, 1), (1, 1, 2)])
def test_geometry_rejects(nb, level, prefix, code):
    X = dev([-1.0, 1.0])
    box = nb.Bbox(X, X, X)
    o = nb.Octree(10)
    nb.allocate_octree_topology(o, 2)
    o.level.copy_(dev([0, level], np.int32))
    o.prefix.copy_(dev(np.array([0, prefix], dtype=np.uint32)))
    o.root = 0
    nb.allocate_octree_geometry(o)
    with pytest.raises(nb.DeviceError) as e:
        nb.compute_octree_geometry(o, box)
    assert e.value.device_code == code


# ----------------------------------------------------------- Barnes-Hut
def hand_octree(nb, t, n_particles):
    M = len(t["parent"])
    o = nb.Octree(10)
    o.n_particles = n_particles
    o.n_leaves = int((t["count"] > 0).sum())
    o.n_nodes = M
    o.root = 0
    o.children = dev(t["children"], np.int32)
    o.parent = dev(t["parent"], np.int32)
    o.first_particle = dev(t["first"], np.int32)
    o.particle_count = dev(t["count"], np.int32)
    for f, k in (("mass", "mass"), ("com_x", "com_x"), ("com_y", "com_y"), ("com_z", "com_z"),
                 ("center_x", "center_x"), ("center_y", "center_y"),
                 ("center_z", "center_z"), ("half_size", "half")):
        setattr(o, f, dev(t[k], np.float64))
    return o


def gpu_bh(nb, o, idx, m, x, y, z, theta, eps, G=1.0, order=0):
    n = len(idx)
    acc = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3)]
    nb.compute_barnes_hut_acceleration(o, dev(idx, np.uint32), dev(m), dev(x), dev(y), dev(z),
                                       *acc, n, nb.BarnesHutParameters(theta, eps, G), order)
    return [host(a) for a in acc]


def test_bh_single_particle(nb):
    t = H.single_particle_tree()
    a = gpu_bh(nb, hand_octree(nb, t, 1), [0], [2.0], [3.0], [4.0], [5.0], 0.5, 1e-3)
    assert [v[0] for v in a] == [0.0, 0.0, 0.0]


@pytest.mark.parametrize("theta,expected", [(1.0, 5.0 / 21.2**2),
                                            (0.85, 2.0 / 400 + 3.0 / 484)])
def test_bh_hand_built_cluster(nb, theta, expected):
    # barnes_hut_test.cu:568-1025 (far cluster accepted / off-centre opened)
    t = H.far_cluster_tree()
    ax, _, _ = gpu_bh(nb, hand_octree(nb, t, 3), [0, 1, 2], [1.0, 2.0, 3.0],
                      [-10.0, 10.0, 12.0], [0.0] * 3, [0.0] * 3, theta, 0.0)
    assert abs(ax[0] - expected) < 1e-12


@pytest.mark.parametrize("clustered", [False, True])
@pytest.mark.parametrize("bits", [10, 21])
@pytest.mark.parametrize("order", [0, 1])
def test_bh_matches_oracle_bitwise(nb, orc, clustered, bits, order):
    m, x, y, z = orc.accuracy_inputs(clustered)
    G = gpu_build(nb, x, y, z, m, bits)
    O = oracle_build(orc, x, y, z, m, bits)
    idx = O["idx"]
    if order == 1:
        xs, ys, zs, ms = x[idx], y[idx], z[idx], m[idx]
    else:
        xs, ys, zs, ms = x, y, z, m
    exact = H.direct_numpy(m, x, y, z, 0.02)
    for theta in (0.0, 0.3, 0.5, 0.7, 1.0, 1.3):
        a = gpu_bh(nb, G["o"], idx, ms, xs, ys, zs, theta, 0.02, 1.0, order)
        b = orc.bh_accel(O["o"], idx, ms, xs, ys, zs, theta, 0.02, 1.0, morton=order == 1)
        for k in range(3):
            assert np.array_equal(a[k], b[k]), (theta, k)
        bound = {0.0: 1e-11, 0.3: 2e-3, 0.5: 1e-2, 0.7: 2e-2, 1.0: 4e-2, 1.3: 1.0}[theta]
        assert H.rel_l2(a, exact) < bound


def test_bh_large_matches_oracle_bitwise(nb, orc):
    n = 40_000
    x, y, z = xyz(5, n, "cubed")
    m = np.random.default_rng(6).uniform(0.5, 2, n)
    G = gpu_build(nb, x, y, z, m, 21)
    O = oracle_build(orc, x, y, z, m, 21)
    a = gpu_bh(nb, G["o"], O["idx"], m, x, y, z, 0.5, 1e-3, 1.0)
    b = orc.bh_accel(O["o"], O["idx"], m, x, y, z, 0.5, 1e-3, 1.0)
    for k in range(3):
        assert np.array_equal(a[k], b[k])


def test_bh_singular_without_softening(nb):
    x = np.array([0.0, 0.0, 1.0])
    m = np.ones(3)
    G = gpu_build(nb, x, x * 0, x * 0, m, 10)
    with pytest.raises(nb.DeviceError) as e:
        gpu_bh(nb, G["o"], host(G["mk"].indices, np.uint32), m, x, x * 0, x * 0, 0.5, 0.0)
    assert e.value.device_code == 6


def test_bh_rejects_parameters(nb):
    t = H.single_particle_tree()
    o = hand_octree(nb, t, 1)
    for theta, eps, G in ((-1.0, 0.0, 1.0), (np.nan, 0.0, 1.0), (0.5, -1.0, 1.0),
                          (0.5, 0.0, 0.0)):
        with pytest.raises(nb.InvalidArgument):
            gpu_bh(nb, o, [0], [2.0], [3.0], [4.0], [5.0], theta, eps, G)


# ----------------------------------------------------------- integrator
def test_kick_drift_known_answer(nb):
    v = [dev([a]) for a in (10.0, 20.0, 30.0, 2.0, 4.0, 6.0, 4.0, -2.0, 8.0)]
    nb.leapfrog_kick_drift(*v, 1, 0.5)
    assert [host(t)[0] for t in v[:6]] == [11.5, 21.75, 34.0, 3.0, 3.5, 8.0]


def test_leapfrog_matches_oracle(nb, orc):
    rng = np.random.default_rng(12)
    s = rng.normal(size=(9, 100_003))
    v = [dev(a) for a in s]
    nb.leapfrog_kick_drift(*v, s.shape[1], 1e-3)
    nb.leapfrog_kick(*v[3:9], s.shape[1], 1e-3)
    r = orc.kick_drift(*s, 1e-3)
    r[3:6] = orc.kick(*r[3:6], *s[6:9], 1e-3)
    for k in range(6):
        assert np.array_equal(host(v[k]), r[k])


def test_leapfrog_rejects(nb):
    v = [dev([1.0]) for _ in range(9)]
    for dt in (0.0, -1.0, float("nan")):
        with pytest.raises(nb.InvalidArgument):
            nb.leapfrog_kick(*v[3:], 1, dt)
    nb.leapfrog_kick_drift(*v, 0, 1.0)  # N = 0 is a no-op


def test_gather_morton(nb):
    rng = np.random.default_rng(2)
    n = 5000
    idx = rng.permutation(n).astype(np.uint32)
    a = rng.normal(size=(4, n))
    out = nb.gather_morton_ordered(dev(idx), *[dev(v) for v in a], n)
    for k in range(4):
        assert np.array_equal(host(out[k]), a[k][idx])
