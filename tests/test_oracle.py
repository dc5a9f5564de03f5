"""The oracle pinned against the reference's own known answers (CPU only).

Cases transcribe tests/unit/cuda/*.cu and tests/integration/cuda/*.cu of the
reference; the independent checks (brute-force octrees, subtree sums, direct
sums, closed forms) are in helpers.py.
"""

import numpy as np
import pytest

import helpers as H


# ---------------------------------------------------------------- bbox
@pytest.mark.parametrize("xyz,expected", [
    (([-2.0, 4.0, 1.0], [10.0, 14.0, -2.0], [-1.0, 0.0, 5.0]), (1.0, 6.0, 2.0, 17.6)),
    (([-2.0, 4.0], [10.0, 14.0], [-1.0, 5.0]), (1.0, 12.0, 2.0, 6.6)),
    (([3.5, 3.5], [-4.0, -4.0], [8.25, 8.25]), (3.5, -4.0, 8.25, 1.0)),
    (([10.0, 14.0], [20.0, 22.0], [30.0, 31.0]), (12.0, 21.0, 30.5, 4.4)),
])
def test_bbox_known_answers(orc, xyz, expected):
    # globalbounding_test.cu:60-119
    box = orc.bbox(*xyz)
    np.testing.assert_allclose(box, expected, rtol=4 * np.finfo(float).eps)


def test_bbox_matches_minmax(orc):
    rng = np.random.default_rng(1)
    x, y, z = rng.normal(size=(3, 1000))
    box = orc.bbox(x, y, z)
    lo = np.array([x.min(), y.min(), z.min()])
    hi = np.array([x.max(), y.max(), z.max()])
    assert np.array_equal(box[:3], (lo + hi) * 0.5)
    assert box[3] == 1.1 * (hi - lo).max()


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_bbox_rejects_non_finite(orc, bad):
    with pytest.raises(orc.OracleError) as e:
        orc.bbox([0.0, bad], [0.0, 1.0], [0.0, 1.0])
    assert e.value.status == -4  # domain_error


def test_bbox_rejects_empty(orc):
    with pytest.raises(orc.OracleError) as e:
        orc.bbox([], [], [])
    assert e.value.status == -1


# -------------------------------------------------------------- Morton
def test_spread_known_values():
    # bit k -> bit 3k: 0b1011 -> 0b1000001000001
    assert H.spread_bits(0b1011, 10) == (1 << 9) | (1 << 3) | 1
    assert H.spread_bits(0x3FF, 10) == 0x09249249  # the reference's last mask
    assert H.spread_bits((1 << 21) - 1, 21) == 0x1249249249249249


@pytest.mark.parametrize("bits", [10, 21])
def test_morton_keys_follow_reference_host_oracle(orc, bits):
    # morton_test.cu:114-206 restated independently: quantise, interleave
    # by bit strings, stable sort by key.
    x = np.array([0.8, -1.0, 0.1, 1.0, -0.4])
    y = np.array([-0.2, 0.9, 0.0, -1.0, 0.3])
    z = np.array([1.0, -0.7, 0.2, 0.4, -1.0])
    box = orc.bbox(x, y, z)
    keys, idx = orc.morton_keys(x, y, z, box, bits)
    grid = float(1 << bits)
    q = lambda a, c: int(min(max((a - (c - box[3] * 0.5)) / box[3] * grid, 0.0), grid - 1))
    expect = [H.morton_key(q(x[i], box[0]), q(y[i], box[1]), q(z[i], box[2]), bits)
              for i in range(5)]
    order = sorted(range(5), key=lambda i: expect[i])  # Python sort is stable
    assert list(idx) == order
    assert [int(k) for k in keys] == [expect[i] for i in order]


@pytest.mark.parametrize("bits", [10, 21])
def test_morton_stable_ties_and_coincident(orc, bits):
    # morton_test.cu:252-272 (perm {2,0,1}) and :338-366 (perm {0,1,2,3})
    x = y = z = np.array([1.0, 1.0, 0.0])
    keys, idx = orc.morton_keys(x, y, z, orc.bbox(x, y, z), bits)
    assert list(idx) == [2, 0, 1] and keys[1] == keys[2]
    x = np.full(4, 3.5)
    y = np.full(4, -4.0)
    z = np.full(4, 8.25)
    keys, idx = orc.morton_keys(x, y, z, orc.bbox(x, y, z), bits)
    assert list(idx) == [0, 1, 2, 3] and len(set(keys.tolist())) == 1


def test_morton_padded_corners(orc):
    # morton_test.cu:274-303: corners (0,0,0) and (1,1,1); the padded box
    # puts them at quantised 46 and 977 on every axis.
    x = y = z = np.array([0.0, 1.0])
    keys, idx = orc.morton_keys(x, y, z, orc.bbox(x, y, z), 10)
    assert list(idx) == [0, 1]
    assert keys[0] == 7 * H.spread_bits(46, 10) and keys[1] == 7 * H.spread_bits(977, 10)


def test_morton_matches_numpy_stable_sort(orc):
    rng = np.random.default_rng(7)
    x, y, z = rng.normal(size=(3, 5000))
    x[::7] = x[0]
    y[::7] = y[0]
    z[::7] = z[0]
    box = orc.bbox(x, y, z)
    for bits in (10, 21):
        keys, idx = orc.morton_keys(x, y, z, box, bits)
        raw = np.empty(5000, dtype=keys.dtype)
        raw[idx] = keys
        assert np.array_equal(np.argsort(raw, kind="stable"), idx)


def test_morton_rejects_empty(orc):
    with pytest.raises(orc.OracleError) as e:
        orc.morton_keys([], [], [], [0, 0, 0, 1], 10)
    assert e.value.status == -1


# --------------------------------------------------------- leaf groups
@pytest.mark.parametrize("keys,uniq,first,count", [
    ([42], [42], [0], [1]),
    ([2, 5, 8, 10], [2, 5, 8, 10], [0, 1, 2, 3], [1, 1, 1, 1]),
    ([7, 7, 7, 7], [7], [0], [4]),
    ([1, 1, 2, 4, 4, 4, 9], [1, 2, 4, 9], [0, 2, 3, 6], [2, 1, 3, 1]),
    ([0, 0, 1, 0xFFFFFFFF, 0xFFFFFFFF], [0, 1, 0xFFFFFFFF], [0, 2, 3], [2, 1, 2]),
])
def test_leaf_groups_known_answers(orc, keys, uniq, first, count):
    # morton_leaf_groups_test.cu:277-316
    u, f, c = orc.leaf_groups(np.array(keys, dtype=np.uint32))
    assert list(u) == uniq and list(f) == first and list(c) == count


def test_leaf_groups_capacity(orc):
    with pytest.raises(orc.OracleError) as e:
        orc.leaf_groups(np.array([1, 2, 3], dtype=np.uint32), capacity=2)
    assert e.value.status == -1


# -------------------------------------------------------- Karras tree
def test_karras_two_leaves(orc):
    # tree_builder_test.cu:529-538
    t = orc.tree_build(np.array([0, 0x80000000], dtype=np.uint32))
    assert list(t["left"]) == [0] and list(t["right"]) == [1]
    assert list(t["parent"]) == [2, 2, -1]
    assert list(t["range_first"]) == [0] and list(t["range_last"]) == [1]
    assert list(t["prefix_length"]) == [0]


def test_karras_balanced_four(orc):
    # tree_builder_test.cu:540-556
    t = orc.tree_build(np.array([0, 0x40000000, 0x80000000, 0xC0000000], dtype=np.uint32))
    assert list(t["left"]) == [5, 0, 2] and list(t["right"]) == [6, 1, 3]
    assert list(t["parent"]) == [5, 5, 6, 6, -1, 4, 4]
    assert list(t["range_first"]) == [0, 0, 2] and list(t["range_last"]) == [3, 1, 3]
    assert list(t["prefix_length"]) == [0, 1, 1]


@pytest.mark.parametrize("n", [1, 3, 5, 7, 16, 31, 64, 257])
def test_karras_random_validator(orc, n):
    # tree_builder_test.cu:602-642: mt19937(0xC0FFEE), sorted, duplicates
    # planted every 11 keys from index 5.
    keys = orc.random_keys(n)
    t = orc.tree_build(keys)
    H.validate_karras(t, keys, 32)


def test_karras_64bit_validator(orc):
    rng = np.random.default_rng(3)
    keys = np.sort(rng.integers(0, 2**63, size=300, dtype=np.uint64))
    keys[10] = keys[9]
    H.validate_karras(orc.tree_build(keys), keys, 64)


def test_legacy_center_of_mass(orc):
    # tree_builder_test.cu:558-570 (permutation and COM) checked by sums
    keys = np.array([0, 0x40000000, 0x80000000, 0xC0000000], dtype=np.uint32)
    idx = np.array([2, 0, 3, 1], dtype=np.uint32)
    m = np.array([1.0, 2.0, 3.0, 4.0])
    x = np.array([10.0, 20.0, 30.0, 40.0])
    t = orc.tree_build(keys)
    c = orc.tree_com(t, 4, idx, m, x, -x / 10, x / 2 - 5)
    root = 4
    assert c["mass"][root] == 10.0
    np.testing.assert_allclose(c["com_x"][root], (m * x).sum() / 10.0, rtol=1e-15)
    assert list(c["visit_count"]) == [2, 2, 2]


# ---------------------------------------------------------------- plan
@pytest.mark.parametrize("keys,levels,counts,offsets,total", [
    ([0], [1], [1], [0], 2),
    ([0, 1 << 29], [1, 1, 0], [1, 1, 0], [0, 1, 2], 3),
    ([0, 1 << 26], [2, 2, 1], [1, 1, 1], [0, 1, 2], 4),
    ([7, 7, 7, 7], [1], [1], [0], 2),
])
def test_plan_known_answers(orc, keys, levels, counts, offsets, total):
    # radix_to_octree_test.cu:310-348
    u, f, c = orc.leaf_groups(np.array(keys, dtype=np.uint32))
    t = orc.tree_build(u)
    lv, cnt, off, m = orc.plan(len(u), t["parent"], t["prefix_length"], u, 10)
    assert list(lv) == levels and list(cnt) == counts and list(off) == offsets
    assert m == total


def test_plan_rejects_keys_outside_layout(orc):
    # radix_to_octree_test.cu:247-302
    u = np.array([0, 0x80000000], dtype=np.uint32)
    t = orc.tree_build(u)
    with pytest.raises(orc.OracleError) as e:
        orc.plan(2, t["parent"], t["prefix_length"], u, 10)
    assert e.value.status == -3 and e.value.code == 1


# -------------------------------------------------------------- octree
def build_octree(orc, keys, bits):
    u, f, c = orc.leaf_groups(keys)
    t = orc.tree_build(u)
    lv, cnt, off, m = orc.plan(len(u), t["parent"], t["prefix_length"], u, bits)
    o = orc.octree(m, len(u), t["parent"], t["range_first"], u, f, c, lv, cnt, off, bits)
    return u, f, c, o


@pytest.mark.parametrize("keys", [
    [0], [7, 7, 7, 7], [0, 1 << 29], [0, 1 << 26],
    [0, 1 << 23, 1 << 26, 1 << 29, (1 << 29) | (1 << 26)],
])
def test_octree_small_cases_match_brute_force(orc, keys):
    # sparse_octree_test.cu:557-598
    u, f, c, o = build_octree(orc, np.array(keys, dtype=np.uint32), 10)
    H.octree_matches_brute(o, u, f, c, 10)


def test_octree_random_257_matches_brute_force(orc):
    # sparse_octree_test.cu:523-548: 30-bit keys, duplicates every 17 from 7
    keys = orc.random_keys(257, mask=0x3FFFFFFF, start=7, stride=17)
    u, f, c, o = build_octree(orc, keys, 10)
    H.octree_matches_brute(o, u, f, c, 10)


def test_octree_21bit_matches_brute_force(orc):
    rng = np.random.default_rng(11)
    x, y, z = rng.normal(size=(3, 3000)) ** 3
    box = orc.bbox(x, y, z)
    keys, idx = orc.morton_keys(x, y, z, box, 21)
    u, f, c, o = build_octree(orc, keys, 21)
    H.octree_matches_brute(o, u, f, c, 21)


# ------------------------------------------------------------ geometry
def test_geometry_known_answers(orc):
    # octree_geometry_test.cu:133-309, box of (-10,10)^3 -> side 22
    box = orc.bbox([-10.0, 10.0], [-10.0, 10.0], [-10.0, 10.0])
    prefix = np.array([0, 0 << 27, 7 << 27, (5 << 27) | (2 << 24) | (7 << 21)], dtype=np.uint32)
    cx, cy, cz, h = orc.geometry(np.array([0, 1, 1, 3]), prefix, box, 10)
    assert (cx[0], cy[0], cz[0], h[0]) == (0.0, 0.0, 0.0, 11.0)
    assert (cx[1], cy[1], cz[1], h[1]) == (-5.5, -5.5, -5.5, 5.5)
    assert (cx[2], cy[2], cz[2], h[2]) == (5.5, 5.5, 5.5, 5.5)
    assert (cx[3], cy[3], cz[3], h[3]) == (4.125, -1.375, 4.125, 1.375)


@pytest.mark.parametrize("level,prefix,code", [(11, 0, 1), (1, 1, 2)])
def test_geometry_rejects_invalid(orc, level, prefix, code):
    # octree_geometry_test.cu:313-375
    box = orc.bbox([-1.0, 1.0], [-1.0, 1.0], [-1.0, 1.0])
    with pytest.raises(orc.OracleError) as e:
        orc.geometry(np.array([0, level]), np.array([0, prefix], dtype=np.uint32), box, 10)
    assert e.value.code == code


# ---------------------------------------------------- mass and COM
@pytest.mark.parametrize("bits", [10, 21])
def test_mass_com_matches_subtree_sums(orc, bits):
    rng = np.random.default_rng(5)
    n = 2000
    x, y, z = rng.normal(size=(3, n))
    m = rng.uniform(0.5, 2.0, size=n)
    box = orc.bbox(x, y, z)
    keys, idx = orc.morton_keys(x, y, z, box, bits)
    u, f, c, o = build_octree(orc, keys, bits)
    r = orc.mass_com(o["children"], o["parent"], o["first"], o["count"], n, idx, m, x, y, z)
    mass, cx, cy, cz = H.subtree_sums(o, idx, m, x, y, z)
    np.testing.assert_allclose(r["mass"], mass, rtol=1e-13)
    np.testing.assert_allclose(r["com_x"], cx, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(r["com_z"], cz, rtol=1e-12, atol=1e-14)


# ---------------------------------------------------------- Barnes-Hut
def test_bh_single_particle_is_zero(orc):
    t = H.single_particle_tree()
    ax, ay, az = orc.bh_accel(t, [0], [2.0], [3.0], [4.0], [5.0], 0.5, 1e-3, 1.0)
    assert (ax[0], ay[0], az[0]) == (0.0, 0.0, 0.0)


def test_bh_far_cluster_accepted(orc):
    # barnes_hut_test.cu:568-782: theta 1, node 2 taken as one mass
    t = H.far_cluster_tree()
    ax, _, _ = orc.bh_accel(t, [0, 1, 2], [1.0, 2.0, 3.0], [-10.0, 10.0, 12.0],
                            [0.0] * 3, [0.0] * 3, 1.0, 0.0, 1.0)
    assert abs(ax[0] - 5.0 / 21.2**2) < 1e-12
    assert abs(ax[0] - (2.0 / 400 + 3.0 / 484)) > 1e-6


def test_bh_off_centre_cluster_opened(orc):
    # barnes_hut_test.cu:784-1025: theta 0.85 opens node 2 (21.2 < 16/0.85+3.2)
    t = H.far_cluster_tree()
    ax, _, _ = orc.bh_accel(t, [0, 1, 2], [1.0, 2.0, 3.0], [-10.0, 10.0, 12.0],
                            [0.0] * 3, [0.0] * 3, 0.85, 0.0, 1.0)
    assert abs(ax[0] - (2.0 / 400 + 3.0 / 484)) < 1e-12


def oracle_tree(orc, m, x, y, z, bits):
    box = orc.bbox(x, y, z)
    keys, idx = orc.morton_keys(x, y, z, box, bits)
    u, f, c, o = build_octree(orc, keys, bits)
    cxg, cyg, czg, h = orc.geometry(o["level"], o["prefix"], box, bits)
    r = orc.mass_com(o["children"], o["parent"], o["first"], o["count"], len(m), idx, m, x, y, z)
    t = dict(children=o["children"], parent=o["parent"], first=o["first"], count=o["count"],
             mass=r["mass"], com_x=r["com_x"], com_y=r["com_y"], com_z=r["com_z"],
             center_x=cxg, center_y=cyg, center_z=czg, half=h)
    return t, idx


@pytest.mark.parametrize("clustered", [False, True])
@pytest.mark.parametrize("bits", [10, 21])
def test_bh_accuracy_bounds(orc, clustered, bits):
    # barnes_hut_accuracy_test.cu:294-397,538-697: 256 particles, G 1,
    # softening 0.02, relative L2 bounds per theta.
    m, x, y, z = orc.accuracy_inputs(clustered)
    exact = H.direct_numpy(m, x, y, z, 0.02)
    t, idx = oracle_tree(orc, m, x, y, z, bits)
    for theta, bound in ((0.0, 1e-11), (0.3, 2e-3), (0.5, 1e-2), (0.7, 2e-2), (1.0, 4e-2)):
        a = orc.bh_accel(t, idx, m, x, y, z, theta, 0.02, 1.0)
        assert H.rel_l2(a, exact) < bound, theta


def test_direct_oracle_matches_numpy(orc):
    m, x, y, z = orc.accuracy_inputs(True, 300)
    a = orc.direct(m, x, y, z, 0.01, 1.0)
    b = H.direct_numpy(m, x, y, z, 0.01)
    assert H.rel_l2(a, b) < 1e-13


# ----------------------------------------------------------- integrator
def test_kick_drift_known_answer(orc):
    # integrator_test.cu:34-119
    x, y, z, vx, vy, vz = orc.kick_drift([10.0], [20.0], [30.0], [2.0], [4.0], [6.0],
                                         [4.0], [-2.0], [8.0], 0.5)
    assert (vx[0], vy[0], vz[0]) == (3.0, 3.5, 8.0)
    assert (x[0], y[0], z[0]) == (11.5, 21.75, 34.0)


def test_constant_acceleration_closed_form(orc):
    # integrator_test.cu:181-305
    dt = 0.25
    x, y, z, vx, vy, vz = orc.kick_drift([1.0], [-2.0], [3.0], [4.0], [5.0], [-6.0],
                                         [2.0], [-4.0], [1.0], dt)
    vx, vy, vz = orc.kick(vx, vy, vz, [2.0], [-4.0], [1.0], dt)
    np.testing.assert_allclose([x[0], y[0], z[0]],
                               [1 + 4 * dt + dt * dt, -2 + 5 * dt - 2 * dt * dt,
                                3 - 6 * dt + 0.5 * dt * dt], atol=1e-14)
    np.testing.assert_allclose([vx[0], vy[0], vz[0]],
                               [4 + 2 * dt, 5 - 4 * dt, -6 + dt], atol=1e-14)


def test_kick_rejects_bad_dt(orc):
    for dt in (0.0, -1.0, np.nan, np.inf):
        with pytest.raises(orc.OracleError):
            orc.kick([0.0], [0.0], [0.0], [0.0], [0.0], [0.0], dt)
