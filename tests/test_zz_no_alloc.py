"""The allocation-free leaf-group build (nbh_leaf_groups_build_noalloc) against
the allocating nbh_leaf_groups_build, through the C ABI, bit for bit: two
calls in a row with the second one smaller (the static scratch keeps the
first call's words), both key widths, the device-only count form without a
host synchronisation, and the size limit."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def sorted_keys(rng, n, key_bytes, distinct):
    hi = (1 << 30) if key_bytes == 4 else (1 << 63)
    pool = np.unique(rng.integers(0, hi, size=distinct, dtype=np.uint64))
    keys = np.sort(rng.choice(pool, size=n))
    if key_bytes == 4:
        return torch.from_numpy(keys.astype(np.uint32).view(np.int32)).cuda()
    return torch.from_numpy(keys.view(np.int64)).cuda()


def groups_of(nb, capacity, key_bytes):
    g = nb.MortonLeafGroups()
    nb.allocate_morton_leaf_groups(g, capacity, key_bytes=key_bytes)
    return g


def same(nb, a, b):
    assert a.n_groups == b.n_groups and a.n_groups > 0
    k = a.n_groups
    assert torch.equal(a.unique_keys[:k], b.unique_keys[:k])
    assert torch.equal(a.first_particle[:k], b.first_particle[:k])
    assert torch.equal(a.particle_count[:k], b.particle_count[:k])


@pytest.mark.parametrize("key_bytes", [4, 8])
def test_noalloc_leaf_groups_equal_the_allocating_build(nb, key_bytes):
    rng = np.random.default_rng(2024 + key_bytes)
    cap = 600_000
    ref = groups_of(nb, cap, key_bytes)
    out = groups_of(nb, cap, key_bytes)
    # the second call is smaller: the status words of the first call's extra
    # tiles are left in the static array and must not matter
    for n, distinct in ((600_000, 200_000), (90_001, 4_000), (1, 1), (2049, 2049)):
        print(f"---- noalloc leaf groups key_bytes={key_bytes} n={n}", flush=True)
        keys = sorted_keys(rng, n, key_bytes, distinct)
        nb.build_morton_leaf_groups(ref, keys, n)
        d_ng = torch.full((1,), -7, dtype=torch.int32, device="cuda")
        nb.build_morton_leaf_groups_noalloc(out, keys, n, d_n_groups=d_ng)
        same(nb, out, ref)
        assert int(d_ng.item()) == ref.n_groups
        # device count only, no host synchronisation inside the call
        out2 = groups_of(nb, cap, key_bytes)
        d_ng2 = torch.zeros(1, dtype=torch.int32, device="cuda")
        nb.build_morton_leaf_groups_noalloc(out2, keys, n, d_n_groups=d_ng2, read_count=False)
        torch.cuda.synchronize()
        out2.n_groups = int(d_ng2.item())
        same(nb, out2, ref)


def test_noalloc_leaf_groups_rejects_sizes_beyond_its_scratch(nb):
    g = groups_of(nb, 16, 4)
    keys = torch.zeros(16, dtype=torch.int32, device="cuda")
    n = nb.LEAF_GROUPS_NOALLOC_MAX + 1
    g.capacity = n  # the limit is checked before any pointer is used
    with pytest.raises(nb.InvalidArgument, match="2\\^28"):
        nb.build_morton_leaf_groups_noalloc(g, keys, n)
    with pytest.raises(nb.InvalidArgument, match="capacity"):
        g.capacity = 8
        nb.build_morton_leaf_groups_noalloc(g, keys, 16)
