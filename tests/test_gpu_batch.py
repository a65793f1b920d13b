"""Batched commit engine vs the plain C-ABI sequence and the oracle.
Bit-exact (tolerance zero)."""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _coset(zk):
    zk.init_coset_rs(0, 24, 7)


def plain_commit(zk, coeffs, cols, lg_n, rate, h, ht):
    """compute_batched_lde -> compute_transpose_rev -> fill_digests (device)."""
    lg = lg_n + rate
    N = 1 << lg
    d_in = zk.DeviceBuffer.from_host(coeffs)
    d_lde = zk.DeviceBuffer(cols * N)
    d_leaves = zk.DeviceBuffer(cols * N)
    zk.lde_batch(0, d_lde, d_in, lg_n, zk.NTTConfig(batches=cols, extension_rate_bits=rate, with_coset=True,
                                                   are_inputs_on_device=True, are_outputs_on_device=True))
    zk.transpose_rev_batch(0, d_leaves, d_lde, lg, zk.TransposeConfig(batches=cols, are_inputs_on_device=True,
                                                                      are_outputs_on_device=True))
    nc = 1 << h
    nd = 0 if nc == N else 2 * (N - nc)
    d_dig = zk.DeviceBuffer(max(nd, 1) * 4)
    d_cap = zk.DeviceBuffer(nc * 4)
    zk.fill_digests_buf_linear_gpu_with_gpu_ptr(d_dig, d_cap, d_leaves, nd, nc, N, cols, h, ht, 0)
    return d_leaves.to_host(), (d_dig.to_host(nd * 4) if nd else np.zeros(0, np.uint64)), d_cap.to_host()


def oracle_commit(coeffs, cols, lg_n, rate, h, ht):
    lde = ob.lde(coeffs, lg_n, rate, cols)
    leaves = ob.transpose_rev(lde, lg_n + rate, cols)
    dig, cap = ob.merkle(leaves, cols, h, ht)
    return leaves, dig, cap


# (cols, lg_n, rate, cap_height, hash)
MIX = [(2, 10, 1, 4, 0), (20, 11, 2, 4, 3), (80, 10, 3, 4, 0), (135, 9, 1, 4, 3), (5, 3, 1, 0, 0),
       (7, 2, 2, 4, 0), (3, 1, 1, 2, 3), (4, 5, 2, 7, 0), (9, 12, 1, 2, 3), (1, 6, 1, 1, 0), (135, 13, 1, 4, 0),
       (17, 9, 2, 4, 1), (34, 8, 1, 3, 4), (3, 7, 1, 2, 1), (12, 6, 3, 4, 4), (9, 6, 1, 3, 2), (5, 4, 2, 1, 2)]


# the MIX plus strided passes of 5..7 stages (one tile per work-group) and
# a 2^19-leaf tree
PIECES_MIX = MIX + [(8, 15, 2, 4, 0), (6, 15, 3, 4, 3), (4, 16, 3, 4, 0)]


# Merkle schedules of a batch (zkx_set_level_launches): the trees finished
# inside the leaf launch, or every level above the leaves one launch over all
# trees; the default picks by the batch's total leaves
LEVELS = {"fused": None, "by_level": 0}


@pytest.fixture
def merkle_schedule(zk, request):
    zk.set_level_launches(LEVELS[request.param])
    yield request.param
    zk.set_level_launches(zk.LEVEL_LAUNCHES_DEFAULT)


@pytest.mark.parametrize("merkle_schedule,ntt_tiles,merkle_groups",
                         [("fused", 1, 1), ("fused", 5, 3), ("fused", 64, 7), ("fused", 0, 2), ("fused", 3, 0),
                          ("by_level", 1, 1), ("by_level", 0, 3)], indirect=["merkle_schedule"])
def test_batch_in_pieces_vs_oracle(zk, merkle_schedule, ntt_tiles, merkle_groups):
    """zkx_commit_batch_pieces: every NTT pass launch split into ranges of
    ntt_tiles tiles and every Merkle launch into ranges of merkle_groups
    work-groups gives the oracle's results (and so the unsplit batch's),
    under both Merkle schedules."""
    commits, inputs = [], []
    for i, (cols, lg_n, rate, h, ht) in enumerate(PIECES_MIX):
        x = ob.canonical(cols << lg_n, 900 + i)
        inputs.append(x)
        commits.append(zk.Commit(x, cols, lg_n, rate, h, ht))
    print(f"pieces {merkle_schedule} ntt_tiles={ntt_tiles} merkle_groups={merkle_groups}", flush=True)
    zk.commit_batch(commits, ntt_tiles=ntt_tiles, merkle_groups=merkle_groups)
    for (cols, lg_n, rate, h, ht), x, c in zip(PIECES_MIX, inputs, commits):
        leaves, dig, cap = c.results()
        el, ed, ec = oracle_commit(x, cols, lg_n, rate, h, ht)
        assert np.array_equal(leaves, el), f"leaves cols={cols} lg_n={lg_n} rate={rate}"
        assert np.array_equal(dig, ed), f"digests cols={cols} lg_n={lg_n} rate={rate} ht={ht}"
        assert np.array_equal(cap, ec)


@pytest.mark.parametrize("merkle_schedule", ["fused", "by_level"], indirect=True)
def test_batch_mixed_vs_oracle(zk, merkle_schedule):
    """The mix (all five hashes, caps == leaves, one-leaf subtrees, leaves
    of at most 4 elements) under both Merkle schedules."""
    commits, inputs = [], []
    for i, (cols, lg_n, rate, h, ht) in enumerate(MIX):
        x = ob.canonical(cols << lg_n, 500 + i)
        inputs.append(x)
        commits.append(zk.Commit(x, cols, lg_n, rate, h, ht))
    zk.commit_batch(commits)
    for (cols, lg_n, rate, h, ht), x, c in zip(MIX, inputs, commits):
        print(f"{merkle_schedule} job cols={cols} lg_n={lg_n} rate={rate} h={h} ht={ht}", flush=True)
        leaves, dig, cap = c.results()
        el, ed, ec = oracle_commit(x, cols, lg_n, rate, h, ht)
        assert np.array_equal(leaves, el)
        assert np.array_equal(dig, ed)
        assert np.array_equal(cap, ec)


@pytest.mark.parametrize("cols,lg_n,rate,h,ht", [(135, 15, 3, 4, 0), (80, 16, 2, 4, 3), (20, 14, 1, 4, 0)])
def test_batch_equals_plain_path(zk, cols, lg_n, rate, h, ht):
    """Same job alone through the plain C-ABI sequence and inside a batch."""
    x = ob.canonical(cols << lg_n, 77 + lg_n)
    pl, pd, pc = plain_commit(zk, x, cols, lg_n, rate, h, ht)
    other = zk.Commit(ob.canonical(20 << 11, 3), 20, 11, 2, 4, 3 - ht)
    c = zk.Commit(x, cols, lg_n, rate, h, ht)
    zk.commit_batch([other, c])
    bl, bd, bc = c.results()
    assert np.array_equal(bl, pl)
    assert np.array_equal(bd, pd)
    assert np.array_equal(bc, pc)


def test_batch_same_result_alone_and_in_mix(zk):
    x = ob.canonical(20 << 12, 9)
    alone = zk.Commit(x, 20, 12, 2, 4, 0)
    zk.commit_batch([alone])
    mix = [zk.Commit(ob.canonical(c << l, 40 + k), c, l, r, 4, ht)
           for k, (c, l, r, ht) in enumerate([(2, 10, 3, 3), (135, 12, 1, 0), (80, 11, 2, 3)])]
    inmix = zk.Commit(x, 20, 12, 2, 4, 0)
    zk.commit_batch(mix[:2] + [inmix] + mix[2:])
    for a, b in zip(alone.results(), inmix.results()):
        assert np.array_equal(a, b)


def test_batch_without_leaf_rows(zk):
    x = ob.canonical(8 << 10, 1)
    c = zk.Commit(x, 8, 10, 1, 3, 0, leaves=False)
    zk.commit_batch([c])
    _, dig, cap = c.results()
    _, ed, ec = oracle_commit(x, 8, 10, 1, 3, 0)
    assert np.array_equal(dig, ed) and np.array_equal(cap, ec)


def test_batch_rejects_bad_jobs(zk):
    x = ob.canonical(8 << 4, 1)
    bad = zk.Commit(x, 8, 4, 1, 3, 0)
    bad.hash_type = 5  # no such hash
    with pytest.raises(zk.ZeknoxError):
        zk.commit_batch([bad])
    bad.hash_type = 0
    bad.cap_height = 9
    with pytest.raises(zk.ZeknoxError):
        zk.commit_batch([bad])


def test_batch_after_trim(zk):
    """Freeing the engine's working memory between batches changes nothing."""
    coeffs = ob.canonical(7 << 9, 71)
    c = zk.Commit(coeffs, 7, 9, 2, 3, 3)
    zk.commit_batch([c])
    first = c.results()
    assert zk.lib.zkx_commit_trim(0).code == 0
    c2 = zk.Commit(coeffs, 7, 9, 2, 3, 3)
    zk.commit_batch([c2])
    for a, b in zip(first, c2.results()):
        assert np.array_equal(a, b)
    assert np.array_equal(first[2], oracle_commit(coeffs, 7, 9, 2, 3, 3)[2])


def test_fill_random_is_canonical_and_seeded(zk):
    n = 1 << 16
    a, b, c = zk.DeviceBuffer(n), zk.DeviceBuffer(n), zk.DeviceBuffer(n)
    a.fill_random(5)
    b.fill_random(5)
    c.fill_random(6)
    x, y, z = a.to_host(), b.to_host(), c.to_host()
    assert np.array_equal(x, y)
    assert not np.array_equal(x, z)
    assert int(x.max()) < ob.P and int(z.max()) < ob.P
    assert len(np.unique(x)) == n
    # uniform on [0, p): the mean of 2^16 draws is p/2 within 1%
    assert abs(float(x.astype(np.float64).mean()) / ob.P - 0.5) < 0.01
