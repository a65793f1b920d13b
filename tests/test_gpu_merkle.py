"""GPU Merkle forest through the C ABI vs the oracle and the reference KATs.
Bit-exact (tolerance zero)."""
import numpy as np
import pytest

import oracle_bind as ob
from test_oracle import FIB12, GO_CAPS, GO_LEAF7, LEAF9, POSEIDON2_KAT, POSEIDON_KAT

pytestmark = pytest.mark.gpu


def gpu_tree(zk, leaves, ls, h, ht):
    return zk.merkle_tree(leaves, ls, h, ht)


def test_poseidon_kat_direct(zk):
    # caps == leaves: cap[i] = hash_one(leaf i); one leaf per k = 1..9
    for k in range(1, 10):
        _, caps = gpu_tree(zk, np.array(LEAF9[:k], np.uint64), k, 0, 0)
        assert [int(v) for v in caps] == POSEIDON_KAT[k]


def test_poseidon2_kat_direct(zk):
    _, caps = gpu_tree(zk, np.array(FIB12, np.uint64), 12, 0, 3)
    assert [int(v) for v in caps] == POSEIDON2_KAT


def test_keccak_kat_direct(zk):
    """tests/tests.cu:162-199: Keccak hash_one of the first `size` words."""
    data = [13421290117754017454, 7401888676587830362, 15316685236050041751, 13588825262671526271,
            13421290117754017454, 7401888676587830362]
    expected = {
        1: [13421290117754017454, 0, 0, 0],
        2: [13421290117754017454, 7401888676587830362, 0, 0],
        3: [13421290117754017454, 7401888676587830362, 15316685236050041751, 0],
        4: [9981707860959651334, 16351366398560378420, 4283762868800363615, 101],
        5: [708367124667950404, 17208681281141108820, 8334320481120086961, 134],
        6: [16109761546392287110, 4918745475135463511, 17110319063854316944, 103],
    }
    for size, want in expected.items():
        _, caps = gpu_tree(zk, np.array(data[:size], np.uint64), size, 0, 1)
        assert [int(v) for v in caps] == want


def test_poseidon_bn128_kat_direct(zk):
    """tests/tests.cu:492-536: Poseidon-BN128 hash_one of 12 words."""
    inp = np.array([8917524657281059100, 13029010200779371910, 16138660518493481604, 17277322750214136960,
                    1441151880423231822, 0, 0, 0, 0, 0, 0, 0], np.uint64)
    _, caps = gpu_tree(zk, inp, 12, 0, 2)
    assert [int(v) for v in caps] == [2163910501769503938, 9976732063159483418, 662985512748194034,
                                      3626198389901409849]


def test_monolith_kat_direct(zk):
    """tests/tests.cu:201-235: Monolith hash_one(0..11)."""
    _, caps = gpu_tree(zk, np.arange(12, dtype=np.uint64), 12, 0, 4)
    assert [int(v) for v in caps] == [0xCB4EF9B3FE5BCA9E, 0xE03C9506D19C8216, 0x2F05CFB355E880C, 0xF614E84BF4DF8342]


@pytest.mark.parametrize("nl,nc,cap", GO_CAPS)
def test_go_cap_kats(zk, nl, nc, cap):
    leaves = np.array(GO_LEAF7 * nl, dtype=np.uint64)
    _, caps = gpu_tree(zk, leaves, 7, nc.bit_length() - 1, 0)
    for j in range(nc):
        assert [int(v) for v in caps[4 * j:4 * j + 4]] == cap


SHAPES = [(1, 7, 0), (1, 7, 1), (2, 9, 1), (3, 5, 1), (4, 2, 2), (5, 4, 3), (6, 3, 2), (7, 8, 0), (8, 16, 4),
          (8, 17, 8), (9, 1, 4), (10, 135, 4), (11, 20, 0), (12, 80, 4), (9, 12, 9), (10, 400, 2),
          # Poseidon's whole-wave (matrix-core) leaf path: one wave of 5-element leaves, a
          # one-element second absorb chunk, caps == leaves
          (6, 5, 0), (6, 9, 1), (7, 6, 7)]


@pytest.mark.parametrize("ht", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("lg_l,ls,h", SHAPES)
def test_tree_vs_oracle(zk, ht, lg_l, ls, h):
    print(f"tree lg_l={lg_l} ls={ls} h={h} ht={ht}", flush=True)
    leaves = ob.canonical(ls << lg_l, 1000 * lg_l + 10 * ls + h + ht)
    dig, caps = gpu_tree(zk, leaves, ls, h, ht)
    edig, ecap = ob.merkle(leaves, ls, h, ht)
    assert np.array_equal(caps, ecap)
    assert np.array_equal(dig, edig)


@pytest.mark.parametrize("ht", [0, 1, 3, 4])
@pytest.mark.parametrize("lg_l,ls,h", [(13, 135, 4),   # 2^13 x 135 leaves, cap height 4
                                       (17, 5, 3),     # two passes over HBM
                                       (18, 5, 0)])    # three passes over HBM
def test_large_tree_vs_oracle(zk, ht, lg_l, ls, h):
    print(f"large tree lg_l={lg_l} ls={ls} h={h} ht={ht}", flush=True)
    leaves = ob.canonical(ls << lg_l, 77 + lg_l + ht)
    dig, caps = gpu_tree(zk, leaves, ls, h, ht)
    edig, ecap = ob.merkle(leaves, ls, h, ht)
    assert np.array_equal(caps, ecap)
    assert np.array_equal(dig, edig)


def test_cpu_entry_point_runs_on_gpu(zk):
    ls, lg_l, h = 9, 8, 2
    leaves = ob.canonical(ls << lg_l, 5)
    nl, nc = 1 << lg_l, 1 << h
    nd = 2 * (nl - nc)
    dig = np.zeros(nd * 4, np.uint64)
    cap = np.zeros(nc * 4, np.uint64)
    zk.fill_digests_buf_linear_cpu(dig, cap, leaves, nd, nc, nl, ls, h, 0)
    edig, ecap = ob.merkle(leaves, ls, h, 0)
    assert np.array_equal(cap, ecap) and np.array_equal(dig, edig)


def test_invalid_arguments_are_reported(zk):
    d = zk.DeviceBuffer(64)
    c = zk.DeviceBuffer(16)
    l = zk.DeviceBuffer(256)
    c.zero()
    with pytest.raises(zk.ZeknoxError):  # digests size inconsistent
        zk.fill_digests_buf_linear_gpu_with_gpu_ptr(d, c, l, 10, 4, 32, 8, 2, 0)
    with pytest.raises(zk.ZeknoxError):  # leaves not a power of two
        zk.fill_digests_buf_linear_gpu_with_gpu_ptr(d, c, l, 2 * (24 - 4), 4, 24, 8, 2, 0)
    with pytest.raises(zk.ZeknoxError):  # hash type not provided
        zk.fill_digests_buf_linear_gpu_with_gpu_ptr(d, c, l, 2 * (32 - 4), 4, 32, 8, 2, 5)
    assert np.all(c.to_host() == 0)  # nothing written


@pytest.fixture
def lone_levels(zk, request):
    """Merkle schedule of one tree through the C ABI: finished inside the
    leaf launch ("fused") or one launch per level above the leaves
    ("by_level"); restored to the library default afterwards."""
    zk.set_lone_level_launches(None if request.param == "fused" else 0)
    yield request.param
    zk.set_lone_level_launches(zk.LONE_LEVEL_LAUNCHES_DEFAULT)


@pytest.mark.parametrize("lone_levels", ["fused", "by_level"], indirect=True)
def test_four_pass_tree(zk, lone_levels):
    """2^25 leaves, one cap, 256-leaf tiles: fused, the leaf pass and three
    upper passes run in one launch (the last work-group of each group of
    tiles goes on); by level, 25 level launches.  Keccak with 2-element
    leaves keeps the oracle's work to the 2^25 inner nodes."""
    print(f"four-pass tree 2^25 x 2, cap height 0, Keccak, {lone_levels}", flush=True)
    leaves = ob.canonical(2 << 25, 2525)
    dig, caps = gpu_tree(zk, leaves, 2, 0, 1)
    edig, ecap = ob.merkle(leaves, 2, 0, 1)
    assert np.array_equal(caps, ecap)
    assert np.array_equal(dig, edig)


@pytest.mark.parametrize("lone_levels", ["by_level"], indirect=True)
@pytest.mark.parametrize("ht", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("lg_l,ls,h", [(1, 7, 0), (2, 9, 1), (5, 4, 3), (7, 6, 7), (9, 12, 9), (10, 135, 4),
                                       (12, 80, 4), (13, 135, 0)])
def test_tree_by_level_vs_oracle(zk, lone_levels, ht, lg_l, ls, h):
    """One tree with its levels launched apart (leaf digests, then one launch
    per level down to the caps): caps == leaves, one-leaf subtrees, width-2
    tops, one and many subtrees, every hash type."""
    print(f"tree by level lg_l={lg_l} ls={ls} h={h} ht={ht}", flush=True)
    leaves = ob.canonical(ls << lg_l, 3000 * lg_l + 10 * ls + h + ht)
    dig, caps = gpu_tree(zk, leaves, ls, h, ht)
    edig, ecap = ob.merkle(leaves, ls, h, ht)
    assert np.array_equal(caps, ecap)
    assert np.array_equal(dig, edig)


@pytest.mark.parametrize("ht,lg_l", [(2, 15), (1, 18), (4, 18), (0, 20), (3, 20)])
def test_lone_schedules_agree_at_switch_sizes(zk, ht, lg_l):
    """At the smallest tree each hash builds by level under the default rule
    (Poseidon-BN128 2^15 leaves, Keccak-256 and Monolith 2^18, Poseidon and
    Poseidon2 2^20), the default, fused and by-level schedules give the same
    digests and caps (the fused schedule is the one checked against the
    oracle at these hashes' smaller sizes)."""
    print(f"lone schedules agree ht={ht} lg_l={lg_l}", flush=True)
    ls, h = 20, 4
    leaves = ob.canonical(ls << lg_l, 7100 + 37 * ht + lg_l)
    out = {}
    try:
        for name, v in (("default", zk.LONE_LEVEL_LAUNCHES_DEFAULT), ("fused", None), ("by_level", 0)):
            zk.set_lone_level_launches(v)
            out[name] = gpu_tree(zk, leaves, ls, h, ht)
    finally:
        zk.set_lone_level_launches(zk.LONE_LEVEL_LAUNCHES_DEFAULT)
    for name in ("fused", "by_level"):
        assert np.array_equal(out[name][1], out["default"][1]), name
        assert np.array_equal(out[name][0], out["default"][0]), name
