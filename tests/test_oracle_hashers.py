"""Oracle Keccak and Monolith (hash types 1 and 4) pinned by a library
routine, the reference's known answers and the reference's own CPU code."""
import hashlib

import numpy as np
import pytest

import oracle_bind as ob


# ---- Keccak permutation and sponge ------------------------------------------

@pytest.mark.parametrize("n", [0, 1, 7, 8, 135, 136, 137, 271, 272, 500])
def test_keccak_sponge_is_sha3_with_sha3_padding(n):
    """The same permutation and sponge with the SHA-3 domain byte gives
    hashlib's SHA3-256: pins keccak-f[1600], the absorb loop and padding."""
    data = bytes((i * 37 + 11) & 0xFF for i in range(n))
    assert ob.keccak(data, 32, 136, 0x06) == hashlib.sha3_256(data).digest()


def test_keccak256_of_empty_input():
    # Keccak-256 (original padding) of the empty string, a published constant
    assert ob.keccak(b"", 32, 136, 0x01).hex() == \
        "c5d2460186f7233c927e7db2dcc703c0e500b653ca82273b7bfad8045d85a470"


def test_keccak_hasher_reference_kats():
    """tests/tests.cu:162-199: hash_one of the first `size` words, size 1..6."""
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
        assert [int(v) for v in ob.hash_one(np.array(data[:size], np.uint64), 1)] == want


def test_keccak_hash_two_uses_25_bytes_of_each_side():
    l = ob.canonical(4, 1)
    r = ob.canonical(4, 2)
    h = ob.hash_two(l, r, 1)
    raw = b"".join(int(v).to_bytes(8, "little") for v in l)[:25] + \
        b"".join(int(v).to_bytes(8, "little") for v in r)[:25]
    md = ob.keccak(raw, 32, 136, 0x01)
    want = [int.from_bytes(md[8 * i:8 * i + 8], "little") for i in range(4)]
    want[3] &= 0xFF
    assert [int(v) for v in h] == want
    # bytes beyond the 25th of either side do not matter
    l2 = l.copy()
    l2[3] ^= np.uint64(0xFFFF_FFFF_FFFF_FF00)
    assert np.array_equal(ob.hash_two(l2, r, 1), h)


# ---- Monolith ----------------------------------------------------------------

def test_monolith_reference_kat():
    """tests/tests.cu:201-235: hash_one(0..11)."""
    want = [0xCB4EF9B3FE5BCA9E, 0xE03C9506D19C8216, 0x2F05CFB355E880C, 0xF614E84BF4DF8342]
    assert [int(v) for v in ob.hash_one(np.arange(12, dtype=np.uint64), 4)] == want


def test_monolith_bar_sbox_is_a_permutation_of_bytes():
    # the 8-bit chi-like S-box of the bars layer is invertible: bars of the
    # 256 words i * 0x0101010101010101 (one repeated byte) hit 256 distinct words
    outs = set()
    for b in range(256):
        s = np.zeros(12, np.uint64)
        s[0] = np.uint64(b * 0x0101010101010101)
        outs.add(int(ob.permute(s, 4)[0]))
    assert len(outs) == 256


# ---- against the reference's own CPU code --------------------------------------

REF = ob.ref_lib()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (no reference tree)")


@needs_ref
@pytest.mark.parametrize("ht", [1, 4])
@pytest.mark.parametrize("n", [1, 3, 4, 5, 8, 9, 16, 17, 18, 34, 135])
def test_hash_one_matches_reference_cpu(ht, n):
    x = ob.canonical(n, 100 + n)
    want = np.zeros(4, np.uint64)
    REF.ref_hash_one(ob._ptr(x), n, ob._ptr(want), ht)
    assert np.array_equal(ob.hash_one(x, ht), want)


@needs_ref
@pytest.mark.parametrize("ht", [1, 4])
def test_hash_two_matches_reference_cpu(ht):
    for seed in range(5):
        l, r = ob.canonical(4, 2 * seed), ob.canonical(4, 2 * seed + 1)
        if ht == 1:  # Keccak digests carry 25 bytes
            l[3] &= np.uint64(0xFF)
            r[3] &= np.uint64(0xFF)
        want = np.zeros(4, np.uint64)
        REF.ref_hash_two(ob._ptr(l), ob._ptr(r), ob._ptr(want), ht)
        assert np.array_equal(ob.hash_two(l, r, ht), want)


@needs_ref
@pytest.mark.parametrize("ht", [1, 4])
@pytest.mark.parametrize("lg,ls,h", [(3, 7, 1), (6, 5, 2), (8, 20, 4), (9, 4, 3), (5, 2, 5), (7, 135, 0)])
def test_merkle_matches_reference_cpu(ht, lg, ls, h):
    leaves = ob.canonical(ls << lg, 7 * lg + ls)
    nl, nc = 1 << lg, 1 << h
    nd = 2 * (nl - nc) if nl != nc else 0
    d = np.zeros(max(nd, 1) * 4, np.uint64)
    c = np.zeros(nc * 4, np.uint64)
    REF.fill_digests_buf_linear_cpu(ob._ptr(d), ob._ptr(c), ob._ptr(leaves), nd, nc, nl, ls, h, ht)
    dig, cap = ob.merkle(leaves, ls, h, ht)
    assert np.array_equal(cap, c)
    assert np.array_equal(dig, d[:nd * 4])


# ---- Poseidon over BN254 (hash type 2) -------------------------------------------

BN_R = 21888242871839275222246405745257275088548364400416034343698204186575808495617


def _limbs(x):
    return np.array([(x >> (64 * i)) & (2**64 - 1) for i in range(4)], np.uint64)


def _int(a):
    return sum(int(v) << (64 * i) for i, v in enumerate(a))


def test_bn254_montgomery_arithmetic_vs_big_integers():
    import random
    rnd = random.Random(7)
    rinv = pow(2**256, -1, BN_R)
    for _ in range(200):
        a, b = rnd.randrange(BN_R), rnd.randrange(BN_R)
        z = np.zeros(4, np.uint64)
        ob.orc.orc_bn_mont_mul(ob._ptr(z), ob._ptr(_limbs(a)), ob._ptr(_limbs(b)))
        assert _int(z) == a * b * rinv % BN_R
        ob.orc.orc_bn_add(ob._ptr(z), ob._ptr(_limbs(a)), ob._ptr(_limbs(b)))
        assert _int(z) == (a + b) % BN_R
    for a, b in [(BN_R - 1, BN_R - 1), (0, BN_R - 1), (1, 1)]:
        z = np.zeros(4, np.uint64)
        ob.orc.orc_bn_mont_mul(ob._ptr(z), ob._ptr(_limbs(a)), ob._ptr(_limbs(b)))
        assert _int(z) == a * b * rinv % BN_R


def test_poseidon_bn128_reference_kat():
    """tests/tests.cu:492-536: hash_one of 12 words."""
    inp = np.array([8917524657281059100, 13029010200779371910, 16138660518493481604, 17277322750214136960,
                    1441151880423231822, 0, 0, 0, 0, 0, 0, 0], np.uint64)
    want = [2163910501769503938, 9976732063159483418, 662985512748194034, 3626198389901409849]
    assert [int(v) for v in ob.hash_one(inp, 2)] == want


@needs_ref
@pytest.mark.parametrize("n", [1, 4, 5, 8, 12, 17, 135])
def test_poseidon_bn128_hash_one_matches_reference_cpu(n):
    x = ob.canonical(n, 300 + n)
    want = np.zeros(4, np.uint64)
    REF.ref_hash_one(ob._ptr(x), n, ob._ptr(want), 2)
    assert np.array_equal(ob.hash_one(x, 2), want)


@needs_ref
def test_poseidon_bn128_hash_two_and_tree_match_reference_cpu():
    for seed in range(3):
        l, r = ob.canonical(4, 40 + seed), ob.canonical(4, 50 + seed)
        want = np.zeros(4, np.uint64)
        REF.ref_hash_two(ob._ptr(l), ob._ptr(r), ob._ptr(want), 2)
        assert np.array_equal(ob.hash_two(l, r, 2), want)
    for lg, ls, h in [(3, 7, 1), (5, 12, 2), (4, 3, 4)]:
        leaves = ob.canonical(ls << lg, 3 * lg + ls)
        nl, nc = 1 << lg, 1 << h
        nd = 2 * (nl - nc) if nl != nc else 0
        d = np.zeros(max(nd, 1) * 4, np.uint64)
        c = np.zeros(nc * 4, np.uint64)
        REF.fill_digests_buf_linear_cpu(ob._ptr(d), ob._ptr(c), ob._ptr(leaves), nd, nc, nl, ls, h, 2)
        dig, cap = ob.merkle(leaves, ls, h, 2)
        assert np.array_equal(cap, c) and np.array_equal(dig, d[:nd * 4])
