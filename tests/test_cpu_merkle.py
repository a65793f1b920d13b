"""fill_digests_buf_linear_cpu: the library's own host Merkle forest
(merkle/merkle.cpp:126-161 in the reference), called through the C ABI on the
CPU with no HIP device.

Pinned by the reference's Go CPU cap KATs (wrappers/go/lib/lib_test.go:310-330),
the Poseidon / Poseidon2 KATs (tests.cu:281-292, 425-428), the committed
fixtures written by the reference's own CPU build, the oracle on random
shapes of every hash type, and, where oracle/_ref is built, the reference's
CPU code itself.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_bind as ob
from test_golden import FIX, check_tree, leaves_of
from test_oracle import GO_CAPS, GO_LEAF7, LEAF9, POSEIDON_KAT, FIB12, POSEIDON2_KAT

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import zeknox_amd as zk  # noqa: E402


def cpu_tree(leaves, leaf_size, cap_height, hash_type):
    leaves = np.ascontiguousarray(leaves, dtype=np.uint64).ravel()
    nl = leaves.size // leaf_size
    nc = 1 << cap_height
    nd = 2 * (nl - nc) if nc != nl else 0
    dig = np.zeros(max(nd, 1) * 4, np.uint64)
    cap = np.zeros(nc * 4, np.uint64)
    zk.fill_digests_buf_linear_cpu(dig, cap, leaves, nd, nc, nl, leaf_size, cap_height, hash_type)
    return dig[: nd * 4], cap


@pytest.mark.parametrize("nl,nc,cap", GO_CAPS)
def test_go_cpu_cap_kats(nl, nc, cap):
    _, caps = cpu_tree(np.array(GO_LEAF7 * nl, dtype=np.uint64), 7, nc.bit_length() - 1, zk.HASH_POSEIDON)
    for j in range(nc):
        assert [int(v) for v in caps[4 * j:4 * j + 4]] == cap


@pytest.mark.parametrize("k", range(1, 10))
def test_poseidon_kat_through_caps_eq_leaves(k):
    # one leaf of k elements, cap height 0: the cap is hash_one(leaf)
    _, cap = cpu_tree(np.array(LEAF9[:k], dtype=np.uint64), k, 0, zk.HASH_POSEIDON)
    assert [int(v) for v in cap] == POSEIDON_KAT[k]


def test_poseidon2_kat_through_caps_eq_leaves():
    _, cap = cpu_tree(np.array(FIB12, dtype=np.uint64), 12, 0, zk.HASH_POSEIDON2)
    assert [int(v) for v in cap] == POSEIDON2_KAT


@pytest.mark.parametrize("rec", FIX["merkle"] + FIX["merkle_sha256"],
                         ids=lambda r: f"L{r['lg_leaves']}x{r['leaf_size']}h{r['cap_height']}H{r['hash']}")
def test_cpu_merkle_matches_reference_fixture(rec):
    dig, cap = cpu_tree(leaves_of(rec), rec["leaf_size"], rec["cap_height"], rec["hash"])
    check_tree(rec, dig, cap)


SHAPES = [  # (log2 leaves, leaf size, cap height)
    (0, 1, 0), (1, 3, 0), (1, 4, 1), (2, 5, 1), (3, 7, 2), (4, 8, 0), (5, 9, 3), (6, 12, 6), (6, 16, 5),
    (7, 17, 2), (8, 2, 4), (9, 20, 4), (10, 135, 4), (11, 24, 9),
]


@pytest.mark.parametrize("ht", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("lg_l,ls,h", SHAPES)
def test_cpu_merkle_vs_oracle(ht, lg_l, ls, h):
    if ht == 2 and (lg_l + (ls + 7) // 8) > 10:
        pytest.skip("Poseidon-BN128 is ~17x a Goldilocks permutation: small trees only")
    leaves = ob.canonical(ls << lg_l, 1000 + 17 * lg_l + ls + ht)
    dig, cap = cpu_tree(leaves, ls, h, ht)
    edig, ecap = ob.merkle(leaves, ls, h, ht)
    assert np.array_equal(cap, ecap)
    assert np.array_equal(dig, edig)


@pytest.mark.parametrize("ht", [0, 1, 2, 3, 4])
def test_cpu_merkle_vs_reference_build(ht):
    ref = ob.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    lg_l, ls, h = (8, 13, 2) if ht == 2 else (11, 45, 3)
    leaves = ob.canonical(ls << lg_l, 77 + ht)
    dig, cap = cpu_tree(leaves, ls, h, ht)
    nl, nc = 1 << lg_l, 1 << h
    nd = 2 * (nl - nc)
    rd, rc = np.zeros(nd * 4, np.uint64), np.zeros(nc * 4, np.uint64)
    ref.fill_digests_buf_linear_cpu(ob._ptr(rd), ob._ptr(rc), ob._ptr(leaves), nd, nc, nl, ls, h, ht)
    assert np.array_equal(cap, rc)
    assert np.array_equal(dig, rd)


@pytest.mark.parametrize("ht", [0, 3])
@pytest.mark.parametrize("lg_l,ls,h", [(12, 135, 4), (10, 9, 10), (11, 8, 2), (9, 5, 6), (8, 200, 0)])
def test_cpu_merkle_vector_and_scalar_paths_agree(monkeypatch, ht, lg_l, ls, h):
    """Poseidon and Poseidon2 hash eight leaves / nodes at a time with AVX-512
    where the CPU has it; ZKX_CPU_NO_AVX512=1 forces the scalar path.  Both
    give the same forest (and the oracle's)."""
    leaves = ob.canonical(ls << lg_l, 31 * lg_l + ls + ht)
    vec = cpu_tree(leaves, ls, h, ht)
    monkeypatch.setenv("ZKX_CPU_NO_AVX512", "1")
    sca = cpu_tree(leaves, ls, h, ht)
    monkeypatch.delenv("ZKX_CPU_NO_AVX512")
    assert all(np.array_equal(a, b) for a, b in zip(vec, sca))
    if lg_l <= 10:
        edig, ecap = ob.merkle(leaves, ls, h, ht)
        assert np.array_equal(vec[1], ecap) and np.array_equal(vec[0], edig)


def test_cpu_merkle_rejects_bad_sizes():
    leaves = np.zeros(8 * 4, np.uint64)
    with pytest.raises(zk.ZeknoxError):
        cpu_tree(leaves[:24], 8, 2, 0)  # 3 leaves
    dig, cap = np.zeros(64, np.uint64), np.zeros(16, np.uint64)
    with pytest.raises(zk.ZeknoxError):  # wrong digest count
        zk.fill_digests_buf_linear_cpu(dig, cap, leaves, 5, 4, 4 * 2, 4, 2, 0)
    with pytest.raises(zk.ZeknoxError):  # unknown hash type
        zk.fill_digests_buf_linear_cpu(dig, cap, leaves, 8, 2, 4, 8, 1, 9)


def test_cpu_merkle_without_any_device():
    """The entry point needs no HIP device: run it in a process that sees none."""
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r); import zeknox_amd as zk\n"
        "leaf=[8395359103262935841,1377884553022145855,2370707998790318766,3651132590097252162,"
        "1141848076261006345,12736915248278257710,9898074228282442027]\n"
        "l=np.array(leaf*16,dtype=np.uint64); d=np.zeros(4*28,np.uint64); c=np.zeros(8,np.uint64)\n"
        "zk.fill_digests_buf_linear_cpu(d,c,l,28,2,16,7,1,0)\n"
        "assert zk.lib.cuda_available() is False, 'a device is visible'\n"
        "print(' '.join(str(int(v)) for v in c[:4]))\n" % ROOT)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="", ROCR_VISIBLE_DEVICES="", CUDA_VISIBLE_DEVICES="")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert [int(v) for v in r.stdout.split()] == GO_CAPS[2][2]
