"""Committed golden fixtures (tests/golden/fixtures.json).

The Merkle expectations were produced by the reference's own CPU code
(native/merkle/merkle.cpp + its Poseidon / Poseidon2 hashers, compiled from
its sources); the LDE expectations by the oracle.  Inputs are regenerated
from (count, seed) with the shared splitmix64 generator.  The CPU tests pin
the oracle to them; the GPU tests pin the library to them.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_bind as ob

FIX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def u64s(xs):
    return np.array([int(v) for v in xs], dtype=np.uint64)


def leaves_of(rec):
    return ob.canonical(rec["leaf_size"] << rec["lg_leaves"], rec["seed"])


def check_tree(rec, dig, cap):
    assert np.array_equal(cap, u64s(rec["caps"]))
    if "digests" in rec:
        assert np.array_equal(dig, u64s(rec["digests"]))
    assert sha(dig) == rec["digests_sha256"]


# ---------------------------------------------------------------- oracle (CPU)

@pytest.mark.parametrize("rec", FIX["merkle"] + FIX["merkle_sha256"],
                         ids=lambda r: f"L{r['lg_leaves']}x{r['leaf_size']}h{r['cap_height']}H{r['hash']}")
def test_oracle_merkle_matches_reference_fixture(rec):
    dig, cap = ob.merkle(leaves_of(rec), rec["leaf_size"], rec["cap_height"], rec["hash"])
    check_tree(rec, dig, cap)


@pytest.mark.parametrize("rec", FIX["lde"], ids=lambda r: f"n{r['lg_n']}r{r['rate']}c{r['cols']}")
def test_oracle_lde_fixture(rec):
    x = ob.canonical(rec["cols"] << rec["lg_n"], rec["seed"])
    lde = ob.lde(x, rec["lg_n"], rec["rate"], rec["cols"], coset=rec["coset"])
    assert sha(lde) == rec["lde_sha256"]
    assert sha(ob.transpose_rev(lde, rec["lg_n"] + rec["rate"], rec["cols"])) == rec["leaves_sha256"]


# ---------------------------------------------------------------- library (GPU)

@pytest.mark.gpu
@pytest.mark.parametrize("rec", FIX["merkle"] + FIX["merkle_sha256"],
                         ids=lambda r: f"L{r['lg_leaves']}x{r['leaf_size']}h{r['cap_height']}H{r['hash']}")
def test_gpu_merkle_matches_reference_fixture(zk, rec):
    dig, cap = zk.merkle_tree(leaves_of(rec), rec["leaf_size"], rec["cap_height"], rec["hash"])
    check_tree(rec, dig, cap)


@pytest.mark.gpu
@pytest.mark.parametrize("rec", FIX["lde"], ids=lambda r: f"n{r['lg_n']}r{r['rate']}c{r['cols']}")
def test_gpu_lde_fixture(zk, rec):
    zk.init_coset_rs(0, 24, 7)
    cols, lg_n, rate = rec["cols"], rec["lg_n"], rec["rate"]
    x = ob.canonical(cols << lg_n, rec["seed"])
    out = np.zeros(cols << (lg_n + rate), np.uint64)
    zk.lde_batch(0, out, x, lg_n, zk.NTTConfig(batches=cols, extension_rate_bits=rate, with_coset=rec["coset"]))
    assert sha(out) == rec["lde_sha256"]
    leaves = np.zeros_like(out)
    zk.transpose_rev_batch(0, leaves, out, lg_n + rate, zk.TransposeConfig(batches=cols))
    assert sha(leaves) == rec["leaves_sha256"]
