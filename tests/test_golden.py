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


# ---------------------------------------------------------------- whole commits

def commit_id(r):
    return f"{r['cols']}x2e{r['lg_n']}r{r['rate']}H{r['hash']}"


def check_commit(rec, leaves, dig, cap):
    assert np.array_equal(cap, u64s(rec["caps"]))
    assert sha(dig) == rec["digests_sha256"]
    if leaves is not None:
        assert sha(leaves) == rec["leaves_sha256"]


SMALL_COMMITS = [r for r in FIX["commits"] if r["cols"] << (r["lg_n"] + r["rate"]) <= 1 << 26]


@pytest.mark.parametrize("rec", SMALL_COMMITS, ids=commit_id)
def test_oracle_whole_commit_matches_reference_fixture(rec):
    """The oracle's coset LDE and leaf rows, then its Merkle, against the
    fixture whose Merkle the reference's merkle.cpp computed."""
    cols, lg_n, rate = rec["cols"], rec["lg_n"], rec["rate"]
    leaves = ob.transpose_rev(ob.lde(ob.canonical(cols << lg_n, rec["seed"]), lg_n, rate, cols), lg_n + rate, cols)
    dig, cap = ob.merkle(leaves, cols, rec["cap_height"], rec["hash"])
    check_commit(rec, leaves, dig, cap)


@pytest.mark.gpu
@pytest.mark.parametrize("rec", FIX["commits"], ids=commit_id)
def test_gpu_whole_commit_matches_reference_fixture(zk, rec):
    """The batched engine at full size (the headline commit and one job of
    the 400 x 2^20 batch), every leaf row, digest and cap."""
    zk.init_coset_rs(0, 24, 7)
    cols, lg_n, rate, h, ht = rec["cols"], rec["lg_n"], rec["rate"], rec["cap_height"], rec["hash"]
    print(f"whole commit {commit_id(rec)}", flush=True)
    c = zk.Commit(ob.canonical(cols << lg_n, rec["seed"]), cols, lg_n, rate, h, ht)
    zk.commit_batch([c])
    check_commit(rec, *c.results())


@pytest.mark.gpu
def test_gpu_plain_headline_commit_matches_reference_fixture(zk):
    """The same headline commit through the plain C-ABI sequence:
    compute_batched_lde -> compute_transpose_rev -> fill_digests (device)."""
    rec = FIX["commits"][0]
    zk.init_coset_rs(0, 24, 7)
    cols, lg_n, rate, h, ht = rec["cols"], rec["lg_n"], rec["rate"], rec["cap_height"], rec["hash"]
    lg, N = lg_n + rate, 1 << (lg_n + rate)
    d_in = zk.DeviceBuffer.from_host(ob.canonical(cols << lg_n, rec["seed"]))
    d_lde, d_leaves = zk.DeviceBuffer(cols * N), zk.DeviceBuffer(cols * N)
    zk.lde_batch(0, d_lde, d_in, lg_n, zk.NTTConfig(batches=cols, extension_rate_bits=rate, with_coset=True,
                                                   are_inputs_on_device=True, are_outputs_on_device=True))
    zk.transpose_rev_batch(0, d_leaves, d_lde, lg, zk.TransposeConfig(batches=cols, are_inputs_on_device=True,
                                                                      are_outputs_on_device=True))
    nc, nd = 1 << h, 2 * (N - (1 << h))
    d_dig, d_cap = zk.DeviceBuffer(nd * 4), zk.DeviceBuffer(nc * 4)
    zk.fill_digests_buf_linear_gpu_with_gpu_ptr(d_dig, d_cap, d_leaves, nd, nc, N, cols, h, ht, 0)
    check_commit(rec, d_leaves.to_host(), d_dig.to_host(nd * 4), d_cap.to_host())


@pytest.mark.gpu
def test_gpu_eight_resident_400x2e20_jobs(zk):
    """Eight 400 x 2^20 rate-1 Poseidon2 jobs resident at once (about 82 GB of
    coefficients, leaf rows and digests) in ONE zkx_commit_batch.  Job 0 holds
    the fixture's coefficients: its leaf rows, digests and caps are the
    fixture's.  Jobs 1-7 hold device-generated coefficients: each one's
    digests and caps equal those of the plain C-ABI sequence
    (compute_batched_lde -> compute_transpose_rev -> fill_digests on the
    device) over the same coefficients; their leaf rows are covered by those
    digests, each leaf digest being the hash of its row."""
    rec = next(r for r in FIX["commits"] if r["cols"] == 400)
    zk.init_coset_rs(0, 24, 7)
    cols, lg_n, rate, h, ht = rec["cols"], rec["lg_n"], rec["rate"], rec["cap_height"], rec["hash"]
    lg, N = lg_n + rate, 1 << (lg_n + rate)
    nc, nd = 1 << h, 2 * (N - (1 << h))
    commits = [zk.Commit(ob.canonical(cols << lg_n, rec["seed"]), cols, lg_n, rate, h, ht)]
    for j in range(1, 8):
        x = zk.DeviceBuffer(cols << lg_n)
        x.fill_random(4000 + j)
        commits.append(zk.Commit(x, cols, lg_n, rate, h, ht))
    print(f"one zkx_commit_batch of 8 resident {commit_id(rec)} jobs", flush=True)
    zk.commit_batch(commits)
    got = [(sha(c.d_digests.to_host(nd * 4)), c.d_caps.to_host()) for c in commits]
    assert sha(commits[0].d_leaves.to_host()) == rec["leaves_sha256"], "job 0 leaf rows"
    assert got[0][0] == rec["digests_sha256"], "job 0 digests"
    assert np.array_equal(got[0][1], u64s(rec["caps"])), "job 0 caps"
    for c in commits[1:]:  # the plain sequence below needs their leaf-row memory
        c.d_leaves.free()
    d_lde, d_leaves = zk.DeviceBuffer(cols * N), zk.DeviceBuffer(cols * N)
    d_dig, d_cap = zk.DeviceBuffer(nd * 4), zk.DeviceBuffer(nc * 4)
    for j, c in enumerate(commits[1:], start=1):
        print(f"job {j}: plain C-ABI sequence over the same coefficients", flush=True)
        zk.lde_batch(0, d_lde, c.d_coeffs, lg_n, zk.NTTConfig(batches=cols, extension_rate_bits=rate, with_coset=True,
                                                             are_inputs_on_device=True, are_outputs_on_device=True))
        zk.transpose_rev_batch(0, d_leaves, d_lde, lg, zk.TransposeConfig(batches=cols, are_inputs_on_device=True,
                                                                          are_outputs_on_device=True))
        zk.fill_digests_buf_linear_gpu_with_gpu_ptr(d_dig, d_cap, d_leaves, nd, nc, N, cols, h, ht, 0)
        assert sha(d_dig.to_host(nd * 4)) == got[j][0], f"job {j} digests"
        assert np.array_equal(d_cap.to_host(), got[j][1]), f"job {j} caps"


# ------------------------------------- benchmark workloads (workloads.json)

WL = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "workloads.json")))


def _oracle_commit_sha(coeffs, cols, lg_n, rate, ht):
    """SHA-256 of the oracle's leaf rows, digests and caps, and the first cap."""
    leaves = ob.transpose_rev(ob.lde(coeffs, lg_n, rate, cols), lg_n + rate, cols)
    dig, cap = ob.merkle(leaves, cols, 4, ht)
    return (sha(leaves), sha(dig), sha(cap)), [str(int(v)) for v in cap[:4]]


def test_workload_jobs_match_their_generators():
    """The fixture's job lists are the ones zeknox_amd.workloads draws."""
    from zeknox_amd.workloads import mixed_batch_jobs, served_trace
    assert [(r["cols"], r["lg_n"], r["rate"], r["hash"], r["seed"]) for r in WL["batch64"]] == mixed_batch_jobs()
    s = WL["served"]
    tr = served_trace(s["seed"], s["seconds"], s["rate_hz"])
    assert [(r["cols"], r["lg_n"], r["rate"], r["hash"]) for r in s["requests"]] == [t[1:] for t in tr]


@pytest.mark.parametrize("j", [j for j, r in enumerate(WL["batch64"]) if r["lg_n"] + r["rate"] <= 13][:6])
def test_oracle_matches_batch64_fixture(j):
    """Small jobs of the mixed batch recomputed by the oracle alone (its Merkle
    forest, not the reference build that wrote the fixture)."""
    r = WL["batch64"][j]
    got, cap0 = _oracle_commit_sha(ob.salt(r["cols"] << r["lg_n"], r["seed"]), r["cols"], r["lg_n"], r["rate"],
                                   r["hash"])
    assert got == (r["leaves_sha256"], r["digests_sha256"], r["caps_sha256"]) and cap0 == r["cap0"]


def test_oracle_matches_served_fixture():
    """The smallest requests of the served trace (both hash types), prefixes
    of one seeded buffer, recomputed by the oracle."""
    s = WL["served"]
    reqs = sorted((r for r in s["requests"] if "leaves_sha256" in r), key=lambda r: r["cols"] << (r["lg_n"] + r["rate"]))
    picked = [next(r for r in reqs if r["hash"] == h) for h in (0, 3)]
    pool = ob.salt(max(r["cols"] << r["lg_n"] for r in picked), s["seed"])
    for r in picked:
        got, cap0 = _oracle_commit_sha(pool[: r["cols"] << r["lg_n"]], r["cols"], r["lg_n"], r["rate"], r["hash"])
        assert got == (r["leaves_sha256"], r["digests_sha256"], r["caps_sha256"]) and cap0 == r["cap0"]
