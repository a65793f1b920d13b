"""The benchmark workloads themselves, checked bit-exactly against expected
outputs written by tests/golden/make_workload_golden.py (the oracle's coset
LDE and leaf rows, then the reference's own CPU Merkle forest,
native/merkle/merkle.cpp:126-161):

* the mixed batch -- 64 Poseidon jobs, log2 degree U[10,16], columns from
  {2,20,80,135}, rate_bits {1,2,3}, cap height 4 (numpy seed 3) -- run in ONE
  zkx_commit_batch under autotuned schedules;
* a replay of the served stream -- Poisson arrivals at 80/s for 4 s (seed 7),
  log2 degree lognormal clipped to [12,20], columns U[2,400], rate_bits
  {1,2,3}, Poseidon / Poseidon2 -- submitted on its own clock to a server
  whose budget holds the largest request and little more, so requests queue
  for memory and engines run side by side.  Requests of at most 5e6
  permutations are compared with the expected outputs; the larger ones with
  the same request run alone (digests and caps), and their leaf rows by the
  Merkle forest of the served rows, which must reproduce those digests.

Leaf rows, digests and caps are compared as SHA-256 of their little-endian
words (tolerance zero)."""
import hashlib
import json
import os
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from zeknox_amd.workloads import CAP_HEIGHT, mixed_batch_jobs, served_trace

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "workloads.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="module", autouse=True)
def _coset(zk):
    zk.init_coset_rs(0, 24, 7)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def test_batch_config3_seeded_64_jobs(zk, golden):
    jobs = mixed_batch_jobs()
    exp = golden["batch64"]
    assert len(exp) == len(jobs) == 64
    commits = []
    for (cols, lg_n, rate, ht, seed), e in zip(jobs, exp):
        assert (cols, lg_n, rate, ht, seed) == (e["cols"], e["lg_n"], e["rate"], e["hash"], e["seed"])
        x = zk.DeviceBuffer(cols << lg_n)
        x.fill_random(seed)
        commits.append(zk.Commit(x, cols, lg_n, rate, CAP_HEIGHT, ht))
    print("autotuning the 64-job batch's shapes", flush=True)
    zk.autotune(commits, iters=1)
    for c in commits:  # the autotuner wrote results: clear them before the real batch
        c.d_leaves.zero()
        c.d_digests.zero()
        c.d_caps.zero()
    print("one zkx_commit_batch of 64 jobs", flush=True)
    zk.commit_batch(commits)
    with ThreadPoolExecutor(8) as ex:
        got = list(ex.map(lambda c: tuple(sha(a) for a in c.results()), commits))
    for j, ((ls, ds, cs), e) in enumerate(zip(got, exp)):
        assert ls == e["leaves_sha256"], f"job {j} leaf rows"
        assert ds == e["digests_sha256"], f"job {j} digests"
        assert cs == e["caps_sha256"], f"job {j} caps"


@pytest.mark.parametrize("streams", [3, 6])
def test_served_config5_replay(zk, golden, streams):
    spec = golden["served"]
    trace = served_trace(spec["seed"], spec["seconds"], spec["rate_hz"])
    exp = spec["requests"]
    assert len(trace) == len(exp)
    pool = zk.DeviceBuffer(max(c << l for _, c, l, _, _ in trace))
    pool.fill_random(spec["seed"])
    feet = [zk.Server.footprint(c, l, r, CAP_HEIGHT) for _, c, l, r, _ in trace]
    budget = int(max(feet) * 1.15)
    print(f"served replay with {streams} engines", flush=True)
    srv = zk.Server(budget, streams=streams, max_batch=16)
    ids = {}
    lock = threading.Lock()
    errors = []

    def submitter():
        t0 = time.monotonic()
        try:
            for i, (ta, c, l, r, ht) in enumerate(trace):
                delay = ta - (time.monotonic() - t0)
                if delay > 0:
                    time.sleep(delay)
                with lock:  # the id is recorded before wait_any can hand it out
                    ids[srv.submit(pool, c, l, r, CAP_HEIGHT, ht)] = i
        except Exception as e:  # surfaced by the main thread
            errors.append(e)

    th = threading.Thread(target=submitter)
    th.start()
    hashes, large, spans = {}, {}, {}
    returned = 0
    with ThreadPoolExecutor(8) as ex:
        while returned < len(trace):
            with lock:
                pending = len(ids) - returned
            if pending == 0:
                assert not errors, errors
                time.sleep(0.0005)
                continue
            rid, res = srv.wait_any()
            returned += 1
            with lock:
                i = ids[rid]
            assert res.state == 2, f"request {i} {trace[i]} failed"
            _, c, l, r, ht = trace[i]
            spans[i] = (res.t_submit, res.t_start, res.t_done)
            if "leaves_sha256" in exp[i]:
                leaves, dig, cap = srv.results(res, c, l, r, CAP_HEIGHT)
                hashes[i] = ex.submit(lambda a, b, d: (sha(a), sha(b), sha(d)), leaves, dig, cap)
            else:
                # the Merkle forest of the served leaf rows, on the device
                N = 1 << (l + r)
                nc = 1 << CAP_HEIGHT
                nd = 2 * (N - nc)
                dd, dc = zk.DeviceBuffer(nd * 4), zk.DeviceBuffer(nc * 4)
                zk.fill_digests_buf_linear_gpu_with_gpu_ptr(dd, dc, res.leaves, nd, nc, N, c, CAP_HEIGHT, ht)
                _, dig, cap = srv.results(res, c, l, r, CAP_HEIGHT, leaves=False)
                large[i] = (dig, cap, dd.to_host(), dc.to_host())
            srv.release(rid)
        th.join()
        assert not errors, errors
        st = srv.stats()
        srv.close()
        waited = sum(1 for s, a, _ in spans.values() if a > s + 1e-3)
        # requests admitted at different times (so in different batches) whose runs overlapped
        iv = sorted((a, d) for _, a, d in spans.values())
        overlapped = sum(1 for k in range(1, len(iv)) if iv[k][0] > iv[k - 1][0] and iv[k][0] < max(
            d for _, d in iv[:k]))
        print(f"served {st['completed']} requests in {st['batches']} batches; {waited} waited for admission; "
              f"{overlapped} started while an earlier batch ran; peak {st['peak_bytes'] / 2**30:.1f} GiB of "
              f"{budget / 2**30:.1f}", flush=True)
        assert st["completed"] == len(trace) and st["failed"] == 0
        assert st["peak_bytes"] <= budget and st["bytes_in_use"] == 0
        assert st["batches"] < len(trace) and waited > 0 and overlapped > 0
        for i, f in hashes.items():
            e = exp[i]
            assert f.result() == (e["leaves_sha256"], e["digests_sha256"], e["caps_sha256"]), f"request {i} {trace[i]}"
    for i, (dig, cap, dig_rows, cap_rows) in large.items():
        _, c, l, r, ht = trace[i]
        alone = zk.Commit(pool, c, l, r, CAP_HEIGHT, ht, leaves=False)  # reads the pool's first c * 2^l words
        zk.commit_batch([alone])
        _, adig, acap = alone.results()
        assert np.array_equal(dig, adig) and np.array_equal(cap, acap), f"request {i} {trace[i]} vs alone"
        assert np.array_equal(dig_rows, adig) and np.array_equal(cap_rows, acap), f"request {i} leaf rows"
