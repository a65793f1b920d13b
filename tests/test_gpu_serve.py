"""Commit server: continuous batching under an HBM budget.  Every result is
bit-identical to the same job run alone (and to the oracle for small ones)."""
import threading
import time

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _coset(zk):
    zk.init_coset_rs(0, 24, 7)


SHAPES = [(20, 12, 2, 4, 0), (2, 10, 3, 4, 3), (135, 11, 1, 4, 0), (80, 13, 1, 4, 3), (7, 9, 2, 2, 0),
          (400, 10, 1, 4, 3), (9, 14, 1, 4, 0), (33, 12, 3, 4, 3)]


def alone(zk, x, cols, lg_n, rate, h, ht):
    c = zk.Commit(x, cols, lg_n, rate, h, ht)
    zk.commit_batch([c])
    return c.results()


def test_served_mix_equals_alone(zk):
    inputs = [ob.canonical(c << l, 900 + i) for i, (c, l, r, h, ht) in enumerate(SHAPES)]
    dev = [zk.DeviceBuffer.from_host(x) for x in inputs]
    # budget fits about three of these at once: forces queueing and several batches
    budget = 3 * max(zk.Server.footprint(c, l, r, h) for c, l, r, h, _ in SHAPES)
    srv = zk.Server(budget, streams=2, max_batch=4)
    ids = []
    for d, (c, l, r, h, ht) in zip(dev, SHAPES):
        ids.append(srv.submit(d, c, l, r, h, ht))
    for rid, x, (c, l, r, h, ht) in zip(ids, inputs, SHAPES):
        print(f"served cols={c} lg_n={l} rate={r} h={h} ht={ht}", flush=True)
        res = srv.wait(rid)
        assert res.state == 2
        got = srv.results(res, c, l, r, h)
        exp = alone(zk, x, c, l, r, h, ht)
        for a, b in zip(got, exp):
            assert np.array_equal(a, b)
        srv.release(rid)
    st = srv.stats()
    assert st["completed"] == len(SHAPES) and st["failed"] == 0
    assert st["batches"] >= 2 and st["peak_bytes"] <= budget and st["bytes_in_use"] == 0
    srv.close()


def test_served_results_match_oracle(zk):
    shapes = [(5, 6, 1, 2, 0), (3, 5, 2, 0, 3), (12, 7, 1, 3, 3)]
    srv = zk.Server(1 << 30, streams=2)
    pend = []
    for i, (c, l, r, h, ht) in enumerate(shapes):
        x = ob.canonical(c << l, 70 + i)
        d = zk.DeviceBuffer.from_host(x)
        pend.append((srv.submit(d, c, l, r, h, ht), x, d, (c, l, r, h, ht)))
    for rid, x, _, (c, l, r, h, ht) in pend:
        got = srv.results(srv.wait(rid), c, l, r, h)
        lde = ob.lde(x, l, r, c)
        leaves = ob.transpose_rev(lde, l + r, c)
        dig, cap = ob.merkle(leaves, c, h, ht)
        assert np.array_equal(got[0], leaves) and np.array_equal(got[1], dig) and np.array_equal(got[2], cap)
        srv.release(rid)
    srv.close()


def test_concurrent_submitters(zk):
    """Requests submitted from two threads while earlier ones run."""
    srv = zk.Server(4 << 30, streams=3, max_batch=8)
    x = ob.canonical(20 << 12, 5)
    d = zk.DeviceBuffer.from_host(x)
    exp = alone(zk, x, 20, 12, 2, 4, 0)
    ids, lock = [], threading.Lock()

    def producer():
        for _ in range(6):
            rid = srv.submit(d, 20, 12, 2, 4, 0)
            with lock:
                ids.append(rid)
            time.sleep(0.002)
    ts = [threading.Thread(target=producer) for _ in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    for rid in ids:
        got = srv.results(srv.wait(rid), 20, 12, 2, 4)
        for a, b in zip(got, exp):
            assert np.array_equal(a, b)
        srv.release(rid)
    assert srv.stats()["completed"] == 12
    srv.close()


def test_request_over_budget_is_rejected(zk):
    srv = zk.Server(1 << 20, streams=1)
    d = zk.DeviceBuffer(135 << 12)
    with pytest.raises(zk.ZeknoxError):
        srv.submit(d, 135, 12, 3, 4, 0)
    srv.close()


def test_wait_any_returns_each_request_once(zk):
    srv = zk.Server(2 << 30, streams=2, max_batch=3)
    shapes = SHAPES[:5]
    inputs = [ob.canonical(c << l, 300 + i) for i, (c, l, r, h, ht) in enumerate(shapes)]
    dev = [zk.DeviceBuffer.from_host(x) for x in inputs]
    ids = {srv.submit(d, c, l, r, h, ht): k for k, (d, (c, l, r, h, ht)) in enumerate(zip(dev, shapes))}
    seen = set()
    for _ in shapes:
        rid, res = srv.wait_any()
        assert rid in ids and rid not in seen and res.state == 2
        seen.add(rid)
        k = ids[rid]
        c, l, r, h, ht = shapes[k]
        for a, b in zip(srv.results(res, c, l, r, h), alone(zk, inputs[k], c, l, r, h, ht)):
            assert np.array_equal(a, b)
        srv.release(rid)
    with pytest.raises(zk.ZeknoxError):
        srv.wait_any()  # nothing pending
    srv.close()


def test_small_requests_run_beside_a_large_one(zk):
    """A large request (LDE above 2^28 elements) runs alone on one engine;
    the small requests submitted after it are admitted to the other engine
    while it runs (how the device interleaves the two streams' work-groups is
    the hardware's choice, so only the first admission is timed), and every
    result equals the same job run alone."""
    big = (135, 19, 2, 4, 0)  # 135 x 2^21 LDE elements
    c, l, r, h, ht = big
    assert (c << (l + r)) > (1 << 28)
    xb = zk.DeviceBuffer(c << l)
    xb.fill_random(31)
    small = [(20, 12, 2, 4, 3), (9, 10, 1, 3, 0), (33, 11, 3, 4, 0)]
    xs = [ob.canonical(sc << sl, 60 + i) for i, (sc, sl, _, _, _) in enumerate(small)]
    ds = [zk.DeviceBuffer.from_host(x) for x in xs]
    srv = zk.Server(24 << 30, streams=2, max_batch=8)
    bid = srv.submit(xb, c, l, r, h, ht)
    sids = [srv.submit(d, *s) for d, s in zip(ds, small)]
    starts = []
    for rid, x, s in zip(sids, xs, small):
        res = srv.wait(rid)
        starts.append(res.t_start)
        got = srv.results(res, *s[:4])
        for a, b in zip(got, alone(zk, x, *s)):
            assert np.array_equal(a, b)
        srv.release(rid)
    bres = srv.wait(bid)
    assert starts[0] < bres.t_done
    got = srv.results(bres, c, l, r, h)
    srv.release(bid)
    ref = zk.Commit(xb, c, l, r, h, ht)
    zk.commit_batch([ref])
    for a, b in zip(got, ref.results()):
        assert np.array_equal(a, b)
    st = srv.stats()
    assert st["completed"] == 4 and st["failed"] == 0 and st["bytes_in_use"] == 0
    srv.close()
