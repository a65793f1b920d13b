"""Schedule autotuning: every schedule computes the same commit (bit-exact
against the oracle), the tuner records a valid schedule per shape, and the
shape table survives save / load."""
import itertools

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

SHAPES = [  # (cols, lg_n, rate, hash)
    (20, 12, 2, 0),
    (7, 14, 1, 3),
    (3, 10, 3, 0),
    (40, 13, 3, 3),
]
SCHEDULES = [(s, e, t, v) for (s, e), t, v in itertools.product([(9, 0), (9, 1), (6, 0), (6, 1), (4, 1)], [8, 9, 10],
                                                                 [0, 1]) if v == 0 or s == 9]


@pytest.fixture(scope="module", autouse=True)
def _coset(zk):
    zk.init_coset_rs(0, 24, 7)
    yield
    zk.set_autotune(False)
    zk.tuning_clear()


def oracle_commit(coeffs, cols, lg_n, rate, h, ht):
    lde = ob.lde(coeffs, lg_n, rate, cols)
    leaves = ob.transpose_rev(lde, lg_n + rate, cols)
    dig, cap = ob.merkle(leaves, cols, h, ht)
    return leaves, dig, cap


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_every_schedule_gives_the_same_commit(zk, shape):
    cols, lg_n, rate, ht = shape
    coeffs = ob.canonical(cols << lg_n, 900 + lg_n)
    want = oracle_commit(coeffs, cols, lg_n, rate, 4, ht)
    d_in = zk.DeviceBuffer.from_host(coeffs)
    for sched in SCHEDULES:
        c = zk.Commit(d_in, cols, lg_n, rate, 4, ht)
        zk.set_tuning(zk.Tuning(*sched), c)
        zk.commit_batch([c])
        got = c.results()
        for a, b, name in zip(got, want, ("leaves", "digests", "caps")):
            assert np.array_equal(a, b), f"{name} differ under schedule {sched}"
    zk.tuning_clear()


def test_c_abi_path_under_a_non_default_schedule(zk):
    cols, lg_n, rate, h = 9, 14, 2, 3
    lg = lg_n + rate
    coeffs = ob.canonical(cols << lg_n, 77)
    zk.set_tuning(zk.Tuning(5, 1, 10))
    try:
        out = np.zeros(cols << lg, np.uint64)
        zk.lde_batch(0, out, coeffs, lg_n, zk.NTTConfig(batches=cols, extension_rate_bits=rate, with_coset=True))
        want = ob.lde(coeffs, lg_n, rate, cols)
        assert np.array_equal(out, want)
        leaves = ob.transpose_rev(want, lg, cols)
        dig, cap = zk.merkle_tree(leaves, cols, h, 0)
        wd, wc = ob.merkle(leaves, cols, h, 0)
        assert np.array_equal(dig, wd) and np.array_equal(cap, wc)
    finally:
        zk.tuning_clear()


@pytest.mark.parametrize("sched", [None, (4, 1, 8, 0), (4, 1, 9, 1), (4, 1, 10, 0)])
def test_c_abi_merkle_default_and_set_defaults(zk, sched):
    """The C-ABI Merkle entry point hashes with 2^10-lane work-groups until a
    default is set, then with the default's; zkx_tuning_clear restores the
    built-in one.  Poseidon, 12-element leaves (matrix-core or VALU form)."""
    cols, lg_l, cap_h = 12, 13, 4
    leaves = ob.canonical(cols << lg_l, 4100)
    want = ob.merkle(leaves, cols, cap_h, 0)
    if sched is not None:
        zk.set_tuning(zk.Tuning(*sched))
    try:
        for rep in range(2):
            got = zk.merkle_tree(leaves, cols, cap_h, 0)
            assert all(np.array_equal(g, w) for g, w in zip(got, want)), f"schedule {sched} call {rep}"
    finally:
        zk.tuning_clear()
    assert zk.get_tuning().astuple() == zk.Tuning().astuple()


def test_autotune_records_a_valid_schedule(zk, tmp_path):
    cols, lg_n, rate, ht = 30, 13, 2, 0
    coeffs = ob.canonical(cols << lg_n, 5)
    c = zk.Commit(coeffs, cols, lg_n, rate, 4, ht)
    zk.autotune([c], iters=2)
    t = zk.get_tuning(c).astuple()
    assert 1 <= t[0] <= 9 and t[1] in (0, 1) and 8 <= t[2] <= 10 and t[3] in (0, 1)
    # the outputs of the tuning runs are the job's result
    want = oracle_commit(coeffs, cols, lg_n, rate, 4, ht)
    for a, b in zip(c.results(), want):
        assert np.array_equal(a, b)
    # save / clear / load
    path = tmp_path / "tune.txt"
    zk.tuning_save(str(path))
    zk.tuning_clear()
    assert zk.get_tuning(c).astuple() == (9, 0, 8, 0)
    zk.tuning_load(str(path))
    assert zk.get_tuning(c).astuple() == t
    zk.tuning_clear()


def test_autotune_on_first_sight(zk):
    cols, lg_n, rate, ht = 11, 12, 1, 3
    coeffs = ob.canonical(cols << lg_n, 6)
    c = zk.Commit(coeffs, cols, lg_n, rate, 2, ht)
    zk.tuning_clear()
    zk.set_autotune(True)
    try:
        zk.commit_batch([c])
    finally:
        zk.set_autotune(False)
    got = c.results()
    want = oracle_commit(coeffs, cols, lg_n, rate, 2, ht)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    t = zk.get_tuning(c).astuple()
    assert 1 <= t[0] <= 9 and 8 <= t[2] <= 10
    zk.tuning_clear()


def test_bad_schedule_rejected(zk):
    with pytest.raises(zk.ZeknoxError):
        zk.set_tuning(zk.Tuning(9, 0, 11))
    with pytest.raises(zk.ZeknoxError):
        zk.set_tuning(zk.Tuning(0, 0, 8))
    with pytest.raises(zk.ZeknoxError):
        zk.set_tuning(zk.Tuning(9, 0, 8, 2))
