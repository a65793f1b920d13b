"""Edge cases and full-size checks through the C ABI: empty batches, salt
only, the largest single transforms and the 400 x 2^20 LDE at full size,
checked through properties that do not depend on size (an inverse undoes a
forward transform, a delta transforms to all ones), and very long leaves."""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _coset(zk):
    zk.init_coset_rs(0, 24, 7)


def dev_cfg(zk, **kw):
    return zk.NTTConfig(are_inputs_on_device=True, are_outputs_on_device=True, **kw)


def test_empty_batches_are_no_ops(zk):
    out = zk.DeviceBuffer(64)
    out.fill_random(1)
    before = out.to_host()
    inp = zk.DeviceBuffer(64)
    zk.lde_batch(0, out, inp, 4, dev_cfg(zk, batches=0, extension_rate_bits=2, with_coset=True))
    zk.ntt_batch(0, out, 4, dev_cfg(zk, batches=0))
    zk.transpose_rev_batch(0, out, inp, 4, zk.TransposeConfig(batches=0, are_inputs_on_device=True,
                                                               are_outputs_on_device=True))
    assert np.array_equal(out.to_host(), before)


def test_lde_salt_only(zk):
    """No polynomial columns, salt_size columns of random canonical elements."""
    lg_n, rate, salt = 6, 2, 3
    out = zk.DeviceBuffer(salt << (lg_n + rate))
    out.zero()
    zk.lde_batch(0, out, zk.DeviceBuffer(1), lg_n,
                 dev_cfg(zk, batches=0, extension_rate_bits=rate, salt_size=salt))
    v = out.to_host()
    assert int(v.max()) < ob.P
    assert len(np.unique(v)) > v.size - 4


@pytest.mark.parametrize("lg", [22, 24])
def test_large_single_column_ntt_roundtrip(zk, lg):
    print(f"NTT round trip 2^{lg}", flush=True)
    n = 1 << lg
    buf = zk.DeviceBuffer(n)
    buf.fill_random(lg)
    x = buf.to_host()
    zk.ntt_batch(0, buf, lg, dev_cfg(zk, batches=1))
    y = buf.to_host()
    assert not np.array_equal(x, y)
    zk.intt_batch(0, buf, lg, dev_cfg(zk, batches=1))
    assert np.array_equal(buf.to_host(), x)


def test_large_ntt_of_delta_is_all_ones(zk):
    lg = 24
    x = np.zeros(1 << lg, np.uint64)
    x[0] = 1
    buf = zk.DeviceBuffer.from_host(x)
    zk.ntt_batch(0, buf, lg, dev_cfg(zk, batches=1))
    assert np.all(buf.to_host() == 1)


def test_full_size_lde_400x2e20_undone_by_inverse_ntt(zk):
    """The 8 x 400 x 2^20 workload's per-job LDE (no coset, rate 1) at full
    size: the inverse NTT of the 2^21 evaluations gives back the 2^20
    coefficients of every column followed by zeros."""
    print("LDE 400 x 2^20 -> 2^21 round trip", flush=True)
    cols, lg_n, rate = 400, 20, 1
    n, N = 1 << lg_n, 1 << (lg_n + rate)
    coeffs = zk.DeviceBuffer(cols * n)
    coeffs.fill_random(400)
    out = zk.DeviceBuffer(cols * N)
    zk.lde_batch(0, out, coeffs, lg_n, dev_cfg(zk, batches=cols, extension_rate_bits=rate, with_coset=False))
    zk.intt_batch(0, out, lg_n + rate, dev_cfg(zk, batches=cols))
    back = out.to_host().reshape(cols, N)
    x = coeffs.to_host().reshape(cols, n)
    assert np.array_equal(back[:, :n], x)
    assert not back[:, n:].any()


@pytest.mark.parametrize("ht", [0, 1, 3])
def test_very_long_leaves(zk, ht):
    """Leaves of 10,000 elements (1250 Poseidon permutations per leaf)."""
    ls, lg, h = 10000, 4, 2
    leaves = ob.canonical(ls << lg, 10000 + ht)
    dig, caps = zk.merkle_tree(leaves, ls, h, ht)
    edig, ecap = ob.merkle(leaves, ls, h, ht)
    assert np.array_equal(caps, ecap) and np.array_equal(dig, edig)
