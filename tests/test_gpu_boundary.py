"""Boundary entry points a wrapper calls first or that complete the
NTT_Config surface: init_cuda_degree / init_cuda_degree_and_generator
(lib.cu:137-175, the first call of wrappers/rust/src/lib.rs:241-250 and
wrappers/go/lib/lib.go:82-86) followed by a coset LDE, and the seeded salted
LDE (the reference's salt, ntt.cu:541-546, is clock-seeded), all against the
oracle bit for bit."""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu


def host_lde(zk, x, lg_n, rate, cols, salt=0, seed=None):
    out = np.zeros((cols + salt) << (lg_n + rate), np.uint64)
    cfg = zk.NTTConfig(batches=cols, extension_rate_bits=rate, with_coset=True, salt_size=salt)
    if seed is None:
        zk.lde_batch(0, out, x, lg_n, cfg)
    else:
        zk.lde_batch_seeded(0, out, x, lg_n, cfg, seed)
    return out


def test_init_cuda_degree_then_coset_lde(zk):
    zk.init_cuda_degree_rs(17)
    assert zk.last_error() == ""
    for lg_n, rate, cols in [(13, 3, 5), (10, 1, 3), (15, 2, 2)]:
        x = ob.canonical(cols << lg_n, 440 + lg_n)
        assert np.array_equal(host_lde(zk, x, lg_n, rate, cols), ob.lde(x, lg_n, rate, cols))


def test_init_cuda_degree_and_generator_sets_the_coset(zk):
    try:
        zk.init_cuda_degree_and_generator_rs(14, 5)
        assert zk.last_error() == ""
        lg_n, rate, cols = 11, 2, 4
        x = ob.canonical(cols << lg_n, 451)
        assert np.array_equal(host_lde(zk, x, lg_n, rate, cols), ob.lde(x, lg_n, rate, cols, gen=5))
    finally:
        zk.init_coset_rs(0, 24, 7)  # the generator the other tests use


@pytest.mark.parametrize("lg_n,rate,cols,salt", [(10, 2, 3, 2), (12, 1, 1, 4), (6, 3, 0, 1), (14, 1, 5, 1)])
def test_seeded_salt_lde(zk, lg_n, rate, cols, salt):
    zk.init_coset_rs(0, 24, 7)
    N = 1 << (lg_n + rate)
    x = ob.canonical(max(cols, 1) << lg_n, 460 + lg_n)[: cols << lg_n]
    a = host_lde(zk, x, lg_n, rate, cols, salt, seed=0xC0FFEE)
    b = host_lde(zk, x, lg_n, rate, cols, salt, seed=0xC0FFEE)
    c = host_lde(zk, x, lg_n, rate, cols, salt, seed=0xC0FFEF)
    assert np.array_equal(a, b)
    if cols:
        assert np.array_equal(a[: cols * N], ob.lde(x, lg_n, rate, cols))
    assert np.array_equal(a[cols * N:], ob.salt(salt * N, 0xC0FFEE))
    assert np.array_equal(c[: cols * N], a[: cols * N]) and not np.array_equal(c[cols * N:], a[cols * N:])


def test_seeded_salt_lde_device_buffers(zk):
    zk.init_coset_rs(0, 24, 7)
    lg_n, rate, cols, salt = 13, 2, 6, 3
    N = 1 << (lg_n + rate)
    x = ob.canonical(cols << lg_n, 470)
    din, dout = zk.DeviceBuffer.from_host(x), zk.DeviceBuffer((cols + salt) * N)
    zk.lde_batch_seeded(0, dout, din, lg_n, zk.NTTConfig(batches=cols, extension_rate_bits=rate, with_coset=True,
                                                         salt_size=salt, are_inputs_on_device=True,
                                                         are_outputs_on_device=True), 99)
    got = dout.to_host()
    assert np.array_equal(got[: cols * N], ob.lde(x, lg_n, rate, cols))
    assert np.array_equal(got[cols * N:], ob.salt(salt * N, 99))
