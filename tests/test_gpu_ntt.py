"""GPU NTT / LDE / transpose through the C ABI vs the oracle. Bit-exact."""
import numpy as np
import pytest

import oracle_bind as ob
from oracle_bind import P, orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _coset(zk):
    zk.init_coset_rs(0, 24, 7)


@pytest.mark.parametrize("lg_n,rate,B", [(1, 1, 2), (2, 1, 3), (3, 2, 1), (4, 3, 5), (5, 1, 2), (8, 2, 3),
                                         (10, 3, 2), (11, 1, 4), (12, 2, 2), (13, 1, 3), (14, 3, 2), (16, 1, 2),
                                         (15, 3, 2)])
@pytest.mark.parametrize("coset", [True, False])
def test_lde_forward_vs_oracle(zk, lg_n, rate, B, coset):
    print(f"lde lg_n={lg_n} rate={rate} B={B} coset={coset}", flush=True)
    x = ob.canonical(B << lg_n, 17 * lg_n + rate + B)
    out = np.zeros(B << (lg_n + rate), np.uint64)
    zk.lde_batch(0, out, x, lg_n, zk.NTTConfig(batches=B, extension_rate_bits=rate, with_coset=coset))
    assert np.array_equal(out, ob.lde(x, lg_n, rate, B, coset=coset))


@pytest.mark.parametrize("lg_n,rate,B", [(3, 1, 2), (9, 2, 3), (13, 1, 2)])
@pytest.mark.parametrize("coset", [True, False])
def test_lde_inverse_vs_oracle(zk, lg_n, rate, B, coset):
    x = ob.canonical(B << lg_n, 5 * lg_n + rate)
    out = np.zeros(B << (lg_n + rate), np.uint64)
    zk.lde_batch(0, out, x, lg_n, zk.NTTConfig(batches=B, extension_rate_bits=rate, with_coset=coset),
                 direction=zk.INVERSE)
    assert np.array_equal(out, ob.lde(x, lg_n, rate, B, inverse=True, coset=coset))


def test_lde_device_pointers_and_salt(zk):
    lg_n, rate, B, salt = 10, 2, 3, 2
    x = ob.canonical(B << lg_n, 3)
    N = 1 << (lg_n + rate)
    din = zk.DeviceBuffer.from_host(x)
    dout = zk.DeviceBuffer((B + salt) * N)
    cfg = zk.NTTConfig(batches=B, extension_rate_bits=rate, with_coset=True, are_inputs_on_device=True,
                       are_outputs_on_device=True, salt_size=salt)
    zk.lde_batch(0, dout, din, lg_n, cfg)
    got = dout.to_host()
    assert np.array_equal(got[:B * N], ob.lde(x, lg_n, rate, B))
    assert np.all(got[B * N:] < np.uint64(P))


@pytest.mark.parametrize("lg,B", [(1, 3), (4, 2), (9, 1), (10, 3), (12, 2), (13, 2), (17, 1), (20, 1)])
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("coset", [False, True])
def test_ntt_vs_oracle(zk, lg, B, inverse, coset):
    print(f"ntt lg={lg} B={B} inverse={inverse} coset={coset}", flush=True)
    x = ob.canonical(B << lg, 7 * lg + B + 2 * inverse + coset)
    buf = x.copy()
    cfg = zk.NTTConfig(batches=B, with_coset=coset)
    (zk.intt_batch if inverse else zk.ntt_batch)(0, buf, lg, cfg)
    assert np.array_equal(buf, ob.ntt(x, lg, B, inverse, coset))


# First passes of 5..8 stages (the two-round strided kernel) over more tiles
# than the device holds work-groups, so every work-group walks several tiles
# (prefetch and twiddle reloads); in place (NTT, with the coset it reads
# through the one-tile kernel) and from coefficients (LDE).
@pytest.mark.parametrize("lg,B", [(17, 40), (18, 16), (19, 8), (20, 8)])
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("coset", [False, True])
def test_ntt_two_round_first_pass_vs_oracle(zk, lg, B, inverse, coset):
    print(f"ntt lg={lg} B={B} inverse={inverse} coset={coset}", flush=True)
    x = ob.canonical(B << lg, 11 * lg + B + 2 * inverse + coset)
    buf = x.copy()
    (zk.intt_batch if inverse else zk.ntt_batch)(0, buf, lg, zk.NTTConfig(batches=B, with_coset=coset))
    assert np.array_equal(buf, ob.ntt(x, lg, B, inverse, coset))


@pytest.mark.parametrize("lg_n,rate,B", [(17, 1, 8), (18, 2, 4), (19, 1, 4), (20, 1, 4)])
@pytest.mark.parametrize("coset", [True, False])
def test_lde_two_round_first_pass_vs_oracle(zk, lg_n, rate, B, coset):
    print(f"lde lg_n={lg_n} rate={rate} B={B} coset={coset}", flush=True)
    x = ob.canonical(B << lg_n, 13 * lg_n + rate + B)
    out = np.zeros(B << (lg_n + rate), np.uint64)
    zk.lde_batch(0, out, x, lg_n, zk.NTTConfig(batches=B, extension_rate_bits=rate, with_coset=coset))
    assert np.array_equal(out, ob.lde(x, lg_n, rate, B, coset=coset))


def test_ntt_roundtrip_on_device(zk):
    lg, B = 18, 2
    x = ob.canonical(B << lg, 42)
    d = zk.DeviceBuffer.from_host(x)
    cfg = zk.NTTConfig(batches=B, are_inputs_on_device=True, are_outputs_on_device=True)
    zk.ntt_batch(0, d, lg, cfg)
    zk.intt_batch(0, d, lg, cfg)
    assert np.array_equal(d.to_host(), x)


@pytest.mark.parametrize("lg,B", [(0, 3), (1, 2), (5, 7), (5, 32), (6, 33), (7, 137), (8, 400), (10, 135), (11, 271),
                                  (12, 85), (13, 2)])
def test_transpose_rev_vs_oracle(zk, lg, B):
    x = ob.canonical(B << lg, 9 * lg + B)
    out = np.zeros_like(x)
    zk.transpose_rev_batch(0, out, x, lg, zk.TransposeConfig(batches=B))
    assert np.array_equal(out, ob.transpose_rev(x, lg, B))


def test_lde_135x2e15_rate3_vs_oracle(zk):
    """135 columns of 2^15 coefficients, rate 3, coset 7 -> 2^18-point LDE."""
    lg_n, rate, B = 15, 3, 135
    x = ob.canonical(B << lg_n, 2024)
    din = zk.DeviceBuffer.from_host(x)
    dout = zk.DeviceBuffer(B << (lg_n + rate))
    cfg = zk.NTTConfig(batches=B, extension_rate_bits=rate, with_coset=True, are_inputs_on_device=True,
                       are_outputs_on_device=True)
    zk.lde_batch(0, dout, din, lg_n, cfg)
    assert np.array_equal(dout.to_host(), ob.lde(x, lg_n, rate, B))


def test_invalid_lde_arguments(zk):
    x = np.zeros(8, np.uint64)
    out = np.zeros(64, np.uint64)
    with pytest.raises(zk.ZeknoxError) as e:
        zk.lde_batch(0, out, x, 0, zk.NTTConfig(batches=1, extension_rate_bits=1))
    assert e.value.code == 1
    with pytest.raises(zk.ZeknoxError) as e:
        zk.lde_batch(0, out, x, 3, zk.NTTConfig(batches=1, extension_rate_bits=0))
    assert e.value.code == 1
