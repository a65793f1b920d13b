"""The reference's multi-GPU entry points on this one-device build:
compute_batched_lde_multi_gpu (lib.h:45-47, lib.cu:83-102) with num_gpu == 1
is compute_batched_lde over the whole input, and other device counts or
element counts that do not match the configuration return code 1;
fill_digests_buf_linear_multigpu_with_gpu_ptr (merkle.h:42-51) builds the
same forest as fill_digests_buf_linear_gpu_with_gpu_ptr.  Bit-exact."""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _coset(zk):
    zk.init_coset_rs(0, 24, 7)


@pytest.mark.parametrize("lg_n,rate,B", [(5, 1, 3), (12, 2, 2), (14, 3, 2)])
def test_lde_multi_gpu_one_device_vs_oracle(zk, lg_n, rate, B):
    x = ob.canonical(B << lg_n, 7 * lg_n + B)
    out = np.zeros(B << (lg_n + rate), np.uint64)
    cfg = zk.NTTConfig(batches=B, extension_rate_bits=rate, with_coset=True)
    zk.lde_batch_multi_gpu(out, x, 1, cfg, lg_n, x.size, out.size)
    assert np.array_equal(out, ob.lde(x, lg_n, rate, B, coset=True))


def test_lde_multi_gpu_rejects_other_device_counts(zk):
    lg_n, rate, B = 6, 1, 2
    x = ob.canonical(B << lg_n, 3)
    out = np.zeros(B << (lg_n + rate), np.uint64)
    cfg = zk.NTTConfig(batches=B, extension_rate_bits=rate, with_coset=True)
    with pytest.raises(zk.ZeknoxError) as e:
        zk.lde_batch_multi_gpu(out, x, 2, cfg, lg_n, x.size, out.size)
    assert e.value.code == 1
    with pytest.raises(zk.ZeknoxError) as e:
        zk.lde_batch_multi_gpu(out, x, 1, cfg, lg_n, x.size - 1, out.size)
    assert e.value.code == 1
    with pytest.raises(zk.ZeknoxError) as e:
        zk.lde_batch_multi_gpu(out, x, 1, cfg, lg_n, x.size, out.size - 1)
    assert e.value.code == 1


@pytest.mark.parametrize("hash_type", [0, 3])
@pytest.mark.parametrize("lg,leaf_size,cap_h", [(10, 135, 4), (6, 7, 0), (9, 20, 9)])
def test_merkle_multigpu_matches_single_gpu(zk, hash_type, lg, leaf_size, cap_h):
    L = 1 << lg
    nc = 1 << cap_h
    nd = 2 * (L - nc) if nc != L else 0
    leaves = zk.DeviceBuffer.from_host(ob.canonical(leaf_size * L, 11 * lg + hash_type))
    outs = []
    for multi in (False, True):
        dig, cap = zk.DeviceBuffer(max(nd, 1) * 4), zk.DeviceBuffer(nc * 4)
        if multi:
            # returns after its work is complete, like the single-GPU entry point
            zk.lib.fill_digests_buf_linear_multigpu_with_gpu_ptr(dig.ptr, cap.ptr, leaves.ptr, nd, nc, L, leaf_size,
                                                                  cap_h, hash_type)
        else:
            zk.fill_digests_buf_linear_gpu_with_gpu_ptr(dig, cap, leaves, nd, nc, L, leaf_size, cap_h, hash_type, 0)
        outs.append((dig.to_host(nd * 4) if nd else np.zeros(0, np.uint64), cap.to_host()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    _, ref_cap = ob.merkle(leaves.to_host(), leaf_size, cap_h, hash_type)
    assert np.array_equal(outs[1][1], ref_cap)
