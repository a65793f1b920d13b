"""zkx_fill_digests_buf_linear_gpu_noalloc (the Merkle forest with its
scratch in a fixed device array) against fill_digests_buf_linear_gpu_with_gpu_ptr
on the same device buffers, bit-exact, two calls in a row with the second on
a smaller input; and its rejection of a forest beyond the fused launch's
maximum (checked on the host, before any launch)."""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu


def _trees(zk, leaves, ls, h, ht):
    """(digests, caps) of the default entry point and of the allocation-free
    one over the same device leaves."""
    nl = leaves.size // ls
    nc = 1 << h
    nd = 2 * (nl - nc) if nc != nl else 0
    dl = zk.DeviceBuffer.from_host(leaves)
    out = []
    for fn in (zk.fill_digests_buf_linear_gpu_with_gpu_ptr, zk.fill_digests_buf_linear_gpu_noalloc):
        dd = zk.DeviceBuffer(max(nd, 1) * 4)
        dc = zk.DeviceBuffer(nc * 4)
        dd.zero()
        dc.zero()
        fn(dd, dc, dl, nd, nc, nl, ls, h, ht, 0)
        out.append(((dd.to_host(nd * 4) if nd else np.zeros(0, np.uint64)), dc.to_host()))
    return out


# (leaves log2, leaf size, cap height, hash): each pair is a larger forest
# (upper passes fused into the launch, so arrival counters in the scratch)
# followed by a smaller one
PAIRS = [
    ((18, 20, 4, 0), (12, 7, 2, 0)),
    ((20, 9, 0, 3), (16, 135, 4, 3)),
    ((17, 135, 3, 0), (9, 4, 1, 0)),
    ((16, 12, 2, 1), (10, 3, 10, 4)),
]


@pytest.mark.parametrize("big,small", PAIRS, ids=[f"{a[0]}x{a[1]}h{a[3]}-{b[0]}x{b[1]}" for a, b in PAIRS])
def test_noalloc_equals_default_entry(zk, big, small):
    for i, (lg, ls, h, ht) in enumerate((big, small)):
        print(f"noalloc call {i}: 2^{lg} leaves x {ls}, cap {h}, hash {ht}", flush=True)
        leaves = ob.canonical(ls << lg, 4000 + 17 * lg + ls)
        (d0, c0), (d1, c1) = _trees(zk, leaves, ls, h, ht)
        assert np.array_equal(d1, d0)
        assert np.array_equal(c1, c0)
        if lg <= 12:
            dig, cap = ob.merkle(leaves, ls, h, ht)
            assert np.array_equal(d1, dig) and np.array_equal(c1, cap)


def test_noalloc_rejects_beyond_maximum(zk):
    """2^44 leaves in one forest need more than the fused launch's four
    passes at every work-group size: rejected with code 1 before any launch
    (the buffers are never touched)."""
    d = zk.DeviceBuffer(4)
    nl, nc = 1 << 44, 1
    with pytest.raises(zk.ZeknoxError) as e:
        zk.fill_digests_buf_linear_gpu_noalloc(d, d, d, 2 * (nl - nc), nc, nl, 8, 0, 0, 0)
    assert e.value.code == 1
    assert "too large" in str(e.value)
    # the entry point still works afterwards
    leaves = ob.canonical(8 << 10, 5)
    (d0, c0), (d1, c1) = _trees(zk, leaves, 8, 2, 0)
    assert np.array_equal(d1, d0) and np.array_equal(c1, c0)
