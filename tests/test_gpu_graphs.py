"""Batches replayed as captured HIP graphs: the same job shapes committed
again and again with new inputs and new output buffers (the graph reads the
re-staged descriptors), with the program's own eager work and device
synchronisations between the replays; every replay equals the oracle and the
kernel-by-kernel launch.  Bit-exact (tolerance zero)."""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

# (cols, lg_n, rate, cap_height, hash): register-round and LDS-round NTT passes,
# leaves of <= 4 elements, caps == leaves, four hash types
SHAPES = [(2, 10, 1, 4, 0), (20, 11, 2, 4, 3), (80, 10, 3, 4, 0), (5, 3, 1, 0, 0), (17, 9, 2, 4, 1),
          (34, 8, 1, 3, 4), (135, 12, 1, 4, 0), (3, 7, 1, 2, 3)]


@pytest.fixture(scope="module", autouse=True)
def _coset(zk):
    zk.init_coset_rs(0, 24, 7)
    yield
    zk.set_commit_graphs(True)


def _oracle(x, cols, lg_n, rate, h, ht):
    leaves = ob.transpose_rev(ob.lde(x, lg_n, rate, cols), lg_n + rate, cols)
    dig, cap = ob.merkle(leaves, cols, h, ht)
    return leaves, dig, cap


def _eager_work(zk, seed):
    """A C-ABI LDE on the device stream, waited for."""
    x = ob.canonical(8 << 9, seed)
    d_in, d_out = zk.DeviceBuffer.from_host(x), zk.DeviceBuffer(8 << 10)
    zk.lde_batch(0, d_out, d_in, 9, zk.NTTConfig(batches=8, extension_rate_bits=1, with_coset=True,
                                                 are_inputs_on_device=True, are_outputs_on_device=True))
    assert np.array_equal(d_out.to_host(), ob.lde(x, 9, 1, 8))


def test_graph_replays_equal_oracle(zk):
    zk.set_commit_graphs(True)
    before = zk.commit_graphs_captured()
    for call in range(6):
        print(f"graph replay call {call}", flush=True)
        inputs = [ob.canonical(c << l, 7000 + 100 * call + i) for i, (c, l, r, h, ht) in enumerate(SHAPES)]
        commits = [zk.Commit(x, c, l, r, h, ht) for x, (c, l, r, h, ht) in zip(inputs, SHAPES)]
        zk.commit_batch(commits)
        for x, c, shape in zip(inputs, commits, SHAPES):
            for got, exp in zip(c.results(), _oracle(x, *shape)):
                assert np.array_equal(got, exp), f"call {call} shape {shape}"
        _eager_work(zk, 50 + call)
        zk.lib.zkx_device_synchronize(0)
    assert zk.commit_graphs_captured() > before  # the third call captured the launch list


def test_graph_and_direct_launches_agree(zk):
    x = [ob.canonical(c << l, 8100 + i) for i, (c, l, r, h, ht) in enumerate(SHAPES)]
    out = []
    for on in (False, True, True, True, True):
        zk.set_commit_graphs(on)
        commits = [zk.Commit(v, c, l, r, h, ht) for v, (c, l, r, h, ht) in zip(x, SHAPES)]
        zk.commit_batch(commits)
        out.append([c.results() for c in commits])
    for rep in out[1:]:
        for a, b in zip(rep, out[0]):
            for u, v in zip(a, b):
                assert np.array_equal(u, v)
