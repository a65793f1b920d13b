"""The device Goldilocks routines the hash and NTT kernels are built from,
against Python integers, on inputs chosen to reach every carry and borrow
case of the reductions (zeknox_ext.h zkx_gl_ops)."""
import itertools

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

P = ob.P
M64 = (1 << 64) - 1
EPS = (1 << 32) - 1
# 64-bit values around the field and word boundaries (lazy values reach 2^64 - 1)
SPECIAL = [0, 1, 2, 3, EPS - 1, EPS, 1 << 32, (1 << 32) + 1, 1 << 48, 1 << 63, P - 2, P - 1, P, P + 1,
           M64 - EPS, M64 - 1, M64, 0x123456789ABCDEF0, (1 << 64) - (1 << 33) + 1]
WORDS32 = [0, 1, 2, 1 << 31, EPS - 1, EPS]


def u64(xs):
    return np.array([x & M64 for x in xs], dtype=np.uint64)


def rand64(n, seed):
    return [int(x) for x in np.random.default_rng(seed).integers(0, 2**64, size=n, dtype=np.uint64)]


def check(zk, op, xs, ys, f):
    got = zk.gl_ops(op, u64(xs), u64(ys))
    exp = u64([f(x, y) for x, y in zip(xs, ys)])
    bad = np.nonzero(got != exp)[0]
    assert bad.size == 0, [(hex(xs[i]), hex(ys[i]), hex(int(got[i])), hex(int(exp[i]))) for i in bad[:5]]


def test_products_of_boundary_values(zk):
    vals = SPECIAL + rand64(40, 1)
    xs, ys = zip(*itertools.product(vals, vals))
    check(zk, zk.GL_MUL_LAZY, xs, ys, lambda x, y: x * y % P)
    check(zk, zk.GL_MUL, xs, ys, lambda x, y: x * y % P)


def test_random_products(zk):
    xs, ys = rand64(1 << 16, 2), rand64(1 << 16, 3)
    check(zk, zk.GL_MUL_LAZY, xs, ys, lambda x, y: x * y % P)


def test_reduce128_every_carry_case(zk):
    """lo + h0 * EPS with and without carry, minus h1 with and without
    borrow, all four combinations (hi = h1 * 2^32 + h0)."""
    los = SPECIAL + rand64(8, 4)
    cases = [(lo, (h1 << 32) | h0) for lo in los for h0 in WORDS32 for h1 in WORDS32]
    xs, ys = zip(*cases)
    check(zk, zk.GL_REDUCE128, xs, ys, lambda lo, hi: (hi * 2**64 + lo) % P)
    seen = set()
    for lo, hi in cases:
        u = lo + (hi & EPS) * EPS
        c, u = u >> 64, u & M64
        seen.add((c, int(u < hi >> 32)))
    assert seen == {(0, 0), (0, 1), (1, 0), (1, 1)}


def test_reduce96(zk):
    los = SPECIAL + rand64(8, 5)
    xs, ys = zip(*[(lo, t) for lo in los for t in WORDS32 + rand64(4, 6)])
    check(zk, zk.GL_REDUCE96, xs, ys, lambda lo, t: ((t & EPS) * 2**64 + lo) % P)


def test_pow7_add_sub_canon(zk):
    vals = SPECIAL + rand64(200, 7)
    check(zk, zk.GL_POW7, vals, vals, lambda x, _: pow(x, 7, P))
    check(zk, zk.GL_CANON, vals, vals, lambda x, _: x % P)
    canon = [v % P for v in vals]
    xs, ys = zip(*itertools.product(canon[:60], canon[:60]))
    check(zk, zk.GL_ADD, xs, ys, lambda x, y: (x + y) % P)
    check(zk, zk.GL_SUB, xs, ys, lambda x, y: (x - y) % P)


def test_bad_op_rejected(zk):
    with pytest.raises(zk.ZeknoxError):
        zk.gl_ops(9, np.zeros(4, np.uint64), np.zeros(4, np.uint64))
