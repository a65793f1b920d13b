"""The one-launch Merkle forest's pass plan and arrival-counter layout
(host logic, CPU only).

The fused launch hashes the leaves tile by tile; a tile that stops at its
root adds itself to the arrival counter of its group, and the work-group
that arrives last hashes the group's roots as a tile of the next pass, up to
the caps.  This test takes the plan from the library (zkx_merkle_plan) and
replays that protocol in random arrival orders, checking that every node of
every subtree is hashed exactly once, every cap written once, the counter
ranges of the passes are disjoint and inside the allocation, and every
counter is back at zero afterwards (merkle.cu:165-244 is the layout the
passes must cover).
"""
import random
import sys
import os

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import zeknox_amd as zk  # noqa: E402


def replay(nl, nc, tile_log, seed):
    plan, total = zk.merkle_plan(nl, nc, tile_log)
    lg = (nl // nc).bit_length() - 1  # per-subtree height
    # plan shape: widths shrink by each pass's tile, the last reaches the caps
    assert plan[0][0] == nl
    for p in range(len(plan) - 1):
        cnt, lw, tl, _ = plan[p]
        assert tl <= lw, plan
        assert plan[p + 1][0] == cnt >> tl and plan[p + 1][1] == lw - tl, plan
    assert plan[-1][1] <= plan[-1][2] or nc == nl, plan
    # counter ranges: pass p >= 1 owns [first_ctr, first_ctr + groups)
    ranges = []
    for p in range(1, len(plan)):
        cnt, _, tl, c0 = plan[p]
        ranges.append((c0, c0 + (cnt >> tl)))
    used = sorted(ranges)
    for (a0, a1), (b0, b1) in zip(used, used[1:]):
        assert a1 <= b0, used
    assert (used[-1][1] if used else 0) == total

    ctr = [0] * max(total, 1)
    hashed = {}  # (subtree, level width log, index) -> times
    caps = [0] * nc

    def run_tile(p, first):
        """Work-group running tile `first` (node index over all subtrees) of pass p."""
        cnt, lw, tl, _ = plan[p]
        cur = min(1 << tl, cnt)
        if lw == 0:  # caps == leaves
            for g in range(first, first + cur):
                caps[g] += 1
            return
        if p == 0:
            for g in range(first, first + cur):
                hashed[(g >> lw, lw, g & ((1 << lw) - 1))] = hashed.get((g >> lw, lw, g & ((1 << lw) - 1)), 0) + 1
        while True:
            if cur == 1:
                if p + 1 >= len(plan):
                    return
                _, _, ntl, c0 = plan[p + 1]
                grp = first >> ntl
                i = c0 + grp
                assert 0 <= i < total
                ctr[i] += 1
                if ctr[i] < (1 << ntl):
                    return
                ctr[i] = 0  # the last arrival resets its counter
                p += 1
                cur = 1 << ntl
                first = grp << ntl
                continue
            if lw == 1:  # width-2 top level: the caps of the subtrees in this tile
                for n in range(cur >> 1):
                    caps[(first >> 1) + n] += 1
                return
            half = cur >> 1
            lw -= 1
            first >>= 1
            for n in range(half):
                g = first + n
                key = (g >> lw, lw, g & ((1 << lw) - 1))
                hashed[key] = hashed.get(key, 0) + 1
            cur = half

    tiles = list(range(plan[0][0] >> plan[0][2])) if plan[0][0] >= (1 << plan[0][2]) else [0]
    random.Random(seed).shuffle(tiles)
    for b in tiles:
        run_tile(0, b << plan[0][2])
    assert all(c == 0 for c in ctr)
    assert caps == [1] * nc
    if nc != nl:
        # every node at every level below the caps (widths 2^lg .. 2^1) once
        want = {(s, w, i) for s in range(nc) for w in range(1, lg + 1) for i in range(1 << w)}
        assert set(hashed) == want
        assert all(v == 1 for v in hashed.values())


@pytest.mark.parametrize("tile_log", [8, 9, 10])
@pytest.mark.parametrize("lg_l,h", [(4, 0), (8, 0), (9, 1), (12, 4), (13, 4), (16, 4), (18, 4), (18, 0),
                                    (20, 2), (17, 17), (10, 9)])
def test_counter_plan_replay(lg_l, h, tile_log):
    replay(1 << lg_l, 1 << h, tile_log, seed=lg_l * 31 + h + tile_log)


def test_four_pass_plan_counters():
    plan, total = zk.merkle_plan(1 << 28, 1, 8)
    assert len(plan) == 4
    assert [c for c, _, _, _ in plan] == [1 << 28, 1 << 20, 1 << 12, 1 << 4]
    assert total == (1 << 12) + (1 << 4) + 1


def test_plan_rejects_bad_forest():
    with pytest.raises(zk.ZeknoxError):
        zk.merkle_plan(12, 4, 8)
    with pytest.raises(zk.ZeknoxError):
        zk.merkle_plan(16, 32, 8)
