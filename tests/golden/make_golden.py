#!/usr/bin/env python3
"""Write tests/golden/fixtures.json: expected outputs for the GPU and oracle
parity tests.

Merkle fixtures come from the reference's OWN CPU code
(native/merkle/merkle.cpp fill_digests_buf_linear_cpu with its Poseidon and
Poseidon2 hashers), compiled from its sources by oracle/Makefile into
oracle/_ref/libzkref.so; this script needs that build (i.e. the reference
tree) and is run once in the build container:

    make -C oracle ref && python3 tests/golden/make_golden.py

LDE / transpose fixtures come from the oracle (the reference has no CPU LDE);
the oracle's LDE is pinned by the brute-force DFT and closed-form tests.
Whole-commit fixtures chain the two: the oracle's coset LDE and leaf rows,
then the reference's Merkle over those rows (the headline 135 x 2^15 rate-3
commit and one 400 x 2^20 rate-1 job, a few minutes of CPU).

Inputs are not stored: they are regenerated from (count, seed) by the shared
splitmix64 generator (oracle orc_fill_canonical), so the fixture file stays
small.  Small outputs are stored in full; large ones as SHA-256 of the
little-endian u64 bytes.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind as ob  # noqa: E402

MERKLE = [  # (log2 leaves, leaf size, cap height, hash, seed)
    (3, 7, 3, 0, 11), (3, 7, 2, 0, 12), (4, 7, 1, 0, 13), (5, 9, 2, 3, 14), (6, 2, 4, 0, 15),
    (6, 4, 0, 3, 16), (7, 5, 3, 0, 17), (8, 8, 4, 3, 18), (8, 20, 4, 0, 19), (9, 135, 4, 3, 20),
    (10, 80, 4, 0, 21), (10, 1, 2, 3, 22), (11, 13, 9, 0, 23), (12, 135, 4, 0, 24),
    # Keccak-256 (1) and Monolith (4) trees
    (3, 7, 1, 1, 31), (5, 17, 2, 1, 32), (6, 4, 0, 1, 33), (8, 135, 4, 1, 34), (4, 2, 4, 1, 35),
    (3, 9, 1, 4, 41), (6, 12, 3, 4, 42), (8, 135, 4, 4, 43), (5, 4, 5, 4, 44),
    # Poseidon over BN254 (2)
    (3, 7, 1, 2, 51), (5, 12, 2, 2, 52), (7, 135, 4, 2, 53), (4, 3, 4, 2, 54),
]
LARGE = [  # hashed outputs: 2^13 x 135, cap height 4, every hash
    (13, 135, 4, 0, 1001), (13, 135, 4, 3, 1002), (13, 135, 4, 1, 1003), (13, 135, 4, 4, 1004),
    (13, 135, 4, 2, 1005),
]
LDE = [  # (lg_n, rate, cols, coset, seed)
    (4, 1, 3, True, 31), (6, 2, 2, True, 32), (8, 3, 4, False, 33), (10, 1, 5, True, 34),
]
COMMITS = [  # whole commits (coset LDE with shift 7 -> leaf rows -> Merkle): (cols, lg_n, rate, cap, hash, seed)
    (135, 15, 3, 4, 0, 2001),  # the headline commit: 135 x 2^15, rate 3, Poseidon
    (135, 15, 3, 4, 3, 2002),  # the same shape, Poseidon2
    (400, 20, 1, 4, 3, 2003),  # one job of the 8 x 400 x 2^20 batch (2^21 leaves of 400), Poseidon2
]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def ref_merkle(ref, leaves, ls, h, ht):
    nl = leaves.size // ls
    nc = 1 << h
    nd = 2 * (nl - nc) if nl != nc else 0
    d = np.zeros(max(nd, 1) * 4, np.uint64)
    c = np.zeros(nc * 4, np.uint64)
    ref.fill_digests_buf_linear_cpu(ob._ptr(d), ob._ptr(c), ob._ptr(leaves), nd, nc, nl, ls, h, ht)
    return d[:nd * 4], c


def main():
    ref = ob.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libzkref.so is not built (make -C oracle ref)")
    out = {"generator": "splitmix64 rejection-sampled into [0, p): oracle orc_fill_canonical(count, seed)",
           "merkle_source": "reference native/merkle/merkle.cpp (compiled from its sources)",
           "lde_source": "oracle/oracle.c orc_lde + orc_transpose_rev",
           "merkle": [], "merkle_sha256": [], "lde": []}
    for lg, ls, h, ht, seed in MERKLE:
        leaves = ob.canonical(ls << lg, seed)
        d, c = ref_merkle(ref, leaves, ls, h, ht)
        rec = {"lg_leaves": lg, "leaf_size": ls, "cap_height": h, "hash": ht, "seed": seed,
               "caps": [str(int(v)) for v in c], "digests_sha256": sha(d)}
        if d.size <= 256:
            rec["digests"] = [str(int(v)) for v in d]
        out["merkle"].append(rec)
    for lg, ls, h, ht, seed in LARGE:
        leaves = ob.canonical(ls << lg, seed)
        d, c = ref_merkle(ref, leaves, ls, h, ht)
        out["merkle_sha256"].append({"lg_leaves": lg, "leaf_size": ls, "cap_height": h, "hash": ht, "seed": seed,
                                     "digests_sha256": sha(d), "caps": [str(int(v)) for v in c]})
    for lg_n, rate, cols, coset, seed in LDE:
        x = ob.canonical(cols << lg_n, seed)
        lde = ob.lde(x, lg_n, rate, cols, coset=coset)
        leaves = ob.transpose_rev(lde, lg_n + rate, cols)
        out["lde"].append({"lg_n": lg_n, "rate": rate, "cols": cols, "coset": coset, "seed": seed,
                           "lde_sha256": sha(lde), "leaves_sha256": sha(leaves),
                           "lde_head": [str(int(v)) for v in lde[:8]]})
    out["commits"] = []
    for cols, lg_n, rate, h, ht, seed in COMMITS:
        print("commit", cols, lg_n, rate, h, ht, flush=True)
        x = ob.canonical(cols << lg_n, seed)
        leaves = ob.transpose_rev(ob.lde(x, lg_n, rate, cols), lg_n + rate, cols)
        del x
        d, c = ref_merkle(ref, leaves, cols, h, ht)
        out["commits"].append({"cols": cols, "lg_n": lg_n, "rate": rate, "cap_height": h, "hash": ht, "seed": seed,
                               "leaves_sha256": sha(leaves), "digests_sha256": sha(d),
                               "caps": [str(int(v)) for v in c]})
        del leaves, d
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "fixtures.json"))


if __name__ == "__main__":
    main()
