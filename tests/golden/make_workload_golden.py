#!/usr/bin/env python3
"""Write tests/golden/workloads.json: expected whole-commit outputs of the
benchmark workloads that the GPU tests replay.

  batch64 -- the mixed batch (zeknox_amd.workloads.mixed_batch_jobs, numpy
             seed 3): 64 Poseidon jobs, log2 degree U[10,16], columns from
             {2,20,80,135}, rate_bits {1,2,3}, cap height 4; job j's
             coefficients are the fill stream of 3000 + j;
  served  -- a Poisson trace of the served stream
             (zeknox_amd.workloads.served_trace, seed 7, 80 requests/s for
             4 s): log2 degree lognormal clipped to [12,20], columns U[2,400],
             rate_bits {1,2,3}, Poseidon / Poseidon2, cap height 4; every
             request's coefficients are a prefix of one buffer filled with
             seed 7.  Requests of at most 5e6 permutations get expected
             outputs; the larger ones are checked on the GPU against the same
             request run alone.

Each commit: the oracle's coset LDE (shift 7) and leaf rows (oracle/oracle.c
orc_lde + orc_transpose_rev), then the reference's own CPU Merkle forest over
those rows (native/merkle/merkle.cpp, built by oracle/Makefile into
oracle/_ref).  Stored as SHA-256 of the little-endian leaf rows, digests and
caps, plus the first cap in full.  Run once in the build container (about
ten minutes of CPU):

    make -C oracle ref && python3 tests/golden/make_workload_golden.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle_bind as ob  # noqa: E402
from zeknox_amd.workloads import CAP_HEIGHT, mixed_batch_jobs, permutations, served_trace  # noqa: E402

SERVED_SEED, SERVED_SECONDS, SERVED_RATE = 7, 4.0, 80.0
ORACLE_PERMS_MAX = 5_000_000


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def commit(ref, coeffs, cols, lg_n, rate, ht):
    leaves = ob.transpose_rev(ob.lde(coeffs, lg_n, rate, cols), lg_n + rate, cols)
    nl = 1 << (lg_n + rate)
    nc = 1 << CAP_HEIGHT
    nd = 2 * (nl - nc)
    d = np.zeros(nd * 4, np.uint64)
    c = np.zeros(nc * 4, np.uint64)
    ref.fill_digests_buf_linear_cpu(ob._ptr(d), ob._ptr(c), ob._ptr(leaves), nd, nc, nl, cols, CAP_HEIGHT, ht)
    return {"leaves_sha256": sha(leaves), "digests_sha256": sha(d), "caps_sha256": sha(c),
            "cap0": [str(int(v)) for v in c[:4]]}


def main():
    ref = ob.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libzkref.so is not built (make -C oracle ref)")
    out = {"source": "oracle orc_lde + orc_transpose_rev (coset shift 7), then reference native/merkle/merkle.cpp",
           "batch64": [], "served": {"seed": SERVED_SEED, "seconds": SERVED_SECONDS, "rate_hz": SERVED_RATE,
                                     "oracle_perms_max": ORACLE_PERMS_MAX, "requests": []}}
    t0 = time.time()
    for j, (cols, lg_n, rate, ht, seed) in enumerate(mixed_batch_jobs()):
        rec = {"cols": cols, "lg_n": lg_n, "rate": rate, "hash": ht, "seed": seed}
        rec.update(commit(ref, ob.salt(cols << lg_n, seed), cols, lg_n, rate, ht))
        out["batch64"].append(rec)
        print(f"batch64 {j}: {cols} x 2^{lg_n} rate {rate} ({time.time() - t0:.0f} s)", flush=True)
    trace = served_trace(SERVED_SEED, SERVED_SECONDS, SERVED_RATE)
    pool = ob.salt(max(c << l for _, c, l, _, _ in trace), SERVED_SEED)
    for i, (ta, cols, lg_n, rate, ht) in enumerate(trace):
        rec = {"i": i, "cols": cols, "lg_n": lg_n, "rate": rate, "hash": ht}
        if permutations(cols, 1 << (lg_n + rate), 1 << CAP_HEIGHT) <= ORACLE_PERMS_MAX:
            rec.update(commit(ref, pool[: cols << lg_n], cols, lg_n, rate, ht))
        out["served"]["requests"].append(rec)
        if i % 20 == 0:
            print(f"served {i}/{len(trace)} ({time.time() - t0:.0f} s)", flush=True)
    with open(os.path.join(HERE, "workloads.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", os.path.join(HERE, "workloads.json"), f"({time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()
