"""A plain C program (tests/c_consumer.c) built with gcc against
include/zeknox.h and linked to libzeknox.so, calling the reference's C ABI
the way its Go wrapper does (wrappers/go/lib/lib.go:65-103, 198-210):
twiddles, coset, host-buffer LDE, transpose, Merkle caps, and the RustError
convention for bad arguments.  The CPU test builds and links it; the GPU test
runs it and checks its leaf rows and caps against the oracle."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "zeknox_amd")

COLS, LG_N, RATE, CAP_H = 7, 9, 2, 2  # as #defined in c_consumer.c


@pytest.fixture(scope="module")
def consumer(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libzeknox.so")):
        pytest.fail("libzeknox.so is not built (run __graft_entry__.build())")
    exe = str(tmp_path_factory.mktemp("c_consumer") / "c_consumer")
    subprocess.run(["gcc", "-std=c99", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "c_consumer.c"), "-L", LIBDIR, "-lzeknox", "-Wl,-rpath," + LIBDIR,
                    "-o", exe], check=True)
    return exe


def test_c_consumer_links(consumer):
    out = subprocess.run(["ldd", consumer], capture_output=True, text=True, check=True).stdout
    assert os.path.join(LIBDIR, "libzeknox.so") in out, out


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [5, 6])
def test_c_consumer_matches_oracle(consumer, seed):
    import oracle_bind as ob
    r = subprocess.run([consumer, str(seed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    lines = r.stdout.split("\n")
    x = ob.canonical(COLS << LG_N, seed)
    leaves = ob.transpose_rev(ob.lde(x, LG_N, RATE, COLS), LG_N + RATE, COLS)
    weights = np.arange(1, leaves.size + 1, dtype=np.uint64)
    checksum = int(np.sum(leaves * weights, dtype=np.uint64))
    _, caps = ob.merkle(leaves, COLS, CAP_H, 0)
    want = [f"leaves {checksum}"] + [f"cap {i} " + " ".join(str(int(v)) for v in caps[4 * i:4 * i + 4])
                                    for i in range(1 << CAP_H)]
    assert lines[:len(want)] == want
