"""The commit server's placement of requests in its device region
(zeknox_amd/csrc/placement.h), checked on the CPU against a brute-force
model of the region: first fit from the bottom under a limit, highest fit
from the top, coalescing on return."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "zeknox_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("placement") / "placement_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "placement_check.cpp"), "-o", exe],
                   check=True)
    return exe


@pytest.mark.parametrize("units,ops,seed", [(64, 100000, 1), (512, 60000, 2), (4096, 6000, 3)])
def test_placement_matches_model(checker, units, ops, seed):
    r = subprocess.run([checker, str(units), str(ops), str(seed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == f"ok {ops}", r.stdout + r.stderr
