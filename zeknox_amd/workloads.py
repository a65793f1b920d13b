"""Seeded synthetic workloads of the commit benchmarks, shared by bench.py,
the tests and the fixture generator so that all three draw the same jobs.

Coefficients come from the device stream of zkx_fill_random (element i of a
buffer filled with seed s is the first value below p of the splitmix64
sequence started at s ^ (i * 0xD1B54A32D192ED03)), so any prefix of a buffer
is the same whatever its length.
"""
import numpy as np

CAP_HEIGHT = 4


def mixed_batch_jobs(seed=3, njobs=64):
    """The mixed batch: `njobs` Poseidon commit jobs, log2 degree uniform in
    [10, 16], columns drawn from {2, 20, 80, 135}, rate_bits in {1, 2, 3},
    cap height 4.  Returns (cols, lg_n, rate, hash_type, coeff_seed) per job;
    job j's coefficients are the fill stream of seed * 1000 + j."""
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(njobs):
        cols = int(rng.choice([2, 20, 80, 135]))
        lg_n = int(rng.integers(10, 17))
        rate = int(rng.integers(1, 4))
        jobs.append((cols, lg_n, rate, 0, seed * 1000 + j))
    return jobs


def served_trace(seed, seconds, rate_hz):
    """Poisson arrivals at `rate_hz` for `seconds`: log2 degree lognormal
    (median 14, sigma 0.15 in log space) rounded and clipped to [12, 20],
    columns uniform in [2, 400], rate_bits uniform in {1, 2, 3}, Poseidon or
    Poseidon2 with probability 1/2 each, cap height 4.  Returns a list of
    (arrival_s, cols, lg_n, rate, hash_type); every request's coefficients
    are the first cols * 2^lg_n words of one buffer filled with `seed`."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t >= seconds:
            break
        lg_n = int(np.clip(np.rint(rng.lognormal(np.log(14.0), 0.15)), 12, 20))
        cols = int(rng.integers(2, 401))
        rate = int(rng.integers(1, 4))
        ht = 0 if rng.random() < 0.5 else 3
        out.append((t, cols, lg_n, rate, ht))
    return out


def permutations(cols, rows, caps):
    """Sponge permutations of one Goldilocks Merkle forest (rate 8): ceil(cols/8)
    per leaf when cols > 4, plus one per internal node and cap."""
    per_leaf = (cols + 7) // 8 if cols > 4 else 0
    return rows * per_leaf + (rows - caps)
