#!/usr/bin/env python3
"""LDE + Merkle commit benchmark on one MI355X.

Workload (one "step" = one commit): 135 polynomials of degree 2^15 with
coefficients uniform in [0, p), rate_bits 3 -> 2^18-point coset LDE (shift
7), bit-reversed rows transposed to 2^18 leaves of 135 elements, Poseidon
Merkle tree with cap height 4.  Inputs are synthetic (seeded splitmix64,
rejection-sampled into [0, p)) and resident in HBM before timing starts.

Two paths are timed:
  plain   -- the C-ABI sequence compute_batched_lde -> compute_transpose_rev
             -> fill_digests_buf_linear_gpu_with_gpu_ptr, one commit per call;
  batched -- the batched commit engine (zeknox_amd.batch), commits of the
             same shape in flight together.
The headline value is the batched rate; both, the per-kernel HIP-event
timings, the roofline of the dominant kernel, a 64-job mixed batch and a
large Poseidon2 batch both ways, a served stream, and a CPU baseline (the
same commit on the host cores: the reference's own CPU Merkle and the
oracle's LDE) go into one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--serve-seconds S]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# the headline commit: 135 polys x 2^15, rate_bits 3, cap height 4
LG_N, RATE, COLS, CAP_H = 15, 3, 135, 4
LG = LG_N + RATE
N = 1 << LG
HASH = 0  # Poseidon

# MI355X published peaks
HBM_PEAK_GBS = 8000.0

# Per-unit figures of the leaf-hash launch on this workload, measured with
# rocprofv3 --pmc on the batched 135 x 2^15 rate-3 launch (4 commits) and
# kept in profiles/merkle_leaves_valu_135x2e18.txt and
# profiles/hbm_traffic_135x2e15_rate3.txt:
#   VALU and MFMA wave-instructions per Poseidon permutation (SQ_INSTS_VALU,
#   SQ_INSTS_MFMA over permutations; leaf loads and in-tile levels included),
#   by Merkle work-group size (log2 lanes).  The leaf permutations run their
#   MDS layers on the matrix cores (v_mfma_i32_32x32x32_i8).
VALU_WAVE_INSTS_PER_PERM = {8: 236.8, 9: 231.9, 10: 229.6}
MFMA_WAVE_INSTS_PER_PERM = {8: 7.97, 9: 7.97, 10: 7.97}
#   issue slots (one per 4 cycles) that one such MFMA takes from the vector
#   issue of its SIMD when it runs beside full-rate VALU work (13 cycles,
#   measured with a microbenchmark of independent chains)
MFMA_ISSUE_SLOTS = 3.25
#   HBM bytes per commit: FETCH_SIZE x 2 (the counter reads half the bytes of
#   a plain 1 GiB copy) + WRITE_SIZE x 1, KiB x 1024, by work-group size
LEAF_TRAFFIC_BYTES_PER_COMMIT = {9: 8.25e8}
#   NTT passes (both), same scaling
NTT_TRAFFIC_BYTES_PER_COMMIT = 8.87e8
# 32-bit VALU issue peak: 256 CUs x 4 SIMDs x one wave64 instruction per 4
# cycles x 2.4 GHz = 39.3 T lane-ops/s (MI355X's published 78.6 TFLOP/s
# vector FP64 counts an FMA as 2 flops on the same issue rate)
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--jobs", type=int, default=4, help="135 x 2^15 rate-3 commits per batched step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="served-stream duration (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=40.0, help="served-stream arrivals per second")
    ap.add_argument("--serve-trace", default=None,
                    help="write one CSV line per served request (times in ms from the first arrival)")
    ap.add_argument("--no-tune", action="store_true", help="batched path on the default schedule (no autotuning)")
    ap.add_argument("--schedule", default=None, metavar="S,E,T",
                    help="batched 135 x 2^15 schedule (ntt_strided_max,ntt_even,merkle_tile_log) instead of autotuning")
    ap.add_argument("--no-extra", action="store_true", help="skip the 64-job and 8 x 400 x 2^20 batches")
    ap.add_argument("--legs", default="batch64,large,hashers,ntt",
                    help="extra legs to run: batch64, large, hashers, ntt (comma-separated)")
    ap.add_argument("--extra-steps", type=int, default=3, help="timed steps of those batches")
    ap.add_argument("--cpu-runs", type=int, default=5, help="whole CPU commits timed (median reported)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain elementwise copy of MIB MiB three times (scale for HBM counters)")
    return ap.parse_args()


def cpu_baseline(np, coeffs, runs, gpu_caps):
    """The same 135 x 2^15 rate-3 commit on the host cores, `runs` times, median
    reported.  Merkle: the reference's own CPU fill_digests_buf_linear_cpu
    (native/merkle/merkle.cpp, OpenMP, compiled from its sources into
    oracle/_ref) when that build is present, else the oracle's restatement of
    it; LDE + transpose: the oracle (the reference has no CPU LDE).  The CPU
    caps are compared with the GPU's."""
    import statistics
    import oracle_bind as ob
    ref = ob.ref_lib()
    nc = 1 << CAP_H
    nd = 2 * (N - nc)
    lde_t, mer_t, same = [], [], True
    for _ in range(runs):
        t0 = time.perf_counter()
        lde = ob.lde(coeffs, LG_N, RATE, COLS)
        leaves = ob.transpose_rev(lde, LG, COLS)
        del lde
        t1 = time.perf_counter()
        if ref is not None:
            dig = np.zeros(nd * 4, np.uint64)
            cap = np.zeros(nc * 4, np.uint64)
            ref.fill_digests_buf_linear_cpu(ob._ptr(dig), ob._ptr(cap), ob._ptr(leaves), nd, nc, N, COLS, CAP_H, HASH)
        else:
            dig, cap = ob.merkle(leaves, COLS, CAP_H, HASH)
        t2 = time.perf_counter()
        lde_t.append(t1 - t0)
        mer_t.append(t2 - t1)
        same = same and bool(np.array_equal(cap, gpu_caps))
        del leaves, dig
    tot = [a + b for a, b in zip(lde_t, mer_t)]
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(1.0 / statistics.median(tot), 4), "unit": "commits/s",
            "cores": ob.orc.orc_num_threads(), "kind": "reference" if ref is not None else "port",
            "sample": f"{runs} whole 135 x 2^15 rate-3 commits of the GPU run's input, median: LDE+transpose "
                      f"{statistics.median(lde_t):.2f} s (oracle C, OpenMP), Poseidon Merkle "
                      f"{statistics.median(mer_t):.2f} s ("
                      + ("reference merkle.cpp compiled from its sources" if ref is not None else "oracle C")
                      + ")",
            "caps_equal_gpu": same, "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "cpu_model": cpu_model, "runs_s": [round(t, 3) for t in tot]}


def _timed(zk, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    zk.lib.zkx_device_synchronize(0)
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    zk.lib.zkx_device_synchronize(0)
    return (time.perf_counter() - t0) / steps


def plain_vs_batched(zk, np, shapes, steps, warmup, seed, tune):
    """Commit jobs of the given (cols, lg_n, rate, hash) shapes, cap height 4,
    coefficients generated on the device: the plain C-ABI sequence one job
    per call (shared LDE / leaf buffers) against one zkx_commit_batch of all
    of them (every job's inputs, LDE and outputs resident at once).  Returns
    commits/s both ways and whether every digest and cap agrees."""
    commits = []
    for j, (cols, lg_n, rate, ht) in enumerate(shapes):
        d = zk.DeviceBuffer(cols << lg_n)
        d.fill_random(seed * 1000 + j)
        commits.append(zk.Commit(d, cols, lg_n, rate, CAP_H, ht))
    big = max(c.cols * c.N for c in commits)
    d_lde, d_leaves = zk.DeviceBuffer(big), zk.DeviceBuffer(big)
    p_dig = [zk.DeviceBuffer(max(c.num_digests, 1) * 4) for c in commits]
    p_cap = [zk.DeviceBuffer(c.num_caps * 4) for c in commits]

    def plain():
        for c, dg, cp in zip(commits, p_dig, p_cap):
            lg = c.lg_n + c.rate
            zk.lde_batch(0, d_lde, c.d_coeffs, c.lg_n,
                         zk.NTTConfig(batches=c.cols, extension_rate_bits=c.rate, with_coset=True,
                                      are_inputs_on_device=True, are_outputs_on_device=True))
            zk.transpose_rev_batch(0, d_leaves, d_lde, lg, zk.TransposeConfig(batches=c.cols,
                                   are_inputs_on_device=True, are_outputs_on_device=True))
            zk.fill_digests_buf_linear_gpu_with_gpu_ptr(dg, cp, d_leaves, c.num_digests, c.num_caps, c.N, c.cols,
                                                       CAP_H, c.hash_type, 0)

    plain_s = _timed(zk, plain, steps, warmup)
    tune_s = 0.0
    if tune:  # one-time schedule search per distinct shape, outside the timing
        t0 = time.perf_counter()
        zk.autotune(commits)
        tune_s = time.perf_counter() - t0
    batch_s = _timed(zk, lambda: zk.commit_batch(commits), steps, warmup)
    same = all(np.array_equal(cp.to_host(), c.d_caps.to_host())
               and np.array_equal(dg.to_host(c.num_digests * 4), c.d_digests.to_host(c.num_digests * 4))
               for c, dg, cp in zip(commits, p_dig, p_cap))
    n = len(commits)
    schedules = sorted({zk.get_tuning(c).astuple() for c in commits})
    for b in [d_lde, d_leaves] + p_dig + p_cap:
        b.free()
    for c in commits:
        for b in (c.d_coeffs, c.d_leaves, c.d_digests, c.d_caps):
            if b is not None:
                b.free()
    zk.lib.zkx_commit_trim(0)
    return {"jobs": n, "plain_commits_per_s": round(n / plain_s, 2), "batched_commits_per_s": round(n / batch_s, 2),
            "batched_over_plain": round(plain_s / batch_s, 3), "batched_equals_plain": bool(same),
            "ms_per_batch": round(batch_s * 1e3, 2), "autotune_s": round(tune_s, 2),
            "schedules (ntt_strided_max, ntt_even, merkle_tile_log)": [list(t) for t in schedules]}


def merkle_hashers(zk, np, steps, lg=20, cols=135, h=4):
    """The Merkle forest alone through fill_digests_buf_linear_gpu_with_gpu_ptr
    (device pointers, row-major leaves resident) for every hash type: 2^lg
    leaves of `cols` elements, cap height h.  Per hash: ms per tree and
    leaves/s; the Goldilocks sponges (Poseidon, Poseidon2, Monolith; rate 8)
    also permutations/s (ceil(cols/8) per leaf + one per internal node)."""
    L = 1 << lg
    leaves = zk.DeviceBuffer(cols * L)
    leaves.fill_random(77)
    nc, nd = 1 << h, 2 * (L - (1 << h))
    dig, cap = zk.DeviceBuffer(nd * 4), zk.DeviceBuffer(nc * 4)
    perms = L * ((cols + 7) // 8) + (L - nc)
    names = {zk.HASH_POSEIDON: "poseidon", zk.HASH_KECCAK: "keccak256", zk.HASH_POSEIDON_BN128: "poseidon_bn128",
             zk.HASH_POSEIDON2: "poseidon2", zk.HASH_MONOLITH: "monolith"}
    out = {"workload": f"Merkle forest only: 2^{lg} leaves x {cols}, cap_height {h}, device pointers"}
    for ht, name in names.items():
        def tree():
            zk.fill_digests_buf_linear_gpu_with_gpu_ptr(dig, cap, leaves, nd, nc, L, cols, h, ht, 0)
        t = _timed(zk, tree, steps, 1)
        rec = {"ms_per_tree": round(t * 1e3, 3), "leaves_per_s": round(L / t)}
        if ht in (zk.HASH_POSEIDON, zk.HASH_POSEIDON2, zk.HASH_MONOLITH):
            rec["perms_per_s"] = round(perms / t)
        out[name] = rec
    return out


def ntt_legs(zk, np, steps, lg=20, cols=64):
    """compute_batched_ntt forward and inverse (natural order in and out) of
    `cols` columns of 2^lg on device buffers, and a salted coset LDE (rate 1,
    4 salt columns) of the same columns at half the length: ms per call and
    the algorithmic rate (each column read and written once)."""
    n = cols << lg
    buf = zk.DeviceBuffer(n)
    buf.fill_random(91)
    dev = dict(are_inputs_on_device=True, are_outputs_on_device=True)
    fwd = _timed(zk, lambda: zk.ntt_batch(0, buf, lg, zk.NTTConfig(batches=cols, **dev)), steps, 1)
    inv = _timed(zk, lambda: zk.intt_batch(0, buf, lg, zk.NTTConfig(batches=cols, **dev)), steps, 1)
    half = zk.DeviceBuffer(cols << (lg - 1))
    half.fill_random(92)
    salt = 4
    lde_out = zk.DeviceBuffer((cols + salt) << lg)
    lde = _timed(zk, lambda: zk.lde_batch(0, lde_out, half, lg - 1, zk.NTTConfig(
        batches=cols, extension_rate_bits=1, with_coset=True, salt_size=salt, **dev)), steps, 1)
    gbs = lambda nbytes, t: round(nbytes / t / 1e9, 1)
    return {"workload": f"{cols} columns x 2^{lg}: NTT forward and inverse (natural order); coset LDE "
                        f"2^{lg - 1} -> 2^{lg}, rate 1, {salt} salt columns",
            "ntt_forward_ms": round(fwd * 1e3, 3), "ntt_forward_gbs": gbs(16 * n, fwd),
            "ntt_inverse_ms": round(inv * 1e3, 3), "ntt_inverse_gbs": gbs(16 * n, inv),
            "lde_salted_ms": round(lde * 1e3, 3), "lde_salted_gbs": gbs(8 * (n // 2) + 8 * ((cols + salt) << lg), lde)}


def batch64_mixed(zk, np, steps, warmup, tune, seed=3):
    rng = np.random.default_rng(seed)
    shapes = [(int(rng.choice([2, 20, 80, 135])), int(rng.integers(10, 17)), int(rng.integers(1, 4)), 0)
              for _ in range(64)]
    r = plain_vs_batched(zk, np, shapes, steps, warmup, seed, tune)
    r["workload"] = ("64 jobs: log2 degree U[10,16], columns from {2,20,80,135}, rate_bits {1,2,3}, "
                     f"Poseidon, cap_height 4 (numpy seed {seed})")
    return r


def large_poseidon2_batch(zk, np, steps, warmup, tune):
    r = plain_vs_batched(zk, np, [(400, 20, 1, 3)] * 8, steps, warmup, 4, tune)
    r["workload"] = "8 jobs of 400 polys x 2^20, rate_bits 1 (2^21-point LDE), Poseidon2, cap_height 4, all resident"
    return r


def calibrate_copy(mib):
    """A plain device copy of known size: three launches of an elementwise
    int64 kernel reading `mib` MiB and writing `mib` MiB.  Under
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE its per-dispatch readings divided
    by the known bytes give the counters' scale on this GPU."""
    import torch
    n = (mib << 20) // 8
    x = torch.arange(n, dtype=torch.int64, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}))


def perms_per_commit(rows, cols, caps):
    # leaf sponge: ceil(cols/8) permutations per leaf (cols > 4), plus one per
    # internal node and cap: rows - caps
    per_leaf = (cols + 7) // 8 if cols > 4 else 0
    return rows * per_leaf + (rows - caps)


def served_stream(zk, np, seconds, rate_hz, budget, streams, seed=5, trace_path=None):
    """Poisson arrivals of heterogeneous commit requests into the commit server
    for `seconds`: log2 degree lognormal clipped to [12, 20], columns uniform
    in [2, 400], rate_bits in {1, 2, 3}, Poseidon / Poseidon2 50/50, cap
    height 4, leaf rows kept.  Returns throughput and latency percentiles
    (latency = completion - submission, steady clock inside the server)."""
    import threading
    rng = np.random.default_rng(seed)
    # one device pool of coefficients serves every request (its first cols*2^lg_n words)
    pool = zk.DeviceBuffer(400 << 20)
    pool.fill_random(seed)
    srv = zk.Server(budget, streams=streams, max_batch=64)
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t >= seconds:
            break
        lg_n = int(np.clip(np.rint(rng.lognormal(np.log(14.0), 0.15)), 12, 20))
        cols = int(rng.integers(2, 401))
        rate = int(rng.integers(1, 4))
        ht = 0 if rng.random() < 0.5 else 3
        arrivals.append((t, cols, lg_n, rate, ht))
    ids, shape_of = [], {}
    done = threading.Event()
    lat, qwait, large, t_last, rows = [], [], [], [None], []
    large_work = 1 << 28  # the server's small/large boundary, LDE elements

    def consumer():
        # results are taken in completion order, as a client would
        got = 0
        while got < len(arrivals):
            while got >= len(ids):
                time.sleep(0.0005)
            rid, r = srv.wait_any()
            lat.append(r.t_done - r.t_submit)
            qwait.append(r.t_start - r.t_submit)
            while rid not in shape_of:  # the submitting thread records it after submit returns
                time.sleep(0.0001)
            cols, lg_n, rate = shape_of[rid]
            large.append((cols << (lg_n + rate)) > large_work)
            rows.append((rid, r.t_submit, r.t_start, r.t_enqueued, r.t_done, cols, lg_n, rate))
            t_last[0] = r.t_done
            srv.release(rid)
            got += 1
        done.set()

    th = threading.Thread(target=consumer)
    th.start()
    t0 = time.perf_counter()
    for (ta, cols, lg_n, rate, ht) in arrivals:
        delay = t0 + ta - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        rid = srv.submit(pool, cols, lg_n, rate, 4, ht)
        shape_of[rid] = (cols, lg_n, rate)
        ids.append(rid)
    done.wait()
    th.join()
    elapsed = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat_ms, q_ms, large = np.array(lat) * 1e3, np.array(qwait) * 1e3, np.array(large)
    if trace_path:
        t_first = min(row[1] for row in rows)
        with open(trace_path, "w") as f:
            f.write("id,submit_ms,start_ms,enqueued_ms,done_ms,cols,lg_n,rate\n")
            for rid, ts, tst, te, td, cols, lg_n, rate in sorted(rows):
                f.write(f"{rid},{(ts - t_first) * 1e3:.3f},{(tst - t_first) * 1e3:.3f},{(te - t_first) * 1e3:.3f},"
                        f"{(td - t_first) * 1e3:.3f},{cols},{lg_n},{rate}\n")

    def pct(a, q):
        return round(float(np.percentile(a, q)), 2) if a.size else None
    return {"workload": f"Poisson {rate_hz}/s for {seconds} s; log2 degree lognormal(ln 14, 0.15) clipped [12,20], "
                        "cols U[2,400], rate_bits {1,2,3}, Poseidon/Poseidon2 50/50, cap_height 4",
            "requests": len(arrivals), "commits_per_s": round(len(arrivals) / elapsed, 2),
            "p50_ms": pct(lat_ms, 50), "p99_ms": pct(lat_ms, 99), "mean_ms": round(float(lat_ms.mean()), 2),
            "queue_p50_ms": pct(q_ms, 50), "queue_p99_ms": pct(q_ms, 99),
            "small_p50_ms": pct(lat_ms[~large], 50), "small_p99_ms": pct(lat_ms[~large], 99),
            "large_requests": int(large.sum()), "large_p50_ms": pct(lat_ms[large], 50),
            "large_p99_ms": pct(lat_ms[large], 99), "batches": st["batches"], "max_batch": st["max_batch"],
            "peak_hbm_gb": round(st["peak_bytes"] / 2**30, 2), "budget_gb": round(budget / 2**30, 1),
            "streams": streams}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build drives one GPU; --gpus {args.gpus} is not supported "
              "(commit jobs are independent and would shard across GPUs one process per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    import numpy as np
    import zeknox_amd as zk
    sys.path.insert(0, os.path.join(ROOT, "tests"))

    ngpu = zk.get_number_of_gpus_rs()
    assert ngpu >= 1, "no HIP device"
    zk.init_coset_rs(0, max(LG, 21), 7)  # 2^21 points: the 400 x 2^20 batch
    zk.init_twiddle_factors_rs(0, LG)

    # synthetic inputs, resident in HBM
    rng = np.random.default_rng(2024)
    coeffs = rng.integers(0, 2**63, size=COLS << LG_N, dtype=np.uint64) * np.uint64(2)
    coeffs %= np.uint64(zk.P)
    d_in = zk.DeviceBuffer.from_host(coeffs)
    d_lde = zk.DeviceBuffer(COLS * N)
    d_leaves = zk.DeviceBuffer(COLS * N)
    nc = 1 << CAP_H
    nd = 2 * (N - nc)
    d_dig = zk.DeviceBuffer(nd * 4)
    d_cap = zk.DeviceBuffer(nc * 4)
    cfg = zk.NTTConfig(batches=COLS, extension_rate_bits=RATE, with_coset=True, are_inputs_on_device=True,
                       are_outputs_on_device=True)
    tcfg = zk.TransposeConfig(batches=COLS, are_inputs_on_device=True, are_outputs_on_device=True)

    def plain_commit():
        zk.lde_batch(0, d_lde, d_in, LG_N, cfg)
        zk.transpose_rev_batch(0, d_leaves, d_lde, LG, tcfg)
        zk.fill_digests_buf_linear_gpu_with_gpu_ptr(d_dig, d_cap, d_leaves, nd, nc, N, COLS, CAP_H, HASH, 0)

    # ---- plain path
    for _ in range(args.warmup):
        plain_commit()
    zk.lib.zkx_device_synchronize(0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plain_commit()
    zk.lib.zkx_device_synchronize(0)
    plain_s = (time.perf_counter() - t0) / args.steps

    # per-kernel HIP-event timings of the same sequence (separate pass)
    zk.kernel_stats_reset()
    zk.profile_enable(True)
    for _ in range(args.steps):
        plain_commit()
    stats = zk.kernel_stats()
    zk.profile_enable(False)
    kernels = {k: {"launches_per_commit": n / args.steps, "ms_per_commit": ms / args.steps,
                   "avg_launch_ms": ms / n} for k, (n, ms) in stats.items()}

    # ---- batched path: JOBS such commits per zkx_commit_batch call
    commits = [zk.Commit(d_in, COLS, LG_N, RATE, CAP_H, HASH) for _ in range(args.jobs)]
    if args.schedule:
        zk.set_tuning(zk.Tuning(*[int(v) for v in args.schedule.split(",")]), commits[0])
    elif not args.no_tune:  # one-time schedule search for the shape, outside the timing
        zk.autotune(commits[:1])
    schedule = dict(zip(("ntt_strided_max", "ntt_even", "merkle_tile_log"), zk.get_tuning(commits[0]).astuple()))
    for _ in range(args.warmup):
        zk.commit_batch(commits)
    zk.lib.zkx_device_synchronize(0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        zk.commit_batch(commits)  # returns when the batch is complete
    zk.lib.zkx_device_synchronize(0)
    batch_step_s = (time.perf_counter() - t0) / args.steps
    batched_s = batch_step_s / args.jobs  # seconds per commit
    zk.kernel_stats_reset()
    zk.profile_enable(True)
    for _ in range(args.steps):
        zk.commit_batch(commits)
    bstats = zk.kernel_stats()
    zk.profile_enable(False)
    batch_kernels = {k: {"launches_per_step": n / args.steps, "ms_per_step": ms / args.steps, "avg_launch_ms": ms / n}
                     for k, (n, ms) in bstats.items()}
    # the batched outputs equal the plain path's (bit-exact) on this input
    pl = d_leaves.to_host(COLS * 64), d_dig.to_host(), d_cap.to_host()
    bl = commits[-1].results()
    bit_exact = bool(np.array_equal(pl[0], bl[0][:COLS * 64]) and np.array_equal(pl[1], bl[1])
                     and np.array_equal(pl[2], bl[2]))
    mode = "batched"

    # ---- roofline of the dominant kernel: the leaf-hash launch of the batch
    # (Poseidon permutations, issue-bound: 64-bit modular arithmetic on the
    # 32-bit integer VALU, the MDS layers as i8 MFMAs that share the SIMD's
    # issue).  achieved = permutations per launch x issue slots per
    # permutation (VALU wave-instructions + MFMA_ISSUE_SLOTS per MFMA) x 64
    # lanes / the launch's HIP-event duration (events on the library's
    # stream); peak = VALU_PEAK_TOPS (one wave-instruction per 4 cycles).
    dom = max(batch_kernels, key=lambda k: batch_kernels[k]["ms_per_step"])
    leaf = batch_kernels.get("merkle_leaves")
    roofline = None
    if leaf:
        tl = schedule["merkle_tile_log"]
        perms = args.jobs * perms_per_commit(N, COLS, nc)
        slots = VALU_WAVE_INSTS_PER_PERM[tl] + MFMA_ISSUE_SLOTS * MFMA_WAVE_INSTS_PER_PERM[tl]
        achieved = perms * slots * 64 / (leaf["avg_launch_ms"] * 1e-3) / 1e12
        alg_bytes = args.jobs * (8 * COLS * N * 2 + 32 * nd)  # LDE read, leaf rows + digests written
        traffic = LEAF_TRAFFIC_BYTES_PER_COMMIT.get(tl)
        roofline = {"bound": "valu+mfma issue", "achieved": round(achieved, 2), "peak": round(VALU_PEAK_TOPS, 2),
                    "unit": "T lane-issue-slots/s (32-bit VALU)", "frac": round(achieved / VALU_PEAK_TOPS, 3),
                    "traffic": round(traffic * args.jobs) if traffic else None, "algorithmic_bytes": alg_bytes,
                    "hbm_gbs": round(traffic * args.jobs / (leaf["avg_launch_ms"] * 1e-3) / 1e9, 1) if traffic else None,
                    "kernel": "merkle_leaves", "avg_launch_ms": round(leaf["avg_launch_ms"], 4),
                    "perms_per_launch": perms, "valu_wave_insts_per_perm": VALU_WAVE_INSTS_PER_PERM[tl],
                    "mfma_wave_insts_per_perm": MFMA_WAVE_INSTS_PER_PERM[tl], "mfma_issue_slots": MFMA_ISSUE_SLOTS}

    # LDE passes against HBM: algorithmic bytes of the LDE (read n, write N
    # per column) over the batch's NTT launches
    ntt = batch_kernels.get("batch_ntt")
    ntt_roofline = None
    if ntt:
        lde_bytes = 8 * COLS * ((1 << LG_N) + N) * args.jobs
        gbs = lde_bytes / (ntt["ms_per_step"] * 1e-3) / 1e9
        ntt_roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 3), "traffic": round(NTT_TRAFFIC_BYTES_PER_COMMIT * args.jobs),
                        "traffic_gbs": round(NTT_TRAFFIC_BYTES_PER_COMMIT * args.jobs / (ntt["ms_per_step"] * 1e-3) / 1e9, 1),
                        "kernel": "batch_ntt (all passes)",
                        "note": "VALU-bound (64-bit modular butterflies), not HBM-bound"}

    # ---- heterogeneous 64-job batch and the large Poseidon2 batch, both ways
    extra = {}
    if not args.no_extra:
        legs = set(args.legs.split(","))
        if "batch64" in legs:
            extra["batch64"] = batch64_mixed(zk, np, args.extra_steps, 1, not args.no_tune)
        if "large" in legs:
            extra["large_poseidon2"] = large_poseidon2_batch(zk, np, args.extra_steps, 1, not args.no_tune)
        if "hashers" in legs:
            extra["merkle_hashers"] = merkle_hashers(zk, np, args.extra_steps)
        if "ntt" in legs:
            extra["ntt"] = ntt_legs(zk, np, args.extra_steps)

    # ---- CPU baseline: whole commits of the same input on the host cores
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(np, coeffs, args.cpu_runs, d_cap.to_host())

    served = None
    if args.serve_seconds > 0:
        served = served_stream(zk, np, args.serve_seconds, args.serve_rate, 64 << 30, 3,
                               trace_path=args.serve_trace)

    value = args.jobs / batch_step_s
    out = {
        "metric": "LDE+Merkle commits/s batched vs plain at bit-exact; p50/p99 commit latency",
        "value": round(value, 2),
        "unit": "commits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batch_step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks field)",
        "data": "synthetic (seeded uniform canonical field elements)",
        "config": {"workload": "LDE+commit: 135 polys x 2^15, rate_bits 3 (2^18-point coset LDE, shift 7), "
                               "transpose to 2^18 leaves x 135, Poseidon Merkle, cap_height 4",
                   "path": mode, "commits_per_step": args.jobs, "parallelism": "single GPU",
                   "schedule": schedule},
        "plain_commits_per_s": round(1.0 / plain_s, 2),
        "batched_over_plain": round(plain_s / batched_s, 3),
        "batched_equals_plain": bit_exact,
        "kernels_plain": kernels,
        "kernels_batched": batch_kernels,
        "dominant_kernel": dom,
        "roofline": roofline,
        "ntt_roofline": ntt_roofline,
        "cpu_baseline": cpu,
        "served": served,
        **extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
