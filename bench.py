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
# profiles/hbm_traffic_135x2e15_rate3.txt.  A batch of that size builds its
# tree levels in launches of their own (zkx_set_level_launches), so the leaf
# launch hashes the leaves only:
#   VALU and MFMA wave-instructions per leaf permutation (SQ_INSTS_VALU,
#   SQ_INSTS_MFMA over the launch's 17 permutations per leaf; leaf loads and
#   leaf-row writes included), by Merkle work-group size (log2 lanes).  The
#   permutations run their MDS layers on the matrix cores
#   (v_mfma_i32_32x32x32_i8).
#   keyed by (merkle_tile_log, merkle_valu)
VALU_WAVE_INSTS_PER_PERM = {(8, 0): 205.49, (9, 0): 205.43, (10, 0): 205.43, (9, 1): 288.12}
MFMA_WAVE_INSTS_PER_PERM = {(8, 0): 8.437, (9, 0): 8.437, (10, 0): 8.437, (9, 1): 0.0}
#   SHA-256 (first 16 hex digits) of the leaf kernel's sources those counts
#   were measured on (leaf_kernel_digest()); a build whose sources differ is
#   flagged in the JSON ("counters_match_build": false)
LEAF_COUNTERS_SOURCES = "62b726b6c1fc2972"
#   issue slots (one per 4 cycles) that one such MFMA takes from the vector
#   issue of its SIMD beside full-rate VALU work: measured live by the
#   "issue" leg (zkx_issue_probe); this value only when that leg is skipped
MFMA_ISSUE_SLOTS = 3.25
#   HBM bytes per commit: FETCH_SIZE x 2 (the counter reads half the bytes of
#   a plain 1 GiB copy) + WRITE_SIZE x 1, KiB x 1024, by work-group size
LEAF_TRAFFIC_BYTES_PER_COMMIT = {(8, 0): 7.636e8, (9, 0): 7.669e8, (10, 0): 8.049e8, (9, 1): 7.912e8}
# Sources of the leaf-hash launch (zeknox_amd/csrc)
LEAF_KERNEL_SOURCES = ("gl64.h", "gl_constants.h", "gmem.h", "hash.h", "poseidon_mfma.h", "merkle_kernels.h",
                       "merkle_launch.h", "merkle_launch_impl.h", "merkle_types.h", "merkle_poseidon.hip")
#   NTT passes (both), same scaling
NTT_TRAFFIC_BYTES_PER_COMMIT = 8.855e8
#   VALU wave-instructions of the NTT launches per commit (both passes,
#   SQ_INSTS_VALU), keyed by the LDE's pass split (stages per pass), measured
#   on the sources NTT_COUNTERS_SOURCES
NTT_VALU_WAVE_INSTS_PER_COMMIT = {(3, 12): 1.406e8, (4, 11): 1.391e8, (5, 10): 1.468e8, (6, 9): 1.428e8,
                                  (7, 8): 1.460e8, (8, 7): 1.453e8}
NTT_COUNTERS_SOURCES = "8bebc20fa8d74bff"
# 32-bit VALU issue peak: 256 CUs x 4 SIMDs x one wave64 instruction per 4
# cycles x 2.4 GHz = 39.3 T lane-ops/s (MI355X's published 78.6 TFLOP/s
# vector FP64 counts an FMA as 2 flops on the same issue rate)
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12


NTT_KERNEL_SOURCES = ("gl64.h", "gmem.h", "ntt_kernels.h", "ntt_types.h")


def leaf_kernel_digest(names=LEAF_KERNEL_SOURCES):
    """SHA-256 (first 16 hex digits) of a kernel's sources (by default the
    leaf-hash launch's)."""
    import hashlib
    h = hashlib.sha256()
    for name in names:
        with open(os.path.join(ROOT, "zeknox_amd", "csrc", name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--jobs", type=int, default=32, help="135 x 2^15 rate-3 commits per batched step")
    ap.add_argument("--no-graphs", action="store_true", help="batched steps launched kernel by kernel "
                    "(no captured graph replay)")
    ap.add_argument("--merkle-schedule", choices=("auto", "fused", "by_level"), default="auto",
                    help="Merkle schedule of batches: the library's rule (auto: tree levels launched apart from "
                         "2^18 leaves), always fused into the leaf launch, or always one launch per level")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="served-stream duration (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=224.0,
                    help="served-stream arrivals per second (default: 80% of the highest rate the server keeps "
                         "up with, 280/s in the sweep)")
    ap.add_argument("--serve-streams", type=int, default=6,
                    help="commit engines (streams) of the server (6: lowest latencies at 200/s, against 3 and 8)")
    ap.add_argument("--serve-trace", default=None,
                    help="write one CSV line per served request (times in ms from the first arrival)")
    ap.add_argument("--no-tune", action="store_true", help="batched path on the default schedule (no autotuning)")
    ap.add_argument("--schedule", default=None, metavar="S,E,T[,V]",
                    help="batched 135 x 2^15 schedule (ntt_strided_max,ntt_even,merkle_tile_log[,merkle_valu]) "
                         "instead of autotuning")
    ap.add_argument("--no-extra", action="store_true", help="skip the 64-job and 8 x 400 x 2^20 batches")
    ap.add_argument("--legs", default="batch64,large,hashers,ntt,config1",
                    help="extra legs to run (comma-separated): batch64, large, hashers, ntt, config1 (by "
                         "default); sweep (served-stream arrival-rate sweep), cpu3 and cpu4 (CPU baselines of the "
                         "64-job batch and of one 400 x 2^20 Poseidon2 job) on request")
    ap.add_argument("--extra-steps", type=int, default=3, help="timed steps of those batches")
    ap.add_argument("--cpu-runs", type=int, default=5, help="whole CPU commits timed (median reported)")
    ap.add_argument("--no-issue", action="store_true", help="skip the issue-rate probe (roofline uses constants)")
    ap.add_argument("--sweep-rates", default="40,80,120,160,200", help="arrival rates of the sweep leg (per s)")
    ap.add_argument("--sweep-seconds", type=float, default=20.0, help="duration of each sweep point")
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


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            return next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        return ""


def _ref_merkle(np, ob, ref, leaves, cols, nl, ht):
    """fill_digests_buf_linear_cpu of the reference build (oracle/_ref), or the
    oracle's restatement where that build is absent: (digests, caps)."""
    nc = 1 << CAP_H
    nd = 2 * (nl - nc)
    if ref is None:
        return ob.merkle(leaves, cols, CAP_H, ht)
    dig = np.zeros(nd * 4, np.uint64)
    cap = np.zeros(nc * 4, np.uint64)
    ref.fill_digests_buf_linear_cpu(ob._ptr(dig), ob._ptr(cap), ob._ptr(leaves), nd, nc, nl, cols, CAP_H, ht)
    return dig, cap


def cpu_config1(np, runs, lg=13, cols=135, seed=13):
    """Config 1 on the host cores: the Poseidon Merkle forest of 2^13 leaves x
    135 (seeded canonical elements), cap height 4, through the reference's own
    fill_digests_buf_linear_cpu (native/merkle/merkle.cpp:126-161, OpenMP),
    `runs` times, median reported."""
    import statistics
    import oracle_bind as ob
    ref = ob.ref_lib()
    nl = 1 << lg
    leaves = ob.canonical(cols << lg, seed)
    ts, caps = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        _, caps = _ref_merkle(np, ob, ref, leaves, cols, nl, 0)
        ts.append(time.perf_counter() - t0)
    med = statistics.median(ts)
    return {"value": round(1.0 / med, 2), "unit": "trees/s", "ms_per_tree": round(med * 1e3, 2),
            "kind": "reference" if ref is not None else "port", "cores": ob.orc.orc_num_threads(),
            "sample": f"{runs} Poseidon Merkle forests of 2^{lg} leaves x {cols}, cap_height {CAP_H}, median",
            "runs_ms": [round(t * 1e3, 2) for t in ts], "caps": caps, "leaves": leaves}


def gpu_config1(zk, np, steps, cpu_caps, leaves, lg=13, cols=135):
    """Config 1 on the GPU and on the library's own host path: the same
    2^13 x 135 Poseidon forest through fill_digests_buf_linear_gpu_with_gpu_ptr
    (device pointers), zkx_fill_digests_buf_linear_gpu_noalloc, and the
    library's fill_digests_buf_linear_cpu (host buffers, no GPU); caps compared
    with the reference CPU's."""
    nl, nc = 1 << lg, 1 << CAP_H
    nd = 2 * (nl - nc)
    dl = zk.DeviceBuffer.from_host(leaves)
    dd, dc = zk.DeviceBuffer(nd * 4), zk.DeviceBuffer(nc * 4)
    gpu = _timed(zk, lambda: zk.fill_digests_buf_linear_gpu_with_gpu_ptr(dd, dc, dl, nd, nc, nl, cols, CAP_H, 0, 0),
                 steps, 2)
    same_gpu = bool(np.array_equal(dc.to_host(), cpu_caps))
    na = _timed(zk, lambda: zk.fill_digests_buf_linear_gpu_noalloc(dd, dc, dl, nd, nc, nl, cols, CAP_H, 0, 0),
                steps, 2)
    same_na = bool(np.array_equal(dc.to_host(), cpu_caps))
    hd, hc = np.zeros(nd * 4, np.uint64), np.zeros(nc * 4, np.uint64)
    t0 = time.perf_counter()
    for _ in range(max(steps // 4, 3)):
        zk.fill_digests_buf_linear_cpu(hd, hc, leaves, nd, nc, nl, cols, CAP_H, 0)
    host = (time.perf_counter() - t0) / max(steps // 4, 3)
    for b in (dl, dd, dc):
        b.free()
    return {"workload": f"config 1: Poseidon Merkle forest of 2^{lg} leaves x {cols}, cap_height {CAP_H}",
            "gpu_ms_per_tree": round(gpu * 1e3, 4), "gpu_trees_per_s": round(1.0 / gpu, 1),
            "gpu_noalloc_ms_per_tree": round(na * 1e3, 4),
            "library_host_ms_per_tree": round(host * 1e3, 2), "library_host_threads": os.cpu_count(),
            "caps_equal_reference_cpu": same_gpu and same_na and bool(np.array_equal(hc, cpu_caps))}


def cpu_config3(np, runs):
    """Config 3 on the host cores: the 64 mixed jobs (zeknox_amd.workloads,
    numpy seed 3) one after another -- oracle LDE + leaf rows, then the
    reference's CPU Merkle forest -- `runs` times, median; every job's caps
    compared with the committed expected outputs (tests/golden/workloads.json)."""
    import hashlib
    import statistics
    import oracle_bind as ob
    from zeknox_amd.workloads import mixed_batch_jobs
    ref = ob.ref_lib()
    with open(os.path.join(ROOT, "tests", "golden", "workloads.json")) as f:
        exp = json.load(f)["batch64"]
    jobs = mixed_batch_jobs()
    inputs = [ob.salt(c << l, seed) for c, l, r, ht, seed in jobs]
    ts, same = [], True
    for _ in range(runs):
        t0 = time.perf_counter()
        caps = []
        for (c, l, r, ht, _), x in zip(jobs, inputs):
            leaves = ob.transpose_rev(ob.lde(x, l, r, c), l + r, c)
            caps.append(_ref_merkle(np, ob, ref, leaves, c, 1 << (l + r), ht)[1])
        ts.append(time.perf_counter() - t0)
        print(f"cpu config 3 run {len(ts)}: {ts[-1]:.1f} s", file=sys.stderr, flush=True)
        same = same and all(hashlib.sha256(np.ascontiguousarray(cp, "<u8").tobytes()).hexdigest() == e["caps_sha256"]
                            for cp, e in zip(caps, exp))
    med = statistics.median(ts)
    return {"value": round(64 / med, 3), "unit": "commits/s", "s_per_batch": round(med, 2),
            "kind": "reference" if ref is not None else "port", "cores": ob.orc.orc_num_threads(),
            "sample": f"{runs} runs of the 64 mixed jobs, sequential, median (oracle LDE + transpose, reference "
                      "CPU Merkle)", "runs_s": [round(t, 2) for t in ts], "caps_equal_expected": same}


def cpu_config4(zk, np, runs):
    """Config 4 per job on the host cores: one 400 x 2^20, rate-1 Poseidon2
    commit (oracle LDE + leaf rows, reference CPU Merkle), `runs` times,
    median; caps compared with the same job through zkx_commit_batch."""
    import statistics
    import oracle_bind as ob
    ref = ob.ref_lib()
    cols, lg_n, rate, ht = 400, 20, 1, 3
    x = ob.salt(cols << lg_n, 4000)
    d = zk.DeviceBuffer.from_host(x)
    c = zk.Commit(d, cols, lg_n, rate, CAP_H, ht, leaves=False)
    zk.commit_batch([c])
    gpu_caps = c.d_caps.to_host()
    for b in (d, c.d_digests, c.d_caps):
        b.free()
    zk.lib.zkx_commit_trim(0)
    ts, same = [], True
    for _ in range(runs):
        t0 = time.perf_counter()
        leaves = ob.transpose_rev(ob.lde(x, lg_n, rate, cols), lg_n + rate, cols)
        _, caps = _ref_merkle(np, ob, ref, leaves, cols, 1 << (lg_n + rate), ht)
        ts.append(time.perf_counter() - t0)
        print(f"cpu config 4 run {len(ts)}: {ts[-1]:.1f} s", file=sys.stderr, flush=True)
        same = same and bool(np.array_equal(caps, gpu_caps))
        del leaves
    med = statistics.median(ts)
    return {"value": round(1.0 / med, 4), "unit": "commits/s", "s_per_commit": round(med, 2),
            "kind": "reference" if ref is not None else "port", "cores": ob.orc.orc_num_threads(),
            "sample": f"{runs} commits of one 400 x 2^20 rate-1 Poseidon2 job, median (oracle LDE + transpose, "
                      "reference CPU Merkle)", "runs_s": [round(t, 2) for t in ts], "caps_equal_gpu": same}


def issue_leg(zk, iters=8192):
    """zkx_issue_probe: the achieved v_mad_u64_u32 rate (four waves per SIMD)
    and the issue slots one i8 MFMA takes beside it."""
    r = zk.issue_probe(iters)
    nominal = r["simds"] * 2.4e9 / 4  # one wave-instruction per 4 cycles per SIMD at 2.4 GHz
    return {"valu_lane_ops_per_s": r["valu_wave_insts_per_s"] * 64,
            "valu_frac_of_nominal": round(r["valu_wave_insts_per_s"] / nominal, 3),
            "mfma_issue_slots": round(r["mfma_issue_slots"], 3),
            "mfma_alone_wave_insts_per_s": r["mfma_wave_insts_per_s"],
            "ms_valu_mfma_mixed": [round(r["ms_valu"], 4), round(r["ms_mfma"], 4), round(r["ms_mixed"], 4)],
            "mads_per_mfma": r["mads_per_mfma"], "waves_per_simd": r["waves_per_simd"]}


def served_sweep(zk, np, rates, seconds, budget, streams):
    """The served stream at several arrival rates: per rate commits/s, latency
    percentiles and batches per request; the knee is the highest rate at
    which the server keeps up (the last request completes within 5% + 0.5 s
    of the stream's end) with a queueing p99 under one second."""
    pts = []
    for i, rate in enumerate(rates):
        r = served_stream(zk, np, seconds, rate, budget, streams, seed=50 + i)
        pts.append({"rate_hz": rate, "commits_per_s": r["commits_per_s"], "p50_ms": r["p50_ms"],
                    "p99_ms": r["p99_ms"], "queue_p99_ms": r["queue_p99_ms"], "small_p99_ms": r["small_p99_ms"],
                    "large_p99_ms": r["large_p99_ms"], "requests": r["requests"],
                    "batches_per_request": round(r["batches"] / max(r["requests"], 1), 3),
                    "max_batch": r["max_batch"], "peak_hbm_gb": r["peak_hbm_gb"], "elapsed_s": r["elapsed_s"]})
        print(f"sweep {rate}/s: {r['commits_per_s']} commits/s, p50 {r['p50_ms']} ms, p99 {r['p99_ms']} ms",
              file=sys.stderr, flush=True)
    ok = [p for p in pts if p["elapsed_s"] <= 1.05 * seconds + 0.5 and (p["queue_p99_ms"] or 0) < 1000.0]
    return {"workload": f"served stream (as below) at each rate for {seconds} s", "points": pts,
            "knee_rate_hz": max((p["rate_hz"] for p in ok), default=None)}


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
            "schedules (ntt_strided_max, ntt_even, merkle_tile_log, merkle_valu)": [list(t) for t in schedules]}


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
    from zeknox_amd.workloads import served_trace
    # one device pool of coefficients serves every request (its first cols*2^lg_n words)
    pool = zk.DeviceBuffer(400 << 20)
    pool.fill_random(seed)
    srv = zk.Server(budget, streams=streams, max_batch=64)
    arrivals = served_trace(seed, seconds, rate_hz)
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
            "elapsed_s": round(elapsed, 2),
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

    # ---- batched path: JOBS such commits per zkx_commit_batch call (the
    # engine captures a repeated batch's launches as a graph on its third
    # call: replay is used only when the warm-up steps cover that capture)
    graphs = not args.no_graphs and args.warmup >= 3
    zk.set_commit_graphs(graphs)
    zk.set_level_launches({"auto": zk.LEVEL_LAUNCHES_DEFAULT, "fused": None, "by_level": 0}[args.merkle_schedule])
    commits = [zk.Commit(d_in, COLS, LG_N, RATE, CAP_H, HASH) for _ in range(args.jobs)]
    if args.schedule:
        zk.set_tuning(zk.Tuning(*[int(v) for v in args.schedule.split(",")]), commits[0])
    elif not args.no_tune:  # one-time schedule search for the shape, outside the timing
        zk.autotune(commits[:1])
    schedule = dict(zip(("ntt_strided_max", "ntt_even", "merkle_tile_log", "merkle_valu"),
                        zk.get_tuning(commits[0]).astuple()))
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

    # ---- issue rates of this GPU (v_mad_u64_u32 alone; an i8 MFMA beside it)
    issue = None if args.no_issue else issue_leg(zk)

    # ---- roofline of the dominant kernel: the leaf-hash launch of the batch
    # (Poseidon permutations, issue-bound: 64-bit modular arithmetic on the
    # 32-bit integer VALU, the MDS layers as i8 MFMAs that share the SIMD's
    # issue).  achieved = permutations per launch x issue slots per
    # permutation (VALU wave-instructions + the measured slots per MFMA) x 64
    # lanes / the launch's HIP-event duration (events on the library's
    # stream); peak = VALU_PEAK_TOPS (one wave-instruction per 4 cycles).
    dom = max(batch_kernels, key=lambda k: batch_kernels[k]["ms_per_step"])
    leaf = batch_kernels.get("merkle_leaves")
    # the batch builds its tree levels in launches of their own (the
    # library's default for batches of >= LEVEL_LAUNCHES_DEFAULT leaves)
    levels_apart = args.merkle_schedule == "by_level" or (args.merkle_schedule == "auto"
                                                          and args.jobs * N >= zk.LEVEL_LAUNCHES_DEFAULT)
    roofline = None
    if leaf and levels_apart and (schedule["merkle_tile_log"], schedule["merkle_valu"]) in VALU_WAVE_INSTS_PER_PERM:
        tl = (schedule["merkle_tile_log"], schedule["merkle_valu"])
        perms = args.jobs * N * ((COLS + 7) // 8)  # the leaf sponges
        t_launch = leaf["avg_launch_ms"] * 1e-3
        mslots = issue["mfma_issue_slots"] if issue else MFMA_ISSUE_SLOTS
        valu = perms * VALU_WAVE_INSTS_PER_PERM[tl] * 64 / t_launch / 1e12
        mfma = perms * mslots * MFMA_WAVE_INSTS_PER_PERM[tl] * 64 / t_launch / 1e12
        achieved = valu + mfma
        alg_bytes = args.jobs * (8 * COLS * N * 2 + 32 * N)  # LDE read, leaf rows + leaf digests written
        traffic = LEAF_TRAFFIC_BYTES_PER_COMMIT.get(tl)
        digest = leaf_kernel_digest()
        roofline = {"bound": "valu+mfma issue", "achieved": round(achieved, 2), "peak": round(VALU_PEAK_TOPS, 2),
                    "unit": "T lane-issue-slots/s (32-bit VALU)", "frac": round(achieved / VALU_PEAK_TOPS, 3),
                    "valu_frac": round(valu / VALU_PEAK_TOPS, 3), "mfma_slot_frac": round(mfma / VALU_PEAK_TOPS, 3),
                    "frac_of_measured_mad_rate": round(achieved * 1e12 / issue["valu_lane_ops_per_s"], 3)
                    if issue else None,
                    "traffic": round(traffic * args.jobs) if traffic else None, "algorithmic_bytes": alg_bytes,
                    "traffic_over_algorithmic": round(traffic * args.jobs / alg_bytes, 3) if traffic else None,
                    "hbm_gbs": round(traffic * args.jobs / t_launch / 1e9, 1) if traffic else None,
                    "hbm_frac": round(traffic * args.jobs / t_launch / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                    "kernel": "merkle_leaves", "avg_launch_ms": round(leaf["avg_launch_ms"], 4),
                    "perms_per_launch": perms, "valu_wave_insts_per_perm": VALU_WAVE_INSTS_PER_PERM[tl],
                    "mfma_wave_insts_per_perm": MFMA_WAVE_INSTS_PER_PERM[tl], "mfma_issue_slots": mslots,
                    "mfma_issue_slots_source": "measured (zkx_issue_probe)" if issue else "constant",
                    "leaf_kernel_sources": digest, "counters_match_build": digest == LEAF_COUNTERS_SOURCES}

    # the tree levels: one launch per level over every tree (internal nodes
    # and caps: N - caps permutations per commit)
    lv = batch_kernels.get("merkle_levels")
    merkle_levels = None
    if lv:
        lperms = args.jobs * (N - nc)
        merkle_levels = {"launches_per_step": lv["launches_per_step"], "ms_per_step": round(lv["ms_per_step"], 4),
                         "perms_per_step": lperms, "perms_per_s": round(lperms / (lv["ms_per_step"] * 1e-3)),
                         "note": "one launch per level over all trees; the narrowest levels are bound by one "
                                 "permutation's latency"}

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
        split = tuple(zk.ntt_pass_sizes(LG_N, zk.get_tuning(commits[0])))
        ntt_valu = NTT_VALU_WAVE_INSTS_PER_COMMIT.get(split)
        ntt_roofline["pass_split"] = list(split)
        if ntt_valu:
            # the bound that holds: VALU issue of the same launches
            valu = ntt_valu * args.jobs * 64 / (ntt["ms_per_step"] * 1e-3) / 1e12
            digest = leaf_kernel_digest(NTT_KERNEL_SOURCES)
            ntt_roofline["valu_issue"] = {"achieved": round(valu, 2), "peak": round(VALU_PEAK_TOPS, 2),
                                          "unit": "T lane-ops/s (32-bit VALU)", "frac": round(valu / VALU_PEAK_TOPS, 3),
                                          "frac_of_measured_mad_rate": round(valu * 1e12 / issue["valu_lane_ops_per_s"], 3)
                                          if issue else None,
                                          "valu_wave_insts_per_commit": ntt_valu,
                                          "ntt_kernel_sources": digest,
                                          "counters_match_build": digest == NTT_COUNTERS_SOURCES}

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
    else:
        legs = set()

    # ---- CPU baseline: whole commits of the same input on the host cores;
    # config 1 (the reference's CPU Merkle forest) beside it, and on request
    # the 64-job batch and one 400 x 2^20 Poseidon2 job
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(np, coeffs, args.cpu_runs, d_cap.to_host())
        c1 = cpu_config1(np, max(args.cpu_runs, 5))
        c1_caps, c1_leaves = c1.pop("caps"), c1.pop("leaves")
        cpu["configs"] = {"1": c1}
        if "config1" in legs:
            extra["config1"] = gpu_config1(zk, np, 4 * args.extra_steps, c1_caps, c1_leaves)
        if "cpu3" in legs:
            cpu["configs"]["3"] = cpu_config3(np, args.cpu_runs)
        if "cpu4" in legs:
            cpu["configs"]["4"] = cpu_config4(zk, np, args.cpu_runs)
    if "sweep" in legs:
        extra["served_sweep"] = served_sweep(zk, np, [float(r) for r in args.sweep_rates.split(",")],
                                             args.sweep_seconds, 64 << 30, args.serve_streams)

    served = None
    if args.serve_seconds > 0:
        served = served_stream(zk, np, args.serve_seconds, args.serve_rate, 64 << 30, args.serve_streams,
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
                   "graphs": graphs, "merkle_levels_apart": levels_apart,
                   "schedule": schedule},
        "plain_commits_per_s": round(1.0 / plain_s, 2),
        "batched_over_plain": round(plain_s / batched_s, 3),
        "batched_equals_plain": bit_exact,
        "kernels_plain": kernels,
        "kernels_batched": batch_kernels,
        "dominant_kernel": dom,
        "roofline": roofline,
        "merkle_levels": merkle_levels,
        "issue_rates": issue,
        "ntt_roofline": ntt_roofline,
        "cpu_baseline": cpu,
        "served": served,
        **extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
