/*
 * zeknox_ext.h -- additions of the MI355X build that have no counterpart in
 * the reference's C ABI.  Callers of the reference interface need none of
 * them; they serve the test harness, benchmarks and batch callers.
 */
#ifndef ZEKNOX_AMD_ZEKNOX_EXT_H
#define ZEKNOX_AMD_ZEKNOX_EXT_H

#include <stddef.h>
#include <stdint.h>

#include "zeknox.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reason of the last failed void entry point on this thread ("" if none).
 * The reference's Merkle entry points assert instead (merkle.cu:258). */
const char *zkx_last_error(void);

/* Device memory helpers (the reference's wrappers bind cudart for this:
 * wrappers/rust/src/device/memory.rs, wrappers/go/device/memory.go). */
RustError zkx_device_malloc(size_t device_id, void **ptr, size_t bytes);
RustError zkx_device_free(size_t device_id, void *ptr);
RustError zkx_memcpy_htod(size_t device_id, void *dst, const void *src, size_t bytes);
RustError zkx_memcpy_dtoh(size_t device_id, void *dst, const void *src, size_t bytes);
RustError zkx_memset(size_t device_id, void *dst, int value, size_t bytes);
RustError zkx_device_synchronize(size_t device_id);
/* Synthetic input on the device: dst[i] = the first value below p of the
 * splitmix64 stream started at seed ^ (i * 0xD1B54A32D192ED03), so uniform
 * canonical field elements (benchmarks; the seeded salt's stream). */
RustError zkx_fill_random(size_t device_id, void *dst, size_t count, uint64_t seed);

/* Goldilocks arithmetic on device vectors through the library's own device
 * routines (those the hash and NTT kernels use): out[i] = op(a[i], b[i]),
 * canonical.  op 0: a * b (lazy product, any a, b < 2^64); 1: a * b;
 * 2: b * 2^64 + a (any 128-bit value); 3: (b mod 2^32) * 2^64 + a; 4: a^7
 * (any a < 2^64); 5: a + b (a, b < p); 6: a - b (a, b < p); 7: a mod p.
 * No counterpart in the reference (its field tests are C++ unit tests). */
RustError zkx_gl_ops(size_t device_id, int op, const void *a, const void *b, void *out, size_t n);

/* compute_batched_lde with a reproducible salt.  compute_batched_lde seeds
 * the salt_size columns that follow the LDE from the clock, as the reference
 * does (ntt/ntt.cu:541-546, ntt/kernels.cu:62-84), so they differ on every
 * call; here element g of the salt block (g < salt_size * 2^(lg+rate),
 * column-major like the LDE) is the first value below p of the splitmix64
 * stream started at salt_seed ^ (g * 0xD1B54A32D192ED03), the stream of
 * zkx_fill_random.  Everything else is compute_batched_lde.  No counterpart
 * in the reference. */
RustError zkx_compute_batched_lde_seeded(size_t device_id, void *output, void *input, uint32_t lg_domain_size,
                                         NTT_Direction ntt_direction, NTT_Config cfg, uint64_t salt_seed);

/* fill_digests_buf_linear_gpu_with_gpu_ptr (same arguments, results and
 * checks) without any device allocation, for callers that cannot allocate:
 * code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc.  No counterpart in the reference.  The fused launch's scratch
 * (a pass descriptor and the upper passes' arrival counters) is a fixed
 * ZKX_NOALLOC_SCRATCH_BYTES array per device, held until the process ends.
 * It holds the arrival counters of every tree of up to 2^32 leaves at any
 * work-group size (at most 65,793 of its 131,008); a tree that needs more is
 * rejected with code 1, as are trees beyond the fused launch's four passes
 * (also rejected by fill_digests_buf_linear_gpu_with_gpu_ptr).  Every call zeroes
 * the counters it uses and returns when its work is complete; calls on one
 * device are serialised by the library, so they never overlap on the array.
 * On failure the message is malloc'd (free() it), as for RustError. */
#define ZKX_NOALLOC_SCRATCH_BYTES (512u * 1024u)
RustError zkx_fill_digests_buf_linear_gpu_noalloc(void *digests_buf, void *cap_buf, void *leaves_buf,
                                                  uint64_t digests_buf_size, uint64_t cap_buf_size,
                                                  uint64_t leaves_buf_size, uint64_t leaf_size, uint64_t cap_height,
                                                  uint64_t hash_type, uint64_t gpu_id);

/* Issue-rate probe of the device's SIMDs (the measured inputs of the hash
 * kernels' roofline), four waves per SIMD on every SIMD, each kernel timed
 * with HIP events over `iters` trips (a positive multiple of 4):
 * full-rate 32-bit multiply-adds (v_mad_u64_u32, independent chains), i8
 * matrix-core instructions alone (v_mfma_i32_32x32x32_i8, four independent
 * accumulators), and both together (mads_per_mfma multiply-adds per MFMA);
 * each the fastest of five launches after a warm-up launch.
 * mfma_issue_slots = multiply-add issue slots one MFMA takes from its SIMD
 * beside them: (t_mixed - t_valu) per MFMA over t_valu per multiply-add.
 * No counterpart in the reference. */
typedef struct {
    double valu_wave_insts_per_s;  /* whole device */
    double mfma_wave_insts_per_s;  /* whole device, MFMAs alone */
    double mfma_issue_slots;
    double ms_valu, ms_mfma, ms_mixed;
    uint32_t simds, waves_per_simd, mads_per_mfma, reserved;
} ZkxIssueRates;
RustError zkx_issue_probe(size_t device_id, uint32_t iters, ZkxIssueRates *out);

/* The stream the library enqueues its work on (a hipStream_t). */
void *zkx_stream(size_t device_id);

/* ---- batched commits ---------------------------------------------------
 * One commit = coset LDE (generator of init_coset) of `cols` polynomials of
 * 2^lg_n coefficients (column-major, device) with blowup 2^rate; the leaf
 * rows leaves[t * cols + b] = LDE_b[bitrev(t)] (compute_transpose_rev of the
 * LDE; optional, NULL to skip); the Merkle forest of those rows with
 * 2^cap_height caps (hash_type as merkle.h: 0 Poseidon, 1 Keccak-256,
 * 2 Poseidon-BN128, 3 Poseidon2, 4 Monolith).  Outputs are
 * bit-identical to compute_batched_lde -> compute_transpose_rev ->
 * fill_digests_buf_linear_gpu_with_gpu_ptr on the same inputs.  All pointers
 * are device pointers; sizes as in merkle.h. */
typedef struct {
    const void *coeffs;
    void *leaves;
    void *digests;
    void *caps;
    uint32_t cols, lg_n, rate, cap_height, hash_type, reserved;
} ZkxCommitJob;

/* All jobs of one call run together (one launch per pass over all jobs);
 * returns when they are complete. */
RustError zkx_commit_batch(size_t device_id, const ZkxCommitJob *jobs, size_t njobs);
/* zkx_commit_batch with its long launches split: every NTT pass launch in
 * ranges of at most `ntt_tiles` 4096-element tiles and every Merkle launch
 * in ranges of at most `merkle_groups` work-groups (0: one launch), in
 * order on the same stream.  For callers that share the GPU with
 * latency-sensitive work on other streams: the device dispatches another
 * stream's kernels between two launches, not while a launch still has
 * work-groups to dispatch (the commit server runs its large requests this
 * way).  Same results. */
RustError zkx_commit_batch_pieces(size_t device_id, const ZkxCommitJob *jobs, size_t njobs, uint64_t ntt_tiles,
                                  uint64_t merkle_groups);
/* Device working memory zkx_commit_batch needs for these jobs, in bytes.
 * The library keeps it between calls (one buffer per device, grown as
 * needed); zkx_commit_trim frees it. */
size_t zkx_commit_scratch_bytes(const ZkxCommitJob *jobs, size_t njobs);
RustError zkx_commit_trim(size_t device_id);

/* ---- schedule autotuning -------------------------------------------------
 * How the commit of a shape (columns, log degree, rate bits, hash type, leaf
 * rows kept) is cut into launches: the LDE's DIF stages per strided pass
 * (1..9; greedy or spread evenly), the leaves per Merkle work-group
 * (2^8..2^10) and, for Poseidon leaves, whether the permutations' MDS layers
 * run on the matrix cores or every step on the VALU.  No setting changes a
 * result.  Shapes without an entry use
 * the default, which the C-ABI entry points (compute_batched_lde,
 * fill_digests_buf_linear_gpu_with_gpu_ptr) use as well, except that until
 * zkx_set_tuning sets a default the Merkle entry points size the work-groups
 * of one tree by its leaf count (2^8 lanes up to 2^16 leaves, 2^(lg-8) at
 * 2^17 and 2^18, then 2^9 for Poseidon and Poseidon2, 2^10 for Poseidon
 * over BN254, 2^8 for the others: the fastest for a tree alone), where
 * batches use the default's 2^8. */
typedef struct {
    uint32_t ntt_strided_max;  /* 1..9 */
    uint32_t ntt_even;         /* 0 greedy, 1 even */
    uint32_t merkle_tile_log;  /* 8..10 */
    uint32_t merkle_valu;      /* Poseidon leaves: 0 MDS on the i8 matrix cores, 1 all-VALU */
} ZkxTuning;

/* Times every distinct shape among the jobs under each schedule (`iters`
 * runs each, fastest kept) and records the best for the shape.  The jobs'
 * outputs are written with their results. */
RustError zkx_autotune(size_t device_id, const ZkxCommitJob *jobs, size_t njobs, int iters);
/* job NULL: the default. */
RustError zkx_get_tuning(size_t device_id, const ZkxCommitJob *job, ZkxTuning *out);
RustError zkx_set_tuning(size_t device_id, const ZkxCommitJob *job, const ZkxTuning *tuning);
/* on: zkx_commit_batch and the server tune each shape they have not seen
 * on its first commit (that commit takes longer). */
RustError zkx_set_autotune(size_t device_id, int on);
/* on (the default): zkx_commit_batch and the server's engines replay the
 * launches of a batch whose job shapes and schedules they have run before
 * as a captured HIP graph (one graph launch per batch; the descriptors are
 * re-staged each call); off: every batch is launched kernel by kernel.  No
 * setting changes a result. */
RustError zkx_set_commit_graphs(size_t device_id, int on);
/* Merkle schedule of zkx_commit_batch and the server's engines: a batch
 * whose trees hold at least min_leaves leaves in total (default 2^18) hashes
 * the leaves in one launch per hash kind and builds the levels above them
 * with one launch per level over every tree; smaller batches finish every
 * tree inside the leaf launch.  0: always by level; UINT64_MAX: never.  No
 * setting changes a result. */
RustError zkx_set_level_launches(size_t device_id, uint64_t min_leaves);
/* The same for one tree through fill_digests_buf_linear_gpu_with_gpu_ptr and
 * the other C-ABI Merkle entry points: Poseidon and Poseidon2 trees of at
 * least min_leaves leaves (default 2^20), Keccak-256 and Monolith trees of at
 * least min_leaves / 4, Poseidon-BN128 trees of at least min_leaves / 32; the
 * allocation-free entry point always finishes the tree in its leaf launch.
 * 0: always by level; UINT64_MAX: never.  No setting changes a result. */
RustError zkx_set_lone_level_launches(size_t device_id, uint64_t min_leaves);
/* Graphs the device's zkx_commit_batch engine holds (tests, benchmarks). */
size_t zkx_commit_graphs_captured(size_t device_id);
/* Stage counts of the DIF passes over 2^lg points under `tuning` (NULL: the
 * built-in default), top stages first; returns the pass count (at most
 * `cap` written), -1 for parameters out of range.  Host-only. */
int zkx_ntt_pass_sizes(uint32_t lg, const ZkxTuning *tuning, uint32_t *sizes, int cap);
/* The pass plan of the one-launch Merkle forest over nl leaves and nc caps
 * with work-groups of at most 2^tile_log lanes (8..10), host-only: per pass
 * (at most `cap` written) its input nodes over all subtrees, log2 width of
 * its input level per subtree and tile log, and for every pass after the
 * first the index of its first arrival counter (the last work-group of a
 * group of 2^tile_log tiles to arrive hashes the group's roots); *counters =
 * the arrival counters the launch uses.  Returns the pass count, -1 for
 * sizes that are not a linear forest. */
int zkx_merkle_plan(uint64_t nl, uint64_t nc, uint32_t tile_log, uint64_t *count, uint32_t *lg_w,
                    uint32_t *tile_logs, uint32_t *first_ctr, int cap, uint64_t *counters);
/* Forget every shape entry and restore the built-in default. */
RustError zkx_tuning_clear(size_t device_id);
/* The shape table as text, one shape per line. */
RustError zkx_tuning_save(size_t device_id, const char *path);
RustError zkx_tuning_load(size_t device_id, const char *path);

/* ---- commit server (continuous batching under an HBM budget) ------------
 * Requests are commit jobs as above whose output pointers are ignored: the
 * server places leaf rows (if want_leaves), digests and caps in device memory
 * it allocated once at creation (budget_bytes) and keeps them until
 * zkx_server_release; serving itself allocates nothing.  Requests that
 * arrive while others run are admitted as soon as an engine (stream) is free
 * and their working set can be placed next to everything admitted and not
 * yet released.
 * Small requests (LDE of at most 2^28 elements) are packed into batches of
 * at most 2^28 elements; large ones run one per batch on at most nstreams - 1
 * engines, and the oldest large request keeps its room in the budget
 * against later small ones. */
typedef struct {
    void *leaves, *digests, *caps;  /* device outputs (leaves NULL if not wanted) */
    double t_submit, t_start, t_done; /* steady-clock seconds: submitted, admitted, finished */
    int state;                        /* 2 done, 3 failed */
    int reserved;
    double t_enqueued;                /* the request's kernels were enqueued */
} ZkxServeResult;

typedef struct {
    uint64_t submitted, completed, failed, batches, max_batch;
    size_t bytes_in_use, peak_bytes;
} ZkxServeStats;

RustError zkx_server_create(size_t device_id, size_t budget_bytes, int nstreams, size_t max_batch_jobs,
                            void **server);
RustError zkx_server_destroy(void *server);
/* Device bytes a request occupies while admitted (scratch + outputs). */
size_t zkx_server_footprint(const ZkxCommitJob *job, int want_leaves);
RustError zkx_server_submit(void *server, const ZkxCommitJob *job, int want_leaves, uint64_t *id);
RustError zkx_server_wait(void *server, uint64_t id, ZkxServeResult *result);
/* Blocks until a request finishes that no zkx_server_wait / _wait_any call
 * has returned yet; writes its id.  Code 1 when no request is pending. */
RustError zkx_server_wait_any(void *server, uint64_t *id, ZkxServeResult *result);
RustError zkx_server_release(void *server, uint64_t id);
RustError zkx_server_stats(void *server, ZkxServeStats *stats);

/* Kernel timing: when enabled, every launch of the library's kernels on a
 * device is bracketed by HIP events on the launching stream; zkx_kernel_stats
 * reports, per kernel family, the launch count and summed duration. */
RustError zkx_profile_enable(int on);
/* Writes up to `cap` entries; returns the number of families. */
int zkx_kernel_stats(char names[][32], uint64_t *launches, double *total_ms, int cap);
void zkx_kernel_stats_reset(void);

#ifdef __cplusplus
}
#endif

#endif /* ZEKNOX_AMD_ZEKNOX_EXT_H */
