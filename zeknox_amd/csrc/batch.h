// Batched LDE + Merkle commit engine.
//
// A commit job = coset LDE of `cols` polynomials of 2^lg_n coefficients with
// blowup 2^rate, its leaf matrix (row t = the evaluations at bitrev(t), i.e.
// compute_transpose_rev of the LDE), and the Poseidon / Poseidon2 Merkle
// forest with 2^cap_height caps over those rows -- exactly what the C-ABI
// sequence compute_batched_lde -> compute_transpose_rev ->
// fill_digests_buf_linear_gpu_with_gpu_ptr produces.
//
// An Engine runs any number of uneven jobs together on one stream: one
// launch per NTT pass index covers that pass of every job, one launch per
// hash kind and work-group size hashes every tree's leaves.  In a small
// batch the same launch finishes the trees (the upper levels of each tree
// by the last work-group of each group of tiles), so the latency-bound top
// levels of different trees overlap with other trees' leaf hashing; a batch
// whose trees hold at least Device::level_launch_leaves leaves in total
// builds the levels with one launch per level over every tree instead, each
// level at full occupancy.  The LDE stays in HBM in its natural DIF output order
// (bit-reversed, column-major) and the leaf kernels read it there directly:
// no bit-reversal or transpose pass; the leaf rows are written by the leaf
// kernel as it reads them.
#pragma once
#include <map>
#include <vector>

#include "merkle.h"
#include "ntt.h"
#include "runtime.h"
#include "tuning.h"

namespace zkx {

struct CommitJob {
    const uint64_t *coeffs = nullptr;  // device, cols x 2^lg_n, column-major
    uint64_t *leaves = nullptr;        // device, 2^(lg_n+rate) x cols rows, or null
    uint64_t *digests = nullptr;       // device, 2 * (N - 2^cap_height) x 4
    uint64_t *caps = nullptr;          // device, 2^cap_height x 4
    uint32_t cols = 0, lg_n = 0, rate = 0, cap_height = 0, hash_type = 0;
};

// Long launches of a commit issued as several launches over ranges of at
// most this many NTT tiles / Merkle work-groups (0: one launch each).
struct LaunchPieces {
    uint64_t ntt_tiles = 0;
    uint64_t merkle_groups = 0;
};

void check_commit_job(const CommitJob &j);
// Device bytes of working memory an Engine needs for these jobs.
size_t commit_scratch_bytes(const CommitJob *jobs, size_t n);

class Engine {
public:
    // stream == nullptr: the engine creates its own (default-priority) stream
    Engine(Device &d, hipStream_t stream = nullptr);
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    // Enqueue the jobs on the engine's stream; returns without waiting.  The
    // tables the jobs need are built on first use.  Calls on one engine are
    // serialised by `mu`.  `scratch` (>= commit_scratch_bytes, device,
    // stream-ordered on this engine's stream) replaces the engine's own
    // working buffer when given.
    // `forced`, when given, replaces every job's schedule parameters.
    // `lde_bufs`, when given, holds one working buffer per job instead
    // (job i: cols << (lg_n + rate) elements, 256-byte aligned).
    // `pieces`: long launches issued as several launches over ranges (a
    // large served request, so that other streams' kernels are dispatched
    // between them).
    void commit(const CommitJob *jobs, size_t n, void *scratch = nullptr, const Tuning *forced = nullptr,
                uint64_t *const *lde_bufs = nullptr, const LaunchPieces &pieces = LaunchPieces{});
    std::mutex mu;
    void sync();
    // After a failed commit or wait: the fused Merkle launches' arrival
    // counters may hold counts, so the next commit zeroes them first.
    void mark_failed() { ctr_dirty_ = true; }
    // Wait for the engine's work, then free its working buffer (it grows
    // back on the next commit that needs it).
    void trim();
    hipStream_t stream() const { return stream_; }
    Device &device() { return dev_; }

    // Graphs captured so far (tests and benchmarks).
    size_t graphs_captured() const { return graphs_.size(); }
    // Whether this engine may replay launch lists as graphs at all (the
    // server's engines do not: their batches' compositions rarely repeat).
    void set_graphs(bool on) { use_graphs_ = on; }

private:
    void stage(const void *src, size_t bytes, size_t &offset);
    // One launch of a batch: a descriptor list in desc_ and its grid.
    struct Launch {
        size_t offset;      // bytes into desc_
        uint32_t count;
        uint64_t blocks;
        int kind;           // 0 NTT, 1 Merkle (leaves, or fused trees), 2 Merkle level
        int hash;           // NTT: contiguous; Merkle: hash type
        int kernel;         // NTT: ntt_kernel_kind of the list
        uint32_t tile_log;  // Merkle work-group size (log2 lanes)
    };
    void issue(hipStream_t s, const std::vector<Launch> &launches, const LaunchPieces &pieces);
    void drop_graphs();

    Device &dev_;
    hipStream_t stream_ = nullptr;
    bool own_stream_ = false;
    DevBuf scratch_;           // LDE of every job, DIF order
    DevBuf desc_;              // descriptor lists (device)
    DevBuf ctr_;               // arrival counters of the fused Merkle passes (zero between launches)
    bool ctr_dirty_ = false;   // ctr_ must be zeroed before the next launch
    void *pinned_ = nullptr;   // descriptor staging (host, pinned)
    size_t pinned_bytes_ = 0;
    hipEvent_t staged_ = nullptr;  // last staging copy
    // A batch's launches depend on its job shapes and schedules and on the
    // device addresses of the descriptor lists, never on the jobs' buffers
    // (those are read from the staged descriptors).  A launch list issued a
    // third time is captured as a HIP graph (on a short-lived stream that
    // only records) and replayed on stream_ from then on.
    struct GraphExec {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    bool use_graphs_ = true;
    std::map<std::vector<uint64_t>, GraphExec> graphs_;
    std::map<std::vector<uint64_t>, int> seen_;
};

// The device's engine on the device stream (used by zkx_commit_batch).
Engine &default_engine(Device &d);

// Shape key of a job for the device's tuning table.
TuneKey tune_key(const CommitJob &j);
// Times the job's shape on `e` under each distinct schedule (NTT pass split,
// then Merkle work-group size; `iters` runs each of a GPU-filling batch of
// copies, fastest kept), records the best for the shape on the device and
// returns it.  The job's outputs are written (every schedule computes the
// same result); the engine's own working buffer grows to the batch's size.
// Call with e.mu held.
Tuning autotune(Engine &e, const CommitJob &job, int iters);

}  // namespace zkx
