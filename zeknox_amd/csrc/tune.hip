// Autotuner of the commit schedule (see batch.h autotune).
//
// The search space is what changes the work decomposition without changing
// a result: how the DIF stages of the LDE are cut into passes (strided passes
// of at most 4..9 stages, greedy or evenly spread), how many leaves one
// Merkle work-group hashes and reduces in LDS (2^8..2^10) and, for Poseidon
// leaves, whether the MDS layers run on the matrix cores or on the VALU.  Candidates that
// give the same launch plan are timed once.  Times are HIP events on the
// engine's stream around whole commits, fastest of `iters`.
//
// A schedule is timed on a batch of copies of the job large enough to fill
// the GPU (at least 2^20 leaf rows, at most 64 copies): that is how shapes
// run in a batch, and a lone small job would favour work-group sizes that
// only pay off when the GPU is half empty.  The copies share the job's
// outputs (they all write the same values).
#include <algorithm>
#include <vector>

#include "batch.h"

namespace zkx {

TuneKey tune_key(const CommitJob &j)
{
    return TuneKey{j.cols, j.lg_n, j.rate, j.hash_type, j.leaves != nullptr};
}

namespace {

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    Timer()
    {
        ZKX_HIP(hipEventCreate(&a));
        ZKX_HIP(hipEventCreate(&b));
    }
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

double time_commit(Engine &e, const std::vector<CommitJob> &jobs, const Tuning &t, int iters, Timer &tm)
{
    double best = 1e30;
    for (int k = 0; k < iters; k++) {
        ZKX_HIP(hipEventRecord(tm.a, e.stream()));
        e.commit(jobs.data(), jobs.size(), nullptr, &t);
        ZKX_HIP(hipEventRecord(tm.b, e.stream()));
        ZKX_HIP(hipEventSynchronize(tm.b));
        float ms = 0.f;
        ZKX_HIP(hipEventElapsedTime(&ms, tm.a, tm.b));
        best = std::min(best, (double)ms);
    }
    return best;
}

}  // namespace

Tuning autotune(Engine &e, const CommitJob &job, int iters)
{
    check_commit_job(job);
    iters = std::max(iters, 1);
    Device &d = e.device();
    Timer tm;
    const uint32_t lg_rows = job.lg_n + job.rate;
    const size_t copies = lg_rows >= 20 ? 1 : std::min<size_t>(64, (size_t)1 << (20 - lg_rows));
    const std::vector<CommitJob> batch(copies, job);
    Tuning best = d.tuning_default();
    // warm-up: builds the tables the shape needs and sizes the engine's scratch
    e.commit(batch.data(), batch.size(), nullptr, &best);
    double best_ms = time_commit(e, batch, best, iters, tm);

    // NTT pass split over the coefficient blocks (2^lg_n points)
    std::vector<std::vector<uint32_t>> seen{ntt_pass_sizes(job.lg_n, best)};
    for (uint32_t even = 0; even <= 1; even++)
        for (uint32_t smax = 9; smax >= 4; smax--) {
            Tuning c = best;
            c.ntt_strided_max = smax;
            c.ntt_even = even;
            const std::vector<uint32_t> plan = ntt_pass_sizes(job.lg_n, c);
            if (std::find(seen.begin(), seen.end(), plan) != seen.end()) continue;
            seen.push_back(plan);
            const double ms = time_commit(e, batch, c, iters, tm);
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }

    // Merkle work-group size (all sizes give one tile when the tree is small)
    const uint32_t lg_leaves = job.lg_n + job.rate;
    const Tuning ntt_best = best;
    for (uint32_t tl = kMerkleMinTileLog; tl <= (uint32_t)kMerkleMaxTileLog; tl++) {
        if (tl == ntt_best.merkle_tile_log || (tl > lg_leaves && ntt_best.merkle_tile_log >= lg_leaves)) continue;
        Tuning c = ntt_best;
        c.merkle_tile_log = tl;
        const double ms = time_commit(e, batch, c, iters, tm);
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }

    // Poseidon leaves: MDS layers on the matrix cores or all-VALU permutations
    if (job.hash_type == kPoseidon && job.cols > 4) {
        Tuning c = best;
        c.merkle_valu = 1 - best.merkle_valu;
        const double ms = time_commit(e, batch, c, iters, tm);
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }

    std::lock_guard<std::mutex> g(d.tune_mu);
    d.tuned[tune_key(job)] = best;
    return best;
}

}  // namespace zkx
