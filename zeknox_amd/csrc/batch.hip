// Batched LDE + Merkle commit engine (see batch.h).
#include "batch.h"

#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <string>

namespace zkx {

namespace {

constexpr size_t kAlign = 256;
constexpr size_t kMaxGraphs = 64;  // captured launch lists kept per engine
// A launch list is captured on its third sight: a capture and instantiation
// costs host time worth many replays, so a list seen twice is not enough.
constexpr int kCaptureAfter = 3;

size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

uint64_t lde_elems(const CommitJob &j) { return (uint64_t)j.cols << (j.lg_n + j.rate); }

}  // namespace

void check_commit_job(const CommitJob &j)
{
    if (j.hash_type > kMonolith)
        invalid("commit job: hash type must be 0 (Poseidon), 1 (Keccak), 2 (Poseidon-BN128), 3 (Poseidon2) or "
                "4 (Monolith)");
    if (j.cols == 0) invalid("commit job: no columns");
    if (j.lg_n == 0 || j.lg_n + j.rate > 30) invalid("commit job: log degree out of range");
    if (j.rate < 1 || j.rate > kMaxRateBits) invalid("commit job: rate_bits must be in [1, 5]");
    if (j.cap_height > j.lg_n + j.rate) invalid("commit job: cap height above the tree height");
    if (!j.coeffs || !j.caps) invalid("commit job: null coefficient or cap buffer");
    if (j.cap_height < j.lg_n + j.rate && !j.digests) invalid("commit job: null digest buffer");
}

size_t commit_scratch_bytes(const CommitJob *jobs, size_t n)
{
    size_t s = 0;
    for (size_t i = 0; i < n; i++) s += align_up(lde_elems(jobs[i]) * sizeof(uint64_t));
    return s;
}

Engine::Engine(Device &d, hipStream_t stream) : dev_(d), stream_(stream)
{
    ZKX_HIP(hipSetDevice(d.id));
    if (!stream_) {
        ZKX_HIP(hipStreamCreate(&stream_));
        own_stream_ = true;
    }
    ZKX_HIP(hipEventCreateWithFlags(&staged_, hipEventDisableTiming));
}

Engine::~Engine()
{
    (void)hipSetDevice(dev_.id);
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graphs();
    if (staged_) (void)hipEventDestroy(staged_);
    if (pinned_) (void)hipHostFree(pinned_);
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

void Engine::sync() { ZKX_HIP(hipStreamSynchronize(stream_)); }

void Engine::trim()
{
    ZKX_HIP(hipSetDevice(dev_.id));
    sync();
    scratch_ = DevBuf();
}

void Engine::commit(const CommitJob *jobs, size_t n, void *scratch, const Tuning *forced, uint64_t *const *lde_bufs,
                    const LaunchPieces &pieces)
{
    if (n == 0) return;
    for (size_t i = 0; i < n; i++) check_commit_job(jobs[i]);
    if (forced && !valid_tuning(*forced)) invalid("schedule parameters out of range");
    ZKX_HIP(hipSetDevice(dev_.id));

    // ---- schedule parameters of each job's shape: forced, the device's
    // entry for the shape, or (autotuning on, shape unseen) tuned now
    std::vector<Tuning> tun(n);
    for (size_t i = 0; i < n; i++) {
        if (forced) {
            tun[i] = *forced;
            continue;
        }
        bool found = false;
        tun[i] = dev_.tuning_for(tune_key(jobs[i]), &found);
        if (!found && dev_.autotune_new_shapes) tun[i] = autotune(*this, jobs[i], 3);
    }

    // ---- working memory: every job's LDE in DIF (bit-reversed) order
    std::vector<uint64_t *> lde(n);
    if (lde_bufs) {
        for (size_t i = 0; i < n; i++) {
            if (!lde_bufs[i] || (reinterpret_cast<uintptr_t>(lde_bufs[i]) & (kAlign - 1)))
                invalid("commit: working buffers must be 256-byte aligned");
            lde[i] = lde_bufs[i];
        }
    } else {
        const size_t need = commit_scratch_bytes(jobs, n);
        if (!scratch && scratch_.bytes < need) {
            sync();
            scratch_ = DevBuf(need);
        }
        char *base = static_cast<char *>(scratch ? scratch : scratch_.p);
        size_t off = 0;
        for (size_t i = 0; i < n; i++) {
            lde[i] = reinterpret_cast<uint64_t *>(base + off);
            off += align_up(lde_elems(jobs[i]) * sizeof(uint64_t));
        }
    }

    // ---- resident tables (built on the device stream, which this engine's
    // stream is ordered after: the null-stream-blocking device stream is
    // synchronised before first use)
    std::vector<const uint64_t *> tw(n), F(n);
    const uint64_t *tw12 = nullptr;
    const uint64_t *const *tw_lg = nullptr;
    {
        std::lock_guard<std::mutex> g(dev_.mu);
        bool built = dev_.twiddles[kNttTileLog] == nullptr || !dev_.twiddle_ptrs.p;
        tw12 = dev_.twiddle_table(kNttTileLog);
        uint32_t lg_max = 0;
        for (size_t i = 0; i < n; i++) lg_max = std::max(lg_max, jobs[i].lg_n + jobs[i].rate);
        for (uint32_t k = 1; k <= lg_max; k++) built |= dev_.twiddles[k] == nullptr;
        tw_lg = dev_.twiddle_tables_upto(lg_max);
        for (size_t i = 0; i < n; i++) {
            const uint32_t lg = jobs[i].lg_n + jobs[i].rate;
            tw[i] = dev_.twiddle_table(lg);
            const size_t before = dev_.lde_factors.size();
            F[i] = lde_factor_table(dev_, jobs[i].lg_n, jobs[i].rate, false, true);
            built |= dev_.lde_factors.size() != before;
        }
        if (built) dev_.sync();
    }

    // ---- Merkle schedule: batches whose trees hold many leaves build the
    // levels above the leaves with one launch per level over every tree
    // (after a leaf launch that writes the leaf digests only); smaller ones
    // finish every tree inside its hash kind's leaf launch
    uint64_t tree_leaves = 0;
    for (size_t i = 0; i < n; i++)
        if (jobs[i].cap_height < jobs[i].lg_n + jobs[i].rate) tree_leaves += 1ull << (jobs[i].lg_n + jobs[i].rate);
    const bool by_level = tree_leaves > 0 && tree_leaves >= dev_.level_launch_leaves.load();

    // ---- plans
    std::vector<std::vector<NttPass>> ntt_plan(n);
    std::vector<uint64_t> ntt_tiles(n);
    std::vector<MerkleArgs> mk(n);  // every pass of a tree in one descriptor
    std::vector<uint64_t> mk_ctr(n);
    uint64_t total_ctr = 0;
    size_t max_ntt = 0;
    for (size_t i = 0; i < n; i++) {
        const CommitJob &j = jobs[i];
        const uint32_t lg = j.lg_n + j.rate;
        ntt_tiles[i] = plan_dif(ntt_plan[i], lde[i], (uint64_t)j.cols << j.rate, j.lg_n, lg, tw[i], tw12, tw_lg, false,
                                kSrcLde, j.coeffs, j.rate, false, F[i], tun[i]);
        MerkleSpec s;
        s.digests = j.digests;
        s.caps = j.caps;
        s.leaves = lde[i];
        s.leaves_out = j.leaves;
        s.num_leaves = 1ull << lg;
        s.num_caps = 1ull << j.cap_height;
        s.num_digests = s.num_caps == s.num_leaves ? 0 : 2 * (s.num_leaves - s.num_caps);
        s.leaf_size = j.cols;
        s.cap_height = j.cap_height;
        s.hash_type = j.hash_type;
        s.leaf_stride = 1;
        s.col_stride = 1ull << lg;
        s.leaf_valu = tun[i].merkle_valu;
        check_merkle(s);
        std::vector<MerkleArgs> passes;
        plan_merkle(s, passes, tun[i].merkle_tile_log);
        mk_ctr[i] = total_ctr;
        if (by_level) {
            mk[i] = passes[0];  // the leaf pass alone: its levels are launched apart
            mk[i].leaves_only = 1;
        } else {
            total_ctr += fuse_merkle_passes(passes, mk[i]);
        }
        max_ntt = std::max(max_ntt, ntt_plan[i].size());
    }

    // ---- descriptor lists: NTT pass k of every job; every tree of each hash
    // kind and work-group size.  All staged with one host->device copy.
    std::vector<Launch> launches;
    std::vector<char> host;
    auto push = [&](const void *p, size_t bytes) {
        const size_t off = host.size();
        host.resize(off + bytes);
        memcpy(host.data() + off, p, bytes);
        return off;
    };
    // arrival counters of the fused upper Merkle passes: zero between launches
    if (ctr_.bytes < total_ctr * sizeof(uint32_t)) {
        sync();
        ctr_ = DevBuf(std::max<size_t>(total_ctr * sizeof(uint32_t), 4096));
        ctr_dirty_ = true;
    }
    if (ctr_dirty_ && ctr_.p) {
        ZKX_HIP(hipMemsetAsync(ctr_.p, 0, ctr_.bytes, stream_));
        ctr_dirty_ = false;
    }
    // NTT pass k of every job, one launch per (contiguity, kernel kind):
    // jobs whose passes the register-round kernels take (ntt_kernel_kind of
    // the single pass) are not held back by others in the same launch (a
    // launch uses such a kernel only if all its passes qualify alike)
    for (size_t k = 0; k < max_ntt; k++) {
        std::map<std::pair<int, int>, std::vector<size_t>> groups;
        for (size_t i = 0; i < n; i++) {
            if (k >= ntt_plan[i].size()) continue;
            const NttPass &p = ntt_plan[i][k];
            groups[{p.s_lo == 0 ? 1 : 0, ntt_kernel_kind(&p, 1)}].push_back(i);
        }
        for (const auto &grp : groups) {
            std::vector<NttPass> list;
            uint64_t blocks = 0;
            for (size_t i : grp.second) {
                NttPass p = ntt_plan[i][k];
                p.tile_begin = blocks;
                blocks += ntt_tiles[i];
                list.push_back(p);
            }
            const size_t off = push(list.data(), list.size() * sizeof(NttPass));
            launches.push_back({off, (uint32_t)list.size(), blocks, 0, grp.first.first, grp.first.second, 0});
            host.resize(align_up(host.size()));
        }
    }
    for (int ht : {kPoseidon, kKeccak, kPoseidonBN128, kPoseidon2, kMonolith})
        for (uint32_t tl = kMerkleMinTileLog; tl <= (uint32_t)kMerkleMaxTileLog; tl++) {
            std::vector<MerkleArgs> list;
            uint64_t blocks = 0;
            for (size_t i = 0; i < n; i++) {
                if ((int)jobs[i].hash_type != ht || tun[i].merkle_tile_log != tl) continue;
                MerkleArgs a = mk[i];
                a.tile_begin = blocks;
                a.ctr = by_level ? nullptr : ctr_.as<uint32_t>() + mk_ctr[i];
                blocks += a.count >> a.tile_log;
                list.push_back(a);
            }
            if (list.empty()) continue;
            const size_t off = push(list.data(), list.size() * sizeof(MerkleArgs));
            launches.push_back({off, (uint32_t)list.size(), blocks, 1, ht, 0, tl});
            host.resize(align_up(host.size()));
        }
    if (by_level) {
        // level k above the leaves of every tree that has it, one launch per
        // hash kind; mk[i].lg_w is the log2 width of tree i's leaf level per
        // subtree (0: caps == leaves, no levels)
        uint32_t depth = 0;
        for (size_t i = 0; i < n; i++) depth = std::max(depth, mk[i].lg_w);
        for (uint32_t k = 1; k <= depth; k++)
            for (int ht : {kPoseidon, kKeccak, kPoseidonBN128, kPoseidon2, kMonolith}) {
                std::vector<MerkleArgs> list;
                uint64_t blocks = 0;
                for (size_t i = 0; i < n; i++) {
                    if ((int)jobs[i].hash_type != ht || mk[i].lg_w < k) continue;
                    MerkleArgs a = merkle_level_args(mk[i], mk[i].lg_w - k);  // the level written (0: the caps)
                    a.tile_begin = blocks;
                    blocks += (a.count + kMerkleLevelLanes - 1) / kMerkleLevelLanes;
                    list.push_back(a);
                }
                if (list.empty()) continue;
                const size_t off = push(list.data(), list.size() * sizeof(MerkleArgs));
                launches.push_back({off, (uint32_t)list.size(), blocks, 2, ht, 0, 0});
                host.resize(align_up(host.size()));
            }
    }
    size_t off0 = 0;
    const void *desc_before = desc_.p;
    stage(host.data(), host.size(), off0);
    (void)off0;
    if (desc_.p != desc_before) drop_graphs();  // captured launches read the old lists

    // ---- launches, in dependency order on one stream (running the second
    // half of a batch's NTT passes beside the first half's hashing on a
    // second stream was measured 7% slower: the co-resident NTT work-groups
    // take the hash kernel's occupancy).  Per-kernel timing (profiling())
    // needs events around each launch, so it runs the launches directly.
    try {
        if (use_graphs_ && dev_.commit_graphs && !profiling() && !forced) {  // (the autotuner times direct launches)
            std::vector<uint64_t> key;
            key.reserve(3 + 7 * launches.size());
            key.push_back((uint64_t)(uintptr_t)desc_.p);
            key.push_back(pieces.ntt_tiles);
            key.push_back(pieces.merkle_groups);
            for (const Launch &l : launches)
                key.insert(key.end(), {(uint64_t)l.offset, l.count, l.blocks, (uint64_t)l.kind, (uint64_t)l.hash,
                                       (uint64_t)l.kernel, l.tile_log});
            auto it = graphs_.find(key);
            if (it == graphs_.end() && seen_.size() >= 64 * kMaxGraphs) seen_.clear();
            if (it == graphs_.end() && ++seen_[key] >= kCaptureAfter) {
                if (graphs_.size() >= kMaxGraphs) drop_graphs();
                // captured on a stream of its own that only records and is
                // destroyed right after: no stream is kept for it
                hipStream_t cap = nullptr;
                ZKX_HIP(hipStreamCreate(&cap));
                GraphExec ge;
                hipError_t ce = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
                if (ce == hipSuccess) {
                    try {
                        issue(cap, launches, pieces);
                    } catch (...) {
                        hipGraph_t g = nullptr;
                        (void)hipStreamEndCapture(cap, &g);
                        if (g) (void)hipGraphDestroy(g);
                        (void)hipStreamDestroy(cap);
                        throw;
                    }
                    ce = hipStreamEndCapture(cap, &ge.graph);
                }
                (void)hipStreamDestroy(cap);
                if (ce != hipSuccess) throw_hip(ce, "stream capture of a batch", __FILE__, __LINE__);
                const hipError_t e = hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0);
                if (e != hipSuccess) {
                    (void)hipGraphDestroy(ge.graph);
                    throw_hip(e, "hipGraphInstantiate", __FILE__, __LINE__);
                }
                it = graphs_.emplace(key, ge).first;
            }
            if (it != graphs_.end()) {
                ZKX_HIP(hipGraphLaunch(it->second.exec, stream_));
                return;
            }
        }
        issue(stream_, launches, pieces);
    } catch (...) {
        ctr_dirty_ = true;
        throw;
    }
}

void Engine::issue(hipStream_t s, const std::vector<Launch> &launches, const LaunchPieces &pieces)
{
    for (const Launch &l : launches) {
        const char *base = static_cast<const char *>(desc_.p) + l.offset;
        if (l.kind == 0) {
            launch_ntt_list(s, reinterpret_cast<const NttPass *>(base), l.count, l.blocks, l.hash == 1, l.kernel,
                            "batch_ntt", pieces.ntt_tiles);
        } else if (l.kind == 1) {
            launch_merkle_pass(s, l.hash, reinterpret_cast<const MerkleArgs *>(base), l.count, MerkleArgs{},
                               l.blocks, l.tile_log, pieces.merkle_groups);
        } else {
            launch_merkle_level(s, l.hash, reinterpret_cast<const MerkleArgs *>(base), l.count, l.blocks,
                                pieces.merkle_groups);
        }
    }
}

void Engine::drop_graphs()
{
    for (auto &kv : graphs_) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    graphs_.clear();
    seen_.clear();
}

Engine &default_engine(Device &d)
{
    static std::mutex mu;
    static std::map<int, std::unique_ptr<Engine>> engines;
    std::lock_guard<std::mutex> g(mu);
    auto &e = engines[d.id];
    if (!e) e = std::make_unique<Engine>(d, d.stream);
    return *e;
}

void Engine::stage(const void *src, size_t bytes, size_t &offset)
{
    // the previous staging copy must have left the pinned buffer
    ZKX_HIP(hipEventSynchronize(staged_));
    if (pinned_bytes_ < bytes) {
        if (pinned_) ZKX_HIP(hipHostFree(pinned_));
        pinned_bytes_ = std::max(bytes, (size_t)1 << 16);
        ZKX_HIP(hipHostMalloc(&pinned_, pinned_bytes_, hipHostMallocDefault));
    }
    if (desc_.bytes < bytes) {
        sync();
        desc_ = DevBuf(std::max(bytes, (size_t)1 << 16));
    }
    memcpy(pinned_, src, bytes);
    ZKX_HIP(hipMemcpyAsync(desc_.p, pinned_, bytes, hipMemcpyHostToDevice, stream_));
    ZKX_HIP(hipEventRecord(staged_, stream_));
    offset = 0;
}

}  // namespace zkx
