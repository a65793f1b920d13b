// Host side of the Merkle forest: argument checks mirroring merkle.cpp:19-124
// and merkle.cu:165-244, the pass schedule of merkle_kernels.h, and launches
// of pass lists (one tree, or many trees of a commit batch).
#include "merkle.h"

#include <algorithm>
#include <string>

#include "merkle_launch.h"

namespace zkx {

namespace {

bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
uint32_t log2u(uint64_t x) { return 63 - __builtin_clzll(x); }

}  // namespace

void check_merkle(const MerkleSpec &s)
{
    if (s.hash_type > kMonolith)
        invalid("hash type " + std::to_string(s.hash_type) +
                " is not provided (Poseidon 0, Keccak 1, Poseidon-BN128 2, Poseidon2 3, Monolith 4)");
    if (!s.caps || !s.leaves) invalid("null cap or leaf buffer");
    if (!is_pow2(s.num_leaves) || !is_pow2(s.num_caps) || s.num_caps > s.num_leaves)
        invalid("leaves and caps must be powers of two with caps <= leaves");
    if (s.leaf_size == 0 || s.leaf_size > 0xFFFFFFFFull) invalid("leaf size out of range");
    if (s.num_caps == s.num_leaves) return;
    if ((1ull << s.cap_height) != s.num_caps) invalid("cap_buf_size must be 2^cap_height");
    if (s.num_digests != 2 * (s.num_leaves - s.num_caps))
        invalid("digests_buf_size must be 2 * (leaves_buf_size - cap_buf_size)");
    if (!s.digests) invalid("null digest buffer");
}

uint32_t plan_merkle(const MerkleSpec &s, MerkleArgs (&passes)[kMerkleMaxPasses], uint32_t tile_log_max)
{
    if (tile_log_max < (uint32_t)kMerkleMinTileLog || tile_log_max > (uint32_t)kMerkleMaxTileLog)
        invalid("Merkle tile size out of range");
    MerkleArgs a{};
    a.leaves = s.leaves;
    a.leaves_out = s.leaves_out;
    a.dig = s.digests;
    a.cap = s.caps;
    a.leaf_size = (uint32_t)s.leaf_size;
    a.leaf_stride = s.leaf_stride;
    a.col_stride = s.col_stride;
    a.leaf_valu = s.leaf_valu;
    if (s.num_caps == s.num_leaves) {
        a.count = s.num_leaves;
        a.lg_w = 0;
        a.tile_log = std::min<uint32_t>(tile_log_max, log2u(a.count));
        passes[0] = a;
        return 1;
    }
    const uint32_t h = log2u(s.num_caps);
    uint32_t lg_w = log2u(s.num_leaves) - h;  // per-subtree width of the input level
    a.sdl = s.num_digests >> h;
    uint32_t n = 0;
    while (true) {
        if (n == (uint32_t)kMerkleMaxPasses) invalid("tree too large");
        a.count = s.num_caps << lg_w;
        a.lg_w = lg_w;
        a.tile_log = std::min<uint32_t>(tile_log_max, log2u(a.count));
        passes[n++] = a;
        if (lg_w <= a.tile_log) break;  // this pass reaches the caps
        lg_w -= a.tile_log;
    }
    return n;
}

void plan_merkle(const MerkleSpec &s, std::vector<MerkleArgs> &passes, uint32_t tile_log_max)
{
    MerkleArgs p[kMerkleMaxPasses];
    const uint32_t n = plan_merkle(s, p, tile_log_max);
    passes.insert(passes.end(), p, p + n);
}

uint64_t fuse_merkle_passes(const MerkleArgs *passes, uint32_t n, MerkleArgs &out)
{
    if (n == 0 || n > (uint32_t)kMerkleMaxPasses) invalid("internal: Merkle pass plan");
    out = passes[0];
    out.n_next = n - 1;
    out.ctr = nullptr;
    uint64_t ctrs = 0;
    for (uint32_t p = 1; p < n; p++) {
        out.next_tile_log[p - 1] = passes[p].tile_log;
        out.next_ctr[p - 1] = (uint32_t)ctrs;
        ctrs += passes[p].count >> passes[p].tile_log;
    }
    if (ctrs > 0xFFFFFFFFull) invalid("tree too large");
    return ctrs;
}

uint64_t fuse_merkle_passes(const std::vector<MerkleArgs> &passes, MerkleArgs &out)
{
    if (passes.size() > (size_t)kMerkleMaxPasses) invalid("internal: Merkle pass plan");
    return fuse_merkle_passes(passes.data(), (uint32_t)passes.size(), out);
}

void launch_merkle_pass(hipStream_t stream, int hash_type, const MerkleArgs *dev_list,
                        uint32_t n, const MerkleArgs &one, uint64_t blocks, uint32_t tile_log_max, uint64_t chunk)
{
    if (blocks == 0) return;
    if (blocks > 0x7FFFFFFFull) invalid("tree too large for one launch");
    if (tile_log_max < (uint32_t)kMerkleMinTileLog || tile_log_max > (uint32_t)kMerkleMaxTileLog)
        invalid("Merkle tile size out of range");
    ProfScope ps("merkle_leaves", stream);
    const uint64_t step = (chunk && n > 0) ? chunk : blocks;
    for (uint64_t b0 = 0; b0 < blocks; b0 += step) {
        const uint64_t nb = std::min(step, blocks - b0);
        switch (hash_type) {
        case kPoseidon2: launch_merkle_hash<kPoseidon2>(stream, dev_list, n, one, nb, tile_log_max, (uint32_t)b0); break;
        case kKeccak: launch_merkle_hash<kKeccak>(stream, dev_list, n, one, nb, tile_log_max, (uint32_t)b0); break;
        case kMonolith: launch_merkle_hash<kMonolith>(stream, dev_list, n, one, nb, tile_log_max, (uint32_t)b0); break;
        case kPoseidonBN128:
            launch_merkle_hash<kPoseidonBN128>(stream, dev_list, n, one, nb, tile_log_max, (uint32_t)b0);
            break;
        default: launch_merkle_hash<kPoseidon>(stream, dev_list, n, one, nb, tile_log_max, (uint32_t)b0); break;
        }
        ZKX_HIP(hipGetLastError());
    }
}

MerkleArgs merkle_level_args(const MerkleArgs &leaf, uint32_t lg_w)
{
    MerkleArgs a{};
    a.dig = leaf.dig;
    a.cap = leaf.cap;
    a.sdl = leaf.sdl;
    a.lg_w = lg_w;
    a.count = (leaf.count >> leaf.lg_w) << lg_w;
    return a;
}

void launch_merkle_level(hipStream_t stream, int hash_type, const MerkleArgs *dev_list, uint32_t n, uint64_t blocks,
                         uint64_t chunk, const MerkleArgs &one)
{
    if (blocks == 0) return;
    if (blocks > 0x7FFFFFFFull) invalid("tree level too large for one launch");
    ProfScope ps("merkle_levels", stream);
    const uint64_t step = chunk ? chunk : blocks;
    for (uint64_t b0 = 0; b0 < blocks; b0 += step) {
        const uint64_t nb = std::min(step, blocks - b0);
        switch (hash_type) {
        case kPoseidon2: launch_merkle_level_hash<kPoseidon2>(stream, dev_list, n, one, nb, (uint32_t)b0); break;
        case kKeccak: launch_merkle_level_hash<kKeccak>(stream, dev_list, n, one, nb, (uint32_t)b0); break;
        case kMonolith: launch_merkle_level_hash<kMonolith>(stream, dev_list, n, one, nb, (uint32_t)b0); break;
        case kPoseidonBN128:
            launch_merkle_level_hash<kPoseidonBN128>(stream, dev_list, n, one, nb, (uint32_t)b0);
            break;
        default: launch_merkle_level_hash<kPoseidon>(stream, dev_list, n, one, nb, (uint32_t)b0); break;
        }
        ZKX_HIP(hipGetLastError());
    }
}

void merkle_build(Device &d, const MerkleSpec &s)
{
    check_merkle(s);
    const Tuning tun = d.cabi_merkle_tuning(s.num_leaves, s.hash_type);
    const uint32_t tl = tun.merkle_tile_log;
    MerkleSpec sv = s;
    sv.leaf_valu = tun.merkle_valu;
    MerkleArgs passes[kMerkleMaxPasses];
    const uint32_t np = plan_merkle(sv, passes, tl);
    if (s.num_caps < s.num_leaves && lone_tree_levels_apart(s.num_leaves, s.hash_type, d.lone_level_leaves.load())) {
        // a large tree: the leaf digests in one launch, then one launch per
        // level (k_merkle_level), each level's descriptor passed by value
        MerkleArgs leaf = passes[0];
        leaf.leaves_only = 1;
        launch_merkle_pass(d.stream, (int)s.hash_type, nullptr, 0, leaf, leaf.count >> leaf.tile_log, tl);
        for (uint32_t w = leaf.lg_w; w-- > 0;) {
            const MerkleArgs lv = merkle_level_args(leaf, w);
            launch_merkle_level(d.stream, (int)s.hash_type, nullptr, 0,
                                (lv.count + kMerkleLevelLanes - 1) / kMerkleLevelLanes, 0, lv);
        }
        return;
    }
    MerkleArgs f;
    const uint64_t nctr = fuse_merkle_passes(passes, np, f);
    f.tile_begin = 0;
    if (!nctr) {
        launch_merkle_pass(d.stream, (int)s.hash_type, nullptr, 0, f, f.count >> f.tile_log, tl);
        return;
    }
    // fused upper passes read their fields from a descriptor in device memory
    f.ctr = d.merkle_counters(nctr);
    MerkleArgs *dj = d.merkle_descriptor();
    ZKX_HIP(hipMemcpyAsync(dj, &f, sizeof(f), hipMemcpyHostToDevice, d.stream));
    launch_merkle_pass(d.stream, (int)s.hash_type, dj, 1, MerkleArgs{}, f.count >> f.tile_log, tl);
}

}  // namespace zkx
