// Host side of the Merkle forest: argument checks mirroring merkle.cpp:19-124
// and merkle.cu:165-244, and the pass schedule of merkle_kernels.h.
#include "merkle.h"

#include <algorithm>
#include <string>

#include "merkle_kernels.h"

namespace zkx {

namespace {

bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
uint32_t log2u(uint64_t x) { return 63 - __builtin_clzll(x); }

template <int HT>
void launch_tree(Device &d, const MerkleSpec &s)
{
    MerkleArgs a{};
    a.leaves = s.leaves;
    a.dig = s.digests;
    a.cap = s.caps;
    a.leaf_size = (uint32_t)s.leaf_size;
    a.leaf_stride = s.leaf_stride;
    a.col_stride = s.col_stride;
    if (s.num_caps == s.num_leaves) {
        a.count = s.num_leaves;
        const uint64_t blocks = (a.count + kMerkleThreads - 1) / kMerkleThreads;
        ProfScope ps("merkle_leaves", d.stream);
        hipLaunchKernelGGL(k_leaf_direct<HT>, dim3((uint32_t)blocks), dim3(kMerkleThreads), 0, d.stream, a);
        ZKX_HIP(hipGetLastError());
        return;
    }
    const uint32_t h = log2u(s.num_caps);
    uint32_t lg_w = log2u(s.num_leaves) - h;  // per-subtree width of the input level
    a.sdl = s.num_digests >> h;
    bool first = true;
    while (true) {
        const uint64_t count = s.num_caps << lg_w;
        const uint32_t tile_log = std::min<uint32_t>(8, log2u(count));
        a.count = count;
        a.lg_w = lg_w;
        a.tile_log = tile_log;
        const uint64_t blocks = count >> tile_log;
        if (blocks > 0x7FFFFFFFull) invalid("tree too large for one launch");
        ProfScope ps(first ? "merkle_leaves" : "merkle_levels", d.stream);
        if (first)
            hipLaunchKernelGGL((k_merkle_tile<HT, true>), dim3((uint32_t)blocks), dim3(kMerkleThreads), 0,
                               d.stream, a);
        else
            hipLaunchKernelGGL((k_merkle_tile<HT, false>), dim3((uint32_t)blocks), dim3(kMerkleThreads), 0,
                               d.stream, a);
        ZKX_HIP(hipGetLastError());
        first = false;
        if (lg_w <= tile_log) break;  // this pass reached the caps
        lg_w -= tile_log;
    }
}

}  // namespace

void check_merkle(const MerkleSpec &s)
{
    if (s.hash_type != 0 && s.hash_type != 3)
        invalid("hash type " + std::to_string(s.hash_type) + " is not provided (Poseidon = 0, Poseidon2 = 3)");
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

void merkle_build(Device &d, const MerkleSpec &s)
{
    check_merkle(s);
    if (s.hash_type == 3) launch_tree<kPoseidon2>(d, s);
    else launch_tree<kPoseidon>(d, s);
}

}  // namespace zkx
