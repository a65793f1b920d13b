// Definition of launch_merkle_hash (merkle_launch.h); included only by the
// per-hash translation units.
#pragma once
#include "merkle_kernels.h"
#include "merkle_launch.h"

namespace zkx {

template <int HT, int TPB>
static void launch_merkle_tpb(hipStream_t stream, bool from_leaves, const MerkleArgs *dev_list, uint32_t n,
                              const MerkleArgs &one, uint64_t blocks)
{
    const dim3 g((uint32_t)blocks), b(TPB);
    if (from_leaves) hipLaunchKernelGGL((k_merkle_tile<HT, true, TPB>), g, b, 0, stream, dev_list, n, one);
    else hipLaunchKernelGGL((k_merkle_tile<HT, false, TPB>), g, b, 0, stream, dev_list, n, one);
}

template <int HT>
void launch_merkle_hash(hipStream_t stream, bool from_leaves, const MerkleArgs *dev_list, uint32_t n,
                        const MerkleArgs &one, uint64_t blocks, uint32_t tile_log_max)
{
    if (tile_log_max == 10) launch_merkle_tpb<HT, 1024>(stream, from_leaves, dev_list, n, one, blocks);
    else if (tile_log_max == 9) launch_merkle_tpb<HT, 512>(stream, from_leaves, dev_list, n, one, blocks);
    else launch_merkle_tpb<HT, 256>(stream, from_leaves, dev_list, n, one, blocks);
}

}  // namespace zkx
