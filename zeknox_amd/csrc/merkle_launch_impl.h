// Definition of launch_merkle_hash (merkle_launch.h); included only by the
// per-hash translation units.
#pragma once
#include "merkle_kernels.h"
#include "merkle_launch.h"

namespace zkx {

template <int HT>
void launch_merkle_hash(hipStream_t stream, const MerkleArgs *dev_list, uint32_t n, const MerkleArgs &one,
                        uint64_t blocks, uint32_t tile_log_max, uint32_t block0)
{
    const dim3 g((uint32_t)blocks);
    if (tile_log_max == 10)
        hipLaunchKernelGGL((k_merkle_tile<HT, 1024>), g, dim3(1024), 0, stream, dev_list, n, one, block0);
    else if (tile_log_max == 9)
        hipLaunchKernelGGL((k_merkle_tile<HT, 512>), g, dim3(512), 0, stream, dev_list, n, one, block0);
    else hipLaunchKernelGGL((k_merkle_tile<HT, 256>), g, dim3(256), 0, stream, dev_list, n, one, block0);
}

template <int HT>
void launch_merkle_level_hash(hipStream_t stream, const MerkleArgs *dev_list, uint32_t n, const MerkleArgs &one,
                              uint64_t blocks, uint32_t block0)
{
    hipLaunchKernelGGL((k_merkle_level<HT>), dim3((uint32_t)blocks), dim3(kMerkleLevelLanes), 0, stream, dev_list, n,
                       one, block0);
}

}  // namespace zkx
