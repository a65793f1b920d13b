// Definition of launch_merkle_hash (merkle_launch.h); included only by the
// per-hash translation units.
#pragma once
#include <algorithm>

#include "merkle_kernels.h"
#include "merkle_launch.h"

namespace zkx {

template <int HT>
void launch_merkle_hash(hipStream_t stream, const MerkleArgs *dev_list, uint32_t n, const MerkleArgs &one,
                        uint64_t blocks, uint32_t tile_log_max, uint64_t max_groups)
{
    if (n > 0 && max_groups && max_groups < blocks) {
        const dim3 g((uint32_t)max_groups);
        if (tile_log_max == 10)
            hipLaunchKernelGGL((k_merkle_tiles_strided<HT, 1024>), g, dim3(1024), 0, stream, dev_list, n, blocks);
        else if (tile_log_max == 9)
            hipLaunchKernelGGL((k_merkle_tiles_strided<HT, 512>), g, dim3(512), 0, stream, dev_list, n, blocks);
        else hipLaunchKernelGGL((k_merkle_tiles_strided<HT, 256>), g, dim3(256), 0, stream, dev_list, n, blocks);
        return;
    }
    const dim3 g((uint32_t)blocks);
    if (tile_log_max == 10) hipLaunchKernelGGL((k_merkle_tile<HT, 1024>), g, dim3(1024), 0, stream, dev_list, n, one);
    else if (tile_log_max == 9) hipLaunchKernelGGL((k_merkle_tile<HT, 512>), g, dim3(512), 0, stream, dev_list, n, one);
    else hipLaunchKernelGGL((k_merkle_tile<HT, 256>), g, dim3(256), 0, stream, dev_list, n, one);
}

template <int HT>
uint64_t merkle_resident_groups(uint32_t tile_log_max)
{
    int per_cu = 0, dev = 0, cus = 0;
    const int tpb = 1 << tile_log_max;
    hipError_t e;
    if (tile_log_max == 10)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_merkle_tiles_strided<HT, 1024>, tpb, 0);
    else if (tile_log_max == 9)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_merkle_tiles_strided<HT, 512>, tpb, 0);
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_merkle_tiles_strided<HT, 256>, tpb, 0);
    if (e != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return (uint64_t)std::max(per_cu, 1) * (uint64_t)std::max(cus, 1);
}

}  // namespace zkx
