// Launchers of the Merkle kernels, one explicit instantiation per hash type
// (merkle_poseidon.hip, merkle_poseidon2.hip, merkle_keccak.hip,
// merkle_monolith.hip: separate translation units, compiled in parallel).
#pragma once
#include <hip/hip_runtime.h>

#include "merkle_types.h"

namespace zkx {

// One launch over `blocks` tiles of 2^tile_log_max lanes of the n
// descriptors at dev_list (n > 0) or the single descriptor `one` (n == 0),
// with min(blocks, max_groups) work-groups (max_groups 0: one per tile).
template <int HT>
void launch_merkle_hash(hipStream_t stream, const MerkleArgs *dev_list, uint32_t n, const MerkleArgs &one,
                        uint64_t blocks, uint32_t tile_log_max, uint64_t max_groups);
// Work-groups of that launch the device holds at once.
template <int HT>
uint64_t merkle_resident_groups(uint32_t tile_log_max);

}  // namespace zkx
