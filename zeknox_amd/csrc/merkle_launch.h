// Launchers of the Merkle kernels, one explicit instantiation per hash type
// (merkle_poseidon.hip, merkle_poseidon2.hip, merkle_keccak.hip,
// merkle_monolith.hip: separate translation units, compiled in parallel).
#pragma once
#include <hip/hip_runtime.h>

#include "merkle_types.h"

namespace zkx {

// One launch of `blocks` work-groups of 2^tile_log_max lanes over the n
// descriptors at dev_list (n > 0) or the single descriptor `one` (n == 0),
// running tiles block0 .. block0 + blocks - 1.
template <int HT>
void launch_merkle_hash(hipStream_t stream, const MerkleArgs *dev_list, uint32_t n, const MerkleArgs &one,
                        uint64_t blocks, uint32_t tile_log_max, uint32_t block0);
// One launch of k_merkle_level over the n level descriptors at dev_list
// (n > 0) or the single descriptor `one` (n == 0), running work-groups
// block0 .. block0 + blocks - 1.
template <int HT>
void launch_merkle_level_hash(hipStream_t stream, const MerkleArgs *dev_list, uint32_t n, const MerkleArgs &one,
                              uint64_t blocks, uint32_t block0);

}  // namespace zkx
