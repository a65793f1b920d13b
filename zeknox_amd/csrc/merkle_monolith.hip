// Merkle kernels hashing with Monolith-64 (hash type of merkle/merkle.h).
#include "merkle_launch_impl.h"

namespace zkx {

template void launch_merkle_hash<kMonolith>(hipStream_t, const MerkleArgs *, uint32_t, const MerkleArgs &, uint64_t,
                                     uint32_t, uint32_t);
template void launch_merkle_level_hash<kMonolith>(hipStream_t, const MerkleArgs *, uint32_t, const MerkleArgs &,
                                         uint64_t, uint32_t);

}  // namespace zkx
