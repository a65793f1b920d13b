// Merkle pass descriptors shared by the kernels (merkle_kernels.h) and the
// host drivers (merkle.hip, batch.hip).
#pragma once
#include <stdint.h>

namespace zkx {

// merkle/merkle.h:21-28 hash type numbers
enum HashKind : int { kPoseidon = 0, kKeccak = 1, kPoseidonBN128 = 2, kPoseidon2 = 3, kMonolith = 4 };

// Work-group sizes of the Merkle kernels: 2^8, 2^9 or 2^10 lanes, one leaf
// (or node) per lane.
constexpr int kMerkleMinTileLog = 8;
constexpr int kMerkleMaxTileLog = 10;

struct MerkleArgs {
    const uint64_t *leaves;  // first pass: leaf source
    uint64_t *leaves_out;    // first pass, optional: leaf rows [t][leaf_size]
    uint64_t *dig;
    uint64_t *cap;
    uint64_t count;          // nodes (or leaves) at the input level, all subtrees
    uint64_t sdl;            // digests per subtree
    uint64_t leaf_stride;    // u64 between consecutive leaves
    uint64_t col_stride;     // u64 between consecutive elements of a leaf
    uint64_t tile_begin;     // first work-group of this job in the launch
    uint32_t lg_w;           // log2 width of the input level per subtree (0: caps == leaves)
    uint32_t tile_log;       // log2 nodes per work-group tile
    uint32_t leaf_size;
    uint32_t pad_;
};

}  // namespace zkx
