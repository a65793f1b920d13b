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
// Passes of one tree: tiles of >= 2^8 nodes reduce 2^32 leaves in 4.
constexpr int kMerkleMaxPasses = 4;
// Work-group size (nodes) of the one-level-per-launch kernel.
constexpr int kMerkleLevelLanes = 256;

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
    // Upper passes run inside the same launch: the last work-group to finish
    // a group of 2^next_tile_log[p] tiles goes on with the tile of their roots.
    uint32_t n_next;
    uint32_t next_tile_log[kMerkleMaxPasses - 1];
    uint32_t next_ctr[kMerkleMaxPasses - 1];  // first arrival counter of each upper pass
    uint32_t *ctr;           // arrival counters, zero before and after the launch (null: n_next == 0)
    uint32_t leaf_valu;      // Poseidon leaves: 1 = all-VALU permutations (no matrix-core MDS)
    uint32_t leaves_only;    // 1: the launch writes the leaf digests and stops (levels launched apart)
};

}  // namespace zkx
