// Schedule parameters of one commit shape, and the per-device table the
// autotuner fills.  None of them changes a result: every setting computes
// the same LDE, digests and caps; they only change how the work is cut into
// launches and work-groups.
#pragma once
#include <stdint.h>

#include <tuple>

namespace zkx {

struct Tuning {
    // NTT: at most this many DIF stages per strided (column-interleaved)
    // pass, 1..9; the last pass takes up to 12 contiguous stages.
    uint32_t ntt_strided_max = 9;
    // NTT: 0 = strided passes as large as allowed and the rest in the last
    // pass; 1 = the stages spread evenly over the fewest passes.
    uint32_t ntt_even = 0;
    // Merkle: log2 lanes (leaves) per work-group tile, 8..10.
    uint32_t merkle_tile_log = 8;
    // Merkle, Poseidon leaves: 0 = the MDS layers of whole-wave leaf
    // permutations on the i8 matrix cores; 1 = every permutation on the VALU.
    uint32_t merkle_valu = 0;
};

// Work-group size (log2 lanes) of the C-ABI Merkle entry points for one tree
// of num_leaves leaves while the caller has not set a default schedule.  A
// tree alone has no other tree's leaves to hide the upper levels, which run
// at low occupancy at the end of the launch, so the largest tile that still
// gives every CU a work-group (2^lg / 2^tile >= 256) wins up to 2^18 leaves;
// above, leaves take several rounds of work-groups and 2^9 lanes win for
// Poseidon and Poseidon2, 2^8 for Keccak-256 and Monolith, 2^10 for Poseidon
// over BN254 (as measured; the cause is not analysed).  Measured on one
// MI355X, leaf launch of 135-element leaves, cap 4, tiles 2^8 / 2^9 / 2^10:
//   Poseidon  2^16 1.07 / 1.31 / 2.09 ms, 2^17 1.51 / 1.38 / 2.12,
//             2^18 2.65 / 2.52 / 2.41, 2^20 9.32 / 9.08 / 9.61;
//   Keccak    2^17 0.39 / 0.33 / 0.45, 2^18 0.57 / 0.58 / 0.50,
//             2^20 1.82 / 1.97 / 1.98 (Poseidon2, Monolith alike);
//   BN254     2^19 178 / 181 / 145, 2^20 324 / 359 / 281.
// In a batch other trees' leaves hide that tail and the batch default or
// the autotuned schedule applies.  hash_type: merkle/merkle.h numbering.
inline uint32_t lone_tree_tile_log(uint64_t num_leaves, uint64_t hash_type)
{
    const uint32_t lg = num_leaves ? 63u - (uint32_t)__builtin_clzll(num_leaves) : 0u;
    if (lg <= 16) return 8;
    if (lg <= 18) return lg - 8;
    if (hash_type == 2) return 10;                   // Poseidon over BN254
    return hash_type == 0 || hash_type == 3 ? 9 : 8;  // Poseidon, Poseidon2
}

// A commit shape: (columns, log2 degree, rate bits, hash type, leaf rows kept).
using TuneKey = std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, bool>;

bool valid_tuning(const Tuning &t);

// Whether one tree through the C-ABI Merkle entry points builds its levels
// with one launch per level (after a leaf-digests-only launch) rather than
// inside its leaf launch, given the device threshold `min_leaves` (default
// 2^20): Poseidon and Poseidon2 trees switch at min_leaves, Keccak-256 and
// Monolith at min_leaves / 4, Poseidon-BN128 at min_leaves / 32 (0: always).
// Measured, fused against by level, 135-element leaves, cap 4, one GPU, ms:
//   leaves  Poseidon      Poseidon2     Keccak-256     Monolith       Poseidon-BN128
//   2^15    0.90 / 1.10   1.10 / 1.27   0.248 / 0.251  0.540 / 0.533  36.2 / 33.5
//   2^17    1.30 / 1.48   1.43 / 1.59   0.330 / 0.333  0.695 / 0.697  48.5 / 46.2
//   2^18    2.32 / 2.45   2.34 / 2.42   0.503 / 0.485  1.078 / 1.056  75.3 / 72.8
//   2^19    4.33 / 4.34   4.02 / 4.01   1.095 / 0.779  2.069 / 1.760
//   2^20    8.28 / 8.16   7.34 / 7.12   1.928 / 1.344  3.639 / 3.143
// Poseidon's fused levels (rows of 16 lanes) are the fastest of the hashes';
// Keccak's and Monolith's permutations are short, so their fused trees'
// tops, one permutation's latency per level on ever fewer lanes, weigh
// more; Poseidon-BN128's levels are slow enough that a full-GPU launch per
// level pays at every size measured.
inline bool lone_tree_levels_apart(uint64_t num_leaves, uint64_t hash_type, uint64_t min_leaves)
{
    const uint32_t shift = hash_type == 1 || hash_type == 4 ? 2 : hash_type == 2 ? 5 : 0;
    return num_leaves >= (min_leaves >> shift);
}

}  // namespace zkx
