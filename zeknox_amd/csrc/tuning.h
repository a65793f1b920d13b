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

// A commit shape: (columns, log2 degree, rate bits, hash type, leaf rows kept).
using TuneKey = std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, bool>;

bool valid_tuning(const Tuning &t);

}  // namespace zkx
