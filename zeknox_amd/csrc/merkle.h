// Host-side Merkle forest driver (device pointers, enqueued on the device's
// stream; the caller synchronises).
#pragma once
#include "runtime.h"

namespace zkx {

struct MerkleSpec {
    uint64_t *digests = nullptr;       // num_digests x 4
    uint64_t *caps = nullptr;          // num_caps x 4
    const uint64_t *leaves = nullptr;  // element k of leaf t at t*leaf_stride + k*col_stride
    uint64_t num_digests = 0, num_caps = 0, num_leaves = 0;
    uint64_t leaf_size = 0, cap_height = 0, hash_type = 0;
    uint64_t leaf_stride = 0, col_stride = 1;
};

// Throws Error(1, ...) when the sizes are not a linear Merkle forest.
void check_merkle(const MerkleSpec &s);
void merkle_build(Device &d, const MerkleSpec &s);

}  // namespace zkx
