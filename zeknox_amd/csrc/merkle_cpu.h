// Host (CPU) linear Merkle forest behind fill_digests_buf_linear_cpu
// (merkle/merkle.cpp:126-161): runs without any HIP device.
#pragma once
#include <stdint.h>

namespace zkx {
namespace cpu {

// Same layout and arguments as the GPU forest (merkle.h): host pointers;
// throws Error(1, ...) when the sizes are not a linear forest.
void fill_digests_linear(uint64_t *digests, uint64_t *caps, const uint64_t *leaves, uint64_t nd, uint64_t nc,
                         uint64_t nl, uint64_t leaf_size, uint64_t cap_height, uint64_t hash_type);

}  // namespace cpu
}  // namespace zkx
