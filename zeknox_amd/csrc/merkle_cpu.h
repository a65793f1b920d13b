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

// Eight hashes at once (merkle_cpu_x8.hip): true when the hash type has an
// AVX-512 form (Poseidon, Poseidon2), the CPU runs it and ZKX_CPU_NO_AVX512
// is unset or 0 (read on every call).
bool have_x8(uint64_t hash_type);
// out + j * out_stride = H(x + j * stride, n elements), j < 8, n > 4.
void hash_one_x8(uint64_t hash_type, const uint64_t *x, uint64_t stride, uint64_t n, uint64_t *out,
                 uint64_t out_stride);
// out + j * out_stride = H(l_j, l_j + 4) with l_j = l + j * stride, j < 8.
void hash_two_x8(uint64_t hash_type, const uint64_t *l, uint64_t stride, uint64_t *out, uint64_t out_stride);

}  // namespace cpu
}  // namespace zkx
