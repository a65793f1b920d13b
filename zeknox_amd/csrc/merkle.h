// Host-side Merkle forest driver (device pointers, enqueued on the device's
// stream; the caller synchronises).
#pragma once
#include <vector>

#include "merkle_types.h"
#include "runtime.h"

namespace zkx {

struct MerkleSpec {
    uint64_t *digests = nullptr;       // num_digests x 4
    uint64_t *caps = nullptr;          // num_caps x 4
    const uint64_t *leaves = nullptr;  // element k of leaf t at t*leaf_stride + k*col_stride
    uint64_t *leaves_out = nullptr;    // optional row-major copy of the leaves
    uint64_t num_digests = 0, num_caps = 0, num_leaves = 0;
    uint64_t leaf_size = 0, cap_height = 0, hash_type = 0;
    uint64_t leaf_stride = 0, col_stride = 1;
    uint32_t leaf_valu = 0;            // Tuning::merkle_valu
};

// Throws Error(1, ...) when the sizes are not a linear Merkle forest.
void check_merkle(const MerkleSpec &s);
// Pass descriptors of one tree, in launch order (tile_begin left 0), with
// tiles of at most 2^tile_log_max nodes (8..10); appended to `passes`, or
// written to a fixed array (returns the pass count; no heap use).
void plan_merkle(const MerkleSpec &s, std::vector<MerkleArgs> &passes, uint32_t tile_log_max);
uint32_t plan_merkle(const MerkleSpec &s, MerkleArgs (&passes)[kMerkleMaxPasses], uint32_t tile_log_max);
// One descriptor that runs every pass of the tree in one launch (the upper
// passes fused into the first); returns the arrival counters it needs
// (its `ctr` is left null for the caller to point at that many zeroed words).
uint64_t fuse_merkle_passes(const std::vector<MerkleArgs> &passes, MerkleArgs &out);
uint64_t fuse_merkle_passes(const MerkleArgs *passes, uint32_t n, MerkleArgs &out);
// One launch over `n` job descriptors in device memory (n > 0), or over the
// single descriptor `one` (n == 0); `blocks` work-groups of 2^tile_log_max
// lanes in total (every descriptor's tile_log <= tile_log_max).
// chunk > 0 (a descriptor list, n > 0): issued as launches of at most
// `chunk` work-groups each, in order on the stream, so kernels of other
// streams can be dispatched between them.
void launch_merkle_pass(hipStream_t stream, int hash_type, const MerkleArgs *dev_list,
                        uint32_t n, const MerkleArgs &one, uint64_t blocks, uint32_t tile_log_max,
                        uint64_t chunk = 0);
// One launch of the level kernel (merkle_kernels.h k_merkle_level) over n
// level descriptors in device memory; `blocks` work-groups of 256 lanes in
// total, issued in ranges of at most `chunk` work-groups when chunk > 0.
void launch_merkle_level(hipStream_t stream, int hash_type, const MerkleArgs *dev_list, uint32_t n, uint64_t blocks,
                         uint64_t chunk = 0, const MerkleArgs &one = MerkleArgs{});
// Level descriptor (k_merkle_level) of the level of width 2^lg_w per subtree
// (0: the caps) of the tree whose leaf pass is `leaf` (plan_merkle's first
// pass); tile_begin 0.
MerkleArgs merkle_level_args(const MerkleArgs &leaf, uint32_t lg_w);
// The whole tree in one launch on the device stream; call with d.mu held.
void merkle_build(Device &d, const MerkleSpec &s);

}  // namespace zkx
