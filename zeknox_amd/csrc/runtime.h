// Device registry, streams, error propagation and per-device resident tables.
//
// Replaces the reference's utils/all_gpus.{hpp,cu} (device registry),
// utils/gpu_t.cuh (streams, copies, temporary buffers) and
// utils/exception.cuh (CUDA_OK -> thrown error, mapped to RustError by the
// entry points) with HIP-native equivalents.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <map>
#include <tuple>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "tuning.h"

namespace zkx {

struct MerkleArgs;

// Thrown inside the library, converted to RustError{code, message} at the
// C boundary.  code 1 = invalid argument; -e = HIP error e.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char *what, const char *file, int line);
[[noreturn]] inline void invalid(const std::string &m) { throw Error(1, m); }

#define ZKX_HIP(call)                                                         \
    do {                                                                      \
        hipError_t zkx_e_ = (call);                                           \
        if (zkx_e_ != hipSuccess) ::zkx::throw_hip(zkx_e_, #call, __FILE__, __LINE__); \
    } while (0)

// A device buffer owned by the library; freed on destruction.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n);
    ~DevBuf();
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            if (p) (void)hipFree(p);
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct Device {
    int id = 0;
    hipStream_t stream = nullptr;  // blocking stream: ordered after the null stream
    std::once_flag stream_once;    // creates `stream` on the first device() call
    hipDeviceProp_t prop{};
    std::mutex mu;                 // serialises table builds and compute calls

    // Resident tables (twiddles per log size, coset powers).
    uint64_t *twiddles[33] = {};   // twiddles[k][i] = w_{2^k}^i, i < 2^k
    // The same 33 table pointers in device memory (null where not built),
    // for kernels that pick a table by log size.
    DevBuf twiddle_ptrs;
    uint64_t *coset = nullptr;     // coset[i] = g^i, i < 2^coset_lg
    uint32_t coset_lg = 0;
    uint64_t coset_gen = 0;
    // LDE factor tables keyed by (lg_n, rate, inverse, generator)
    std::map<std::tuple<uint32_t, uint32_t, bool, uint64_t>, uint64_t *> lde_factors;

    // Zeroed arrival counters of the fused Merkle launches on the device
    // stream (each completed launch leaves them zero); grows on demand, under
    // mu.  A launch that fails part-way may leave counts behind: the caller
    // then sets merkle_ctr_dirty and the next merkle_counters() zeroes them.
    DevBuf merkle_ctr;
    bool merkle_ctr_dirty = false;
    uint32_t *merkle_counters(uint64_t n);
    // One device-side descriptor slot for those launches (written with a
    // stream-ordered copy before each launch), under mu.
    DevBuf merkle_desc;
    struct MerkleArgs *merkle_descriptor();

    const uint64_t *twiddle_table(uint32_t lg);  // builds on first use
    // Builds the tables of every log size 1..lg; returns the device array of
    // table pointers.
    const uint64_t *const *twiddle_tables_upto(uint32_t lg);
    void build_coset(uint32_t lg, uint64_t gen);
    void sync();

    // Schedule parameters: the default, and per commit shape (set by the
    // caller or found by the autotuner).  Guarded by tune_mu.
    std::mutex tune_mu;
    Tuning tune_default;
    bool tune_default_set = false;  // the caller replaced the built-in default
    std::map<TuneKey, Tuning> tuned;
    bool autotune_new_shapes = false;  // tune unseen shapes in zkx_commit_batch
    // Batch engines replay launch lists they have issued before as HIP
    // graphs (zkx_set_commit_graphs); read at each commit.
    std::atomic<bool> commit_graphs{true};
    // Batches whose trees hold at least this many leaves in total build the
    // levels above the leaves with one launch per level over all trees
    // (zkx_set_level_launches); smaller ones, one fused launch per hash kind.
    // Measured on 1-16 commits of 135 x 2^15 rate 3 and 20 x 2^12 rate 2
    // (Poseidon, Poseidon2): by level 0.93-0.99 of the fused time from 2^18
    // leaves up, 1.14-1.16 below (each level launch costs at least one
    // permutation's latency, about 43 us).
    std::atomic<uint64_t> level_launch_leaves{1ull << 18};
    // The same for one tree through the C-ABI Merkle entry points
    // (fill_digests_buf_linear_gpu_with_gpu_ptr; the allocation-free entry
    // point always runs one fused launch); Keccak-256 and Monolith trees
    // from a quarter of this, Poseidon-BN128 trees from a 32nd
    // (lone_tree_levels_apart): a lone tree's narrow levels are
    // latency-bound either way, so only wide trees gain.
    std::atomic<uint64_t> lone_level_leaves{1ull << 20};
    Tuning tuning_default();
    // The schedule of the C-ABI Merkle entry points for one tree: the
    // default, with lone_tree_tile_log()'s work-groups unless the caller set
    // a default.
    Tuning cabi_merkle_tuning(uint64_t num_leaves, uint64_t hash_type);
    // The shape's entry, or the default (found = false).
    Tuning tuning_for(const TuneKey &k, bool *found = nullptr);
};

int device_count();
// Makes the device current and returns it; id is taken modulo 2^32 (Rust
// callers pass i32 in a size_t slot).
Device &device(size_t id);

// Primitive 2^k-th root of unity (gl64_params.hpp OMEGA[k]).
uint64_t omega(uint32_t k);
uint64_t host_mul(uint64_t a, uint64_t b);
uint64_t host_pow(uint64_t b, uint64_t e);
uint64_t host_inv(uint64_t a);

// Last error recorded by a void entry point (thread-local).
void set_last_error(const std::string &m);
const char *last_error();

// Optional per-kernel timing: a ProfScope around a launch records HIP events
// on the launching stream when profiling is enabled (zkx_profile_enable).
bool profiling();
struct ProfScope {
    hipEvent_t e0 = nullptr;
    const char *name;
    hipStream_t stream;
    ProfScope(const char *kernel_family, hipStream_t s);
    ~ProfScope();
};

}  // namespace zkx
