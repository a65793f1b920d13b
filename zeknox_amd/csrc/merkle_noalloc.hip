// fill_digests_buf_linear_gpu_with_gpu_ptr for callers that cannot allocate
// device memory (code under stream capture, tight serving loops, harnesses
// that forbid hipMalloc).  The one-launch Merkle forest needs scratch beyond
// its inputs and outputs: a pass descriptor that the fused upper passes read,
// and one arrival counter per group of tiles of every upper pass.  The
// default entry point keeps both in per-device library buffers grown with
// hipMalloc on demand; here they live in a file-scope device array of fixed
// size, and the entry point allocates nothing.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>

#include "../../include/zeknox_ext.h"
#include "merkle.h"

namespace zkx {

namespace {

// Scratch layout: [descriptor | arrival counters].
constexpr size_t kDescBytes = 256;
constexpr size_t kScratchBytes = ZKX_NOALLOC_SCRATCH_BYTES;
constexpr uint64_t kMaxCounters = (kScratchBytes - kDescBytes) / sizeof(uint32_t);
static_assert(sizeof(MerkleArgs) <= kDescBytes, "Merkle descriptor does not fit its scratch slot");

// Each GPU the library runs on holds its own copy until the process ends.
// In host code this name is only a handle for hipGetSymbolAddress: kernels
// get the device address kept in g_scratch[device].
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
// Device address of `scratch` per device, set on the first call on that
// device (under the device's lock) and never changed afterwards.
unsigned char *g_scratch[kMaxDevices] = {};

// Writes the launch's descriptor and zeroes its arrival counters: every call
// starts from zero whatever an earlier (possibly failed) launch left behind.
// A fill kernel rather than hipMemsetAsync, so that captured streams record
// kernel nodes only.
__global__ void k_noalloc_setup(MerkleArgs *desc, MerkleArgs f, uint32_t *ctr, uint32_t nctr)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *desc = f;
    for (uint32_t k = i; k < nctr; k += gridDim.x * blockDim.x) ctr[k] = 0;
}

unsigned char *scratch_of(Device &d)
{
    if (d.id < 0 || d.id >= kMaxDevices) invalid("device index beyond the scratch table");
    if (!g_scratch[d.id]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) throw_hip(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        if (!p) invalid("hipGetSymbolAddress(scratch) returned null");
        g_scratch[d.id] = static_cast<unsigned char *>(p);
    }
    return g_scratch[d.id];
}

void merkle_noalloc(void *dig, void *cap, void *leaves, uint64_t nd, uint64_t nc, uint64_t nl, uint64_t ls,
                    uint64_t h, uint64_t ht, uint64_t gpu_id)
{
    MerkleSpec s;
    s.num_digests = nd;
    s.num_caps = nc;
    s.num_leaves = nl;
    s.leaf_size = ls;
    s.cap_height = h;
    s.hash_type = ht;
    s.leaf_stride = ls;
    s.col_stride = 1;
    s.digests = static_cast<uint64_t *>(dig);
    s.caps = static_cast<uint64_t *>(cap);
    s.leaves = static_cast<const uint64_t *>(leaves);
    check_merkle(s);
    Device &d = device(gpu_id);
    std::lock_guard<std::mutex> g(d.mu);
    unsigned char *const base = scratch_of(d);
    const Tuning tun = d.cabi_merkle_tuning(s.num_leaves, s.hash_type);
    const uint32_t tl = tun.merkle_tile_log;
    s.leaf_valu = tun.merkle_valu;
    MerkleArgs passes[kMerkleMaxPasses];
    const uint32_t np = plan_merkle(s, passes, tl);
    MerkleArgs f;
    const uint64_t nctr = fuse_merkle_passes(passes, np, f);
    if (nctr > kMaxCounters)
        invalid("tree needs " + std::to_string(nctr) + " arrival counters; the allocation-free entry point holds " +
                std::to_string(kMaxCounters) + " (" + std::to_string(kScratchBytes) +
                " bytes of scratch); use fill_digests_buf_linear_gpu_with_gpu_ptr");
    f.tile_begin = 0;
    if (!nctr) {
        launch_merkle_pass(d.stream, (int)s.hash_type, nullptr, 0, f, f.count >> f.tile_log, tl);
    } else {
        MerkleArgs *desc = reinterpret_cast<MerkleArgs *>(base);
        f.ctr = reinterpret_cast<uint32_t *>(base + kDescBytes);
        const uint32_t blocks = (uint32_t)std::min<uint64_t>((nctr + 255) / 256, 256);
        hipLaunchKernelGGL(k_noalloc_setup, dim3(blocks), dim3(256), 0, d.stream, desc, f, f.ctr, (uint32_t)nctr);
        ZKX_HIP(hipGetLastError());
        launch_merkle_pass(d.stream, (int)s.hash_type, desc, 1, MerkleArgs{}, f.count >> f.tile_log, tl);
    }
    d.sync();
}

}  // namespace

}  // namespace zkx

extern "C" RustError zkx_fill_digests_buf_linear_gpu_noalloc(void *digests_buf, void *cap_buf, void *leaves_buf,
                                                             uint64_t digests_buf_size, uint64_t cap_buf_size,
                                                             uint64_t leaves_buf_size, uint64_t leaf_size,
                                                             uint64_t cap_height, uint64_t hash_type, uint64_t gpu_id)
{
    try {
        zkx::merkle_noalloc(digests_buf, cap_buf, leaves_buf, digests_buf_size, cap_buf_size, leaves_buf_size,
                            leaf_size, cap_height, hash_type, gpu_id);
        return RustError{0, nullptr};
    } catch (const zkx::Error &e) {
        return RustError{e.code, strdup(e.what())};
    } catch (const std::exception &e) {
        return RustError{-1, strdup(e.what())};
    }
}
