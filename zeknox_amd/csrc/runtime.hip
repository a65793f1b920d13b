// Device registry and resident tables (see runtime.h).
#include "runtime.h"

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <memory>

#include "merkle_types.h"
#include "ntt.h"

namespace zkx {

void throw_hip(hipError_t e, const char *what, const char *file, int line)
{
    char buf[512];
    snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    throw Error(-(int)e, buf);
}

DevBuf::DevBuf(size_t n) : bytes(n)
{
    if (n) ZKX_HIP(hipMalloc(&p, n));
}

DevBuf::~DevBuf()
{
    if (p) (void)hipFree(p);
}

namespace {

struct Registry {
    std::vector<std::unique_ptr<Device>> devs;
    Registry()
    {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        for (int i = 0; i < n; i++) {
            auto d = std::make_unique<Device>();
            d->id = i;
            if (hipGetDeviceProperties(&d->prop, i) != hipSuccess) continue;
            devs.push_back(std::move(d));
        }
    }
};

Registry &registry()
{
    static Registry r;
    return r;
}

thread_local std::string g_last_error;

}  // namespace

int device_count() { return (int)registry().devs.size(); }

Device &device(size_t id)
{
    const uint32_t i = (uint32_t)id;
    auto &r = registry();
    if (i >= r.devs.size()) invalid("device id " + std::to_string(i) + " out of range (" +
                                    std::to_string(r.devs.size()) + " devices)");
    Device &d = *r.devs[i];
    ZKX_HIP(hipSetDevice(d.id));
    // once per device, without holding d.mu (a compute call may hold it);
    // a throwing creation leaves the flag unset for the next caller
    std::call_once(d.stream_once, [&] { ZKX_HIP(hipStreamCreate(&d.stream)); });
    return d;
}

void Device::sync() { ZKX_HIP(hipStreamSynchronize(stream)); }

uint32_t *Device::merkle_counters(uint64_t n)
{
    const size_t bytes = n * sizeof(uint32_t);
    if (merkle_ctr.bytes < bytes) {
        sync();  // no launch may still use the old counters
        merkle_ctr = DevBuf(std::max<size_t>(bytes, 4096));
        merkle_ctr_dirty = true;
    }
    if (merkle_ctr_dirty) {
        ZKX_HIP(hipMemsetAsync(merkle_ctr.p, 0, merkle_ctr.bytes, stream));
        merkle_ctr_dirty = false;
    }
    return merkle_ctr.as<uint32_t>();
}

MerkleArgs *Device::merkle_descriptor()
{
    if (!merkle_desc.p) merkle_desc = DevBuf(sizeof(MerkleArgs));
    return merkle_desc.as<MerkleArgs>();
}

bool valid_tuning(const Tuning &t)
{
    return t.ntt_strided_max >= 1 && t.ntt_strided_max <= 9 && t.ntt_even <= 1 && t.merkle_tile_log >= 8 &&
           t.merkle_tile_log <= 10 && t.merkle_valu <= 1;
}

Tuning Device::tuning_default()
{
    std::lock_guard<std::mutex> g(tune_mu);
    return tune_default;
}

Tuning Device::cabi_merkle_tuning(uint64_t num_leaves, uint64_t hash_type)
{
    std::lock_guard<std::mutex> g(tune_mu);
    Tuning t = tune_default;
    if (!tune_default_set) t.merkle_tile_log = lone_tree_tile_log(num_leaves, hash_type);
    return t;
}

Tuning Device::tuning_for(const TuneKey &k, bool *found)
{
    std::lock_guard<std::mutex> g(tune_mu);
    auto it = tuned.find(k);
    if (found) *found = it != tuned.end();
    return it != tuned.end() ? it->second : tune_default;
}

const uint64_t *Device::twiddle_table(uint32_t lg)
{
    if (lg > 32) invalid("log size " + std::to_string(lg) + " exceeds 32");
    if (!twiddles[lg]) {
        const uint64_t n = 1ull << lg;
        uint64_t *p = nullptr;
        ZKX_HIP(hipMalloc(&p, n * sizeof(uint64_t)));
        fill_powers(p, n, omega(lg), stream);
        twiddles[lg] = p;
        if (twiddle_ptrs.p)
            ZKX_HIP(hipMemcpyAsync(twiddle_ptrs.as<uint64_t *>() + lg, &twiddles[lg], sizeof(uint64_t *),
                                   hipMemcpyHostToDevice, stream));
    }
    return twiddles[lg];
}

const uint64_t *const *Device::twiddle_tables_upto(uint32_t lg)
{
    if (!twiddle_ptrs.p) {
        twiddle_ptrs = DevBuf(sizeof(twiddles));
        ZKX_HIP(hipMemcpyAsync(twiddle_ptrs.p, twiddles, sizeof(twiddles), hipMemcpyHostToDevice, stream));
    }
    for (uint32_t k = 1; k <= lg; k++) twiddle_table(k);
    return twiddle_ptrs.as<const uint64_t *const>();
}

void Device::build_coset(uint32_t lg, uint64_t gen)
{
    if (lg > 32) invalid("coset log size " + std::to_string(lg) + " exceeds 32");
    const uint64_t n = 1ull << lg;
    if (coset && (coset_lg < lg || coset_gen != gen)) {
        sync();
        ZKX_HIP(hipFree(coset));
        coset = nullptr;
    }
    if (!coset) {
        ZKX_HIP(hipMalloc(&coset, n * sizeof(uint64_t)));
        coset_lg = lg;
    }
    coset_gen = gen;
    fill_powers(coset, 1ull << coset_lg, gen, stream);
}

// ---------------------------------------------------------------- host field

static constexpr uint64_t P = 0xFFFFFFFF00000001ull;
static constexpr uint64_t OMEGA32 = 1753635133440165772ull;  // gl64_params.hpp:48

uint64_t host_mul(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) % P); }

uint64_t host_pow(uint64_t b, uint64_t e)
{
    uint64_t r = 1;
    b %= P;
    while (e) {
        if (e & 1) r = host_mul(r, b);
        b = host_mul(b, b);
        e >>= 1;
    }
    return r;
}

uint64_t host_inv(uint64_t a) { return host_pow(a, P - 2); }

uint64_t omega(uint32_t k)
{
    uint64_t w = OMEGA32;
    for (uint32_t i = k; i < 32; i++) w = host_mul(w, w);
    return w;
}

void set_last_error(const std::string &m) { g_last_error = m; }
const char *last_error() { return g_last_error.c_str(); }

}  // namespace zkx
