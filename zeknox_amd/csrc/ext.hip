// zeknox_ext.h entry points: memory helpers, last error, kernel timing.
#include <string.h>

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/zeknox_ext.h"
#include "gl64.h"
#include "ntt.h"
#include "runtime.h"

using namespace zkx;

namespace zkx {

namespace {

std::atomic<bool> g_prof{false};

struct Pending {
    std::string name;
    hipEvent_t e0, e1;
};

struct ProfState {
    std::mutex mu;
    std::vector<Pending> pending;
    std::map<std::string, std::pair<uint64_t, double>> totals;
};

ProfState &prof()
{
    static ProfState s;
    return s;
}

}  // namespace

bool profiling() { return g_prof.load(std::memory_order_relaxed); }

ProfScope::ProfScope(const char *kernel_family, hipStream_t s) : name(kernel_family), stream(s)
{
    if (!profiling()) return;
    if (hipEventCreate(&e0) != hipSuccess) { e0 = nullptr; return; }
    (void)hipEventRecord(e0, stream);
}

ProfScope::~ProfScope()
{
    if (!e0) return;
    hipEvent_t e1 = nullptr;
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return; }
    (void)hipEventRecord(e1, stream);
    std::lock_guard<std::mutex> g(prof().mu);
    prof().pending.push_back({name, e0, e1});
}

static void resolve_pending()
{
    std::lock_guard<std::mutex> g(prof().mu);
    for (auto &p : prof().pending) {
        (void)hipEventSynchronize(p.e1);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.e0, p.e1) == hipSuccess) {
            auto &t = prof().totals[p.name];
            t.first += 1;
            t.second += ms;
        }
        (void)hipEventDestroy(p.e0);
        (void)hipEventDestroy(p.e1);
    }
    prof().pending.clear();
}

}  // namespace zkx

namespace {

RustError ok() { return RustError{0, nullptr}; }

template <class F>
RustError guarded(F &&f)
{
    try {
        f();
        return ok();
    } catch (const Error &e) {
        return RustError{e.code, strdup(e.what())};
    } catch (const std::exception &e) {
        return RustError{-1, strdup(e.what())};
    }
}

}  // namespace

extern "C" {

const char *zkx_last_error(void) { return last_error(); }

RustError zkx_device_malloc(size_t device_id, void **ptr, size_t bytes)
{
    return guarded([&] {
        if (!ptr) invalid("null pointer");
        device(device_id);
        *ptr = nullptr;
        if (bytes) ZKX_HIP(hipMalloc(ptr, bytes));
    });
}

RustError zkx_device_free(size_t device_id, void *ptr)
{
    return guarded([&] {
        device(device_id);
        if (ptr) ZKX_HIP(hipFree(ptr));
    });
}

RustError zkx_memcpy_htod(size_t device_id, void *dst, const void *src, size_t bytes)
{
    return guarded([&] {
        Device &d = device(device_id);
        if (bytes) ZKX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, d.stream));
        d.sync();
    });
}

RustError zkx_memcpy_dtoh(size_t device_id, void *dst, const void *src, size_t bytes)
{
    return guarded([&] {
        Device &d = device(device_id);
        if (bytes) ZKX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, d.stream));
        d.sync();
    });
}

RustError zkx_memset(size_t device_id, void *dst, int value, size_t bytes)
{
    return guarded([&] {
        Device &d = device(device_id);
        if (bytes) ZKX_HIP(hipMemsetAsync(dst, value, bytes, d.stream));
        d.sync();
    });
}

RustError zkx_fill_random(size_t device_id, void *dst, size_t count, uint64_t seed)
{
    return guarded([&] {
        Device &d = device(device_id);
        if (count && !dst) invalid("null pointer");
        if (count > (0x7FFFFFFFull << 8)) invalid("fill too large for one launch");
        fill_salt(static_cast<uint64_t *>(dst), count, seed, d.stream);
        d.sync();
    });
}

namespace {

// out[i] = op(a[i], b[i]) through the device field routines the kernels use;
// every result canonical.
__global__ void k_gl_ops(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t x = a[i], y = b[i];
    uint64_t r;
    switch (op) {
    case 0: r = gl::canon(gl::mul_lazy(x, y)); break;
    case 1: r = gl::mul(x, y); break;
    case 2: r = gl::canon(gl::reduce128_lazy(x, y)); break;
    case 3: r = gl::canon(gl::reduce96_lazy(x, gl::lo32(y))); break;
    case 4: r = gl::canon(gl::pow7_lazy(x)); break;
    case 5: r = gl::add(x, y); break;
    case 6: r = gl::sub(x, y); break;
    default: r = gl::canon(x); break;
    }
    out[i] = r;
}

}  // namespace

RustError zkx_gl_ops(size_t device_id, int op, const void *a, const void *b, void *out, size_t n)
{
    return guarded([&] {
        Device &d = device(device_id);
        if (op < 0 || op > 7) invalid("gl ops: op must be 0..7");
        if (!n) return;
        if (!a || !b || !out) invalid("null pointer");
        if (n > (1ull << 31)) invalid("gl ops: n above 2^31");
        const unsigned blocks = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_gl_ops, dim3(blocks), dim3(256), 0, d.stream, op, static_cast<const uint64_t *>(a),
                           static_cast<const uint64_t *>(b), static_cast<uint64_t *>(out), (uint64_t)n);
        ZKX_HIP(hipGetLastError());
        d.sync();
    });
}

RustError zkx_device_synchronize(size_t device_id)
{
    return guarded([&] {
        device(device_id);
        ZKX_HIP(hipDeviceSynchronize());
    });
}

void *zkx_stream(size_t device_id)
{
    try {
        return (void *)device(device_id).stream;
    } catch (...) {
        return nullptr;
    }
}

RustError zkx_profile_enable(int on)
{
    return guarded([&] { g_prof.store(on != 0); });
}

int zkx_kernel_stats(char names[][32], uint64_t *launches, double *total_ms, int cap)
{
    resolve_pending();
    std::lock_guard<std::mutex> g(prof().mu);
    int i = 0;
    for (auto &kv : prof().totals) {
        if (i < cap) {
            if (names) {
                strncpy(names[i], kv.first.c_str(), 31);
                names[i][31] = 0;
            }
            if (launches) launches[i] = kv.second.first;
            if (total_ms) total_ms[i] = kv.second.second;
        }
        i++;
    }
    return i;
}

void zkx_kernel_stats_reset(void)
{
    resolve_pending();
    std::lock_guard<std::mutex> g(prof().mu);
    prof().totals.clear();
}

}  // extern "C"
