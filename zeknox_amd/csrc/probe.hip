// Issue-rate probe of the gfx950 SIMD, the measured inputs of the hash
// kernels' roofline:
//   * the achieved rate of v_mad_u64_u32, the 32x32+64 multiply-add all
//     Goldilocks products here are built from (independent chains, four
//     waves per SIMD);
//   * how many of those instruction slots one v_mfma_i32_32x32x32_i8 (the
//     Poseidon MDS layer's matrix-core instruction) takes from its SIMD when
//     it runs beside full-rate multiply-adds: (t(mixed) - t(VALU only)) per
//     MFMA over t(VALU only) per multiply-add.
// Every lane writes one word, so nothing is optimised away; the loop bodies
// are exactly kMads multiply-adds (and kMfma MFMAs), the loop control scalar.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>

#include "../../include/zeknox_ext.h"
#include "runtime.h"

namespace zkx {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kChains = 8;     // independent multiply-add chains per lane
constexpr int kMadsPerIter = 32;
constexpr int kAcc = 4;        // independent MFMA accumulators
constexpr int kThreads = 256;
constexpr int kWavesPerSimd = 4;

template <bool VALU, int MFMA>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kWavesPerSimd)))
k_issue_probe(uint64_t *out, uint32_t iters, uint32_t m)
{
    const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    uint64_t acc[kChains];
#pragma unroll
    for (int k = 0; k < kChains; k++) acc[k] = t + (uint64_t)k * 0x9E3779B97F4A7C15ull;
    v16i d[kAcc];
#pragma unroll
    for (int q = 0; q < kAcc; q++) d[q] = v16i{};
    const v4i a = {(int)t, (int)m, 3, 5}, b = {(int)m, 7, (int)t, 11};
    // kAcc iterations per trip, each on its own accumulator (no dynamic
    // indexing of the register arrays); iters is a multiple of kAcc
    for (uint32_t it = 0; it < iters; it += kAcc) {
#pragma unroll
        for (int s = 0; s < kAcc; s++) {
            if constexpr (VALU) {
#pragma unroll
                for (int r = 0; r < kMadsPerIter / kChains; r++)
#pragma unroll
                    for (int k = 0; k < kChains; k++) acc[k] = (uint64_t)(uint32_t)acc[k] * m + acc[k];
            }
#pragma unroll
            for (int q = 0; q < MFMA; q++) d[s] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, d[s], 0, 0, 0);
        }
    }
    uint64_t x = 0;
#pragma unroll
    for (int k = 0; k < kChains; k++) x ^= acc[k];
#pragma unroll
    for (int q = 0; q < kAcc; q++)
#pragma unroll
        for (int i = 0; i < 16; i++) x += (uint32_t)d[q][i];
    out[t] = x;
}

}  // namespace

}  // namespace zkx

extern "C" RustError zkx_issue_probe(size_t device_id, uint32_t iters, ZkxIssueRates *out)
{
    using namespace zkx;
    try {
        if (!out) invalid("null output");
        if (iters == 0 || iters % kAcc) invalid("iters must be a positive multiple of 4");
        Device &d = device(device_id);
        std::lock_guard<std::mutex> g(d.mu);
        // four waves on every SIMD
        const uint32_t blocks = (uint32_t)d.prop.multiProcessorCount * 4 * kWavesPerSimd * 64 / kThreads;
        DevBuf buf((size_t)blocks * kThreads * sizeof(uint64_t));
        hipEvent_t e0, e1;
        ZKX_HIP(hipEventCreate(&e0));
        ZKX_HIP(hipEventCreate(&e1));
        // one warm-up launch, then the fastest of kRuns timed launches (the
        // MFMA cost is a small difference of two times)
        constexpr int kRuns = 5;
        auto run = [&](auto kernel, uint32_t n) {
            hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kThreads), 0, d.stream, buf.as<uint64_t>(), n, 0x9E3779B1u);
            ZKX_HIP(hipGetLastError());
            double best = 1e30;
            for (int r = 0; r < kRuns; r++) {
                ZKX_HIP(hipEventRecord(e0, d.stream));
                hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kThreads), 0, d.stream, buf.as<uint64_t>(), n,
                                   0x9E3779B1u);
                ZKX_HIP(hipGetLastError());
                ZKX_HIP(hipEventRecord(e1, d.stream));
                ZKX_HIP(hipEventSynchronize(e1));
                float ms = 0;
                ZKX_HIP(hipEventElapsedTime(&ms, e0, e1));
                best = std::min(best, (double)ms);
            }
            return best;
        };
        const double waves = (double)blocks * kThreads / 64;
        const double t_valu = run(k_issue_probe<true, 0>, iters);
        const double t_mfma = run(k_issue_probe<false, 1>, iters);
        const double t_mixed = run(k_issue_probe<true, 1>, iters);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        memset(out, 0, sizeof(*out));
        out->ms_valu = t_valu;
        out->ms_mfma = t_mfma;
        out->ms_mixed = t_mixed;
        const double mads = waves * iters * kMadsPerIter, mfmas = waves * iters;
        out->valu_wave_insts_per_s = mads / (t_valu * 1e-3);
        out->mfma_wave_insts_per_s = mfmas / (t_mfma * 1e-3);
        out->mfma_issue_slots = ((t_mixed - t_valu) / mfmas) / (t_valu / mads);
        out->simds = (uint32_t)d.prop.multiProcessorCount * 4;
        out->waves_per_simd = kWavesPerSimd;
        out->mads_per_mfma = kMadsPerIter;
        return RustError{0, nullptr};
    } catch (const Error &e) {
        return RustError{e.code, strdup(e.what())};
    } catch (const std::exception &e) {
        return RustError{-1, strdup(e.what())};
    }
}
