// Host side of the NTT / LDE / transpose path: pass planning and launches.
// Entry semantics follow ntt/ntt.cu:182-226 (transpose), 253-309 (batched
// NTT) and 477-564 (batched LDE).
#include "ntt.h"

#include <algorithm>
#include <string>
#include <vector>

#include "ntt_kernels.h"

namespace zkx {

namespace {

// Launch one pass (a single descriptor passed by value).
void launch_pass(Device &d, const NttPass &p, uint64_t tiles, const char *name)
{
    if (tiles > 0x7FFFFFFFull) invalid("transform too large for one launch");
    ProfScope ps(name, d.stream);
    if (p.s_lo == 0)
        hipLaunchKernelGGL(k_ntt_tiles<true>, dim3((uint32_t)tiles), dim3(kNttThreads), 0, d.stream,
                           (const NttPass *)nullptr, 0u, p);
    else
        hipLaunchKernelGGL(k_ntt_tiles<false>, dim3((uint32_t)tiles), dim3(kNttThreads), 0, d.stream,
                           (const NttPass *)nullptr, 0u, p);
    ZKX_HIP(hipGetLastError());
}

// DIF passes over `nblk` blocks of 2^L held in buf; the first pass reads its
// tile through `mode0` (LDE coefficients, coset scaling or in place).
void run_dif(Device &d, uint64_t *buf, uint64_t nblk, uint32_t L, uint32_t lg_tw, const uint64_t *tw, bool inverse,
             uint32_t mode0, const uint64_t *src, uint32_t rate, bool with_coset, const uint64_t *coset)
{
    std::vector<NttPass> passes;
    const uint64_t *tw12 = d.twiddle_table(kNttTileLog);
    uint64_t tiles = plan_dif(passes, buf, nblk, L, lg_tw, tw, tw12, inverse, mode0, src, rate, with_coset, coset);
    for (const NttPass &p : passes) launch_pass(d, p, tiles, p.mode == kSrcLde ? "ntt_pass_lde" : "ntt_pass");
}

}  // namespace

void launch_ntt_list(hipStream_t stream, const NttPass *dev_list, uint32_t n, uint64_t blocks, bool contiguous,
                     const char *name)
{
    if (blocks == 0) return;
    if (blocks > 0x7FFFFFFFull) invalid("transform batch too large for one launch");
    ProfScope ps(name, stream);
    if (contiguous)
        hipLaunchKernelGGL(k_ntt_tiles<true>, dim3((uint32_t)blocks), dim3(kNttThreads), 0, stream, dev_list, n,
                           NttPass{});
    else
        hipLaunchKernelGGL(k_ntt_tiles<false>, dim3((uint32_t)blocks), dim3(kNttThreads), 0, stream, dev_list, n,
                           NttPass{});
    ZKX_HIP(hipGetLastError());
}

uint64_t plan_dif(std::vector<NttPass> &out, uint64_t *buf, uint64_t nblk, uint32_t L, uint32_t lg_tw,
                  const uint64_t *tw, const uint64_t *tw12, bool inverse, uint32_t mode0, const uint64_t *src,
                  uint32_t rate, bool with_coset, const uint64_t *coset)
{
    const uint64_t total = nblk << L;
    const uint64_t tiles = (total + (1ull << kNttTileLog) - 1) >> kNttTileLog;
    int hi = (int)L - 1;  // highest stage not yet planned
    bool first = true;
    do {
        NttPass a{};
        a.buf = buf;
        a.src = src;
        a.tw = tw;
        a.coset = coset;
        a.tw12 = tw12;
        a.total = total;
        a.L = L;
        a.lg_tw = lg_tw;
        a.rate = rate;
        a.inverse = inverse ? 1u : 0u;
        a.with_coset = (with_coset && first) ? 1u : 0u;
        a.mode = first ? mode0 : (uint32_t)kSrcInPlace;
        uint32_t m, s_lo;
        if (hi + 1 <= kNttTileLog) {
            // last pass: stages hi..0 on contiguous groups (hi = -1: source mapping only)
            m = (uint32_t)(hi + 1);
            s_lo = 0;
        } else {
            // strided pass: stages hi..s_lo, at least 2^3 columns per tile, leaving
            // <= 12 stages for the final contiguous pass
            const int rest = hi + 1 - kNttTileLog;
            m = (uint32_t)std::min(rest, kNttTileLog - 3);
            s_lo = (uint32_t)(hi + 1) - m;
        }
        a.s_lo = s_lo;
        a.m = m;
        a.c = kNttTileLog - m;
        out.push_back(a);
        first = false;
        hi = (int)s_lo - 1;
    } while (hi >= 0);
    return tiles;
}

void bitrev_columns(Device &d, uint64_t *buf, uint32_t lg, uint32_t cols, bool use_scale, uint64_t scale,
                    const uint64_t *coset)
{
    BitrevArgs a{};
    a.buf = buf;
    a.coset = coset;
    a.scale = scale;
    a.lg = lg;
    a.cols = cols;
    a.use_scale = use_scale ? 1u : 0u;
    a.use_coset = coset ? 1u : 0u;
    ProfScope ps("bitrev", d.stream);
    if (lg >= 2 * kBrTile) {
        const uint64_t blocks = (uint64_t)cols << (lg - 2 * kBrTile);
        if (blocks > 0x7FFFFFFFull) invalid("bit reversal too large for one launch");
        hipLaunchKernelGGL(k_bitrev_tiles, dim3((uint32_t)blocks), dim3(kNttThreads), 0, d.stream, a);
    } else {
        const uint64_t n = ((uint64_t)cols) << lg;
        hipLaunchKernelGGL(k_bitrev_small, dim3((uint32_t)((n + kNttThreads - 1) / kNttThreads)),
                           dim3(kNttThreads), 0, d.stream, a);
    }
    ZKX_HIP(hipGetLastError());
}

const uint64_t *lde_factor_table(Device &d, uint32_t lg_n, uint32_t rate, bool inverse, bool with_coset)
{
    if (rate > kMaxRateBits) invalid("extension_rate_bits above 5 are not supported by this build");
    const bool coset = with_coset && !inverse;
    if (coset && !d.coset) invalid("coset table not initialised (init_coset)");
    const uint64_t gen = coset ? d.coset_gen : 1;
    const auto key = std::make_tuple(lg_n, rate, inverse, gen);
    auto it = d.lde_factors.find(key);
    if (it != d.lde_factors.end()) return it->second;
    const uint32_t lg = lg_n + rate;
    const uint64_t w = inverse ? host_inv(omega(lg)) : omega(lg);
    FactorBases fb{};
    for (uint32_t b = 0; b < (1u << rate); b++) {
        uint32_t rb = 0;
        for (uint32_t i = 0; i < rate; i++) rb |= ((b >> i) & 1u) << (rate - 1 - i);
        fb.c[b] = host_mul(gen, host_pow(w, rb));
    }
    const uint64_t total = 1ull << lg;
    uint64_t *F = nullptr;
    ZKX_HIP(hipMalloc(&F, total * sizeof(uint64_t)));
    const uint64_t chunk = lg_n < 6 ? (1ull << lg_n) : 64;
    const uint64_t threads = total / chunk;
    hipLaunchKernelGGL(k_lde_factors, dim3((uint32_t)((threads + kNttThreads - 1) / kNttThreads)),
                       dim3(kNttThreads), 0, d.stream, F, lg_n, rate, fb);
    ZKX_HIP(hipGetLastError());
    d.lde_factors[key] = F;
    return F;
}

void lde_bitrev(Device &d, uint64_t *out, const uint64_t *in, uint32_t lg_n, uint32_t rate, uint32_t cols,
                bool inverse, bool with_coset)
{
    const uint32_t lg = lg_n + rate;
    const uint64_t *tw = d.twiddle_table(lg);
    const uint64_t *F = lde_factor_table(d, lg_n, rate, inverse, with_coset);
    run_dif(d, out, (uint64_t)cols << rate, lg_n, lg, tw, inverse, kSrcLde, in, rate, false, F);
}

void lde(Device &d, uint64_t *out, const uint64_t *in, uint32_t lg_n, uint32_t rate, uint32_t cols,
         bool inverse, bool with_coset)
{
    const uint32_t lg = lg_n + rate;
    if (with_coset && inverse && (!d.coset || d.coset_lg < lg))
        invalid("coset table not initialised for log size " + std::to_string(lg));
    lde_bitrev(d, out, in, lg_n, rate, cols, inverse, with_coset);
    // natural order; inverse transforms are scaled by N^-1 and (reference
    // semantics, ntt.cu:137-139) by g^t when with_coset
    bitrev_columns(d, out, lg, cols, inverse, inverse ? host_inv((1ull << lg) % 0xFFFFFFFF00000001ull) : 1,
                   (inverse && with_coset) ? d.coset : nullptr);
}

void ntt(Device &d, uint64_t *inout, uint32_t lg, uint32_t cols, bool inverse, bool with_coset)
{
    if (with_coset && (!d.coset || d.coset_lg < lg))
        invalid("coset table not initialised for log size " + std::to_string(lg));
    const uint64_t *tw = d.twiddle_table(lg);
    const uint32_t mode = (with_coset && !inverse) ? kSrcCoset : kSrcInPlace;
    run_dif(d, inout, cols, lg, lg, tw, inverse, mode, nullptr, 0, false, d.coset);
    bitrev_columns(d, inout, lg, cols, inverse, inverse ? host_inv((1ull << lg) % 0xFFFFFFFF00000001ull) : 1,
                   (inverse && with_coset) ? d.coset : nullptr);
}

void transpose_rev(Device &d, uint64_t *out, const uint64_t *in, uint32_t lg, uint32_t cols, bool rev)
{
    const uint64_t n = 1ull << lg;
    const uint64_t gx = (n + kTrTile - 1) / kTrTile;
    const uint32_t gy = (cols + kTrTile - 1) / kTrTile;
    if (gx > 0x7FFFFFFFull || gy > 65535) invalid("transpose too large for one launch");
    ProfScope ps("transpose_rev", d.stream);
    hipLaunchKernelGGL(k_transpose_rev, dim3((uint32_t)gx, gy), dim3(kNttThreads), 0, d.stream, out, in, lg, cols,
                       rev ? 1u : 0u);
    ZKX_HIP(hipGetLastError());
}

void fill_powers(uint64_t *dst, uint64_t n, uint64_t base, hipStream_t s)
{
    const uint64_t threads = (n + 63) / 64;
    const uint32_t blocks = (uint32_t)((threads + kNttThreads - 1) / kNttThreads);
    hipLaunchKernelGGL(k_powers, dim3(blocks), dim3(kNttThreads), 0, s, dst, n, base);
    ZKX_HIP(hipGetLastError());
}

void fill_salt(uint64_t *dst, uint64_t count, uint64_t seed, hipStream_t s)
{
    if (!count) return;
    hipLaunchKernelGGL(k_salt, dim3((uint32_t)((count + kNttThreads - 1) / kNttThreads)), dim3(kNttThreads), 0, s,
                       dst, count, seed);
    ZKX_HIP(hipGetLastError());
}

}  // namespace zkx
