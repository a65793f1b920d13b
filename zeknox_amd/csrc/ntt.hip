// Host side of the NTT / LDE / transpose path: pass planning and launches.
// Entry semantics follow ntt/ntt.cu:182-226 (transpose), 253-309 (batched
// NTT) and 477-564 (batched LDE).
#include "ntt.h"

#include <algorithm>
#include <string>

#include "ntt_kernels.h"

namespace zkx {

namespace {

// DIF passes over `nblk` blocks of 2^L held in buf.  The first pass reads
// its tile through `mode0` (LDE coefficients, coset scaling or in place).
void run_dif(Device &d, uint64_t *buf, uint64_t nblk, uint32_t L, uint32_t lg_tw, const uint64_t *tw,
             bool inverse, int mode0, const uint64_t *src, uint32_t rate, bool with_coset,
             const uint64_t *coset)
{
    const uint64_t total = nblk << L;
    int hi = (int)L - 1;
    bool first = true;
    if (L == 0) {
        // a length-1 transform is the identity; only the source mapping runs
        hi = -1;
    }
    do {
        PassArgs a{};
        a.buf = buf;
        a.src = src;
        a.tw = tw;
        a.coset = coset;
        a.total = total;
        a.L = L;
        a.lg_tw = lg_tw;
        a.rate = rate;
        a.inverse = inverse ? 1u : 0u;
        a.with_coset = with_coset ? 1u : 0u;
        uint32_t m, s_lo, c;
        if (hi < 0) {
            m = 0;
            s_lo = 0;
        } else if (hi + 1 <= kNttTileLog) {
            // last pass: stages hi..0 on contiguous groups
            m = (uint32_t)hi + 1;
            s_lo = 0;
        } else {
            // strided pass: stages hi..s_lo with at least 2^3 columns per tile,
            // leaving <= 12 stages for the final contiguous pass
            const int rest = hi + 1 - kNttTileLog;  // stages that must go before the last pass
            m = (uint32_t)std::min(rest, kNttTileLog - 3);
            s_lo = (uint32_t)(hi + 1) - m;
        }
        if (s_lo > 0) c = std::min<uint32_t>(s_lo, kNttTileLog - m);
        else c = kNttTileLog - m;  // 2^c contiguous groups per tile
        a.s_lo = s_lo;
        a.m = m;
        a.c = c;
        const uint64_t tiles = (total + (1ull << (m + c)) - 1) >> (m + c);
        if (tiles > 0x7FFFFFFFull) invalid("transform too large for one launch");
        const int mode = first ? mode0 : kSrcInPlace;
        dim3 grid((uint32_t)tiles), block(kNttThreads);
        ProfScope ps(mode == kSrcLde ? "ntt_pass_lde" : "ntt_pass", d.stream);
        if (mode == kSrcLde) hipLaunchKernelGGL(k_dif_pass<kSrcLde>, grid, block, 0, d.stream, a);
        else if (mode == kSrcCoset) hipLaunchKernelGGL(k_dif_pass<kSrcCoset>, grid, block, 0, d.stream, a);
        else hipLaunchKernelGGL(k_dif_pass<kSrcInPlace>, grid, block, 0, d.stream, a);
        ZKX_HIP(hipGetLastError());
        first = false;
        hi = (int)s_lo - 1;
    } while (hi >= 0);
}

}  // namespace

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

void lde_bitrev(Device &d, uint64_t *out, const uint64_t *in, uint32_t lg_n, uint32_t rate, uint32_t cols,
                bool inverse, bool with_coset)
{
    const uint32_t lg = lg_n + rate;
    const uint64_t *tw = d.twiddle_table(lg);
    const uint64_t *cs = nullptr;
    if (with_coset && !inverse) {
        if (!d.coset || d.coset_lg < lg_n) invalid("coset table not initialised for log size " + std::to_string(lg));
        cs = d.coset;
    }
    run_dif(d, out, (uint64_t)cols << rate, lg_n, lg, tw, inverse, kSrcLde, in, rate, with_coset && !inverse, cs);
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
    const int mode = (with_coset && !inverse) ? kSrcCoset : kSrcInPlace;
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
