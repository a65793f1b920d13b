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

// One work-group per tile of [first, end).
template <int MODE>
void launch_strided1(int m, hipStream_t stream, const NttPass *dev_list, uint32_t n, const NttPass &one,
                     uint64_t first, uint64_t end)
{
    const dim3 grid((uint32_t)(end - first)), block(kNttThreads);
    switch (m) {
    case 1: hipLaunchKernelGGL((k_ntt_strided1_tiles<MODE, 1>), grid, block, 0, stream, dev_list, n, one, first); break;
    case 2: hipLaunchKernelGGL((k_ntt_strided1_tiles<MODE, 2>), grid, block, 0, stream, dev_list, n, one, first); break;
    case 3: hipLaunchKernelGGL((k_ntt_strided1_tiles<MODE, 3>), grid, block, 0, stream, dev_list, n, one, first); break;
    default: hipLaunchKernelGGL((k_ntt_strided1_tiles<MODE, 4>), grid, block, 0, stream, dev_list, n, one, first); break;
    }
}

// One work-group per tile of [first, end).
template <int MODE>
void launch_strided2(int m, hipStream_t stream, const NttPass *dev_list, uint32_t n, const NttPass &one,
                     uint64_t first, uint64_t end)
{
    const dim3 grid((uint32_t)(end - first)), block(kNttThreads);
    switch (m) {
    case 5: hipLaunchKernelGGL((k_ntt_strided2_tiles<MODE, 5>), grid, block, 0, stream, dev_list, n, one, first); break;
    case 6: hipLaunchKernelGGL((k_ntt_strided2_tiles<MODE, 6>), grid, block, 0, stream, dev_list, n, one, first); break;
    case 7: hipLaunchKernelGGL((k_ntt_strided2_tiles<MODE, 7>), grid, block, 0, stream, dev_list, n, one, first); break;
    default: hipLaunchKernelGGL((k_ntt_strided2_tiles<MODE, 8>), grid, block, 0, stream, dev_list, n, one, first); break;
    }
}

// A launch over the tiles [first, end) of the passes (device list, or `one`
// when n == 0) with the kernel ntt_kernel_kind chose; false for kind 0.
bool launch_kind(int kind, hipStream_t stream, const NttPass *dev_list, uint32_t n, const NttPass &one,
                 uint64_t first, uint64_t end)
{
    if (kind == 1) {
        hipLaunchKernelGGL(k_ntt_contig_reg, dim3((uint32_t)(end - first)), dim3(kNttThreads), 0, stream, dev_list, n,
                           one, first);
        return true;
    }
    if (kind >= 2 && kind < 10) {
        const int m = (kind - 2) % 4 + 1;
        if (kind >= 6) launch_strided1<kSrcLde>(m, stream, dev_list, n, one, first, end);
        else launch_strided1<kSrcInPlace>(m, stream, dev_list, n, one, first, end);
        return true;
    }
    if (kind >= 10 && kind < 18) {
        const int m = (kind - 10) % 4 + 5;
        if (kind >= 14) launch_strided2<kSrcLde>(m, stream, dev_list, n, one, first, end);
        else launch_strided2<kSrcInPlace>(m, stream, dev_list, n, one, first, end);
        return true;
    }
    return false;
}

// Launch one pass (a single descriptor passed by value).
void launch_pass(Device &d, const NttPass &p, uint64_t tiles, const char *name)
{
    if (tiles > 0x7FFFFFFFull) invalid("transform too large for one launch");
    ProfScope ps(name, d.stream);
    if (launch_kind(ntt_kernel_kind(&p, 1), d.stream, nullptr, 0u, p, 0, tiles)) {
    } else if (p.s_lo == 0)
        hipLaunchKernelGGL(k_ntt_tiles<true>, dim3((uint32_t)tiles), dim3(kNttThreads), 0, d.stream,
                           (const NttPass *)nullptr, 0u, p, (uint64_t)0);
    else
        hipLaunchKernelGGL(k_ntt_tiles<false>, dim3((uint32_t)tiles), dim3(kNttThreads), 0, d.stream,
                           (const NttPass *)nullptr, 0u, p, (uint64_t)0);
    ZKX_HIP(hipGetLastError());
}

// DIF passes over `nblk` blocks of 2^L held in buf; the first pass reads its
// tile through `mode0` (LDE coefficients, coset scaling or in place).
void run_dif(Device &d, uint64_t *buf, uint64_t nblk, uint32_t L, uint32_t lg_tw, const uint64_t *tw, bool inverse,
             uint32_t mode0, const uint64_t *src, uint32_t rate, bool with_coset, const uint64_t *coset)
{
    std::vector<NttPass> passes;
    const uint64_t *tw12 = d.twiddle_table(kNttTileLog);
    const uint64_t *const *tw_lg = d.twiddle_tables_upto(L);
    uint64_t tiles = plan_dif(passes, buf, nblk, L, lg_tw, tw, tw12, tw_lg, inverse, mode0, src, rate, with_coset,
                              coset, d.tuning_default());
    for (const NttPass &p : passes) launch_pass(d, p, tiles, p.mode == kSrcLde ? "ntt_pass_lde" : "ntt_pass");
}

}  // namespace

int ntt_kernel_kind(const NttPass *host_list, size_t n)
{
    if (n == 0) return 0;
    const NttPass &p0 = host_list[0];
    if (p0.s_lo == 0) {
        for (size_t i = 0; i < n; i++) {
            const NttPass &p = host_list[i];
            if (p.s_lo != 0 || p.mode != kSrcInPlace || (p.total & ((1ull << kNttTileLog) - 1)) != 0) return 0;
        }
        return 1;
    }
    if (p0.m < 1 || p0.m > 8 || (p0.mode != kSrcInPlace && p0.mode != kSrcLde)) return 0;
    for (size_t i = 0; i < n; i++) {
        const NttPass &p = host_list[i];
        if (p.s_lo == 0 || p.m != p0.m || p.mode != p0.mode) return 0;
    }
    if (p0.m >= 5) return 10 + (p0.mode == kSrcLde ? 4 : 0) + (int)(p0.m - 5);
    return 2 + (p0.mode == kSrcLde ? 4 : 0) + (int)(p0.m - 1);
}

void launch_ntt_list(hipStream_t stream, const NttPass *dev_list, uint32_t n, uint64_t blocks, bool contiguous,
                     int kernel_kind, const char *name, uint64_t chunk)
{
    if (blocks == 0) return;
    if (blocks > 0x7FFFFFFFull) invalid("transform batch too large for one launch");
    ProfScope ps(name, stream);
    const uint64_t step = chunk ? chunk : blocks;
    for (uint64_t first = 0; first < blocks; first += step) {
        const uint64_t end = std::min(blocks, first + step);
        if (launch_kind(kernel_kind, stream, dev_list, n, NttPass{}, first, end)) {
        } else if (contiguous)
            hipLaunchKernelGGL(k_ntt_tiles<true>, dim3((uint32_t)(end - first)), dim3(kNttThreads), 0, stream,
                               dev_list, n, NttPass{}, first);
        else
            hipLaunchKernelGGL(k_ntt_tiles<false>, dim3((uint32_t)(end - first)), dim3(kNttThreads), 0, stream,
                               dev_list, n, NttPass{}, first);
        ZKX_HIP(hipGetLastError());
    }
}

// Stage counts of the passes, top stages first: every pass but the last is
// strided (<= strided_max stages, and its stride s_lo at least its column
// count 12 - m); the last is contiguous (<= 12 stages).
std::vector<uint32_t> ntt_pass_sizes(uint32_t L, const Tuning &tun)
{
    const uint32_t smax = std::min<uint32_t>(std::max<uint32_t>(tun.ntt_strided_max, 1), kNttTileLog - 3);
    std::vector<uint32_t> sizes;
    if (L <= (uint32_t)kNttTileLog) {
        sizes.push_back(L);
        return sizes;
    }
    if (tun.ntt_even) {
        const uint32_t P = std::max((L + kNttTileLog - 1) / kNttTileLog, 1 + (L - kNttTileLog + smax - 1) / smax);
        std::vector<uint32_t> e(P, L / P);
        for (uint32_t k = 0; k < L % P; k++) e[k]++;
        // strided passes over their limit hand stages to the last pass
        for (uint32_t k = 0; k + 1 < P; k++)
            while (e[k] > smax && e[P - 1] < (uint32_t)kNttTileLog) {
                e[k]--;
                e[P - 1]++;
            }
        bool ok = e[P - 1] <= (uint32_t)kNttTileLog;
        uint32_t below = L;
        for (uint32_t k = 0; ok && k + 1 < P; k++) {
            below -= e[k];
            ok = e[k] >= 1 && e[k] <= smax && below >= kNttTileLog - e[k];
        }
        if (ok) return e;
    }
    uint32_t rest = L;
    while (rest > (uint32_t)kNttTileLog) {
        const uint32_t m = std::min(rest - kNttTileLog, smax);
        sizes.push_back(m);
        rest -= m;
    }
    sizes.push_back(rest);
    return sizes;
}

uint64_t plan_dif(std::vector<NttPass> &out, uint64_t *buf, uint64_t nblk, uint32_t L, uint32_t lg_tw,
                  const uint64_t *tw, const uint64_t *tw12, const uint64_t *const *tw_lg, bool inverse,
                  uint32_t mode0, const uint64_t *src, uint32_t rate, bool with_coset, const uint64_t *coset,
                  const Tuning &tun)
{
    const uint64_t total = nblk << L;
    const uint64_t tiles = (total + (1ull << kNttTileLog) - 1) >> kNttTileLog;
    const std::vector<uint32_t> sizes = ntt_pass_sizes(L, tun);
    uint32_t hi = L;  // stages [0, hi) not yet planned
    for (size_t k = 0; k < sizes.size(); k++) {
        NttPass a{};
        a.buf = buf;
        a.src = src;
        a.tw = tw;
        a.coset = coset;
        a.tw12 = tw12;
        a.tw_lg = tw_lg;
        a.total = total;
        a.L = L;
        a.lg_tw = lg_tw;
        a.rate = rate;
        a.inverse = inverse ? 1u : 0u;
        a.with_coset = (with_coset && k == 0) ? 1u : 0u;
        a.mode = k == 0 ? mode0 : (uint32_t)kSrcInPlace;
        a.m = sizes[k];
        a.s_lo = hi - a.m;  // 0 exactly for the last (contiguous) pass
        a.c = kNttTileLog - a.m;
        if (a.s_lo > 0 && a.s_lo < a.c) invalid("internal: NTT pass plan with stride below its column count");
        out.push_back(a);
        hi = a.s_lo;
    }
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

namespace {

template <int NE>
void launch_transpose_rows(uint32_t pw, dim3 grid, hipStream_t s, uint64_t *out, const uint64_t *in, uint32_t lg,
                           uint32_t B, uint32_t rev, uint32_t cb)
{
    switch (pw) {
    case 32: hipLaunchKernelGGL((k_transpose_rows<NE, 32>), grid, dim3(kNttThreads), 0, s, out, in, lg, B, rev, cb); break;
    case 64: hipLaunchKernelGGL((k_transpose_rows<NE, 64>), grid, dim3(kNttThreads), 0, s, out, in, lg, B, rev, cb); break;
    case 128: hipLaunchKernelGGL((k_transpose_rows<NE, 128>), grid, dim3(kNttThreads), 0, s, out, in, lg, B, rev, cb); break;
    default: hipLaunchKernelGGL((k_transpose_rows<NE, 256>), grid, dim3(kNttThreads), 0, s, out, in, lg, B, rev, cb); break;
    }
}

}  // namespace

void transpose_rev(Device &d, uint64_t *out, const uint64_t *in, uint32_t lg, uint32_t cols, bool rev)
{
    const uint64_t n = 1ull << lg;
    if (cols >= 32 && n >= kTrRows) {
        // whole-row tiles: columns split evenly into chunks of at most kTrMaxCols
        const uint32_t chunks = (cols + kTrMaxCols - 1) / kTrMaxCols;
        const uint32_t cb = (cols + chunks - 1) / chunks;
        const uint32_t ne = (cb * kTrRows + kNttThreads - 1) / kNttThreads;
        uint32_t pw = 32;
        while (pw < cb) pw <<= 1;
        const uint64_t gx = n / kTrRows;
        if (gx > 0x7FFFFFFFull || chunks > 65535) invalid("transpose too large for one launch");
        ProfScope ps("transpose_rev", d.stream);
        const dim3 grid((uint32_t)gx, chunks);
        if (ne <= 4) launch_transpose_rows<4>(pw, grid, d.stream, out, in, lg, cols, rev ? 1u : 0u, cb);
        else if (ne <= 8) launch_transpose_rows<8>(pw, grid, d.stream, out, in, lg, cols, rev ? 1u : 0u, cb);
        else if (ne <= 12) launch_transpose_rows<12>(pw, grid, d.stream, out, in, lg, cols, rev ? 1u : 0u, cb);
        else launch_transpose_rows<17>(pw, grid, d.stream, out, in, lg, cols, rev ? 1u : 0u, cb);
        ZKX_HIP(hipGetLastError());
        return;
    }
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
