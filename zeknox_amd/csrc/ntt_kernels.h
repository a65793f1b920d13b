// Goldilocks NTT / LDE / transpose kernels for gfx950.
//
// Layout in HBM: column-major batches, column b of a batch of length 2^lg at
// [b * 2^lg, (b+1) * 2^lg), u64 little-endian canonical elements -- the
// layout of compute_batched_ntt / compute_batched_lde (ntt/ntt.cu:253-309,
// 477-564).
//
// Algorithm.  The forward transform is decimation-in-frequency (natural
// order in, bit-reversed order out), run as a few passes over HBM.  Each pass
// performs up to 12 consecutive radix-2 stages on a tile of 4096 elements in
// LDS: 2^m elements at stride 2^s_lo (the butterfly partners of stages
// s_lo..s_lo+m-1) times 2^c consecutive columns of that stride (m + c = 12),
// so every global load and store is a run of 2^c consecutive u64.  Inside
// the tile, each lane holds 16 elements in VGPRs and runs 4 stages per round
// (radix 16) between LDS exchanges.
//
// LDE.  Zero-padding n = 2^lg_n coefficients to N = 2^(lg_n+r) and running the
// first r DIF stages leaves, in block b of length n (b < 2^r, memory order),
// the values a_i * F[b*n + i] with F[b*n + i] = (g * w_N^{bitrev_r(b)})^i
// (g = 1 without coset; w^-1 for the inverse): the first pass reads the n
// coefficients directly and multiplies by the resident factor table F, so
// the zero half is never read, written or butterflied.
#pragma once
#include <type_traits>

#include "gl64.h"
#include "gmem.h"
#include "ntt_types.h"

namespace zkx {

__device__ __forceinline__ uint64_t tw_at(const uint64_t *tw, uint32_t lg_tw, uint64_t k, bool inverse)
{
    const uint64_t mask = (1ull << lg_tw) - 1;
    return gmem(tw)[inverse ? ((0 - k) & mask) : k];
}

// Stage twiddle of a contiguous pass: w_4096^k (w^-k for the inverse),
// k < 2048, read from the resident 4096-entry table (16 KB of it hot in L1).
__device__ __forceinline__ uint64_t tw12_at(const NttPass &a, uint32_t k)
{
    return gmem(a.tw12)[a.inverse ? ((0u - k) & ((1u << kNttTileLog) - 1)) : k];
}

// Twiddle table of a strided pass's stage d (butterfly span 2^(s_lo+d+1)):
// w_{2^(s_lo+d+1)}^k sits at index k, so the 2^c consecutive columns of a
// tile read consecutive words (rather than words 2^(lg_tw-s_lo-d-1) apart
// in the largest table).
__device__ __forceinline__ const uint64_t *stage_table(const NttPass &a, uint32_t lg)
{
    return gmem(a.tw_lg)[lg];
}

__device__ __forceinline__ uint32_t brev_bits(uint32_t x, uint32_t bits)
{
    return bits ? (__builtin_bitreverse32(x) >> (32 - bits)) : 0u;
}

// LDS slot of tile position u (one pad word per 32 breaks power-of-two strides)
__device__ __forceinline__ uint32_t lds_slot(uint32_t u) { return u + (u >> 5); }

// Tile positions are numbered u in memory order: u = j*2^c + q when the
// pass is strided (s_lo > 0: column q of the stride, butterfly index j),
// u = q*2^m + j when it is contiguous (s_lo == 0: group q, index j).
__device__ __forceinline__ uint64_t tile_index(const NttPass &a, uint64_t tile, uint32_t u)
{
    if (a.s_lo > 0) {
        const uint32_t q = u & ((1u << a.c) - 1), j = u >> a.c;
        const uint32_t chunk_bits = a.s_lo - a.c;  // column chunks per group
        const uint64_t g = tile >> chunk_bits, ch = tile & ((1ull << chunk_bits) - 1);
        return (g << (a.s_lo + a.m)) + ((uint64_t)j << a.s_lo) + (ch << a.c) + q;
    }
    return (tile << kNttTileLog) + u;
}

// Index of tile position u = t + 256 i as tile_index(a, tile, t) plus this
// wave-uniform step: the lane id and i land in separate bit fields of the
// position (contiguous: u itself; strided: q = u mod 2^c, j = u >> c).
__device__ __forceinline__ uint64_t tile_step(const NttPass &a, int i)
{
    constexpr uint32_t lt = 8;  // log2 kNttThreads
    static_assert((1 << lt) == kNttThreads, "256 lanes per tile");
    if (a.s_lo == 0) return (uint64_t)i << lt;
    if (a.c >= lt) {
        const uint32_t cb = a.c - lt;
        return ((uint64_t)(i & ((1 << cb) - 1)) << lt) + ((uint64_t)(i >> cb) << a.s_lo);
    }
    return (uint64_t)i << (lt - a.c + a.s_lo);
}

// One register round: stages on j-bits [b0, b0+K) (u-bits [RB, RB+K)).
// Lane t holds the 2^K elements that differ in those bits, repeated over
// 2^(4-K) groups taken from the extra element bits; the other u-bits come
// from the lane id.  The twiddle of stage d for element e depends only on
// (j mod 2^d) << s_lo | lo.
template <bool CONTIG, int K>
__device__ __forceinline__ void ntt_round(uint64_t *lds, const NttPass &a, uint32_t t, uint32_t b0, uint32_t jpos,
                                          uint64_t lo_base)
{
    constexpr int E = kNttPerThread;
    const uint32_t RB = jpos + b0;
    uint64_t x[E];
    uint32_t uu[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint32_t rb = (uint32_t)e & ((1u << K) - 1);
        const uint32_t other = t | (((uint32_t)e >> K) << 8);
        const uint32_t u = ((other >> RB) << (RB + K)) | (rb << RB) | (other & ((1u << RB) - 1));
        uu[e] = u;
        x[e] = lds[lds_slot(u)];
    }
#pragma unroll
    for (int dd = K - 1; dd >= 0; dd--) {
        const uint32_t d = b0 + (uint32_t)dd;
        const uint32_t lg = a.s_lo + d + 1;
        const uint64_t *tab = !CONTIG ? stage_table(a, lg) : nullptr;
#pragma unroll
        for (int e = 0; e < E; e++) {
            if ((e >> dd) & 1) continue;
            const int f = e | (1 << dd);
            const uint32_t u = uu[e];
            const uint32_t j = !CONTIG ? (u >> a.c) : (u & ((1u << a.m) - 1));
            uint64_t w;
            if constexpr (CONTIG) {
                w = tw12_at(a, (j & ((1u << d) - 1)) << (kNttTileLog - 1 - d));
            } else {
                const uint64_t lo_part = lo_base + (u & ((1u << a.c) - 1));
                const uint64_t ex = (((uint64_t)(j & ((1u << d) - 1))) << a.s_lo) | lo_part;
                w = tw_at(tab, lg, ex, a.inverse);
            }
            const uint64_t y0 = x[e], y1 = x[f];
            x[e] = gl::add(y0, y1);
            x[f] = gl::mul(gl::sub(y0, y1), w);
        }
    }
#pragma unroll
    for (int e = 0; e < E; e++) lds[lds_slot(uu[e])] = x[e];
}

// Full-width round (K = 4, no extra element bits): the lane's other u-bits
// are its id, so the 15 distinct twiddles of the round are loaded once.
template <bool CONTIG>
__device__ __forceinline__ void ntt_round4(uint64_t *lds, const NttPass &a, uint32_t t, uint32_t b0, uint32_t jpos,
                                           uint64_t lo_base)
{
    constexpr int E = kNttPerThread;
    const uint32_t RB = jpos + b0;
    const uint32_t ubase = ((t >> RB) << (RB + 4)) | (t & ((1u << RB) - 1));
    uint64_t x[E];
#pragma unroll
    for (int e = 0; e < E; e++) x[e] = lds[lds_slot(ubase | ((uint32_t)e << RB))];
    // lane constants: j-bits below b0 and the column's low in-block bits
    const uint32_t j0 = (!CONTIG ? (ubase >> a.c) : (ubase & ((1u << a.m) - 1))) & ((1u << b0) - 1);
    const uint64_t lo_part = !CONTIG ? lo_base + (ubase & ((1u << a.c) - 1)) : 0;
    const uint64_t base = ((uint64_t)j0 << a.s_lo) | lo_part;
#pragma unroll
    for (int dd = 3; dd >= 0; dd--) {
        const uint32_t d = b0 + (uint32_t)dd;
        const uint32_t lg = a.s_lo + d + 1;
        uint64_t w[8];  // 2^dd used
        if constexpr (CONTIG) {
            // contiguous pass: w_{2^(d+1)}^x = w_4096^(x << (11 - d))
#pragma unroll
            for (int v = 0; v < (1 << dd); v++)
                w[v] = tw12_at(a, (j0 | ((uint32_t)v << b0)) << (kNttTileLog - 1 - d));
        } else {
            const uint64_t *tab = stage_table(a, lg);
#pragma unroll
            for (int v = 0; v < (1 << dd); v++) w[v] = tw_at(tab, lg, base + ((uint64_t)v << (b0 + a.s_lo)), a.inverse);
        }
#pragma unroll
        for (int e = 0; e < E; e++) {
            if ((e >> dd) & 1) continue;
            const int f = e | (1 << dd);
            const uint64_t y0 = x[e], y1 = x[f];
            x[e] = gl::add(y0, y1);
            x[f] = gl::mul(gl::sub(y0, y1), w[e & ((1 << dd) - 1)]);
        }
    }
#pragma unroll
    for (int e = 0; e < E; e++) lds[lds_slot(ubase | ((uint32_t)e << RB))] = x[e];
}

// ---- last register round of a contiguous pass
//
// With b0 = 0 the round covers the K lowest stages; stage d's twiddle for
// element e is w_{2^(d+1)}^v with v = e mod 2^d, a root of order <= 16:
// w = 1 (no product) or a power of two (a shift, mul_pow2).

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// x * w_{2^(D+1)}^V (w^-V for the inverse), x canonical
template <int D, int V, bool INV>
__device__ __forceinline__ uint64_t small_root_mul(uint64_t x)
{
    constexpr int L = gl::clog2(gl::comega(D + 1));
    static_assert(L >= 0, "the 2^(D+1)-th root of unity of the transform is a power of two");
    constexpr int E0 = (V * L) % 192;
    constexpr int E = INV ? (192 - E0) % 192 : E0;
    if constexpr (E == 0) return x;
    else return gl::mul_pow2<E>(x);
}

template <int DD, bool INV>
__device__ __forceinline__ void last_stages(uint64_t (&x)[kNttPerThread])
{
    static_for<0, kNttPerThread>([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        if constexpr (!((e >> DD) & 1)) {
            constexpr int f = e | (1 << DD);
            const uint64_t y0 = x[e], y1 = x[f];
            x[e] = gl::add(y0, y1);
            x[f] = small_root_mul<DD, (e & ((1 << DD) - 1)), INV>(gl::sub(y0, y1));
        }
    });
    if constexpr (DD > 0) last_stages<DD - 1, INV>(x);
}

// The round on stages K-1..0 of a contiguous pass (lane t holds the 2^K
// elements u = (other << K) | rb, other = t | group << 8): ntt_round /
// ntt_round4 with b0 = 0, twiddles folded in at compile time.
template <int K, bool INV>
__device__ __forceinline__ void ntt_round_last(uint64_t *lds, uint32_t t)
{
    constexpr int E = kNttPerThread;
    uint64_t x[E];
    uint32_t uu[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint32_t rb = (uint32_t)e & ((1u << K) - 1);
        const uint32_t other = t | (((uint32_t)e >> K) << 8);
        uu[e] = (other << K) | rb;
        x[e] = lds[lds_slot(uu[e])];
    }
    last_stages<K - 1, INV>(x);
#pragma unroll
    for (int e = 0; e < E; e++) lds[lds_slot(uu[e])] = x[e];
}

template <int K>
__device__ __forceinline__ void ntt_round_last(uint64_t *lds, uint32_t t, bool inverse)
{
    if (inverse) ntt_round_last<K, true>(lds, t);
    else ntt_round_last<K, false>(lds, t);
}

// `passes` (npasses > 0, device memory) lists independent passes sorted by
// tile_begin; with npasses == 0 the launch runs the single pass `one`.
// CONTIG: every pass of the launch is contiguous (s_lo == 0), else strided.
// Work-group k runs tile first + k of the list (a launch may cover a range).
// Four work-groups per CU: 33 KB of LDS and at most 128 VGPRs each.
template <bool CONTIG>
__global__ void __launch_bounds__(kNttThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_ntt_tiles(const NttPass *__restrict__ passes, uint32_t npasses, NttPass one, uint64_t first)
{
    __shared__ uint64_t lds[(1 << kNttTileLog) + (1 << kNttTileLog) / 32];

    // locate this work-group's pass (uniform)
    const uint64_t bid = first + blockIdx.x;
    uint32_t lo = 0, hi = npasses;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (passes[mid].tile_begin <= bid) lo = mid;
        else hi = mid;
    }
    const NttPass a = npasses ? passes[lo] : one;
    const uint64_t tile = bid - a.tile_begin;
    const uint32_t t = threadIdx.x;

    // ---- load, coalesced: lane t takes positions t + 256 i, all 16 in flight,
    // at its own index plus a uniform step.  Only a contiguous pass over
    // blocks shorter than a tile can end inside a tile (strided passes have
    // L > 12, so every tile is full).
    const uint64_t ibase = tile_index(a, tile, t);
    const bool ragged = CONTIG && (a.total & ((1ull << kNttTileLog) - 1)) != 0;
    const gptr<uint64_t> bufp = gmem(a.buf) + ibase;
    // A strided pass reading through a mode is the first pass (s_lo + m = L):
    // its tile lies in one block of 2^L, so the LDE source row and the factor
    // (or coset) row follow the index at fixed offsets.
    gptr<const uint64_t> srcp = nullptr, facp = nullptr;
    if (!CONTIG && a.mode != kSrcInPlace) {
        const uint64_t blk = ibase >> a.L;
        const uint64_t pos0 = ibase - (blk << a.L);
        if (a.mode == kSrcLde) {
            // first `rate` DIF stages of the zero-padded column: block b of
            // column col holds a_i * F[b*n + i], F the resident factor table
            srcp = gmem(a.src) + ((blk >> a.rate) << a.L) + pos0;
            facp = gmem(a.coset) + ((blk & ((1ull << a.rate) - 1)) << a.L) + pos0;
        } else {
            facp = gmem(a.coset) + pos0;
        }
    }
#pragma unroll
    for (int i = 0; i < kNttPerThread; i++) {
        const uint32_t u = t + kNttThreads * i;
        const uint64_t st = tile_step(a, i);
        uint64_t v = 0;
        if constexpr (CONTIG) {
            const uint64_t idx = ibase + st;
            if (!ragged || idx < a.total) {
                if (a.mode == kSrcLde) {
                    const uint64_t n = 1ull << a.L;
                    const uint64_t col = idx >> (a.L + a.rate), pos = idx & (n - 1);
                    v = gl::mul(gmem(a.src)[(col << a.L) + pos], gmem(a.coset)[idx & ((n << a.rate) - 1)]);
                } else {
                    v = bufp[st];
                    if (a.mode == kSrcCoset) v = gl::mul(v, gmem(a.coset)[idx & ((1ull << a.L) - 1)]);
                }
            }
        } else {
            if (a.mode == kSrcLde) {
                v = gl::mul(srcp[st], facp[st]);
            } else {
                v = bufp[st];
                if (a.mode == kSrcCoset) v = gl::mul(v, facp[st]);
            }
        }
        lds[lds_slot(u)] = v;
    }
    __syncthreads();

    // low in-block bits shared by the tile's columns (strided passes)
    const uint64_t lo_base = !CONTIG ? ((tile & ((1ull << (a.s_lo - a.c)) - 1)) << a.c) : 0;
    const uint32_t jpos = !CONTIG ? a.c : 0;  // u-bit of j-bit 0

    // ---- register rounds, top j-bits first: 4 stages each, then the rest
    uint32_t rem = a.m;
    while (rem >= 4) {
        rem -= 4;
        if (CONTIG && rem == 0) ntt_round_last<4>(lds, t, a.inverse);
        else ntt_round4<CONTIG>(lds, a, t, rem, jpos, lo_base);
        __syncthreads();
    }
    if constexpr (CONTIG) {
        if (rem == 3) ntt_round_last<3>(lds, t, a.inverse);
        else if (rem == 2) ntt_round_last<2>(lds, t, a.inverse);
        else if (rem == 1) ntt_round_last<1>(lds, t, a.inverse);
    } else {
        if (rem == 3) ntt_round<CONTIG, 3>(lds, a, t, 0, jpos, lo_base);
        else if (rem == 2) ntt_round<CONTIG, 2>(lds, a, t, 0, jpos, lo_base);
        else if (rem == 1) ntt_round<CONTIG, 1>(lds, a, t, 0, jpos, lo_base);
    }
    if (rem) __syncthreads();

    // ---- store, coalesced
#pragma unroll
    for (int i = 0; i < kNttPerThread; i++) {
        const uint32_t u = t + kNttThreads * i;
        const uint64_t st = tile_step(a, i);
        if (!ragged || ibase + st < a.total) bufp[st] = lds[lds_slot(u)];
    }
}

// ---- descriptors of the one-tile-per-work-group kernels
//
// A launch's descriptor list is read through the constant address space
// (scalar loads into SGPRs; as vector loads they cost 32 VGPRs).

__device__ __forceinline__ NttPass pass_at(const NttPass *passes, uint32_t npasses, const NttPass &one, uint64_t b)
{
    if (!npasses) return one;
    const cptr<NttPass> cp = cmem(passes);
    uint32_t lo = 0, hi = npasses;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cp[mid].tile_begin <= b) lo = mid;
        else hi = mid;
    }
    static_assert(sizeof(NttPass) % 8 == 0, "descriptor copied as 64-bit words");
    constexpr int W = sizeof(NttPass) / 8;
    const cptr<uint64_t> src = (cptr<uint64_t>)(cp + lo);
    uint64_t w[W];
#pragma unroll
    for (int i = 0; i < W; i++) w[i] = src[i];
    NttPass r;
    __builtin_memcpy(&r, w, sizeof r);
    return r;
}

// ---- contiguous in-place passes over whole tiles, one tile per work-group
//
// Lane t loads the tile's words t + 256 i, which are, for the stages 8..m-1
// of the pass (m > 8), the butterfly groups of those stages: element i
// differs from its partner of stage 8 + dd in bit dd of i, and its twiddle
// w_{2^(9+dd)}^(j mod 2^(8+dd)), j mod 2^(8+dd) = (i mod 2^dd) << 8 | t, is
// w_4096^(that << (3 - dd)) from the global table (16 KB of it in use).  So
// those stages run on the loaded registers and only the stages below 8 go
// through LDS (two register rounds, the last with its small roots as
// shifts).  Four work-groups per CU at 127 VGPRs.  This replaced a
// persistent kernel that prefetched the next tile and staged the stage
// twiddles in LDS (three waves per SIMD, 164-168 VGPRs): batched LDE NTT
// 5.13 -> 4.75 ms per 16 commits, the plain 12-stage pass 0.216 -> 0.198 ms,
// same GPU, alternating builds.  The kernel takes in-place passes over whole
// tiles only (ntt_kernel_kind 1); ragged, coset and LDE-source contiguous passes
// run on k_ntt_tiles.
__global__ void __launch_bounds__(kNttThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_ntt_contig_reg(const NttPass *__restrict__ passes, uint32_t npasses, NttPass one, uint64_t first)
{
    __shared__ uint64_t lds[(1 << kNttTileLog) + (1 << kNttTileLog) / 32];
    const uint32_t t = threadIdx.x;
    const uint64_t b = first + blockIdx.x;
    const NttPass a = pass_at(passes, npasses, one, b);
    const uint64_t tile = b - a.tile_begin;
    const gptr<uint64_t> p = gmem(a.buf) + (tile << kNttTileLog) + t;
    uint64_t x[kNttPerThread];
#pragma unroll
    for (int i = 0; i < kNttPerThread; i++) x[i] = p[i << 8];
#pragma unroll
    for (int dd = 3; dd >= 0; dd--) {
        if ((uint32_t)dd + 8 >= a.m) continue;
        uint64_t w[8];
#pragma unroll
        for (int v = 0; v < (1 << dd); v++) w[v] = tw12_at(a, (((uint32_t)v << 8) | t) << (3 - dd));
#pragma unroll
        for (int e = 0; e < kNttPerThread; e++) {
            if ((e >> dd) & 1) continue;
            const int f = e | (1 << dd);
            const uint64_t y0 = x[e], y1 = x[f];
            x[e] = gl::add(y0, y1);
            x[f] = gl::mul(gl::sub(y0, y1), w[e & ((1 << dd) - 1)]);
        }
    }
#pragma unroll
    for (int i = 0; i < kNttPerThread; i++) lds[lds_slot(t + kNttThreads * i)] = x[i];
    __syncthreads();
    uint32_t rem = a.m < 8 ? a.m : 8u;
    while (rem >= 4) {
        rem -= 4;
        if (rem == 0) ntt_round_last<4>(lds, t, a.inverse);
        else ntt_round4<true>(lds, a, t, rem, 0, 0);
        __syncthreads();
    }
    if (rem == 3) ntt_round_last<3>(lds, t, a.inverse);
    else if (rem == 2) ntt_round_last<2>(lds, t, a.inverse);
    else if (rem == 1) ntt_round_last<1>(lds, t, a.inverse);
    if (rem) __syncthreads();
#pragma unroll
    for (int i = 0; i < kNttPerThread; i++) p[i << 8] = lds[lds_slot(t + kNttThreads * i)];
}

// ---- strided single-round passes, one tile per work-group, registers only
//
// A strided pass of m = K <= 4 stages is one register round (b0 = 0, RB = c =
// 12 - K >= 8): lane t's loaded positions t + 256 i are u = (rb << c) | o,
// o = t | g << 8, i = rb << (4 - K) | g (rb < 2^K, g < 2^(4-K)), so every
// butterfly group of the round lies in one lane's registers: no LDS and no
// barrier.  The twiddle of stage dd for rb is
//   W(dd, g, v) = w_{2^(s_lo+dd+1)}^((v << s_lo) | (lo_base + o)),  v = rb mod 2^dd,
// at most 15 values per lane, loaded with the tile's words.  A persistent
// form (next tile prefetched, the tile staged in LDS) held three waves per
// SIMD and ran the 2^18 x 135 LDE pass 10% slower than this one at four or
// five.  MODE is the pass's source (kSrcInPlace or kSrcLde), K its stage
// count; a launch takes passes of one mode and one stage count
// (ntt_kernel_kind).

// tile_step of a strided pass with m = K <= 4 stages (c = 12 - K >= 8): the
// branches of the general form decided at compile time.
template <int K>
__device__ __forceinline__ uint64_t strided_step(const NttPass &a, int i)
{
    constexpr int cb = 4 - K;  // c - log2(256)
    return ((uint64_t)(i & ((1 << cb) - 1)) << 8) + ((uint64_t)(i >> cb) << a.s_lo);
}

template <int K>
constexpr int strided_ntw()
{
    return ((1 << K) - 1) << (4 - K);  // sum over dd of 2^dd values x 2^(4-K) groups
}

// offset of stage dd's twiddles (stages stored K-1 first, then g-major, v-minor)
template <int K>
constexpr int strided_tw_off(int dd)
{
    int off = 0;
    for (int d = K - 1; d > dd; d--) off += (1 << d) << (4 - K);
    return off;
}

template <int K>
__device__ __forceinline__ void strided_tw_load(const NttPass &a, uint64_t tile, uint32_t t,
                                                uint64_t (&w)[strided_ntw<K>()])
{
    const uint64_t lo_base = (tile & ((1ull << (a.s_lo - a.c)) - 1)) << a.c;
    const cptr<const uint64_t *> tabs = cmem(a.tw_lg);
#pragma unroll
    for (int dd = K - 1; dd >= 0; dd--) {
        const uint32_t lg = a.s_lo + (uint32_t)dd + 1;
        const uint64_t *tab = tabs[lg];
#pragma unroll
        for (int g = 0; g < (1 << (4 - K)); g++)
#pragma unroll
            for (int v = 0; v < (1 << dd); v++) {
                const uint64_t ex = ((uint64_t)v << a.s_lo) | (lo_base + (t | ((uint32_t)g << 8)));
                w[strided_tw_off<K>(dd) + (g << dd) + v] = tw_at(tab, lg, ex, a.inverse);
            }
    }
}

// Row pointers of a strided tile: the buffer, and for an LDE pass the
// coefficient row and factor row (the tile lies in one block of 2^L).
struct StridedRows {
    gptr<uint64_t> buf;
    gptr<const uint64_t> src, fac;
};

template <int MODE>
__device__ __forceinline__ StridedRows strided_rows(const NttPass &a, uint64_t tile, uint32_t t)
{
    const uint64_t ibase = tile_index(a, tile, t);
    StridedRows r;
    r.buf = gmem(a.buf) + ibase;
    if constexpr (MODE == kSrcLde) {
        const uint64_t blk = ibase >> a.L;
        const uint64_t pos0 = ibase - (blk << a.L);
        r.src = gmem(a.src) + ((blk >> a.rate) << a.L) + pos0;
        r.fac = gmem(a.coset) + ((blk & ((1ull << a.rate) - 1)) << a.L) + pos0;
    } else {
        r.src = r.buf;
        r.fac = nullptr;
    }
    return r;
}

// Work-group k runs tile first + k of the list.
template <int MODE, int K>
__global__ void __launch_bounds__(kNttThreads)
k_ntt_strided1_tiles(const NttPass *__restrict__ passes, uint32_t npasses, NttPass one, uint64_t first)
{
    constexpr int CB = 4 - K;
    const uint32_t t = threadIdx.x;
    const uint64_t b = first + blockIdx.x;
    const NttPass a = pass_at(passes, npasses, one, b);
    const uint64_t tile = b - a.tile_begin;
    uint64_t w[strided_ntw<K>()];
    strided_tw_load<K>(a, tile, t, w);
    const StridedRows r = strided_rows<MODE>(a, tile, t);
    uint64_t x[kNttPerThread];
#pragma unroll
    for (int i = 0; i < kNttPerThread; i++) x[i] = r.src[strided_step<K>(a, i)];
    if constexpr (MODE == kSrcLde) {
#pragma unroll
        for (int i = 0; i < kNttPerThread; i++) x[i] = gl::mul(x[i], r.fac[strided_step<K>(a, i)]);
    }
#pragma unroll
    for (int dd = K - 1; dd >= 0; dd--) {
#pragma unroll
        for (int i = 0; i < kNttPerThread; i++) {
            const int rb = i >> CB, g = i & ((1 << CB) - 1);
            if ((rb >> dd) & 1) continue;
            const int f = i | (1 << (dd + CB));
            const uint64_t wv = w[strided_tw_off<K>(dd) + (g << dd) + (rb & ((1 << dd) - 1))];
            const uint64_t y0 = x[i], y1 = x[f];
            x[i] = gl::add(y0, y1);
            x[f] = gl::mul(gl::sub(y0, y1), wv);
        }
    }
#pragma unroll
    for (int i = 0; i < kNttPerThread; i++) r.buf[strided_step<K>(a, i)] = x[i];
}

// ---- strided two-round passes, one tile per work-group
//
// A strided pass of m = M in 5..8 stages has c = 12 - M < 8 columns per
// tile: lane t loads positions u = t + 256 i, i.e. column q = t mod 2^c and
// butterfly index j = (i << (M - 4)) | j0, j0 = t >> c.  Its 16 registers
// therefore already hold the 16 elements that differ in the top four j-bits,
// so the first radix-16 round (stages M-1..M-4) runs on the loaded registers
// with no LDS exchange; the second (stages K-1..0, K = M - 4) runs through
// LDS as ntt_round does.  Every twiddle depends only on the lane and the
// tile (15 for the first round, 2^K - 1 for the second; none on the
// register group of the second round, whose column bits are the lane's):
//   first:  w_{2^(s_lo+d+1)}^(((v << K | j0) << s_lo) | lo),  d = K + dd, v < 2^dd
//   second: w_{2^(s_lo+dd+1)}^((v << s_lo) | lo),             v < 2^dd
// with lo = lo_base + q, all loaded with the tile's words.  Per tile: loads |
// round 1 in registers | write | barrier | round 2 | barrier | stores: 20%
// fewer VALU instructions than ntt_round4 + ntt_round over LDS, at four
// work-groups per CU (a persistent, prefetching form needs ~230 VGPRs for
// the words in flight and the 30 twiddles, two waves per SIMD, and ran 9%
// slower).

template <int M>
constexpr int strided2_nwb()
{
    return (1 << (M - 4)) - 1;
}

// offset of stage dd's values in a table of stages n-1..0 (2^dd values each)
constexpr int stage_off(int n, int dd)
{
    int off = 0;
    for (int d = n - 1; d > dd; d--) off += 1 << d;
    return off;
}

template <int M>
__device__ __forceinline__ void strided2_tw_load(const NttPass &a, uint64_t tile, uint32_t t, uint64_t (&wa)[15],
                                                 uint64_t (&wb)[strided2_nwb<M>()])
{
    constexpr int K = M - 4, C = kNttTileLog - M;
    const uint64_t lo = ((tile & ((1ull << (a.s_lo - C)) - 1)) << C) + (t & ((1u << C) - 1));
    const uint64_t j0 = t >> C;
    const cptr<const uint64_t *> tabs = cmem(a.tw_lg);
#pragma unroll
    for (int dd = 3; dd >= 0; dd--) {
        const uint32_t lg = a.s_lo + (uint32_t)(K + dd) + 1;
        const uint64_t *tab = tabs[lg];
#pragma unroll
        for (int v = 0; v < (1 << dd); v++)
            wa[stage_off(4, dd) + v] = tw_at(tab, lg, ((((uint64_t)v << K) | j0) << a.s_lo) | lo, a.inverse);
    }
#pragma unroll
    for (int dd = K - 1; dd >= 0; dd--) {
        const uint32_t lg = a.s_lo + (uint32_t)dd + 1;
        const uint64_t *tab = tabs[lg];
#pragma unroll
        for (int v = 0; v < (1 << dd); v++)
            wb[stage_off(K, dd) + v] = tw_at(tab, lg, ((uint64_t)v << a.s_lo) | lo, a.inverse);
    }
}

// Butterfly product of the two-round kernels: the carry form (gl::mul) but
// in the in-place 8-stage kernel, where its schedule exceeds the 128 VGPRs
// of four waves per SIMD (20 VGPRs spilled, against 6 with gl::mul_nc; the
// 64 x 2^20 forward NTT took 1.03 against 0.95 ms).
template <bool NC>
__device__ __forceinline__ uint64_t strided2_mul(uint64_t a, uint64_t b)
{
    if constexpr (NC) return gl::mul_nc(a, b);
    else return gl::mul(a, b);
}

// First round: stages K+3..K on the registers (element i = top j-bits).
template <bool NC>
__device__ __forceinline__ void strided2_round1(uint64_t (&x)[kNttPerThread], const uint64_t (&wa)[15])
{
#pragma unroll
    for (int dd = 3; dd >= 0; dd--) {
#pragma unroll
        for (int e = 0; e < kNttPerThread; e++) {
            if ((e >> dd) & 1) continue;
            const int f = e | (1 << dd);
            const uint64_t y0 = x[e], y1 = x[f];
            x[e] = gl::add(y0, y1);
            x[f] = strided2_mul<NC>(gl::sub(y0, y1), wa[stage_off(4, dd) + (e & ((1 << dd) - 1))]);
        }
    }
}

// Second round: stages K-1..0 through LDS (ntt_round with b0 = 0, RB = c).
template <int M, bool NC>
__device__ __forceinline__ void strided2_round2(uint64_t *lds, uint32_t t, const uint64_t (&wb)[strided2_nwb<M>()])
{
    constexpr int K = M - 4, C = kNttTileLog - M, E = kNttPerThread;
    uint64_t x[E];
    uint32_t sl[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint32_t rb = (uint32_t)e & ((1u << K) - 1);
        const uint32_t other = t | (((uint32_t)e >> K) << 8);
        sl[e] = lds_slot(((other >> C) << (C + K)) | (rb << C) | (other & ((1u << C) - 1)));
        x[e] = lds[sl[e]];
    }
#pragma unroll
    for (int dd = K - 1; dd >= 0; dd--) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            if ((e >> dd) & 1) continue;
            const int f = e | (1 << dd);
            const uint64_t y0 = x[e], y1 = x[f];
            x[e] = gl::add(y0, y1);
            x[f] = strided2_mul<NC>(gl::sub(y0, y1), wb[stage_off(K, dd) + (e & ((1 << dd) - 1))]);
        }
    }
#pragma unroll
    for (int e = 0; e < E; e++) lds[sl[e]] = x[e];
}

// tile_step of a strided pass with m = M >= 5 stages (c = 12 - M < 8)
template <int M>
__device__ __forceinline__ uint64_t strided2_step(const NttPass &a, int i)
{
    return (uint64_t)i << (M - 4 + a.s_lo);
}

// Rows of a two-round strided tile split into wave-uniform tile bases and
// the lane's 32-bit offset ((j0 << s_lo) + q < 2^(L-4) + 2^s_lo): element i
// of lane t sits at base + (i << (K + s_lo)) + off.
struct Strided2Rows {
    const uint64_t *src, *fac;
    uint64_t *buf;
};

template <int MODE, int M>
__device__ __forceinline__ Strided2Rows strided2_rows(const NttPass &a, uint64_t tile)
{
    constexpr int C = kNttTileLog - M;
    const uint32_t chunk_bits = a.s_lo - C;
    const uint64_t g = tile >> chunk_bits, ch = tile & ((1ull << chunk_bits) - 1);
    Strided2Rows r;
    // group g of 2^(s_lo + M) elements (one block of 2^L in the first pass,
    // where an LDE pass reads: s_lo + M = L)
    r.buf = a.buf + (g << (a.s_lo + M)) + (ch << C);
    if constexpr (MODE == kSrcLde) {
        r.src = a.src + ((g >> a.rate) << a.L) + (ch << C);
        r.fac = a.coset + ((g & ((1ull << a.rate) - 1)) << a.L) + (ch << C);
    } else {
        r.src = r.buf;
        r.fac = nullptr;
    }
    return r;
}

template <int M>
__device__ __forceinline__ uint32_t strided2_off(const NttPass &a, uint32_t t)
{
    constexpr int C = kNttTileLog - M;
    return ((t >> C) << a.s_lo) + (t & ((1u << C) - 1));
}

// Work-group k runs tile first + k of the list (passes of one source mode
// and M stages; grid = the tiles of the range).
template <int MODE, int M>
__global__ void __launch_bounds__(kNttThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_ntt_strided2_tiles(const NttPass *__restrict__ passes, uint32_t npasses, NttPass one, uint64_t first)
{
    __shared__ uint64_t lds[(1 << kNttTileLog) + (1 << kNttTileLog) / 32];
    const uint32_t t = threadIdx.x;
    const uint64_t b = first + blockIdx.x;
    const NttPass a = pass_at(passes, npasses, one, b);
    const uint64_t tile = b - a.tile_begin;
    uint64_t wa[15], wb[strided2_nwb<M>()];
    strided2_tw_load<M>(a, tile, t, wa, wb);
    uint64_t x[kNttPerThread];
    const Strided2Rows r = strided2_rows<MODE, M>(a, tile);
    const uint32_t off = strided2_off<M>(a, t);
#pragma unroll
    for (int i = 0; i < kNttPerThread; i++) x[i] = gmem(r.src + strided2_step<M>(a, i))[off];
    if constexpr (MODE == kSrcLde) {
#pragma unroll
        for (int i = 0; i < kNttPerThread; i++) x[i] = gl::mul(x[i], gmem(r.fac + strided2_step<M>(a, i))[off]);
    }
    constexpr bool NC = MODE == kSrcInPlace && M == 8;
    strided2_round1<NC>(x, wa);
#pragma unroll
    for (int i = 0; i < kNttPerThread; i++) lds[lds_slot(t + kNttThreads * i)] = x[i];
    __syncthreads();
    strided2_round2<M, NC>(lds, t, wb);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kNttPerThread; i++)
        gmem(r.buf + strided2_step<M>(a, i))[off] = lds[lds_slot(t + kNttThreads * i)];
}

// In-place bit reversal of every column of length 2^lg (lg >= 2*kBrTile),
// with optional post-scaling of element at natural index t by
// scale * (coset ? g^t : 1).  Index bits are (h: 5 | mid | l: 5); the tile for
// `mid` and the tile for bitrev(mid) swap, transposed in LDS.
constexpr int kBrTile = 5;

struct BitrevArgs {
    uint64_t *buf;
    const uint64_t *coset;
    uint64_t scale;       // 1 = none
    uint32_t lg;
    uint32_t cols;
    uint32_t use_scale;
    uint32_t use_coset;
};

__device__ __forceinline__ uint64_t br_post(const BitrevArgs &a, uint64_t v, uint64_t t)
{
    if (a.use_scale) v = gl::mul(v, a.scale);
    if (a.use_coset) v = gl::mul(v, gmem(a.coset)[t]);
    return v;
}

// One lane's share of a tile pair: the 2 x 4 loads in flight before the
// first LDS write, and no branch inside the loops (the pair / self-paired
// case and the post-scaling are decided outside them; a branch between
// loads would leave one load per lane in flight).
template <bool PAIR, bool POST>
__device__ __forceinline__ void bitrev_pair(const BitrevArgs &a, gptr<uint64_t> base, uint32_t hs, uint32_t mid,
                                            uint32_t rmid, uint64_t (*ta)[33], uint64_t (*tb)[33])
{
    constexpr uint32_t T = 1u << kBrTile;
    constexpr int PER = T * T / kNttThreads;
    const uint32_t t = threadIdx.x;
    uint64_t va[PER], vb[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t e = t + kNttThreads * k, h = e >> kBrTile, l = e & (T - 1);
        va[k] = base[((uint64_t)h << hs) | ((uint64_t)mid << kBrTile) | l];
        if constexpr (PAIR) vb[k] = base[((uint64_t)h << hs) | ((uint64_t)rmid << kBrTile) | l];
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t e = t + kNttThreads * k, h = e >> kBrTile, l = e & (T - 1);
        ta[h][l] = va[k];
        if constexpr (PAIR) tb[h][l] = vb[k];
    }
    __syncthreads();
    // element (h, l) of tile(mid) goes to (rev(l), rmid, rev(h))
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint32_t e = t + kNttThreads * k;
        const uint32_t oh = e >> kBrTile, ol = e & (T - 1);  // destination (h', l')
        const uint32_t sh = brev_bits(ol, kBrTile), sl = brev_bits(oh, kBrTile);
        const uint64_t t1 = ((uint64_t)oh << hs) | ((uint64_t)rmid << kBrTile) | ol;
        uint64_t v1 = ta[sh][sl];
        if constexpr (POST) v1 = br_post(a, v1, t1);
        base[t1] = v1;
        if constexpr (PAIR) {
            const uint64_t t2 = ((uint64_t)oh << hs) | ((uint64_t)mid << kBrTile) | ol;
            uint64_t v2 = tb[sh][sl];
            if constexpr (POST) v2 = br_post(a, v2, t2);
            base[t2] = v2;
        }
    }
}

__global__ void __launch_bounds__(kNttThreads) k_bitrev_tiles(BitrevArgs a)
{
    constexpr uint32_t T = 1u << kBrTile;
    __shared__ uint64_t ta[T][T + 1];
    __shared__ uint64_t tb[T][T + 1];
    const uint32_t mid_bits = a.lg - 2 * kBrTile;
    const uint64_t n = 1ull << a.lg;
    const uint64_t col = blockIdx.x >> mid_bits;
    const uint32_t mid = blockIdx.x & ((1u << mid_bits) - 1);
    const uint32_t rmid = brev_bits(mid, mid_bits);
    if (rmid < mid) return;  // the partner tile's block does both
    const gptr<uint64_t> base = gmem(a.buf) + col * n;
    const uint32_t hs = a.lg - kBrTile;
    const bool post = a.use_scale || a.use_coset;
    if (rmid != mid) {
        if (post) bitrev_pair<true, true>(a, base, hs, mid, rmid, ta, tb);
        else bitrev_pair<true, false>(a, base, hs, mid, rmid, ta, tb);
    } else {
        if (post) bitrev_pair<false, true>(a, base, hs, mid, rmid, ta, tb);
        else bitrev_pair<false, false>(a, base, hs, mid, rmid, ta, tb);
    }
}

// Small columns (lg < 2*kBrTile): one thread per element pair.
__global__ void __launch_bounds__(kNttThreads) k_bitrev_small(BitrevArgs a)
{
    const uint64_t n = 1ull << a.lg;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * a.cols) return;
    const uint64_t col = g >> a.lg, i = g & (n - 1);
    const uint64_t r = brev_bits((uint32_t)i, a.lg);
    if (r < i) return;
    const gptr<uint64_t> base = gmem(a.buf) + col * n;
    const uint64_t x = base[i], y = base[r];
    base[i] = br_post(a, y, i);
    if (r != i) base[r] = br_post(a, x, r);
}

// out[t * B + b] = in[b * 2^lg + bitrev_lg(t)]  (transpose_rev_kernel,
// ntt/kernels.cu:94-135).  Tile: 32 consecutive source positions j of 32
// columns; loads are 256 B runs of one column, stores are runs of 32
// consecutive columns of row bitrev(j).
constexpr int kTrTile = 32;

__global__ void __launch_bounds__(kNttThreads) k_transpose_rev(uint64_t *__restrict__ out,
                                                             const uint64_t *__restrict__ in,
                                                             uint32_t lg, uint32_t B, uint32_t rev)
{
    __shared__ uint64_t t[kTrTile][kTrTile + 1];
    const uint64_t n = 1ull << lg;
    const uint64_t j0 = (uint64_t)blockIdx.x * kTrTile;
    const uint32_t b0 = blockIdx.y * kTrTile;
    for (uint32_t e = threadIdx.x; e < kTrTile * kTrTile; e += kNttThreads) {
        const uint32_t bl = e / kTrTile, jl = e % kTrTile;
        const uint64_t j = j0 + jl;
        if (b0 + bl < B && j < n) t[bl][jl] = in[(uint64_t)(b0 + bl) * n + j];
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < kTrTile * kTrTile; e += kNttThreads) {
        const uint32_t jl = e / kTrTile, bl = e % kTrTile;
        const uint64_t j = j0 + jl;
        if (b0 + bl < B && j < n) {
            const uint64_t row = rev ? brev_bits((uint32_t)j, lg) : j;
            out[row * B + b0 + bl] = t[bl][jl];
        }
    }
}

// The same transpose for 2^lg >= 32 rows, whole output rows at a time: a
// tile is 32 consecutive source positions j of `cb` columns (all of B when
// B <= 136), so each output row bitrev(j) receives one contiguous run of cb
// words and every lane keeps NE loads in flight (NE = cb * 32 / 256 rounded
// up; positions past the tile's columns re-read its last word, in bounds).
// Stores: PW lanes per row (the power of two >= cb, at most 256), 256 / PW
// rows per step.
constexpr uint32_t kTrRows = 32;
constexpr uint32_t kTrMaxCols = 136;

template <int NE, int PW>
__global__ void __launch_bounds__(kNttThreads) k_transpose_rows(uint64_t *__restrict__ out,
                                                              const uint64_t *__restrict__ in, uint32_t lg,
                                                              uint32_t B, uint32_t rev, uint32_t cb)
{
    static_assert(NE * kNttThreads >= kTrMaxCols * kTrRows || NE < 17, "lanes cover the tile");
    __shared__ uint64_t tl[kTrMaxCols][kTrRows + 1];
    const uint64_t n = 1ull << lg;
    const uint64_t j0 = (uint64_t)blockIdx.x * kTrRows;
    const uint32_t b0 = blockIdx.y * cb;
    const uint32_t nc = B - b0 < cb ? B - b0 : cb;
    const uint32_t tot = nc * kTrRows;
    const uint32_t t = threadIdx.x;
    const gptr<const uint64_t> src = gmem(in) + (uint64_t)b0 * n + j0;
    uint64_t v[NE];
#pragma unroll
    for (int k = 0; k < NE; k++) {
        uint32_t e = t + kNttThreads * k;
        e = e < tot ? e : tot - 1;
        v[k] = src[(uint64_t)(e >> 5) * n + (e & 31)];
    }
#pragma unroll
    for (int k = 0; k < NE; k++) {
        const uint32_t e = t + kNttThreads * k;
        if (e < tot) tl[e >> 5][e & 31] = v[k];
    }
    __syncthreads();
    constexpr uint32_t RPS = kNttThreads / PW;  // rows per step
    const uint32_t r = t / PW, c = t % PW;
    if (c >= nc) return;
#pragma unroll
    for (uint32_t k = 0; k < kTrRows / RPS; k++) {
        const uint32_t jl = r + k * RPS;
        const uint64_t j = j0 + jl;
        const uint64_t row = rev ? brev_bits((uint32_t)j, lg) : j;
        gmem(out)[row * B + b0 + c] = tl[c][jl];
    }
}

// Salt columns appended after an LDE (gen_random_salt_kernel,
// ntt/kernels.cu:62-84): salt_size columns of n field elements drawn from a
// per-call seed (the reference seeds from the clock, so salt is never part
// of a parity comparison).  splitmix64, rejection into [0, p).
__global__ void __launch_bounds__(kNttThreads) k_salt(uint64_t *out, uint64_t count, uint64_t seed)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= count) return;
    uint64_t x = seed ^ (g * 0xD1B54A32D192ED03ull);
    uint64_t z;
    do {
        x += 0x9E3779B97F4A7C15ull;
        z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
    } while (z >= gl::P);
    out[g] = z;
}

// LDE factor table: F[b*n + i] = c_b^i, c_b = bases[b], b < 2^rate, i < n.
struct FactorBases {
    uint64_t c[1 << kMaxRateBits];
};

__global__ void __launch_bounds__(kNttThreads) k_lde_factors(uint64_t *F, uint32_t lg_n, uint32_t rate,
                                                           FactorBases bases)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = 1ull << lg_n;
    const uint64_t chunk = n < 64 ? n : 64;  // runs never cross a block
    const uint64_t start = t * chunk;
    if (start >= (n << rate)) return;
    const uint32_t b = (uint32_t)(start >> lg_n);
    const uint64_t i0 = start & (n - 1);
    const uint64_t c = bases.c[b];
    uint64_t v = gl::pow(c, i0);
    const uint64_t end = i0 + chunk;
    for (uint64_t i = i0; i < end; i++) {
        F[((uint64_t)b << lg_n) + i] = v;
        v = gl::mul(v, c);
    }
}

// table[i] = base^i for i < n, 64 consecutive powers per thread.
__global__ void __launch_bounds__(kNttThreads) k_powers(uint64_t *table, uint64_t n, uint64_t base)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t start = t * 64;
    if (start >= n) return;
    uint64_t v = gl::pow(base, start);
    const uint64_t end = start + 64 < n ? start + 64 : n;
    for (uint64_t i = start; i < end; i++) {
        table[i] = v;
        v = gl::mul(v, base);
    }
}

}  // namespace zkx
