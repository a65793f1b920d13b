// Goldilocks NTT / LDE / transpose kernels for gfx950.
//
// Layout in HBM: column-major batches, column b of a batch of length 2^lg at
// [b * 2^lg, (b+1) * 2^lg), u64 little-endian canonical elements -- the
// layout of compute_batched_ntt / compute_batched_lde (ntt/ntt.cu:253-309,
// 477-564).
//
// Algorithm.  The forward transform is decimation-in-frequency (natural
// order in, bit-reversed order out), run as a few passes over HBM.  Each pass
// performs up to 12 consecutive radix-2 stages on tiles held in LDS: a tile is
// 2^m elements at stride 2^s_lo (the butterfly partners of stages
// s_lo..s_lo+m-1) times 2^c consecutive columns of that stride, so every
// global load and store is a run of 2^c consecutive u64.
//
// LDE.  Zero-padding n = 2^lg_n coefficients to N = 2^(lg_n+r) and running the
// first r DIF stages leaves, in block b of length n (b < 2^r, memory order),
// the values a_i * (w_N^{bitrev_r(b)})^i: the first pass reads the n
// coefficients directly and applies that factor (and the coset factor g^i),
// so the zero half is never read, written or butterflied.
#pragma once
#include "gl64.h"

namespace zkx {

constexpr int kNttTileLog = 12;  // 4096 u64 = 32 KB of LDS per work-group
constexpr int kNttThreads = 256;

enum SrcMode : int {
    kSrcInPlace = 0,  // tile source is the buffer itself
    kSrcCoset = 1,    // buffer itself, times g^i (forward coset NTT)
    kSrcLde = 2,      // coefficients of the un-extended columns (LDE first pass)
};

struct PassArgs {
    uint64_t *buf;          // flat array of nblk blocks of 2^L
    const uint64_t *src;    // kSrcLde: coefficients [cols][2^L]
    const uint64_t *tw;     // w_T^k, k < 2^lg_tw (w_T = primitive 2^lg_tw-th root)
    const uint64_t *coset;  // g^i table (kSrcCoset / kSrcLde with coset)
    uint64_t total;         // nblk << L
    uint32_t L;             // log2 block length
    uint32_t lg_tw;         // log2 of twiddle table length (>= L + rate)
    uint32_t rate;          // kSrcLde: log2 of blocks per column
    uint32_t s_lo, m, c;    // stages s_lo .. s_lo+m-1; 2^c columns per tile
    uint32_t inverse;       // use w^-1
    uint32_t with_coset;    // kSrcLde forward: multiply by g^i
};

__device__ __forceinline__ uint64_t tw_at(const uint64_t *tw, uint32_t lg_tw, uint64_t k, bool inverse)
{
    const uint64_t mask = (1ull << lg_tw) - 1;
    return tw[inverse ? ((0 - k) & mask) : k];
}

__device__ __forceinline__ uint32_t brev_bits(uint32_t x, uint32_t bits)
{
    return bits ? (__builtin_bitreverse32(x) >> (32 - bits)) : 0u;
}

// Memory index (flat) of tile element (j, q) for tile id `tile`.
__device__ __forceinline__ uint64_t pass_index(const PassArgs &a, uint64_t tile, uint32_t j, uint32_t q)
{
    if (a.s_lo > 0) {
        // groups of 2^(s_lo+m) elements; in a group, (j, lo) -> j*2^s_lo + lo;
        // 2^(s_lo-c) column chunks per group
        const uint32_t chunk_bits = a.s_lo - a.c;
        const uint64_t g = tile >> chunk_bits, ch = tile & ((1ull << chunk_bits) - 1);
        return (g << (a.s_lo + a.m)) + ((uint64_t)j << a.s_lo) + (ch << a.c) + q;
    }
    // s_lo == 0: 2^c consecutive contiguous groups of 2^m
    return (((tile << a.c) + q) << a.m) + j;
}

// In-block position modulo 2^s needed for the twiddle exponent of the upper
// butterfly element: ((j mod 2^d) << s_lo) | lo.
template <int MODE>
__global__ void __launch_bounds__(kNttThreads) k_dif_pass(PassArgs a)
{
    __shared__ uint64_t lds[1 << kNttTileLog];
    const uint32_t tsz_log = a.m + a.c;
    const uint32_t tsz = 1u << tsz_log;
    const uint64_t tile = blockIdx.x;
    const uint32_t cmask = (1u << a.c) - 1;
    const uint32_t mmask = (1u << a.m) - 1;

    // ---- load (consecutive threads -> consecutive memory)
    for (uint32_t e = threadIdx.x; e < tsz; e += kNttThreads) {
        uint32_t j, q;
        if (a.s_lo > 0) { q = e & cmask; j = e >> a.c; }
        else { j = e & mmask; q = e >> a.m; }
        const uint64_t idx = pass_index(a, tile, j, q);
        uint64_t v = 0;
        if (idx < a.total) {
            if constexpr (MODE == kSrcLde) {
                const uint64_t n = 1ull << a.L;
                const uint64_t blk = idx >> a.L, i = idx & (n - 1);
                const uint64_t col = blk >> a.rate;
                const uint32_t b = (uint32_t)(blk & ((1u << a.rate) - 1));
                v = a.src[(col << a.L) + i];
                // factor (w_N^{bitrev_r(b)})^i, plus g^i for a forward coset LDE
                const uint64_t e_tw = ((uint64_t)brev_bits(b, a.rate) * i) << (a.lg_tw - a.L - a.rate);
                if (e_tw) v = gl::mul(v, tw_at(a.tw, a.lg_tw, e_tw, a.inverse));
                if (a.with_coset && !a.inverse) v = gl::mul(v, a.coset[i]);
            } else {
                v = a.buf[idx];
                if constexpr (MODE == kSrcCoset) v = gl::mul(v, a.coset[idx & ((1ull << a.L) - 1)]);
            }
        }
        lds[(j << a.c) + q] = v;
    }
    __syncthreads();

    // ---- stages s = s_lo+m-1 .. s_lo
    const uint32_t nbf = tsz >> 1;
    // low in-block bits shared by the tile's columns (strided passes only)
    const uint64_t lo_base = a.s_lo > 0 ? ((tile & ((1ull << (a.s_lo - a.c)) - 1)) << a.c) : 0;
    for (int d = (int)a.m - 1; d >= 0; d--) {
        const uint32_t s = a.s_lo + (uint32_t)d;
        const uint32_t tw_shift = a.lg_tw - s - 1;
        for (uint32_t u = threadIdx.x; u < nbf; u += kNttThreads) {
            const uint32_t q = u & cmask, jj = u >> a.c;
            const uint32_t j = ((jj >> d) << (d + 1)) | (jj & ((1u << d) - 1));
            const uint64_t lo_part = a.s_lo > 0 ? lo_base + q : 0;
            const uint64_t ex = (((uint64_t)(j & ((1u << d) - 1)) << a.s_lo) | lo_part) << tw_shift;
            const uint32_t i0 = (j << a.c) + q, i1 = ((j + (1u << d)) << a.c) + q;
            const uint64_t x = lds[i0], y = lds[i1];
            lds[i0] = gl::add(x, y);
            const uint64_t diff = gl::sub(x, y);
            lds[i1] = ex ? gl::mul(diff, tw_at(a.tw, a.lg_tw, ex, a.inverse)) : diff;
        }
        __syncthreads();
    }

    // ---- store
    for (uint32_t e = threadIdx.x; e < tsz; e += kNttThreads) {
        uint32_t j, q;
        if (a.s_lo > 0) { q = e & cmask; j = e >> a.c; }
        else { j = e & mmask; q = e >> a.m; }
        const uint64_t idx = pass_index(a, tile, j, q);
        if (idx < a.total) a.buf[idx] = lds[(j << a.c) + q];
    }
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
    if (a.use_coset) v = gl::mul(v, a.coset[t]);
    return v;
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
    uint64_t *base = a.buf + col * n;
    const uint32_t hs = a.lg - kBrTile;

    for (uint32_t e = threadIdx.x; e < T * T; e += kNttThreads) {
        const uint32_t h = e >> kBrTile, l = e & (T - 1);
        ta[h][l] = base[((uint64_t)h << hs) | ((uint64_t)mid << kBrTile) | l];
        if (rmid != mid) tb[h][l] = base[((uint64_t)h << hs) | ((uint64_t)rmid << kBrTile) | l];
    }
    __syncthreads();
    // element (h, l) of tile(mid) goes to (rev(l), rmid, rev(h))
    for (uint32_t e = threadIdx.x; e < T * T; e += kNttThreads) {
        const uint32_t oh = e >> kBrTile, ol = e & (T - 1);  // destination (h', l')
        const uint32_t sh = brev_bits(ol, kBrTile), sl = brev_bits(oh, kBrTile);
        const uint64_t t1 = ((uint64_t)oh << hs) | ((uint64_t)rmid << kBrTile) | ol;
        base[t1] = br_post(a, ta[sh][sl], t1);
        if (rmid != mid) {
            const uint64_t t2 = ((uint64_t)oh << hs) | ((uint64_t)mid << kBrTile) | ol;
            base[t2] = br_post(a, tb[sh][sl], t2);
        }
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
    uint64_t *base = a.buf + col * n;
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
