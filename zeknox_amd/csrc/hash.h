// Poseidon and Poseidon2 permutations (Goldilocks, width 12, rate 8,
// 8 full + 22 partial rounds, x^7 S-box) and the plonky2 sponge, for one
// lane (the 12-element state in VGPRs), and on a pair of lanes with six
// elements each for the tree levels (poseidon_hash_two_pair,
// poseidon2_hash_two_pair).  Rounds run as a rolled loop
// (the unrolled permutation is ~150 KB of code, more than the instruction
// cache); round constants are read through the uniform round counter from
// constant memory, i.e. as scalar loads into SGPRs.
//
// Poseidon      : poseidon/poseidon.cu:563-838 (naive rounds; same function
//                 as the fast-partial-round form of poseidon/poseidon.cpp).
// Poseidon2     : poseidon2/poseidon2.cu:65-209.
// Monolith      : monolith/monolith.cu:104-227 (monolith.h).
// Poseidon-BN128: poseidon/poseidon_bn128.cu:116-204 (poseidon_bn128.h).
// Keccak-256    : keccak/keccak.cu (keccak.h; its own sponge).
// Sponge        : poseidon/poseidon_permutation.hpp:83-139 /
//                 poseidon_permutation.cuh:83-134.
#pragma once
#include "gl64.h"
#include "gl_constants.h"
#include "keccak.h"
#include "merkle_types.h"
#include "monolith.h"
#include "poseidon_bn128.h"

namespace zkx {

// Round constants, [round][12].  The Poseidon table has one extra all-zero
// row: after round r the next round's constants (row r+1) are folded into
// the MDS accumulators, and the last round folds the zero row.
static __constant__ uint64_t c_poseidon_rc[31 * 12] = {ZK_POSEIDON_RC_LIST 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
static __constant__ uint64_t c_poseidon2_rc[30 * 12] = {ZK_POSEIDON2_RC_LIST};

// ---------------------------------------------------------------- Poseidon

// value = A + B * 2^32 (A, B < 2^64 - 2^32, value < 2^96), reduced lazily
// (below 2^64, congruent mod p): the middle word is B + hi(A) formed with
// one v_mad_u64_u32 (x * 1 + y).
__device__ __forceinline__ uint64_t limb_combine(uint64_t A, uint64_t B)
{
    const uint64_t T = gl::add_small(B, gl::hi32(A));  // < 2^64
    return gl::reduce96_lazy(gl::pack(gl::lo32(A), gl::lo32(T)), gl::hi32(T));
}

// state <- MDS * state + k (k: 12 constants, the next round's), with
// MDS = circ(17,15,41,16,2,28,13,13,39,18,34,20) + diag(8,0,...,0).
// Each output row is accumulated in 32-bit limbs, A = k_lo + sum c*lo(x),
// B = k_hi + sum c*hi(x), one v_mad_u64_u32 per product (coefficients are
// inline constants; row coefficient sums <= 264 keep A, B < 2^41 for any
// x < 2^64).  Inputs and outputs are lazy (below 2^64, congruent mod p).
__device__ __forceinline__ void poseidon_mds(uint64_t s[12], const uint64_t *k)
{
    uint32_t lo[12], hi[12];
#pragma unroll
    for (int i = 0; i < 12; i++) {
        lo[i] = gl::lo32(s[i]);
        hi[i] = gl::hi32(s[i]);
    }
    // two rows at a time: four accumulation chains interleaved, so no
    // multiply-add reads an accumulator written by the instruction before
    // last (which would cost a wait state)
#pragma unroll
    for (int r = 0; r < 12; r += 2) {
        uint64_t A[2], B[2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            A[q] = gl::lo32(k[r + q]);
            B[q] = gl::hi32(k[r + q]);
        }
#pragma unroll
        for (int i = 0; i < 12; i++)
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int row = r + q, j = (i + row) % 12;
                const uint32_t c0 =
                    (uint32_t)ZK_POSEIDON_MDS_CIRC[i] + (i == 0 ? (uint32_t)ZK_POSEIDON_MDS_DIAG[row] : 0u);
                // the first product takes the SGPR constant pair as its addend,
                // so its coefficient stays an inline constant (one SGPR per VALU op)
                const uint32_t c = i == 0 ? c0 : gl::opaque(c0);
                // each step pinned: one v_mad_u64_u32 per product, no regrouping
                A[q] = gl::vopaque64((uint64_t)lo[j] * c + A[q]);
                B[q] = gl::vopaque64((uint64_t)hi[j] * c + B[q]);
            }
#pragma unroll
        for (int q = 0; q < 2; q++) s[r + q] = limb_combine(A[q], B[q]);
    }
}

// Lazy in and out: the state stays below 2^64 and congruent mod p; the
// sponge canonicalises what it outputs.  (gl::add(a, b) with a < 2^64 and
// b < p returns a value below 2^64 congruent to a + b.)
__device__ __forceinline__ void poseidon_permute(uint64_t s[12])
{
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = gl::add(s[i], c_poseidon_rc[i]);
    // rounds 0-3 and 26-29 are full (S-box on all 12), 4-25 partial (s0
    // only); the partial rounds run two per iteration, so each round's MDS
    // writes where the next reads (no loop-carried register copies)
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 12; i += 2) gl::pow7_lazy2(s[i], s[i + 1]);
        poseidon_mds(s, c_poseidon_rc + 12 * (r + 1));
    }
#pragma unroll 1
    for (int r = 4; r < 26; r += 2) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            s[0] = gl::pow7_lazy(s[0]);
            poseidon_mds(s, c_poseidon_rc + 12 * (r + h + 1));
        }
    }
#pragma unroll 1
    for (int r = 26; r < 30; r++) {
#pragma unroll
        for (int i = 0; i < 12; i += 2) gl::pow7_lazy2(s[i], s[i + 1]);
        poseidon_mds(s, c_poseidon_rc + 12 * (r + 1));
    }
}

// ------------------------------------------------- Poseidon on a lane pair
//
// Tree levels run at low occupancy: once the leaves are hashed, every level
// of a tile is one permutation deep, so its time is one permutation's
// latency on nearly idle SIMDs, and half the lanes of a level sit idle
// anyway.  There two lanes (t, t ^ 1; h = t & 1) hash one node: lane h holds
// the six state elements x'_m = x_{(m + 6h) mod 12}, m < 6.  The MDS is
// circulant, so in this rotated indexing row q of lane h (state row
// 6h + q) is sum_m CIRC[(m - q) mod 12] x'_m (+ DIAG[6h + q] x'_q): the same
// coefficients on both lanes, except a diagonal term that differs between
// rows q and 6 + q.  Each lane takes its partner's six words (x'_{6..11})
// with a DPP lane swap and accumulates its own six rows: about 58% of the
// single-lane instructions per lane.  Both lanes of a pair must be active.

// The partner lane's copy of v (lanes 2k and 2k+1 exchange).
__device__ __forceinline__ uint32_t lane_swap(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1 /* quad_perm [1,0,3,2] */, 0xF, 0xF, false);
}

// Row seeds (the next round's constants) and the MDS of the rotated state.
__device__ __forceinline__ void poseidon_mds_pair(uint64_t x[6], const uint64_t *k, bool h)
{
    uint32_t lo[12], hi[12];
#pragma unroll
    for (int m = 0; m < 6; m++) {
        lo[m] = gl::lo32(x[m]);
        hi[m] = gl::hi32(x[m]);
    }
#pragma unroll
    for (int m = 0; m < 6; m++) {
        lo[6 + m] = lane_swap(lo[m]);
        hi[6 + m] = lane_swap(hi[m]);
    }
#pragma unroll
    for (int q0 = 0; q0 < 6; q0 += 2) {
        uint64_t A[2], B[2];
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const uint64_t kk = h ? k[6 + q0 + d] : k[q0 + d];
            A[d] = gl::lo32(kk);
            B[d] = gl::hi32(kk);
        }
#pragma unroll
        for (int i = 0; i < 12; i++)
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const int q = q0 + d, m = (i + q) % 12;
                uint32_t c;
                if (i == 0 && ZK_POSEIDON_MDS_DIAG[q] != ZK_POSEIDON_MDS_DIAG[6 + q]) {
                    const uint32_t c0 = (uint32_t)(ZK_POSEIDON_MDS_CIRC[0] + ZK_POSEIDON_MDS_DIAG[q]);
                    const uint32_t c1 = (uint32_t)(ZK_POSEIDON_MDS_CIRC[0] + ZK_POSEIDON_MDS_DIAG[6 + q]);
                    c = h ? c1 : c0;
                } else {
                    c = gl::opaque((uint32_t)(ZK_POSEIDON_MDS_CIRC[i] + (i == 0 ? ZK_POSEIDON_MDS_DIAG[q] : 0)));
                }
                A[d] = gl::vopaque64((uint64_t)lo[m] * c + A[d]);
                B[d] = gl::vopaque64((uint64_t)hi[m] * c + B[d]);
            }
#pragma unroll
        for (int d = 0; d < 2; d++) x[q0 + d] = limb_combine(A[d], B[d]);
    }
}

// poseidon_permute on the lane pair's rotated halves.
__device__ __forceinline__ void poseidon_permute_pair(uint64_t x[6], bool h)
{
#pragma unroll
    for (int m = 0; m < 6; m++) x[m] = gl::add(x[m], h ? c_poseidon_rc[6 + m] : c_poseidon_rc[m]);
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int m = 0; m < 6; m += 2) gl::pow7_lazy2(x[m], x[m + 1]);
        poseidon_mds_pair(x, c_poseidon_rc + 12 * (r + 1), h);
    }
#pragma unroll 1
    for (int r = 4; r < 26; r++) {
        const uint64_t s0 = gl::pow7_lazy(x[0]);  // state element 0 lives on lane 0
        x[0] = h ? x[0] : s0;
        poseidon_mds_pair(x, c_poseidon_rc + 12 * (r + 1), h);
    }
#pragma unroll 1
    for (int r = 26; r < 30; r++) {
#pragma unroll
        for (int m = 0; m < 6; m += 2) gl::pow7_lazy2(x[m], x[m + 1]);
        poseidon_mds_pair(x, c_poseidon_rc + 12 * (r + 1), h);
    }
}

// hash_two<kPoseidon>(l, r) on a lane pair: state [l, r, 0]; the digest is
// valid on the even lane (h = 0).
__device__ __forceinline__ void poseidon_hash_two_pair(const uint64_t l[4], const uint64_t r[4], uint64_t out[4],
                                                       bool h)
{
    uint64_t x[6];
#pragma unroll
    for (int m = 0; m < 4; m++) x[m] = h ? (m < 2 ? r[2 + m] : 0) : l[m];
    x[4] = h ? 0 : r[0];
    x[5] = h ? 0 : r[1];
    poseidon_permute_pair(x, h);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = gl::canon(x[i]);
}

// -------------------------------------------------- Poseidon on a lane row
//
// The narrowest levels of a tile (a few nodes per work-group) are bound by
// one permutation's latency even on a lane pair.  Sixteen lanes of a wave
// (a row: lanes 16r .. 16r + 15) hash one node: lane k < 12 holds state
// element k, lanes 12..15 ride along on element k - 12 (never read by the
// others).  MDS row k is sum_m CIRC[m] x_{(k + m) mod 12} (+ DIAG[k] x_k):
// each lane fetches the eleven other elements in that rotated order with
// ds_bpermute (two 32-bit words each), so every coefficient but the first is
// the same immediate on all lanes.  All 16 lanes of a row must be active.

__device__ __forceinline__ uint32_t row_fetch(uint32_t byte_addr, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)byte_addr, (int)v);
}

// x <- row k of MDS * state + kc (kc: this lane's constant of the next
// round); src[m] = 4 * (the lane holding element (k + m) mod 12); c0 =
// CIRC[0] + DIAG[k].  Lazy in and out, as poseidon_mds.
__device__ __forceinline__ uint64_t poseidon_mds_row(uint64_t x, const uint32_t (&src)[12], uint64_t kc, uint32_t c0)
{
    uint32_t lo[12], hi[12];
    lo[0] = gl::lo32(x);
    hi[0] = gl::hi32(x);
#pragma unroll
    for (int m = 1; m < 12; m++) {
        lo[m] = row_fetch(src[m], lo[0]);
        hi[m] = row_fetch(src[m], hi[0]);
    }
    uint64_t A = gl::vopaque64((uint64_t)lo[0] * c0 + gl::lo32(kc));
    uint64_t B = gl::vopaque64((uint64_t)hi[0] * c0 + gl::hi32(kc));
#pragma unroll
    for (int m = 1; m < 12; m++) {
        const uint32_t c = gl::opaque((uint32_t)ZK_POSEIDON_MDS_CIRC[m]);
        A = gl::vopaque64((uint64_t)lo[m] * c + A);
        B = gl::vopaque64((uint64_t)hi[m] * c + B);
    }
    return limb_combine(A, B);
}

// poseidon_permute with the state spread over a row (x: this lane's element,
// k = lane mod 16).  Lazy in and out.
__device__ __forceinline__ uint64_t poseidon_permute_row(uint64_t x, uint32_t k)
{
    const uint32_t kk = k < 12 ? k : k - 12;
    const uint32_t base = (threadIdx.x & 63u) & ~15u;  // first lane of the row within the wave
    uint32_t src[12];
#pragma unroll
    for (int m = 0; m < 12; m++) {
        const uint32_t e = kk + (uint32_t)m;
        src[m] = 4u * (base + (e >= 12 ? e - 12 : e));
    }
    uint32_t c0 = (uint32_t)ZK_POSEIDON_MDS_CIRC[0];
#pragma unroll
    for (int i = 0; i < 12; i++)
        if (ZK_POSEIDON_MDS_DIAG[i] != 0 && kk == (uint32_t)i) c0 += (uint32_t)ZK_POSEIDON_MDS_DIAG[i];
    x = gl::add(x, c_poseidon_rc[kk]);
    // rounds 0-3 and 26-29 full, 4-25 partial (the S-box on element 0 only)
#pragma unroll 1
    for (int r = 0; r < 30; r++) {
        const uint64_t s = gl::pow7_lazy(x);
        x = (r < 4 || r >= 26 || kk == 0) ? s : x;
        x = poseidon_mds_row(x, src, c_poseidon_rc[12 * (r + 1) + kk], c0);
    }
    return x;
}

// hash_two<kPoseidon>(l, r) on a row: state [l, r, 0]; lane k < 4 returns
// element k of the digest (canonical).
__device__ __forceinline__ uint64_t poseidon_hash_two_row(const uint64_t l[4], const uint64_t r[4], uint32_t k)
{
    const uint64_t x = k < 4 ? l[k] : k < 8 ? r[k - 4] : 0;
    return gl::canon(poseidon_permute_row(x, k));
}

// --------------------------------------------------------------- Poseidon2

// External layer (lazy in and out) M_E = [[2M4, M4, M4], [M4, 2M4, M4], [M4, M4, 2M4]] with
// M4 = [[5,7,1,3],[4,6,1,1],[1,3,5,7],[1,1,4,6]] (poseidon2.cu:65-145: M4 on
// each 4-block, then each position plus the sum of that position over the
// blocks).  out_b = M4 * (s_b + u), u[c] = s[c] + s[4+c] + s[8+c]; the shared
// M4 * u seeds every block's limb accumulators.
__device__ __forceinline__ void p2_external(uint64_t s[12])
{
    constexpr uint32_t M4[4][4] = {{5, 7, 1, 3}, {4, 6, 1, 1}, {1, 3, 5, 7}, {1, 1, 4, 6}};
    uint32_t lo[12], hi[12];
#pragma unroll
    for (int i = 0; i < 12; i++) {
        lo[i] = gl::lo32(s[i]);
        hi[i] = gl::hi32(s[i]);
    }
    uint64_t VA[4], VB[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        uint64_t A = 0, B = 0;
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int b = 0; b < 12; b += 4) {
                A = gl::vopaque64((uint64_t)lo[b + c] * gl::opaque(M4[r][c]) + A);
                B = gl::vopaque64((uint64_t)hi[b + c] * gl::opaque(M4[r][c]) + B);
            }
        VA[r] = A;
        VB[r] = B;
    }
#pragma unroll
    for (int b = 0; b < 12; b += 4)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            uint64_t A = VA[r], B = VB[r];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                A = gl::vopaque64((uint64_t)lo[b + c] * gl::opaque(M4[r][c]) + A);
                B = gl::vopaque64((uint64_t)hi[b + c] * gl::opaque(M4[r][c]) + B);
            }
            s[b + r] = limb_combine(A, B);
        }
}

// Sum of the 12 state elements (each below 2^64), lazy: 32-bit carry chain
// into three words, one reduction.
__device__ __forceinline__ uint64_t sum12(const uint64_t s[12])
{
    uint32_t lo = 0, hi = 0, top = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) {
        uint32_t c1, c2, c3;
        lo = __builtin_addc(lo, gl::lo32(s[i]), 0u, &c1);
        hi = __builtin_addc(hi, gl::hi32(s[i]), c1, &c2);
        top = __builtin_addc(top, 0u, c2, &c3);
    }
    return gl::reduce96_lazy(gl::pack(lo, hi), top);
}

// Internal layer: s_i <- s_i * d_i + sum(s) (d = MATRIX_DIAG_12 = diag - 1),
// each as one 128-bit value reduced once (lazy in and out).  The product's
// middle carry (gl::mul128c) enters the reduction as a borrow: the sum added
// to the low word carries at most 1 into the high word, and the total stays
// below (2^64 - 1)^2 + 2^64, so the high word plus the carry fits 64 bits.
__device__ __forceinline__ void p2_internal(uint64_t s[12])
{
    const uint64_t sum = sum12(s);
#pragma unroll
    for (int i = 0; i < 12; i++) {
        uint64_t lo, hi, pc;
        gl::mul128c(s[i], ZK_POSEIDON2_DIAG[i], lo, hi, pc);
        uint32_t c1, c2, c3, c4;
        const uint32_t l0 = __builtin_addc(gl::lo32(lo), gl::lo32(sum), 0u, &c1);
        const uint32_t l1 = __builtin_addc(gl::hi32(lo), gl::hi32(sum), c1, &c2);
        const uint32_t h0 = __builtin_addc(gl::lo32(hi), 0u, c2, &c3);
        const uint32_t h1 = __builtin_addc(gl::hi32(hi), 0u, c3, &c4);
        s[i] = gl::reduce128c_lazy(gl::pack(l0, l1), gl::pack(h0, h1), pc);
    }
}

// Lazy in and out, like poseidon_permute.
__device__ __forceinline__ void poseidon2_permute(uint64_t s[12])
{
    p2_external(s);
    // rounds 0-3 and 26-29 external (all 12 S-boxes, in pairs), 4-25
    // internal (s0 only)
    auto external_round = [&](int r) {
        const uint64_t *rc = c_poseidon2_rc + 12 * r;
#pragma unroll
        for (int i = 0; i < 12; i += 2) {
            s[i] = gl::add(s[i], rc[i]);
            s[i + 1] = gl::add(s[i + 1], rc[i + 1]);
            gl::pow7_lazy2(s[i], s[i + 1]);
        }
        p2_external(s);
    };
#pragma unroll 1
    for (int r = 0; r < 4; r++) external_round(r);
#pragma unroll 1
    for (int r = 4; r < 26; r += 2) {  // two per iteration: no loop-carried copies
#pragma unroll
        for (int h = 0; h < 2; h++) {
            s[0] = gl::pow7_lazy(gl::add(s[0], c_poseidon2_rc[12 * (r + h)]));
            p2_internal(s);
        }
    }
#pragma unroll 1
    for (int r = 26; r < 30; r++) external_round(r);
}

// poseidon2_permute on a lane pair (as poseidon_permute_pair: tree levels,
// both lanes active, h = lane & 1).  External rounds keep the full state on
// both lanes: each lane applies the S-box to its six elements (6h .. 6h+5),
// the halves are exchanged, and both lanes run the external layer.  Internal
// rounds update only the lane's own six elements (s_i d_i + sum, the sum
// from both lanes' partial sums), and the halves are exchanged again before
// the next external round.
__device__ __forceinline__ uint64_t lane_swap64(uint64_t v)
{
    return gl::pack(lane_swap(gl::lo32(v)), lane_swap(gl::hi32(v)));
}

__device__ __forceinline__ void poseidon2_permute_pair(uint64_t s[12], bool h)
{
    p2_external(s);
    // own[m] = s[6h + m]; the partner's half back into the full state
    auto own_of = [&](uint64_t own[6]) {
#pragma unroll
        for (int m = 0; m < 6; m++) own[m] = h ? s[6 + m] : s[m];
    };
    auto merge = [&](const uint64_t own[6]) {
#pragma unroll
        for (int m = 0; m < 6; m++) {
            const uint64_t other = lane_swap64(own[m]);
            s[m] = h ? other : own[m];
            s[6 + m] = h ? own[m] : other;
        }
    };
    auto external_round = [&](int r) {
        const uint64_t *rc = c_poseidon2_rc + 12 * r;
        uint64_t own[6];
#pragma unroll
        for (int m = 0; m < 6; m++) own[m] = gl::add(h ? s[6 + m] : s[m], h ? rc[6 + m] : rc[m]);
#pragma unroll
        for (int m = 0; m < 6; m += 2) gl::pow7_lazy2(own[m], own[m + 1]);
        merge(own);
        p2_external(s);
    };
#pragma unroll 1
    for (int r = 0; r < 4; r++) external_round(r);
    uint64_t own[6], d[6];
    own_of(own);
#pragma unroll
    for (int m = 0; m < 6; m++) d[m] = h ? ZK_POSEIDON2_DIAG[6 + m] : ZK_POSEIDON2_DIAG[m];
#pragma unroll 1
    for (int r = 4; r < 26; r++) {
        const uint64_t t0 = gl::pow7_lazy(gl::add(own[0], c_poseidon2_rc[12 * r]));  // element 0: lane 0
        own[0] = h ? own[0] : t0;
        // sum of the lane's six (three words), plus the partner's
        uint32_t lo = 0, hi = 0, top = 0;
#pragma unroll
        for (int m = 0; m < 6; m++) {
            uint32_t c1, c2, c3;
            lo = __builtin_addc(lo, gl::lo32(own[m]), 0u, &c1);
            hi = __builtin_addc(hi, gl::hi32(own[m]), c1, &c2);
            top = __builtin_addc(top, 0u, c2, &c3);
        }
        {
            const uint32_t plo = lane_swap(lo), phi = lane_swap(hi), ptop = lane_swap(top);
            uint32_t c1, c2, c3;
            lo = __builtin_addc(lo, plo, 0u, &c1);
            hi = __builtin_addc(hi, phi, c1, &c2);
            top = __builtin_addc(top, ptop, c2, &c3);
        }
        const uint64_t sum = gl::reduce96_lazy(gl::pack(lo, hi), top);
#pragma unroll
        for (int m = 0; m < 6; m++) {
            uint64_t plo, phi, pc;
            gl::mul128c(own[m], d[m], plo, phi, pc);
            uint32_t c1, c2, c3, c4;
            const uint32_t l0 = __builtin_addc(gl::lo32(plo), gl::lo32(sum), 0u, &c1);
            const uint32_t l1 = __builtin_addc(gl::hi32(plo), gl::hi32(sum), c1, &c2);
            const uint32_t h0 = __builtin_addc(gl::lo32(phi), 0u, c2, &c3);
            const uint32_t h1 = __builtin_addc(gl::hi32(phi), 0u, c3, &c4);
            own[m] = gl::reduce128c_lazy(gl::pack(l0, l1), gl::pack(h0, h1), pc);
        }
    }
    merge(own);
#pragma unroll 1
    for (int r = 26; r < 30; r++) external_round(r);
}

// hash_two<kPoseidon2>(l, r) on a lane pair; both lanes end with the digest.
__device__ __forceinline__ void poseidon2_hash_two_pair(const uint64_t l[4], const uint64_t r[4], uint64_t out[4],
                                                        bool h)
{
    uint64_t s[12];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        s[i] = l[i];
        s[4 + i] = r[i];
        s[8 + i] = 0;
    }
    poseidon2_permute_pair(s, h);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = gl::canon(s[i]);
}

template <int HT>
__device__ __forceinline__ void permute(uint64_t s[12])
{
    if constexpr (HT == kPoseidon2) poseidon2_permute(s);
    else if constexpr (HT == kMonolith) monolith_permute(s);
    else if constexpr (HT == kPoseidonBN128) poseidon_bn128_permute(s);
    else poseidon_permute(s);
}

// hash_two: state = [l(4), r(4), 0(4)], permute, output state[0..4).
template <int HT>
__device__ __forceinline__ void hash_two(const uint64_t l[4], const uint64_t r[4], uint64_t out[4])
{
    if constexpr (HT == kKeccak) {
        keccak_hash_two(l, r, out);
        return;
    }
    uint64_t s[12];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        s[i] = l[i];
        s[4 + i] = r[i];
        s[8 + i] = 0;
    }
    permute<HT>(s);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = gl::canon(s[i]);
}

// hash_one over a leaf of n elements fetched through load(k):
//   n <= 4 : the elements themselves, zero-padded to 4;
//   n  > 4 : zero state; per chunk of 8, overwrite state[0..len) and permute.
template <int HT, class Load>
__device__ __forceinline__ void hash_one(uint32_t n, Load load, uint64_t out[4])
{
    if constexpr (HT == kKeccak) {
        keccak_hash_one(n, load, out);
        return;
    }
    if (n <= 4) {
#pragma unroll
        for (int i = 0; i < 4; i++) out[i] = (uint32_t)i < n ? load(i) : 0;
        return;
    }
    uint64_t s[12];
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = 0;
    for (uint32_t idx = 0; idx < n; idx += 8) {
#pragma unroll
        for (int i = 0; i < 8; i++)
            if (idx + i < n) s[i] = load(idx + i);
        permute<HT>(s);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = gl::canon(s[i]);
}

}  // namespace zkx
