// Poseidon and Poseidon2 permutations (Goldilocks, width 12, rate 8,
// 8 full + 22 partial rounds, x^7 S-box) and the plonky2 sponge, for one
// lane: the 12-element state lives in VGPRs, round constants are
// compile-time immediates after full unrolling (wave-uniform by
// construction, no memory traffic).
//
// Poseidon      : poseidon/poseidon.cu:563-838 (naive rounds; same function
//                 as the fast-partial-round form of poseidon/poseidon.cpp).
// Poseidon2     : poseidon2/poseidon2.cu:65-209.
// Sponge        : poseidon/poseidon_permutation.hpp:83-139 /
//                 poseidon_permutation.cuh:83-134.
#pragma once
#include "gl64.h"
#include "gl_constants.h"

namespace zkx {

enum HashKind : int { kPoseidon = 0, kPoseidon2 = 3 };

// ---------------------------------------------------------------- Poseidon

// state <- MDS * state, MDS = circ(17,15,41,16,2,28,13,13,39,18,34,20) +
// diag(8,0,...).  Every coefficient is < 64 and a row's coefficients sum to
// at most 264, so each element is split into 22/22/20-bit limbs and the 12x12
// products are accumulated per limb in 32 bits with v_mad_u32_u24
// (limb < 2^22, sum < 2^22 * 264 < 2^31).
__device__ __forceinline__ void poseidon_mds(uint64_t s[12])
{
    uint32_t l0[12], l1[12], l2[12];
#pragma unroll
    for (int i = 0; i < 12; i++) {
        l0[i] = (uint32_t)s[i] & 0x3FFFFFu;
        l1[i] = (uint32_t)(s[i] >> 22) & 0x3FFFFFu;
        l2[i] = (uint32_t)(s[i] >> 44);
    }
#pragma unroll
    for (int r = 0; r < 12; r++) {
        uint32_t a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
        for (int i = 0; i < 12; i++) {
            const int k = (i + r) % 12;
            const uint32_t c = (uint32_t)ZK_POSEIDON_MDS_CIRC[i] + (i == 0 ? (uint32_t)ZK_POSEIDON_MDS_DIAG[r] : 0u);
            a0 = __umul24(l0[k], c) + a0;
            a1 = __umul24(l1[k], c) + a1;
            a2 = __umul24(l2[k], c) + a2;
        }
        // value = a0 + a1*2^22 + a2*2^44  (< 2^75)
        uint64_t lo = (uint64_t)a0 + ((uint64_t)a1 << 22);
        uint64_t t = (uint64_t)a2 << 44;
        uint64_t sum = lo + t;
        uint32_t hi = (a2 >> 20) + (sum < t ? 1u : 0u);
        s[r] = gl::reduce96(sum, hi);
    }
}

__device__ __forceinline__ void poseidon_permute(uint64_t s[12])
{
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 12; i++) s[i] = gl::pow7(gl::add(s[i], ZK_POSEIDON_RC[12 * r + i]));
        poseidon_mds(s);
    }
#pragma unroll
    for (int r = 4; r < 26; r++) {
#pragma unroll
        for (int i = 0; i < 12; i++) s[i] = gl::add(s[i], ZK_POSEIDON_RC[12 * r + i]);
        s[0] = gl::pow7(s[0]);
        poseidon_mds(s);
    }
#pragma unroll
    for (int r = 26; r < 30; r++) {
#pragma unroll
        for (int i = 0; i < 12; i++) s[i] = gl::pow7(gl::add(s[i], ZK_POSEIDON_RC[12 * r + i]));
        poseidon_mds(s);
    }
}

// --------------------------------------------------------------- Poseidon2

// One 4-block times M4 = [[5,7,1,3],[4,6,1,1],[1,3,5,7],[1,1,4,6]].
__device__ __forceinline__ void p2_m4(uint64_t &x0, uint64_t &x1, uint64_t &x2, uint64_t &x3)
{
    uint64_t t0 = gl::add(x0, x1);
    uint64_t t1 = gl::add(x2, x3);
    uint64_t t2 = gl::add(gl::add(x1, x1), t1);
    uint64_t t3 = gl::add(gl::add(x3, x3), t0);
    uint64_t t1x2 = gl::add(t1, t1);
    uint64_t t0x2 = gl::add(t0, t0);
    uint64_t t4 = gl::add(gl::add(t1x2, t1x2), t3);
    uint64_t t5 = gl::add(gl::add(t0x2, t0x2), t2);
    x0 = gl::add(t3, t5);
    x1 = t5;
    x2 = gl::add(t2, t4);
    x3 = t4;
}

__device__ __forceinline__ void p2_external(uint64_t s[12])
{
#pragma unroll
    for (int b = 0; b < 12; b += 4) p2_m4(s[b], s[b + 1], s[b + 2], s[b + 3]);
#pragma unroll
    for (int c = 0; c < 4; c++) {
        uint64_t sum = gl::add(gl::add(s[c], s[4 + c]), s[8 + c]);
        s[c] = gl::add(s[c], sum);
        s[4 + c] = gl::add(s[4 + c], sum);
        s[8 + c] = gl::add(s[8 + c], sum);
    }
}

__device__ __forceinline__ void poseidon2_permute(uint64_t s[12])
{
    p2_external(s);
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 12; i++) s[i] = gl::pow7(gl::add(s[i], ZK_POSEIDON2_RC[12 * r + i]));
        p2_external(s);
    }
#pragma unroll
    for (int r = 4; r < 26; r++) {
        s[0] = gl::pow7(gl::add(s[0], ZK_POSEIDON2_RC[12 * r]));
        uint64_t sum = s[0];
#pragma unroll
        for (int i = 1; i < 12; i++) sum = gl::add(sum, s[i]);
#pragma unroll
        for (int i = 0; i < 12; i++) {
            uint64_t lo, hi;
            gl::mul128(s[i], ZK_POSEIDON2_DIAG[i], lo, hi);
            // s_i * d_i + sum as one 128-bit value, reduced once
            uint64_t l2 = lo + sum;
            hi += (l2 < lo) ? 1 : 0;
            s[i] = gl::reduce128(l2, hi);
        }
    }
#pragma unroll
    for (int r = 26; r < 30; r++) {
#pragma unroll
        for (int i = 0; i < 12; i++) s[i] = gl::pow7(gl::add(s[i], ZK_POSEIDON2_RC[12 * r + i]));
        p2_external(s);
    }
}

template <int HT>
__device__ __forceinline__ void permute(uint64_t s[12])
{
    if constexpr (HT == kPoseidon2) poseidon2_permute(s);
    else poseidon_permute(s);
}

// hash_two: state = [l(4), r(4), 0(4)], permute, output state[0..4).
template <int HT>
__device__ __forceinline__ void hash_two(const uint64_t l[4], const uint64_t r[4], uint64_t out[4])
{
    uint64_t s[12];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        s[i] = l[i];
        s[4 + i] = r[i];
        s[8 + i] = 0;
    }
    permute<HT>(s);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = s[i];
}

// hash_one over a leaf of n elements fetched through load(k):
//   n <= 4 : the elements themselves, zero-padded to 4;
//   n  > 4 : zero state; per chunk of 8, overwrite state[0..len) and permute.
template <int HT, class Load>
__device__ __forceinline__ void hash_one(uint32_t n, Load load, uint64_t out[4])
{
    if (n <= 4) {
#pragma unroll
        for (int i = 0; i < 4; i++) out[i] = (uint32_t)i < n ? load(i) : 0;
        return;
    }
    uint64_t s[12];
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = 0;
    // one permutation call site keeps the unrolled permutation in the
    // instruction cache once per kernel
    for (uint32_t idx = 0; idx < n; idx += 8) {
#pragma unroll
        for (int i = 0; i < 8; i++)
            if (idx + i < n) s[i] = load(idx + i);
        permute<HT>(s);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = s[i];
}

}  // namespace zkx
