// Monolith-64 permutation (Goldilocks, width 12) for one lane, state in
// VGPRs.  Reference: monolith/monolith.cu:104-227.
//
//   concrete(rc row 0), then 6 rounds of
//     bars     : s[0..4) through the byte S-box y = x ^ (rotl1(~x) & rotl2(x)
//                & rotl3(x)), out rotl1(y) (8-bit rotations, all 8 bytes at once);
//     bricks   : s[i] += s[i-1]^2 (i = 11..1, squares mod p, the sum as an
//                integer: bit 64 kept as a carry);
//     concrete : state * circulant matrix + rc row r over the 65-bit values,
//                reduced to canonical (the next bars act on the bit pattern).
//
// concrete runs in 32-bit limbs like the Poseidon MDS: A = sum m*lo(x),
// B = sum m*hi(x), carries sum m*c, one v_mad_u64_u32 per product.
#pragma once
#include "gl64.h"
#include "gl_constants.h"

namespace zkx {

static __constant__ uint64_t c_monolith_rc[7 * 12] = {ZK_MONOLITH_RC_LIST};

// The byte S-box on every byte of a 32-bit word.
__device__ __forceinline__ uint32_t monolith_bar32(uint32_t x)
{
    const uint32_t nx = ~x;
    const uint32_t l1 = ((nx & 0x80808080u) >> 7) | ((nx & 0x7F7F7F7Fu) << 1);
    const uint32_t l2 = ((x & 0xC0C0C0C0u) >> 6) | ((x & 0x3F3F3F3Fu) << 2);
    const uint32_t l3 = ((x & 0xE0E0E0E0u) >> 5) | ((x & 0x1F1F1F1Fu) << 3);
    const uint32_t y = x ^ (l1 & l2 & l3);
    return ((y & 0x80808080u) >> 7) | ((y & 0x7F7F7F7Fu) << 1);
}

// s <- M s + k with carries c (s[i] + c[i] * 2^64 is the input value).
__device__ __forceinline__ void monolith_concrete(uint64_t s[12], const uint32_t c[12], const uint64_t *k)
{
    uint32_t lo[12], hi[12];
#pragma unroll
    for (int i = 0; i < 12; i++) {
        lo[i] = gl::lo32(s[i]);
        hi[i] = gl::hi32(s[i]);
    }
#pragma unroll
    for (int r = 0; r < 12; r++) {
        const uint64_t kr = k[r];
        uint64_t A = gl::lo32(kr), B = gl::hi32(kr);
        uint32_t top = 0;
#pragma unroll
        for (int j = 0; j < 12; j++) {
            const uint32_t m = (uint32_t)ZK_MONOLITH_MAT[12 * r + j];
            A = gl::vopaque64((uint64_t)lo[j] * gl::opaque(m) + A);
            B = gl::vopaque64((uint64_t)hi[j] * gl::opaque(m) + B);
            top += c[j] * m;
        }
        // value = A + B * 2^32 + top * 2^64 (A, B < 2^41, top < 2^9)
        const uint64_t T = gl::add_small(B, gl::hi32(A));
        s[r] = gl::reduce96(gl::pack(gl::lo32(A), gl::lo32(T)), gl::hi32(T) + top);
    }
}

__device__ __forceinline__ void monolith_permute(uint64_t s[12])
{
    uint32_t c[12];
#pragma unroll
    for (int i = 0; i < 12; i++) c[i] = 0;
    monolith_concrete(s, c, c_monolith_rc);
#pragma unroll 1
    for (int r = 1; r <= 6; r++) {
#pragma unroll
        for (int i = 0; i < 4; i++) s[i] = gl::pack(monolith_bar32(gl::lo32(s[i])), monolith_bar32(gl::hi32(s[i])));
#pragma unroll
        for (int i = 11; i > 0; i--) {
            const uint64_t sq = gl::mul(s[i - 1], s[i - 1]);
            uint32_t c1, c2;
            const uint32_t tlo = __builtin_addc(gl::lo32(s[i]), gl::lo32(sq), 0u, &c1);
            const uint32_t thi = __builtin_addc(gl::hi32(s[i]), gl::hi32(sq), c1, &c2);
            s[i] = gl::pack(tlo, thi);
            c[i] = c2;
        }
        c[0] = 0;
        monolith_concrete(s, c, c_monolith_rc + 12 * r);
    }
}

}  // namespace zkx
