// Keccak-f[1600] and the Keccak-256 Merkle hasher (hash type 1) for one
// lane, the 25-lane state in VGPRs (two 32-bit halves per lane).
//
// Reference: keccak/keccak.cu:39-157 (KeccakHasher): Keccak-256 with the
// original padding (byte 0x01 after the message, 0x80 in the last byte of
// the 136-byte block) over the little-endian bytes of the input words; the
// digest keeps 25 bytes (word 3 masked to its low byte).  hash_two hashes the
// 50 bytes made of the first 25 bytes of each child digest.  Fewer than four
// input words are copied (zero-padded; the reference leaves the rest of its
// output buffer untouched, which equals zero padding for zeroed buffers).
//
// The permutation follows FIPS 202 3.2 (theta, rho+pi, chi, iota); rotation
// offsets and lane moves are compile-time constants, round constants are read
// through the uniform round counter from constant memory.  Rotations are two
// v_alignbit_b32 each; nothing runs on the 64-bit shifter.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace zkx {

// iota round constants (FIPS 202 3.2.5; the same numbers as
// keccak/keccak.cu:15-24)
static __constant__ uint64_t c_keccak_rc[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808aull, 0x8000000080008000ull,
    0x000000000000808bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000aull,
    0x000000008000808bull, 0x800000000000008bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800aull, 0x800000008000000aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};

struct Lane {
    uint32_t lo, hi;
};

template <int K>
__device__ __forceinline__ Lane rotl(Lane x)
{
    constexpr int k = K & 63;
    if constexpr (k == 0) {
        return x;
    } else if constexpr (k == 32) {
        return Lane{x.hi, x.lo};
    } else if constexpr (k < 32) {
        // (hi:lo) << k: hi' = hi << k | lo >> (32-k), lo' = lo << k | hi >> (32-k)
        return Lane{__builtin_amdgcn_alignbit(x.lo, x.hi, 32 - k), __builtin_amdgcn_alignbit(x.hi, x.lo, 32 - k)};
    } else {
        return rotl<k - 32>(Lane{x.hi, x.lo});
    }
}

__device__ __forceinline__ Lane lxor(Lane a, Lane b) { return Lane{a.lo ^ b.lo, a.hi ^ b.hi}; }

// rho offset of lane x + 5y (FIPS 202 table 2)
constexpr int kRho[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};

template <int I>
__device__ __forceinline__ void rho_pi_one(Lane (&A)[25], Lane (&B)[25])
{
    constexpr int x = I % 5, y = I / 5;
    B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl<kRho[I]>(A[I]);
}

template <int... I>
__device__ __forceinline__ void rho_pi(Lane (&A)[25], Lane (&B)[25], std::integer_sequence<int, I...>)
{
    (rho_pi_one<I>(A, B), ...);
}

__device__ __forceinline__ void keccak_f1600(Lane (&A)[25])
{
#pragma unroll 1
    for (int r = 0; r < 24; r++) {
        Lane C[5], D[5], B[25];
#pragma unroll
        for (int x = 0; x < 5; x++) C[x] = lxor(lxor(lxor(A[x], A[x + 5]), lxor(A[x + 10], A[x + 15])), A[x + 20]);
#pragma unroll
        for (int x = 0; x < 5; x++) D[x] = lxor(C[(x + 4) % 5], rotl<1>(C[(x + 1) % 5]));
#pragma unroll
        for (int i = 0; i < 25; i++) A[i] = lxor(A[i], D[i % 5]);
        rho_pi(A, B, std::make_integer_sequence<int, 25>{});
#pragma unroll
        for (int y = 0; y < 25; y += 5)
#pragma unroll
            for (int x = 0; x < 5; x++) {
                const Lane b0 = B[y + x], b1 = B[y + (x + 1) % 5], b2 = B[y + (x + 2) % 5];
                A[y + x] = Lane{b0.lo ^ (~b1.lo & b2.lo), b0.hi ^ (~b1.hi & b2.hi)};
            }
        const uint64_t rc = c_keccak_rc[r];
        A[0].lo ^= (uint32_t)rc;
        A[0].hi ^= (uint32_t)(rc >> 32);
    }
}

__device__ __forceinline__ Lane to_lane(uint64_t v) { return Lane{(uint32_t)v, (uint32_t)(v >> 32)}; }
__device__ __forceinline__ uint64_t from_lane(Lane l) { return ((uint64_t)l.hi << 32) | l.lo; }

__device__ __forceinline__ void keccak_out(const Lane (&A)[25], uint64_t out[4])
{
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = from_lane(A[i]);
    out[3] &= 0xFFull;
}

// hash_one: absorb n words (n >= 4) in blocks of 17 words (136 bytes).
template <class Load>
__device__ __forceinline__ void keccak_hash_one(uint32_t n, Load load, uint64_t out[4])
{
    if (n < 4) {
#pragma unroll
        for (int i = 0; i < 4; i++) out[i] = (uint32_t)i < n ? load(i) : 0;
        return;
    }
    Lane A[25];
#pragma unroll
    for (int i = 0; i < 25; i++) A[i] = Lane{0u, 0u};
    uint32_t idx = 0;
    for (; n - idx >= 17; idx += 17) {
#pragma unroll
        for (int i = 0; i < 17; i++) A[i] = lxor(A[i], to_lane(load(idx + i)));
        keccak_f1600(A);
    }
    // last block: r < 17 words, then 0x01 and the final 0x80
    const uint32_t r = n - idx;
#pragma unroll
    for (int i = 0; i < 17; i++) {
        if ((uint32_t)i < r) A[i] = lxor(A[i], to_lane(load(idx + i)));
        else if ((uint32_t)i == r) A[i].lo ^= 1u;
    }
    A[16].hi ^= 0x80000000u;
    keccak_f1600(A);
    keccak_out(A, out);
}

// hash_two: the 50 bytes l[0..25) || r[0..25), one block.
__device__ __forceinline__ void keccak_hash_two(const uint64_t l[4], const uint64_t r[4], uint64_t out[4])
{
    Lane A[25];
#pragma unroll
    for (int i = 0; i < 25; i++) A[i] = Lane{0u, 0u};
    A[0] = to_lane(l[0]);
    A[1] = to_lane(l[1]);
    A[2] = to_lane(l[2]);
    A[3] = to_lane((l[3] & 0xFFull) | (r[0] << 8));
    A[4] = to_lane((r[0] >> 56) | (r[1] << 8));
    A[5] = to_lane((r[1] >> 56) | (r[2] << 8));
    A[6] = to_lane((r[2] >> 56) | ((r[3] & 0xFFull) << 8) | (1ull << 16));
    A[16].hi = 0x80000000u;
    keccak_f1600(A);
    keccak_out(A, out);
}

}  // namespace zkx
