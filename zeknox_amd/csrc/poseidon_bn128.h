// Poseidon over the BN254 scalar field (hash type 2) for one lane: t = 5,
// x^5 S-box, 8 full rounds and 60 partial rounds in the optimised form
// (full rounds mix with M, the last first-half round with P, partial rounds
// with the sparse rows S).  Reference: poseidon/poseidon_bn128.cu:116-204
// and poseidon/element_bn128.cu (Montgomery arithmetic, R = 2^256).
//
// Field elements are eight 32-bit limbs (least significant first) in
// Montgomery form; multiplication is CIOS on 32-bit words, each step one
// v_mad_u64_u32.  The 12 Goldilocks words of the sponge state map to
// state[1..4] as element i = w[3i] 2^128 + w[3i+1] 2^64 + w[3i+2]; the
// output words of state[0..4) come back the same way (the top 64 bits
// dropped) and each is reduced once mod p.  The constant tables are the
// reference's, in Montgomery form (gl_constants.h).
#pragma once
#include "gl64.h"
#include "gl_constants.h"

namespace zkx {

static __constant__ uint64_t c_bn128_c[100 * 4] = {ZK_BN128_C_LIST};
static __constant__ uint64_t c_bn128_s[540 * 4] = {ZK_BN128_S_LIST};
static __constant__ uint64_t c_bn128_m[25 * 4] = {ZK_BN128_M_LIST};
static __constant__ uint64_t c_bn128_p[25 * 4] = {ZK_BN128_P_LIST};

struct Fr {
    uint32_t w[8];
};

// r, the BN254 scalar field modulus, and -r^-1 mod 2^32
constexpr uint32_t kBnR[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                              0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
constexpr uint32_t kBnNinv = 0xefffffffu;
// R^2 mod r (R = 2^256), for the conversion into Montgomery form
constexpr uint32_t kBnR2[8] = {0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u,
                               0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u};

__device__ __forceinline__ Fr fr_load(const uint64_t *p)
{
    Fr a;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        a.w[2 * i] = gl::lo32(p[i]);
        a.w[2 * i + 1] = gl::hi32(p[i]);
    }
    return a;
}

// a - r if a >= r (a < 2r), as a borrow chain and a select
__device__ __forceinline__ Fr fr_reduce_once(const Fr &a, uint32_t carry)
{
    Fr d;
    uint32_t b = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.w[i] = __builtin_subc(a.w[i], kBnR[i], b, &b);
    // a >= r when the subtraction did not borrow, or the sum carried out
    const bool ge = carry || !b;
    Fr r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.w[i] = ge ? d.w[i] : a.w[i];
    return r;
}

__device__ __forceinline__ Fr fr_add(const Fr &a, const Fr &b)
{
    Fr s;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s.w[i] = __builtin_addc(a.w[i], b.w[i], c, &c);
    return fr_reduce_once(s, c);
}

// a * b * R^-1 mod r (CIOS, 32-bit words)
__device__ __noinline__ Fr fr_mul(Fr a, Fr b)
{
    uint32_t t[10];
#pragma unroll
    for (int i = 0; i < 10; i++) t[i] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint64_t v = (uint64_t)a.w[i] * b.w[j] + t[j] + (c >> 32);
            t[j] = (uint32_t)v;
            c = v;
        }
        uint64_t v = (uint64_t)t[8] + (c >> 32);
        t[8] = (uint32_t)v;
        t[9] = (uint32_t)(v >> 32);
        const uint32_t m = t[0] * kBnNinv;
        c = (uint64_t)m * kBnR[0] + t[0];
#pragma unroll
        for (int j = 1; j < 8; j++) {
            const uint64_t u = (uint64_t)m * kBnR[j] + t[j] + (c >> 32);
            t[j - 1] = (uint32_t)u;
            c = u;
        }
        v = (uint64_t)t[8] + (c >> 32);
        t[7] = (uint32_t)v;
        t[8] = t[9] + (uint32_t)(v >> 32);
    }
    Fr z;
#pragma unroll
    for (int i = 0; i < 8; i++) z.w[i] = t[i];
    return fr_reduce_once(z, t[8]);
}

__device__ __forceinline__ Fr fr_exp5(const Fr &x)
{
    const Fr x2 = fr_mul(x, x);
    const Fr x4 = fr_mul(x2, x2);
    return fr_mul(x4, x);
}

// state <- mat^T * state (new[i] = sum_j mat[j][i] * state[j])
__device__ __forceinline__ void fr_mix(Fr st[5], const uint64_t *mat)
{
    Fr ns[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        Fr acc = fr_mul(fr_load(mat + 4 * i), st[0]);
#pragma unroll
        for (int j = 1; j < 5; j++) acc = fr_add(acc, fr_mul(fr_load(mat + 4 * (5 * j + i)), st[j]));
        ns[i] = acc;
    }
#pragma unroll
    for (int i = 0; i < 5; i++) st[i] = ns[i];
}

__device__ __forceinline__ void fr_ark(Fr st[5], int off)
{
#pragma unroll
    for (int k = 0; k < 5; k++) st[k] = fr_add(st[k], fr_load(c_bn128_c + 4 * (off + k)));
}

__device__ __forceinline__ void fr_full_round(Fr st[5])
{
#pragma unroll
    for (int k = 0; k < 5; k++) st[k] = fr_exp5(st[k]);
}

// The 12-word Goldilocks sponge state, lazy in (any word below 2^64 is
// made canonical first) and canonical out.
__device__ __forceinline__ void poseidon_bn128_permute(uint64_t s[12])
{
    Fr r2;
#pragma unroll
    for (int i = 0; i < 8; i++) r2.w[i] = kBnR2[i];
    Fr st[5];
#pragma unroll
    for (int i = 0; i < 8; i++) st[0].w[i] = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint64_t w2 = gl::canon(s[3 * i]), w1 = gl::canon(s[3 * i + 1]), w0 = gl::canon(s[3 * i + 2]);
        Fr v;
        v.w[0] = gl::lo32(w0);
        v.w[1] = gl::hi32(w0);
        v.w[2] = gl::lo32(w1);
        v.w[3] = gl::hi32(w1);
        v.w[4] = gl::lo32(w2);
        v.w[5] = gl::hi32(w2);
        v.w[6] = 0;
        v.w[7] = 0;
        st[i + 1] = fr_mul(v, r2);
    }
    fr_ark(st, 0);
#pragma unroll 1
    for (int r = 0; r < 3; r++) {
        fr_full_round(st);
        fr_ark(st, (r + 1) * 5);
        fr_mix(st, c_bn128_m);
    }
    fr_full_round(st);
    fr_ark(st, 20);
    fr_mix(st, c_bn128_p);
#pragma unroll 1
    for (int r = 0; r < 60; r++) {
        st[0] = fr_add(fr_exp5(st[0]), fr_load(c_bn128_c + 4 * (25 + r)));
        const uint64_t *row = c_bn128_s + 4 * (9 * r);
        Fr n0 = fr_mul(fr_load(row), st[0]);
#pragma unroll
        for (int j = 1; j < 5; j++) n0 = fr_add(n0, fr_mul(fr_load(row + 4 * j), st[j]));
#pragma unroll
        for (int k = 1; k < 5; k++) st[k] = fr_add(st[k], fr_mul(st[0], fr_load(row + 4 * (4 + k))));
        st[0] = n0;
    }
#pragma unroll 1
    for (int r = 0; r < 3; r++) {
        fr_full_round(st);
        fr_ark(st, 85 + 5 * r);
        fr_mix(st, c_bn128_m);
    }
    fr_full_round(st);
    fr_mix(st, c_bn128_m);
    Fr one;
#pragma unroll
    for (int i = 0; i < 8; i++) one.w[i] = i == 0 ? 1u : 0u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const Fr v = fr_mul(st[i], one);
        s[3 * i] = gl::canon(gl::pack(v.w[4], v.w[5]));
        s[3 * i + 1] = gl::canon(gl::pack(v.w[2], v.w[3]));
        s[3 * i + 2] = gl::canon(gl::pack(v.w[0], v.w[1]));
    }
}

}  // namespace zkx
