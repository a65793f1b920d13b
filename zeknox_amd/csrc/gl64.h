// Goldilocks field p = 2^64 - 2^32 + 1 on the CDNA VALU.
//
// Values are u64 held in two VGPRs.  Every function takes and returns
// canonical values in [0, p) -- the representation the reference's digests
// and LDE outputs are compared in (ff/gl64_t.cuh:485-506 reduces to
// canonical; ff/goldilocks.hpp:139-177 keeps canonical values) -- unless its
// comment says otherwise.
//
// Instruction selection.  On gfx950 the 32-bit integer multiply-adds
// (v_mad_u64_u32, v_mul_hi_u32) issue at full rate, while the 64-bit VALU
// forms (v_lshl_add_u64, v_cmp_*_u64) run at about a third of it.  All
// additions and comparisons here are therefore written as 32-bit carry
// chains (__builtin_addc / __builtin_subc -> v_add_co_u32 / v_addc_co_u32,
// carries in SGPR masks, results picked with v_cndmask_b32), and products
// are built from four v_mad_u64_u32.  Reduction uses 2^64 = 2^32 - 1 and
// 2^96 = -1 (mod p).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gl {

constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p

__device__ __forceinline__ uint32_t lo32(uint64_t x) { return (uint32_t)x; }
__device__ __forceinline__ uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }
__device__ __forceinline__ uint64_t pack(uint32_t lo, uint32_t hi) { return ((uint64_t)hi << 32) | lo; }

// A wave-uniform constant the compiler cannot see through: it stays an SGPR
// operand of v_mad_u64_u32 instead of being strength-reduced into 64-bit
// shift-adds (which issue at about a third of the multiply-add rate).
__device__ __forceinline__ uint32_t opaque(uint32_t c)
{
    uint32_t r;
    asm("; opaque %0" : "=s"(r) : "0"(c));
    return r;
}

// A per-lane 64-bit value the compiler cannot see through (kept in a VGPR
// pair).  Placed after the first multiply-add of an accumulation chain it
// pins the chain's initial addend to that multiply-add, instead of letting
// the adds be reassociated into a trailing 64-bit add (v_lshl_add_u64, a
// third of the rate).
__device__ __forceinline__ uint64_t vopaque64(uint64_t x)
{
    asm("; vopaque %0" : "+v"(x));
    return x;
}

// x + y for y < 2^32 as one full-rate v_mad_u64_u32 (y * 1 + x); valid when
// the sum fits in 64 bits.
__device__ __forceinline__ uint64_t add_small(uint64_t x, uint32_t y) { return (uint64_t)y * opaque(1u) + x; }

// (a + b) mod p for a < 2^64 and b < p, canonical when either
//   a < p, or b <= 2^64 - 2^33 + 1 (then a may be any u64).
// s = a + b; t = s + EPS = s - p (mod 2^64); a + b >= p iff either carries.
__device__ __forceinline__ uint64_t add(uint64_t a, uint64_t b)
{
    uint32_t c1, c2, c3, c4;
    const uint32_t slo = __builtin_addc(lo32(a), lo32(b), 0u, &c1);
    const uint32_t shi = __builtin_addc(hi32(a), hi32(b), c1, &c2);
    const uint32_t tlo = __builtin_addc(slo, 0xFFFFFFFFu, 0u, &c3);
    const uint32_t thi = __builtin_addc(shi, 0u, c3, &c4);
    return (c2 | c4) ? pack(tlo, thi) : pack(slo, shi);
}

// (a - b) mod p for a, b < p: borrow -> add p, i.e. subtract EPS.
__device__ __forceinline__ uint64_t sub(uint64_t a, uint64_t b)
{
    uint32_t b1, b2, b3, b4;
    const uint32_t dlo = __builtin_subc(lo32(a), lo32(b), 0u, &b1);
    const uint32_t dhi = __builtin_subc(hi32(a), hi32(b), b1, &b2);
    const uint32_t tlo = __builtin_subc(dlo, 0xFFFFFFFFu, 0u, &b3);
    const uint32_t thi = __builtin_subc(dhi, 0u, b3, &b4);
    return b2 ? pack(tlo, thi) : pack(dlo, dhi);
}

__device__ __forceinline__ uint64_t neg(uint64_t a) { return sub(0, a); }

// h * EPS = h * 2^32 - h for h < 2^32 (exact, <= 2^64 - 2^33 + 1).
__device__ __forceinline__ uint64_t mul_eps(uint32_t h)
{
    uint32_t b1, b2;
    const uint32_t ylo = __builtin_subc(0u, h, 0u, &b1);
    const uint32_t yhi = __builtin_subc(h, 0u, b1, &b2);
    return pack(ylo, yhi);
}

// Canonical residue of hi * 2^64 + lo (any 128-bit value):
//   hi * 2^64 = h0 * 2^64 + h1 * 2^96 = h0 * EPS - h1  (mod p).
__device__ __forceinline__ uint64_t reduce128(uint64_t lo, uint64_t hi)
{
    const uint32_t h0 = lo32(hi), h1 = hi32(hi);
    uint32_t b1, b2, b3, b4;
    const uint32_t xlo = __builtin_subc(lo32(lo), h1, 0u, &b1);
    const uint32_t xhi = __builtin_subc(hi32(lo), 0u, b1, &b2);
    // borrow: x = lo - h1 + 2^64, take off 2^64 = EPS (no underflow: x >= 2^64 - 2^32 + 1)
    const uint32_t elo = __builtin_subc(xlo, 0xFFFFFFFFu, 0u, &b3);
    const uint32_t ehi = __builtin_subc(xhi, 0u, b3, &b4);
    const uint64_t x = b2 ? pack(elo, ehi) : pack(xlo, xhi);
    return add(x, mul_eps(h0));
}

// Canonical residue of top * 2^64 + lo with top < 2^32.
__device__ __forceinline__ uint64_t reduce96(uint64_t lo, uint32_t top) { return add(lo, mul_eps(top)); }

// Lazy residue (below 2^64, congruent mod p) of top * 2^64 + lo, top < 2^32:
// y = lo + top * EPS (< 2^65, carry c), then y + c * EPS = y + c * 2^32 - c,
// which cannot carry again (c leaves y <= 2^64 - 2^33).
__device__ __forceinline__ uint64_t reduce96_lazy(uint64_t lo, uint32_t top)
{
    const uint64_t m = (uint64_t)top * opaque(0xFFFFFFFFu);
    uint32_t c1, c, k, k2, k3;
    const uint32_t ylo = __builtin_addc(lo32(lo), lo32(m), 0u, &c1);
    const uint32_t yhi = __builtin_addc(hi32(lo), hi32(m), c1, &c);
    const uint32_t rlo = __builtin_subc(ylo, 0u, c, &k);
    const uint32_t rhi = __builtin_subc(__builtin_addc(yhi, 0u, c, &k2), 0u, k, &k3);
    return pack(rlo, rhi);
}

// 64x64 -> 128 product from five 32x32+64 multiply-adds (any a, b < 2^64).
// The last one adds hi(t2) to t3 (t3 + hi(t2) <= 2^64 - 1), which keeps the
// high word off the slow 64-bit adder.
__device__ __forceinline__ void mul128(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi)
{
    const uint32_t a0 = lo32(a), a1 = hi32(a), b0 = lo32(b), b1 = hi32(b);
    const uint64_t t0 = (uint64_t)a0 * b0;
    const uint64_t t1 = (uint64_t)a1 * b0 + hi32(t0);
    const uint64_t t2 = (uint64_t)a0 * b1 + lo32(t1);
    const uint64_t t3 = (uint64_t)a1 * b1 + hi32(t1);  // <= 2^64 - 2^32
    hi = add_small(t3, hi32(t2));
    lo = pack(lo32(t0), lo32(t2));
}

// A residue of hi * 2^64 + lo (any 128-bit value) below 2^64 but not
// necessarily below p ("lazy"):  hi * 2^64 = h0 * EPS - h1 (mod p).
//   y = lo + h0 * EPS       (< 2^65; carry c)
//   z = y - h1 (mod 2^64)   (borrow b)
//   r = z + (c - b) * EPS   -- each case lands in [0, 2^64): c without b
//       leaves z <= 2^64 - 2^33, b without c leaves z >= 2^64 - 2^32 + 1,
//       so the wrapped 64-bit sum is the exact value.
__device__ __forceinline__ uint64_t reduce128_lazy(uint64_t lo, uint64_t hi)
{
    const uint32_t h0 = lo32(hi), h1 = hi32(hi);
    const uint64_t m = (uint64_t)h0 * opaque(0xFFFFFFFFu);
    uint32_t c1, c, b1, b;
    const uint32_t ylo = __builtin_addc(lo32(lo), lo32(m), 0u, &c1);
    const uint32_t yhi = __builtin_addc(hi32(lo), hi32(m), c1, &c);
    const uint32_t zlo = __builtin_subc(ylo, h1, 0u, &b1);
    const uint32_t zhi = __builtin_subc(yhi, 0u, b1, &b);
    // (c - b) * EPS as a 64-bit two's-complement addend: +EPS, -EPS or 0
    const uint32_t dlo = c > b ? 0xFFFFFFFFu : (b > c ? 1u : 0u);
    const uint32_t dhi = b > c ? 0xFFFFFFFFu : 0u;
    uint32_t k1, k2;
    const uint32_t rlo = __builtin_addc(zlo, dlo, 0u, &k1);
    const uint32_t rhi = __builtin_addc(zhi, dhi, k1, &k2);
    return pack(rlo, rhi);
}

// x mod p for any x < 2^64 (x >= p iff x + EPS carries).
__device__ __forceinline__ uint64_t canon(uint64_t x)
{
    uint32_t c1, c2;
    const uint32_t tlo = __builtin_addc(lo32(x), 0xFFFFFFFFu, 0u, &c1);
    const uint32_t thi = __builtin_addc(hi32(x), 0u, c1, &c2);
    return c2 ? pack(tlo, thi) : x;
}

// a * b, lazy (below 2^64, congruent mod p) for any a, b < 2^64.
__device__ __forceinline__ uint64_t mul_lazy(uint64_t a, uint64_t b)
{
    uint64_t lo, hi;
    mul128(a, b, lo, hi);
    return reduce128_lazy(lo, hi);
}

__device__ __forceinline__ uint64_t mul(uint64_t a, uint64_t b)
{
    uint64_t lo, hi;
    mul128(a, b, lo, hi);
    return reduce128(lo, hi);
}

__device__ __forceinline__ uint64_t sqr(uint64_t a) { return mul(a, a); }

// x^7 with the addition chain x2 = x^2, x3 = x*x2, x4 = x2^2, x7 = x3*x4;
// lazy in and out (any x < 2^64; result below 2^64, congruent mod p).
__device__ __forceinline__ uint64_t pow7_lazy(uint64_t x)
{
    const uint64_t x2 = mul_lazy(x, x);
    const uint64_t x3 = mul_lazy(x2, x);
    const uint64_t x4 = mul_lazy(x2, x2);
    return mul_lazy(x3, x4);
}

// x^7, canonical.
__device__ __forceinline__ uint64_t pow7(uint64_t x)
{
    const uint64_t x2 = mul(x, x);
    const uint64_t x3 = mul(x2, x);
    const uint64_t x4 = mul(x2, x2);
    return mul(x3, x4);
}

__device__ __forceinline__ uint64_t pow(uint64_t b, uint64_t e)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = mul(r, b);
        b = mul(b, b);
        e >>= 1;
    }
    return r;
}

}  // namespace gl
