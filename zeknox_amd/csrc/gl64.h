// Goldilocks field p = 2^64 - 2^32 + 1 on the CDNA VALU.
//
// Values are u64 held in two VGPRs.  Every function here takes and returns
// canonical values in [0, p), which is the representation the reference's
// digests and LDE outputs are compared in (ff/gl64_t.cuh:485-506 reduces to
// canonical; ff/goldilocks.hpp:139-177 keeps canonical values).
//
// Multiplication is built from 32x32->64 pieces (v_mad_u64_u32) and reduced
// with 2^64 = 2^32 - 1 (mod p) and 2^96 = -1 (mod p).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gl {

constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p

__device__ __forceinline__ uint64_t add(uint64_t a, uint64_t b)
{
    uint64_t s = a + b;
    // a + b >= 2^64 (carry) or p <= s < 2^64: subtract p, i.e. add EPS mod 2^64
    return (s < a || s >= P) ? s + EPS : s;
}

__device__ __forceinline__ uint64_t sub(uint64_t a, uint64_t b)
{
    uint64_t d = a - b;
    return (a < b) ? d - EPS : d;  // borrow: d + p = d - EPS mod 2^64
}

__device__ __forceinline__ uint64_t neg(uint64_t a) { return a ? P - a : 0; }

// Reduce hi * 2^64 + lo (any 128-bit value) to [0, p).
__device__ __forceinline__ uint64_t reduce128(uint64_t lo, uint64_t hi)
{
    uint32_t h0 = (uint32_t)hi, h1 = (uint32_t)(hi >> 32);
    // hi*2^64 = h0*2^64 + h1*2^96 = h0*EPS - h1
    uint64_t x = lo - h1;
    if (lo < h1) x -= EPS;
    uint64_t y = ((uint64_t)h0 << 32) - h0;  // h0 * EPS, exact
    uint64_t z = x + y;
    if (z < y) z += EPS;
    return z >= P ? z - P : z;
}

// Reduce hi * 2^64 + lo with hi < 2^32.
__device__ __forceinline__ uint64_t reduce96(uint64_t lo, uint32_t hi)
{
    uint64_t y = ((uint64_t)hi << 32) - hi;
    uint64_t z = lo + y;
    if (z < y) z += EPS;
    return z >= P ? z - P : z;
}

// 64x64 -> 128 product from four 32x32 multiply-adds.
__device__ __forceinline__ void mul128(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi)
{
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t t0 = (uint64_t)a0 * b0;
    uint64_t t1 = (uint64_t)a0 * b1 + (t0 >> 32);
    uint64_t t2 = (uint64_t)a1 * b0 + (uint32_t)t1;
    uint64_t t3 = (uint64_t)a1 * b1 + (t1 >> 32);
    hi = t3 + (t2 >> 32);
    lo = (t2 << 32) | (uint32_t)t0;
}

__device__ __forceinline__ uint64_t mul(uint64_t a, uint64_t b)
{
    uint64_t lo, hi;
    mul128(a, b, lo, hi);
    return reduce128(lo, hi);
}

__device__ __forceinline__ uint64_t sqr(uint64_t a) { return mul(a, a); }

__device__ __forceinline__ uint64_t pow7(uint64_t x)
{
    uint64_t x2 = mul(x, x);
    uint64_t x3 = mul(x2, x);
    uint64_t x4 = mul(x2, x2);
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
