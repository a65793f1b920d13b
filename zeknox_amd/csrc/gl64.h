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

// ---- carry-out forms
//
// v_mad_u64_u32 writes bit 64 of S0 * S1 + S2 to an SGPR lane mask; the
// compiler does not expose that carry, so the lazy reductions below reach it
// through inline asm.  gfx950 needs two wait states between a VALU write of
// a lane mask and a VALU read of it (the compiler inserts the same for its
// own carry chains): every mask is created and consumed inside one asm
// block, behind an s_nop 1 -- except the product carry of mul128c, which is
// read by the second asm block of reduce128c_lazy(2): that block depends on
// the first one, whose inputs depend on the multiply-add writing the mask,
// so at least the first block's three instructions lie between them.  The
// inline constant -1 is 0xFFFFFFFF = EPS as the 32-bit multiplicand.

// Lazy residue (below 2^64, congruent mod p) of top * 2^64 + lo, top < 2^32:
// u = lo + top * EPS (mod 2^64, carry c: one v_mad_u64_u32), then
// u + c * EPS, which cannot carry again (c leaves u <= 2^64 - 2^33).
__device__ __forceinline__ uint64_t reduce96_lazy(uint64_t lo, uint32_t top)
{
    uint64_t u, r, s0, s1;
    uint32_t cv;
    asm("v_mad_u64_u32 %[u], %[s0], %[top], -1, %[lo]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[cv], 0, 1, %[s0]\n\t"
        "v_mad_u64_u32 %[r], %[s1], %[cv], -1, %[u]"
        : [u] "=&v"(u), [r] "=&v"(r), [s0] "=&s"(s0), [s1] "=&s"(s1), [cv] "=&v"(cv)
        : [top] "v"(top), [lo] "v"(lo));
    return r;
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

// 64x64 -> 128 product from four multiply-adds, the carry of the middle sum
// left in a lane mask: a * b = lo + (hi + c * 2^32) * 2^64, c in {0, 1}
// (any a, b < 2^64).  u = a0 b0; v = a0 b1 + hi(u) (< 2^64); w = a1 b0 + v,
// whose bit 64 is c; hi = a1 b1 + hi(w) <= 2^64 - 2^32 (no carry).  Two
// instructions fewer than mul128 (one multiply-add and one register-pair
// copy): the carry goes into the reduction as a borrow (reduce128c_lazy).
__device__ __forceinline__ void mul128c(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi, uint64_t &c)
{
    const uint32_t a0 = lo32(a), a1 = hi32(a), b0 = lo32(b), b1 = hi32(b);
    const uint64_t u = (uint64_t)a0 * b0;
    const uint64_t v = (uint64_t)a0 * b1 + hi32(u);
    uint64_t w;
    asm("v_mad_u64_u32 %[w], %[c], %[x], %[y], %[v]"
        : [w] "=&v"(w), [c] "=&s"(c)
        : [x] "v"(a1), [y] "v"(b0), [v] "v"(v));
    hi = (uint64_t)a1 * b1 + hi32(w);
    lo = pack(lo32(u), lo32(w));
}

// A residue of hi * 2^64 + lo (any 128-bit value) below 2^64 but not
// necessarily below p ("lazy"):  hi * 2^64 = h0 * EPS - h1 (mod p).
//   u = lo + h0 * EPS (mod 2^64; carry c)      one v_mad_u64_u32
//   z = u - h1 (mod 2^64; borrow b)
//   r = z + (c - b) * EPS   -- each case lands in [0, 2^64): c without b
//       leaves z <= 2^64 - 2^33, b without c leaves z >= 2^64 - 2^32 + 1,
//       so the wrapped 64-bit sum is the exact value.  With cb = c - b in
//       {-1, 0, 1}: z + cb * EPS = (z - cb) + cb * 2^32, i.e. one signed
//       v_mad_i64_i32 (cb * -1 + z) and a 32-bit add to the high word.
__device__ __forceinline__ uint64_t reduce128_lazy(uint64_t lo, uint64_t hi)
{
    const uint32_t h0 = lo32(hi), h1 = hi32(hi);
    uint64_t u, sc;
    uint32_t cv;
    asm("v_mad_u64_u32 %[u], %[sc], %[h0], -1, %[lo]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[cv], 0, 1, %[sc]"
        : [u] "=&v"(u), [sc] "=&s"(sc), [cv] "=&v"(cv)
        : [h0] "v"(h0), [lo] "v"(lo));
    uint32_t zlo, zhi, cb;
    uint64_t b1, b, bj;
    asm("v_sub_co_u32_e64 %[zlo], %[b1], %[ulo], %[h1]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %[zhi], %[b], %[uhi], 0, %[b1]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %[cb], %[bj], %[cv], 0, %[b]"
        : [zlo] "=&v"(zlo), [zhi] "=&v"(zhi), [cb] "=&v"(cb), [b1] "=&s"(b1), [b] "=&s"(b), [bj] "=&s"(bj)
        : [ulo] "v"(lo32(u)), [uhi] "v"(hi32(u)), [h1] "v"(h1), [cv] "v"(cv));
    uint64_t t, tj;
    asm("v_mad_i64_i32 %[t], %[tj], %[cb], -1, %[z]"
        : [t] "=&v"(t), [tj] "=&s"(tj)
        : [cb] "v"(cb), [z] "v"(pack(zlo, zhi)));
    // The high-word add stays in asm: written in C, the compiler folds it
    // into a following carry chain (canon's hi + 0 + carry becomes
    // hi(t) + cb + carry), whose carry-out then counts the wrap of this add.
    uint32_t rhi;
    asm("v_add_u32 %[r], %[a], %[b]" : [r] "=v"(rhi) : [a] "v"(hi32(t)), [b] "v"(cb));
    return pack(lo32(t), rhi);
}

// reduce128_lazy of lo + (hi + cin * 2^32) * 2^64 (cin a lane mask, the
// high word plus cin below 2^64): cin is the borrow into u - h1.
__device__ __forceinline__ uint64_t reduce128c_lazy(uint64_t lo, uint64_t hi, uint64_t cin)
{
    const uint32_t h0 = lo32(hi), h1 = hi32(hi);
    uint64_t u, sc;
    uint32_t cv;
    asm("v_mad_u64_u32 %[u], %[sc], %[h0], -1, %[lo]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[cv], 0, 1, %[sc]"
        : [u] "=&v"(u), [sc] "=&s"(sc), [cv] "=&v"(cv)
        : [h0] "v"(h0), [lo] "v"(lo));
    uint32_t zlo, zhi, cb;
    uint64_t b1, b, bj;
    asm("v_subb_co_u32_e64 %[zlo], %[b1], %[ulo], %[h1], %[cin]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %[zhi], %[b], %[uhi], 0, %[b1]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %[cb], %[bj], %[cv], 0, %[b]"
        : [zlo] "=&v"(zlo), [zhi] "=&v"(zhi), [cb] "=&v"(cb), [b1] "=&s"(b1), [b] "=&s"(b), [bj] "=&s"(bj)
        : [ulo] "v"(lo32(u)), [uhi] "v"(hi32(u)), [h1] "v"(h1), [cv] "v"(cv), [cin] "s"(cin));
    uint64_t t, tj;
    asm("v_mad_i64_i32 %[t], %[tj], %[cb], -1, %[z]"
        : [t] "=&v"(t), [tj] "=&s"(tj)
        : [cb] "v"(cb), [z] "v"(pack(zlo, zhi)));
    // The high-word add stays in asm: written in C, the compiler folds it
    // into a following carry chain (canon's hi + 0 + carry becomes
    // hi(t) + cb + carry), whose carry-out then counts the wrap of this add.
    uint32_t rhi;
    asm("v_add_u32 %[r], %[a], %[b]" : [r] "=v"(rhi) : [a] "v"(hi32(t)), [b] "v"(cb));
    return pack(lo32(t), rhi);
}

// Two independent reduce128_lazy at once: the same instructions with the two
// reductions interleaved, so every lane-mask read already sits two
// instructions after its write (one s_nop 0 per pair of steps instead of an
// s_nop 1 per step).
__device__ __forceinline__ void reduce128_lazy2(uint64_t lo0, uint64_t hi0, uint64_t lo1, uint64_t hi1, uint64_t &r0,
                                                uint64_t &r1)
{
    uint64_t u0, u1, sc0, sc1;
    uint32_t cv0, cv1;
    asm("v_mad_u64_u32 %[u0], %[sc0], %[a0], -1, %[l0]\n\t"
        "v_mad_u64_u32 %[u1], %[sc1], %[a1], -1, %[l1]\n\t"
        "s_nop 0\n\t"
        "v_cndmask_b32_e64 %[cv0], 0, 1, %[sc0]\n\t"
        "v_cndmask_b32_e64 %[cv1], 0, 1, %[sc1]"
        : [u0] "=&v"(u0), [u1] "=&v"(u1), [sc0] "=&s"(sc0), [sc1] "=&s"(sc1), [cv0] "=&v"(cv0), [cv1] "=&v"(cv1)
        : [a0] "v"(lo32(hi0)), [l0] "v"(lo0), [a1] "v"(lo32(hi1)), [l1] "v"(lo1));
    uint32_t zl0, zh0, cb0, zl1, zh1, cb1;
    uint64_t b10, b0, j0, b11, b1, j1;
    asm("v_sub_co_u32_e64 %[zl0], %[b10], %[ul0], %[h10]\n\t"
        "v_sub_co_u32_e64 %[zl1], %[b11], %[ul1], %[h11]\n\t"
        "s_nop 0\n\t"
        "v_subb_co_u32_e64 %[zh0], %[b0], %[uh0], 0, %[b10]\n\t"
        "v_subb_co_u32_e64 %[zh1], %[b1], %[uh1], 0, %[b11]\n\t"
        "s_nop 0\n\t"
        "v_subb_co_u32_e64 %[cb0], %[j0], %[cv0], 0, %[b0]\n\t"
        "v_subb_co_u32_e64 %[cb1], %[j1], %[cv1], 0, %[b1]"
        : [zl0] "=&v"(zl0), [zh0] "=&v"(zh0), [cb0] "=&v"(cb0), [zl1] "=&v"(zl1), [zh1] "=&v"(zh1),
          [cb1] "=&v"(cb1), [b10] "=&s"(b10), [b0] "=&s"(b0), [j0] "=&s"(j0), [b11] "=&s"(b11), [b1] "=&s"(b1),
          [j1] "=&s"(j1)
        : [ul0] "v"(lo32(u0)), [uh0] "v"(hi32(u0)), [h10] "v"(hi32(hi0)), [cv0] "v"(cv0), [ul1] "v"(lo32(u1)),
          [uh1] "v"(hi32(u1)), [h11] "v"(hi32(hi1)), [cv1] "v"(cv1));
    uint64_t t0, t1, tj0, tj1;
    asm("v_mad_i64_i32 %[t0], %[tj0], %[cb0], -1, %[z0]\n\t"
        "v_mad_i64_i32 %[t1], %[tj1], %[cb1], -1, %[z1]"
        : [t0] "=&v"(t0), [t1] "=&v"(t1), [tj0] "=&s"(tj0), [tj1] "=&s"(tj1)
        : [cb0] "v"(cb0), [z0] "v"(pack(zl0, zh0)), [cb1] "v"(cb1), [z1] "v"(pack(zl1, zh1)));
    uint32_t rh0, rh1;  // high-word adds kept in asm (see reduce128_lazy)
    asm("v_add_u32 %[r], %[a], %[b]" : [r] "=v"(rh0) : [a] "v"(hi32(t0)), [b] "v"(cb0));
    asm("v_add_u32 %[r], %[a], %[b]" : [r] "=v"(rh1) : [a] "v"(hi32(t1)), [b] "v"(cb1));
    r0 = pack(lo32(t0), rh0);
    r1 = pack(lo32(t1), rh1);
}

// reduce128_lazy2 with the borrows cin0, cin1 (reduce128c_lazy).
__device__ __forceinline__ void reduce128c_lazy2(uint64_t lo0, uint64_t hi0, uint64_t cin0, uint64_t lo1, uint64_t hi1,
                                                 uint64_t cin1, uint64_t &r0, uint64_t &r1)
{
    uint64_t u0, u1, sc0, sc1;
    uint32_t cv0, cv1;
    asm("v_mad_u64_u32 %[u0], %[sc0], %[a0], -1, %[l0]\n\t"
        "v_mad_u64_u32 %[u1], %[sc1], %[a1], -1, %[l1]\n\t"
        "s_nop 0\n\t"
        "v_cndmask_b32_e64 %[cv0], 0, 1, %[sc0]\n\t"
        "v_cndmask_b32_e64 %[cv1], 0, 1, %[sc1]"
        : [u0] "=&v"(u0), [u1] "=&v"(u1), [sc0] "=&s"(sc0), [sc1] "=&s"(sc1), [cv0] "=&v"(cv0), [cv1] "=&v"(cv1)
        : [a0] "v"(lo32(hi0)), [l0] "v"(lo0), [a1] "v"(lo32(hi1)), [l1] "v"(lo1));
    uint32_t zl0, zh0, cb0, zl1, zh1, cb1;
    uint64_t b10, b0, j0, b11, b1, j1;
    asm("v_subb_co_u32_e64 %[zl0], %[b10], %[ul0], %[h10], %[ci0]\n\t"
        "v_subb_co_u32_e64 %[zl1], %[b11], %[ul1], %[h11], %[ci1]\n\t"
        "s_nop 0\n\t"
        "v_subb_co_u32_e64 %[zh0], %[b0], %[uh0], 0, %[b10]\n\t"
        "v_subb_co_u32_e64 %[zh1], %[b1], %[uh1], 0, %[b11]\n\t"
        "s_nop 0\n\t"
        "v_subb_co_u32_e64 %[cb0], %[j0], %[cv0], 0, %[b0]\n\t"
        "v_subb_co_u32_e64 %[cb1], %[j1], %[cv1], 0, %[b1]"
        : [zl0] "=&v"(zl0), [zh0] "=&v"(zh0), [cb0] "=&v"(cb0), [zl1] "=&v"(zl1), [zh1] "=&v"(zh1),
          [cb1] "=&v"(cb1), [b10] "=&s"(b10), [b0] "=&s"(b0), [j0] "=&s"(j0), [b11] "=&s"(b11), [b1] "=&s"(b1),
          [j1] "=&s"(j1)
        : [ul0] "v"(lo32(u0)), [uh0] "v"(hi32(u0)), [h10] "v"(hi32(hi0)), [cv0] "v"(cv0), [ul1] "v"(lo32(u1)),
          [uh1] "v"(hi32(u1)), [h11] "v"(hi32(hi1)), [cv1] "v"(cv1), [ci0] "s"(cin0),
          [ci1] "s"(cin1));
    uint64_t t0, t1, tj0, tj1;
    asm("v_mad_i64_i32 %[t0], %[tj0], %[cb0], -1, %[z0]\n\t"
        "v_mad_i64_i32 %[t1], %[tj1], %[cb1], -1, %[z1]"
        : [t0] "=&v"(t0), [t1] "=&v"(t1), [tj0] "=&s"(tj0), [tj1] "=&s"(tj1)
        : [cb0] "v"(cb0), [z0] "v"(pack(zl0, zh0)), [cb1] "v"(cb1), [z1] "v"(pack(zl1, zh1)));
    uint32_t rh0, rh1;  // high-word adds kept in asm (see reduce128_lazy)
    asm("v_add_u32 %[r], %[a], %[b]" : [r] "=v"(rh0) : [a] "v"(hi32(t0)), [b] "v"(cb0));
    asm("v_add_u32 %[r], %[a], %[b]" : [r] "=v"(rh1) : [a] "v"(hi32(t1)), [b] "v"(cb1));
    r0 = pack(lo32(t0), rh0);
    r1 = pack(lo32(t1), rh1);
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
    uint64_t lo, hi, c;
    mul128c(a, b, lo, hi, c);
    return reduce128c_lazy(lo, hi, c);
}

// mul_lazy through mul128 and reduce128_lazy: two instructions more, for
// kernels where the carry form's schedule spills (its products' operands
// and partial sums live longer).
__device__ __forceinline__ uint64_t mul_lazy_nc(uint64_t a, uint64_t b)
{
    uint64_t lo, hi;
    mul128(a, b, lo, hi);
    return reduce128_lazy(lo, hi);
}

// Two independent lazy products (their multiply-add chains and reductions
// interleave).
__device__ __forceinline__ void mul_lazy2(uint64_t a0, uint64_t b0, uint64_t a1, uint64_t b1, uint64_t &r0,
                                          uint64_t &r1)
{
    uint64_t lo0, hi0, c0, lo1, hi1, c1;
    mul128c(a0, b0, lo0, hi0, c0);
    mul128c(a1, b1, lo1, hi1, c1);
    reduce128c_lazy2(lo0, hi0, c0, lo1, hi1, c1, r0, r1);
}

// a * b, canonical (the lazy product's 7-instruction reduction plus the
// 4-instruction canonical fold is shorter than reduce128's chains).
__device__ __forceinline__ uint64_t mul(uint64_t a, uint64_t b) { return canon(mul_lazy(a, b)); }
__device__ __forceinline__ uint64_t mul_nc(uint64_t a, uint64_t b) { return canon(mul_lazy_nc(a, b)); }

__device__ __forceinline__ uint64_t sqr(uint64_t a) { return mul(a, a); }

// ---- multiplication by a power of two known at compile time
//
// 2 has order 192 mod p (2^96 = -1), and the 2^k-th roots of unity of the
// NTT for k <= 4 are powers of two (OMEGA[1..4] = 2^96, 2^48, 2^120, 2^156),
// so a butterfly whose twiddle is such a root multiplies by shifting.

constexpr uint64_t cmul(uint64_t a, uint64_t b) { return (uint64_t)((unsigned __int128)a * b % P); }
constexpr uint64_t cpow(uint64_t b, uint64_t e)
{
    uint64_t r = 1;
    for (; e; e >>= 1, b = cmul(b, b))
        if (e & 1) r = cmul(r, b);
    return r;
}
// OMEGA[k] = OMEGA[32]^(2^(32-k)) (gl64_params.hpp:48 gives OMEGA[32])
constexpr uint64_t kOmega32 = 1753635133440165772ull;
constexpr uint64_t comega(int k) { return cpow(kOmega32, 1ull << (32 - k)); }
// e with 2^e = w (mod p), -1 if w is not a power of two
constexpr int clog2(uint64_t w)
{
    for (int e = 0; e < 192; e++)
        if (cpow(2, e) == w) return e;
    return -1;
}

// x * 2^T, lazy, for 0 <= T < 32 and any x < 2^64: the 96-bit value
// (x << T) + (x >> (64 - T)) * 2^64.
template <int T>
__device__ __forceinline__ uint64_t shl_lazy(uint64_t x)
{
    static_assert(T >= 0 && T < 32, "shift below 32");
    if constexpr (T == 0) return x;
    else return reduce96_lazy(x << T, (uint32_t)(x >> (64 - T)));
}

// x * 2^32, lazy: lo(x) * 2^32 + hi(x) * 2^64 = pack(0, lo(x)) + hi(x) * EPS.
__device__ __forceinline__ uint64_t shl32_lazy(uint64_t x) { return reduce96_lazy(pack(0u, lo32(x)), hi32(x)); }

// x * 2^E mod p, canonical, for 0 <= E < 192 and any x < 2^64.
template <int E>
__device__ __forceinline__ uint64_t mul_pow2(uint64_t x)
{
    static_assert(E >= 0 && E < 192, "exponent below the order of 2");
    constexpr int t = E % 96;
    uint64_t r = shl_lazy<t % 32>(x);
    if constexpr (t >= 32) r = shl32_lazy(r);
    if constexpr (t >= 64) r = shl32_lazy(r);
    r = canon(r);
    if constexpr (E >= 96) r = neg(r);
    return r;
}

// x^7 with the addition chain x2 = x^2, x3 = x*x2, x4 = x2^2, x7 = x3*x4;
// lazy in and out (any x < 2^64; result below 2^64, congruent mod p).  x3
// and x4 are independent products, computed as a pair.
__device__ __forceinline__ uint64_t pow7_lazy(uint64_t x)
{
    const uint64_t x2 = mul_lazy(x, x);
    uint64_t x3, x4;
    mul_lazy2(x2, x, x2, x2, x3, x4);
    return mul_lazy(x3, x4);
}

// x^7 and y^7 in lockstep (every product paired).
__device__ __forceinline__ void pow7_lazy2(uint64_t &x, uint64_t &y)
{
    uint64_t x2, y2, x3, y3, x4, y4;
    mul_lazy2(x, x, y, y, x2, y2);
    mul_lazy2(x2, x, y2, y, x3, y3);
    mul_lazy2(x2, x2, y2, y2, x4, y4);
    mul_lazy2(x3, x4, y3, y4, x, y);
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
