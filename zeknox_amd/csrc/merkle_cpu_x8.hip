// Host (CPU) Poseidon and Poseidon2 hashing of eight leaves or eight tree
// nodes at once with AVX-512: the vectorised form of the host forest's
// leaf and level hashing (the reference's optional AVX CPU Merkle,
// merkle/merkle_avx.cpp, builds only against an external Goldilocks AVX
// library).  Host-only code; used when the CPU reports AVX-512F and DQ and
// ZKX_CPU_NO_AVX512 is unset or 0, else the scalar path runs.
//
// Lane j of every vector holds element e of permutation j.  Goldilocks
// arithmetic on canonical 64-bit lanes: a product from four 32 x 32 -> 64
// multiplies (vpmuludq), reduced with 2^64 = 2^32 - 1 and 2^96 = -1 (mod p);
// sums with the carry folded back as 2^32 - 1.  The MDS and Poseidon2's
// external layer have coefficients below 2^7: their rows are accumulated as
// two 64-bit sums of 32-bit halves and reduced once.  Poseidon runs every
// round in its textbook form (constants on all twelve elements, S-box on all
// or on element 0, full MDS), Poseidon2 as in poseidon2.cu:65-209; the
// arithmetic is exact mod p, so the results are the scalar path's.
#include <immintrin.h>
#include <stdint.h>
#include <stdlib.h>

#include "gl_constants.h"
#include "merkle_cpu.h"
#include "merkle_types.h"

namespace zkx {
namespace cpu {

namespace {

#define ZKX_X8 __attribute__((target("avx512f,avx512dq"))) inline

typedef __m512i V;

constexpr uint64_t kP = 0xFFFFFFFF00000001ull;
constexpr uint64_t kEps = 0xFFFFFFFFull;

ZKX_X8 V vset(uint64_t c) { return _mm512_set1_epi64((long long)c); }

// a + b mod p, a and b canonical
ZKX_X8 V vadd(V a, V b)
{
    V s = _mm512_add_epi64(a, b);
    const __mmask8 c = _mm512_cmplt_epu64_mask(s, a);  // wrapped: + 2^64 = + EPS
    s = _mm512_mask_add_epi64(s, c, s, vset(kEps));
    const __mmask8 ge = _mm512_cmpge_epu64_mask(s, vset(kP));
    return _mm512_mask_sub_epi64(s, ge, s, vset(kP));
}

// Canonical residue of hi * 2^64 + lo: hi * 2^64 = h0 * EPS - h1 (mod p).
ZKX_X8 V vreduce(V lo, V hi)
{
    const V h1 = _mm512_srli_epi64(hi, 32), h0 = _mm512_and_si512(hi, vset(kEps));
    V t = _mm512_sub_epi64(lo, h1);
    const __mmask8 b = _mm512_cmplt_epu64_mask(lo, h1);  // borrowed 2^64 = EPS
    t = _mm512_mask_sub_epi64(t, b, t, vset(kEps));
    const V m = _mm512_sub_epi64(_mm512_slli_epi64(h0, 32), h0);  // h0 * EPS < 2^64
    V r = _mm512_add_epi64(t, m);
    const __mmask8 c = _mm512_cmplt_epu64_mask(r, m);
    r = _mm512_mask_add_epi64(r, c, r, vset(kEps));
    const __mmask8 ge = _mm512_cmpge_epu64_mask(r, vset(kP));
    return _mm512_mask_sub_epi64(r, ge, r, vset(kP));
}

// a * b mod p for any a, b < 2^64
ZKX_X8 V vmul(V a, V b)
{
    const V ah = _mm512_srli_epi64(a, 32), bh = _mm512_srli_epi64(b, 32);
    const V ll = _mm512_mul_epu32(a, b), lh = _mm512_mul_epu32(a, bh);
    const V hl = _mm512_mul_epu32(ah, b), hh = _mm512_mul_epu32(ah, bh);
    const V mid = _mm512_add_epi64(lh, hl);
    const __mmask8 c1 = _mm512_cmplt_epu64_mask(mid, lh);  // mid's 2^64: the product's 2^96
    const V lo = _mm512_add_epi64(ll, _mm512_slli_epi64(mid, 32));
    const __mmask8 c2 = _mm512_cmplt_epu64_mask(lo, ll);
    V hi = _mm512_add_epi64(hh, _mm512_srli_epi64(mid, 32));
    hi = _mm512_mask_add_epi64(hi, c2, hi, vset(1));
    hi = _mm512_mask_add_epi64(hi, c1, hi, vset(1ull << 32));
    return vreduce(lo, hi);
}

ZKX_X8 V vpow7(V x)
{
    const V x2 = vmul(x, x), x3 = vmul(x2, x), x4 = vmul(x2, x2);
    return vmul(x3, x4);
}

// A sum of products of canonical elements with coefficients < 2^7, kept as
// the sums of their 32-bit halves' products (each below 2^46 for up to 2^7
// terms): value = lo + hi * 2^32.
struct Acc {
    V lo, hi;
};

ZKX_X8 Acc acc_zero() { return Acc{_mm512_setzero_si512(), _mm512_setzero_si512()}; }

ZKX_X8 void acc_mac(Acc &a, V xlo, V xhi, uint64_t c)
{
    const V cv = vset(c);
    a.lo = _mm512_add_epi64(a.lo, _mm512_mul_epu32(xlo, cv));
    a.hi = _mm512_add_epi64(a.hi, _mm512_mul_epu32(xhi, cv));
}

ZKX_X8 Acc acc_add(Acc a, Acc b) { return Acc{_mm512_add_epi64(a.lo, b.lo), _mm512_add_epi64(a.hi, b.hi)}; }

ZKX_X8 V acc_reduce(Acc a)
{
    // lo + hi * 2^32 = (lo + (hi << 32)) + (hi >> 32) * 2^64, with the low
    // sum's carry into the high word
    const V l = _mm512_add_epi64(a.lo, _mm512_slli_epi64(a.hi, 32));
    const __mmask8 c = _mm512_cmplt_epu64_mask(l, a.lo);
    V h = _mm512_srli_epi64(a.hi, 32);
    h = _mm512_mask_add_epi64(h, c, h, vset(1));
    return vreduce(l, h);
}

// ---- Poseidon

ZKX_X8 void poseidon_mds8(V *s)
{
    V lo[12], hi[12];
    for (int i = 0; i < 12; i++) {
        lo[i] = s[i];  // vpmuludq reads the low halves
        hi[i] = _mm512_srli_epi64(s[i], 32);
    }
    for (int r = 0; r < 12; r++) {
        Acc a = acc_zero();
        acc_mac(a, lo[r], hi[r], ZK_POSEIDON_MDS_DIAG[r]);
        for (int i = 0; i < 12; i++) acc_mac(a, lo[(i + r) % 12], hi[(i + r) % 12], ZK_POSEIDON_MDS_CIRC[i]);
        s[r] = acc_reduce(a);
    }
}

ZKX_X8 void poseidon_permute8(V *s)
{
    for (int r = 0; r < 30; r++) {
        const bool full = r < 4 || r >= 26;
        for (int i = 0; i < 12; i++) s[i] = vadd(s[i], vset(ZK_POSEIDON_RC[12 * r + i]));
        if (full)
            for (int i = 0; i < 12; i++) s[i] = vpow7(s[i]);
        else
            s[0] = vpow7(s[0]);
        poseidon_mds8(s);
    }
}

// ---- Poseidon2

ZKX_X8 void p2_external8(V *s)
{
    static const uint8_t M4[4][4] = {{5, 7, 1, 3}, {4, 6, 1, 1}, {1, 3, 5, 7}, {1, 1, 4, 6}};
    V lo[12], hi[12];
    for (int i = 0; i < 12; i++) {
        lo[i] = s[i];
        hi[i] = _mm512_srli_epi64(s[i], 32);
    }
    Acc y[12];
    for (int b = 0; b < 12; b += 4)
        for (int r = 0; r < 4; r++) {
            y[b + r] = acc_zero();
            for (int c = 0; c < 4; c++) acc_mac(y[b + r], lo[b + c], hi[b + c], M4[r][c]);
        }
    for (int c = 0; c < 4; c++) {
        const Acc col = acc_add(acc_add(y[c], y[4 + c]), y[8 + c]);
        for (int b = 0; b < 12; b += 4) s[b + c] = acc_reduce(acc_add(y[b + c], col));
    }
}

ZKX_X8 void poseidon2_permute8(V *s)
{
    p2_external8(s);
    for (int r = 0; r < 30; r++) {
        if (r < 4 || r >= 26) {
            for (int i = 0; i < 12; i++) s[i] = vpow7(vadd(s[i], vset(ZK_POSEIDON2_RC[12 * r + i])));
            p2_external8(s);
        } else {
            s[0] = vpow7(vadd(s[0], vset(ZK_POSEIDON2_RC[12 * r])));
            // sum of the twelve canonical elements: halves summed, one reduction
            Acc sum = acc_zero();
            for (int i = 0; i < 12; i++) acc_mac(sum, s[i], _mm512_srli_epi64(s[i], 32), 1);
            const V sm = acc_reduce(sum);
            for (int i = 0; i < 12; i++) s[i] = vadd(vmul(s[i], vset(ZK_POSEIDON2_DIAG[i])), sm);
        }
    }
}

ZKX_X8 void permute8(uint64_t hash_type, V *s)
{
    if (hash_type == kPoseidon2) poseidon2_permute8(s);
    else poseidon_permute8(s);
}

ZKX_X8 V lane_offsets(uint64_t stride)
{
    return _mm512_mullo_epi64(_mm512_set_epi64(7, 6, 5, 4, 3, 2, 1, 0), vset(stride));
}

__attribute__((target("avx512f,avx512dq"))) void one_x8(uint64_t hash_type, const uint64_t *x, uint64_t stride,
                                                      uint64_t n, uint64_t *out, uint64_t out_stride)
{
    const V in_off = lane_offsets(stride), out_off = lane_offsets(out_stride);
    V s[12];
    for (int i = 0; i < 12; i++) s[i] = _mm512_setzero_si512();
    for (uint64_t k = 0; k < n; k += 8) {
        const uint64_t len = n - k < 8 ? n - k : 8;
        for (uint64_t e = 0; e < len; e++) s[e] = _mm512_i64gather_epi64(in_off, (const void *)(x + k + e), 8);
        permute8(hash_type, s);
    }
    for (int e = 0; e < 4; e++) _mm512_i64scatter_epi64((void *)(out + e), out_off, s[e], 8);
}

__attribute__((target("avx512f,avx512dq"))) void two_x8(uint64_t hash_type, const uint64_t *l, uint64_t stride,
                                                      uint64_t *out, uint64_t out_stride)
{
    const V in_off = lane_offsets(stride), out_off = lane_offsets(out_stride);
    V s[12];
    for (int e = 0; e < 8; e++) s[e] = _mm512_i64gather_epi64(in_off, (const void *)(l + e), 8);
    for (int e = 8; e < 12; e++) s[e] = _mm512_setzero_si512();
    permute8(hash_type, s);
    for (int e = 0; e < 4; e++) _mm512_i64scatter_epi64((void *)(out + e), out_off, s[e], 8);
}

}  // namespace

bool have_x8(uint64_t hash_type)
{
    if (hash_type != kPoseidon && hash_type != kPoseidon2) return false;
    if (const char *e = getenv("ZKX_CPU_NO_AVX512"))
        if (e[0] && e[0] != '0') return false;
    static const bool cpu = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq");
    return cpu;
}

void hash_one_x8(uint64_t hash_type, const uint64_t *x, uint64_t stride, uint64_t n, uint64_t *out,
                 uint64_t out_stride)
{
    one_x8(hash_type, x, stride, n, out, out_stride);
}

void hash_two_x8(uint64_t hash_type, const uint64_t *l, uint64_t stride, uint64_t *out, uint64_t out_stride)
{
    two_x8(hash_type, l, stride, out, out_stride);
}

}  // namespace cpu
}  // namespace zkx
