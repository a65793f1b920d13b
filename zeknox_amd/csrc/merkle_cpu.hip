// Host (CPU) Merkle forest: fill_digests_buf_linear_cpu.
//
// The reference computes this entry point on the host with OpenMP
// (merkle/merkle.cpp:19-161, hashers in poseidon/poseidon.cpp,
// poseidon2/poseidon2.cu, monolith/monolith.cu, keccak/keccak.cu,
// poseidon/poseidon_bn128.cu); callers run it on machines without a GPU
// (wrappers/go/lib/lib_test.go:310-330).  This file is that path, host-only
// C++ (no device code, no HIP call): the same linear digest layout as the
// GPU forest, every hasher behind HashType, leaves of every subtree hashed
// together across a pool of std::threads (OMP_NUM_THREADS of them when set,
// else one per hardware thread), then each tree level across all subtrees.
//
// Poseidon and Poseidon2 leaves and levels go eight at a time through the
// AVX-512 forms in merkle_cpu_x8.hip where the CPU has them.
//
// Field arithmetic: canonical Goldilocks values, 128-bit products reduced
// with 2^64 = 2^32 - 1 and 2^96 = -1 (mod p).  Poseidon runs in plonky2's
// "fast partial round" form (the reference CPU's, poseidon.cpp:112-203): the
// 22 partial rounds as one dense 11x11 layer followed by sparse rounds.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "gl_constants.h"
#include "merkle_cpu.h"
#include "merkle_types.h"
#include "runtime.h"

namespace zkx {
namespace cpu {

namespace {

typedef unsigned __int128 u128;

constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p

inline uint64_t canon(uint64_t x) { return x >= P ? x - P : x; }

inline uint64_t fadd(uint64_t a, uint64_t b)  // a, b < p
{
    uint64_t s = a + b;
    // a wrap past 2^64 is worth EPS; s + EPS cannot wrap again since a, b < p
    s += (s < a) ? EPS : 0;
    return canon(s);
}

// x mod p for any 128-bit x: x = lo + 2^64 (h0 + 2^32 h1) = lo - h1 + h0 * EPS.
inline uint64_t reduce128(u128 x)
{
    const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    const uint64_t h0 = hi & EPS, h1 = hi >> 32;
    uint64_t t = lo - h1;
    if (lo < h1) t -= EPS;  // borrow of 2^64 = EPS (mod p); t >= 2^64 - 2^32 > EPS here
    const uint64_t m = h0 * EPS;  // < 2^64
    uint64_t r = t + m;
    if (r < m) r += EPS;  // carry: r < m < 2^64 - EPS after the wrap
    return canon(r);
}

inline uint64_t fmul(uint64_t a, uint64_t b) { return reduce128((u128)a * b); }

inline uint64_t pow7(uint64_t x)
{
    const uint64_t x2 = fmul(x, x), x3 = fmul(x2, x), x4 = fmul(x2, x2);
    return fmul(x3, x4);
}

// ---- Poseidon (width 12, 8 full + 22 partial rounds, x^7)

// Full-round MDS: y_r = sum_i CIRC[i] x_{(i+r) mod 12} + DIAG[r] x_r; the
// coefficients are below 2^7, so each row fits 128 bits before one reduction.
inline void poseidon_mds(uint64_t *s)
{
    uint64_t y[12];
    for (int r = 0; r < 12; r++) {
        u128 acc = (u128)s[r] * ZK_POSEIDON_MDS_DIAG[r];
        for (int i = 0; i < 12; i++) acc += (u128)s[(i + r) % 12] * ZK_POSEIDON_MDS_CIRC[i];
        y[r] = reduce128(acc);
    }
    memcpy(s, y, sizeof y);
}

inline void poseidon_full_round(uint64_t *s, int round)
{
    for (int i = 0; i < 12; i++) s[i] = pow7(fadd(s[i], ZK_POSEIDON_RC[12 * round + i]));
    poseidon_mds(s);
}

void poseidon_permute(uint64_t *s)
{
    for (int r = 0; r < 4; r++) poseidon_full_round(s, r);
    // partial rounds 4..25, factored: one constant layer, a dense 11x11 map
    // of elements 1..11, then per round an S-box on s0, its constant, and
    // the sparse layer  s0' = <w_hat_r, s[1..12)> + (CIRC[0]+DIAG[0]) s0,
    // s_i' = s_i + v_{r,i} s0.
    for (int i = 0; i < 12; i++) s[i] = fadd(s[i], ZK_POSEIDON_FAST_FIRST_RC[i]);
    {
        u128 acc[12] = {};
        for (int r = 1; r < 12; r++)
            for (int c = 1; c < 12; c++) {
                // 11 products of two 64-bit values: reduce every product's
                // high part as it comes to stay within 128 bits
                const u128 p = (u128)s[r] * ZK_POSEIDON_FAST_INIT_MATRIX[11 * (r - 1) + (c - 1)];
                acc[c] += reduce128(p);
            }
        for (int c = 1; c < 12; c++) s[c] = reduce128(acc[c]);
    }
    const uint64_t m00 = ZK_POSEIDON_MDS_CIRC[0] + ZK_POSEIDON_MDS_DIAG[0];
    for (int r = 0; r < 22; r++) {
        const uint64_t x0 = fadd(pow7(s[0]), ZK_POSEIDON_FAST_RC[r]);
        u128 d = (u128)x0 * m00;
        for (int i = 1; i < 12; i++) d += (u128)reduce128((u128)s[i] * ZK_POSEIDON_FAST_W_HATS[11 * r + (i - 1)]);
        for (int i = 1; i < 12; i++) s[i] = fadd(s[i], fmul(x0, ZK_POSEIDON_FAST_VS[11 * r + (i - 1)]));
        s[0] = reduce128(d);
    }
    for (int r = 26; r < 30; r++) poseidon_full_round(s, r);
}

// ---- Poseidon2 (width 12, external M4 blocks, diagonal internal layer)

inline void p2_external(uint64_t *s)
{
    // M4 = [[5,7,1,3],[4,6,1,1],[1,3,5,7],[1,1,4,6]] on each 4-block, then
    // every element gains the sum of its column across the three blocks.
    static const uint8_t M4[4][4] = {{5, 7, 1, 3}, {4, 6, 1, 1}, {1, 3, 5, 7}, {1, 1, 4, 6}};
    u128 y[12];
    for (int b = 0; b < 12; b += 4)
        for (int r = 0; r < 4; r++) {
            u128 acc = 0;
            for (int c = 0; c < 4; c++) acc += (u128)M4[r][c] * s[b + c];
            y[b + r] = acc;
        }
    for (int c = 0; c < 4; c++) {
        const u128 col = y[c] + y[4 + c] + y[8 + c];
        for (int b = 0; b < 12; b += 4) s[b + c] = reduce128(y[b + c] + col);
    }
}

void poseidon2_permute(uint64_t *s)
{
    p2_external(s);
    for (int r = 0; r < 4; r++) {
        for (int i = 0; i < 12; i++) s[i] = pow7(fadd(s[i], ZK_POSEIDON2_RC[12 * r + i]));
        p2_external(s);
    }
    for (int r = 4; r < 26; r++) {
        s[0] = pow7(fadd(s[0], ZK_POSEIDON2_RC[12 * r]));
        u128 sum = 0;
        for (int i = 0; i < 12; i++) sum += s[i];
        const uint64_t sm = reduce128(sum);
        for (int i = 0; i < 12; i++) s[i] = reduce128((u128)s[i] * ZK_POSEIDON2_DIAG[i] + sm);
    }
    for (int r = 26; r < 30; r++) {
        for (int i = 0; i < 12; i++) s[i] = pow7(fadd(s[i], ZK_POSEIDON2_RC[12 * r + i]));
        p2_external(s);
    }
}

// ---- Monolith-64 (width 12, 6 rounds of bars, bricks and concrete)

inline uint64_t monolith_bar(uint64_t x)
{
    // per byte: y = x ^ (rotl1(~x) & rotl2(x) & rotl3(x)), out rotl1(y),
    // all eight bytes at once with byte-lane masks
    auto rotl = [](uint64_t v, int k) {
        const uint64_t lo_mask = 0x0101010101010101ull * ((1u << k) - 1);  // low k bits of each byte
        return ((v << k) & ~lo_mask) | ((v >> (8 - k)) & lo_mask);
    };
    const uint64_t y = x ^ (rotl(~x, 1) & rotl(x, 2) & rotl(x, 3));
    return rotl(y, 1);
}

void monolith_permute(uint64_t *s)
{
    uint64_t carry[12] = {};  // bit 64 of each brick output
    for (int r = 0; r <= 6; r++) {
        if (r > 0) {
            for (int i = 0; i < 4; i++) s[i] = monolith_bar(s[i]);
            for (int i = 11; i > 0; i--) {
                const uint64_t sq = fmul(s[i - 1], s[i - 1]);
                const uint64_t t = s[i] + sq;
                carry[i] = t < sq;
                s[i] = t;
            }
        }
        uint64_t y[12];
        for (int row = 0; row < 12; row++) {
            // 65-bit inputs times coefficients < 2^5: 12 terms fit 128 bits
            u128 acc = ZK_MONOLITH_RC[12 * r + row];
            for (int c = 0; c < 12; c++) acc += (((u128)carry[c] << 64) | s[c]) * ZK_MONOLITH_MAT[12 * row + c];
            y[row] = reduce128(acc);
        }
        memcpy(s, y, sizeof y);
        memset(carry, 0, sizeof carry);
    }
}

// ---- Keccak-256 (original Keccak padding 0x01 ... 0x80, rate 136 bytes)

const uint64_t kKeccakRC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};
// rotation of lane x + 5y
const int kKeccakRho[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                            25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};

inline uint64_t rotl64(uint64_t v, int k) { return k ? (v << k) | (v >> (64 - k)) : v; }

void keccak_f(uint64_t *A)
{
    for (int round = 0; round < 24; round++) {
        uint64_t C[5], B[25];
        for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
        for (int x = 0; x < 5; x++) {
            const uint64_t D = C[(x + 4) % 5] ^ rotl64(C[(x + 1) % 5], 1);
            for (int y = 0; y < 25; y += 5) A[x + y] ^= D;
        }
        // rho and pi: lane (x, y) moves to (y, 2x + 3y)
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++) B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(A[x + 5 * y], kKeccakRho[x + 5 * y]);
        for (int y = 0; y < 25; y += 5)
            for (int x = 0; x < 5; x++) A[x + y] = B[x + y] ^ (~B[(x + 1) % 5 + y] & B[(x + 2) % 5 + y]);
        A[0] ^= kKeccakRC[round];
    }
}

// Keccak-256 of `len` bytes; the digest's first 25 bytes in out[0..4)
// (out[3] keeps its low byte only: the reference's 25-byte digests).
void keccak25(const uint8_t *in, size_t len, uint64_t *out)
{
    constexpr size_t R = 136;
    uint64_t A[25] = {};
    uint8_t block[R];
    while (true) {
        const size_t take = std::min(len, R);
        memset(block, 0, R);
        memcpy(block, in, take);
        const bool last = take < R;
        if (last) {
            block[take] ^= 0x01;
            block[R - 1] ^= 0x80;
        }
        for (size_t w = 0; w < R / 8; w++) {
            uint64_t v;
            memcpy(&v, block + 8 * w, 8);  // little-endian host
            A[w] ^= v;
        }
        keccak_f(A);
        if (last) break;
        in += R;
        len -= R;
    }
    out[0] = A[0];
    out[1] = A[1];
    out[2] = A[2];
    out[3] = A[3] & 0xFF;
}

// ---- Poseidon over BN254 (t = 5, x^5), Montgomery form with R = 2^256

const uint64_t kBnR[4] = {0x43e1f593f0000001ull, 0x2833e84879b97091ull, 0xb85045b68181585dull,
                          0x30644e72e131a029ull};

struct Bn {
    uint64_t v[4];
};

inline bool bn_ge_r(const uint64_t *a)
{
    for (int i = 3; i >= 0; i--)
        if (a[i] != kBnR[i]) return a[i] > kBnR[i];
    return true;
}

inline void bn_sub_r(uint64_t *a)
{
    uint64_t borrow = 0;
    for (int i = 0; i < 4; i++) {
        const u128 d = (u128)a[i] - kBnR[i] - borrow;
        a[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 64) & 1;
    }
}

inline Bn bn_add(const Bn &a, const Bn &b)
{
    Bn z;
    uint64_t c = 0;
    for (int i = 0; i < 4; i++) {
        const u128 s = (u128)a.v[i] + b.v[i] + c;
        z.v[i] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
    if (c || bn_ge_r(z.v)) bn_sub_r(z.v);
    return z;
}

struct BnConst {
    uint64_t ninv;  // -r^-1 mod 2^64
    Bn r2;          // 2^512 mod r
    BnConst()
    {
        uint64_t x = 1;  // Newton: x = r0^-1 mod 2^64, doubling the correct bits each step
        for (int i = 0; i < 6; i++) x *= 2 - kBnR[0] * x;
        ninv = 0 - x;
        Bn t{{1, 0, 0, 0}};
        for (int i = 0; i < 512; i++) t = bn_add(t, t);
        r2 = t;
    }
};

const BnConst &bn_const()
{
    static const BnConst c;
    return c;
}

// a * b * 2^-256 mod r, word-by-word Montgomery reduction interleaved with
// the product (a, b < r).
inline Bn bn_mul(const Bn &a, const Bn &b)
{
    const uint64_t ninv = bn_const().ninv;
    uint64_t t[6] = {};
    for (int i = 0; i < 4; i++) {
        uint64_t c = 0;
        for (int j = 0; j < 4; j++) {
            const u128 s = (u128)a.v[i] * b.v[j] + t[j] + c;
            t[j] = (uint64_t)s;
            c = (uint64_t)(s >> 64);
        }
        u128 s4 = (u128)t[4] + c;
        t[4] = (uint64_t)s4;
        t[5] = (uint64_t)(s4 >> 64);
        const uint64_t m = t[0] * ninv;
        u128 s = (u128)m * kBnR[0] + t[0];
        c = (uint64_t)(s >> 64);
        for (int j = 1; j < 4; j++) {
            s = (u128)m * kBnR[j] + t[j] + c;
            t[j - 1] = (uint64_t)s;
            c = (uint64_t)(s >> 64);
        }
        s4 = (u128)t[4] + c;
        t[3] = (uint64_t)s4;
        t[4] = t[5] + (uint64_t)(s4 >> 64);
    }
    Bn z{{t[0], t[1], t[2], t[3]}};
    if (t[4] || bn_ge_r(z.v)) bn_sub_r(z.v);
    return z;
}

inline Bn bn_at(const uint64_t *table, int k) { return Bn{{table[4 * k], table[4 * k + 1], table[4 * k + 2], table[4 * k + 3]}}; }

inline void bn_pow5(Bn &x)
{
    const Bn x2 = bn_mul(x, x), x4 = bn_mul(x2, x2);
    x = bn_mul(x4, x);
}

// st' = M^T st for a 5x5 matrix stored as m[5 j + i] (row j = input j)
inline void bn_mix(Bn *st, const uint64_t *m)
{
    Bn out[5];
    for (int i = 0; i < 5; i++) {
        Bn acc{{0, 0, 0, 0}};
        for (int j = 0; j < 5; j++) acc = bn_add(acc, bn_mul(bn_at(m, 5 * j + i), st[j]));
        out[i] = acc;
    }
    for (int i = 0; i < 5; i++) st[i] = out[i];
}

// The 12 Goldilocks words enter as four BN254 elements (words 3i, 3i+1, 3i+2
// most significant first) in state[1..4]; state[0..4) leave the same way,
// each word reduced once mod p (poseidon_bn128.cu:116-204).
void poseidon_bn128_permute(uint64_t *s)
{
    constexpr int T = 5, NP = 60;
    const Bn r2 = bn_const().r2;
    Bn st[T];
    st[0] = Bn{{0, 0, 0, 0}};
    for (int i = 0; i < 4; i++) st[i + 1] = bn_mul(Bn{{s[3 * i + 2], s[3 * i + 1], s[3 * i], 0}}, r2);
    auto ark = [&](int off) {
        for (int k = 0; k < T; k++) st[k] = bn_add(st[k], bn_at(ZK_BN128_C, off + k));
    };
    ark(0);
    for (int r = 0; r < 3; r++) {
        for (int k = 0; k < T; k++) bn_pow5(st[k]);
        ark((r + 1) * T);
        bn_mix(st, ZK_BN128_M);
    }
    for (int k = 0; k < T; k++) bn_pow5(st[k]);
    ark(4 * T);
    bn_mix(st, ZK_BN128_P);
    for (int r = 0; r < NP; r++) {
        bn_pow5(st[0]);
        st[0] = bn_add(st[0], bn_at(ZK_BN128_C, 5 * T + r));
        const uint64_t *row = ZK_BN128_S + 4 * ((2 * T - 1) * r);
        Bn n0{{0, 0, 0, 0}};
        for (int j = 0; j < T; j++) n0 = bn_add(n0, bn_mul(bn_at(row, j), st[j]));
        for (int k = 1; k < T; k++) st[k] = bn_add(st[k], bn_mul(st[0], bn_at(row, T + k - 1)));
        st[0] = n0;
    }
    for (int r = 0; r < 3; r++) {
        for (int k = 0; k < T; k++) bn_pow5(st[k]);
        ark(5 * T + NP + r * T);
        bn_mix(st, ZK_BN128_M);
    }
    for (int k = 0; k < T; k++) bn_pow5(st[k]);
    bn_mix(st, ZK_BN128_M);
    const Bn one{{1, 0, 0, 0}};
    for (int i = 0; i < 4; i++) {
        const Bn v = bn_mul(st[i], one);  // out of Montgomery form
        s[3 * i] = canon(v.v[2]);
        s[3 * i + 1] = canon(v.v[1]);
        s[3 * i + 2] = canon(v.v[0]);
    }
}

// ---- sponge (poseidon_permutation.hpp:83-139; keccak.cu:136-157)

using Permute = void (*)(uint64_t *);

Permute permutation_of(uint64_t hash_type)
{
    switch (hash_type) {
    case kPoseidon2: return poseidon2_permute;
    case kMonolith: return monolith_permute;
    case kPoseidonBN128: return poseidon_bn128_permute;
    default: return poseidon_permute;
    }
}

struct Hasher {
    uint64_t type;
    Permute perm;

    // out = H(x[0..n)): up to 4 elements are the digest themselves (zero
    // padded); otherwise 8-element chunks overwrite the state's rate part
    // before each permutation.  Keccak hashes the elements' bytes (fewer
    // than 4 elements are copied).
    void one(const uint64_t *x, uint64_t n, uint64_t *out) const
    {
        if (type == kKeccak) {
            if (n < 4) {
                for (uint64_t i = 0; i < 4; i++) out[i] = i < n ? x[i] : 0;
                return;
            }
            keccak25(reinterpret_cast<const uint8_t *>(x), n * 8, out);
            return;
        }
        if (n <= 4) {
            for (uint64_t i = 0; i < 4; i++) out[i] = i < n ? x[i] : 0;
            return;
        }
        uint64_t s[12] = {};
        for (uint64_t k = 0; k < n; k += 8) {
            memcpy(s, x + k, std::min<uint64_t>(8, n - k) * 8);
            perm(s);
        }
        memcpy(out, s, 32);
    }

    void two(const uint64_t *l, const uint64_t *r, uint64_t *out) const
    {
        if (type == kKeccak) {
            uint8_t b[50];
            memcpy(b, l, 25);  // each 25-byte digest, little-endian
            memcpy(b + 25, r, 25);
            keccak25(b, 50, out);
            return;
        }
        uint64_t s[12] = {};
        memcpy(s, l, 32);
        memcpy(s + 4, r, 32);
        perm(s);
        memcpy(out, s, 32);
    }
};

// ---- work sharing

unsigned pool_threads()
{
    static const unsigned n = [] {
        unsigned t = std::thread::hardware_concurrency();
        if (const char *e = getenv("OMP_NUM_THREADS")) {
            const long v = strtol(e, nullptr, 10);
            if (v > 0) t = (unsigned)v;
        }
        return std::max(1u, std::min(t, 256u));
    }();
    return n;
}

// f(i) for i in [0, n), split into chunks taken by up to pool_threads()
// threads; `cost` (permutations per item) keeps small levels on one thread.
template <class F>
void parallel_for(uint64_t n, uint64_t cost, F &&f)
{
    const uint64_t work = n * std::max<uint64_t>(cost, 1);
    unsigned nt = (unsigned)std::min<uint64_t>(pool_threads(), std::max<uint64_t>(1, work / 256));
    nt = (unsigned)std::min<uint64_t>(nt, n);
    if (nt <= 1) {
        for (uint64_t i = 0; i < n; i++) f(i);
        return;
    }
    const uint64_t chunk = std::max<uint64_t>(1, n / (8ull * nt));
    std::atomic<uint64_t> next{0};
    auto worker = [&] {
        while (true) {
            const uint64_t b = next.fetch_add(chunk);
            if (b >= n) break;
            const uint64_t e = std::min(n, b + chunk);
            for (uint64_t i = b; i < e; i++) f(i);
        }
    };
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (unsigned k = 1; k < nt; k++) th.emplace_back(worker);
    worker();
    for (auto &t : th) t.join();
}

}  // namespace

void fill_digests_linear(uint64_t *digests, uint64_t *caps, const uint64_t *leaves, uint64_t nd, uint64_t nc,
                         uint64_t nl, uint64_t leaf_size, uint64_t cap_height, uint64_t hash_type)
{
    if (hash_type > kMonolith)
        invalid("hash type " + std::to_string(hash_type) +
                " is not provided (Poseidon 0, Keccak 1, Poseidon-BN128 2, Poseidon2 3, Monolith 4)");
    if (!caps || !leaves) invalid("null cap or leaf buffer");
    if (nl == 0 || nc == 0 || (nl & (nl - 1)) || (nc & (nc - 1)) || nc > nl)
        invalid("leaves and caps must be powers of two with caps <= leaves");
    if (leaf_size == 0) invalid("leaf size out of range");
    const Hasher H{hash_type, permutation_of(hash_type)};
    const uint64_t chunks = (leaf_size + 7) / 8;
    // eight leaves or nodes per call where the hash has an AVX-512 form:
    // groups of 8 aligned items, inside one subtree whenever the subtree
    // width is a multiple of 8 (smaller widths run one at a time)
    const bool x8 = have_x8(hash_type);
    if (nc == nl) {  // every leaf is its own cap
        if (x8 && leaf_size > 4 && nl >= 8) {
            parallel_for(nl / 8, 8 * chunks, [&](uint64_t g) {
                hash_one_x8(hash_type, leaves + 8 * g * leaf_size, leaf_size, leaf_size, caps + 32 * g, 4);
            });
            return;
        }
        parallel_for(nl, chunks, [&](uint64_t t) { H.one(leaves + t * leaf_size, leaf_size, caps + 4 * t); });
        return;
    }
    if ((1ull << cap_height) != nc) invalid("cap_buf_size must be 2^cap_height");
    if (nd != 2 * (nl - nc)) invalid("digests_buf_size must be 2 * (leaves_buf_size - cap_buf_size)");
    if (!digests) invalid("null digest buffer");

    // subtree k: leaves [k sl, (k+1) sl), digests [k sdl, (k+1) sdl) with
    // digest i = heap node i + 2 (children 2i + 2, 2i + 3); the level of 2^j
    // nodes sits at [2^j - 2, 2^(j+1) - 2), the leaf hashes at [sl - 2, 2 sl - 2)
    const uint64_t sl = nl / nc, sdl = nd / nc;
    if (x8 && leaf_size > 4 && sl % 8 == 0) {
        parallel_for(nl / 8, 8 * chunks, [&](uint64_t g) {
            const uint64_t t = 8 * g, k = t / sl, i = t % sl;
            hash_one_x8(hash_type, leaves + t * leaf_size, leaf_size, leaf_size, digests + 4 * (k * sdl + sl - 2 + i),
                        4);
        });
    } else {
        parallel_for(nl, chunks, [&](uint64_t t) {
            const uint64_t k = t / sl, i = t % sl;
            H.one(leaves + t * leaf_size, leaf_size, digests + 4 * (k * sdl + sl - 2 + i));
        });
    }
    for (uint64_t w = sl / 2; w >= 2; w /= 2) {
        if (x8 && w % 8 == 0) {
            // nodes j..j+7 of a level: children at 8-word steps, outputs at 4
            parallel_for(nc * w / 8, 8, [&](uint64_t g) {
                const uint64_t q = 8 * g, k = q / w, j = w - 2 + q % w;
                uint64_t *d = digests + 4 * k * sdl;
                hash_two_x8(hash_type, d + 4 * (2 * j + 2), 8, d + 4 * j, 4);
            });
            continue;
        }
        parallel_for(nc * w, 1, [&](uint64_t q) {
            const uint64_t k = q / w, j = w - 2 + q % w;
            uint64_t *d = digests + 4 * k * sdl;
            H.two(d + 4 * (2 * j + 2), d + 4 * (2 * j + 3), d + 4 * j);
        });
    }
    parallel_for(nc, 1, [&](uint64_t k) {
        const uint64_t *d = digests + 4 * k * sdl;
        H.two(d, d + 4, caps + 4 * k);
    });
}

}  // namespace cpu
}  // namespace zkx
