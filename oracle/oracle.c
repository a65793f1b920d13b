/*
 * oracle.c -- CPU restatement of the zeknox LDE + Merkle-commit path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against; it is never linked into, loaded by or called from the
 * product library.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline use it.
 *
 * Clarity over speed: field arithmetic uses unsigned __int128 and the textbook
 * definitions; Poseidon uses the naive partial rounds (full 12x12 MDS every
 * round), which is the permutation's definition.  The reference's CPU path
 * computes the same function through the "fast partial round" factorisation.
 *
 * What each function restates (reference file:line) and what pins it:
 *   gl_add/gl_sub/gl_mul     ff/goldilocks.hpp:139-177 (modulo_add/sub, reduce128)
 *                            pinned: brute force vs Python big ints on edge values
 *   gl_omega                 ff/gl64_params.hpp:15-49 (OMEGA[k]) from OMEGA[32]
 *                            pinned: closed forms (order 2^k exactly, OMEGA[1]=p-1,
 *                            OMEGA[2]=2^48) and the table's OMEGA[32]
 *   poseidon_permute         poseidon/poseidon.cu:613-838 (naive rounds),
 *                            poseidon/poseidon.cpp:75-110,401-408
 *                            pinned: KATs tests/tests.cu:281-292, Go caps
 *                            wrappers/go/lib/lib_test.go:197-213, reference CPU
 *   poseidon2_permute        poseidon2/poseidon2.cu:65-209
 *                            pinned: KAT tests/tests.cu:425-428, reference CPU
 *   monolith_permute         monolith/monolith.cu:104-227 (Monolith-64)
 *                            pinned: KAT tests/tests.cu:201-235 (hash of 0..11),
 *                            reference CPU (oracle/_ref)
 *   keccak_f1600 / keccak    FIPS 202 (rho offsets and round constants generated
 *                            by their defining rules); keccak/keccak.cu:39-157
 *                            pinned: SHA3-256 (ds 0x06) vs Python hashlib, KATs
 *                            tests/tests.cu:162-199, reference CPU (oracle/_ref)
 *   bn_mont_mul / bn_add     poseidon/element_bn128.cu (BN254 Fr, Montgomery R = 2^256)
 *                            pinned: Python big integers (a*b*R^-1 mod r, a+b mod r)
 *   poseidon_bn128_permute   poseidon/poseidon_bn128.cu:116-204 (t = 5, x^5, optimised rounds)
 *                            pinned: KAT tests/tests.cu:492-536, reference CPU (oracle/_ref)
 *   hash_one / hash_two      poseidon/poseidon_permutation.hpp:83-139
 *   fill_digests             merkle/merkle.cpp:19-161 (linear forest layout)
 *                            pinned: Go caps KATs, reference CPU (oracle/_ref)
 *   ntt / lde                ntt/ntt.cu:106-158,253-309,477-564; kernels.cu:14-212
 *                            (natural-order coset NTT, zero-padded LDE; its
 *                            field product uses the Goldilocks identities)
 *                            pinned: brute-force O(n^2) DFT, closed forms,
 *                            inverse-of-forward identity, Horner spot checks
 *   transpose_rev            ntt/kernels.cu:94-135; wrappers/rust/tests/utils.rs:15-27
 *                            pinned: brute-force permutation formula in Python
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../zeknox_amd/csrc/gl_constants.h"

typedef uint64_t u64;
typedef unsigned __int128 u128;

#define GL_P 0xFFFFFFFF00000001ull
#define GL_OMEGA32 1753635133440165772ull /* gl64_params.hpp:48 */

/* ------------------------------------------------------------------ field */

u64 orc_gl_add(u64 a, u64 b) { return (u64)(((u128)a + b) % GL_P); }
u64 orc_gl_sub(u64 a, u64 b) { return (u64)(((u128)a + GL_P - (b % GL_P)) % GL_P); }
u64 orc_gl_mul(u64 a, u64 b) { return (u64)(((u128)a * b) % GL_P); }

u64 orc_gl_pow(u64 b, u64 e)
{
    u64 r = 1;
    b %= GL_P;
    while (e) {
        if (e & 1) r = orc_gl_mul(r, b);
        b = orc_gl_mul(b, b);
        e >>= 1;
    }
    return r;
}

u64 orc_gl_inv(u64 a) { return orc_gl_pow(a, GL_P - 2); }

/* primitive 2^k-th root of unity, OMEGA[k] of gl64_params.hpp */
u64 orc_gl_omega(unsigned k)
{
    u64 w = GL_OMEGA32;
    for (unsigned i = k; i < 32; i++) w = orc_gl_mul(w, w);
    return w;
}

static u64 sbox7(u64 x)
{
    u64 x2 = orc_gl_mul(x, x);
    u64 x4 = orc_gl_mul(x2, x2);
    return orc_gl_mul(orc_gl_mul(x4, x2), x);
}

/* --------------------------------------------------------------- poseidon */

static void poseidon_mds(u64 *s)
{
    u64 out[12];
    for (int r = 0; r < 12; r++) {
        u128 acc = 0;
        for (int i = 0; i < 12; i++) acc += (u128)s[(i + r) % 12] * ZK_POSEIDON_MDS_CIRC[i];
        acc += (u128)s[r] * ZK_POSEIDON_MDS_DIAG[r];
        out[r] = (u64)(acc % GL_P);
    }
    memcpy(s, out, sizeof(out));
}

void orc_poseidon_permute(u64 *s)
{
    int rc = 0;
    for (int r = 0; r < 4; r++, rc++) {
        for (int i = 0; i < 12; i++) s[i] = sbox7(orc_gl_add(s[i], ZK_POSEIDON_RC[12 * rc + i]));
        poseidon_mds(s);
    }
    for (int r = 0; r < 22; r++, rc++) {
        for (int i = 0; i < 12; i++) s[i] = orc_gl_add(s[i], ZK_POSEIDON_RC[12 * rc + i]);
        s[0] = sbox7(s[0]);
        poseidon_mds(s);
    }
    for (int r = 0; r < 4; r++, rc++) {
        for (int i = 0; i < 12; i++) s[i] = sbox7(orc_gl_add(s[i], ZK_POSEIDON_RC[12 * rc + i]));
        poseidon_mds(s);
    }
}

/* -------------------------------------------------------------- poseidon2 */

/* M4 = [[5,7,1,3],[4,6,1,1],[1,3,5,7],[1,1,4,6]] applied to one 4-block,
 * the matrix that poseidon2.cu:65-82 computes with additions. */
static void p2_m4(u64 *x)
{
    static const u64 M4[4][4] = {{5, 7, 1, 3}, {4, 6, 1, 1}, {1, 3, 5, 7}, {1, 1, 4, 6}};
    u64 y[4];
    for (int r = 0; r < 4; r++) {
        u128 acc = 0;
        for (int c = 0; c < 4; c++) acc += (u128)M4[r][c] * x[c];
        y[r] = (u64)(acc % GL_P);
    }
    memcpy(x, y, sizeof(y));
}

static void p2_external(u64 *s)
{
    for (int b = 0; b < 12; b += 4) p2_m4(s + b);
    for (int c = 0; c < 4; c++) {
        u64 sum = orc_gl_add(orc_gl_add(s[c], s[4 + c]), s[8 + c]);
        for (int b = 0; b < 12; b += 4) s[b + c] = orc_gl_add(s[b + c], sum);
    }
}

void orc_poseidon2_permute(u64 *s)
{
    p2_external(s);
    for (int r = 0; r < 4; r++) {
        for (int i = 0; i < 12; i++) s[i] = sbox7(orc_gl_add(s[i], ZK_POSEIDON2_RC[12 * r + i]));
        p2_external(s);
    }
    for (int r = 4; r < 26; r++) {
        s[0] = sbox7(orc_gl_add(s[0], ZK_POSEIDON2_RC[12 * r]));
        u64 sum = 0;
        for (int i = 0; i < 12; i++) sum = orc_gl_add(sum, s[i]);
        for (int i = 0; i < 12; i++) s[i] = orc_gl_add(orc_gl_mul(s[i], ZK_POSEIDON2_DIAG[i]), sum);
    }
    for (int r = 26; r < 30; r++) {
        for (int i = 0; i < 12; i++) s[i] = sbox7(orc_gl_add(s[i], ZK_POSEIDON2_RC[12 * r + i]));
        p2_external(s);
    }
}

/* --------------------------------------------------------------- monolith */

/* One byte of Monolith-64's bar S-box: y = x ^ (rotl1(~x) & rotl2(x) &
 * rotl3(x)), output rotl1(y) (8-bit rotations), monolith.cu:107-117. */
static unsigned mono_sbox8(unsigned x)
{
#define ROTL8(v, k) ((((v) << (k)) | ((v) >> (8 - (k)))) & 0xFFu)
    unsigned y = x ^ (ROTL8((~x) & 0xFFu, 1) & ROTL8(x, 2) & ROTL8(x, 3));
    return ROTL8(y, 1);
#undef ROTL8
}

/* Monolith-64 permutation, monolith.cu:119-227: concrete (rc row 0 = 0),
 * then 6 x [bars on s[0..4) bytewise, bricks s[i] += s[i-1]^2 (integer add,
 * carry kept as bit 64), concrete with rc row r].  concrete: row-times-matrix
 * over the 65-bit values plus rc, reduced mod p. */
void orc_monolith_permute(u64 *s)
{
    unsigned hi[12] = {0};
    for (int r = 0; r <= 6; r++) {
        if (r > 0) {
            for (int i = 0; i < 4; i++) {
                u64 v = 0;
                for (int b = 0; b < 8; b++) v |= (u64)mono_sbox8((unsigned)(s[i] >> (8 * b)) & 0xFFu) << (8 * b);
                s[i] = v;
            }
            for (int i = 11; i > 0; i--) {
                u64 sq = orc_gl_mul(s[i - 1], s[i - 1]);
                u64 t = s[i] + sq;
                hi[i] = t < sq;
                s[i] = t;
            }
        }
        u64 out[12];
        for (int row = 0; row < 12; row++) {
            u128 acc = ZK_MONOLITH_RC[12 * r + row];
            for (int col = 0; col < 12; col++) {
                u128 v = ((u128)hi[col] << 64) | s[col];
                acc += v * ZK_MONOLITH_MAT[12 * row + col];
            }
            out[row] = (u64)(acc % GL_P);
        }
        memcpy(s, out, sizeof(out));
        memset(hi, 0, sizeof(hi));
    }
}

/* ----------------------------------------------------------------- keccak */

/* Keccak-f[1600] from its definition (FIPS 202 3.2): theta, rho, pi, chi,
 * iota on lanes A[x + 5y].  The rho offsets and iota round constants are
 * generated here by their defining rules (3.2.2 and the rc(t) LFSR of 3.2.5)
 * rather than copied, so the permutation is pinned independently. */
static void keccak_f1600(u64 *A)
{
    static int ready = 0;
    static unsigned rho[25];
    static u64 RC[24];
    if (!ready) {
        unsigned x = 1, y = 0;
        rho[0] = 0;
        for (unsigned t = 0; t < 24; t++) {
            rho[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
            unsigned nx = y, ny = (2 * x + 3 * y) % 5;
            x = nx;
            y = ny;
        }
        unsigned lfsr = 1; /* rc(t): x^8 + x^6 + x^5 + x^4 + 1 over GF(2) */
        for (int ir = 0; ir < 24; ir++) {
            u64 c = 0;
            for (int j = 0; j <= 6; j++) {
                if (lfsr & 1) c |= 1ull << ((1u << j) - 1);
                lfsr = (lfsr & 0x80) ? ((lfsr << 1) ^ 0x171) : (lfsr << 1);
            }
            RC[ir] = c;
        }
        ready = 1;
    }
#define ROTL64(v, k) ((k) ? (((v) << (k)) | ((v) >> (64 - (k)))) : (v))
    for (int ir = 0; ir < 24; ir++) {
        u64 C[5], D[5], B[25];
        for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
        for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ ROTL64(C[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) A[i] ^= D[i % 5];
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++) B[y + 5 * ((2 * x + 3 * y) % 5)] = ROTL64(A[x + 5 * y], rho[x + 5 * y]);
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                A[x + 5 * y] = B[x + 5 * y] ^ ((~B[(x + 1) % 5 + 5 * y]) & B[(x + 2) % 5 + 5 * y]);
        A[0] ^= RC[ir];
    }
#undef ROTL64
}

/* Keccak sponge over bytes with rate `rate` bytes and padding byte `ds`
 * (0x01: the original Keccak padding the reference uses, keccak.cu:98-120;
 * 0x06: SHA-3), output `outlen` bytes. */
void orc_keccak(const unsigned char *in, u64 len, unsigned char *out, unsigned outlen, unsigned rate, unsigned ds)
{
    u64 A[25] = {0};
    unsigned char block[200];
    for (;;) {
        unsigned take = len < rate ? (unsigned)len : rate;
        memset(block, 0, sizeof(block));
        memcpy(block, in, take);
        if (take < rate) {
            block[take] ^= (unsigned char)ds;
            block[rate - 1] ^= 0x80;
        }
        for (unsigned i = 0; i < rate / 8; i++) {
            u64 w = 0;
            for (int b = 0; b < 8; b++) w |= (u64)block[8 * i + b] << (8 * b);
            A[i] ^= w;
        }
        keccak_f1600(A);
        if (take < rate) break;
        in += rate;
        len -= rate;
    }
    for (unsigned i = 0; i < outlen; i++) out[i] = (unsigned char)(A[i / 8] >> (8 * (i % 8)));
}

/* KeccakHasher, keccak.cu:136-157: Keccak-256 of the input words' bytes,
 * keeping 25 bytes (the top 7 bytes of word 3 cleared); fewer than 4 words
 * are copied (zero-padded here; the reference leaves the rest of its output
 * buffer as it was).  hash_two hashes the first 25 bytes of each digest. */
static void keccak_hash_one(const u64 *in, u64 n, u64 *out)
{
    if (n < 4) {
        for (u64 i = 0; i < 4; i++) out[i] = i < n ? in[i] : 0;
        return;
    }
    unsigned char *bytes = (unsigned char *)malloc(n * 8), md[32];
    for (u64 i = 0; i < n * 8; i++) bytes[i] = (unsigned char)(in[i / 8] >> (8 * (i % 8)));
    orc_keccak(bytes, n * 8, md, 32, 136, 0x01);
    free(bytes);
    for (int i = 0; i < 4; i++) {
        u64 w = 0;
        for (int b = 0; b < 8; b++) w |= (u64)md[8 * i + b] << (8 * b);
        out[i] = w;
    }
    out[3] &= 0xFF;
}

static void keccak_hash_two(const u64 *l, const u64 *r, u64 *out)
{
    unsigned char bytes[50], md[32];
    for (int i = 0; i < 25; i++) {
        bytes[i] = (unsigned char)(l[i / 8] >> (8 * (i % 8)));
        bytes[25 + i] = (unsigned char)(r[i / 8] >> (8 * (i % 8)));
    }
    orc_keccak(bytes, 50, md, 32, 136, 0x01);
    for (int i = 0; i < 4; i++) {
        u64 w = 0;
        for (int b = 0; b < 8; b++) w |= (u64)md[8 * i + b] << (8 * b);
        out[i] = w;
    }
    out[3] &= 0xFF;
}

/* ---------------------------------------------------------- bn254 poseidon */

/* The BN254 scalar field r (four 64-bit limbs, least significant first) and
 * Montgomery multiplication with R = 2^256 by the CIOS method (Koc, Acar,
 * Kaliski 1996).  R^2 mod r and -r^-1 mod 2^64 are computed here from r. */
static const u64 BN_R[4] = {0x43e1f593f0000001ull, 0x2833e84879b97091ull, 0xb85045b68181585dull,
                            0x30644e72e131a029ull};

static int bn_geq(const u64 *a, const u64 *b)
{
    for (int i = 3; i >= 0; i--)
        if (a[i] != b[i]) return a[i] > b[i];
    return 1;
}

static void bn_sub_r(u64 *a)
{
    u128 borrow = 0;
    for (int i = 0; i < 4; i++) {
        u128 d = (u128)a[i] - BN_R[i] - borrow;
        a[i] = (u64)d;
        borrow = (d >> 64) ? 1 : 0;
    }
}

/* z = a + b mod r for a, b < r */
void orc_bn_add(u64 *z, const u64 *a, const u64 *b)
{
    u128 c = 0;
    u64 t[4];
    for (int i = 0; i < 4; i++) {
        c += (u128)a[i] + b[i];
        t[i] = (u64)c;
        c >>= 64;
    }
    if (c || bn_geq(t, BN_R)) bn_sub_r(t);
    memcpy(z, t, sizeof(t));
}

static u64 bn_ninv(void)
{
    /* -r^-1 mod 2^64 by Newton iteration on the inverse of r[0] (odd) */
    u64 x = 1;
    for (int i = 0; i < 7; i++) x *= 2 - BN_R[0] * x;
    return (u64)0 - x;
}

/* z = a * b * R^-1 mod r for a, b < r */
void orc_bn_mont_mul(u64 *z, const u64 *a, const u64 *b)
{
    const u64 ninv = bn_ninv();
    u64 t[6] = {0};
    for (int i = 0; i < 4; i++) {
        u128 c = 0;
        for (int j = 0; j < 4; j++) {
            c += (u128)t[j] + (u128)a[i] * b[j];
            t[j] = (u64)c;
            c >>= 64;
        }
        c += t[4];
        t[4] = (u64)c;
        t[5] = (u64)(c >> 64);
        const u64 m = t[0] * ninv;
        c = (u128)t[0] + (u128)m * BN_R[0];
        c >>= 64;
        for (int j = 1; j < 4; j++) {
            c += (u128)t[j] + (u128)m * BN_R[j];
            t[j - 1] = (u64)c;
            c >>= 64;
        }
        c += t[4];
        t[3] = (u64)c;
        t[4] = t[5] + (u64)(c >> 64);
    }
    if (t[4] || bn_geq(t, BN_R)) bn_sub_r(t);
    memcpy(z, t, 4 * sizeof(u64));
}

static void bn_r2(u64 *out) /* R^2 mod r = 2^512 mod r, by doubling */
{
    u64 x[4] = {1, 0, 0, 0};
    for (int i = 0; i < 512; i++) orc_bn_add(x, x, x);
    memcpy(out, x, sizeof(x));
}

static void bn_exp5(u64 *x)
{
    u64 x2[4], x4[4];
    orc_bn_mont_mul(x2, x, x);
    orc_bn_mont_mul(x4, x2, x2);
    orc_bn_mont_mul(x, x4, x);
}

/* Poseidon over BN254 with t = 5 in the optimised form (full rounds with M,
 * the last first-half round with P, 60 partial rounds with the sparse S
 * rows), poseidon/poseidon_bn128.cu:116-192.  The 12 Goldilocks words map to
 * state[1..4] as element i = w[3i] 2^128 + w[3i+1] 2^64 + w[3i+2]; the
 * output words of state[0..4) come back the same way (the top limb dropped)
 * and each is reduced once mod p (:195-204). */
void orc_poseidon_bn128_permute(u64 *s)
{
    const int t = 5, nrp = 60;
    u64 r2[4], st[5][4], ns[5][4], tmp[4];
    bn_r2(r2);
    memset(st, 0, sizeof(st));
    for (int i = 0; i < 4; i++) {
        u64 v[4] = {s[3 * i + 2], s[3 * i + 1], s[3 * i], 0};
        orc_bn_mont_mul(st[i + 1], v, r2);
    }
#define ARK(off)                                                                     \
    for (int k = 0; k < t; k++) orc_bn_add(st[k], st[k], ZK_BN128_C + 4 * ((off) + k))
#define MIX(mat)                                                                     \
    do {                                                                             \
        for (int i = 0; i < t; i++) {                                                \
            u64 acc[4] = {0, 0, 0, 0};                                               \
            for (int j = 0; j < t; j++) {                                            \
                orc_bn_mont_mul(tmp, (mat) + 4 * (t * j + i), st[j]);                \
                orc_bn_add(acc, acc, tmp);                                           \
            }                                                                        \
            memcpy(ns[i], acc, sizeof(acc));                                         \
        }                                                                            \
        memcpy(st, ns, sizeof(st));                                                  \
    } while (0)
    ARK(0);
    for (int r = 0; r < 3; r++) {
        for (int k = 0; k < t; k++) bn_exp5(st[k]);
        ARK((r + 1) * t);
        MIX(ZK_BN128_M);
    }
    for (int k = 0; k < t; k++) bn_exp5(st[k]);
    ARK(4 * t);
    MIX(ZK_BN128_P);
    for (int r = 0; r < nrp; r++) {
        bn_exp5(st[0]);
        orc_bn_add(st[0], st[0], ZK_BN128_C + 4 * (5 * t + r));
        const u64 *row = ZK_BN128_S + 4 * ((2 * t - 1) * r);
        u64 n0[4] = {0, 0, 0, 0};
        for (int j = 0; j < t; j++) {
            orc_bn_mont_mul(tmp, row + 4 * j, st[j]);
            orc_bn_add(n0, n0, tmp);
        }
        for (int k = 1; k < t; k++) {
            orc_bn_mont_mul(tmp, st[0], row + 4 * (t + k - 1));
            orc_bn_add(st[k], st[k], tmp);
        }
        memcpy(st[0], n0, sizeof(n0));
    }
    for (int r = 0; r < 3; r++) {
        for (int k = 0; k < t; k++) bn_exp5(st[k]);
        ARK(5 * t + nrp + r * t);
        MIX(ZK_BN128_M);
    }
    for (int k = 0; k < t; k++) bn_exp5(st[k]);
    MIX(ZK_BN128_M);
#undef ARK
#undef MIX
    const u64 one[4] = {1, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        u64 v[4];
        orc_bn_mont_mul(v, st[i], one);
        const u64 w[3] = {v[2], v[1], v[0]};
        for (int k = 0; k < 3; k++) s[3 * i + k] = w[k] >= GL_P ? w[k] - GL_P : w[k];
    }
}

/* ---------------------------------------------------------------- sponge */

enum { HASH_POSEIDON = 0, HASH_KECCAK = 1, HASH_POSEIDON_BN128 = 2, HASH_POSEIDON2 = 3, HASH_MONOLITH = 4 };

static void permute(u64 *s, int hash_type)
{
    if (hash_type == HASH_POSEIDON2) orc_poseidon2_permute(s);
    else if (hash_type == HASH_MONOLITH) orc_monolith_permute(s);
    else if (hash_type == HASH_POSEIDON_BN128) orc_poseidon_bn128_permute(s);
    else orc_poseidon_permute(s);
}

void orc_hash_one(const u64 *in, u64 n, u64 *out, int hash_type)
{
    if (hash_type == HASH_KECCAK) {
        keccak_hash_one(in, n, out);
        return;
    }
    if (n <= 4) {
        for (u64 i = 0; i < 4; i++) out[i] = i < n ? in[i] : 0;
        return;
    }
    u64 s[12] = {0};
    for (u64 idx = 0; idx < n; idx += 8) {
        u64 len = n - idx < 8 ? n - idx : 8;
        memcpy(s, in + idx, len * sizeof(u64)); /* overwrite, no padding */
        permute(s, hash_type);
    }
    memcpy(out, s, 4 * sizeof(u64));
}

void orc_hash_two(const u64 *l, const u64 *r, u64 *out, int hash_type)
{
    if (hash_type == HASH_KECCAK) {
        keccak_hash_two(l, r, out);
        return;
    }
    u64 s[12] = {0};
    memcpy(s, l, 4 * sizeof(u64));
    memcpy(s + 4, r, 4 * sizeof(u64));
    permute(s, hash_type);
    memcpy(out, s, 4 * sizeof(u64));
}

/* ----------------------------------------------------------------- merkle */

/* Linear Merkle forest, merkle.cpp:19-124.  Returns 0, or -1 when the sizes
 * are inconsistent (the reference asserts). */
int orc_fill_digests(u64 *dig, u64 *cap, const u64 *leaves, u64 nd, u64 nc, u64 nl,
                     u64 ls, u64 h, int hash_type)
{
    if (nc == nl) {
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)nl; i++) orc_hash_one(leaves + i * ls, ls, cap + 4 * i, hash_type);
        return 0;
    }
    u64 sdl = nd >> h, sl = nl >> h;
    if (sdl == 0 || sl == 0 || nd / sdl != nc || nl / sl != nc) return -1;
    for (u64 k = 0; k < nc; k++) {
        const u64 *lv = leaves + k * sl * ls;
        u64 *dg = dig + k * sdl * 4;
        u64 *cp = cap + k * 4;
        if (sl == 1) {
            orc_hash_one(lv, ls, dg, hash_type);
            memcpy(cp, dg, 32);
            continue;
        }
        if (sl == 2) {
            orc_hash_one(lv, ls, dg, hash_type);
            orc_hash_one(lv + ls, ls, dg + 4, hash_type);
            orc_hash_two(dg, dg + 4, cp, hash_type);
            continue;
        }
        u64 *lh = dg + (sdl - sl) * 4;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)sl; i++) orc_hash_one(lv + i * ls, ls, lh + 4 * i, hash_type);
        /* node j (heap index j+2) has children 2j+2, 2j+3; level with 2^r
         * nodes starts at 2^r - 2 */
        for (u64 w = sl / 2; w >= 2; w /= 2) {
            u64 base = w - 2;
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < (long long)w; i++) {
                u64 j = base + i;
                orc_hash_two(dg + 4 * (2 * j + 2), dg + 4 * (2 * j + 3), dg + 4 * j, hash_type);
            }
        }
        orc_hash_two(dg, dg + 4, cp, hash_type);
    }
    return 0;
}

/* -------------------------------------------------------------------- ntt */

/* the low `bits` bits of x in reverse order (bits <= 32) */
static unsigned bitrev(u64 x, unsigned bits)
{
    uint32_t v = (uint32_t)x;
    v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
    v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
    v = ((v >> 4) & 0x0F0F0F0Fu) | ((v & 0x0F0F0F0Fu) << 4);
    v = ((v >> 8) & 0x00FF00FFu) | ((v & 0x00FF00FFu) << 8);
    v = (v >> 16) | (v << 16);
    return bits ? v >> (32 - bits) : 0u;
}

/* Field product with the Goldilocks identities 2^64 = 2^32 - 1 and
 * 2^96 = -1 (mod p) (ff/goldilocks.hpp:164-177 reduce128) for the NTT's
 * inner loop; equal to orc_gl_mul, only faster than a 128-bit division. */
static inline u64 ntt_mul(u64 a, u64 b)
{
    const u128 x = (u128)a * b;
    const u64 lo = (u64)x, hi = (u64)(x >> 64);
    const u64 h0 = hi & 0xFFFFFFFFull, h1 = hi >> 32;
    const u64 t = (lo - h1) - (((u64)0 - (u64)(lo < h1)) & 0xFFFFFFFFull); /* borrow: lo - h1 + p */
    const u64 u = h0 * 0xFFFFFFFFull;
    u64 r = t + u;
    r += ((u64)0 - (u64)(r < t)) & 0xFFFFFFFFull; /* carry: 2^64 = 2^32 - 1 */
    return r - (((u64)0 - (u64)(r >= GL_P)) & GL_P);
}

static inline u64 ntt_add(u64 a, u64 b) /* a, b < p */
{
    const u64 s = a + b;
    return s - (((u64)0 - (u64)((s < a) | (s >= GL_P))) & GL_P);
}

static inline u64 ntt_sub(u64 a, u64 b) /* a, b < p */
{
    return (a - b) + (((u64)0 - (u64)(a < b)) & GL_P);
}

/* Per-stage twiddles of a 2^lg transform: stage s (butterfly span 2^s) uses
 * w_{2^s}^k, k < 2^(s-1), at tw[2^(s-1) - 1 + k] (w or w^-1). */
static u64 *ntt_twiddles(unsigned lg, int inverse)
{
    u64 *tw = (u64 *)malloc(((1ull << lg) > 1 ? (1ull << lg) : 2) * sizeof(u64));
    for (unsigned s = 1; s <= lg; s++) {
        u64 half = 1ull << (s - 1), *t = tw + half - 1;
        u64 wm = orc_gl_omega(s);
        if (inverse) wm = orc_gl_inv(wm);
        t[0] = 1;
        for (u64 k = 1; k < half; k++) t[k] = ntt_mul(t[k - 1], wm);
    }
    return tw;
}

/* In-place natural-order -> natural-order DFT of size 2^lg over one column:
 * out[j] = sum_i in[i] * w^(i*j), w = OMEGA[lg] (or its inverse).
 * Bit-reverse + iterative radix-2 Cooley-Tukey; tw from ntt_twiddles. */
static void ntt_column(u64 *a, unsigned lg, const u64 *tw)
{
    u64 n = 1ull << lg;
    for (u64 i = 0; i < n; i++)
        if (a[i] >= GL_P) a[i] -= GL_P; /* ntt_add / ntt_sub take canonical operands */
    for (u64 i = 0; i < n; i++) {
        u64 j = bitrev(i, lg);
        if (j > i) { u64 t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (unsigned s = 1; s <= lg; s++) {
        u64 m = 1ull << s, half = m >> 1;
        const u64 *t = tw + half - 1;
        for (u64 base = 0; base < n; base += m)
            for (u64 k = 0; k < half; k++) {
                u64 u = a[base + k], v = ntt_mul(a[base + k + half], t[k]);
                a[base + k] = ntt_add(u, v);
                a[base + k + half] = ntt_sub(u, v);
            }
    }
}

/* compute_batched_ntt semantics (ntt.cu:253-309 / 106-158) on host memory:
 * forward: [x_i *= g^i if coset] then NTT, natural order;
 * inverse: iNTT, [x_i *= g^i if coset -- the reference multiplies by g^i, not
 * g^-i, ntt.cu:137-139], then x_i *= 2^-lg. */
void orc_ntt(u64 *data, unsigned lg, unsigned batches, int inverse, int coset, u64 gen)
{
    u64 n = 1ull << lg;
    u64 *tw = ntt_twiddles(lg, inverse);
#pragma omp parallel for schedule(dynamic)
    for (long long b = 0; b < (long long)batches; b++) {
        u64 *a = data + b * n;
        if (!inverse && coset) {
            u64 g = 1;
            for (u64 i = 0; i < n; i++) { a[i] = ntt_mul(a[i], g); g = ntt_mul(g, gen); }
        }
        ntt_column(a, lg, tw);
        if (inverse) {
            u64 ninv = orc_gl_inv(n % GL_P), g = 1;
            for (u64 i = 0; i < n; i++) {
                if (coset) { a[i] = ntt_mul(a[i], g); g = ntt_mul(g, gen); }
                a[i] = ntt_mul(a[i], ninv);
            }
        }
    }
    free(tw);
}

/* compute_batched_lde semantics (ntt.cu:477-564): per column zero-pad 2^lg_n
 * coefficients to 2^(lg_n+rate), then the batched NTT above on size
 * 2^(lg_n+rate).  Output column-major [batches][2^(lg_n+rate)]. */
void orc_lde(u64 *out, const u64 *in, unsigned lg_n, unsigned rate, unsigned batches,
             int inverse, int coset, u64 gen)
{
    u64 n = 1ull << lg_n, N = n << rate;
#pragma omp parallel for schedule(static)
    for (long long b = 0; b < (long long)batches; b++) {
        memcpy(out + b * N, in + b * n, n * sizeof(u64));
        memset(out + b * N + n, 0, (N - n) * sizeof(u64));
    }
    orc_ntt(out, lg_n + rate, batches, inverse, coset, gen);
}

/* compute_transpose_rev semantics (kernels.cu:94-135):
 * out[t * batches + b] = in[b * 2^lg_n + bitrev_lg_n(t)]. */
void orc_transpose_rev(u64 *out, const u64 *in, unsigned lg_n, unsigned batches)
{
    u64 n = 1ull << lg_n;
#pragma omp parallel for schedule(static)
    for (long long t = 0; t < (long long)n; t++) {
        u64 src = bitrev((u64)t, lg_n);
        for (u64 b = 0; b < batches; b++) out[t * batches + b] = in[b * n + src];
    }
}

/* ---------------------------------------------------- synthetic inputs */

/* splitmix64 stream rejection-sampled into [0, p): the canonical-element
 * generator the tests, the bench and the fixtures share. */
void orc_fill_canonical(u64 *out, u64 count, u64 seed)
{
    u64 x = seed;
    for (u64 i = 0; i < count;) {
        x += 0x9E3779B97F4A7C15ull;
        u64 z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        if (z < GL_P) out[i++] = z;
    }
}

/* Seeded salt / zkx_fill_random stream: element g is the first value below p
 * of the splitmix64 sequence (Steele, Lea and Flood 2014; the published
 * mixer with increment 0x9E3779B97F4A7C15) started at seed ^ (g *
 * 0xD1B54A32D192ED03).  The reference's salt is clock-seeded curand
 * (ntt/kernels.cu:62-84, ntt.cu:541-546) and reproducible by nothing, so
 * this seeded form is the library's own definition; the mixer is pinned by
 * splitmix64's published outputs. */
void orc_salt(u64 *out, u64 count, u64 seed)
{
    for (u64 g = 0; g < count; g++) {
        u64 x = seed ^ (g * 0xD1B54A32D192ED03ull), z;
        do {
            x += 0x9E3779B97F4A7C15ull;
            z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
        } while (z >= GL_P);
        out[g] = z;
    }
}

/* The raw splitmix64 outputs of `seed` (no rejection): pins the mixer
 * against the published sequence. */
void orc_splitmix64(u64 *out, u64 count, u64 seed)
{
    u64 x = seed;
    for (u64 i = 0; i < count; i++) {
        x += 0x9E3779B97F4A7C15ull;
        u64 z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        out[i] = z ^ (z >> 31);
    }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
