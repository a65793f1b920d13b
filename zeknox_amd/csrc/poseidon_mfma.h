// Poseidon (Goldilocks, width 12; poseidon/poseidon.cu:563-838) with the MDS
// layer on the matrix cores, for leaf hashing by whole waves.
//
// The MDS is a 12 x 12 matrix with small coefficients applied to every
// permutation of a wave: a batch of matrix-vector products.  Its integer
// form is exact on the i8 MFMA once the 64-bit elements are cut into their
// 8 bytes: with u_{C,b} the b-th byte of element C,
//   y_R = sum_C CIRC[(C - R) mod 12] x_C = sum_b 2^(8b) D_b[R],
//   D_b[R] = sum_C CIRC[(C - R) mod 12] u_{C,b}  (< 2^16),
// so one MFMA output row per (output element, byte) and a block-diagonal
// coefficient matrix (byte b of an input pairs only with row byte b).  The
// vector unit then only recombines eight partials per element, instead of
// 24 multiply-adds per element.
//
// Wave layout.  Two lanes hold one permutation, six elements each; a wave
// holds two groups t of 32 permutations.  Lane l (column j = l & 31, half
// h = l >> 5) holds, for group t, the elements
//   E(t, h, e) = 6 ((h + t) mod 2) + e,  e = 0..5,
// of permutation (t, j).  v_mfma_i32_32x32x32_i8 (gfx950, layout measured)
// computes D[i][j] = sum_{h, q} A_{32h+i}[q] * B_{32h+j}[q] over the 16
// bytes q each lane supplies, and lane 32h' + j receives D[i][j] for
// i = 8k + 4h' + r in its register 4k + r.  So lane (j, h) supplies slots
// 2kappa, 2kappa + 1 as B of K-chunk kappa (0..2, chained in place), and
// M-block beta (0..2) row 8k + 4h' + r is output element
// E(t, h', 2 beta + k / 2), byte 4 (k % 2) + r: every lane receives the
// eight byte partials of exactly its own six output elements.  The
// coefficient of that row and input slot e of half h is
//   CIRC[(6 (h - h') + e - 2 beta - k / 2) mod 12]
// for both groups (the rotation by 6t cancels in a circulant), so one set of
// A operands (9 x 4 VGPRs) serves both.  Element 0 of group t sits in half
// t: in a partial round each lane applies exactly one S-box.  MDS_DIAG (8 on
// element 0) is added on the vector unit by the lanes that hold element 0.
// (Measured on gfx950 beside full-rate vector work: 18 of these MFMAs per
// round cost about half the issue time of 48 16x16x64 ones doing the same
// products.  Accumulator chains stay within one opcode, in place: a
// 16x16x64 result taken as the accumulator of a 16x16x32 product, issued
// with no wait states between them, came out wrong.)
//
// Signed bytes.  The MFMA takes signed i8: the state enters XORed with 0x80
// in every byte (u - 128), so D_b = (byte-plane sum) - 128 * 256, in
// [-32768, 32512]; the 32768 per byte plane goes back in with the folded
// round constants (below), which keeps every partial sum non-negative.
//
// Recombination (per output element, lazy in and out, as hash.h):
//   L = KL + (D0 + 2^8 D1) + 2^16 (D2 + 2^8 D3),  H = KH + (same on D4..D7),
// with KL, KH = the low and high 32 bits of the next round's constant plus
// 32768 * (1 + 2^8 + 2^16 + 2^24) each, then limb_combine(L, H)
// (0 <= L, H < 2^42).  The chains run in signed 32x32+64 multiply-adds.
#pragma once
#include "hash.h"

namespace zkx {

typedef int mf_v4i __attribute__((ext_vector_type(4)));
typedef int mf_v16i __attribute__((ext_vector_type(16)));

static __constant__ uint32_t c_poseidon_circ[12] = {ZK_POSEIDON_MDS_CIRC_LIST};

// Round-constant table in LDS: row r (0..30), element E -> {lo32, hi32} of
// c_poseidon_rc[r][E], each widened to u64 (a multiply-add addend pair).
constexpr int kMfmaRcWords = 31 * 12 * 2;

// 32768 per byte plane of one 32-bit half: 32768 * (1 + 2^8 + 2^16 + 2^24)
constexpr uint64_t kMfmaPlaneBias = 32768ull * 0x01010101ull;

// Row 0 holds the first round's constants as they are (added before the
// first S-box); rows 1..30 also carry kMfmaPlaneBias in each half.
__device__ __forceinline__ void mfma_fill_rc(uint64_t *kt, uint32_t t, uint32_t nthreads)
{
    for (uint32_t i = t; i < 31 * 12; i += nthreads) {
        const uint64_t c = c_poseidon_rc[i], b = i < 12 ? 0 : kMfmaPlaneBias;
        kt[2 * i] = gl::lo32(c) + b;
        kt[2 * i + 1] = gl::hi32(c) + b;
    }
}

struct MfmaOps {
    mf_v4i a[3][3];  // coefficient operands [M-block beta][K-chunk kappa]
    uint32_t h;      // lane half
};

__device__ __forceinline__ MfmaOps mfma_setup()
{
    MfmaOps o;
    const uint32_t l = threadIdx.x & 63, i = l & 31, h = l >> 5, k = i >> 3, hp = (i >> 2) & 1, r = i & 3;
    const uint32_t sh = 8 * (4 * (k & 1) + r);  // bit offset of the row's byte within a slot
    o.h = h;
#pragma unroll
    for (int beta = 0; beta < 3; beta++)
#pragma unroll
        for (int kappa = 0; kappa < 3; kappa++) {
            uint64_t w[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const uint32_t c = (6 * (h + 2 - hp) + 2 * kappa + u + 12 - 2 * beta - (k >> 1)) % 12;
                w[u] = (uint64_t)c_poseidon_circ[c] << sh;
            }
            o.a[beta][kappa] = mf_v4i{(int)gl::lo32(w[0]), (int)gl::hi32(w[0]), (int)gl::lo32(w[1]), (int)gl::hi32(w[1])};
        }
    return o;
}

// Lanes of half t hold element 0 of group t.
__device__ __forceinline__ bool mfma_own(uint32_t h, int t) { return h == (uint32_t)t; }

// s <- MDS * s + RC[row] (krow: the next round's row of the LDS table).
__device__ __forceinline__ void mds_mfma(uint64_t s[2][6], const MfmaOps &o, const uint64_t *krow)
{
    constexpr uint64_t XM = 0x8080808080808080ull;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        mf_v4i b[3];
#pragma unroll
        for (int kappa = 0; kappa < 3; kappa++) {
            const uint64_t x0 = s[t][2 * kappa] ^ XM, x1 = s[t][2 * kappa + 1] ^ XM;
            b[kappa] = mf_v4i{(int)gl::lo32(x0), (int)gl::hi32(x0), (int)gl::lo32(x1), (int)gl::hi32(x1)};
        }
        const bool own = mfma_own(o.h, t);
        const uint32_t dl = own ? gl::lo32(s[t][0]) : 0u, dh = own ? gl::hi32(s[t][0]) : 0u;
        const uint64_t *k = krow + 12 * ((o.h + t) & 1);
        // one M-block at a time, recombined as soon as its chain is done: 16
        // accumulator VGPRs live instead of 48 (the operands b and dl, dh
        // hold what the later blocks need of the old state)
#pragma unroll
        for (int beta = 0; beta < 3; beta++) {
            mf_v16i dd = __builtin_amdgcn_mfma_i32_32x32x32_i8(o.a[beta][0], b[0], mf_v16i{}, 0, 0, 0);
            dd = __builtin_amdgcn_mfma_i32_32x32x32_i8(o.a[beta][1], b[1], dd, 0, 0, 0);
            dd = __builtin_amdgcn_mfma_i32_32x32x32_i8(o.a[beta][2], b[2], dd, 0, 0, 0);
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int e = 2 * beta + u;
                const int q = 8 * u;  // registers q..q+3: bytes 0-3, q+4..q+7: bytes 4-7
                const int32_t p01 = dd[q] + (int32_t)((uint32_t)dd[q + 1] << 8);
                const int32_t p23 = dd[q + 2] + (int32_t)((uint32_t)dd[q + 3] << 8);
                const int32_t q01 = dd[q + 4] + (int32_t)((uint32_t)dd[q + 5] << 8);
                const int32_t q23 = dd[q + 6] + (int32_t)((uint32_t)dd[q + 7] << 8);
                const int64_t one = (int32_t)gl::opaque(1u), m16 = (int32_t)gl::opaque(65536u);
                uint64_t L = gl::vopaque64((uint64_t)((int64_t)p01 * one + (int64_t)k[2 * e]));
                uint64_t H = gl::vopaque64((uint64_t)((int64_t)q01 * one + (int64_t)k[2 * e + 1]));
                if (e == 0) {
                    L = gl::vopaque64((uint64_t)dl * gl::opaque(8u) + L);
                    H = gl::vopaque64((uint64_t)dh * gl::opaque(8u) + H);
                }
                L = gl::vopaque64((uint64_t)((int64_t)p23 * m16 + (int64_t)L));
                H = gl::vopaque64((uint64_t)((int64_t)q23 * m16 + (int64_t)H));
                s[t][e] = limb_combine(L, H);
            }
        }
    }
}

// The permutation of 64 states (layout above); kt: the LDS round-constant
// table.  Lazy in and out.  All 64 lanes of the wave must be active.
__device__ __forceinline__ void poseidon_permute_mfma(uint64_t s[2][6], const MfmaOps &o, const uint64_t *kt)
{
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const uint64_t *k = kt + 12 * ((o.h + t) & 1);
#pragma unroll
        for (int e = 0; e < 6; e++) s[t][e] = gl::add(s[t][e], gl::pack(gl::lo32(k[2 * e]), gl::lo32(k[2 * e + 1])));
    }
    // rounds 0-3 and 26-29 full, 4-25 partial (two per iteration)
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int e = 0; e < 6; e += 2) gl::pow7_lazy2(s[t][e], s[t][e + 1]);
        mds_mfma(s, o, kt + 24 * (r + 1));
    }
#pragma unroll 1
    for (int r = 4; r < 26; r += 2) {
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
            // each lane holds element 0 of exactly one group
            uint64_t x = mfma_own(o.h, 1) ? s[1][0] : s[0][0];
            x = gl::pow7_lazy(x);
            s[0][0] = mfma_own(o.h, 0) ? x : s[0][0];
            s[1][0] = mfma_own(o.h, 1) ? x : s[1][0];
            mds_mfma(s, o, kt + 24 * (r + hh + 1));
        }
    }
#pragma unroll 1
    for (int r = 26; r < 30; r++) {
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int e = 0; e < 6; e += 2) gl::pow7_lazy2(s[t][e], s[t][e + 1]);
        mds_mfma(s, o, kt + 24 * (r + 1));
    }
}

}  // namespace zkx
