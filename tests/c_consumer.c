/*
 * A C consumer of the reference's C ABI, calling it as the reference's Go
 * wrapper does (wrappers/go/lib/lib.go:65-103 InitTwiddleFactors, InitCoset,
 * LDEBatch; lib.go:198-210 FillDigestsBufLinearCPU): only include/zeknox.h,
 * host buffers and the library itself (-lzeknox).  Builds the coset
 * LDE of COLS polynomials of 2^LG_N coefficients (rate RATE), its
 * bit-reversed transpose into leaf rows and the Poseidon Merkle caps, and
 * checks the error convention (code 1 + a malloc'd message on bad
 * arguments).  Prints "leaves <sum>" and "cap i <4 words>" for the test.
 *
 *   c_consumer <seed>
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "zeknox.h"

#define COLS 7u
#define LG_N 9u
#define RATE 2u
#define CAP_H 2u

static uint64_t splitmix(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int main(int argc, char **argv)
{
    const uint64_t P = 0xFFFFFFFF00000001ull;
    uint64_t seed = argc > 1 ? strtoull(argv[1], 0, 10) : 1;
    const uint32_t lg = LG_N + RATE;
    const size_t n = (size_t)1 << LG_N, N = (size_t)1 << lg;
    uint64_t *coeffs = malloc(COLS * n * sizeof(uint64_t));
    uint64_t *lde = malloc(COLS * N * sizeof(uint64_t));
    uint64_t *leaves = malloc(COLS * N * sizeof(uint64_t));
    const size_t nc = (size_t)1 << CAP_H, nd = 2 * (N - nc);
    uint64_t *digests = malloc(nd * 4 * sizeof(uint64_t));
    uint64_t *caps = malloc(nc * 4 * sizeof(uint64_t));
    if (!coeffs || !lde || !leaves || !digests || !caps) return 2;
    for (size_t i = 0; i < COLS * n; i++) {
        uint64_t v;
        do v = splitmix(&seed); while (v >= P);
        coeffs[i] = v;
    }

    size_t ngpus = 0;
    RustError e = get_number_of_gpus(&ngpus);
    if (e.code || ngpus == 0) return 3;
    e = init_twiddle_factors(0, lg);
    if (e.code) return 4;
    e = init_coset(0, lg, 7);
    if (e.code) return 4;

    NTT_Config cfg = {0};
    cfg.batches = COLS;
    cfg.extension_rate_bits = RATE;
    cfg.with_coset = 1;
    e = compute_batched_lde(0, lde, coeffs, LG_N, forward, cfg);
    if (e.code) return 5;
    NTT_TransposeConfig tcfg = {0};
    tcfg.batches = COLS;
    e = compute_transpose_rev(0, leaves, lde, lg, tcfg);
    if (e.code) return 6;
    fill_digests_buf_linear_cpu(digests, caps, leaves, nd, nc, N, COLS, CAP_H, HashPoseidon);

    /* bad arguments: code 1 and a message the caller frees */
    cfg.extension_rate_bits = 40;
    e = compute_batched_lde(0, lde, coeffs, LG_N, forward, cfg);
    if (e.code != 1 || !e.message) return 7;
    free(e.message);

    uint64_t sum = 0;
    for (size_t i = 0; i < COLS * N; i++) sum += leaves[i] * (i + 1);
    printf("leaves %llu\n", (unsigned long long)sum);
    for (size_t i = 0; i < nc; i++)
        printf("cap %zu %llu %llu %llu %llu\n", i, (unsigned long long)caps[4 * i], (unsigned long long)caps[4 * i + 1],
               (unsigned long long)caps[4 * i + 2], (unsigned long long)caps[4 * i + 3]);
    free(coeffs);
    free(lde);
    free(leaves);
    free(digests);
    free(caps);
    return 0;
}
