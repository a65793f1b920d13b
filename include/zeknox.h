/*
 * zeknox.h -- the C ABI of the MI355X (gfx950) LDE + Merkle-commit library.
 *
 * Drop-in for the reference's C interface: same symbol names, argument
 * types, by-value struct layouts and error convention, so the reference's
 * Rust (wrappers/rust/src/lib.rs:12-96), Go (wrappers/go/lib/lib.go:89-212)
 * and C++ (wrappers/cpp/test_gl64.cpp) callers link against libzeknox
 * unchanged.  Each declaration names the reference declaration it replaces.
 *
 * Field: Goldilocks p = 2^64 - 2^32 + 1; every buffer holds canonical u64
 * elements (< p), little-endian.
 */
#ifndef ZEKNOX_AMD_ZEKNOX_H
#define ZEKNOX_AMD_ZEKNOX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define ZK_EXTERN extern "C"
#define ZK_BOOL bool
#else
#define ZK_EXTERN
#define ZK_BOOL char
#endif

/* ---- native/utils/rusterror.h:15-41 ------------------------------------
 * Returned by value.  code 0 = success, 1 (hipErrorInvalidValue, same value
 * as cudaErrorInvalidValue) = bad arguments, -e = HIP runtime error e.
 * message is NULL or a malloc'd string the caller releases with free(). */
typedef struct RustError {
    int code;
    char *message;
} RustError;

/* ---- native/ntt/ntt.h:17-62 ------------------------------------------- */
typedef enum { NN, NR, RN, RR } NTT_InputOutputOrder;
typedef enum { forward, inverse } NTT_Direction;
typedef enum { standard, coset } NTT_Type;
typedef enum { GS, CT } NTT_Algorithm;

typedef struct {
    uint32_t batches;             /* columns */
    NTT_InputOutputOrder order;   /* ignored, as in the reference (always NN) */
    NTT_Type ntt_type;            /* ignored, as in the reference */
    uint32_t extension_rate_bits; /* LDE: log2 blowup, >= 1 */
    ZK_BOOL are_inputs_on_device;
    ZK_BOOL are_outputs_on_device;
    ZK_BOOL with_coset;
    ZK_BOOL is_multi_gpu;         /* ignored, as in the reference */
    uint32_t salt_size;           /* LDE: extra columns of random salt */
} NTT_Config;

typedef struct {
    uint32_t batches;
    ZK_BOOL are_inputs_on_device;
    ZK_BOOL are_outputs_on_device;
} NTT_TransposeConfig;

/* ---- native/merkle/merkle.h:21-28 ------------------------------------- */
enum HashType {
    HashPoseidon = 0,
    HashKeccak = 1,        /* Keccak-256, 25-byte digests */
    HashPoseidonBN128 = 2, /* Poseidon over BN254, t = 5 */
    HashPoseidon2 = 3,
    HashMonolith = 4
};

/* ---- native/lib.h:21-51, native/lib.cu:36-175 ------------------------- */

/* lib.h:21 -- number of usable HIP devices. */
ZK_EXTERN RustError get_number_of_gpus(size_t *ngpus);

/* lib.h:23 -- prints the properties of every device to stdout. */
ZK_EXTERN RustError list_devices_info(void);

/* lib.h:25 / lib.cu:168-175 -- init_cuda_degree(24). */
ZK_EXTERN void init_cuda(void);

/* lib.h:27 / lib.cu:157-166 -- coset table of 2^max_degree powers of the
 * generator 7 and twiddle tables for log sizes 2..max_degree on every device. */
ZK_EXTERN void init_cuda_degree(const uint32_t max_degree);

/* lib.cu:137-155 (exported, not in lib.h) -- as above with a chosen generator. */
ZK_EXTERN void init_cuda_degree_and_generator(const uint32_t max_degree, const uint64_t group_generator);

/* lib.h:29 / ntt.cu:160-173 -- twiddle table w^i, i < 2^lg_n, on device_id.
 * (Compute calls also build a missing table on first use.) */
ZK_EXTERN RustError init_twiddle_factors(size_t device_id, size_t lg_n);

/* lib.h:31 / ntt.cu:228-239 -- coset table coset_gen^i, i < 2^lg_domain_size. */
ZK_EXTERN RustError init_coset(size_t device_id, size_t lg_domain_size, const uint64_t coset_gen);

/* lib.h:33-35 / ntt.cu:253-309 -- in-place batched NTT of cfg.batches
 * columns of 2^lg_domain_size, natural order in and out.  forward: optional
 * coset scaling x_i *= g^i first; inverse: iNTT then optional x_i *= g^i,
 * then x_i *= 2^-lg (the reference's semantics, ntt.cu:128-143). */
ZK_EXTERN RustError compute_batched_ntt(size_t device_id, void *inout, uint32_t lg_domain_size,
                                        NTT_Direction ntt_direction, NTT_Config cfg);

/* lib.h:37-39 / ntt.cu:477-564 -- low-degree extension: each of cfg.batches
 * columns of 2^lg_domain_size coefficients is zero-padded to
 * 2^(lg_domain_size + extension_rate_bits) and transformed as above (coset
 * g^i when with_coset); output column-major, natural order, followed by
 * salt_size salt columns. */
ZK_EXTERN RustError compute_batched_lde(size_t device_id, void *output, void *input, uint32_t lg_domain_size,
                                        NTT_Direction ntt_direction, NTT_Config cfg);

/* lib.h:41-46 / ntt.cu:315-471 -- multi-device LDE.  This build is
 * single-device: num_gpu must be 1 (device 0), otherwise code 1. */
ZK_EXTERN RustError compute_batched_lde_multi_gpu(void *output, void *input, uint32_t num_gpu,
                                                  NTT_Direction ntt_direction, NTT_Config cfg,
                                                  uint32_t lg_domain_size, size_t total_num_input_elements,
                                                  size_t total_num_output_elements);

/* lib.h:48-49 / ntt.cu:182-226 -- out[t * batches + b] = in[b * 2^lg_n + bitrev(t)]. */
ZK_EXTERN RustError compute_transpose_rev(size_t device_id, void *output, void *input, uint32_t lg_n,
                                          NTT_TransposeConfig cfg);

/* ---- native/merkle/merkle.h:30-62 -------------------------------------
 * Sizes count digests, caps and leaves (not bytes); leaf_size counts u64.
 * The GPU entry points take device pointers, return after the work is
 * complete, and on bad arguments leave the buffers untouched and record the
 * reason (zkx_last_error in zeknox_ext.h). */

/* merkle.h:30-40 / merkle.cu:246-280 */
ZK_EXTERN void fill_digests_buf_linear_gpu_with_gpu_ptr(void *digests_buf_gpu_ptr, void *cap_buf_gpu_ptr,
                                                       void *leaves_buf_gpu_ptr, uint64_t digests_buf_size,
                                                       uint64_t cap_buf_size, uint64_t leaves_buf_size,
                                                       uint64_t leaf_size, uint64_t cap_height,
                                                       uint64_t hash_type, uint64_t gpu_id);

/* merkle.h:42-51 / merkle.cu:283-459 -- single-device build: runs on device 0. */
ZK_EXTERN void fill_digests_buf_linear_multigpu_with_gpu_ptr(void *digests_buf_gpu_ptr, void *cap_buf_gpu_ptr,
                                                            void *leaves_buf_gpu_ptr, uint64_t digests_buf_size,
                                                            uint64_t cap_buf_size, uint64_t leaves_buf_size,
                                                            uint64_t leaf_size, uint64_t cap_height,
                                                            uint64_t hash_type);

/* merkle.h:53-62 / merkle.cpp:126-161 -- the same tree computed by the
 * library's GPU kernels from host buffers (copied in and out); the name is
 * kept so CPU callers link unchanged. */
ZK_EXTERN void fill_digests_buf_linear_cpu(void *digests_buf_ptr, void *cap_buf_ptr, const void *leaves_buf_ptr,
                                          uint64_t digests_buf_size, uint64_t cap_buf_size,
                                          uint64_t leaves_buf_size, uint64_t leaf_size, uint64_t cap_height,
                                          uint64_t hash_type);

/* all_gpus.cu:68-71 -- true when at least one HIP device is present. */
ZK_EXTERN ZK_BOOL cuda_available(void);

#undef ZK_BOOL
#endif /* ZEKNOX_AMD_ZEKNOX_H */
