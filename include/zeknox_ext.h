/*
 * zeknox_ext.h -- additions of the MI355X build that have no counterpart in
 * the reference's C ABI.  Callers of the reference interface need none of
 * them; they serve the test harness, benchmarks and batch callers.
 */
#ifndef ZEKNOX_AMD_ZEKNOX_EXT_H
#define ZEKNOX_AMD_ZEKNOX_EXT_H

#include <stddef.h>
#include <stdint.h>

#include "zeknox.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reason of the last failed void entry point on this thread ("" if none).
 * The reference's Merkle entry points assert instead (merkle.cu:258). */
const char *zkx_last_error(void);

/* Device memory helpers (the reference's wrappers bind cudart for this:
 * wrappers/rust/src/device/memory.rs, wrappers/go/device/memory.go). */
RustError zkx_device_malloc(size_t device_id, void **ptr, size_t bytes);
RustError zkx_device_free(size_t device_id, void *ptr);
RustError zkx_memcpy_htod(size_t device_id, void *dst, const void *src, size_t bytes);
RustError zkx_memcpy_dtoh(size_t device_id, void *dst, const void *src, size_t bytes);
RustError zkx_memset(size_t device_id, void *dst, int value, size_t bytes);
RustError zkx_device_synchronize(size_t device_id);

/* The stream the library enqueues its work on (a hipStream_t). */
void *zkx_stream(size_t device_id);

/* Kernel timing: when enabled, every launch of the library's kernels on a
 * device is bracketed by HIP events on the launching stream; zkx_kernel_stats
 * reports, per kernel family, the launch count and summed duration. */
RustError zkx_profile_enable(int on);
/* Writes up to `cap` entries; returns the number of families. */
int zkx_kernel_stats(char names[][32], uint64_t *launches, double *total_ms, int cap);
void zkx_kernel_stats_reset(void);

#ifdef __cplusplus
}
#endif

#endif /* ZEKNOX_AMD_ZEKNOX_EXT_H */
