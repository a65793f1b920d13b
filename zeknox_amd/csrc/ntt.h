// Host-side NTT / LDE / transpose drivers (device pointers, enqueued on the
// device's stream; the caller synchronises).
#pragma once
#include <vector>

#include "ntt_types.h"
#include "runtime.h"
#include "tuning.h"

namespace zkx {

// LDE of `cols` columns of 2^lg_n coefficients into `out` (cols x 2^(lg_n+rate)),
// each output column in bit-reversed order (DIF order, no reversal pass).
void lde_bitrev(Device &d, uint64_t *out, const uint64_t *in, uint32_t lg_n, uint32_t rate, uint32_t cols,
                bool inverse, bool with_coset);
// Resident LDE factor table F[b*n + i] = (g * w_N^{bitrev_r(b)})^i (g = 1
// without coset; w^-1 for the inverse), built on first use.
const uint64_t *lde_factor_table(Device &d, uint32_t lg_n, uint32_t rate, bool inverse, bool with_coset);
// compute_batched_lde semantics: natural-order output.
void lde(Device &d, uint64_t *out, const uint64_t *in, uint32_t lg_n, uint32_t rate, uint32_t cols,
         bool inverse, bool with_coset);
// compute_batched_ntt semantics: in place, natural order.
void ntt(Device &d, uint64_t *inout, uint32_t lg, uint32_t cols, bool inverse, bool with_coset);
// In-place bit reversal of each column with optional post-scaling.
void bitrev_columns(Device &d, uint64_t *buf, uint32_t lg, uint32_t cols, bool use_scale, uint64_t scale,
                    const uint64_t *coset);
// out[t * cols + b] = in[b * 2^lg + (rev ? bitrev(t) : t)].
void transpose_rev(Device &d, uint64_t *out, const uint64_t *in, uint32_t lg, uint32_t cols, bool rev);

// Pass descriptors of the DIF over nblk blocks of 2^L in buf (in launch
// order; each pass must finish before the next starts), cut into passes as
// `tun` says.  Returns the tile count of every pass.
// tw_lg: the device array of twiddle tables by log size, built up to L
// (Device::twiddle_tables_upto).
uint64_t plan_dif(std::vector<NttPass> &out, uint64_t *buf, uint64_t nblk, uint32_t L, uint32_t lg_tw,
                  const uint64_t *tw, const uint64_t *tw12, const uint64_t *const *tw_lg, bool inverse,
                  uint32_t mode0, const uint64_t *src, uint32_t rate, bool with_coset, const uint64_t *coset,
                  const Tuning &tun);

// Stage counts of the DIF passes over blocks of 2^L under `tun`, top
// stages first (the last pass contiguous).
std::vector<uint32_t> ntt_pass_sizes(uint32_t L, const Tuning &tun);

// Which register-round kernel a launch over these passes can use (each runs
// one tile per work-group): 0 none (k_ntt_tiles); 1 contiguous (every pass
// in place and a whole number of tiles; stages above 8 in registers);
// 2 + 4 * lde + (m - 1) strided single-round (every pass strided with the
// same m <= 4 stages and the same source, in place or LDE; registers only);
// 10 + 4 * lde + (m - 5) strided two-round (the same with 5 <= m <= 8, the
// first round in registers).
int ntt_kernel_kind(const NttPass *host_list, size_t n);

// One launch over n pass descriptors in device memory (sorted by
// tile_begin), all contiguous (s_lo == 0) or all strided; `kernel_kind` as
// ntt_kernel_kind of the same list.  chunk > 0: launches over ranges of at
// most `chunk` tiles instead, in order (other streams' kernels can be
// dispatched between them).
void launch_ntt_list(hipStream_t stream, const NttPass *dev_list, uint32_t n, uint64_t blocks, bool contiguous,
                     int kernel_kind, const char *name, uint64_t chunk = 0);

// table[i] = base^i, i < n.
void fill_powers(uint64_t *dst, uint64_t n, uint64_t base, hipStream_t s);
// `count` pseudo-random field elements from `seed` (LDE salt columns).
void fill_salt(uint64_t *dst, uint64_t count, uint64_t seed, hipStream_t s);

}  // namespace zkx
