// NTT pass descriptors shared by the kernels (ntt_kernels.h) and the host
// drivers (ntt.hip, batch.hip).
#pragma once
#include <stdint.h>

namespace zkx {

constexpr int kNttTileLog = 12;  // 4096 u64 = 32 KB of LDS per work-group
constexpr int kNttThreads = 256;
constexpr int kNttPerThread = (1 << kNttTileLog) / kNttThreads;  // 16 elements per lane
constexpr int kNttRadixLog = 4;                                 // stages per register round
constexpr uint32_t kMaxRateBits = 5;                             // LDE blowup up to 32

enum SrcMode : uint32_t {
    kSrcInPlace = 0,  // tile source is the buffer itself
    kSrcCoset = 1,    // buffer itself, times g^i (forward coset NTT)
    kSrcLde = 2,      // coefficients of the un-extended columns (LDE first pass)
};

// One pass over one transform batch: stages s_lo .. s_lo+m-1 of the DIF on
// every block of 2^L in buf.  A launch covers a list of independent passes
// (of different jobs); work-group b runs tile b - tile_begin of the pass
// with the largest tile_begin <= b.
struct NttPass {
    uint64_t *buf;          // flat array of (total >> L) blocks of 2^L
    const uint64_t *src;    // kSrcLde: coefficients [cols][2^L]
    const uint64_t *tw;     // w_T^k, k < 2^lg_tw (w_T = OMEGA[lg_tw])
    const uint64_t *coset;  // kSrcCoset: g^i table; kSrcLde: factor table F (see ntt_kernels.h)
    const uint64_t *tw12;   // w_4096^k, k < 4096 (stage twiddles of contiguous passes)
    uint64_t total;         // elements in buf
    uint64_t tile_begin;    // first work-group of this pass in the launch
    uint32_t L;             // log2 block length
    uint32_t lg_tw;         // log2 twiddle table length (>= L + rate)
    uint32_t rate;          // kSrcLde: log2 blocks per column
    uint32_t s_lo, m, c;    // stages s_lo .. s_lo+m-1; 2^c columns per tile (m + c = 12)
    uint32_t mode;          // SrcMode of the tile load
    uint32_t inverse;       // use w^-1
    uint32_t with_coset;    // unused by the kernel (the factor table carries the coset)
    // device array: tw_lg[k][i] = w_{2^k}^i (strided stages: table of the
    // stage's butterfly span, so consecutive columns read consecutive words)
    const uint64_t *const *tw_lg;
};

}  // namespace zkx
