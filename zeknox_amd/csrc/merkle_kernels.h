// Linear Merkle forest kernels (merkle/merkle.cu:165-244 layout, built
// MI355X-first).
//
// Layout (merkle/merkle.h, native/README.md:64-71): L leaves, 2^h caps,
// 2(L - 2^h) digests of 4 u64.  Subtree s (sl = L >> h leaves) owns digests
// [s*sdl, (s+1)*sdl), sdl = 2*sl - 2; inside it digest i is heap node i+2,
// so node i has children 2i+2 and 2i+3, the level of width w sits at
// [w-2, 2w-2), leaf hashes at [sl-2, 2sl-2), cap[s] = H(d_s[0], d_s[1]).
//
// Schedule.  A work-group of TPB (256..1024) lanes owns a tile of consecutive
// leaves.  Each lane hashes its leaf, then the tile is reduced level by level
// in LDS (Poseidon and Poseidon2 on lane pairs, two lanes per node, since a
// level is one permutation deep and runs at low occupancy): every level's
// digests are written to HBM on the way up, and a tile
// that reaches the width-2 top level of its subtree also writes the cap.  A
// tile that stops below the top leaves its root for an upper pass (a tile of
// such roots), and so on until the caps are written (no upper pass when a
// subtree has <= TPB leaves, one up to TPB^2, ...).
//
// Leaf sources.  Element k of leaf t is read at t*leaf_stride + k*col_stride:
// row-major leaves (the C ABI) use (leaf_size, 1); the commit engine hashes
// the LDE's bit-reversed column-major output directly with (1, N), so lanes
// of a wave read consecutive words of each column, and optionally writes the
// leaf rows (the transpose) on the way.
//
// Upper passes in the same launch.  A tile that stops at its root below the
// subtree top counts itself in an arrival counter of its group (the 2^tl
// tiles whose roots form one tile of the next level); the work-group that
// arrives last, after the others' roots are visible, loads the group's roots
// and reduces that tile too, and so on up to the caps.  Nothing waits: every
// work-group either continues or exits.  The last arrival resets its counter,
// so counters are zero between launches.
//
// A launch covers a list of independent jobs (trees or tree passes);
// work-group b runs tile b - tile_begin of the job with the largest
// tile_begin <= b, b = blockIdx.x + block0: a list may be issued as several
// launches over consecutive ranges of b (nothing waits across work-groups,
// and the arrival counters work across launches).
#pragma once
#include "gmem.h"
#include "hash.h"
#include "merkle_types.h"
#include "poseidon_mfma.h"

namespace zkx {

// Occupancy hint per hash type: four waves per SIMD (128 VGPRs) for Poseidon,
// whose matrix-core leaf path keeps 36 VGPRs of coefficient operands and 48
// of MFMA accumulators (measured on the 4 x 135 x 2^18 leaf launch: the
// operands read from LDS inside the round loop, at five or six waves, 9%
// slower); six waves (80 VGPRs) for Poseidon2 (8 VGPRs spilled outside the
// round loop, and as fast as five waves without the spill); the others would
// spill more, and keep the compiler's choice.
constexpr int merkle_waves_per_eu(int ht) { return ht == kPoseidon ? 4 : ht == kPoseidon2 ? 6 : 1; }

// Digest slot of node `idx` of the level of width 2^lg_w in subtree `sub`.
__device__ __forceinline__ gptr<uint64_t> node_ptr(const MerkleArgs &a, uint64_t sub, uint32_t lg_w, uint64_t idx)
{
    return gmem(a.dig) + 4 * (sub * a.sdl + ((1ull << lg_w) - 2) + idx);
}

__device__ __forceinline__ const MerkleArgs *find_job(const MerkleArgs *jobs, uint32_t njobs, uint32_t bid)
{
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].tile_begin <= bid) lo = mid;
        else hi = mid;
    }
    return jobs + lo;
}

// Poseidon leaf hashing by whole waves with the MDS on the matrix cores
// (poseidon_mfma.h): wave w hashes leaves first + 64w + [0, 64), leaf
// 32 t + j of the wave in group t, column j.  Needs whole waves (cur >= 64)
// and leaf_size > 4; leaves are absorbed as in hash_one.  The digests land in
// lds[leaf - first]; kt receives the round-constant table.  Every lane of the
// work-group calls it.
template <int TPB>
__device__ __forceinline__ void leaf_hash_mfma(const MerkleArgs &a, uint64_t first, uint32_t cur, uint64_t (*lds)[4],
                                               uint64_t *kt)
{
    const uint32_t t = threadIdx.x, w = t >> 6, j = t & 31, h = (t >> 5) & 1;
    mfma_fill_rc(kt, t, TPB);
    __syncthreads();
    if (64 * w < cur) {
        const MfmaOps o = mfma_setup();
        const uint32_t ls = a.leaf_size;
        const uint64_t cs = a.col_stride;
        uint64_t s[2][6];
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int e = 0; e < 6; e++) s[q][e] = 0;
        for (uint32_t idx = 0; idx < ls; idx += 8) {
            // addresses recomputed per chunk from an opaque chunk offset:
            // running pointers (24 VGPRs) would otherwise live through the
            // permutation and spill
            uint32_t kb = idx;
            asm volatile("; chunk %0" : "+s"(kb));
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const uint64_t leaf = first + 64 * w + 32 * q + j;
                const uint32_t el0 = 6 * ((h + q) & 1);  // this lane's first element of group q
                const gptr<const uint64_t> row = gmem(a.leaves) + leaf * a.leaf_stride;
#pragma unroll
                for (int e = 0; e < 6; e++) {
                    const uint32_t k = kb + el0 + e;
                    if (el0 + e < 8 && k < ls) {
                        const uint64_t v = row[(uint64_t)k * cs];
                        if (a.leaves_out) gmem(a.leaves_out)[leaf * ls + k] = v;
                        s[q][e] = v;
                    }
                }
            }
            poseidon_permute_mfma(s, o, kt);
        }
        // elements 0-3 of group q sit in slots 0-3 of half q
#pragma unroll
        for (int q = 0; q < 2; q++)
            if (h == (uint32_t)q) {
                uint64_t *d = lds[64 * w + 32 * q + j];
#pragma unroll
                for (int e = 0; e < 4; e++) d[e] = gl::canon(s[q][e]);
            }
    }
    __syncthreads();
}

template <int HT, int TPB>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(merkle_waves_per_eu(HT))))
k_merkle_tile(const MerkleArgs *__restrict__ jobs, uint32_t njobs, MerkleArgs one, uint32_t block0)
{
    __shared__ uint64_t lds[TPB][4];
    __shared__ uint32_t lds_last;
    __shared__ uint64_t kt[HT == kPoseidon ? kMfmaRcWords : 1];
    const uint32_t bid = blockIdx.x + block0;  // this work-group's tile in the whole launch list
    const MerkleArgs *job = njobs ? find_job(jobs, njobs, bid) : nullptr;
    const MerkleArgs a = njobs ? *job : one;
    uint32_t up = 0;  // upper passes done
    const uint32_t t = threadIdx.x;
    uint32_t cur = 1u << a.tile_log;  // nodes of this tile at the current level
    uint32_t lg_w = a.lg_w;           // log2 per-subtree width of the current level
    uint64_t first = ((uint64_t)bid - a.tile_begin) << a.tile_log;  // first node of the tile

    uint64_t h[4];
    const bool mfma = HT == kPoseidon && cur >= 64 && a.leaf_size > 4 && !a.leaf_valu;
    if constexpr (HT == kPoseidon)
        if (mfma) leaf_hash_mfma<TPB>(a, first, cur, lds, kt);
    if (t < cur) {
        const uint64_t g = first + t;
        const uint64_t sub = g >> lg_w, idx = g & ((1ull << lg_w) - 1);
        if (mfma) {
#pragma unroll
            for (int i = 0; i < 4; i++) h[i] = lds[t][i];
        } else {
            const gptr<const uint64_t> row = gmem(a.leaves) + g * a.leaf_stride;
            const uint64_t cs = a.col_stride;
            const gptr<uint64_t> out = a.leaves_out ? gmem(a.leaves_out) + g * a.leaf_size : nullptr;
            hash_one<HT>(a.leaf_size,
                         [&](uint32_t k) {
                             const uint64_t v = row[(uint64_t)k * cs];
                             if (out) out[k] = v;
                             return v;
                         },
                         h);
#pragma unroll
            for (int i = 0; i < 4; i++) lds[t][i] = h[i];
        }
        const gptr<uint64_t> dst = lg_w ? node_ptr(a, sub, lg_w, idx) : gmem(a.cap) + 4 * g;
#pragma unroll
        for (int i = 0; i < 4; i++) dst[i] = h[i];
    }
    if (lg_w == 0 || a.leaves_only) return;  // caps == leaves, or levels launched apart
    __syncthreads();

    while (true) {
        if (cur == 1) {
            // tile root (at any level, the width-2 one included): continue
            // with the group's tile if this work-group is the last of its
            // group to get here
            // The upper-pass fields are read here, from the descriptor list
            // (fused launches always have one), rather than kept in
            // registers through the hashing.
            if (!job) return;
            const volatile MerkleArgs *vj = job;
            if (up >= vj->n_next) return;
            const uint32_t tl = vj->next_tile_log[up];
            const uint32_t c0 = vj->next_ctr[up];
            uint32_t *const ctr = vj->ctr;
            const uint64_t grp = first >> tl;
            __threadfence();  // this tile's digests, before the arrival
            __syncthreads();
            if (t == 0) {
                uint32_t *c = ctr + c0 + grp;
                const uint32_t arrived = atomicAdd(c, 1u);
                const bool last = arrived == (1u << tl) - 1;
                if (last) atomicExch(c, 0u);
                lds_last = last ? 1u : 0u;
            }
            __syncthreads();
            if (!lds_last) return;
            __threadfence();  // the other tiles' roots, after their arrivals
            up++;
            cur = 1u << tl;
            first = grp << tl;
            if (t < cur) {
                const uint64_t g = first + t;
                const gptr<uint64_t> src = node_ptr(a, g >> lg_w, lg_w, g & ((1ull << lg_w) - 1));
#pragma unroll
                for (int i = 0; i < 4; i++) lds[t][i] = src[i];
            }
            __syncthreads();
            continue;
        }
        // Nodes of a level: with PAIR (Poseidon, Poseidon2) lanes 2n and 2n+1
        // hash node n together (a level never has more than TPB / 2 nodes)
        // and lane 2n writes the digest; a Poseidon level of at most TPB / 16
        // nodes runs on rows (ROW: lanes 16n .. 16n + 15, lane 16n + k
        // writes digest element k < 4); otherwise lane n hashes node n.
        constexpr bool PAIR = HT == kPoseidon || HT == kPoseidon2;
        const bool ROW = HT == kPoseidon && (cur >> 1) * 16 <= TPB;
        const uint32_t k = t & 15;
        const uint32_t n = ROW ? t >> 4 : PAIR ? t >> 1 : t;
        const bool writer = ROW ? k < 4 : PAIR ? (t & 1) == 0 : true;
        auto hash_node = [&](uint64_t out[4]) {
            if constexpr (HT == kPoseidon) poseidon_hash_two_pair(lds[2 * n], lds[2 * n + 1], out, (t & 1) != 0);
            else if constexpr (HT == kPoseidon2) poseidon2_hash_two_pair(lds[2 * n], lds[2 * n + 1], out, (t & 1) != 0);
            else hash_two<HT>(lds[2 * n], lds[2 * n + 1], out);
        };
        if (lg_w == 1) {
            // width-2 top level (cur >= 2 here): caps of the subtrees in this tile
            const uint32_t pairs = cur >> 1;
            if (n < pairs) {
                const uint64_t sub = (first >> 1) + n;
                if constexpr (HT == kPoseidon) {
                    if (ROW) {
                        const uint64_t v = poseidon_hash_two_row(lds[2 * n], lds[2 * n + 1], k);
                        if (writer) gmem(a.cap)[4 * sub + k] = v;
                        return;
                    }
                }
                uint64_t out[4];
                hash_node(out);
                if (writer) {
#pragma unroll
                    for (int i = 0; i < 4; i++) gmem(a.cap)[4 * sub + i] = out[i];
                }
            }
            return;
        }
        const uint32_t half = cur >> 1;
        uint64_t out[4];
        uint64_t v = 0;  // ROW: this lane's digest element
        if (n < half) {
            if constexpr (HT == kPoseidon) {
                if (ROW) v = poseidon_hash_two_row(lds[2 * n], lds[2 * n + 1], k);
                else hash_node(out);
            } else {
                hash_node(out);
            }
        }
        __syncthreads();
        lg_w -= 1;
        first >>= 1;
        if (n < half && writer) {
            const uint64_t g = first + n;
            const uint64_t sub = g >> lg_w, idx = g & ((1ull << lg_w) - 1);
            const gptr<uint64_t> dst = node_ptr(a, sub, lg_w, idx);
            if (ROW) {
                dst[k] = v;
                lds[n][k] = v;
            } else {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    dst[i] = out[i];
                    lds[n][i] = out[i];
                }
            }
        }
        cur = half;
        __syncthreads();
    }
}

// ---- one level of many trees per launch
//
// For large batches the trees' levels above the leaves run as one launch per
// level over every tree (after a leaves_only leaf launch): a level of n
// nodes is then n / 256 work-groups of full waves, where the fused kernel
// reduces each tile's levels on ever fewer lanes, one permutation latency
// each, holding the work-group's slot meanwhile.  The descriptor's lg_w is
// the OUTPUT level here (width 2^lg_w per subtree; 0: the caps), count its
// nodes over all subtrees, tile_begin its first work-group; node g =
// hash_two of the 8 consecutive words of its two children at level lg_w + 1.
// Poseidon hashes 64 nodes per wave on the matrix cores (the state [l, r, 0]
// laid out as leaf_hash_mfma lays out a leaf's first chunk), the other hashes
// one node per lane.  njobs == 0: the single descriptor `one` (a lone tree).
template <int HT>
__global__ void __launch_bounds__(kMerkleLevelLanes) __attribute__((amdgpu_waves_per_eu(merkle_waves_per_eu(HT))))
k_merkle_level(const MerkleArgs *__restrict__ jobs, uint32_t njobs, MerkleArgs one, uint32_t block0)
{
    __shared__ uint64_t kt[HT == kPoseidon ? kMfmaRcWords : 1];
    const uint32_t bid = blockIdx.x + block0;
    const MerkleArgs a = njobs ? *find_job(jobs, njobs, bid) : one;
    const uint32_t t = threadIdx.x, lw = a.lg_w;
    const uint64_t g0 = (uint64_t)(bid - a.tile_begin) * kMerkleLevelLanes;
    const uint64_t wmask = (1ull << lw) - 1;
    // first word of node g's two children (level lw + 1 >= 1)
    auto children = [&](uint64_t g) { return node_ptr(a, g >> lw, lw + 1, 2 * (g & wmask)); };
    const uint64_t g = g0 + t;
    uint64_t out[4];
    if constexpr (HT == kPoseidon) {
        mfma_fill_rc(kt, t, kMerkleLevelLanes);
        __syncthreads();
        const uint32_t w = t >> 6, j = t & 31, h = (t >> 5) & 1;
        const MfmaOps o = mfma_setup();
        uint64_t s[2][6];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const uint64_t n = g0 + 64 * w + 32 * q + j;
            const uint32_t el0 = 6 * ((h + q) & 1);  // this lane's first element of group q
            const bool real = n < a.count;           // lanes past the level's end hash zeros
            const gptr<const uint64_t> c = children(real ? n : 0);
#pragma unroll
            for (int e = 0; e < 6; e++) s[q][e] = real && el0 + e < 8 ? c[el0 + e] : 0;
        }
        poseidon_permute_mfma(s, o, kt);
        // elements 0-3 of group h sit in slots 0-3 of this lane: node g0 + 64w + 32h + j = g
#pragma unroll
        for (int i = 0; i < 4; i++) out[i] = gl::canon(h ? s[1][i] : s[0][i]);
    } else {
        if (g < a.count) {
            const gptr<const uint64_t> c = children(g);
            uint64_t l[4], r[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                l[i] = c[i];
                r[i] = c[4 + i];
            }
            hash_two<HT>(l, r, out);
        }
    }
    if (g < a.count) {
        const gptr<uint64_t> d = lw ? node_ptr(a, g >> lw, lw, g & wmask) : gmem(a.cap) + 4 * g;
#pragma unroll
        for (int i = 0; i < 4; i++) d[i] = out[i];
    }
}

}  // namespace zkx
