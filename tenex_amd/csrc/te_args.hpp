// te_args.hpp -- launch constants and per-pair output helpers shared by the
// TE kernels (te_hist.hip) and the persistent single launch (te_persist.hip).
#pragma once
#include "common.hpp"

namespace tnx {

constexpr int kWavesPerWG = 4;
constexpr int kBlockThreads = 512;
#ifndef TNX_DEFAULT_CAP_WAVE
// 3136 cells + 64 key-0 words = 12.5 KiB per wave: three workgroups of four
// waves (plus each workgroup's 2 KiB c log2 c table) fit one CU's 160 KiB
#define TNX_DEFAULT_CAP_WAVE 3136
#endif
constexpr int kDefaultCapWave = TNX_DEFAULT_CAP_WAVE;
constexpr int kDefaultCapBlock = 16384;  // 64 KiB + the b_t b_s marginal
// LDS the planner aims for per block-kernel workgroup, so that two
// workgroups fit one CU's 160 KiB: histogram + marginal + reduction scratch.
constexpr int kBlockLdsTarget = 80 * 1024 - 512 - 8 * kSmallTbl;
constexpr int kWaveItemPairs = 256;  // pairs per wave-kernel work item
constexpr int kXcds = 8;             // gfx950: 8 XCDs, workgroups dealt round-robin
constexpr int kChunkSrc = 64;        // sources per locality chunk

// Words of one wave's LDS slice: cap histogram words and the 64 lanes' key-0
// words, rounded to 16 bytes (the copy fold reads four copies at a time).
__host__ __device__ inline int wave_slice(int cap) { return (cap + kWave + 3) & ~3; }

// Hot target states (persistent launch, hot_cw > 0): words of a wave slice's
// hot cells (4 states x hot_maxb source levels x hot_cw copies), and the
// words before the slice's 64 lane words (histogram, then the hot cells
// from a 16-byte boundary).
constexpr int kHotCB = 16;  // block-role copies of a hot cell
__host__ __device__ inline int hot_wave_words(const tnx_te_args_t &a) {
    return a.hot_cw > 0 ? 4 * a.hot_cw * a.hot_maxb : 0;
}
__host__ __device__ inline int hot_block_words(const tnx_te_args_t &a) {
    return a.hot_cw > 0 ? 4 * kHotCB * a.hot_maxb : 0;
}
__host__ __device__ inline int hot_base(const tnx_te_args_t &a) { return (a.cap_wave + 3) & ~3; }
__host__ __device__ inline int wave_area(const tnx_te_args_t &a) {
    return a.hot_cw > 0 ? hot_base(a) + hot_wave_words(a) : a.cap_wave;
}

// Item of workgroup slot w in [0, n): workgroup w runs on XCD w % 8, and XCD
// x takes the x-th contiguous slice of the items, in order (a bijection of
// [0, n)).
// Writes TE of the pair at slot: the value, or (surrogate statistics) its
// float32 rounding added to the FP64 sum and sum of squares and the count of
// values >= the observed TE (float32 compare), as the reference accumulates
// its per-surrogate float32 TE matrices.
// Whether a pair of b_x^2 b_y cells is left to the occupied-key kernel.
__device__ __forceinline__ bool sort_pair(const tnx_te_args_t &a, int bx, int by) {
    return a.cap_sort > 0 && (int64_t)bx * bx * by > a.cap_sort;
}

__device__ __forceinline__ void emit_te(const tnx_te_args_t &a, int64_t slot, double te) {
    if (a.out_mode == TNX_OUT_STATS) {
        const float tf = (float)te;
        const double td = (double)tf;
        a.sum_te[slot] += td;
        a.sumsq_te[slot] += td * td;
        a.count_ge[slot] += (tf >= a.observed[slot]) ? 1 : 0;
    } else {
        a.out[slot] = (float)te;
    }
}

__device__ __forceinline__ int64_t xcd_item(int64_t w, int64_t n) {
    const int64_t x = w % kXcds, i = w / kXcds;
    const int64_t q = n / kXcds, r = n % kXcds;
    return x * q + (x < r ? x : r) + i;
}
constexpr int kBlockItemPairs = 16;  // pairs per block-kernel work item (at most)
// Small plans split their pairs into more, shorter items so that every wave
// slot gets work: about four items per resident workgroup (256 CUs x 3
// wave-kernel or 2 block-kernel workgroups).
constexpr int64_t kFillWaveItems = 256 * 3 * 4;
constexpr int64_t kFillBlockItems = 256 * 2 * 4;

// Output slot of pair (x, y); returns false if the pair is filtered out.
__device__ __forceinline__ bool out_slot(const tnx_te_args_t &a, int x, int y, int64_t &slot) {
    if (a.out_mode == TNX_OUT_MATRIX) {
        slot = (int64_t)x * a.ld_out + y;
        return true;
    }
    if (a.out_mode == TNX_OUT_STATS) {
        slot = (int64_t)y * a.ld_out + x;  // public layout [source, target]
        return true;
    }
    if (a.out_mode == TNX_OUT_STAGE) {  // x, y in the same copy x / stage_vars
        const int k = x / a.stage_vars;
        slot = (int64_t)k * a.stage_stride + (int64_t)(y - k * a.stage_vars) * a.ld_out +
               (x - k * a.stage_vars);
        return true;
    }
    int64_t id = (int64_t)x * (a.n_vars - 1) + (y - (y > x ? 1 : 0));
    slot = id - a.pair_offset;
    return slot >= 0 && slot < a.n_pairs_local;
}

// The persistent launch (te_persist.hip): one grid of resident workgroups
// over the plan's work queues; returns TNX_OK or an error status.
int launch_persist(const tnx_te_args_t &a, int ni, hipStream_t s);
// Longest series the persistent launch takes (register variants NI <= 12).
constexpr int64_t kPersistMaxL = 12 * 1024;

}  // namespace tnx
