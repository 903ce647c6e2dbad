// sort_lds.hpp -- in-LDS bitonic sort of 64-bit keys for one workgroup and
// the run-length helpers built on it (TRACE entropies, occupied-key TE).
//
// A workgroup of nt = P / 16 threads sorts P keys (a power of two), thread
// tid holding elements 16 tid .. 16 tid + 15 in registers.  Stages whose
// partner element is in the same thread run in registers, those whose
// partner is in the same wave by lane shuffles, the rest through LDS (one
// store, one load and two barriers per stage).
#pragma once
#include "common.hpp"

namespace tnx {

constexpr int kSortE = 16;          // keys per thread
constexpr int kSortMinP = 1024;     // sort sizes: powers of two in [1024, 16384]
constexpr int kSortMaxP = 16384;

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}

// Ascending bitonic sort of P = 16 * nt keys, thread tid holding elements
// 16 tid .. 16 tid + 15 in r; lds holds P keys (element i of thread t at
// lds[(i % 16) * nt + t]).
__device__ void bitonic_sort(uint64_t (&r)[kSortE], uint64_t *lds, int P, int nt, int tid) {
    const int base = tid * kSortE;
    for (int k = 2; k <= P; k <<= 1) {
        // partners in other threads: j >= 16
        for (int j = k >> 1; j >= kSortE; j >>= 1) {
            const int m = j / kSortE;             // partner thread = tid ^ m
            const bool keep_min = (((tid & m) == 0) == ((base & k) == 0));
            if (m < kWave) {
#pragma unroll
                for (int e = 0; e < kSortE; ++e) {
                    const uint64_t o = shfl_xor_u64(r[e], m);
                    r[e] = keep_min ? (o < r[e] ? o : r[e]) : (o > r[e] ? o : r[e]);
                }
            } else {
                __syncthreads();
#pragma unroll
                for (int e = 0; e < kSortE; ++e) lds[e * nt + tid] = r[e];
                __syncthreads();
#pragma unroll
                for (int e = 0; e < kSortE; ++e) {
                    const uint64_t o = lds[e * nt + (tid ^ m)];
                    r[e] = keep_min ? (o < r[e] ? o : r[e]) : (o > r[e] ? o : r[e]);
                }
            }
        }
        // partners in the same thread: j < 16
#pragma unroll
        for (int j = kSortE / 2; j >= 1; j >>= 1) {
            if (j < k) {
#pragma unroll
                for (int e = 0; e < kSortE; ++e) {
                    if ((e & j) == 0) {
                        const int p = e | j;
                        const bool asc = ((base + e) & k) == 0;
                        const uint64_t x = r[e], y = r[p];
                        const uint64_t lo = x < y ? x : y, hi = x < y ? y : x;
                        r[e] = asc ? lo : hi;
                        r[p] = asc ? hi : lo;
                    }
                }
            }
        }
    }
}

__device__ __forceinline__ int64_t wave_sum_i64_t(int64_t v) {
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) v += (int64_t)shfl_xor_u64((uint64_t)v, m);
    return v;
}


}  // namespace tnx

namespace tnx {

// Sum of tbl[c] over the runs of equal (key >> sh) among the first L
// sorted elements (the run holding element L - 1 is cut there), as this
// thread's share: the caller sums the shares over the workgroup.  lds holds
// the sorted keys in the transposed layout of bitonic_sort; s_last is a
// 16-int workgroup scratch.  Run starts are found by a max-scan over the
// workgroup (within a wave by shuffles, across waves through s_last).
__device__ inline int64_t run_sum(const uint64_t (&key)[kSortE], const uint64_t *lds, int nt,
                                  int tid, int L, int sh, const int64_t *__restrict__ tbl,
                                  int32_t *s_last) {
    const int lane = tid & (kWave - 1), wid = tid / kWave, base = tid * kSortE;
    const uint64_t prev = tid > 0 ? lds[(kSortE - 1) * nt + tid - 1] >> sh : ~(key[0] >> sh);
    const uint64_t next = tid + 1 < nt ? lds[tid + 1] >> sh : ~(key[kSortE - 1] >> sh);
    int last = -1;
#pragma unroll
    for (int e = 0; e < kSortE; ++e) {
        const bool start = (e == 0 ? (key[0] >> sh) != prev : (key[e] >> sh) != (key[e - 1] >> sh));
        if (start) last = base + e;
    }
    int incl = last;
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
        const int o = __shfl_up(incl, m);
        if (lane >= m) incl = o > incl ? o : incl;
    }
    __syncthreads();  // s_last may still be read from a previous call
    if (lane == kWave - 1) s_last[wid] = incl;
    __syncthreads();
    int cur = __shfl_up(incl, 1);
    if (lane == 0) cur = -1;
    for (int w = 0; w < wid; ++w) cur = s_last[w] > cur ? s_last[w] : cur;
    int64_t acc = 0;
#pragma unroll
    for (int e = 0; e < kSortE; ++e) {
        const int i = base + e;
        const uint64_t k = key[e] >> sh;
        const bool start = (e == 0 ? k != prev : k != (key[e - 1] >> sh));
        if (start) cur = i;
        const bool end = (i == L - 1) || (e + 1 < kSortE ? (key[e + 1] >> sh) != k : next != k);
        if (i < L && end) acc += tbl[i - cur + 1];
    }
    return acc;
}

// Workgroup sum of v (valid in thread 0); s_sum is a 16-entry scratch.
__device__ inline int64_t block_sum_i64(int64_t v, int64_t *s_sum, int nt, int tid) {
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) v += (int64_t)shfl_xor_u64((uint64_t)v, m);
    __syncthreads();
    if ((tid & (kWave - 1)) == 0) s_sum[tid / kWave] = v;
    __syncthreads();
    int64_t s = 0;
    if (tid == 0)
        for (int w = 0; w < nt / kWave; ++w) s += s_sum[w];
    return s;
}

}  // namespace tnx
