// surrogate.hip -- fused block-shuffle surrogate statistics (gfx950).
//
// For surrogate k every variable v is read through its own block order
// perm[k, v, :]: position t < head = n_blocks * block_length maps to
// perm[k, v, t / bl] * bl + t % bl, the tail [head, T) is untouched
// (tenex/inference/surrogate_test.py:107-148, the same indexing as the
// reference's fused kernel tenex/csrc/full_smem_surrogate_test.cu:112-176).
// The shuffled target is shifted after shuffling, so its (x'_{t+tau}, x'_t)
// states, S1' and S2a' differ per k and are rebuilt per (target, k).
//
// One workgroup = one target x and a chunk of sources (plan items with
// tgt_cnt == 1).  Per k:
//   1. all 256 threads build states'[t] = x'[t+tau] * b_x + x'[t] in LDS and
//      count them; wave 0 reduces S1'_k and S2a'_k;
//   2. each wave takes its sources: stages y'_k (the shuffled source row)
//      in LDS, counts key = states'[t] * b_y + y'[t] into its private u16
//      histogram (two counters per 32-bit word), reduces S3 and S2b, and
//      adds TE_k to the pair's FP64 sum and sum of squares and the count of
//      TE_k >= observed (float32 compare, as the reference).
// Statistics go to the public slot [source * n + target]
// (full_smem_surrogate_test.cu:83) and are accumulated (+=), so surrogates
// can be split across calls.
#include <algorithm>

#include "common.hpp"
#include "te_device.hpp"

namespace tnx {

constexpr int kSurrWaves = 4;
constexpr int kSurrMaxSrcPerWave = 64;

struct BlockMap {
    const int32_t *prow;  // perm[k, v, :]
    int32_t bl, head, nb;
    uint32_t magic;       // floor((2^32 - 1) / bl): umulhi(t, magic) is t / bl or one less
};

// Position in the original row of shuffled position t.  Block indices that
// are not in [0, n_blocks) (an invalid perm) read block 0 instead of
// leaving the row.
__device__ __forceinline__ int32_t bmap(const BlockMap &m, int32_t t) {
    if (t >= m.head) return t;
    int32_t b = (int32_t)__umulhi((uint32_t)t, m.magic);
    int32_t r = t - b * m.bl;
    if (r >= m.bl) { b += 1; r -= m.bl; }
    uint32_t pv = (uint32_t)m.prow[b];
    if (pv >= (uint32_t)m.nb) pv = 0;
    return (int32_t)pv * m.bl + r;
}

__device__ __forceinline__ void lds_inc16(uint32_t *words, uint32_t cell) {
    __hip_atomic_fetch_add(words + (cell >> 1), 1u << (16 * (cell & 1)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ uint32_t get16(const uint32_t *words, uint32_t cell) {
    return (words[cell >> 1] >> (16 * (cell & 1))) & 0xFFFFu;
}

__global__ __launch_bounds__(256) void surrogate_wave_kernel(tnx_surr_args_t a, int32_t lds_len,
                                                             int32_t ystage_len) {
    extern __shared__ uint32_t lds_u32[];
    __shared__ double sk[2];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int tid = threadIdx.x;
    const int cap_words = a.cap_wave / 2;
    // LDS: [hist words x 4 waves][states' u16 lds_len][y' bytes x 4 waves]
    uint32_t *hist = lds_u32 + wave * cap_words;
    uint16_t *stp = reinterpret_cast<uint16_t *>(lds_u32 + kSurrWaves * cap_words);
    int8_t *ystage = reinterpret_cast<int8_t *>(stp + lds_len) + wave * ystage_len;
    const int L = (int)(a.T - a.tau);
    const int n = (int)a.n_vars;
    const float2 *tbl = reinterpret_cast<const float2 *>(a.xlogx);
    const double invN = 1.0 / (double)L;
    const uint32_t magic = 0xFFFFFFFFu / (uint32_t)a.block_length;
    const int32_t head = a.n_blocks * a.block_length;
    for (int i = tid; i < kSurrWaves * cap_words; i += 256) lds_u32[i] = 0;
    __syncthreads();
    for (int64_t w = blockIdx.x; w < a.n_wave_items; w += gridDim.x) {
        const tnx_work_t it = a.work[w];
        for (int ti = 0; ti < it.tgt_cnt; ti++)
        for (int s0 = it.src_beg; s0 < it.src_end; s0 += kSurrWaves * kSurrMaxSrcPerWave) {
        const int x = a.tgt_order[it.tgt_beg + ti];
        const int bx = a.n_per_var[x];
        if (bx * bx > a.cap_wave) continue;  // every pair of x exceeds the histogram
        const int8_t *xr = a.bins + (int64_t)x * a.ld;
        const int nsrc = min(it.src_end - s0, kSurrWaves * kSurrMaxSrcPerWave);
        // per-source accumulators: lane s of wave w holds source w + 4 s
        double acc_s = 0.0, acc_q = 0.0;
        int32_t acc_c = 0;
        for (int k = 0; k < a.n_surr; k++) {
            // ── 1. shuffled target states and S1'_k, S2a'_k ────────────────
            BlockMap mx{a.perm + ((int64_t)k * n + x) * a.n_blocks, a.block_length, head, a.n_blocks,
                        magic};
            uint32_t *h0 = lds_u32;  // wave 0's histogram, bx^2 cells
            for (int t = tid; t < L; t += 256) {
                const int32_t c = xr[bmap(mx, t)];
                const int32_t aa = xr[bmap(mx, t + a.tau)];
                const uint32_t s = (uint32_t)(aa * bx + c);
                stp[t] = (uint16_t)s;
                lds_inc16(h0, s);
            }
            __syncthreads();
            if (wave == 0) {
                DF s2a = {0.f, 0.f}, s1 = {0.f, 0.f};
                for (int b = lane; b < bx; b += kWave) {
                    uint32_t c1 = 0;
                    for (int aa = 0; aa < bx; aa++) {
                        const uint32_t v = get16(h0, (uint32_t)(aa * bx + b));
                        c1 += v;
                        xlogx_add(s2a, tbl, v);
                    }
                    xlogx_add(s1, tbl, c1);
                }
                s2a = df_wave_sum(s2a);
                s1 = df_wave_sum(s1);
                wave_lds_fence();
                for (int i = lane; i < (bx * bx + 1) / 2; i += kWave) h0[i] = 0;
                if (lane == 0) {
                    sk[0] = df_to_double(s1);
                    sk[1] = df_to_double(s2a);
                }
            }
            __syncthreads();
            const double S1k = sk[0], S2ak = sk[1];
            // ── 2. this wave's sources ────────────────────────────────────
            for (int sj = wave; sj < nsrc; sj += kSurrWaves) {
                const int y = a.src_order[s0 + sj];
                if (y == x) continue;
                const int by = a.n_per_var[y];
                if (bx * bx * by > a.cap_wave) continue;  // excluded by the host
                const int8_t *yr = a.bins + (int64_t)y * a.ld;
                BlockMap my{a.perm + ((int64_t)k * n + y) * a.n_blocks, a.block_length, head,
                            a.n_blocks, magic};
                // stage y'_k[t], t < L, four bytes per lane step
                for (int t0 = lane * 4; t0 < L; t0 += 256) {
                    uint32_t wv = 0;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int t = t0 + q;
                        const uint32_t v = t < L ? (uint32_t)(uint8_t)yr[bmap(my, t)] : 0u;
                        wv |= v << (8 * q);
                    }
                    *reinterpret_cast<uint32_t *>(ystage + t0) = wv;
                }
                wave_lds_fence();
                // count
                for (int t0 = lane * 4; t0 < L; t0 += 256) {
                    const uint32_t yw = *reinterpret_cast<const uint32_t *>(ystage + t0);
                    const uint2 sw = *reinterpret_cast<const uint2 *>(stp + t0);
                    const uint32_t s4[4] = {sw.x & 0xFFFFu, sw.x >> 16, sw.y & 0xFFFFu, sw.y >> 16};
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        if (t0 + q < L) lds_inc16(hist, __umul24(s4[q], (uint32_t)by) + ((yw >> (8 * q)) & 0xFFu));
                }
                wave_lds_fence();
                // reduce
                const int ncol = bx * by;
                DF a3 = {0.f, 0.f}, a2 = {0.f, 0.f};
                for (int col = lane; col < ncol; col += kWave) {
                    uint32_t c2 = 0;
                    for (int aa = 0; aa < bx; aa++) {
                        const uint32_t v = get16(hist, (uint32_t)(aa * ncol + col));
                        c2 += v;
                        xlogx_add(a3, tbl, v);
                    }
                    xlogx_add(a2, tbl, c2);
                }
                a3 = df_wave_sum(a3);
                a2 = df_wave_sum(a2);
                wave_lds_fence();
                for (int i = lane; i < (bx * ncol + 1) / 2; i += kWave) hist[i] = 0;
                wave_lds_fence();
                const double te = (df_to_double(a3) + S1k - S2ak - df_to_double(a2)) * invN;
                const int64_t slot = (int64_t)y * n + x;
                if (lane == (sj / kSurrWaves)) {
                    acc_s += te;
                    acc_q += te * te;
                    acc_c += ((float)te >= a.observed[slot]) ? 1 : 0;
                }
                if (a.te_k_dbg && lane == 0) a.te_k_dbg[slot * a.n_surr + k] = (float)te;
            }
            __syncthreads();  // states' is rebuilt for the next k
        }
        // write the per-source accumulators
        const int sj = wave + kSurrWaves * lane;
        if (lane < kSurrMaxSrcPerWave && sj < nsrc) {
            const int y = a.src_order[s0 + sj];
            const int by = a.n_per_var[y];
            if (y != x && bx * bx * by <= a.cap_wave) {
                const int64_t slot = (int64_t)y * n + x;
                a.sum_te[slot] += acc_s;
                a.sumsq_te[slot] += acc_q;
                a.count_ge[slot] += acc_c;
            }
        }
        __syncthreads();
        }
    }
}

}  // namespace tnx

using namespace tnx;

extern "C" int tnx_surrogate_run(const tnx_surr_args_t *args, void *stream) {
    clear_error();
    TNX_CHECK_ARG(args != nullptr, "tnx_surrogate_run: args is NULL");
    const tnx_surr_args_t &a = *args;
    TNX_CHECK_ARG(a.bins && a.n_per_var && a.xlogx && a.perm && a.tgt_order && a.src_order &&
                      a.work && a.observed && a.sum_te && a.sumsq_te && a.count_ge,
                  "tnx_surrogate_run: null pointer");
    TNX_CHECK_ARG(a.tau >= 1 && a.tau < a.T, "tnx_surrogate_run: need 1 <= tau < T");
    TNX_CHECK_ARG(a.ld >= a.T && a.ld % 16 == 0, "tnx_surrogate_run: ld must be >= T, multiple of 16");
    TNX_CHECK_ARG(a.block_length >= 1 && a.n_blocks >= 1 &&
                      (int64_t)a.n_blocks * a.block_length <= a.T,
                  "tnx_surrogate_run: need n_blocks * block_length <= T");
    TNX_CHECK_ARG(a.T * (int64_t)a.block_length < (1LL << 32), "tnx_surrogate_run: T too large");
    TNX_CHECK_ARG(a.T - a.tau <= 65535, "tnx_surrogate_run: L must be <= 65535 (16-bit counters)");
    TNX_CHECK_ARG(a.n_surr >= 1, "tnx_surrogate_run: n_surr must be >= 1");
    TNX_CHECK_ARG(a.cap_wave >= 128 * 2 && a.cap_wave <= 32768 && a.cap_wave % 2 == 0,
                  "tnx_surrogate_run: cap_wave must be even, in [256, 32768] (>= b_max^2 cells)");
    TNX_CHECK_ARG(a.n_block_items == 0,
                  "tnx_surrogate_run: pairs with b_t^2 * b_s > cap_wave are not supported");
    if (a.n_wave_items == 0) return TNX_OK;
    const int64_t L = a.T - a.tau;
    const int32_t lds_len = (int32_t)((L + 63) / 64 * 64);
    const int32_t ystage_len = lds_len;
    const size_t shm = (size_t)kSurrWaves * (a.cap_wave / 2) * 4 + (size_t)lds_len * 2 +
                       (size_t)kSurrWaves * ystage_len;
    TNX_CHECK_ARG(shm <= 160 * 1024, "tnx_surrogate_run: LDS %zu bytes exceeds 160 KiB "
                  "(reduce cap_wave or T)", shm);
    TNX_CHECK_HIP(hipFuncSetAttribute((const void *)surrogate_wave_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    const unsigned grid = (unsigned)std::min<int64_t>(a.n_wave_items, 1 << 20);
    hipLaunchKernelGGL(surrogate_wave_kernel, dim3(grid), dim3(256), shm, as_stream(stream), a,
                       lds_len, ystage_len);
    TNX_CHECK_LAUNCH("surrogate_wave_kernel");
    return TNX_OK;
}
