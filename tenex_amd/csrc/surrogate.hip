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

// y'_k[t0 .. t0 + 4) of the block-shuffled row yr as one packed word, in
// two steps so the loads overlap other work: shuffle_issue requests the
// bytes, shuffle_finish assembles them.  Four consecutive positions inside
// one block are four consecutive bytes of the original row: two aligned
// dword loads and a byte funnel shift.  Groups that straddle a block
// boundary, reach the tail or pass L take the per-byte path at finish.
// Rows are padded to a multiple of 16 bytes, so the second dword of the
// fast path stays inside the row.
struct ShufWord {
    uint32_t lo, hi;
    int32_t sh;  // byte shift, or -1: per-byte path
};

__device__ __forceinline__ ShufWord shuffle_issue(const int8_t *__restrict__ yr, const BlockMap &m,
                                                  int32_t t0, int32_t L) {
    ShufWord w{0u, 0u, -1};
    const int32_t b = (int32_t)__umulhi((uint32_t)t0, m.magic);
    int32_t bb = b, r = t0 - b * m.bl;
    if (r >= m.bl) { bb += 1; r -= m.bl; }
    if (r + 3 < m.bl && t0 + 3 < m.head && t0 + 3 < L) {
        uint32_t pv = (uint32_t)m.prow[bb];
        if (pv >= (uint32_t)m.nb) pv = 0;
        const int32_t src = (int32_t)pv * m.bl + r;
        const uint32_t *p = reinterpret_cast<const uint32_t *>(yr + (src & ~3));
        w.sh = src & 3;
        w.lo = p[0];
        w.hi = w.sh ? p[1] : 0u;
    }
    return w;
}

// The rare groups off the fast path, one byte at a time (kept out of line
// so the common path stays small).
__device__ __noinline__ uint32_t shuffle_bytes(const int8_t *__restrict__ yr, BlockMap m,
                                               int32_t t0, int32_t L) {
    uint32_t wv = 0;
    for (int q = 0; q < 4; q++) {
        const int32_t t = t0 + q;
        const uint32_t v = t < L ? (uint32_t)(uint8_t)yr[bmap(m, t)] : 0u;
        wv |= v << (8 * q);
    }
    return wv;
}

__device__ __forceinline__ uint32_t shuffle_finish(const ShufWord &w, const int8_t *__restrict__ yr,
                                                   const BlockMap &m, int32_t t0, int32_t L) {
    if (w.sh >= 0) return __builtin_amdgcn_alignbyte(w.hi, w.lo, (uint32_t)w.sh);  // byte shift
    return shuffle_bytes(yr, m, t0, L);
}

// Stage y'_k into LDS: the words t0 = first, first + step, ... < L of this
// thread (at most NW), requested by stage_issue and written by
// stage_finish.
template <int NW>
struct Stage {
    ShufWord w[NW];
};

template <int NW>
__device__ __forceinline__ void stage_issue(Stage<NW> &st, const int8_t *__restrict__ yr,
                                            const BlockMap &m, int first, int step, int L) {
#pragma unroll
    for (int i = 0; i < NW; i++) {
        const int t0 = first + i * step;
        st.w[i] = t0 < L ? shuffle_issue(yr, m, t0, L) : ShufWord{0u, 0u, 0};
    }
}

template <int NW>
__device__ __forceinline__ void stage_finish(const Stage<NW> &st, int8_t *yb,
                                             const int8_t *__restrict__ yr, const BlockMap &m,
                                             int first, int step, int L) {
#pragma unroll
    for (int i = 0; i < NW; i++) {
        const int t0 = first + i * step;
        if (t0 < L) *reinterpret_cast<uint32_t *>(yb + t0) = shuffle_finish(st.w[i], yr, m, t0, L);
    }
}

// Count key = states'[t] * b_y + y'[t], t < L, into a u16-packed histogram,
// except key 0 (all three values at level 0): with zero-inflated data a
// quarter of the samples can share that key, and their atomics would
// serialise on one address.  Its count is L minus the others'
// (reduce_pair16).
// Four consecutive states from LDS (ST = uint8_t when every target has at
// most 16 levels, uint16_t otherwise).
__device__ __forceinline__ void load_states4(const uint8_t *stp, int t0, uint32_t *s4) {
    const uint32_t w = *reinterpret_cast<const uint32_t *>(stp + t0);
    s4[0] = w & 0xFFu;
    s4[1] = (w >> 8) & 0xFFu;
    s4[2] = (w >> 16) & 0xFFu;
    s4[3] = w >> 24;
}

__device__ __forceinline__ void load_states4(const uint16_t *stp, int t0, uint32_t *s4) {
    const uint2 sw = *reinterpret_cast<const uint2 *>(stp + t0);
    s4[0] = sw.x & 0xFFFFu;
    s4[1] = sw.x >> 16;
    s4[2] = sw.y & 0xFFFFu;
    s4[3] = sw.y >> 16;
}

template <typename ST>
__device__ __forceinline__ void count_pair16(uint32_t *hist, const ST *stp,
                                             const int8_t *yb, uint32_t by, int L, int lane) {
#pragma unroll 4
    for (int t0 = lane * 4; t0 < L; t0 += 256) {
        const uint32_t yw = *reinterpret_cast<const uint32_t *>(yb + t0);
        uint32_t s4[4];
        load_states4(stp, t0, s4);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t cell = __umul24(s4[q], by) + ((yw >> (8 * q)) & 0xFFu);
            if (t0 + q < L && cell != 0u) lds_inc16(hist, cell);
        }
    }
}

// S3 - S2b of a u16-packed pair histogram (cells (a * b_x + b) * b_y + c)
// counted by count_pair16 from L samples, summed over the wave; zeroes the
// histogram.  Cell 0 (not counted) is L minus the rest; it belongs to
// column 0, which lane 0 finishes once that count is known.
__device__ __forceinline__ DF reduce_pair16(uint32_t *hist, int bx, int by, int L,
                                            const float2 *ltbl, const float2 *__restrict__ tbl,
                                            int lane) {
    const int ncol = bx * by;
    DF a3 = {0.f, 0.f}, a2 = {0.f, 0.f};
    uint32_t mine = 0, c2_col0 = 0;
    for (int col = lane; col < ncol; col += kWave) {
        uint32_t c2 = 0;
        for (int aa = 0; aa < bx; aa++) {
            const uint32_t v = get16(hist, (uint32_t)(aa * ncol + col));
            c2 += v;
            xlogx_add_l(a3, ltbl, tbl, v);
        }
        mine += c2;
        if (col == 0) c2_col0 = c2;
        else xlogx_add_l(a2, ltbl, tbl, c2);
    }
    const uint32_t c0 = (uint32_t)L - wave_sum_u32(mine);
    if (lane == 0) {
        xlogx_add_l(a3, ltbl, tbl, c0);
        xlogx_add_l(a2, ltbl, tbl, c2_col0 + c0);
    }
    df_add(a3, -a2.hi, -a2.lo);
    const DF d = df_wave_sum(a3);
    wave_lds_fence();
    for (int i = lane; i < (bx * ncol + 1) / 2; i += kWave) hist[i] = 0;
    wave_lds_fence();
    return d;
}

// One workgroup per plan item.  With tg > 1 targets of the item handled
// together (tg_max from the host, by LDS), wave w owns target w: per
// surrogate k every target's shuffled states and S1'_k, S2a'_k are built
// once, then per source the shuffled row y'_k is staged once by the whole
// workgroup (double buffer, one barrier per source) and every wave counts
// its pair from it.  With one target, the four waves take every fourth
// source and stage their own rows.  Per-pair FP64 sums, sums of squares and
// counts of TE_k >= observed stay in registers (lane = source index / step)
// and are added to the public slots once per item.
template <int NW, typename ST>
__global__ __launch_bounds__(256) void surrogate_wave_kernel(tnx_surr_args_t a, int32_t lds_len,
                                                             int32_t tg_max, int32_t nbatch,
                                                             int32_t ksplit) {
    extern __shared__ uint32_t lds_u32[];
    __shared__ double s12_sh;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int tid = threadIdx.x;
    const int cap_words = a.cap_wave / 2;
    // LDS: [hist words x 4 waves][states' ST lds_len x tg_max]
    //      [y' bytes lds_len x (nbatch, or 4 with one target)][small c log2 c table]
    uint32_t *hist = lds_u32 + wave * cap_words;
    ST *stp0 = reinterpret_cast<ST *>(lds_u32 + kSurrWaves * cap_words);
    int8_t *ybuf = reinterpret_cast<int8_t *>(stp0 + (int64_t)lds_len * tg_max);
    float2 *ltbl = reinterpret_cast<float2 *>(ybuf +
                                              (int64_t)(tg_max > 1 ? nbatch : kSurrWaves) * lds_len);
    const int L = (int)(a.T - a.tau);
    const int n = (int)a.n_vars;
    const float2 *tbl = reinterpret_cast<const float2 *>(a.xlogx);
    const double invN = 1.0 / (double)L;
    const uint32_t magic = 0xFFFFFFFFu / (uint32_t)a.block_length;
    const int32_t head = a.n_blocks * a.block_length;
    for (int i = tid; i < kSurrWaves * cap_words; i += 256) lds_u32[i] = 0;
    for (int i = tid; i < kSmallTbl; i += 256)
        ltbl[i] = i <= L ? tbl[i] : make_float2(0.f, 0.f);
    __syncthreads();
    Stage<NW> stg;
    // work unit g = (item g % n_items, surrogate slice g / n_items): slices
    // of the k range give small plans enough workgroups; the statistics of
    // different slices meet in the public slots through atomic adds
    const int64_t n_units = a.n_wave_items * ksplit;
    for (int64_t g = blockIdx.x; g < n_units; g += gridDim.x) {
        const int64_t w = g % a.n_wave_items;
        const int kc = (int)(g / a.n_wave_items);
        const int k_beg = (int)((int64_t)a.n_surr * kc / ksplit);
        const int k_end = (int)((int64_t)a.n_surr * (kc + 1) / ksplit);
        const tnx_work_t it = a.work[w];
        for (int ti0 = 0; ti0 < it.tgt_cnt; ti0 += tg_max) {
        const int tg = min(tg_max, it.tgt_cnt - ti0);
        // batched staging whenever the launch has several-target items (a
        // leftover item with one target then counts on wave 0 only)
        const bool shared = tg_max > 1;
        const int chunk = shared ? kWave : kSurrWaves * kWave;  // sources per pass
        // this wave's target (shared mode) or the item's target
        const int my_t = shared ? (wave < tg ? wave : -1) : 0;
        const int x = my_t >= 0 ? a.tgt_order[it.tgt_beg + ti0 + my_t] : -1;
        const int bx = x >= 0 ? a.n_per_var[x] : 0;
        const ST *my_stp = stp0 + (int64_t)(my_t > 0 ? my_t : 0) * lds_len;
        for (int s0 = it.src_beg; s0 < it.src_end; s0 += chunk) {
        const int nsrc = min(it.src_end - s0, chunk);
        double acc_s = 0.0, acc_q = 0.0;
        int32_t acc_c = 0;
        for (int k = k_beg; k < k_end; k++) {
            // ── 1. shuffled states of the tg targets, S1'_k and S2a'_k ─────
            for (int j = 0; j < tg; j++) {
                const int xj = a.tgt_order[it.tgt_beg + ti0 + j];
                const int bxj = a.n_per_var[xj];
                const int8_t *xr = a.bins + (int64_t)xj * a.ld;
                BlockMap mx{a.perm + ((int64_t)k * n + xj) * a.n_blocks, a.block_length, head,
                            a.n_blocks, magic};
                ST *stp = stp0 + (int64_t)j * lds_len;
                uint32_t *hj = lds_u32 + j * cap_words;  // wave j's histogram
                for (int t = tid; t < L; t += 256) {
                    const int32_t c = xr[bmap(mx, t)];
                    const int32_t aa = xr[bmap(mx, t + a.tau)];
                    const uint32_t st = (uint32_t)(aa * bxj + c);
                    stp[t] = (ST)st;
                    if (bxj * bxj <= a.cap_wave) lds_inc16(hj, st);
                }
            }
            __syncthreads();
            double S12k = 0.0;  // S1'_k - S2a'_k of this wave's target
            if (x >= 0 && (shared || wave == 0)) {
                DF s2a = {0.f, 0.f}, s1 = {0.f, 0.f};
                for (int b = lane; b < bx; b += kWave) {
                    uint32_t c1 = 0;
                    for (int aa = 0; aa < bx; aa++) {
                        const uint32_t v = get16(hist, (uint32_t)(aa * bx + b));
                        c1 += v;
                        xlogx_add_l(s2a, ltbl, tbl, v);
                    }
                    xlogx_add_l(s1, ltbl, tbl, c1);
                }
                df_add(s1, -s2a.hi, -s2a.lo);
                S12k = df_to_double(df_wave_sum(s1));
                wave_lds_fence();
                for (int i = lane; i < (bx * bx + 1) / 2; i += kWave) hist[i] = 0;
                wave_lds_fence();
            }
            // ── 2. sources ─────────────────────────────────────────────────
            if (shared) {
                // sources in batches of nbatch: the workgroup stages the
                // batch's shuffled rows, one barrier, then every wave counts
                // its target against each of them
                for (int sb0 = 0; sb0 < nsrc; sb0 += nbatch) {
                    const int nbt = min(nbatch, nsrc - sb0);
                    for (int j = 0; j < nbt; j++) {
                        const int y = a.src_order[s0 + sb0 + j];
                        const int8_t *yr = a.bins + (int64_t)y * a.ld;
                        const BlockMap my{a.perm + ((int64_t)k * n + y) * a.n_blocks,
                                          a.block_length, head, a.n_blocks, magic};
                        stage_issue<NW>(stg, yr, my, tid * 4, 1024, L);
                        stage_finish<NW>(stg, ybuf + (int64_t)j * lds_len, yr, my, tid * 4, 1024,
                                         L);
                    }
                    __syncthreads();
                    for (int j = 0; j < nbt; j++) {
                        const int sj = sb0 + j;
                        const int y = a.src_order[s0 + sj];
                        const int by = a.n_per_var[y];
                        if (x < 0 || y == x || bx * bx * by > a.cap_wave) continue;
                        count_pair16(hist, my_stp, ybuf + (int64_t)j * lds_len, (uint32_t)by, L,
                                     lane);
                        wave_lds_fence();
                        const DF d = reduce_pair16(hist, bx, by, L, ltbl, tbl, lane);
                        const double te = (df_to_double(d) + S12k) * invN;
                        const int64_t slot = (int64_t)y * n + x;
                        if (lane == sj) {  // the float32 TE_k, as emit_te accumulates it
                            const float tf = (float)te;
                            const double td = (double)tf;
                            acc_s += td;
                            acc_q += td * td;
                            acc_c += (tf >= a.observed[slot]) ? 1 : 0;
                        }
                        if (a.te_k_dbg && lane == 0) a.te_k_dbg[slot * a.n_surr + k] = (float)te;
                    }
                    __syncthreads();  // the batch's rows are replaced next
                }
            } else {
                // one target: wave 0 holds S1'_k - S2a'_k; hand it to all waves
                if (wave == 0 && lane == 0) s12_sh = S12k;
                __syncthreads();
                S12k = s12_sh;
                // each wave its own sources and its own staging buffer
                int8_t *yb = ybuf + (int64_t)wave * lds_len;
                for (int sj = wave; sj < nsrc; sj += kSurrWaves) {
                    const int y = a.src_order[s0 + sj];
                    const int by = a.n_per_var[y];
                    if (y == x || bx * bx * by > a.cap_wave) continue;
                    const int8_t *yr = a.bins + (int64_t)y * a.ld;
                    const BlockMap my{a.perm + ((int64_t)k * n + y) * a.n_blocks, a.block_length,
                                      head, a.n_blocks, magic};
                    Stage<4> sb;
                    for (int f = lane * 4; f < L; f += 4 * 256) {
                        stage_issue<4>(sb, yr, my, f, 256, L);
                        stage_finish<4>(sb, yb, yr, my, f, 256, L);
                    }
                    wave_lds_fence();  // y' of sj written by this wave
                    {
                        count_pair16(hist, my_stp, yb, (uint32_t)by, L, lane);
                        wave_lds_fence();
                        const DF d = reduce_pair16(hist, bx, by, L, ltbl, tbl, lane);
                        const double te = (df_to_double(d) + S12k) * invN;
                        const int64_t slot = (int64_t)y * n + x;
                        if (lane == (sj / kSurrWaves)) {  // the float32 TE_k
                            const float tf = (float)te;
                            const double td = (double)tf;
                            acc_s += td;
                            acc_q += td * td;
                            acc_c += (tf >= a.observed[slot]) ? 1 : 0;
                        }
                        if (a.te_k_dbg && lane == 0) a.te_k_dbg[slot * a.n_surr + k] = (float)te;
                    }
                }
            }
            __syncthreads();  // states' and y' are rebuilt for the next k
        }
        // the per-pair accumulators of this pass
        const int sj = shared ? lane : wave + kSurrWaves * lane;
        if (x >= 0 && sj < nsrc) {
            const int y = a.src_order[s0 + sj];
            const int by = a.n_per_var[y];
            if (y != x && bx * bx * by <= a.cap_wave) {
                const int64_t slot = (int64_t)y * n + x;
                atomicAdd(a.sum_te + slot, acc_s);
                atomicAdd(a.sumsq_te + slot, acc_q);
                atomicAdd(a.count_ge + slot, acc_c);
            }
        }
        __syncthreads();
        }
        }
    }
}

// One shuffled copy of every output row: thread i of the grid writes 16
// bytes of one row (four words, zero-padded past T).  Output row v reads
// input row v % n_src (n_rows = copies x n_src, below 2^31).
__global__ __launch_bounds__(256) void block_shuffle_kernel(const int8_t *__restrict__ bins,
                                                            int64_t n_rows, int32_t n_src,
                                                            int32_t T, int64_t ld,
                                                            const int32_t *__restrict__ perm,
                                                            int32_t n_blocks, int32_t bl,
                                                            int8_t *__restrict__ out) {
    const int64_t per_row = ld / 16;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_rows * per_row) return;
    const int64_t v = i / per_row;
    const int32_t t16 = (int32_t)(i - v * per_row) * 16;
    const int8_t *yr = bins + (int64_t)((uint32_t)v % (uint32_t)n_src) * ld;
    const BlockMap m{perm + v * n_blocks, bl, n_blocks * bl, n_blocks, 0xFFFFFFFFu / (uint32_t)bl};
    uint32_t w[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int32_t t0 = t16 + 4 * j;
        if (t0 >= m.head && t0 + 3 < T) {
            w[j] = *reinterpret_cast<const uint32_t *>(yr + t0);  // tail: unshuffled
        } else if (t0 < T) {
            const ShufWord sw = shuffle_issue(yr, m, t0, T);
            w[j] = shuffle_finish(sw, yr, m, t0, T);
        } else {
            w[j] = 0u;
        }
    }
    *reinterpret_cast<uint4 *>(out + v * ld + t16) = make_uint4(w[0], w[1], w[2], w[3]);
}

}  // namespace tnx

using namespace tnx;

extern "C" int tnx_block_shuffle_copies(const int8_t *bins, int64_t n_vars, int32_t n_copies,
                                        int64_t T, int64_t ld, const int32_t *perm,
                                        int32_t n_blocks, int32_t block_length, int8_t *out,
                                        void *stream) {
    clear_error();
    TNX_CHECK_ARG(bins && perm && out && n_vars >= 0 && n_copies >= 1 && T >= 1,
                  "tnx_block_shuffle: bad arguments");
    TNX_CHECK_ARG(ld >= T && ld % 16 == 0, "tnx_block_shuffle: ld must be >= T, multiple of 16");
    TNX_CHECK_ARG(block_length >= 1 && n_blocks >= 1 && (int64_t)n_blocks * block_length <= T,
                  "tnx_block_shuffle: need n_blocks * block_length <= T");
    TNX_CHECK_ARG(T < (1LL << 31) / 2, "tnx_block_shuffle: T too large");
    TNX_CHECK_ARG(n_vars * n_copies < (1LL << 31), "tnx_block_shuffle: too many rows");
    TNX_CHECK_ARG(bins != out, "tnx_block_shuffle: out must not alias bins");
    if (n_vars == 0) return TNX_OK;
    const int64_t rows = n_vars * n_copies;
    const int64_t threads = rows * (ld / 16);
    TNX_CHECK_ARG((threads + 255) / 256 < (1LL << 32), "tnx_block_shuffle: too many rows");
    hipLaunchKernelGGL(block_shuffle_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                       as_stream(stream), bins, rows, (int32_t)n_vars, (int32_t)T, ld, perm,
                       n_blocks, block_length, out);
    TNX_CHECK_LAUNCH("block_shuffle_kernel");
    return TNX_OK;
}

extern "C" int tnx_block_shuffle(const int8_t *bins, int64_t n_vars, int64_t T, int64_t ld,
                                 const int32_t *perm, int32_t n_blocks, int32_t block_length,
                                 int8_t *out, void *stream) {
    return tnx_block_shuffle_copies(bins, n_vars, 1, T, ld, perm, n_blocks, block_length, out,
                                    stream);
}

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
    // several targets per workgroup share each staged source row when their
    // shuffled states fit beside the histograms; else one target at a time.
    // States are bytes when every target has at most 16 levels, else 16 bit.
    const bool st8 = a.max_levels >= 1 && a.max_levels <= 16;
    const int st_bytes = st8 ? 1 : 2;
    // staged rows: nb in several-target mode, one per wave with one target
    auto lds_for = [&](int tg, int nb) {
        const int rows = tg > 1 ? nb : kSurrWaves;
        return (size_t)kSurrWaves * (a.cap_wave / 2) * 4 + (size_t)lds_len * st_bytes * tg +
               (size_t)rows * lds_len + (size_t)kSmallTbl * sizeof(float2);
    };
    const int items_tg = a.max_tgt_cnt > 0 ? std::min(a.max_tgt_cnt, kSurrWaves) : kSurrWaves;
    const int tg_max = (items_tg > 1 && lds_for(kSurrWaves, 2) <= 160 * 1024) ? kSurrWaves : 1;
    // the largest batch of staged rows (2..6) that keeps three workgroups per
    // CU, else the largest that fits
    int nbatch = 2;
    for (int nb = 6; nb >= 2; nb--)
        if (3 * lds_for(tg_max, nb) <= 160 * 1024) { nbatch = nb; break; }
    if (3 * lds_for(tg_max, nbatch) > 160 * 1024)
        for (int nb = 6; nb >= 2; nb--)
            if (lds_for(tg_max, nb) <= 160 * 1024) { nbatch = nb; break; }
    const size_t shm = lds_for(tg_max, nbatch);
    TNX_CHECK_ARG(shm <= 160 * 1024, "tnx_surrogate_run: LDS %zu bytes exceeds 160 KiB "
                  "(reduce cap_wave or T)", shm);
    // words staged per thread and source with several targets (L / 1024);
    // the one-target mode stages in groups of four words
    const int64_t words = tg_max > 1 ? (L + 1023) / 1024 : 1;
    // enough work units to fill the chip: slice the surrogates of small plans
    const int64_t want_units = 3 * 256;
    const int ksplit = (int)std::max<int64_t>(
        1, std::min<int64_t>(a.n_surr, (want_units + a.n_wave_items - 1) / a.n_wave_items));
    const unsigned grid = (unsigned)std::min<int64_t>(a.n_wave_items * ksplit, 1 << 20);
    const void *kfn = nullptr;
#define TNX_SURR(NWV, STT)                                                                   \
    if (!kfn && words <= NWV && (sizeof(STT) == 1) == st8) {                                \
        kfn = (const void *)surrogate_wave_kernel<NWV, STT>;                                  \
        TNX_CHECK_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize,    \
                                          (int)shm));                                         \
        hipLaunchKernelGGL((surrogate_wave_kernel<NWV, STT>), dim3(grid), dim3(256), shm,     \
                           as_stream(stream), a, lds_len, tg_max, nbatch, ksplit);            \
    }
    TNX_SURR(1, uint8_t) TNX_SURR(2, uint8_t) TNX_SURR(4, uint8_t) TNX_SURR(8, uint8_t)
    TNX_SURR(16, uint8_t)
    TNX_SURR(1, uint16_t) TNX_SURR(2, uint16_t) TNX_SURR(4, uint16_t) TNX_SURR(8, uint16_t)
    TNX_SURR(16, uint16_t)
#undef TNX_SURR
    TNX_CHECK_ARG(kfn != nullptr, "tnx_surrogate_run: L too long");
    TNX_CHECK_LAUNCH("surrogate_wave_kernel");
    return TNX_OK;
}
