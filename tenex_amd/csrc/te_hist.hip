// te_hist.hip -- pairwise transfer entropy from LDS-resident joint histograms.
//
// For a pair (target x, source y) with b_x, b_y levels and L = T - tau:
//   key(t) = (x[t+tau] * b_x + x[t]) * b_y + y[t]      (state * b_y + y)
//   S3  = sum over the b_x^2 b_y cells of c log2 c
//   S2b = sum over the b_x b_y cells of (x_t, y_t) of c log2 c
//   TE(y -> x) = (S3 + S1[x] - S2a[x] - S2b) / L
// which is the reference formula (tenex/kernels/scatter_add.py:98-117,
// tenex/csrc/full_smem.cu:113-128) with the log expanded; S1, S2a depend on
// the target only and come from tnx_gene_prep.
//
// Two kernels share the accumulation / reduction code:
//   te_wave_kernel   one wavefront per pair, private LDS histogram of
//                    cap_wave cells (pairs with b_x^2 b_y <= cap_wave);
//                    four waves of a workgroup work on up to four targets
//                    against the same source rows (L1 reuse).
//   te_block_kernel  one workgroup per pair, histogram of cap_block cells,
//                    several passes over ranges of x_{t+tau} when the pair
//                    has more cells (exact counts, no coarsening).
// Replaces te_smem_kernel(_pairfree) (tenex/csrc/full_smem.cu:31-276),
// te_adaptive_smem_kernel(_pairfree) (tenex/csrc/adaptive_smem.cu:30-273) and
// te_gmem_kernel (tenex/csrc/gmem.cu:27-155).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "te_args.hpp"
#include "te_device.hpp"

namespace tnx {


// Wave-per-pair histogram kernel.  The four waves of a workgroup take the
// (up to four) targets of a work item and stream the same source rows, which
// stay in L1.  NI > 0: each wave keeps its target's states in registers
// (L <= NI * 1024) so only the source row is read per pair; NI == 0: states
// are re-read from memory (any L).  Every pair counts into privatised copies
// of its histogram (wave_copies) inside the wave's cap_wave-cell LDS slice.
// The register variants are limited to 3 waves per SIMD so the state
// registers do not spill; the LDS slice is sized for that occupancy.
template <int NI, bool HZ>
__device__ __forceinline__ void te_wave_body(const tnx_te_args_t &a) {
    extern __shared__ uint32_t lds_u32[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // slice: cap_wave histogram words, then the lanes' key-0 words
    uint32_t *hist = lds_u32 + wave * wave_slice(a.cap_wave);
    uint32_t *zero_row = hist + a.cap_wave;
    for (int i = lane; i < wave_slice(a.cap_wave); i += kWave) hist[i] = 0;
    wave_lds_fence();
    const int L = (int)(a.T - a.tau);
    const float2 *tbl = reinterpret_cast<const float2 *>(a.xlogx);
    const double invN = 1.0 / (double)L;
    TargetRegs<(NI > 0 ? NI : 1)> regs;
    int cur_x = -1;
    double s12 = 0.0;  // S1[x] - S2a[x] of the target in registers
    // short series (NI <= 4): the next pair's source chunks are requested
    // before the current pair's histogram is reduced, so their latency
    // overlaps the reduction; pf_y names the row they hold (-1: none)
    constexpr bool kPrefetch = NI > 0 && NI <= 4;
    SourceChunks<(kPrefetch ? NI : 1)> ysrc;
    int pf_y = -1;
    // Items are dealt to workgroups (XCD-sliced order).  By default the four
    // waves split an item's pairs; with item_per_wave each wave takes its
    // own item of a group of four (plans of one-target items, where a wave
    // then keeps its target in registers for all of the item's pairs).
    constexpr bool kPerWaveOK = NI > 0 && NI <= 4;   // a hint, honoured for short series
    const bool per_wave = kPerWaveOK && a.item_per_wave != 0;
    const int64_t n_groups =
        per_wave ? (a.n_wave_items + kWavesPerWG - 1) / kWavesPerWG : a.n_wave_items;
    const int pfirst = per_wave ? 0 : wave;
    const int pstep = per_wave ? 1 : kWavesPerWG;
    for (int64_t w = blockIdx.x; w < n_groups; w += gridDim.x) {
        int64_t wi = xcd_item(w, n_groups);
        if (per_wave) {
            wi = wi * kWavesPerWG + wave;
            if (wi >= a.n_wave_items) continue;
        }
        const tnx_work_t it = a.work[wi];
        const int cnt = it.tgt_cnt;
        const int npair = cnt * (it.src_end - it.src_beg);
        // pair q = sj * cnt + ti; this wave takes q = pfirst, pfirst + pstep,
        // ... ((ti, sj) advance without division), and the next pair's
        // indices and level counts are loaded one pair ahead.
        int ti = 0, sj = 0;
        for (int u = 0; u < pfirst; u++)
            if (++ti == cnt) { ti = 0; ++sj; }
        int xn = 0, yn = 0, bxn = 0, byn = 0;
        if (pfirst < npair) {
            xn = a.tgt_order[it.tgt_beg + ti];
            yn = a.src_order[it.src_beg + sj];
            bxn = a.n_per_var[xn];
            byn = a.n_per_var[yn];
        }
        for (int q = pfirst; q < npair; q += pstep) {
            const int x = xn, y = yn, bx = bxn, by = byn;
            for (int u = 0; u < pstep; u++)
                if (++ti == cnt) { ti = 0; ++sj; }
            if (q + pstep < npair) {
                xn = a.tgt_order[it.tgt_beg + ti];
                yn = a.src_order[it.src_beg + sj];
                bxn = a.n_per_var[xn];
                byn = a.n_per_var[yn];
            }
            int64_t slot;
            const bool want = (x != y) && out_slot(a, x, y, slot);
            const bool fits = bx * bx * by <= a.cap_wave;
            if (want && !fits && lane == 0) emit_te(a, slot, __builtin_nan(""));  // plan mismatch
            if (!(want && fits)) continue;
            const WaveCopies cp = wave_copies(bx * bx * by, a.cap_wave, TNX_WAVE_COPY_BUDGET);
            uint32_t *hc = hist + (lane & (cp.G - 1));
            const uint32_t kmul = (uint32_t)by << (2 + cp.gsh), ysh = 2 + cp.gsh;
            const int8_t *yr = a.bins + (int64_t)y * a.ld;
            if (x != cur_x) {
                if constexpr (NI > 0) load_target_regs<NI>(regs, a.states + (int64_t)x * a.lds, L, lane);
                s12 = a.S1[x] - a.S2a[x];
                cur_x = x;
            }
            if constexpr (kPrefetch) {
                if (pf_y != y) load_source_chunks<NI>(ysrc, yr, L, lane);
                hist_count_chunks<NI, HZ>(hc, regs, ysrc, kmul, ysh, zero_row + lane, L, lane);
                pf_y = -1;
                if (q + pstep < npair) {
                    load_source_chunks<NI>(ysrc, a.bins + (int64_t)yn * a.ld, L, lane);
                    pf_y = yn;
                }
            } else if constexpr (NI > 0) {
                hist_accumulate_regs<NI, HZ>(hc, regs, yr, kmul, ysh, zero_row + lane, L, lane);
            } else {
                hist_accumulate_mem<HZ>(hc, a.states + (int64_t)x * a.lds, yr, kmul, ysh,
                                        zero_row + lane, L, lane);
            }
            wave_lds_fence();
            fold_zero_key<HZ>(hist, zero_row, lane);
            const DF d = hist_reduce_wave_copies(hist, bx, by, cp, XlTable{tbl}, lane);
            wave_lds_fence();
            if (lane == 0) emit_te(a, slot, (df_to_double(d) + s12) * invN);
        }
    }
}

template <bool HZ>
__global__ __launch_bounds__(256) void te_wave_kernel(tnx_te_args_t a) { te_wave_body<0, HZ>(a); }

template <int NI, bool HZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void te_wave_reg_kernel(
    tnx_te_args_t a) {
    te_wave_body<NI, HZ>(a);
}

// Workgroup-per-pair kernel for pairs above cap_wave cells (multi-pass, exact).
// The target's states are loaded into registers once per target and reused
// for every source and pass.  Requires L <= 16 * kBlockThreads * NB.
// Loads this thread's NB chunks of 16 source bytes (zero past L).
template <int NB>
__device__ __forceinline__ void load_block_row(uint4 *ys, const int8_t *__restrict__ yr, int L) {
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int t = 16 * threadIdx.x + 16 * kBlockThreads * i;
        ys[i] = t < L ? *reinterpret_cast<const uint4 *>(yr + t) : make_uint4(0u, 0u, 0u, 0u);
    }
}

template <int NB, bool HZ>
__global__ __launch_bounds__(kBlockThreads) void te_block_kernel(tnx_te_args_t a) {
    extern __shared__ uint32_t lds_u32[];
    __shared__ float red[2 * (kBlockThreads / kWave)];
    __shared__ float2 ltbl[kSmallTbl];  // c log2 c of small counts
    uint32_t *hist = lds_u32;
    uint32_t *c2b = lds_u32 + a.cap_block;
    uint32_t *zero_row = c2b + a.max_block_cols;
    const int tid = threadIdx.x;
    const int L = (int)(a.T - a.tau);
    const float2 *tbl = reinterpret_cast<const float2 *>(a.xlogx);
    for (int i = tid; i < kSmallTbl; i += kBlockThreads)
        ltbl[i] = i <= L ? tbl[i] : make_float2(0.f, 0.f);
    for (int i = tid; i < a.cap_block + a.max_block_cols + kWave; i += kBlockThreads)
        lds_u32[i] = 0;
    __syncthreads();
    const double invN = 1.0 / (double)L;
    const tnx_work_t *work = a.work + a.n_wave_items;
    TargetRegs<NB> regs;
    // long series (NB > 4) skip the row prefetch to stay within registers
    constexpr bool kPipe = NB <= 4;
    uint4 ycur[NB], ynext[kPipe ? NB : 1];
    int cur_x = -1;
    double s12 = 0.0;
    for (int64_t w = blockIdx.x; w < a.n_block_items; w += gridDim.x) {
        const tnx_work_t it = work[w];
        for (int ti = 0; ti < it.tgt_cnt; ti++) {
            const int x = a.tgt_order[it.tgt_beg + ti];
            const int bx = a.n_per_var[x];
            if (x != cur_x) {
                const uint16_t *st = a.states + (int64_t)x * a.lds;
#pragma unroll
                for (int i = 0; i < NB; i++) {
                    const int t = 16 * tid + 16 * kBlockThreads * i;
                    if (t < L) {
                        const uint4 u = *reinterpret_cast<const uint4 *>(st + t);
                        const uint4 v = *reinterpret_cast<const uint4 *>(st + t + 8);
                        regs.s[8 * i + 0] = u.x; regs.s[8 * i + 1] = u.y;
                        regs.s[8 * i + 2] = u.z; regs.s[8 * i + 3] = u.w;
                        regs.s[8 * i + 4] = v.x; regs.s[8 * i + 5] = v.y;
                        regs.s[8 * i + 6] = v.z; regs.s[8 * i + 7] = v.w;
                    }
                }
                s12 = a.S1[x] - a.S2a[x];
                cur_x = x;
            }
            // sources are software-pipelined: the next source's index, level
            // count and row are requested while the current pair is counted
            int y = a.src_order[it.src_beg];
            int by = a.n_per_var[y];
            if constexpr (kPipe) load_block_row<NB>(ycur, a.bins + (int64_t)y * a.ld, L);
            for (int sj = it.src_beg; sj < it.src_end; sj++) {
                int yn = 0, byn = 0;
                if (sj + 1 < it.src_end) {
                    yn = a.src_order[sj + 1];
                    byn = a.n_per_var[yn];
                    if constexpr (kPipe) load_block_row<NB>(ynext, a.bins + (int64_t)yn * a.ld, L);
                }
                if constexpr (!kPipe) load_block_row<NB>(ycur, a.bins + (int64_t)y * a.ld, L);
                int64_t slot;
                // pairs above cap_sort are left to the occupied-key kernel
                const bool want = (x != y) && !sort_pair(a, bx, by) && out_slot(a, x, y, slot);
                const bool fits = bx * by <= a.max_block_cols;
                if (want && !fits && tid == 0) emit_te(a, slot, __builtin_nan(""));  // plan mismatch
                if (want && fits) {
                    const DF d = hist_pair_block<NB, HZ>(hist, c2b, zero_row, a.cap_block, regs, ycur,
                                                     bx, by, L, XlSmall{ltbl, tbl}, red);
                    if (tid == 0) emit_te(a, slot, (df_to_double(d) + s12) * invN);
                }
                y = yn;
                by = byn;
                if constexpr (kPipe) {
#pragma unroll
                    for (int i = 0; i < NB; i++) ycur[i] = ynext[kPipe ? i : 0];
                }
            }
        }
    }
}

// Debug: counts of explicit pairs through the same accumulation code as the
// wave kernel tnx_te_run launches at this L (NI as in wave_variant()).
template <int NI>
__global__ __launch_bounds__(256) void te_counts_kernel(
    const int8_t *__restrict__ bins, const uint16_t *__restrict__ states,
    const int32_t *__restrict__ npv, const double *__restrict__ S1,
    const double *__restrict__ S2a, const float *__restrict__ xlogx, int64_t T, int64_t ld,
    int64_t lds, int32_t tau, const int32_t *__restrict__ pairs, int64_t n_pairs, int32_t cap,
    int32_t *__restrict__ counts, float *__restrict__ te_out) {
    extern __shared__ uint32_t lds_u32[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    uint32_t *hist = lds_u32 + wave * wave_slice(cap);
    uint32_t *zero_row = hist + cap;
    for (int i = lane; i < wave_slice(cap); i += kWave) hist[i] = 0;
    wave_lds_fence();
    const int L = (int)(T - tau);
    const float2 *tbl = reinterpret_cast<const float2 *>(xlogx);
    int64_t p = (int64_t)blockIdx.x * kWavesPerWG + wave;
    if (p >= n_pairs) return;
    const int x = pairs[2 * p], y = pairs[2 * p + 1];
    const int bx = npv[x], by = npv[y];
    if (bx * bx * by > cap) {
        if (lane == 0) te_out[p] = __builtin_nanf("");
        return;
    }
    const int cells = bx * bx * by;
    const WaveCopies cp = wave_copies(cells, cap, TNX_WAVE_COPY_BUDGET);
    uint32_t *hc = hist + (lane & (cp.G - 1));
    const uint32_t kmul = (uint32_t)by << (2 + cp.gsh), ysh = 2 + cp.gsh;
    if constexpr (NI > 0) {
        TargetRegs<NI> regs;
        load_target_regs<NI>(regs, states + (int64_t)x * lds, L, lane);
        hist_accumulate_regs<NI, true>(hc, regs, bins + (int64_t)y * ld, kmul, ysh, zero_row + lane,
                                       L, lane);
    } else {
        hist_accumulate_mem<true>(hc, states + (int64_t)x * lds, bins + (int64_t)y * ld, kmul,
                                  ysh, zero_row + lane, L, lane);
    }
    wave_lds_fence();
    fold_zero_key<true>(hist, zero_row, lane);
    for (int i = lane; i < cells; i += kWave) {
        uint32_t v = 0;
        for (int r = 0; r < cp.G; r++) v += hist[(i << cp.gsh) + r];
        counts[p * cap + i] = (int32_t)v;
    }
    wave_lds_fence();
    const DF d = hist_reduce_wave_copies(hist, bx, by, cp, XlTable{tbl}, lane);
    if (lane == 0) te_out[p] = (float)((df_to_double(d) + (S1[x] - S2a[x])) / (double)L);
}

// Register-variant choice for a series of L samples: the smallest
// instantiated NI with NI * 1024 >= L, 0 (memory variant) above 12288.
static int wave_variant(int64_t L) {
    static const int kNI[] = {1, 2, 3, 4, 6, 8, 10, 12};
    for (int ni : kNI)
        if (L <= (int64_t)ni * 1024) return ni;
    return 0;
}

}  // namespace tnx

using namespace tnx;

// One thread per slot, copies in order: the additions of K statistics runs.
__global__ __launch_bounds__(256) void stats_accumulate_kernel(
    const float *__restrict__ stage, int K, int64_t stride, const int64_t *__restrict__ slots,
    int64_t n_slots, const float *__restrict__ observed, double *__restrict__ sum,
    double *__restrict__ sumsq, int32_t *__restrict__ count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_slots) return;
    const int64_t sl = slots[i];
    const bool counting = count != nullptr;
    const float ob = counting ? observed[sl] : 0.f;
    double s = sum[sl], s2 = sumsq[sl];
    int32_t c = counting ? count[sl] : 0;
    for (int k = 0; k < K; k++) {
        const float tf = stage[(int64_t)k * stride + sl];
        const double td = (double)tf;
        s += td;
        s2 += td * td;
        c += (tf >= ob) ? 1 : 0;
    }
    sum[sl] = s;
    sumsq[sl] = s2;
    if (counting) count[sl] = c;
}

template <bool HZ>
static int launch_wave(int ni, unsigned grid, size_t shm, hipStream_t s, const tnx_te_args_t &a) {
    const void *kfn = (const void *)te_wave_kernel<HZ>;
    switch (ni) {
#define TNX_WAVE_FN(NI) \
    case NI: kfn = (const void *)te_wave_reg_kernel<NI, HZ>; break;
        TNX_WAVE_FN(1) TNX_WAVE_FN(2) TNX_WAVE_FN(3) TNX_WAVE_FN(4) TNX_WAVE_FN(6)
        TNX_WAVE_FN(8) TNX_WAVE_FN(10) TNX_WAVE_FN(12)
#undef TNX_WAVE_FN
        default: break;
    }
    TNX_CHECK_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    switch (ni) {
#define TNX_WAVE_LAUNCH(NI) \
    case NI: hipLaunchKernelGGL((te_wave_reg_kernel<NI, HZ>), dim3(grid), dim3(256), shm, s, a); break;
        TNX_WAVE_LAUNCH(1) TNX_WAVE_LAUNCH(2) TNX_WAVE_LAUNCH(3) TNX_WAVE_LAUNCH(4)
        TNX_WAVE_LAUNCH(6) TNX_WAVE_LAUNCH(8) TNX_WAVE_LAUNCH(10) TNX_WAVE_LAUNCH(12)
#undef TNX_WAVE_LAUNCH
        default: hipLaunchKernelGGL(te_wave_kernel<HZ>, dim3(grid), dim3(256), shm, s, a); break;
    }
    return TNX_OK;
}

template <bool HZ>
static int launch_block(int nbi, unsigned grid, size_t shm, hipStream_t s,
                         const tnx_te_args_t &a) {
    switch (nbi) {
#define TNX_BLOCK(NB)                                                                          \
    case NB:                                                                                   \
        TNX_CHECK_HIP(hipFuncSetAttribute((const void *)te_block_kernel<NB, HZ>,               \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm)); \
        hipLaunchKernelGGL((te_block_kernel<NB, HZ>), dim3(grid), dim3(kBlockThreads), shm, s, a); \
        break;
        TNX_BLOCK(1) TNX_BLOCK(2) TNX_BLOCK(4) TNX_BLOCK(8) TNX_BLOCK(16)
#undef TNX_BLOCK
        default: break;
    }
    return TNX_OK;
}

extern "C" {

int tnx_te_plan(const int32_t *npv, int64_t n_vars, int64_t L, const int32_t *tgt_list,
                int64_t n_tgt, const int32_t *src_list, int64_t n_src, int32_t cap_wave,
                int32_t cap_block, const int32_t *src_skip, int32_t *tgt_order,
                int32_t *src_order, tnx_work_t *work, int64_t work_cap, tnx_plan_info_t *info) {
    clear_error();
    TNX_CHECK_ARG(npv && info, "tnx_te_plan: null pointer");
    TNX_CHECK_ARG(n_vars >= 0 && n_vars < (1LL << 31), "tnx_te_plan: bad n_vars");
    if (cap_wave <= 0) cap_wave = kDefaultCapWave;
    std::vector<int32_t> tg, sr;
    if (tgt_list) tg.assign(tgt_list, tgt_list + n_tgt);
    else { tg.resize(n_vars); std::iota(tg.begin(), tg.end(), 0); }
    if (src_list) sr.assign(src_list, src_list + n_src);
    else { sr.resize(n_vars); std::iota(sr.begin(), sr.end(), 0); }
    for (int32_t v : tg) TNX_CHECK_ARG(v >= 0 && v < n_vars, "tnx_te_plan: target %d out of range", v);
    for (int32_t v : sr) TNX_CHECK_ARG(v >= 0 && v < n_vars, "tnx_te_plan: source %d out of range", v);
    for (int64_t v = 0; v < n_vars; v++)
        TNX_CHECK_ARG(npv[v] >= 1 && npv[v] <= 127, "tnx_te_plan: n_per_var[%lld]=%d not in [1,127]",
                      (long long)v, npv[v]);
    // Sources ascending by level count, targets descending (heavy first).
    std::stable_sort(sr.begin(), sr.end(), [&](int32_t a, int32_t b) { return npv[a] < npv[b]; });
    std::stable_sort(tg.begin(), tg.end(), [&](int32_t a, int32_t b) { return npv[a] > npv[b]; });
    const int64_t nt = (int64_t)tg.size(), ns = (int64_t)sr.size();
    std::vector<int32_t> sb(ns);
    for (int64_t j = 0; j < ns; j++) sb[j] = npv[sr[j]];
    // kw[i]: number of sources handled by the wave kernel for target tg[i];
    // sk[i]: sources left to a one-hot plan (a prefix of the sorted sources).
    std::vector<int64_t> kw(nt), sk(nt, 0);
    for (int64_t i = 0; i < nt; i++) {
        int64_t bx = npv[tg[i]];
        int64_t lim = cap_wave / (bx * bx);
        kw[i] = std::upper_bound(sb.begin(), sb.end(), (int32_t)std::min<int64_t>(lim, 1 << 30)) -
                sb.begin();
        if (src_skip) {
            TNX_CHECK_ARG(src_skip[tg[i]] >= 0 && src_skip[tg[i]] <= ns,
                          "tnx_te_plan: src_skip[%d] = %d not in [0, %lld]", tg[i], src_skip[tg[i]],
                          (long long)ns);
            sk[i] = src_skip[tg[i]];
        }
    }
    // Widest (x_t, y_t) marginal among block-kernel pairs: a pass of the
    // block kernel holds at least one value of x_{t+tau}, i.e. b_x * b_y cells.
    int32_t maxcols = 0;
    for (int64_t i = 0; i < nt; i++)
        if (std::max(kw[i], sk[i]) < ns) maxcols = std::max<int32_t>(maxcols, npv[tg[i]] * sb[ns - 1]);
    if (cap_block <= 0) {
        cap_block = std::min(kDefaultCapBlock, kBlockLdsTarget / 4 - maxcols) & ~3;
        cap_block = std::max(cap_block, (maxcols + 3) & ~3);
    }
    TNX_CHECK_ARG(cap_block >= maxcols && cap_block >= 4 && cap_block % 4 == 0,
                  "tnx_te_plan: cap_block %d must be a multiple of 4 and >= the widest block "
                  "pair's b_t*b_s = %d", cap_block, maxcols);
    std::vector<tnx_work_t> wv, bk;
    double cw = 0, cb = 0;
    int64_t pw = 0, pb = 0;
    for (int64_t i = 0; i < nt; i++) {
        pw += std::max<int64_t>(0, kw[i] - sk[i]);
        pb += ns - std::max(kw[i], sk[i]);
    }
    // pairs per item: as many as kWaveItemPairs / kBlockItemPairs, fewer when
    // the plan is too small to fill the chip otherwise
    const int64_t wave_pairs = std::clamp<int64_t>((pw + kFillWaveItems - 1) / kFillWaveItems,
                                                   kWavesPerWG, kWaveItemPairs);
    const int64_t block_pairs = std::clamp<int64_t>((pb + kFillBlockItems - 1) / kFillBlockItems,
                                                    1, kBlockItemPairs);
    auto wave_range = [&](int64_t i, int64_t j0, int64_t j1) {
        for (int64_t j = j0; j < j1; j += wave_pairs)
            wv.push_back({(int32_t)i, 1, (int32_t)j, (int32_t)std::min(j1, j + wave_pairs)});
    };
    for (int64_t g0 = 0; g0 < nt; g0 += kWavesPerWG) {
        int64_t g1 = std::min<int64_t>(nt, g0 + kWavesPerWG);
        int32_t cnt = (int32_t)(g1 - g0);
        // the group's common wave range [smax, kmin); each target's own
        // sources outside it go to one-target items
        int64_t kmin = ns, smax = 0;
        for (int64_t i = g0; i < g1; i++) {
            kmin = std::min(kmin, kw[i]);
            smax = std::max(smax, sk[i]);
        }
        int64_t ch = std::max<int64_t>(1, wave_pairs / cnt);
        for (int64_t j = smax; j < kmin; j += ch)
            wv.push_back({(int32_t)g0, cnt, (int32_t)j, (int32_t)std::min(kmin, j + ch)});
        for (int64_t i = g0; i < g1; i++) {
            int64_t bx = npv[tg[i]];
            if (smax < kmin) {
                wave_range(i, sk[i], smax);
                wave_range(i, kmin, kw[i]);
            } else {
                wave_range(i, sk[i], kw[i]);
            }
            const int64_t b0 = std::max(kw[i], sk[i]);
            for (int64_t j = sk[i]; j < kw[i]; j++) cw += (double)L + (double)(bx * bx * sb[j]);
            for (int64_t j = b0; j < ns; j += block_pairs)
                bk.push_back({(int32_t)i, 1, (int32_t)j, (int32_t)std::min<int64_t>(ns, j + block_pairs)});
            for (int64_t j = b0; j < ns; j++) {
                int64_t cols = bx * sb[j];
                int64_t passes = (bx * cols + cap_block - 1) / cap_block;
                cb += (double)L * (double)passes + (double)(bx * cols);
            }
        }
    }
    // L2 locality: the wave kernel hands item ranges to the 8 XCDs in
    // contiguous slices (xcd_item), so order its items by source chunk,
    // chunks dealt to slices round-robin: each XCD's L2 then holds the
    // source rows its items share, and every slice mixes cheap and dear
    // sources.  Block items keep the heavy-target-first order, which
    // balances their far more uneven costs across the chip.
    auto by_chunk = [](const tnx_work_t &u, const tnx_work_t &v) {
        const int cu = u.src_beg / kChunkSrc, cv = v.src_beg / kChunkSrc;
        if (cu % kXcds != cv % kXcds) return cu % kXcds < cv % kXcds;
        if (cu != cv) return cu < cv;
        return u.tgt_beg < v.tgt_beg;
    };
    std::stable_sort(wv.begin(), wv.end(), by_chunk);
    // pairs (excluding x == y and the skipped sources)
    std::vector<int64_t> src_pos(n_vars, -1);
    for (int64_t j = 0; j < ns; j++) src_pos[sr[j]] = j;
    int64_t kept = 0;
    for (int64_t i = 0; i < nt; i++) {
        kept += ns - sk[i];
        if (src_pos[tg[i]] >= sk[i]) kept -= 1;
    }
    info->n_wave_items = (int64_t)wv.size();
    info->n_block_items = (int64_t)bk.size();
    info->n_pairs = kept;
    info->cap_wave = cap_wave;
    info->cap_block = cap_block;
    info->max_block_cols = maxcols;
    info->n_tgt = (int32_t)nt;
    info->n_src = (int32_t)ns;
    info->est_cost_wave = cw;
    info->est_cost_block = cb;
    if (!work) return TNX_OK;  // query mode
    if ((int64_t)(wv.size() + bk.size()) > work_cap) {
        set_error("tnx_te_plan: work_cap %lld < %lld items", (long long)work_cap,
                  (long long)(wv.size() + bk.size()));
        return TNX_E_CAP;
    }
    TNX_CHECK_ARG(tgt_order && src_order, "tnx_te_plan: null order buffers");
    memcpy(tgt_order, tg.data(), sizeof(int32_t) * nt);
    memcpy(src_order, sr.data(), sizeof(int32_t) * ns);
    memcpy(work, wv.data(), sizeof(tnx_work_t) * wv.size());
    memcpy(work + wv.size(), bk.data(), sizeof(tnx_work_t) * bk.size());
    return TNX_OK;
}

int tnx_te_run(const tnx_te_args_t *args, void *stream) {
    clear_error();
    TNX_CHECK_ARG(args != nullptr, "tnx_te_run: args is NULL");
    const tnx_te_args_t &a = *args;
    TNX_CHECK_ARG(a.bins && a.states && a.n_per_var && a.S1 && a.S2a && a.xlogx && a.tgt_order &&
                      a.src_order,
                  "tnx_te_run: null pointer");
    TNX_CHECK_ARG(a.out_mode == TNX_OUT_STATS
                      ? (a.sum_te && a.sumsq_te && a.count_ge && a.observed &&
                         a.ld_out >= a.n_vars)
                      : (a.out != nullptr),
                  "tnx_te_run: null output pointer for out_mode %d", a.out_mode);
    TNX_CHECK_ARG(a.tau >= 1 && a.tau < a.T, "tnx_te_run: need 1 <= tau < T");
    TNX_CHECK_ARG(a.ld >= a.T && a.ld % 16 == 0, "tnx_te_run: ld must be >= T and a multiple of 16");
    TNX_CHECK_ARG(a.lds >= a.T - a.tau && a.lds % 16 == 0,
                  "tnx_te_run: lds must be >= T - tau and a multiple of 16");
    TNX_CHECK_ARG(a.T - a.tau < (1LL << 31), "tnx_te_run: L too large");
    TNX_CHECK_ARG(a.out_mode == TNX_OUT_MATRIX || a.out_mode == TNX_OUT_PAIRFREE ||
                      a.out_mode == TNX_OUT_STATS || a.out_mode == TNX_OUT_STAGE,
                  "tnx_te_run: bad out_mode");
    TNX_CHECK_ARG(a.out_mode != TNX_OUT_STAGE ||
                      (a.stage_vars >= 1 && a.ld_out >= a.stage_vars &&
                       a.stage_stride >= (int64_t)a.stage_vars * a.ld_out),
                  "tnx_te_run: TNX_OUT_STAGE needs stage_vars >= 1, ld_out >= stage_vars and "
                  "stage_stride >= stage_vars * ld_out");
    TNX_CHECK_ARG(a.cap_wave >= 8 && a.cap_wave <= 8192, "tnx_te_run: cap_wave out of range");
    TNX_CHECK_ARG(a.cap_sort == 0 || a.cap_sort > a.cap_wave,
                  "tnx_te_run: cap_sort must be 0 or exceed cap_wave");
    TNX_CHECK_ARG((size_t)kWavesPerWG * wave_slice(a.cap_wave) * 4 <= 160 * 1024,
                  "tnx_te_run: cap_wave too large");
    hipStream_t s = as_stream(stream);
    if (a.queue_items != nullptr) {
        if (a.n_wave_items + a.n_block_items == 0) return TNX_OK;
        const int ni = wave_variant(a.T - a.tau);
        return launch_persist(a, ni, s);
    }
    TNX_CHECK_ARG(a.counts_map == nullptr, "tnx_te_run: counts_map needs the persistent launch");
    TNX_CHECK_ARG(a.var_L == nullptr, "tnx_te_run: var_L needs the persistent launch");
    if (a.n_wave_items > 0) {
        TNX_CHECK_ARG(a.work, "tnx_te_run: null work");
        size_t shm = (size_t)kWavesPerWG * wave_slice(a.cap_wave) * 4;
        const int ni = wave_variant(a.T - a.tau);
        const bool per_wave = a.item_per_wave && ni > 0 && ni <= 4;  // as te_wave_body
        const int64_t groups =
            per_wave ? (a.n_wave_items + kWavesPerWG - 1) / kWavesPerWG : a.n_wave_items;
        unsigned grid = (unsigned)std::min<int64_t>(groups, 1 << 20);
        const int rc = a.hot_zero ? launch_wave<true>(ni, grid, shm, s, a)
                                  : launch_wave<false>(ni, grid, shm, s, a);
        if (rc != TNX_OK) return rc;
        TNX_CHECK_LAUNCH("te_wave_kernel");
    }
    if (a.n_block_items > 0) {
        TNX_CHECK_ARG(a.work, "tnx_te_run: null work");
        TNX_CHECK_ARG(a.max_block_cols >= 1 && a.max_block_cols <= 127 * 127,
                      "tnx_te_run: bad max_block_cols");
        TNX_CHECK_ARG(a.cap_block >= a.max_block_cols && a.cap_block % 4 == 0,
                      "tnx_te_run: cap_block %d must be a multiple of 4 and >= max_block_cols %d",
                      a.cap_block, a.max_block_cols);
        size_t shm = (size_t)(a.cap_block + a.max_block_cols + kWave) * 4;
        TNX_CHECK_ARG(shm + 8 * kSmallTbl + 256 <= 160 * 1024,
                      "tnx_te_run: block kernel LDS %zu exceeds 160 KiB", shm);
        unsigned grid = (unsigned)std::min<int64_t>(a.n_block_items, 1 << 20);
        const int64_t L = a.T - a.tau;
        const int nb = (int)((L + 16 * kBlockThreads - 1) / (16 * kBlockThreads));
        TNX_CHECK_ARG(nb <= 16, "tnx_te_run: L = %lld too long for the block kernel (max %d)",
                      (long long)L, 16 * 16 * kBlockThreads);
        const int nbi = nb <= 1 ? 1 : nb <= 2 ? 2 : nb <= 4 ? 4 : nb <= 8 ? 8 : 16;
        const int rc = a.hot_zero ? launch_block<true>(nbi, grid, shm, s, a)
                                  : launch_block<false>(nbi, grid, shm, s, a);
        if (rc != TNX_OK) return rc;
        TNX_CHECK_LAUNCH("te_block_kernel");
    }
    return TNX_OK;
}

int tnx_te_counts_debug(const int8_t *bins, const uint16_t *states, const int32_t *n_per_var,
                        const double *S1, const double *S2a, const float *xlogx, int64_t n_vars,
                        int64_t T, int64_t ld, int64_t lds, int32_t tau, const int32_t *pairs,
                        int64_t n_pairs, int32_t max_cells, int32_t *counts_out, float *te_out,
                        void *stream) {
    clear_error();
    TNX_CHECK_ARG(bins && states && n_per_var && S1 && S2a && xlogx && pairs && counts_out && te_out,
                  "tnx_te_counts_debug: null pointer");
    TNX_CHECK_ARG(max_cells >= 8 && max_cells <= 8192 && max_cells % 2 == 0, "tnx_te_counts_debug: max_cells in [8, 8192]");
    TNX_CHECK_ARG(ld % 16 == 0 && lds % 16 == 0 && lds >= T - tau && tau >= 1 && tau < T,
                  "tnx_te_counts_debug: bad strides / tau");
    if (n_pairs == 0) return TNX_OK;
    size_t shm = (size_t)kWavesPerWG * wave_slice(max_cells) * 4;
    unsigned grid = (unsigned)((n_pairs + kWavesPerWG - 1) / kWavesPerWG);
    const int ni = wave_variant(T - tau);
    switch (ni) {
#define TNX_COUNTS(NI)                                                                         \
    case NI:                                                                                   \
        TNX_CHECK_HIP(hipFuncSetAttribute((const void *)te_counts_kernel<NI>,                  \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm)); \
        hipLaunchKernelGGL(te_counts_kernel<NI>, dim3(grid), dim3(256), shm, as_stream(stream),  \
                           bins, states, n_per_var, S1, S2a, xlogx, T, ld, lds, tau, pairs,      \
                           n_pairs, max_cells, counts_out, te_out);                            \
        break;
        TNX_COUNTS(0) TNX_COUNTS(1) TNX_COUNTS(2) TNX_COUNTS(3) TNX_COUNTS(4) TNX_COUNTS(6)
        TNX_COUNTS(8) TNX_COUNTS(10) TNX_COUNTS(12)
#undef TNX_COUNTS
        default: break;
    }
    TNX_CHECK_LAUNCH("te_counts_kernel");
    return TNX_OK;
}

int tnx_stats_accumulate(const float *stage, int32_t K, int64_t stage_stride,
                         const int64_t *slots, int64_t n_slots, const float *observed,
                         double *sum_te, double *sumsq_te, int32_t *count_ge, void *stream) {
    clear_error();
    TNX_CHECK_ARG(K >= 0 && n_slots >= 0 && stage_stride >= 1, "tnx_stats_accumulate: bad sizes");
    if (K == 0 || n_slots == 0) return TNX_OK;
    TNX_CHECK_ARG(stage && slots && sum_te && sumsq_te && (count_ge == nullptr || observed),
                  "tnx_stats_accumulate: null pointer");
    const unsigned grid = (unsigned)((n_slots + 255) / 256);
    hipLaunchKernelGGL(stats_accumulate_kernel, dim3(grid), dim3(256), 0, as_stream(stream), stage,
                       (int)K, stage_stride, slots, n_slots, observed, sum_te, sumsq_te,
                       count_ge);
    TNX_CHECK_LAUNCH("stats_accumulate_kernel");
    return TNX_OK;
}

}  // extern "C"
