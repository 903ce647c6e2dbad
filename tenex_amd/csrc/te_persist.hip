// te_persist.hip -- every pair of a plan in one persistent launch.
//
// One grid of resident workgroups (256 threads, four waves, three per CU)
// takes entries from sixteen work queues (a wave queue and a block queue per
// XCD) with an atomic take counter each.  A workgroup drains the queues of
// its own XCD first (the planner deals entries
// to queues by source chunk, so an XCD's workgroups share source rows in its
// L2), then the other queues, so the launch ends when the last entry does.
// An entry is either
//   a wave item: the four waves split its pairs, one pair per wave at a time,
//                each wave in a private LDS slice (te_hist.hip's wave kernel
//                layout), or
//   a block item: the whole workgroup on one pair at a time, the four
//                slices as one histogram, several passes over ranges of
//                x_{t+tau} when the pair has more cells.
// Both kinds share CUs, so the latency-bound block pairs (barriers, scans)
// overlap the LDS-atomic-bound wave pairs instead of running after them.
//
// Per wave pair the source row is read through a buffer resource in groups
// of 1024-sample chunks that are software-pipelined across pairs: before a
// group is counted the next one (after the last group, the next pair's
// first) is requested.  The reduction reads each column's cells four at a
// time and takes c*log2(c) of counts below 256 from an LDS copy of the
// table.  Counts are those of the two-launch kernels (te_wave_reg_kernel /
// te_block_kernel); S3 - S2b is an exact 64-bit fixed-point sum of the
// table's entries, so TE does not depend on the order of the additions.
//
// Replaces the pair-free launches te_smem_launch_pairfree
// (tenex/csrc/full_smem.cu:350-381) and te_adaptive_smem_launch_pairfree
// (tenex/csrc/adaptive_smem.cu:351-376) for series of up to 12,288 samples.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "te_args.hpp"
#include "te_device.hpp"

namespace tnx {

constexpr int kQueueStride = 16;  // take counters 64 bytes apart
static_assert(2 * TNX_QUEUES * kQueueStride <= TNX_QUEUE_CTR_INTS, "queue counters");
constexpr int kPersistThreads = kWavesPerWG * kWave;

__global__ void queue_reset_kernel(int32_t *ctr) {
    if (threadIdx.x < TNX_QUEUE_CTR_INTS) ctr[threadIdx.x] = 0;
}

// Source chunks [C0, C0 + CN) of one pair for lane `lane`: chunk i covers
// bytes [16 lane + 1024 i, +16) of the row (rsrc spans the padded row, whose
// bytes past T are zero; chunks past the row read zero).
template <int CN>
__device__ __forceinline__ void load_group(uint4 *ys, __amdgpu_buffer_rsrc_t yr, int C0,
                                           int lane) {
#pragma unroll
    for (int u = 0; u < CN; u++) ys[u] = row_load16(yr, 16u * lane, 1024u * (C0 + u));
}

// Target states of a register-resident wave pass (load_target_regs through
// a buffer resource: zero past the padded row).
template <int NI>
__device__ __forceinline__ void load_target_buf(TargetRegs<NI> &r, __amdgpu_buffer_rsrc_t st,
                                                int L, int lane) {
#pragma unroll
    for (int i = 0; i < NI; i++) {
        const int t = 16 * lane + 1024 * i;
        uint4 a = make_uint4(0u, 0u, 0u, 0u), b = a;  // state 0 past L
        if (1024 * i < L) {  // wave-uniform; lanes past L read the zero padding or zero
            a = row_load16(st, 32u * lane, 2048u * i);
            b = row_load16(st, 32u * lane + 16u, 2048u * i);
        }
        (void)t;
        r.s[8 * i + 0] = a.x; r.s[8 * i + 1] = a.y; r.s[8 * i + 2] = a.z; r.s[8 * i + 3] = a.w;
        r.s[8 * i + 4] = b.x; r.s[8 * i + 5] = b.y; r.s[8 * i + 6] = b.z; r.s[8 * i + 7] = b.w;
    }
}

// Hot target states (HS): a target's states (a, b) = (x_{t+tau}, x_t) with
// a < 2 and b < 2 hold most samples of zero-inflated data, so many lanes of
// a wave would add to the same few cells.  Their samples go to lane-private
// copies instead: in the target's registers such a state a * b_x + b
// becomes 0x8000 | (h << csh), h = 2a + b, and source level c of it is
// counted in copy r = lane mod C of hot cell (c, h), word (4c + h) C + r of
// the hot cells (csh = 2 + log2 C, so h << csh is its byte offset).  Other
// states stay below 0x8000 (b_x <= 127).  The copies are summed into the
// histogram before the reduction, so the counts are unchanged.
__device__ __forceinline__ uint32_t hot_encode16(uint32_t v, uint32_t bx, uint32_t csh) {
    const uint32_t d = v - bx;  // a = 1: d = b
    const uint32_t h = v < 2u ? v : (d < 2u ? 2u + d : 4u);
    return h < 4u ? (0x8000u | (h << csh)) : v;
}

template <int NR>
__device__ __forceinline__ void hot_encode_regs(uint32_t (&s)[NR], int bx, uint32_t csh) {
#pragma unroll
    for (int k = 0; k < NR; k++)
        s[k] = hot_encode16(s[k] & 0xFFFFu, (uint32_t)bx, csh) |
               (hot_encode16(s[k] >> 16, (uint32_t)bx, csh) << 16);
}

// Byte offset of sample (state sv, source level yv) from the lane's
// histogram base: the cell (kmul = 4 G b_y, ysh = 2 + log2 G), or for a hot
// state its hot copy (hmul = 16 C bytes per source level; hadj = the lane's
// copy offset from the base, less the 0x8000 flag that sv carries).
struct HotK {
    uint32_t hmul, hadj;
};
template <bool HS>
__device__ __forceinline__ uint32_t key_off(uint32_t sv, uint32_t yv, uint32_t kmul,
                                            uint32_t ysh, const HotK &hk) {
    const uint32_t off = __umul24(sv, kmul) + (yv << ysh);
    if constexpr (!HS) return off;
    const uint32_t hof = __umul24(yv, hk.hmul) + sv + hk.hadj;
    return sv >= 0x8000u ? hof : off;
}

// Count chunks [C0, C0 + CN) of the pair (target states in regs, source
// chunks in ys; C0 is a constant once the caller's group loop is unrolled)
// into the privatised histogram (with HS, hot states into their copies).
// Only the chunk holding sample L - 1 can be partial; its samples past L
// are not predicated one by one (per-sample exec masks cost scalar
// registers the register-resident target cannot spare) but sent by a select
// to `trash`, a lane-private word the caller clears.  Chunks wholly past L,
// or past NI, are skipped (wave-uniform).
template <int NI, int CN, bool HS>
__device__ __forceinline__ void count_group(uint32_t *hist, TargetRegs<NI> &r, const uint4 *ys,
                                            int C0, uint32_t kmul, uint32_t ysh, const HotK &hk,
                                            uint32_t trash_off, int L, int lane) {
#pragma unroll
    for (int k = 0; k < NI * 8; k++) asm volatile("" : "+v"(r.s[k]));
    const uint32_t *sreg = r.s;
#pragma unroll
    for (int u = 0; u < CN; u++) {
        const int i = C0 + u;
        if (i >= NI || 1024 * i >= L) break;  // wave-uniform
        const uint32_t yw[4] = {ys[u].x, ys[u].y, ys[u].z, ys[u].w};
        if (1024 * i + 1024 <= L) {
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const uint32_t sv = (sreg[8 * i + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
                const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                lds_inc(reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(hist) +
                                                     key_off<HS>(sv, yv, kmul, ysh, hk)));
            }
        } else {
            // samples past L (state 0 past L, but the source bytes between
            // L and T are real) go to the trash word
            const int nv = L - (16 * lane + 1024 * i);  // valid samples of this lane's 16
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const uint32_t sv = (sreg[8 * i + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
                const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                const uint32_t off = key_off<HS>(sv, yv, kmul, ysh, hk);
                lds_inc(reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(hist) +
                                                     (q < nv ? off : trash_off)));
            }
        }
    }
}

// Source-row pipeline of a wave pair: NG groups of GS chunks in two
// alternating buffers; before group g is counted, group g + 1 (after the
// last group: group 0 of the next pair) is requested into the other buffer.
// NG is even, so a pair always starts in buffer 0.  Long series use four
// groups to keep the in-flight chunks (two groups) within registers.
#ifndef TNX_PERSIST_NG_LONG
#define TNX_PERSIST_NG_LONG 4  // source groups per pair for NI >= 10
#endif
template <int NI>
struct SrcPipe {
    static constexpr int NG = NI >= 10 ? TNX_PERSIST_NG_LONG : 2;
    static constexpr int GS = (NI + NG - 1) / NG;
};

template <int NI, int G, bool HS>
__device__ __forceinline__ void pipe_count(uint4 (&b0)[SrcPipe<NI>::GS],
                                           uint4 (&b1)[SrcPipe<NI>::GS], TargetRegs<NI> &regs,
                                           __amdgpu_buffer_rsrc_t yr, __amdgpu_buffer_rsrc_t yr_next,
                                           bool more, bool run,
                                           uint32_t *hc, uint32_t kmul, uint32_t ysh,
                                           const HotK &hk, uint32_t trash_off, int L, int lane) {
    constexpr int NG = SrcPipe<NI>::NG, GS = SrcPipe<NI>::GS;
    if constexpr (G < NG) {
        uint4(&cur)[GS] = (G & 1) ? b1 : b0;
        uint4(&nxt)[GS] = (G & 1) ? b0 : b1;
        if constexpr (G + 1 < NG) {
            load_group<GS>(nxt, yr, (G + 1) * GS, lane);
        } else {
            if (more) load_group<GS>(nxt, yr_next, 0, lane);
        }
        if (run) count_group<NI, GS, HS>(hc, regs, cur, G * GS, kmul, ysh, hk, trash_off, L, lane);
        pipe_count<NI, G + 1, HS>(b0, b1, regs, yr, yr_next, more, run, hc, kmul, ysh, hk,
                                  trash_off, L, lane);
    }
}

// c*log2(c) as 64-bit fixed point (tnx_xlogx_table_fx): counts below
// kSmallTbl from the workgroup's LDS copy, read by every lane without a
// branch; the rare larger counts from the table in memory.  Sums of these
// integers are exact, so a pair's S3 - S2b does not depend on the order
// the cells are visited in (wave or block role, any launch).
struct XfLookup {
    const int64_t *lfx;          // LDS, kSmallTbl entries
    __amdgpu_buffer_rsrc_t gfx;  // the N + 1 entries in memory
};

// Sum of the fixed-point c log2 c of N counts: the LDS reads for all of
// them; then, only if a lane of the wave holds a count >= kSmallTbl, one
// buffer load per count (a small count's offset lies past the table and
// reads zero) and a single wait.  Branch-free per lane; a buffer load
// cannot be merged with the LDS read into a generic (flat) load.
template <int N>
__device__ __forceinline__ int64_t xf_sum(const XfLookup &xf, const uint32_t (&c)[N]) {
    // every count reads entry min(c, kSmallTbl - 1) from LDS (one min, no
    // compare and select); a count >= kSmallTbl then swaps that entry for
    // its own, read from memory
    int64_t s = 0;
    uint32_t any = 0;
#pragma unroll
    for (int u = 0; u < N; u++) {
        s += xf.lfx[min(c[u], (uint32_t)kSmallTbl - 1u)];
        any |= c[u];
    }
    if (__builtin_amdgcn_ballot_w64(any >= (uint32_t)kSmallTbl)) {  // wave-uniform
        int64_t w[N];
#pragma unroll
        for (int u = 0; u < N; u++)
            w[u] = (int64_t)__builtin_amdgcn_raw_buffer_load_b64(
                xf.gfx, (int)(c[u] < (uint32_t)kSmallTbl ? 0x80000000u : c[u] * 8u), 0, 0);
#pragma unroll
        for (int u = 0; u < N; u++)
            s += c[u] < (uint32_t)kSmallTbl ? 0 : w[u] - xf.lfx[kSmallTbl - 1];
    }
    return s;
}

__device__ __forceinline__ int64_t xf_one(const XfLookup &xf, uint32_t c) {
    const uint32_t cc[1] = {c};
    return xf_sum<1>(xf, cc);
}

// Sum of an int64 over the 64 lanes (every lane gets it): rotate-and-add
// inside each row of 16 lanes on the two 32-bit halves (DPP), then the four
// row totals.
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
    uint64_t u = (uint64_t)v;
#define TNX_ROW_ROR_ADD(N)                                                                      \
    {                                                                                          \
        const uint32_t lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x120 + N, 0xF,   \
                                                        0xF, false);                            \
        const uint32_t hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x120 + N, \
                                                        0xF, 0xF, false);                       \
        u += ((uint64_t)hi << 32) | lo;                                                        \
    }
    TNX_ROW_ROR_ADD(1) TNX_ROW_ROR_ADD(2) TNX_ROW_ROR_ADD(4) TNX_ROW_ROR_ADD(8)
#undef TNX_ROW_ROR_ADD
    auto lane_u64 = [&](int l) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
        return ((uint64_t)hi << 32) | lo;
    };
    return (int64_t)(lane_u64(0) + lane_u64(16) + lane_u64(32) + lane_u64(48));
}

// S3 - S2b of the pair in the wave's slice, exact in fixed point (every
// lane): with copies, lane l first folds cells k = l, l + 64, ... (the G
// copies of cell k read in a rotated order that spreads a half-wave over
// the banks) and writes their totals after the copies; then per (x_t, y_t)
// column the b_x totals are read four at a time (their latencies overlap)
// and give S2b (without copies the same pass gives S3).  Leaves the slice
// zero.
__device__ __forceinline__ int64_t reduce_wave_fx(uint32_t *hist, int bx, int by,
                                                  const WaveCopies &cp, const XfLookup &xf,
                                                  int lane) {
    const int ncol = bx * by;
    const int cells = bx * ncol;
    int64_t a3 = 0, a2 = 0;
    uint32_t *col_base = hist;  // where the column pass reads the cell totals
    if (cp.G > 1) {
        uint32_t *tot = hist + cp.G * cells;
        const int per_sh = TNX_LDS_BANK_SHIFT - cp.gsh;
        // four cells per step, so that their folds and lookups overlap
        for (int k0 = lane; k0 < cells; k0 += 4 * kWave) {
            uint32_t t[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int k = k0 + u * kWave;
                t[u] = 0;
                if (k < cells) {
                    if (cp.G >= 4) {
                        const int G4 = cp.G >> 2;
                        const int f = (k >> per_sh) & (G4 - 1);
                        uint4 *row = reinterpret_cast<uint4 *>(hist + (k << cp.gsh));
                        for (int j = 0; j < G4; j++) {
                            uint4 *q = row + ((j + f) & (G4 - 1));
                            const uint4 v = *q;
                            *q = make_uint4(0u, 0u, 0u, 0u);
                            t[u] += (v.x + v.y) + (v.z + v.w);
                        }
                    } else {
                        const int f = k >> per_sh;
                        uint32_t *row = hist + (k << cp.gsh);
                        for (int j = 0; j < cp.G; j++)
                            t[u] += lds_take(row + ((j + f) & (cp.G - 1)));
                    }
                    tot[k] = t[u];
                }
            }
            a3 += xf_sum<4>(xf, t);
        }
        wave_lds_fence();
        col_base = tot;
    }
    const bool s3_here = cp.G == 1;  // without copies the column pass also gives S3
    // four columns per step; each column's b_x cells read four at a time
    for (int c0 = lane; c0 < ncol; c0 += 4 * kWave) {
        uint32_t c2[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int col = c0 + u * kWave;
            c2[u] = 0;
            if (col >= ncol) continue;
            int aa = 0;
            for (; aa + 4 <= bx; aa += 4) {
                uint32_t v[4];
#pragma unroll
                for (int w = 0; w < 4; w++) v[w] = lds_take(col_base + (aa + w) * ncol + col);
                c2[u] += (v[0] + v[1]) + (v[2] + v[3]);
                if (s3_here) a3 += xf_sum<4>(xf, v);
            }
            for (; aa < bx; aa++) {
                const uint32_t v = lds_take(col_base + aa * ncol + col);
                c2[u] += v;
                if (s3_here) a3 += xf_one(xf, v);
            }
        }
        a2 += xf_sum<4>(xf, c2);
    }
    return wave_sum_i64(a3 - a2);
}

// Test hook: this pair's folded joint counts (the G copies summed) to
// counts_dbg, read-only on the slice (one wavefront).
__device__ __forceinline__ void dump_wave_counts(const tnx_te_args_t &a, int x, int y,
                                                 const uint32_t *hist, int cells,
                                                 const WaveCopies &cp, int lane) {
    const int r = a.counts_map[(int64_t)x * a.n_vars + y];
    if (r < 0 || cells > a.counts_ld) return;
    int32_t *dst = a.counts_dbg + (int64_t)r * a.counts_ld;
    for (int k = lane; k < cells; k += kWave) {
        uint32_t v = 0;
        for (int j = 0; j < cp.G; j++) v += hist[(k << cp.gsh) + j];
        dst[k] = (int32_t)v;
    }
}

// The pairs q = pfirst, pfirst + pstep, ... of a wave item (pair q is
// target tgt_order[tgt_beg + q % cnt] against source src_order[src_beg +
// q / cnt]), one wave.
template <int NI, bool HS>
__device__ __forceinline__ void persist_wave_item(const tnx_te_args_t &a, const tnx_work_t it,
                                                  int pfirst, int pstep, uint32_t *hist,
                                                  uint32_t *zero_row, const XfLookup &xf, int L,
                                                  double invN, int lane) {
    constexpr int GS = SrcPipe<NI>::GS;
    const int cnt = it.tgt_cnt;
    const int npair = cnt * (it.src_end - it.src_beg);
    if (pfirst >= npair) return;
    int ti = 0, sj = 0;
    for (int u = 0; u < pfirst; u++)
        if (++ti == cnt) { ti = 0; ++sj; }
    int x = a.tgt_order[it.tgt_beg + ti];
    int y = a.src_order[it.src_beg + sj];
    int bx = a.n_per_var[x], by = a.n_per_var[y];
    TargetRegs<NI> regs;
    int cur_x = -1;
    double s12 = 0.0;
    int Lx = L;            // the current target's samples (var_L) and normaliser
    double invNx = invN;
    const uint32_t row_bytes = (uint32_t)a.ld;
    uint4 b0[GS], b1[GS];
    load_group<GS>(b0, row_rsrc(a.bins + (int64_t)y * a.ld, row_bytes), 0, lane);
    for (int q = pfirst; q < npair; q += pstep) {
        for (int u = 0; u < pstep; u++)
            if (++ti == cnt) { ti = 0; ++sj; }
        const bool more = q + pstep < npair;
        int xn = x, yn = y, bxn = bx, byn = by;
        if (more) {
            xn = a.tgt_order[it.tgt_beg + ti];
            yn = a.src_order[it.src_beg + sj];
            bxn = a.n_per_var[xn];
            byn = a.n_per_var[yn];
        }
        if (x != cur_x) {
            if (a.var_L) {
                Lx = min(max(a.var_L[x], 1), L);
                invNx = 1.0 / (double)Lx;
            }
            load_target_buf<NI>(regs, row_rsrc(a.states + (int64_t)x * a.lds, 2u * (uint32_t)a.lds),
                                Lx, lane);
            if constexpr (HS) hot_encode_regs(regs.s, bx, 2u + (uint32_t)__builtin_ctz(a.hot_cw));
            s12 = a.S1[x] - a.S2a[x];
            cur_x = x;
        }
        int64_t slot = 0;
        const bool want = (x != y) && out_slot(a, x, y, slot);
        const int cells = bx * bx * by;
        const bool fits = cells <= a.cap_wave;
        const bool run = want && fits;
        const WaveCopies cp = wave_copies(cells, a.cap_wave, TNX_WAVE_COPY_BUDGET);
        const int gl = lane & (cp.G - 1);
        uint32_t *hc = hist + gl;
        const uint32_t kmul = (uint32_t)by << (2 + cp.gsh), ysh = 2 + cp.gsh;
        // the lane's word after the slice's histogram (and hot cells) takes
        // the padded samples
        const uint32_t trash_off = (uint32_t)((zero_row - hist) + lane - gl) * 4u;
        HotK hk{0u, 0u};
        if constexpr (HS)
            hk = HotK{16u * (uint32_t)a.hot_cw,
                      (uint32_t)(hot_base(a) + (lane & (a.hot_cw - 1)) - gl) * 4u - 0x8000u};
        pipe_count<NI, 0, HS>(b0, b1, regs, row_rsrc(a.bins + (int64_t)y * a.ld, row_bytes),
                              row_rsrc(a.bins + (int64_t)yn * a.ld, row_bytes), more, run, hc,
                              kmul, ysh, hk, trash_off, Lx, lane);
        if (run) {
            wave_lds_fence();
            if constexpr (HS) {  // hot copies into copy 0 of their cells (empty until now)
                uint32_t *hot = hist + hot_base(a);
                const int C = a.hot_cw;
                for (int m = lane; m < 4 * by; m += kWave) {
                    uint4 *q = reinterpret_cast<uint4 *>(hot + m * C);
                    uint32_t v = 0;
                    for (int r = 0; r < C / 4; r++) {
                        const uint4 u = q[r];
                        q[r] = make_uint4(0u, 0u, 0u, 0u);
                        v += (u.x + u.y) + (u.z + u.w);
                    }
                    const int h = m & 3, svh = (h >> 1) * bx + (h & 1);
                    if (v) hist[(svh * by + (m >> 2)) << cp.gsh] += v;
                }
                wave_lds_fence();
            }
            zero_row[lane] = 0u;  // the padded samples' counts
            if (a.counts_map) dump_wave_counts(a, x, y, hist, cells, cp, lane);
            const int64_t d = reduce_wave_fx(hist, bx, by, cp, xf, lane);
            wave_lds_fence();
            if (lane == 0) emit_te(a, slot, (ldexp((double)d, -a.fx_shift) + s12) * invNx);
        } else if (want && lane == 0) {
            emit_te(a, slot, __builtin_nan(""));  // plan mismatch: the pair is not counted
        }
        x = xn; y = yn; bx = bxn; by = byn;
    }
}

// One pair, whole workgroup (hist_pair_block's passes over ranges of
// x_{t+tau}) with a branch-free accumulate and a cell-parallel scan: samples
// of the pair's other passes and samples past L are sent to `trash`, a
// thread-private word, by a select instead of an exec mask per sample.
// With HS the hot states' samples are counted in the first pass, in kHotCB
// copies of their hot cells at `hot`, and each pass adds the copies of its
// hot cells to its histogram before the scan.  hist, c2b and the hot cells
// are zero on entry and exit; trash holds garbage (the caller clears it
// before the region serves wave slices again).
template <int NB, bool HS>
__device__ __forceinline__ int64_t block_pair(uint32_t *hist, uint32_t *c2b, uint32_t *hot,
                                         uint32_t *trash, int cap, const TargetRegs<NB> &regs,
                                         const uint4 *ys, int bx, int by, int L,
                                         const XfLookup &xf, int64_t *red, int32_t *cdbg) {
    const int tid = threadIdx.x, nth = kPersistThreads;
    const int ncol = bx * by;
    const uint32_t by4 = (uint32_t)by << 2;
    int apass = cap / ncol;
    if (apass < 1) apass = 1;
    int64_t a3 = 0;
    for (int a0 = 0; a0 < bx; a0 += apass) {
        const int a1 = (a0 + apass < bx) ? (a0 + apass) : bx;
        const uint32_t lo = (uint32_t)(a0 * bx), span = (uint32_t)((a1 - a0) * bx);
        uint32_t *hc = hist - a0 * ncol;  // cell index of the full pair
        // byte offset from hc of this thread's trash word
        const uint32_t trash_rel = (uint32_t)((trash + (tid & 63)) - hc) * 4u;
        const bool single = a1 - a0 == bx;
        HotK hk{0u, 0u};
        if constexpr (HS)
            hk = HotK{16u * kHotCB,
                      (uint32_t)((hot + (tid & (kHotCB - 1))) - hc) * 4u - 0x8000u};
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int c0 = 16 * nth * i;  // first sample of chunk row i
            if (c0 >= L) break;           // wave-uniform
            uint32_t yw[4] = {ys[i].x, ys[i].y, ys[i].z, ys[i].w};
            uint32_t sr[8];
#pragma unroll
            for (int k = 0; k < 8; k++) sr[k] = regs.s[8 * i + k];
            // one chunk row at a time: keeps the compiler from computing the
            // addresses of several rows ahead in registers
            asm volatile("" : "+v"(yw[0]), "+v"(yw[1]), "+v"(yw[2]), "+v"(yw[3]));
            asm volatile("" : "+v"(sr[0]), "+v"(sr[1]), "+v"(sr[2]), "+v"(sr[3]), "+v"(sr[4]),
                         "+v"(sr[5]), "+v"(sr[6]), "+v"(sr[7]));
            // a wave whose 1024 samples of this row all lie past L skips it
            if (c0 + 16 * (tid & ~(kWave - 1)) >= L) continue;  // wave-uniform
            if (single && c0 + 16 * nth <= L) {
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const uint32_t sv = (sr[q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
                    const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                    lds_inc(reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(hc) +
                                                         key_off<HS>(sv, yv, by4, 2u, hk)));
                }
            } else {
                const int nv = L - (16 * tid + c0);  // valid samples of this thread's 16
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const uint32_t sv = (sr[q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
                    const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                    const uint32_t off = __umul24(sv, by4) + (yv << 2);
                    uint32_t rel = (sv - lo < span && q < nv) ? off : trash_rel;
                    if constexpr (HS) {  // hot states: counted once, in the first pass
                        const uint32_t hof = __umul24(yv, hk.hmul) + sv + hk.hadj;
                        rel = (a0 == 0 && sv >= 0x8000u && q < nv) ? hof : rel;
                    }
                    lds_inc(reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(hc) + rel));
                }
            }
        }
        __syncthreads();
        if constexpr (HS) {  // this pass's hot cells: copies summed into the histogram
            const int m = tid, h = m & 3, ah = h >> 1;
            if (m < 4 * by && ah >= a0 && ah < a1) {
                uint4 *q = reinterpret_cast<uint4 *>(hot + m * kHotCB);
                uint32_t v = 0;
#pragma unroll
                for (int r = 0; r < kHotCB / 4; r++) {
                    const uint4 u = q[r];
                    q[r] = make_uint4(0u, 0u, 0u, 0u);
                    v += (u.x + u.y) + (u.z + u.w);
                }
                if (v) hist[(ah * bx + (h & 1)) * by + (m >> 2) - a0 * ncol] += v;
            }
            __syncthreads();
        }
        if (cdbg) {  // test hook: the pass's counts, read-only, before the scan clears them
            const int pc = (a1 - a0) * ncol;
            for (int k = tid; k < pc; k += nth) cdbg[a0 * ncol + k] = (int32_t)hist[k];
            __syncthreads();
        }
        // cell-parallel scan: thread tid takes cells k = tid + nth j of the
        // pass with four reads (read and clear) in flight; a nonzero cell
        // adds its count to its (x_t, y_t) column, c2b[k mod ncol], the
        // column index kept incrementally (nth mod ncol per step)
        const int pc = (a1 - a0) * ncol;
        const int step = nth % ncol;
        int col = tid % ncol;
        int k = tid;
        for (; k + 3 * nth < pc; k += 4 * nth) {
            uint32_t v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) v[u] = lds_take(hist + k + u * nth);
            a3 += xf_sum<4>(xf, v);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (v[u])
                    __hip_atomic_fetch_add(c2b + col, v[u], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                col += step;
                if (col >= ncol) col -= ncol;
            }
        }
        for (; k < pc; k += nth) {
            const uint32_t v = lds_take(hist + k);
            if (v) {
                __hip_atomic_fetch_add(c2b + col, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                a3 += xf_one(xf, v);
            }
            col += step;
            if (col >= ncol) col -= ncol;
        }
        __syncthreads();
    }
    int64_t a2 = 0;
    for (int c0 = tid; c0 < ncol; c0 += 4 * nth) {
        uint32_t v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = c0 + u * nth < ncol ? lds_take(c2b + c0 + u * nth) : 0u;
        a2 += xf_sum<4>(xf, v);
    }
    // workgroup sum: waves, then thread 0 over the waves' partials (read
    // before the caller's next barrier, so none is needed after)
    const int64_t w = wave_sum_i64(a3 - a2);
    if ((tid & 63) == 0) red[tid >> 6] = w;
    __syncthreads();
    int64_t tot = 0;
    if (tid == 0)
        for (int i = 0; i < kWavesPerWG; i++) tot += red[i];
    return tot;
}

// Block-role source row: thread tid holds samples 16 tid + 4096 i, i < NB.
template <int NB>
__device__ __forceinline__ void load_row256(uint4 *ys, const int8_t *__restrict__ yr, int L) {
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int t = 16 * threadIdx.x + 16 * kPersistThreads * i;
        ys[i] = t < L ? *reinterpret_cast<const uint4 *>(yr + t) : make_uint4(0u, 0u, 0u, 0u);
    }
}

// A block item: its targets one by one against sources [src_beg, src_end),
// the whole workgroup on each pair (hist_pair_block), the next source row
// loaded while the current pair is counted.
template <int NI, bool HS>
__device__ __forceinline__ void persist_block_item(const tnx_te_args_t &a, const tnx_work_t it,
                                                   uint32_t *region, int capB,
                                                   const XfLookup &xf, int64_t *red, int L,
                                                   double invN) {
    constexpr int NB = (NI + 3) / 4;  // 16-sample chunks per thread (4096 samples per chunk row)
    const int tid = threadIdx.x;
    uint32_t *hist = region;
    uint32_t *hot = region + capB;  // hot cells (HS): 4 states x b_y levels x kHotCB copies
    uint32_t *c2b = hot + hot_block_words(a);
    uint32_t *trash = c2b + a.max_block_cols + kWave;
    TargetRegs<NB> regs;
    uint4 ycur[NB], ynext[NB];
    for (int ti = 0; ti < it.tgt_cnt; ti++) {
        const int x = a.tgt_order[it.tgt_beg + ti];
        const int bx = a.n_per_var[x];
        const uint16_t *st = a.states + (int64_t)x * a.lds;
        const int Lx = a.var_L ? min(max(a.var_L[x], 1), L) : L;  // this target's samples
        const double invNx = a.var_L ? 1.0 / (double)Lx : invN;
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int t = 16 * tid + 16 * kPersistThreads * i;
            if (t < Lx) {
                const uint4 u = *reinterpret_cast<const uint4 *>(st + t);
                const uint4 v = *reinterpret_cast<const uint4 *>(st + t + 8);
                regs.s[8 * i + 0] = u.x; regs.s[8 * i + 1] = u.y;
                regs.s[8 * i + 2] = u.z; regs.s[8 * i + 3] = u.w;
                regs.s[8 * i + 4] = v.x; regs.s[8 * i + 5] = v.y;
                regs.s[8 * i + 6] = v.z; regs.s[8 * i + 7] = v.w;
            }
        }
        if constexpr (HS) hot_encode_regs(regs.s, bx, 2u + (uint32_t)__builtin_ctz(kHotCB));
        const double s12 = a.S1[x] - a.S2a[x];
        int y = a.src_order[it.src_beg];
        int by = a.n_per_var[y];
        load_row256<NB>(ycur, a.bins + (int64_t)y * a.ld, Lx);
        for (int sj = it.src_beg; sj < it.src_end; sj++) {
            int yn = 0, byn = 0;
            if (sj + 1 < it.src_end) {
                yn = a.src_order[sj + 1];
                byn = a.n_per_var[yn];
                load_row256<NB>(ynext, a.bins + (int64_t)yn * a.ld, Lx);
            }
            int64_t slot = 0;
            // pairs above cap_sort are left to the occupied-key kernel
            const bool want = (x != y) && !sort_pair(a, bx, by) && out_slot(a, x, y, slot);
            const bool fits = bx * by <= a.max_block_cols;
            if (want && fits) {
                int32_t *cdbg = nullptr;
                if (a.counts_map) {
                    const int r = a.counts_map[(int64_t)x * a.n_vars + y];
                    if (r >= 0 && (int64_t)bx * bx * by <= a.counts_ld)
                        cdbg = a.counts_dbg + (int64_t)r * a.counts_ld;
                }
                const int64_t d = block_pair<NB, HS>(hist, c2b, hot, trash, capB, regs, ycur,
                                                     bx, by, Lx, xf, red, cdbg);
                if (tid == 0) emit_te(a, slot, (ldexp((double)d, -a.fx_shift) + s12) * invNx);
            } else if (want && !fits && tid == 0) {
                emit_te(a, slot, __builtin_nan(""));  // plan mismatch
            }
            y = yn;
            by = byn;
#pragma unroll
            for (int i = 0; i < NB; i++) ycur[i] = ynext[i];
        }
    }
    __syncthreads();  // every thread's trash increments are done
    if (tid < kWave) trash[tid] = 0u;  // the region serves wave slices next
}

// Next entry of the queues [q0, q0 + TNX_QUEUES) (wave queues q0 = 0, block
// queues q0 = TNX_QUEUES), starting at queue q0 + home, or -1 when all are
// drained; `tries` counts the queues found empty (kept by the caller).
__device__ __forceinline__ int take_entry(const tnx_te_args_t &a, int q0, int home, int &tries) {
    while (tries < TNX_QUEUES) {
        const int qq = q0 + (home + tries) % TNX_QUEUES;
        const int beg = a.queue_off[qq], n_q = a.queue_off[qq + 1] - beg;
        if (n_q > 0) {
            const int pos = atomicAdd(a.queue_ctr + qq * kQueueStride, 1);
            if (pos < n_q) return a.queue_items[beg + pos];
        }
        ++tries;
    }
    return -1;
}

template <int NI, bool HS>
__global__ __launch_bounds__(kPersistThreads) __attribute__((amdgpu_waves_per_eu(3, 8))) void
te_persist_kernel(tnx_te_args_t a) {
    // 16-byte aligned: the copy fold reads four copies per ds_read_b128
    extern __shared__ __attribute__((aligned(16))) uint32_t lds_u32[];
    __shared__ int64_t red[kWavesPerWG];
    __shared__ int s_entry;
    __shared__ int64_t lfx[kSmallTbl];  // fixed-point c log2 c of counts < kSmallTbl
    uint32_t *region = lds_u32;         // four wave slices
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int L = (int)(a.T - a.tau);
    const double invN = 1.0 / (double)L;
    for (int i = tid; i < kSmallTbl; i += kPersistThreads) lfx[i] = i <= L ? a.xlogx_fx[i] : 0;
    const XfLookup xf{lfx, row_rsrc(a.xlogx_fx, (uint32_t)(8 * (L + 1)))};
    const int slice = wave_slice(wave_area(a));
    for (int i = tid; i < kWavesPerWG * slice; i += kPersistThreads) region[i] = 0;
    __syncthreads();
    const int capB =
        (kWavesPerWG * slice - a.max_block_cols - 2 * kWave - hot_block_words(a)) & ~3;
    const int home = (int)(blockIdx.x % TNX_QUEUES);  // this workgroup's XCD
    // Block entries: the whole workgroup takes one at a time (thread 0 takes,
    // LDS broadcast), its own XCD's block queue first.
    auto block_loop = [&]() {
        int tries = 0;  // thread 0: block queues found empty so far
        while (true) {
            if (tid == 0) s_entry = take_entry(a, TNX_QUEUES, home, tries);
            __syncthreads();
            const int e = __builtin_amdgcn_readfirstlane(s_entry);  // uniform: scalar loads
            __syncthreads();  // s_entry is rewritten by the next take
            if (e < 0) break;
            persist_block_item<NI, HS>(a, a.work[a.n_wave_items + e], region, capB, xf, red, L,
                                       invN);
        }
    };
    // One workgroup in three per XCD starts on block entries, so the
    // latency-bound block pairs run beside LDS-bound wave pairs on most CUs.
    if ((blockIdx.x / TNX_QUEUES) % 3 == 0) block_loop();
    // Wave entries: each wave takes its own, no workgroup barrier (the waves
    // of a workgroup drift apart instead of waiting for the slowest).  An
    // entry is item e / 4, pairs q = e % 4, e % 4 + 4, ... (or, with
    // item_per_wave, item e whole).
    {
        uint32_t *hist = region + wave * slice;
        uint32_t *zero_row = hist + wave_area(a);  // the lanes' trash words
        const bool per_wave = a.item_per_wave != 0;
        int tries = 0;
        while (true) {
            int e = -1;
            if (lane == 0) e = take_entry(a, 0, home, tries);
            e = __builtin_amdgcn_readfirstlane(e);
            tries = __builtin_amdgcn_readfirstlane(tries);
            if (e < 0) break;
            const int64_t wi = per_wave ? e : e / kWavesPerWG;
            persist_wave_item<NI, HS>(a, a.work[wi], per_wave ? 0 : e % kWavesPerWG,
                                      per_wave ? 1 : kWavesPerWG, hist, zero_row, xf, L, invN,
                                      lane);
        }
    }
    __syncthreads();  // every wave slice is clear: the region may serve block pairs
    block_loop();     // block entries still queued
}

// Resident workgroups per CU of a persistent kernel times the CUs of the
// current device (cached per device, kernel and LDS size).
static int persist_grid(const void *kfn, size_t shm) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    static std::mutex mu;
    static std::map<std::tuple<int, const void *, size_t>, int> cache;
    const auto key = std::make_tuple(dev, kfn, shm);
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -1;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kPersistThreads, shm) !=
        hipSuccess)
        return -1;
    const int grid = std::max(1, per_cu) * n_cu;
    std::lock_guard<std::mutex> g(mu);
    cache[key] = grid;
    return grid;
}

template <bool HS>
static const void *persist_fn(int ni) {
    switch (ni) {
#define TNX_PERSIST_FN(NI) \
    case NI: return (const void *)te_persist_kernel<NI, HS>;
        TNX_PERSIST_FN(1) TNX_PERSIST_FN(2) TNX_PERSIST_FN(3) TNX_PERSIST_FN(4)
        TNX_PERSIST_FN(6) TNX_PERSIST_FN(8) TNX_PERSIST_FN(10) TNX_PERSIST_FN(12)
#undef TNX_PERSIST_FN
        default: return nullptr;
    }
}

int launch_persist(const tnx_te_args_t &a, int ni, hipStream_t s) {
    TNX_CHECK_ARG(a.queue_off && a.queue_ctr && a.work && a.xlogx_fx,
                  "tnx_te_run: persistent launch needs queue_off, queue_ctr, work and xlogx_fx");
    TNX_CHECK_ARG(a.fx_shift >= 0 && a.fx_shift <= 62, "tnx_te_run: bad fx_shift %d", a.fx_shift);
    TNX_CHECK_ARG(ni >= 1 && ni <= 12 && a.T - a.tau <= kPersistMaxL,
                  "tnx_te_run: the persistent launch takes T - tau <= %lld",
                  (long long)kPersistMaxL);
    TNX_CHECK_ARG(a.hot_cw == 0 || ((a.hot_cw == 4 || a.hot_cw == 8 || a.hot_cw == 16) &&
                                    a.hot_maxb >= 1 && a.hot_maxb <= 32),
                  "tnx_te_run: hot_cw must be 0, 4, 8 or 16 and 1 <= hot_maxb <= 32 (got %d, %d)",
                  a.hot_cw, a.hot_maxb);
    const size_t slice = (size_t)wave_slice(wave_area(a));
    const size_t shm = (size_t)kWavesPerWG * slice * 4;  // + the static 2 KiB table
    TNX_CHECK_ARG(shm <= 156 * 1024,
                  "tnx_te_run: cap_wave %d too large for the persistent launch", a.cap_wave);
    if (a.n_block_items > 0)
        TNX_CHECK_ARG(a.max_block_cols >= 1 &&
                          kWavesPerWG * (int64_t)slice - a.max_block_cols - 2 * kWave -
                                  hot_block_words(a) >=
                              std::max<int64_t>(a.max_block_cols, 4),
                      "tnx_te_run: max_block_cols %d too large for the persistent launch",
                      a.max_block_cols);
    TNX_CHECK_ARG(!a.counts_map || (a.counts_dbg && a.counts_ld >= 1),
                  "tnx_te_run: counts_map needs counts_dbg and counts_ld");
    const void *kfn = a.hot_cw > 0 ? persist_fn<true>(ni) : persist_fn<false>(ni);
    TNX_CHECK_ARG(kfn != nullptr, "tnx_te_run: no persistent variant for NI = %d", ni);
    TNX_CHECK_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    const int grid = persist_grid(kfn, shm);
    TNX_CHECK_ARG(grid > 0, "tnx_te_run: occupancy query failed");
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(TNX_QUEUE_CTR_INTS), 0, s, a.queue_ctr);
    TNX_CHECK_LAUNCH("queue_reset_kernel");
    tnx_te_args_t args = a;
    void *kargs[] = {&args};
    TNX_CHECK_HIP(hipLaunchKernel(kfn, dim3((unsigned)grid), dim3(kPersistThreads), kargs, shm, s));
    TNX_CHECK_LAUNCH("te_persist_kernel");
    return TNX_OK;
}

// ── tnx_te_run_noalloc: the queue take counters in a device-side array ────
// One copy per device (each GPU the library runs on holds it until the
// process ends).  Host code never passes the name `scratch` to a kernel or a
// copy: the name is a host-side stand-in; its device address comes from
// hipGetSymbolAddress, looked up once per device with that device current.
static __device__ __align__(256) int32_t scratch[TNX_NOALLOC_SCRATCH_INTS];
static_assert(2 * TNX_QUEUES * kQueueStride <= TNX_NOALLOC_SCRATCH_INTS,
              "the no-alloc scratch holds every queue's take counter");

constexpr int kMaxDevices = 64;

static int32_t *scratch_address() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("tnx_te_run_noalloc: no current device (or device index >= %d)", kMaxDevices);
        return nullptr;
    }
    static std::mutex mu;
    static int32_t *addr[kMaxDevices] = {};
    std::lock_guard<std::mutex> g(mu);
    if (addr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || p == nullptr) {
            set_error("tnx_te_run_noalloc: hipGetSymbolAddress failed on device %d: %s", dev,
                      hipGetErrorString(rc));
            return nullptr;
        }
        addr[dev] = static_cast<int32_t *>(p);
    }
    return addr[dev];
}

}  // namespace tnx

using namespace tnx;

extern "C" int tnx_te_run_noalloc(const tnx_te_args_t *args, void *stream) {
    clear_error();
    TNX_CHECK_ARG(args != nullptr, "tnx_te_run_noalloc: args is NULL");
    TNX_CHECK_ARG(args->queue_items != nullptr,
                  "tnx_te_run_noalloc: needs a persistent plan (queue_items from tnx_te_queue)");
    TNX_CHECK_ARG(args->queue_ctr == nullptr,
                  "tnx_te_run_noalloc: queue_ctr must be NULL (the counters live in the "
                  "library's device array)");
    const int64_t need = 2 * TNX_QUEUES * kQueueStride;
    if (need > TNX_NOALLOC_SCRATCH_INTS) {
        set_error("tnx_te_run_noalloc: %lld queue counters exceed the scratch maximum of %d",
                  (long long)need, TNX_NOALLOC_SCRATCH_INTS);
        return TNX_E_CAP;
    }
    int32_t *ctr = scratch_address();
    if (ctr == nullptr) return TNX_E_HIP;
    tnx_te_args_t a = *args;
    a.queue_ctr = ctr;  // cleared by queue_reset_kernel on the stream, as in tnx_te_run
    return tnx_te_run(&a, stream);
}

namespace {

// Relative cost of a pair in sample-update units: the wave path's atomics
// slowed by the bank conflicts of G copies plus a per-pair reduction; the
// block path's atomics, its passes' sample scans and its cell scan.
double wave_pair_cost(int64_t bx, int64_t by, int64_t L, int32_t cap) {
    const int64_t cells = bx * bx * by;
    int G = 1;
    while (G < TNX_MAX_COPIES && 2 * G * cells <= TNX_WAVE_COPY_BUDGET &&
           (2 * G + 1) * cells <= cap)
        G *= 2;
    static const double conflict[6] = {2.4, 2.0, 1.7, 1.4, 1.15, 1.0};  // G = 1 .. 32
    int lg = 0;
    while ((1 << lg) < G) lg++;
    return (double)L * conflict[std::min(lg, 5)] + 4.0 * (double)(G * cells) + 1500.0;
}

double block_pair_cost(int64_t bx, int64_t by, int64_t L, int64_t capB) {
    const int64_t ncol = bx * by, cells = bx * ncol;
    const int64_t per = std::max<int64_t>(1, capB / ncol);
    const int64_t passes = (bx + per - 1) / per;
    return (double)L * (1.8 + 0.35 * (double)passes) + 2.0 * (double)cells + 6000.0;
}

}  // namespace

extern "C" int tnx_te_queue(const tnx_work_t *work, int64_t n_wave_items, int64_t n_block_items,
                            const int32_t *npv, int64_t n_vars, const int32_t *tgt_order,
                            const int32_t *src_order, int64_t L, int32_t cap_wave,
                            int32_t item_per_wave, int32_t *qitems, int64_t qcap, int32_t *qoff,
                            int64_t *n_wave_entries) {
    clear_error();
    TNX_CHECK_ARG(npv && tgt_order && src_order && n_wave_entries &&
                      (work || n_wave_items + n_block_items == 0),
                  "tnx_te_queue: null pointer");
    TNX_CHECK_ARG(n_wave_items >= 0 && n_block_items >= 0 && L >= 1,
                  "tnx_te_queue: bad sizes");
    if (cap_wave <= 0) cap_wave = kDefaultCapWave;
    // wave entries: item * 4 + slot (a wave takes the item's pairs q = slot,
    // slot + 4, ...), or with item_per_wave the item itself
    std::vector<int32_t> wave_e;
    wave_e.reserve(item_per_wave ? n_wave_items : 4 * n_wave_items);
    for (int64_t i = 0; i < n_wave_items; i++) {
        if (item_per_wave) {
            wave_e.push_back((int32_t)i);
            continue;
        }
        const int64_t np = (int64_t)work[i].tgt_cnt * (work[i].src_end - work[i].src_beg);
        for (int slot = 0; slot < kWavesPerWG && slot < np; slot++)
            wave_e.push_back((int32_t)(i * kWavesPerWG + slot));
    }
    const int64_t nwe = (int64_t)wave_e.size();
    const int64_t n_entries = nwe + n_block_items;
    TNX_CHECK_ARG(n_entries < (1LL << 31) && n_wave_items * kWavesPerWG < (1LL << 31),
                  "tnx_te_queue: too many entries");
    *n_wave_entries = nwe;
    if (!qitems) return TNX_OK;  // size query
    TNX_CHECK_ARG(qoff != nullptr, "tnx_te_queue: null qoff");
    if (n_entries > qcap) {
        set_error("tnx_te_queue: qcap %lld < %lld entries", (long long)qcap, (long long)n_entries);
        return TNX_E_CAP;
    }
    // the block role's histogram: four default slices minus the widest marginal
    int32_t maxcols = 1;
    for (int64_t b = 0; b < n_block_items; b++) {
        const tnx_work_t &w = work[n_wave_items + b];
        for (int32_t t = 0; t < w.tgt_cnt; t++)
            for (int32_t j = w.src_beg; j < w.src_end; j++)
                maxcols = std::max(maxcols, npv[tgt_order[w.tgt_beg + t]] * npv[src_order[j]]);
    }
    const int64_t capB =
        std::max<int64_t>(4, kWavesPerWG * wave_slice(cap_wave) - maxcols - 2 * kWave);
    auto pair_cost = [&](const tnx_work_t &w, int32_t t, int32_t j, bool block) {
        const int64_t bx = npv[tgt_order[w.tgt_beg + t]], by = npv[src_order[j]];
        return block ? block_pair_cost(bx, by, L, capB) : wave_pair_cost(bx, by, L, cap_wave);
    };
    struct Entry {
        int32_t id, chunk;
        int64_t item;
        double cost, item_cost;
    };
    // cost of a wave entry: its pairs (every 4th of the item, or all)
    std::vector<Entry> went, bent;
    went.reserve(nwe);
    for (int32_t e : wave_e) {
        const int64_t i = item_per_wave ? e : e / kWavesPerWG;
        const tnx_work_t &w = work[i];
        const int32_t first = item_per_wave ? 0 : e % kWavesPerWG;
        const int32_t step = item_per_wave ? 1 : kWavesPerWG;
        const int64_t np = (int64_t)w.tgt_cnt * (w.src_end - w.src_beg);
        double c = 0;
        for (int64_t q = first; q < np; q += step)
            c += pair_cost(w, (int32_t)(q % w.tgt_cnt), w.src_beg + (int32_t)(q / w.tgt_cnt), false);
        went.push_back({e, w.src_beg / kChunkSrc, i, c, 0.0});
    }
    // an item's slots stay adjacent: order by the item's total cost
    {
        std::vector<double> icost(n_wave_items, 0.0);
        for (const Entry &x : went) icost[x.item] += x.cost;
        for (Entry &x : went) x.item_cost = icost[x.item];
    }
    bent.reserve(n_block_items);
    for (int64_t b = 0; b < n_block_items; b++) {
        const tnx_work_t &w = work[n_wave_items + b];
        double c = 0;
        for (int32_t t = 0; t < w.tgt_cnt; t++)
            for (int32_t j = w.src_beg; j < w.src_end; j++) c += pair_cost(w, t, j, true);
        bent.push_back({(int32_t)b, w.src_beg / kChunkSrc, b, c, c});
    }
    // source chunks to XCD queues: heaviest chunk first onto the least loaded
    int32_t n_chunks = 1;
    for (const Entry &x : went) n_chunks = std::max(n_chunks, x.chunk + 1);
    for (const Entry &x : bent) n_chunks = std::max(n_chunks, x.chunk + 1);
    std::vector<double> chunk_cost(n_chunks, 0.0);
    for (const Entry &x : went) chunk_cost[x.chunk] += x.cost;
    for (const Entry &x : bent) chunk_cost[x.chunk] += x.cost;
    std::vector<int32_t> order(n_chunks);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t p, int32_t q) { return chunk_cost[p] > chunk_cost[q]; });
    std::vector<int32_t> chunk_q(n_chunks, 0);
    double load[TNX_QUEUES] = {0};
    for (int32_t c : order) {
        int best = 0;
        for (int qq = 1; qq < TNX_QUEUES; qq++)
            if (load[qq] < load[best]) best = qq;
        chunk_q[c] = best;
        load[best] += chunk_cost[c];
    }
    // inside a queue: chunk by chunk (ascending), heaviest items first
    auto by_queue = [&](const Entry &u, const Entry &v) {
        const int32_t qu = chunk_q[u.chunk], qv = chunk_q[v.chunk];
        if (qu != qv) return qu < qv;
        if (u.chunk != v.chunk) return u.chunk < v.chunk;
        if (u.item_cost != v.item_cost) return u.item_cost > v.item_cost;
        return u.id < v.id;
    };
    std::stable_sort(went.begin(), went.end(), by_queue);
    std::stable_sort(bent.begin(), bent.end(), by_queue);
    int64_t pos = 0;
    for (int part = 0; part < 2; part++) {
        const std::vector<Entry> &ent = part ? bent : went;
        size_t k = 0;
        for (int qq = 0; qq < TNX_QUEUES; qq++) {
            qoff[part * TNX_QUEUES + qq] = (int32_t)pos;
            while (k < ent.size() && chunk_q[ent[k].chunk] == qq) qitems[pos++] = ent[k++].id;
        }
    }
    qoff[2 * TNX_QUEUES] = (int32_t)pos;
    return TNX_OK;
}
