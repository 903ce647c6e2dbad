// te_device.hpp -- device building blocks shared by the TE and surrogate kernels.
#pragma once
#include "common.hpp"

// Cells of privatised histogram copies a pair may use (see wave_copies).
#ifndef TNX_WAVE_COPY_BUDGET
#define TNX_WAVE_COPY_BUDGET 2048
#endif
// Most copies of a wave histogram (a power of two <= 2^TNX_LDS_BANK_SHIFT).
#ifndef TNX_MAX_COPIES
#define TNX_MAX_COPIES 32
#endif
// log2 of the LDS banks the fold's rotation spreads a half-wave over.
#ifndef TNX_LDS_BANK_SHIFT
#define TNX_LDS_BANK_SHIFT 5
#endif

namespace tnx {

// Raw buffer access to one row: the base address and size live in four
// scalar registers and each load adds a per-lane 32-bit offset, so a wave
// streaming a row keeps no 64-bit address per chunk; loads past `bytes`
// return zero.  (Buffer resource word 3 of the gfx9 family, raw addressing.)
typedef int32_t i32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                             0x00020000);
}
__device__ __forceinline__ uint4 row_load16(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                            uint32_t soff) {
    const i32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
    return make_uint4((uint32_t)v.x, (uint32_t)v.y, (uint32_t)v.z, (uint32_t)v.w);
}

__device__ __forceinline__ void lds_inc(uint32_t *p) {
    __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Read an LDS word and leave zero behind (one ds_wrxchg_rtn_b32).
__device__ __forceinline__ uint32_t lds_take_w(uint32_t *p) {
    return __hip_atomic_exchange(p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Byte offset of key state * by + y in a histogram whose cells are `G`
// words apart (cell-major privatised copies, below): kmul = 4 * G * by and
// ysh = 2 + log2 G.  States < 2^16 and by <= 127, so the product fits the
// full-rate 24-bit multiplier (v_mul_u32_u24, which also takes the 16-bit
// half of a packed state register directly).  Three VALU operations per
// sample: multiply, shift of the source byte, three-way add with the copy
// base.
__device__ __forceinline__ void lds_inc_key(uint32_t *hist, uint32_t sv, uint32_t kmul,
                                            uint32_t yv, uint32_t ysh) {
    const uint32_t off = __umul24(sv, kmul) + (yv << ysh);
    lds_inc(reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(hist) + off));
}

// With HZ (hot zero key), as lds_inc_key but key 0 (state 0 and source
// level 0) goes to `zero`, a word private to the lane.  In zero-inflated
// count data level 0 is each variable's most frequent level and key 0 can
// hold a quarter of all samples; counted in one cell it would serialise the
// wave's atomics on one address.  fold_zero_key adds the lanes' words back
// into cell 0.  The diversion costs two VALU operations per sample (compare,
// select), so data without a hot key 0 run the plain form (HZ false).
template <bool HZ>
__device__ __forceinline__ void lds_inc_key0(uint32_t *hist, uint32_t sv, uint32_t kmul,
                                             uint32_t yv, uint32_t ysh, uint32_t *zero) {
    const uint32_t off = __umul24(sv, kmul) + (yv << ysh);
    uint32_t *cell = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(hist) + off);
    if constexpr (HZ) {
        lds_inc(off ? cell : zero);
    } else {
        lds_inc(cell);
    }
}

// Cell 0 (copy 0, the first word of the slice) += the 64 lanes' key-0 words
// zero_row[0..63], which are cleared.  One wavefront; nothing without HZ.
template <bool HZ>
__device__ __forceinline__ void fold_zero_key(uint32_t *hist, uint32_t *zero_row, int lane) {
    if constexpr (!HZ) return;
    const uint32_t c0 = wave_sum_u32(lds_take_w(zero_row + lane));
    if (lane == 0 && c0) hist[0] += c0;
    wave_lds_fence();
}

// 16 samples of one lane: 8 packed 16-bit states, 4 source dwords.
template <bool HZ>
__device__ __forceinline__ void acc16(uint32_t *hist, const uint32_t *sreg, const uint4 y4,
                                      uint32_t kmul, uint32_t ysh, uint32_t *zero) {
    const uint32_t yw[4] = {y4.x, y4.y, y4.z, y4.w};
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const uint32_t sv = (sreg[q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
        const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
        lds_inc_key0<HZ>(hist, sv, kmul, yv, ysh, zero);
    }
}

template <bool HZ>
__device__ __forceinline__ void acc16_tail(uint32_t *hist, const uint32_t *sreg, const uint4 y4,
                                           uint32_t kmul, uint32_t ysh, uint32_t *zero,
                                           int n_valid) {
    const uint32_t yw[4] = {y4.x, y4.y, y4.z, y4.w};
#pragma unroll
    for (int q = 0; q < 16; q++) {
        if (q < n_valid) {
            const uint32_t sv = (sreg[q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
            const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
            lds_inc_key0<HZ>(hist, sv, kmul, yv, ysh, zero);
        }
    }
}

// Phase 1 (one wavefront), target states read from memory: key =
// state[t] * by + y[t] for t < L.  Lane `lane` takes 16 consecutive samples
// per step (one dwordx4 of source bytes, two of states), two steps in
// flight.  Rows must be padded to a multiple of 64 elements.
template <bool HZ>
__device__ __forceinline__ void hist_accumulate_mem(uint32_t *hist, const uint16_t *__restrict__ st,
                                                    const int8_t *__restrict__ yr, uint32_t kmul,
                                                    uint32_t ysh, uint32_t *zero, int L,
                                                    int lane) {
    // two whole 1024-sample steps per iteration while every lane's samples
    // are inside L (a wave-uniform loop test); the tail is predicated
    int t = 16 * lane;
    for (int tb = 0; tb + 2048 <= L; tb += 2048, t += 2048) {
        const uint4 ya = *reinterpret_cast<const uint4 *>(yr + t);
        const uint4 sa0 = *reinterpret_cast<const uint4 *>(st + t);
        const uint4 sa1 = *reinterpret_cast<const uint4 *>(st + t + 8);
        const uint4 yb = *reinterpret_cast<const uint4 *>(yr + t + 1024);
        const uint4 sb0 = *reinterpret_cast<const uint4 *>(st + t + 1024);
        const uint4 sb1 = *reinterpret_cast<const uint4 *>(st + t + 1032);
        const uint32_t sa[8] = {sa0.x, sa0.y, sa0.z, sa0.w, sa1.x, sa1.y, sa1.z, sa1.w};
        const uint32_t sb[8] = {sb0.x, sb0.y, sb0.z, sb0.w, sb1.x, sb1.y, sb1.z, sb1.w};
        acc16<HZ>(hist, sa, ya, kmul, ysh, zero);
        acc16<HZ>(hist, sb, yb, kmul, ysh, zero);
    }
    for (; t < L; t += 1024) {
        const uint4 ya = *reinterpret_cast<const uint4 *>(yr + t);
        const uint4 sa0 = *reinterpret_cast<const uint4 *>(st + t);
        const uint4 sa1 = *reinterpret_cast<const uint4 *>(st + t + 8);
        const uint32_t sa[8] = {sa0.x, sa0.y, sa0.z, sa0.w, sa1.x, sa1.y, sa1.z, sa1.w};
        acc16_tail<HZ>(hist, sa, ya, kmul, ysh, zero, L - t);
    }
}

// Register-resident target: lane `lane` owns samples
// t = 16 * lane + 1024 * i + q (q < 16) for chunks i < NI, holding their
// 16-bit states as 8 packed dwords per chunk.  Rows must be padded so that
// a 16-sample chunk starting below L stays inside the row (true for rows
// padded to a multiple of 64 elements).
template <int NI>
struct TargetRegs {
    uint32_t s[NI * 8];
};

template <int NI>
__device__ __forceinline__ void load_target_regs(TargetRegs<NI> &r,
                                                 const uint16_t *__restrict__ st, int L,
                                                 int lane) {
#pragma unroll
    for (int i = 0; i < NI; i++) {
        const int t = 16 * lane + 1024 * i;
        uint4 a = make_uint4(0u, 0u, 0u, 0u), b = a;  // state 0 past L
        if (t < L) {
            a = *reinterpret_cast<const uint4 *>(st + t);
            b = *reinterpret_cast<const uint4 *>(st + t + 8);
        }
        r.s[8 * i + 0] = a.x; r.s[8 * i + 1] = a.y; r.s[8 * i + 2] = a.z; r.s[8 * i + 3] = a.w;
        r.s[8 * i + 4] = b.x; r.s[8 * i + 5] = b.y; r.s[8 * i + 6] = b.z; r.s[8 * i + 7] = b.w;
    }
}

// Phase 1 with the target in registers: 16 source bytes per lane per chunk
// (one dwordx4 load), 16 LDS increments.
template <int NI, bool HZ>
__device__ __forceinline__ void hist_accumulate_regs(uint32_t *hist, TargetRegs<NI> &r,
                                                     const int8_t *__restrict__ yr, uint32_t kmul,
                                                     uint32_t ysh, uint32_t *zero, int L,
                                                     int lane) {
    // Keep the packed states opaque per call: otherwise the compiler hoists
    // the per-sample 16-bit extractions out of the caller's source loop and
    // spends a register on every sample.
#pragma unroll
    for (int k = 0; k < NI * 8; k++) asm volatile("" : "+v"(r.s[k]));
    const uint32_t *sreg = r.s;
    // the source bytes of PF chunks are requested before the first of them is
    // counted: one memory latency per PF chunks instead of one per chunk
    // (PF = NI up to 10 chunks, half of NI above, to stay within 3 waves'
    // worth of registers)
    constexpr int PF = NI <= 10 ? NI : (NI + 1) / 2;
#pragma unroll
    for (int g0 = 0; g0 < NI; g0 += PF) {
        uint4 ys[PF];
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int t = 16 * lane + 1024 * (g0 + u);
            ys[u] = (g0 + u < NI && t < L) ? *reinterpret_cast<const uint4 *>(yr + t)
                                           : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int i = g0 + u;
            const int t = 16 * lane + 1024 * i;
            if (i >= NI || t >= L) break;
            const uint32_t yw[4] = {ys[u].x, ys[u].y, ys[u].z, ys[u].w};
            // wave-uniform test (every lane's 16 samples inside L): a per-lane
            // test here lets the compiler fold this path into the tail's
            // per-sample predication
            if (1024 * i + 1024 <= L) {
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const uint32_t sv = (sreg[8 * i + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
                    const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                    lds_inc_key0<HZ>(hist, sv, kmul, yv, ysh, zero);
                }
            } else {
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    if (t + q < L) {
                        const uint32_t sv = (sreg[8 * i + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
                        const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                        lds_inc_key0<HZ>(hist, sv, kmul, yv, ysh, zero);
                    }
                }
            }
        }
    }
}

// The source chunks of one pair for the register path: chunk i of lane l is
// bytes [16 l + 1024 i, +16) of the source row (zero past L).
template <int NI>
struct SourceChunks {
    uint4 c[NI];
};

template <int NI>
__device__ __forceinline__ void load_source_chunks(SourceChunks<NI> &ys,
                                                   const int8_t *__restrict__ yr, int L,
                                                   int lane) {
#pragma unroll
    for (int i = 0; i < NI; i++) {
        const int t = 16 * lane + 1024 * i;
        ys.c[i] = t < L ? *reinterpret_cast<const uint4 *>(yr + t) : make_uint4(0u, 0u, 0u, 0u);
    }
}

// hist_accumulate_regs on source chunks already in registers (loaded by
// load_source_chunks, possibly while the previous pair was reduced).
template <int NI, bool HZ>
__device__ __forceinline__ void hist_count_chunks(uint32_t *hist, TargetRegs<NI> &r,
                                                  const SourceChunks<NI> &ys, uint32_t kmul,
                                                  uint32_t ysh, uint32_t *zero, int L, int lane) {
#pragma unroll
    for (int k = 0; k < NI * 8; k++) asm volatile("" : "+v"(r.s[k]));
    const uint32_t *sreg = r.s;
#pragma unroll
    for (int i = 0; i < NI; i++) {
        const int t = 16 * lane + 1024 * i;
        if (t >= L) break;
        const uint32_t yw[4] = {ys.c[i].x, ys.c[i].y, ys.c[i].z, ys.c[i].w};
        if (1024 * i + 1024 <= L) {   // wave-uniform: every lane's 16 samples inside L
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const uint32_t sv = (sreg[8 * i + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
                const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                lds_inc_key0<HZ>(hist, sv, kmul, yv, ysh, zero);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 16; q++) {
                if (t + q < L) {
                    const uint32_t sv = (sreg[8 * i + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
                    const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                    lds_inc_key0<HZ>(hist, sv, kmul, yv, ysh, zero);
                }
            }
        }
    }
}

// Privatised histograms, cell-major: a pair with `cells` cells gets G copies
// (G a power of two <= 32); cell k of copy r is word k * G + r and lane l
// counts into copy l % G.  LDS has 32 banks of one word, so with G = 32 a
// lane always hits bank l % 32 whatever its key (no bank conflicts, and
// lanes l and l + 32, the only ones sharing a copy, are served in different
// halves of the wave).  Smaller G leaves 32 / G banks per copy and a 1 / cells
// chance of a same-address pair.  The copies take G * cells words and the
// folded totals `cells` more, within the wave's cap-word slice; budget
// bounds G * cells because the reduction reads every copy.
struct WaveCopies {
    int G;    // copies
    int gsh;  // log2 G
};

__device__ __forceinline__ WaveCopies wave_copies(int cells, int cap, int budget) {
    WaveCopies c{1, 0};
    while (c.G < TNX_MAX_COPIES && 2 * c.G * cells <= budget && (2 * c.G + 1) * cells <= cap) {
        c.G *= 2;
        c.gsh += 1;
    }
    return c;
}

// Read an LDS cell and leave zero behind (one ds_wrxchg_rtn_b32).
__device__ __forceinline__ uint32_t lds_take(uint32_t *p) {
    return __hip_atomic_exchange(p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Phases 2-4 for one wavefront: returns S3 - S2b (valid in every lane).
// With copies, lane l first folds cells k = l, l + 64, ...: the G copies of
// cell k are read from rotation (k / (32 / G)) on, so the 32 lanes of a
// half-wave hit 32 distinct banks; the totals go to the words after the
// copies and give S3.  Then per (x_t, y_t) column the b_x totals of
// x_{t+tau} give S2b (without copies, this column pass gives S3 too).
// Leaves the slice zero.
template <class XL>
__device__ __forceinline__ DF hist_reduce_wave_copies(uint32_t *hist, int bx, int by,
                                                      const WaveCopies &cp, const XL &xl,
                                                      int lane) {
    const int ncol = bx * by;
    const int cells = bx * ncol;
    DF a3 = {0.f, 0.f}, a2 = {0.f, 0.f};
    if (cp.G >= 4) {
        // four copies per 16-byte read: the 8 lanes of a read group (32
        // banks) take chunks (j + f) of consecutive cells, f = k / (8 / (G/4))
        // mod G/4, which lands them on 8 distinct 4-bank groups
        const int G4 = cp.G >> 2;
        uint32_t *tot = hist + cp.G * cells;
        const int per_sh = TNX_LDS_BANK_SHIFT - cp.gsh;  // log2(8 / G4)
        for (int k = lane; k < cells; k += kWave) {
            const int f = (k >> per_sh) & (G4 - 1);
            uint4 *row = reinterpret_cast<uint4 *>(hist + (k << cp.gsh));
            uint32_t t = 0;
            for (int j = 0; j < G4; j++) {
                uint4 *q = row + ((j + f) & (G4 - 1));
                const uint4 v = *q;
                *q = make_uint4(0u, 0u, 0u, 0u);
                t += (v.x + v.y) + (v.z + v.w);
            }
            tot[k] = t;
            xl(a3, t);
        }
        wave_lds_fence();
        for (int col = lane; col < ncol; col += kWave) {
            uint32_t c2 = 0;
            for (int aa = 0; aa < bx; aa++) c2 += lds_take(tot + aa * ncol + col);
            xl(a2, c2);
        }
    } else if (cp.G > 1) {
        const int G = cp.G;
        uint32_t *tot = hist + G * cells;
        const int per_sh = TNX_LDS_BANK_SHIFT - cp.gsh;  // log2(banks / G)
        for (int k = lane; k < cells; k += kWave) {
            const int f = k >> per_sh;
            uint32_t *row = hist + (k << cp.gsh);
            uint32_t t = 0;
            for (int j = 0; j < G; j++) t += lds_take(row + ((j + f) & (G - 1)));
            tot[k] = t;
            xl(a3, t);
        }
        wave_lds_fence();
        for (int col = lane; col < ncol; col += kWave) {
            uint32_t c2 = 0;
            for (int aa = 0; aa < bx; aa++) c2 += lds_take(tot + aa * ncol + col);
            xl(a2, c2);
        }
    } else {
        for (int col = lane; col < ncol; col += kWave) {
            uint32_t c2 = 0;
            for (int aa = 0; aa < bx; aa++) {
                const uint32_t v = lds_take(hist + aa * ncol + col);
                c2 += v;
                xl(a3, v);
            }
            xl(a2, c2);
        }
    }
    df_add(a3, -a2.hi, -a2.lo);
    return df_wave_sum(a3);
}

// Workgroup sum of a DF; result valid in thread 0.  red: shared
// float[2 * n_waves].  The partials in red are read by thread 0 only, before
// it reaches the caller's next barrier, so no trailing barrier is needed.
__device__ __forceinline__ DF df_block_sum(DF u, float *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    u = df_wave_sum(u);
    if (lane == 0) {
        red[2 * w + 0] = u.hi;
        red[2 * w + 1] = u.lo;
    }
    __syncthreads();
    DF su = {0.f, 0.f};
    if (threadIdx.x == 0)
        for (int i = 0; i < nw; i++) df_add(su, red[2 * i + 0], red[2 * i + 1]);
    return su;
}

// Whole workgroup on one pair with a histogram of cap cells: passes over
// ranges [a0, a1) of x_{t+tau} so any b_x^2 b_y fits; c2b (b_x b_y cells)
// accumulates the (x_t, y_t) marginal across passes.  The target's states
// are in registers (thread tid owns samples 16 tid + 16 nth i + q, NB chunks,
// L <= 16 nth NB) and so is the source row (ys, same samples).  Returns
// S3 - S2b, valid in thread 0.  hist and c2b must be zero on entry and are
// zero on exit; the caller's next pair may start right away.  cdbg (test
// hook, may be NULL) receives the pair's joint counts, pass by pass.
template <int NB, bool HZ, class XL>
__device__ __forceinline__ DF hist_pair_block(uint32_t *hist, uint32_t *c2b, uint32_t *zero_row,
                                              int cap, const TargetRegs<NB> &regs,
                                              const uint4 *ys, int bx, int by, int L,
                                              const XL &xl, float *red,
                                              int32_t *cdbg = nullptr) {
    const int tid = threadIdx.x, nth = blockDim.x;
    uint32_t *zero = zero_row + (tid & 63);
    const int ncol = bx * by;
    const uint32_t by4 = (uint32_t)by << 2;
    int apass = cap / ncol;
    if (apass < 1) apass = 1;
    DF a3 = {0.f, 0.f};
    for (int a0 = 0; a0 < bx; a0 += apass) {
        const int a1 = (a0 + apass < bx) ? (a0 + apass) : bx;
        const uint32_t lo = (uint32_t)(a0 * bx), span = (uint32_t)((a1 - a0) * bx);
        uint32_t *hc = hist - a0 * ncol;  // cell index of the full pair
        const bool single = a1 - a0 == bx;  // one pass: no range test
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int t = 16 * tid + 16 * nth * i;
            if (t >= L) continue;
            const uint32_t yw[4] = {ys[i].x, ys[i].y, ys[i].z, ys[i].w};
            // the wave's last lane decides whole-chunk paths (wave-uniform;
            // a per-lane test lets the compiler predicate every sample)
            const bool full = __builtin_amdgcn_readfirstlane(16 * (tid | 63) + 16 * nth * i) + 16 <= L;
            if (single && full) {
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const uint32_t sv = (regs.s[8 * i + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
                    const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                    lds_inc_key0<HZ>(hc, sv, by4, yv, 2u, zero);
                }
            } else if (full) {
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const uint32_t sv = (regs.s[8 * i + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
                    const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                    if (sv - lo < span) lds_inc_key0<HZ>(hc, sv, by4, yv, 2u, zero);
                }
            } else {
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const uint32_t sv = (regs.s[8 * i + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
                    const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                    if (t + q < L && sv - lo < span) lds_inc_key0<HZ>(hc, sv, by4, yv, 2u, zero);
                }
            }
        }
        __syncthreads();
        if (HZ && a0 == 0) {  // key 0 only occurs in the first pass
            if (tid < kWave) fold_zero_key<HZ>(hist, zero_row, tid);
            __syncthreads();
        }
        if (cdbg) {  // test hook: the pass's counts, read-only, before the scan clears them
            const int pc = (a1 - a0) * ncol;
            for (int i = tid; i < pc; i += nth) cdbg[a0 * ncol + i] = (int32_t)hist[i];
            __syncthreads();
        }
        // column-owner scan: cell (a, col) of the pass is word a * ncol + col;
        // a thread takes whole columns (or, when there are fewer columns than
        // threads, every R-th row of one), so the b_t * b_s marginal of its
        // column sums in a register and every cell costs one exchange (read
        // and clear), a compare and an add; lanes of a wave take consecutive
        // columns (conflict-free) and loop the same number of rows
        const int na = a1 - a0;
        if (ncol >= nth) {
            for (int col = tid; col < ncol; col += nth) {
                uint32_t c2 = 0;
                for (int r = 0; r < na; r++) {
                    const uint32_t v = lds_take(hist + r * ncol + col);
                    c2 += v;
                    xl(a3, v);
                }
                c2b[col] += c2;  // the column's only writer in this pass
            }
        } else {
            const int R = nth / ncol;
            if (tid < R * ncol) {
                const int col = tid % ncol;
                uint32_t c2 = 0;
                for (int r = tid / ncol; r < na; r += R) {
                    const uint32_t v = lds_take(hist + r * ncol + col);
                    c2 += v;
                    xl(a3, v);
                }
                if (c2)
                    __hip_atomic_fetch_add(c2b + col, c2, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        __syncthreads();
    }
    DF a2 = {0.f, 0.f};
    for (int c = tid; c < ncol; c += nth) xl(a2, lds_take(c2b + c));
    df_add(a3, -a2.hi, -a2.lo);
    return df_block_sum(a3, red);
}

}  // namespace tnx
