// te_device.hpp -- device building blocks shared by the TE and surrogate kernels.
#pragma once
#include "common.hpp"

// Cells of privatised histogram copies a pair may use (see wave_copies).
#ifndef TNX_WAVE_COPY_BUDGET
#define TNX_WAVE_COPY_BUDGET 2048
#endif
#ifndef TNX_BLOCK_COPY_BUDGET
#define TNX_BLOCK_COPY_BUDGET 1
#endif

namespace tnx {

__device__ __forceinline__ void lds_inc(uint32_t *p) {
    __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Cell of key state * by + y as a byte offset: states < 2^16 and by <= 127,
// so the product fits the full-rate 24-bit multiplier (v_mul_u32_u24, which
// also takes the 16-bit half of a packed state register directly); by4 is
// 4 * by.  Three VALU operations per sample: multiply, shift of the source
// byte, three-way add with the copy base.
__device__ __forceinline__ void lds_inc_key(uint32_t *hist, uint32_t sv, uint32_t by4,
                                            uint32_t yv) {
    const uint32_t off = __umul24(sv, by4) + (yv << 2);
    lds_inc(reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(hist) + off));
}

// 16 samples of one lane: 8 packed 16-bit states, 4 source dwords.
__device__ __forceinline__ void acc16(uint32_t *hist, const uint32_t *sreg, const uint4 y4,
                                      uint32_t by) {
    const uint32_t by4 = by << 2;
    const uint32_t yw[4] = {y4.x, y4.y, y4.z, y4.w};
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const uint32_t sv = (sreg[q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
        const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
        lds_inc_key(hist, sv, by4, yv);
    }
}

__device__ __forceinline__ void acc16_tail(uint32_t *hist, const uint32_t *sreg, const uint4 y4,
                                           uint32_t by, int n_valid) {
    const uint32_t by4 = by << 2;
    const uint32_t yw[4] = {y4.x, y4.y, y4.z, y4.w};
#pragma unroll
    for (int q = 0; q < 16; q++) {
        if (q < n_valid) {
            const uint32_t sv = (sreg[q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
            const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
            lds_inc_key(hist, sv, by4, yv);
        }
    }
}

// Phase 1 (one wavefront), target states read from memory: key =
// state[t] * by + y[t] for t < L.  Lane `lane` takes 16 consecutive samples
// per step (one dwordx4 of source bytes, two of states), two steps in
// flight.  Rows must be padded to a multiple of 64 elements.
__device__ __forceinline__ void hist_accumulate_mem(uint32_t *hist, const uint16_t *__restrict__ st,
                                                    const int8_t *__restrict__ yr, uint32_t by,
                                                    int L, int lane) {
    int t = 16 * lane;
    for (; t + 1024 + 16 <= L; t += 2048) {
        const uint4 ya = *reinterpret_cast<const uint4 *>(yr + t);
        const uint4 sa0 = *reinterpret_cast<const uint4 *>(st + t);
        const uint4 sa1 = *reinterpret_cast<const uint4 *>(st + t + 8);
        const uint4 yb = *reinterpret_cast<const uint4 *>(yr + t + 1024);
        const uint4 sb0 = *reinterpret_cast<const uint4 *>(st + t + 1024);
        const uint4 sb1 = *reinterpret_cast<const uint4 *>(st + t + 1032);
        const uint32_t sa[8] = {sa0.x, sa0.y, sa0.z, sa0.w, sa1.x, sa1.y, sa1.z, sa1.w};
        const uint32_t sb[8] = {sb0.x, sb0.y, sb0.z, sb0.w, sb1.x, sb1.y, sb1.z, sb1.w};
        acc16(hist, sa, ya, by);
        acc16(hist, sb, yb, by);
    }
    for (; t < L; t += 1024) {
        const uint4 ya = *reinterpret_cast<const uint4 *>(yr + t);
        const uint4 sa0 = *reinterpret_cast<const uint4 *>(st + t);
        const uint4 sa1 = *reinterpret_cast<const uint4 *>(st + t + 8);
        const uint32_t sa[8] = {sa0.x, sa0.y, sa0.z, sa0.w, sa1.x, sa1.y, sa1.z, sa1.w};
        acc16_tail(hist, sa, ya, by, L - t);
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
        if (t < L) {
            const uint4 a = *reinterpret_cast<const uint4 *>(st + t);
            const uint4 b = *reinterpret_cast<const uint4 *>(st + t + 8);
            r.s[8 * i + 0] = a.x; r.s[8 * i + 1] = a.y; r.s[8 * i + 2] = a.z; r.s[8 * i + 3] = a.w;
            r.s[8 * i + 4] = b.x; r.s[8 * i + 5] = b.y; r.s[8 * i + 6] = b.z; r.s[8 * i + 7] = b.w;
        }
    }
}

// Phase 1 with the target in registers: 16 source bytes per lane per chunk
// (one dwordx4 load), 16 LDS increments.
template <int NI>
__device__ __forceinline__ void hist_accumulate_regs(uint32_t *hist, TargetRegs<NI> &r,
                                                     const int8_t *__restrict__ yr, uint32_t by,
                                                     int L, int lane) {
    // Keep the packed states opaque per call: otherwise the compiler hoists
    // the per-sample 16-bit extractions out of the caller's source loop and
    // spends a register on every sample.
#pragma unroll
    for (int k = 0; k < NI * 8; k++) asm volatile("" : "+v"(r.s[k]));
    const uint32_t *sreg = r.s;
    const uint32_t by4 = by << 2;
#pragma unroll
    for (int i = 0; i < NI; i++) {
        const int t = 16 * lane + 1024 * i;
        if (t >= L) break;
        const uint4 y4 = *reinterpret_cast<const uint4 *>(yr + t);
        const uint32_t yw[4] = {y4.x, y4.y, y4.z, y4.w};
        if (t + 16 <= L) {
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const uint32_t sv = (sreg[8 * i + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
                const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                lds_inc_key(hist, sv, by4, yv);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 16; q++) {
                if (t + q < L) {
                    const uint32_t sv = (sreg[8 * i + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
                    const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                    lds_inc_key(hist, sv, by4, yv);
                }
            }
        }
    }
}

// Privatised histograms: a pair with `cells` cells gets R copies at an odd
// stride S >= cells inside the wave's cap-cell LDS slice; lane l counts into
// copy l % R.  Lanes whose samples share a key then mostly hit different
// copies, and the odd stride puts one key's copies in different LDS banks,
// which removes the same-address serialisation that dominates small
// histograms (a pair with 8 cells would otherwise see ~20 lanes of a wave
// on one address).
struct WaveCopies {
    int S, R;
};

// budget bounds R * S: the reduction reads and zeroes R * S cells per pair,
// so copies are only worth it for small histograms.
__device__ __forceinline__ WaveCopies wave_copies(int cells, int cap, int budget) {
    WaveCopies c;
    c.S = cells | 1;
    int lim = cap < budget ? cap : budget;
    c.R = lim / c.S;
    if (c.R > 32) c.R = 32;
    if (c.R < 1) c.R = 1;
    return c;
}

// Sum of the R copies of a cell; zeroes them.
__device__ __forceinline__ uint32_t take_cell(uint32_t *hist, int cell, const WaveCopies &cp) {
    uint32_t v = 0;
    uint32_t *p = hist + cell;
    for (int r = 0; r < cp.R; r++, p += cp.S) {
        v += *p;
        *p = 0;
    }
    return v;
}

// Read an LDS cell and leave zero behind (one ds_wrxchg_rtn_b32).
__device__ __forceinline__ uint32_t lds_take(uint32_t *p) {
    return __hip_atomic_exchange(p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Fold the R copies of a pair's histogram onto copy 0 with every lane busy.
// Small histograms (S < 64) split the lanes into G = min(64 / S, R) groups:
// lane (k, g) takes cell k of copies g, g + G, ...; the G partial sums then
// meet in copies 0 .. G-1 and group 0 adds them up.  Larger ones walk the
// cells 64 at a time.  On exit copy 0 holds the totals of cells [0, cells)
// and every other copy is zero.  One wavefront; wave_lds_fence between
// dependent LDS phases.
__device__ __forceinline__ void fold_wave_copies(uint32_t *hist, int cells, const WaveCopies &cp,
                                                 int lane) {
    const int S = cp.S, R = cp.R;
    if (R <= 1) return;
    if (S < kWave) {
        int G = kWave / S;
        if (G > R) G = R;
        const int k = lane % S, g = lane / S;
        const bool act = g < G && k < cells;
        uint32_t v = 0;
        if (act)
            for (int r = g; r < R; r += G) v += lds_take(hist + r * S + k);
        wave_lds_fence();
        if (act) hist[g * S + k] = v;
        wave_lds_fence();
        if (act && g == 0) {
            uint32_t t = v;
            for (int gg = 1; gg < G; gg++) t += lds_take(hist + gg * S + k);
            hist[k] = t;
        }
    } else {
        for (int k = lane; k < cells; k += kWave) {
            uint32_t t = hist[k];
            for (int r = 1; r < R; r++) t += lds_take(hist + r * S + k);
            hist[k] = t;
        }
    }
    wave_lds_fence();
}

// Phases 2-4 with copies: fold the copies, then per (x_t, y_t) column the
// b_x cells of x_{t+tau}: S3 from every cell, S2b from the column sums.
// Leaves the histogram zero.
__device__ __forceinline__ void hist_reduce_wave_copies(uint32_t *hist, int bx, int by,
                                                        const WaveCopies &cp,
                                                        const float2 *__restrict__ tbl, int lane,
                                                        DF &s3, DF &s2b) {
    const int ncol = bx * by;
    fold_wave_copies(hist, bx * ncol, cp, lane);
    DF a3 = {0.f, 0.f}, a2 = {0.f, 0.f};
    for (int col = lane; col < ncol; col += kWave) {
        uint32_t c2 = 0;
        for (int aa = 0; aa < bx; aa++) {
            const uint32_t v = lds_take(hist + aa * ncol + col);
            c2 += v;
            xlogx_add(a3, tbl, v);
        }
        xlogx_add(a2, tbl, c2);
    }
    s3 = df_wave_sum(a3);
    s2b = df_wave_sum(a2);
}

// Workgroup reduction of two DF values; result valid in thread 0.
// red: shared float[4 * n_waves].
__device__ __forceinline__ void df_block_sum2(DF &u, DF &v, float *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    u = df_wave_sum(u);
    v = df_wave_sum(v);
    if (lane == 0) {
        red[4 * w + 0] = u.hi;
        red[4 * w + 1] = u.lo;
        red[4 * w + 2] = v.hi;
        red[4 * w + 3] = v.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DF su = {0.f, 0.f}, sv = {0.f, 0.f};
        for (int i = 0; i < nw; i++) {
            df_add(su, red[4 * i + 0], red[4 * i + 1]);
            df_add(sv, red[4 * i + 2], red[4 * i + 3]);
        }
        u = su;
        v = sv;
    }
    __syncthreads();
}

// Whole workgroup on one pair with a histogram of cap cells: passes over
// ranges [a0, a1) of x_{t+tau} so any b_x^2 b_y fits; c2b (b_x b_y cells)
// accumulates the (x_t, y_t) marginal across passes.  The target's states
// are in registers (thread tid owns samples 16 tid + 16 nth i + q, NB chunks,
// L <= 16 nth NB); a pass with few cells counts into privatised copies.
// S3, S2b valid in thread 0.  hist must be zero on entry and is zero on exit.
template <int NB>
__device__ __forceinline__ void hist_pair_block(uint32_t *hist, uint32_t *c2b, int cap,
                                                const TargetRegs<NB> &regs,
                                                const int8_t *__restrict__ yr, int bx, int by,
                                                int L, const float2 *__restrict__ tbl, float *red,
                                                DF &s3, DF &s2b) {
    const int tid = threadIdx.x, nth = blockDim.x;
    const int ncol = bx * by;
    const uint32_t by4 = (uint32_t)by << 2;
    for (int i = tid; i < ncol; i += nth) c2b[i] = 0;
    int apass = cap / ncol;
    if (apass < 1) apass = 1;
    DF a3 = {0.f, 0.f}, a2 = {0.f, 0.f};
    for (int a0 = 0; a0 < bx; a0 += apass) {
        const int a1 = (a0 + apass < bx) ? (a0 + apass) : bx;
        const uint32_t lo = (uint32_t)(a0 * bx), hi = (uint32_t)(a1 * bx);
        const int cells = (a1 - a0) * ncol;
        const WaveCopies cp = wave_copies(cells, cap, TNX_BLOCK_COPY_BUDGET);
        uint32_t *hc = hist + (tid % cp.R) * cp.S - a0 * ncol;  // cell index of the full pair
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int t = 16 * tid + 16 * nth * i;
            if (t < L) {
                const uint4 y4 = *reinterpret_cast<const uint4 *>(yr + t);
                const uint32_t yw[4] = {y4.x, y4.y, y4.z, y4.w};
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const uint32_t sv = (regs.s[8 * i + (q >> 1)] >> (16 * (q & 1))) & 0xFFFFu;
                    const uint32_t yv = (yw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                    if (t + q < L && sv >= lo && sv < hi) lds_inc_key(hc, sv, by4, yv);
                }
            }
        }
        __syncthreads();
        // flattened scan over the pass's cells: coalesced, independent lookups
        const int step = nth % ncol;
        int col = tid % ncol;
#pragma unroll 4
        for (int i = tid; i < cells; i += nth) {
            const uint32_t v = take_cell(hist, i, cp);
            const float2 tv = tbl[v];  // tbl[0] = tbl[1] = 0
            df_add(a3, tv.x, tv.y);
            if (v) __hip_atomic_fetch_add(c2b + col, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            col += step;
            if (col >= ncol) col -= ncol;
        }
    }
    __syncthreads();
    for (int col = tid; col < ncol; col += nth) xlogx_add(a2, tbl, c2b[col]);
    df_block_sum2(a3, a2, red);
    s3 = a3;
    s2b = a2;
}

}  // namespace tnx
