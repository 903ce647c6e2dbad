// te_device.hpp -- device building blocks shared by the TE and surrogate kernels.
#pragma once
#include "common.hpp"

namespace tnx {

__device__ __forceinline__ void lds_inc(uint32_t *p) {
    __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Phase 1 (one wavefront): histogram of key = state[t] * by + y[t], t < L.
// st: target states (uint16, 8-byte aligned row), yr: source bins (int8,
// 4-byte aligned row).  Each lane takes 4 consecutive samples per step so a
// wave reads 256 contiguous source bytes and 512 state bytes per step.
__device__ __forceinline__ void hist_accumulate_wave(uint32_t *hist,
                                                     const uint16_t *__restrict__ st,
                                                     const int8_t *__restrict__ yr, int by, int L,
                                                     int lane) {
    const int Lfull = L & ~255;
    int t0 = lane * 4;
#pragma unroll 2
    for (; t0 < Lfull; t0 += 256) {
        const uint32_t yw = *reinterpret_cast<const uint32_t *>(yr + t0);
        const uint2 sw = *reinterpret_cast<const uint2 *>(st + t0);
        lds_inc(hist + (sw.x & 0xFFFFu) * by + (yw & 0xFFu));
        lds_inc(hist + (sw.x >> 16) * by + ((yw >> 8) & 0xFFu));
        lds_inc(hist + (sw.y & 0xFFFFu) * by + ((yw >> 16) & 0xFFu));
        lds_inc(hist + (sw.y >> 16) * by + (yw >> 24));
    }
    if (t0 < L) {
        const uint32_t yw = *reinterpret_cast<const uint32_t *>(yr + t0);
        const uint2 sw = *reinterpret_cast<const uint2 *>(st + t0);
        lds_inc(hist + (sw.x & 0xFFFFu) * by + (yw & 0xFFu));
        if (t0 + 1 < L) lds_inc(hist + (sw.x >> 16) * by + ((yw >> 8) & 0xFFu));
        if (t0 + 2 < L) lds_inc(hist + (sw.y & 0xFFFFu) * by + ((yw >> 16) & 0xFFu));
        if (t0 + 3 < L) lds_inc(hist + (sw.y >> 16) * by + (yw >> 24));
    }
}

// Phases 2-4 (one wavefront): S3 over the b_x^2 b_y cells and S2b over the
// b_x b_y (x_t, y_t) columns, zeroing the histogram for the next pair.
// Cell (a, b, c) sits at a * (b_x b_y) + (b * b_y + c).
__device__ __forceinline__ void hist_reduce_wave(uint32_t *hist, int bx, int by,
                                                 const float2 *__restrict__ tbl, int lane, DF &s3,
                                                 DF &s2b) {
    const int ncol = bx * by;
    DF a3 = {0.f, 0.f}, a2 = {0.f, 0.f};
    for (int col = lane; col < ncol; col += kWave) {
        uint32_t c2 = 0;
        uint32_t *p = hist + col;
        for (int aa = 0; aa < bx; aa++, p += ncol) {
            const uint32_t v = *p;
            if (v) {
                *p = 0;
                c2 += v;
                xlogx_add(a3, tbl, v);
            }
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
// accumulates the (x_t, y_t) marginal across passes.  S3, S2b valid in
// thread 0.  hist must be zero on entry and is zero on exit.
__device__ __forceinline__ void hist_pair_block(uint32_t *hist, uint32_t *c2b, int cap,
                                                const uint16_t *__restrict__ st,
                                                const int8_t *__restrict__ yr, int bx, int by,
                                                int L, const float2 *__restrict__ tbl, float *red,
                                                DF &s3, DF &s2b) {
    const int tid = threadIdx.x, nth = blockDim.x;
    const int ncol = bx * by;
    for (int i = tid; i < ncol; i += nth) c2b[i] = 0;
    int apass = cap / ncol;
    if (apass < 1) apass = 1;
    DF a3 = {0.f, 0.f}, a2 = {0.f, 0.f};
    for (int a0 = 0; a0 < bx; a0 += apass) {
        const int a1 = (a0 + apass < bx) ? (a0 + apass) : bx;
        const uint32_t lo = (uint32_t)(a0 * bx), hi = (uint32_t)(a1 * bx);
        const uint32_t base = (uint32_t)(a0 * ncol);
        __syncthreads();
        for (int t0 = tid * 4; t0 < L; t0 += 4 * nth) {
            const uint32_t yw = *reinterpret_cast<const uint32_t *>(yr + t0);
            const uint2 sw = *reinterpret_cast<const uint2 *>(st + t0);
            const uint32_t s[4] = {sw.x & 0xFFFFu, sw.x >> 16, sw.y & 0xFFFFu, sw.y >> 16};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (t0 + q < L && s[q] >= lo && s[q] < hi)
                    lds_inc(hist + s[q] * by + ((yw >> (8 * q)) & 0xFFu) - base);
            }
        }
        __syncthreads();
        const int rows = a1 - a0;
        for (int col = tid; col < ncol; col += nth) {
            uint32_t c2 = c2b[col];
            uint32_t *p = hist + col;
            for (int aa = 0; aa < rows; aa++, p += ncol) {
                const uint32_t v = *p;
                if (v) {
                    *p = 0;
                    c2 += v;
                    xlogx_add(a3, tbl, v);
                }
            }
            c2b[col] = c2;
        }
    }
    __syncthreads();
    for (int col = tid; col < ncol; col += nth) xlogx_add(a2, tbl, c2b[col]);
    df_block_sum2(a3, a2, red);
    s3 = a3;
    s2b = a2;
}

}  // namespace tnx
