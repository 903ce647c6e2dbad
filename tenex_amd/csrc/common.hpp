// common.hpp -- shared helpers for the gfx950 transfer-entropy kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/tenex_hip.h"

namespace tnx {

constexpr int kWave = 64;

// ── error reporting (thread-local last error) ───────────────────────────────
void set_error(const char *fmt, ...);
void clear_error();

#define TNX_CHECK_ARG(cond, ...)                 \
    do {                                         \
        if (!(cond)) {                           \
            ::tnx::set_error(__VA_ARGS__);       \
            return TNX_E_ARG;                    \
        }                                        \
    } while (0)

#define TNX_CHECK_HIP(expr)                                                        \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            ::tnx::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                             __FILE__, __LINE__);                                  \
            return TNX_E_HIP;                                                      \
        }                                                                          \
    } while (0)

// After kernel launches: report launch errors without synchronising.
#define TNX_CHECK_LAUNCH(name)                                                     \
    do {                                                                           \
        hipError_t _e = hipGetLastError();                                         \
        if (_e != hipSuccess) {                                                    \
            ::tnx::set_error("launch of %s failed: %s", name, hipGetErrorString(_e)); \
            return TNX_E_HIP;                                                      \
        }                                                                          \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// ── double-float (hi + lo) accumulation in fp32 ─────────────────────────────
// Knuth TwoSum: s + e == a + b exactly.
struct DF {
    float hi, lo;
};

__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// acc += (xh + xl), compensated.
__device__ __forceinline__ void df_add(DF &acc, float xh, float xl) {
    float s, e;
    two_sum(acc.hi, xh, s, e);
    e += acc.lo + xl;
    float s2 = s + e;               // renormalise (Fast2Sum, |s| >= |e|)
    acc.lo = e - (s2 - s);
    acc.hi = s2;
}

// v from lane (l - N) mod 16 of the same row of 16 lanes (DPP row_ror:N).
template <int N>
__device__ __forceinline__ float row_ror(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x120 + N, 0xF, 0xF,
                                                      false));
}

// Sum of a DF over the 64 lanes of a wavefront; result valid in every lane.
// Rotate-and-add inside each row of 16 lanes (DPP, no LDS traffic), then
// the four row totals through readlane.
__device__ __forceinline__ DF df_wave_sum(DF v) {
    df_add(v, row_ror<1>(v.hi), row_ror<1>(v.lo));
    df_add(v, row_ror<2>(v.hi), row_ror<2>(v.lo));
    df_add(v, row_ror<4>(v.hi), row_ror<4>(v.lo));
    df_add(v, row_ror<8>(v.hi), row_ror<8>(v.lo));
    auto lane_f = [](float f, int l) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(f), l));
    };
    DF t = {lane_f(v.hi, 0), lane_f(v.lo, 0)};
    df_add(t, lane_f(v.hi, 16), lane_f(v.lo, 16));
    df_add(t, lane_f(v.hi, 32), lane_f(v.lo, 32));
    df_add(t, lane_f(v.hi, 48), lane_f(v.lo, 48));
    return t;
}

__device__ __forceinline__ double df_to_double(DF v) { return (double)v.hi + (double)v.lo; }

// Integer sum over the 64 lanes (every lane gets it).
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    v += __builtin_amdgcn_update_dpp(0, (int)v, 0x121, 0xF, 0xF, false);  // row_ror:1
    v += __builtin_amdgcn_update_dpp(0, (int)v, 0x122, 0xF, 0xF, false);  // row_ror:2
    v += __builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false);  // row_ror:4
    v += __builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);  // row_ror:8
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 0) +
           (uint32_t)__builtin_amdgcn_readlane((int)v, 16) +
           (uint32_t)__builtin_amdgcn_readlane((int)v, 32) +
           (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
}

// c*log2(c) lookup from the hi/lo table.
__device__ __forceinline__ void xlogx_add(DF &acc, const float2 *__restrict__ tbl, uint32_t c) {
    if (c > 1) {
        float2 t = tbl[c];
        df_add(acc, t.x, t.y);
    }
}

// c*log2(c) entries kept in LDS by kernels that look up many small counts
// (the global table is L + 1 entries; small counts dominate the lookups).
constexpr int kSmallTbl = 256;

// As xlogx_add, with counts below kSmallTbl read from the LDS copy ltbl.
__device__ __forceinline__ void xlogx_add_l(DF &acc, const float2 *ltbl,
                                            const float2 *__restrict__ tbl, uint32_t c) {
    if (c > 1) {
        // the LDS copy through an LDS-typed pointer: a select between the two
        // tables' pointers would become one flat load
        typedef const float __attribute__((address_space(3))) *lds_f;
        float hi, lo;
        if (c < (uint32_t)kSmallTbl) {
            const lds_f lp = (lds_f)(const void *)ltbl;
            hi = lp[2 * c];
            lo = lp[2 * c + 1];
        } else {
            const float2 t = tbl[c];
            hi = t.x;
            lo = t.y;
        }
        df_add(acc, hi, lo);
    }
}

// c*log2(c) lookups as objects, so one reduction body serves every table
// placement: XlTable reads the global hi/lo table, XlSmall takes counts below
// kSmallTbl from an LDS copy.
struct XlTable {
    const float2 *__restrict__ tbl;
    __device__ __forceinline__ void operator()(DF &acc, uint32_t c) const { xlogx_add(acc, tbl, c); }
};

struct XlSmall {
    const float2 *ltbl;
    const float2 *__restrict__ tbl;
    __device__ __forceinline__ void operator()(DF &acc, uint32_t c) const {
        xlogx_add_l(acc, ltbl, tbl, c);
    }
};

// Wave-local ordering of LDS traffic: all of this wave's LDS operations
// issued so far have completed before any later one is issued.
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
}

}  // namespace tnx
