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

// Sum of a DF over the 64 lanes of a wavefront; result valid in every lane.
__device__ __forceinline__ DF df_wave_sum(DF v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        float h = __shfl_xor(v.hi, off, kWave);
        float l = __shfl_xor(v.lo, off, kWave);
        df_add(v, h, l);
    }
    return v;
}

__device__ __forceinline__ double df_to_double(DF v) { return (double)v.hi + (double)v.lo; }

// c*log2(c) lookup from the hi/lo table.
__device__ __forceinline__ void xlogx_add(DF &acc, const float2 *__restrict__ tbl, uint32_t c) {
    if (c > 1) {
        float2 t = tbl[c];
        df_add(acc, t.x, t.y);
    }
}

// Wave-local ordering of LDS traffic: all of this wave's LDS operations
// issued so far have completed before any later one is issued.
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
}

}  // namespace tnx
