// microbench.hip -- hardware ceilings used as roofline peaks.
//
// tnx_bench_lds_atomics: non-returning ds_add_u32 throughput over the whole
// chip.  mode 0: conflict-free (lane i hits bank i); mode 1: random keys in a
// 2048-cell histogram (the access pattern of a TE histogram with uniform
// keys).  Each of the n_blocks x 256 threads performs iters x 8 increments.
//
// tnx_bench_mfma: matrix-core throughput from registers alone (four
// independent accumulators per wave, no memory in the loop).  mode 0:
// v_mfma_scale_f32_32x32x64_f8f6f4 on FP4 operands; mode 1: the same on
// FP8; mode 2: v_mfma_i32_32x32x32_i8.  Each of the n_blocks x 4 waves
// issues iters x 4 MFMAs.
#include "common.hpp"
#include "te_device.hpp"

namespace tnx {

__global__ __launch_bounds__(256) void lds_atomic_kernel(uint32_t *__restrict__ out, int iters,
                                                         int mode) {
    __shared__ uint32_t h[4096];
    const int tid = threadIdx.x;
    for (int i = tid; i < 4096; i += 256) h[i] = 0;
    __syncthreads();
    uint32_t k = (uint32_t)tid;
    uint32_t rng = (uint32_t)(tid * 2654435761u) ^ (uint32_t)(blockIdx.x * 40503u);
    if (mode == 0) {
        for (int it = 0; it < iters; it++) {
#pragma unroll
            for (int u = 0; u < 8; u++) lds_inc(h + ((k + 256u * u) & 4095u));
            k += 64u;
        }
    } else {
        for (int it = 0; it < iters; it++) {
#pragma unroll
            for (int u = 0; u < 8; u++) {
                rng = rng * 1664525u + 1013904223u;
                lds_inc(h + ((rng >> 21) & 2047u));
            }
        }
    }
    __syncthreads();
    if (tid == 0) out[blockIdx.x] = h[0] + h[2047];
}

typedef int mb_v4i __attribute__((ext_vector_type(4)));
typedef int mb_v8i __attribute__((ext_vector_type(8)));
typedef int mb_v16i __attribute__((ext_vector_type(16)));
typedef float mb_v16f __attribute__((ext_vector_type(16)));

template <int MODE>
__global__ __launch_bounds__(256) void mfma_rate_kernel(float *__restrict__ out, int iters) {
    mb_v16f f[4] = {};
    mb_v16i g[4] = {};
    const int t = (int)threadIdx.x, b = (int)blockIdx.x;
    const mb_v8i a = {b, t, 3, 4, 5, 6, 7, 8};
    const mb_v8i c = {t, b, 1, 2, 3, 4, 5, 6};
    const mb_v4i a4 = {b, t, 3, 4}, c4 = {t, 5, 6, b};
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if constexpr (MODE == 0)
                f[j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, c, f[j], 4, 4, 0,
                                                                      0x7F7F7F7F, 0, 0x7F7F7F7F);
            else if constexpr (MODE == 1)
                f[j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, c, f[j], 0, 0, 0,
                                                                      0x7F7F7F7F, 0, 0x7F7F7F7F);
            else
                g[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a4, c4, g[j], 0, 0, 0);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int q = 0; q < 16; q++) s += f[j][q] + (float)g[j][q];
    out[(int64_t)b * 256 + t] = s;  // keeps the products live
}

}  // namespace tnx

using namespace tnx;

extern "C" int tnx_bench_mfma(int32_t mode, int64_t n_blocks, int32_t iters, float *out,
                              void *stream) {
    clear_error();
    TNX_CHECK_ARG(out && n_blocks > 0 && n_blocks < (1LL << 31) && iters > 0 && mode >= 0 &&
                      mode <= 2,
                  "tnx_bench_mfma: bad arguments (out needs n_blocks x 256 floats)");
    hipStream_t s = as_stream(stream);
    if (mode == 0) hipLaunchKernelGGL(mfma_rate_kernel<0>, dim3((unsigned)n_blocks), dim3(256), 0, s, out, iters);
    else if (mode == 1) hipLaunchKernelGGL(mfma_rate_kernel<1>, dim3((unsigned)n_blocks), dim3(256), 0, s, out, iters);
    else hipLaunchKernelGGL(mfma_rate_kernel<2>, dim3((unsigned)n_blocks), dim3(256), 0, s, out, iters);
    TNX_CHECK_LAUNCH("mfma_rate_kernel");
    return TNX_OK;
}

extern "C" int tnx_bench_lds_atomics(int32_t mode, int64_t n_blocks, int32_t iters, uint32_t *out,
                                     void *stream) {
    clear_error();
    TNX_CHECK_ARG(out && n_blocks > 0 && n_blocks < (1LL << 31) && iters > 0,
                  "tnx_bench_lds_atomics: bad arguments");
    hipLaunchKernelGGL(lds_atomic_kernel, dim3((unsigned)n_blocks), dim3(256), 0,
                       as_stream(stream), out, iters, mode);
    TNX_CHECK_LAUNCH("lds_atomic_kernel");
    return TNX_OK;
}
