// microbench.hip -- hardware ceilings used as roofline peaks.
//
// tnx_bench_lds_atomics: non-returning ds_add_u32 throughput over the whole
// chip.  mode 0: conflict-free (lane i hits bank i); mode 1: random keys in a
// 2048-cell histogram (the access pattern of a TE histogram with uniform
// keys).  Each of the n_blocks x 256 threads performs iters x 8 increments.
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

}  // namespace tnx

using namespace tnx;

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
