// te_sort.hip -- transfer entropy of pairs by sorting their occupied keys.
//
// For pairs whose joint histogram b_t^2 b_s is far larger than the series
// (L samples), and for variables with more than 127 levels (int16 bins),
// counting into a dense LDS histogram would scan mostly empty cells.  Here
// the L joint states of a pair are keys
//   key(t) = (x_t << 30) | (y_t << 15) | x_{t+tau}      (levels < 32768)
// sorted in LDS by one workgroup (sort_lds.hpp); the runs of equal keys give
// S3 = sum c log2 c over (x_{t+tau}, x_t, y_t) and, read once more with the
// low 15 bits dropped, the runs of (x_t, y_t) give S2b.  Both sums are exact
// 64-bit fixed point (tnx_xlogx_table_fx), so
//   TE = (2^-shift (S3 - S2b) + S1 - S2a) / L
// as in the histogram kernels (S1 over x_t, S2a over (x_{t+tau}, x_t), from
// tnx_gene_prep); for wide variables tnx_gene_prep_sort gives S1 and S2a
// from the same machinery on keys (x_t << 30) | x_{t+tau}.  This is the
// occupied-key counterpart of the reference's global-memory overflow path
// (tenex/csrc/gmem.cu:27-155), which keeps b_t^2 b_s dense counters per
// pair in DRAM, and of its hash / sort entropy kernels
// (tenex/csrc/trace_hash_entropy.cu:24-46, trace_sort_entropy.cu:134-192).
#include "common.hpp"
#include "sort_lds.hpp"

namespace tnx {

// Keys of the samples 16 tid .. 16 tid + 15: (a << 30) | (b << 15) | c with
// a = x_t, b = y_t (ry NULL: no y), c = x_{t + tau}; past L the largest key.
__device__ __forceinline__ void sort_keys(uint64_t (&key)[kSortE], const int16_t *__restrict__ rx,
                                          const int16_t *__restrict__ ry, int tau, int L,
                                          int base) {
#pragma unroll
    for (int e = 0; e < kSortE; ++e) {
        const int t = base + e;
        uint64_t k = ~0ull;
        if (t < L) {
            const uint64_t a = (uint16_t)rx[t], c = (uint16_t)rx[t + tau];
            const uint64_t b = ry ? (uint64_t)(uint16_t)ry[t] : 0ull;
            k = (a << 30) | (b << 15) | c;
        }
        key[e] = k;
    }
}

__global__ void __launch_bounds__(1024)
te_sort_pair_kernel(const int16_t *__restrict__ bins, int64_t n_vars, int64_t ld, int32_t tau,
                    int32_t L, int32_t P, const int32_t *__restrict__ pairs,
                    const double *__restrict__ S1, const double *__restrict__ S2a,
                    const int64_t *__restrict__ tbl, int32_t shift, float *__restrict__ out,
                    int64_t ld_out) {
    extern __shared__ __attribute__((aligned(16))) uint64_t keys_lds[];
    __shared__ int32_t s_last[16];
    __shared__ int64_t s_sum[16];
    const int64_t p = blockIdx.x;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int x = pairs[2 * p], y = pairs[2 * p + 1];
    if (x < 0 || x >= n_vars || y < 0 || y >= n_vars) return;  // uniform
    uint64_t key[kSortE];
    sort_keys(key, bins + (int64_t)x * ld, bins + (int64_t)y * ld, tau, L, tid * kSortE);
    bitonic_sort(key, keys_lds, P, nt, tid);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < kSortE; ++e) keys_lds[e * nt + tid] = key[e];
    __syncthreads();
    const int64_t s3 = run_sum(key, keys_lds, nt, tid, L, 0, tbl, s_last);
    const int64_t s2b = run_sum(key, keys_lds, nt, tid, L, 15, tbl, s_last);
    const int64_t d = block_sum_i64(s3 - s2b, s_sum, nt, tid);
    if (tid == 0) {
        const double te = (ldexp((double)d, -shift) + (S1[x] - S2a[x])) / (double)L;
        out[ld_out > 0 ? (int64_t)x * ld_out + y : p] = (float)te;
    }
}

__global__ void __launch_bounds__(1024)
sort_marginal_kernel(const int16_t *__restrict__ bins, int64_t n_vars, int64_t ld, int32_t tau,
                     int32_t L, int32_t P, const int32_t *__restrict__ vars,
                     const int64_t *__restrict__ tbl, int32_t shift, double *__restrict__ S1,
                     double *__restrict__ S2a) {
    extern __shared__ __attribute__((aligned(16))) uint64_t keys_lds[];
    __shared__ int32_t s_last[16];
    __shared__ int64_t s_sum[16];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int x = vars[blockIdx.x];
    if (x < 0 || x >= n_vars) return;
    uint64_t key[kSortE];
    sort_keys(key, bins + (int64_t)x * ld, nullptr, tau, L, tid * kSortE);
    bitonic_sort(key, keys_lds, P, nt, tid);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < kSortE; ++e) keys_lds[e * nt + tid] = key[e];
    __syncthreads();
    // S1 over x_t: runs of key >> 30 (the zero y field dropped with
    // x_{t+tau}); S2a over (x_{t+tau}, x_t): runs of the whole key
    const int64_t s1 = block_sum_i64(run_sum(key, keys_lds, nt, tid, L, 30, tbl, s_last), s_sum,
                                     nt, tid);
    const int64_t s2 = block_sum_i64(run_sum(key, keys_lds, nt, tid, L, 0, tbl, s_last), s_sum,
                                     nt, tid);
    if (tid == 0) {
        S1[x] = ldexp((double)s1, -shift);
        S2a[x] = ldexp((double)s2, -shift);
    }
}

static int sort_size(int64_t L) {
    int P = kSortMinP;
    while (P < L) P <<= 1;
    return P;
}

}  // namespace tnx

using namespace tnx;

extern "C" int tnx_te_sort_pairs(const int16_t *bins16, int64_t n_vars, int64_t T, int64_t ld,
                                 int32_t tau, const int32_t *pairs, int64_t n_pairs,
                                 const double *S1, const double *S2a, const int64_t *xlogx_fx,
                                 int32_t fx_shift, float *out, int64_t ld_out, void *stream) {
    clear_error();
    TNX_CHECK_ARG(n_pairs >= 0 && n_pairs <= 0x7fffffffLL, "tnx_te_sort_pairs: bad n_pairs");
    if (n_pairs == 0) return TNX_OK;
    TNX_CHECK_ARG(bins16 && pairs && S1 && S2a && xlogx_fx && out,
                  "tnx_te_sort_pairs: null pointer");
    TNX_CHECK_ARG(tau >= 1 && tau < T && ld >= T, "tnx_te_sort_pairs: need 1 <= tau < T <= ld");
    TNX_CHECK_ARG(T - tau <= kSortMaxP, "tnx_te_sort_pairs: T - tau = %lld exceeds %d samples",
                  (long long)(T - tau), kSortMaxP);
    TNX_CHECK_ARG(fx_shift >= 0 && fx_shift <= 62, "tnx_te_sort_pairs: bad fx_shift");
    TNX_CHECK_ARG(ld_out >= 0, "tnx_te_sort_pairs: ld_out < 0");
    const int P = sort_size(T - tau);
    const size_t shm = (size_t)P * sizeof(uint64_t);
    hipStream_t s = as_stream(stream);
    TNX_CHECK_HIP(hipFuncSetAttribute((const void *)te_sort_pair_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(te_sort_pair_kernel, dim3((unsigned)n_pairs), dim3(P / kSortE), shm, s,
                       bins16, n_vars, ld, tau, (int32_t)(T - tau), P, pairs, S1, S2a, xlogx_fx,
                       fx_shift, out, ld_out);
    TNX_CHECK_LAUNCH("te_sort_pair_kernel");
    return TNX_OK;
}

extern "C" int tnx_gene_prep_sort(const int16_t *bins16, int64_t n_vars, int64_t T, int64_t ld,
                                  int32_t tau, const int32_t *vars, int64_t n_sel,
                                  const int64_t *xlogx_fx, int32_t fx_shift, double *S1,
                                  double *S2a, void *stream) {
    clear_error();
    TNX_CHECK_ARG(n_sel >= 0 && n_sel <= 0x7fffffffLL, "tnx_gene_prep_sort: bad n_sel");
    if (n_sel == 0) return TNX_OK;
    TNX_CHECK_ARG(bins16 && vars && xlogx_fx && S1 && S2a, "tnx_gene_prep_sort: null pointer");
    TNX_CHECK_ARG(tau >= 1 && tau < T && ld >= T, "tnx_gene_prep_sort: need 1 <= tau < T <= ld");
    TNX_CHECK_ARG(T - tau <= kSortMaxP, "tnx_gene_prep_sort: T - tau = %lld exceeds %d samples",
                  (long long)(T - tau), kSortMaxP);
    TNX_CHECK_ARG(fx_shift >= 0 && fx_shift <= 62, "tnx_gene_prep_sort: bad fx_shift");
    const int P = sort_size(T - tau);
    const size_t shm = (size_t)P * sizeof(uint64_t);
    hipStream_t s = as_stream(stream);
    TNX_CHECK_HIP(hipFuncSetAttribute((const void *)sort_marginal_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(sort_marginal_kernel, dim3((unsigned)n_sel), dim3(P / kSortE), shm, s,
                       bins16, n_vars, ld, tau, (int32_t)(T - tau), P, vars, xlogx_fx, fx_shift,
                       S1, S2a);
    TNX_CHECK_LAUNCH("sort_marginal_kernel");
    return TNX_OK;
}
