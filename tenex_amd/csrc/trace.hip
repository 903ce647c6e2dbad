// trace.hip -- TRACE OutTE / InTE: entropies of multivariate states.
//
// For a variable X with d connected variables Y_1..Y_d (its network targets
// for OutTE, its network sources for InTE), both quantities are sums of
// four joint entropies over the L = T - tau samples:
//   OutTE(X) = H(Yf,Yp) - H(Yp) - H(Yf,Xp,Yp) + H(Xp,Yp)
//   InTE(X)  = H(Xf,Xp) - H(Xp) - H(Xf,Xp,Yp) + H(Xp,Yp)
// (f = sample t + tau, p = sample t, Y = the d connected variables jointly).
// A state is keyed by the reference's 64-bit polynomial hash with wrapping
// arithmetic, h(A, B) = h(A) * P^|B| + h(B), P = 1,000,000,007
// (tenex/csrc/trace_sort_entropy.cu:27-96, tenex/inference/trace.py:475-547),
// so the multiplicities counted here are those of the reference's keys.
//
// One workgroup per (variable, entropy term): each thread builds the keys of
// 16 consecutive samples in registers (the bytes of each connected row read
// as aligned 16-byte buffer loads), the workgroup sorts the keys with a
// bitonic network whose stages run in registers (partner in the same
// thread), across lanes (partner in the same wave) or through LDS (partner
// in another wave), and every run of equal keys adds c*log2(c) from a
// 64-bit fixed-point table.  Sums are exact integers, so
//   TE * L = (-S0 + S1 + S2 - S3) * 2^-shift
// is exact up to the table's rounding and the result does not depend on the
// order of anything.  Replaces the hash kernels, the CUB segmented radix
// sort and the one-thread-per-segment entropy loop of
// fused_outte_4way_launch / fused_inte_4way_launch
// (tenex/csrc/trace_sort_entropy.cu:191-304).
#include "common.hpp"
#include "sort_lds.hpp"
#include "te_device.hpp"

namespace tnx {

constexpr uint64_t kTraceHashP = 1000000007ull;
constexpr int kTraceE = kSortE;          // keys per thread
constexpr int kTraceMinP = kSortMinP;
constexpr int kTraceMaxP = kSortMaxP;

// Bytes [sh, sh + 16) of the 32-byte window (a, b); sh uniform in [0, 16).
__device__ __forceinline__ void window16(const uint4 &a, const uint4 &b, int sh, uint32_t (&w)[4]) {
    const uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    const uint32_t s = (uint32_t)(sh & 3);
    switch (sh >> 2) {
        case 0:
#pragma unroll
            for (int i = 0; i < 4; ++i) w[i] = __builtin_amdgcn_alignbyte(v[i + 1], v[i], s);
            break;
        case 1:
#pragma unroll
            for (int i = 0; i < 4; ++i) w[i] = __builtin_amdgcn_alignbyte(v[i + 2], v[i + 1], s);
            break;
        case 2:
#pragma unroll
            for (int i = 0; i < 4; ++i) w[i] = __builtin_amdgcn_alignbyte(v[i + 3], v[i + 2], s);
            break;
        default:
#pragma unroll
            for (int i = 0; i < 4; ++i) w[i] = __builtin_amdgcn_alignbyte(v[i + 4], v[i + 3], s);
            break;
    }
}

// h[e] = h[e] * P + row[t0 + e], e < 16, for the row `var` of bins (t0 =
// 16 tid + off).  Rows past n_vars read as zero columns (no access).
__device__ __forceinline__ void hash_row(uint64_t (&h)[kTraceE], const int8_t *bins, int64_t ld,
                                         int64_t n_vars, int64_t var, int base, int off) {
    uint32_t w[4] = {0u, 0u, 0u, 0u};
    if (var >= 0 && var < n_vars) {
        const __amdgpu_buffer_rsrc_t r = row_rsrc(bins + var * ld, (uint32_t)ld);
        const int a = base + (off & ~15);
        const uint4 lo = row_load16(r, (uint32_t)a, 0u);
        const uint4 hi = row_load16(r, (uint32_t)a + 16u, 0u);
        window16(lo, hi, off & 15, w);
    }
#pragma unroll
    for (int e = 0; e < kTraceE; ++e) {
        const uint64_t v = (w[e >> 2] >> (8 * (e & 3))) & 0xffu;
        h[e] = h[e] * kTraceHashP + v;
    }
}

__device__ __forceinline__ void hash_single(uint64_t (&h)[kTraceE], const int8_t *bins,
                                            int64_t ld, int64_t n_vars, int64_t var, int base,
                                            int off) {
#pragma unroll
    for (int e = 0; e < kTraceE; ++e) h[e] = 0;
    hash_row(h, bins, ld, n_vars, var, base, off);
}

// One (job, entropy term): blockIdx.x = job, blockIdx.y = term.  mode 0:
// OutTE (var = X, conn = its targets), mode 1: InTE (conn = its sources).
__global__ void __launch_bounds__(1024)
trace_entropy_kernel(const int8_t *__restrict__ bins, int64_t n_vars, int64_t ld, int32_t tau,
                     int32_t L, int32_t P, int32_t mode, const int32_t *__restrict__ var_ids,
                     const int32_t *__restrict__ conn, const int64_t *__restrict__ conn_off,
                     const int64_t *__restrict__ xlogx_fx, int64_t *__restrict__ sums) {
    extern __shared__ __attribute__((aligned(16))) uint64_t keys_lds[];
    __shared__ int32_t s_wave_last[16];
    __shared__ int64_t s_wave_sum[16];
    const int job = blockIdx.x, term = blockIdx.y;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    const int base = tid * kTraceE;
    const int64_t c0 = conn_off[job], d = conn_off[job + 1] - c0;
    const int64_t x = var_ids[job];
    uint64_t Pd = 1;
    for (int64_t j = 0; j < d; ++j) Pd *= kTraceHashP;

    uint64_t key[kTraceE], aux[kTraceE];
    if (mode == 0) {
        // OutTE: Yf (conn at +tau), Yp (conn at 0), Xp (x at 0)
        //   0: h(Yf) Pd + h(Yp)   1: h(Yp)   2: (h(Yf) P + Xp) Pd + h(Yp)   3: Xp Pd + h(Yp)
#pragma unroll
        for (int e = 0; e < kTraceE; ++e) key[e] = aux[e] = 0;
        if (term == 0 || term == 2)
            for (int64_t j = 0; j < d; ++j) hash_row(key, bins, ld, n_vars, conn[c0 + j], base, tau);
        if (term >= 2) hash_row(key, bins, ld, n_vars, x, base, 0);
        for (int64_t j = 0; j < d; ++j) hash_row(aux, bins, ld, n_vars, conn[c0 + j], base, 0);
        if (term == 1) {
#pragma unroll
            for (int e = 0; e < kTraceE; ++e) key[e] = aux[e];
        } else {
#pragma unroll
            for (int e = 0; e < kTraceE; ++e) key[e] = key[e] * Pd + aux[e];
        }
    } else {
        // InTE: Xf (x at +tau), Xp (x at 0), Yp (conn at 0)
        //   0: Xf P + Xp   1: Xp   2: (Xf P + Xp) Pd + h(Yp)   3: Xp Pd + h(Yp)
#pragma unroll
        for (int e = 0; e < kTraceE; ++e) key[e] = aux[e] = 0;
        if (term == 0 || term == 2) hash_row(key, bins, ld, n_vars, x, base, tau);
        hash_row(key, bins, ld, n_vars, x, base, 0);
        if (term >= 2) {
            for (int64_t j = 0; j < d; ++j) hash_row(aux, bins, ld, n_vars, conn[c0 + j], base, 0);
#pragma unroll
            for (int e = 0; e < kTraceE; ++e) key[e] = key[e] * Pd + aux[e];
        }
    }
    // samples past L sort last as the largest key (a real key equal to it
    // joins their run, which is cut at L below)
#pragma unroll
    for (int e = 0; e < kTraceE; ++e)
        if (base + e >= L) key[e] = ~0ull;

    bitonic_sort(key, keys_lds, P, nt, tid);

    // runs of equal keys: start index of the run holding each element by a
    // max-scan of run starts; every run ending below L adds tbl[length]
    __syncthreads();
#pragma unroll
    for (int e = 0; e < kTraceE; ++e) keys_lds[e * nt + tid] = key[e];
    __syncthreads();
    const uint64_t prev = tid > 0 ? keys_lds[(kTraceE - 1) * nt + tid - 1] : ~key[0];
    const uint64_t next = tid + 1 < nt ? keys_lds[tid + 1] : ~key[kTraceE - 1];
    int last = -1;
#pragma unroll
    for (int e = 0; e < kTraceE; ++e) {
        const bool start = (e == 0 ? key[0] != prev : key[e] != key[e - 1]);
        if (start) last = base + e;
    }
    int incl = last;  // inclusive max-scan over threads
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
        const int o = __shfl_up(incl, m);
        if (lane >= m) incl = o > incl ? o : incl;
    }
    if (lane == kWave - 1) s_wave_last[wid] = incl;
    __syncthreads();
    int carry = __shfl_up(incl, 1);
    if (lane == 0) carry = -1;
    for (int w = 0; w < wid; ++w) carry = s_wave_last[w] > carry ? s_wave_last[w] : carry;
    int64_t acc = 0;
    int cur = carry;
#pragma unroll
    for (int e = 0; e < kTraceE; ++e) {
        const int i = base + e;
        const bool start = (e == 0 ? key[0] != prev : key[e] != key[e - 1]);
        if (start) cur = i;
        const bool end = (i == L - 1) || (e + 1 < kTraceE ? key[e + 1] != key[e] : next != key[e]);
        if (i < L && end) acc += xlogx_fx[i - cur + 1];
    }
    acc = wave_sum_i64_t(acc);
    if (lane == 0) s_wave_sum[wid] = acc;
    __syncthreads();
    if (tid == 0) {
        int64_t s = 0;
        for (int w = 0; w < nt / kWave; ++w) s += s_wave_sum[w];
        sums[(int64_t)job * 4 + term] = s;
    }
}

// te[job] = max(0, -S0 + S1 + S2 - S3) * 2^-shift / L (clamped at zero as
// the reference's max(..., 0)).
__global__ void trace_combine_kernel(const int64_t *__restrict__ sums, int64_t n_jobs,
                                     int32_t shift, int32_t L, float *__restrict__ te) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_jobs) return;
    const int64_t *s = sums + 4 * j;
    const int64_t v = -s[0] + s[1] + s[2] - s[3];
    te[j] = v > 0 ? (float)(ldexp((double)v, -shift) / (double)L) : 0.0f;
}

}  // namespace tnx

using namespace tnx;

extern "C" int tnx_trace_entropy(const int8_t *bins, int64_t n_vars, int64_t T, int64_t ld,
                                 int32_t tau, int32_t mode, const int32_t *var_ids,
                                 const int32_t *conn, const int64_t *conn_off, int64_t n_jobs,
                                 const int64_t *xlogx_fx, int32_t fx_shift, int64_t *sums_out,
                                 float *te_out, void *stream) {
    clear_error();
    TNX_CHECK_ARG(n_jobs >= 0, "tnx_trace_entropy: n_jobs < 0");
    if (n_jobs == 0) return TNX_OK;
    TNX_CHECK_ARG(bins && var_ids && conn_off && xlogx_fx && sums_out,
                  "tnx_trace_entropy: null pointer");
    TNX_CHECK_ARG(mode == 0 || mode == 1, "tnx_trace_entropy: mode must be 0 (OutTE) or 1 (InTE)");
    TNX_CHECK_ARG(tau >= 1 && tau < T, "tnx_trace_entropy: need 1 <= tau < T");
    TNX_CHECK_ARG(ld >= T && ld % 16 == 0 && ld <= (1LL << 31),
                  "tnx_trace_entropy: ld must be >= T, a multiple of 16 and < 2^31");
    const int64_t L = T - tau;
    TNX_CHECK_ARG(L <= kTraceMaxP,
                  "tnx_trace_entropy: T - tau = %lld exceeds the in-LDS sort's %d samples",
                  (long long)L, kTraceMaxP);
    TNX_CHECK_ARG(fx_shift >= 0 && fx_shift <= 62, "tnx_trace_entropy: bad fx_shift %d", fx_shift);
    TNX_CHECK_ARG(n_jobs <= 0x7fffffffLL, "tnx_trace_entropy: too many jobs");
    int P = kTraceMinP;
    while (P < L) P <<= 1;
    const int nt = P / kTraceE;
    const size_t shm = (size_t)P * sizeof(uint64_t);
    hipStream_t s = as_stream(stream);
    TNX_CHECK_HIP(hipFuncSetAttribute((const void *)trace_entropy_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(trace_entropy_kernel, dim3((unsigned)n_jobs, 4), dim3(nt), shm, s, bins,
                       n_vars, ld, tau, (int32_t)L, P, mode, var_ids, conn, conn_off, xlogx_fx,
                       sums_out);
    TNX_CHECK_LAUNCH("trace_entropy_kernel");
    if (te_out != nullptr) {
        const unsigned grid = (unsigned)((n_jobs + 255) / 256);
        hipLaunchKernelGGL(trace_combine_kernel, dim3(grid), dim3(256), 0, s, sums_out, n_jobs,
                           fx_shift, (int32_t)L, te_out);
        TNX_CHECK_LAUNCH("trace_combine_kernel");
    }
    return TNX_OK;
}
