// prep.hip -- discretisation and per-variable preparation kernels (gfx950).
//
//  tnx_fsbw_l       FSBW-L with numpy-exact float32 statistics
//                   (tenex/preprocess.py:61-88 bin_fsbw_l_numpy, :622 clamp)
//  tnx_remap_dense  per-variable dense relabelling into int8
//                   (tenex/preprocess.py:401-459 remap_bins, :335-354)
//  tnx_coarsen_bins lossy coarsening (tenex/preprocess.py:462-551,
//                   tenex/csrc/adaptive_smem.cu:557-592)
//  tnx_xlogx_table  c*log2(c) hi/lo table used by every entropy reduction
//  tnx_gene_prep    target states, S1, S2a and learned-null features
#include <math.h>
#include <stdarg.h>
#include <string.h>

#include <algorithm>

#include "common.hpp"

namespace tnx {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
void clear_error() { g_err[0] = 0; }

// ── numpy float32 pairwise summation, bit-exact ───────────────────────────
// numpy/_core/src/umath/loops_utils.h.src (pairwise_sum_FLOAT): blocks of
// <= 128 with 8 accumulators, recursive halving (rounded to a multiple of
// 8) above that; np.add.reduce feeds a contiguous row in buffers of 8192.
// Evaluated here without recursion (explicit stack, depth <= 7 per buffer).
#pragma clang fp contract(off)

// Correctly rounded float32 square root and quotient (numpy's np.sqrt and
// true_divide on float32).  HIP's __fsqrt_rn is the native v_sqrt_f32 (not
// correctly rounded) unless OCML_BASIC_ROUNDED_OPERATIONS is defined, so
// both are computed in FP64 and rounded once: the FP64 result is correctly
// rounded, and rounding it again to float32 gives the correctly rounded
// float32 result because 53 >= 2 * 24 + 2 (double rounding is innocuous
// for sqrt and division at these precisions).
__device__ __forceinline__ float sqrt_rn(float v) { return (float)sqrt((double)v); }
__device__ __forceinline__ float div_rn(float a, float b) { return (float)((double)a / (double)b); }

template <typename F>
__device__ float pw_leaf(const F &f, int64_t off, int64_t n) {
    if (n < 8) {
        float res = 0.0f;
        for (int64_t i = 0; i < n; i++) res = res + f(off + i);
        return res;
    }
    float r0 = f(off + 0), r1 = f(off + 1), r2 = f(off + 2), r3 = f(off + 3);
    float r4 = f(off + 4), r5 = f(off + 5), r6 = f(off + 6), r7 = f(off + 7);
    int64_t i;
    for (i = 8; i < n - (n % 8); i += 8) {
        r0 = r0 + f(off + i + 0);
        r1 = r1 + f(off + i + 1);
        r2 = r2 + f(off + i + 2);
        r3 = r3 + f(off + i + 3);
        r4 = r4 + f(off + i + 4);
        r5 = r5 + f(off + i + 5);
        r6 = r6 + f(off + i + 6);
        r7 = r7 + f(off + i + 7);
    }
    float res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
    for (; i < n; i++) res = res + f(off + i);
    return res;
}

template <typename F>
__device__ float pw_sum(const F &f, int64_t off0, int64_t n0) {
    // Post-order traversal of the halving tree.
    int64_t so[16], sn[16];
    int stage[16];
    float left[16];
    int d = 0;
    so[0] = off0;
    sn[0] = n0;
    stage[0] = 0;
    float ret = 0.0f;
    while (d >= 0) {
        int64_t n = sn[d];
        if (n <= 128) {
            ret = pw_leaf(f, so[d], n);
            d--;
        } else {
            int64_t n2 = n / 2;
            n2 -= n2 % 8;
            if (stage[d] == 0) {
                stage[d] = 1;
                so[d + 1] = so[d];
                sn[d + 1] = n2;
                stage[d + 1] = 0;
                d++;
                continue;
            } else if (stage[d] == 1) {
                left[d] = ret;
                stage[d] = 2;
                so[d + 1] = so[d] + n2;
                sn[d + 1] = n - n2;
                stage[d + 1] = 0;
                d++;
                continue;
            } else {
                ret = left[d] + ret;
                d--;
            }
        }
    }
    return ret;
}

template <typename F>
__device__ float np_rowsum(const F &f, int64_t n) {
    float res = 0.0f;
    for (int64_t s = 0; s < n; s += 8192) {
        int64_t m = (n - s < 8192) ? (n - s) : 8192;
        float p = pw_sum(f, s, m);
        res = (s == 0) ? p : (res + p);
    }
    return res;
}

__global__ void fsbw_stats_kernel(const float *__restrict__ x, int64_t n_vars, int64_t T,
                                  float kp, float *__restrict__ sd_out,
                                  float *__restrict__ smin_out, int32_t *__restrict__ nb_out) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_vars) return;
    const float *row = x + g * T;
    auto ident = [row](int64_t i) { return row[i]; };
    float s = np_rowsum(ident, T);
    float mean = (float)((double)s / (double)T);
    auto sq = [row, mean](int64_t i) {
        float d = row[i] - mean;
        return d * d;
    };
    float ss = np_rowsum(sq, T);
    float var = (float)((double)ss / (double)(T - 1));
    float sd = sqrt_rn(var);
    if (sd == 0.0f || isnan(sd)) sd = 1.0f;
    float mn = row[0], mx = row[0];
    for (int64_t t = 1; t < T; t++) {
        float v = row[t];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    float smin = mn - kp * sd;
    sd_out[g] = sd;
    smin_out[g] = smin;
    if (nb_out) nb_out[g] = (int32_t)ceilf(div_rn(mx - mn, sd));
}

__global__ void fsbw_bins_kernel(const float *__restrict__ x, int64_t n_vars, int64_t T,
                                 const float *__restrict__ sd, const float *__restrict__ smin,
                                 int32_t *__restrict__ raw) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_vars * T) return;
    int64_t g = i / T;
    float q = div_rn(x[i] - smin[g], sd[g]);
    int32_t b = (int32_t)floorf(q);
    raw[i] = b < 0 ? 0 : b;
}

#pragma clang fp contract(on)

// ── dense remap: bitmap of used raw levels -> prefix rank ──────────────────
constexpr int kRemapBits = 65536;
constexpr int kRemapWords = kRemapBits / 32;

// OUT = int8_t: levels above 127 set status bit 2 (values clamped);
// OUT = int16_t: levels above 32767 set status bit 4.
template <typename OUT>
__global__ __launch_bounds__(256) void remap_kernel(const int32_t *__restrict__ raw,
                                                    int64_t n_vars, int64_t T,
                                                    OUT *__restrict__ out, int64_t ld,
                                                    int32_t *__restrict__ npv,
                                                    int32_t *__restrict__ status) {
    constexpr int32_t kMaxLevel = sizeof(OUT) == 1 ? 127 : 32767;
    __shared__ uint32_t bm[kRemapWords];
    __shared__ uint32_t base[kRemapWords];
    __shared__ uint32_t part[256];
    const int64_t g = blockIdx.x;
    const int tid = threadIdx.x;
    const int32_t *r = raw + g * T;
    for (int i = tid; i < kRemapWords; i += 256) bm[i] = 0;
    __syncthreads();
    int bad = 0;
    for (int64_t t = tid; t < T; t += 256) {
        int32_t v = r[t];
        if (v < 0 || v >= kRemapBits) {
            bad = 1;
            continue;
        }
        atomicOr(&bm[v >> 5], 1u << (v & 31));
    }
    if (bad) atomicOr(status, 1);
    __syncthreads();
    // exclusive prefix of popcounts: 8 words per thread
    constexpr int per = kRemapWords / 256;
    uint32_t loc = 0;
    for (int k = 0; k < per; k++) loc += __popc(bm[tid * per + k]);
    part[tid] = loc;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = (tid >= off) ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = part[tid] - loc;
    for (int k = 0; k < per; k++) {
        base[tid * per + k] = run;
        run += __popc(bm[tid * per + k]);
    }
    __syncthreads();
    const uint32_t total = part[255];
    if (tid == 0) {
        npv[g] = (int32_t)total;
        if (total > (uint32_t)kMaxLevel) atomicOr(status, sizeof(OUT) == 1 ? 2 : 4);
    }
    OUT *o = out + g * ld;
    for (int64_t t = tid; t < ld; t += 256) {
        int32_t d = 0;
        if (t < T) {
            int32_t v = r[t];
            if (v >= 0 && v < kRemapBits) {
                uint32_t w = (uint32_t)v >> 5, b = (uint32_t)v & 31;
                d = (int32_t)(base[w] + __popc(bm[w] & ((1u << b) - 1u)));
            }
        }
        o[t] = (OUT)(d > kMaxLevel ? kMaxLevel : d);
    }
}

__global__ void coarsen_kernel(int8_t *__restrict__ bins, int64_t n_vars, int64_t T, int64_t ld,
                               const int32_t *__restrict__ npv, int32_t limit) {
    const int64_t g = blockIdx.y;
    const int32_t ng = npv[g];
    if (ng <= limit) return;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T;
         t += (int64_t)gridDim.x * blockDim.x) {
        int32_t v = bins[g * ld + t];
        bins[g * ld + t] = (int8_t)((v * limit) / ng);
    }
}

// int16 rows -> int8 rows: new = old * limit / ng where ng > limit, else
// old (every variable then has at most max(limit, 127) levels).
__global__ void coarsen16_kernel(const int16_t *__restrict__ in, int64_t T, int64_t ld,
                                 const int32_t *__restrict__ npv, int32_t limit,
                                 int8_t *__restrict__ out) {
    const int64_t g = blockIdx.y;
    const int32_t ng = npv[g];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ld;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = t < T ? (int32_t)in[g * ld + t] : 0;
        out[g * ld + t] = (int8_t)(ng > limit ? (v * limit) / ng : v);
    }
}

__global__ void coarsen_npv_kernel(int32_t *__restrict__ npv, int64_t n_vars, int32_t limit) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n_vars && npv[g] > limit) npv[g] = limit;
}

// ── c log2 c table ─────────────────────────────────────────────────────────
__global__ void xlogx_kernel(float *__restrict__ tbl, int64_t N) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > N) return;
    double v = c > 1 ? (double)c * log2((double)c) : 0.0;
    float hi = (float)v;
    float lo = (float)(v - (double)hi);
    tbl[2 * c] = hi;
    tbl[2 * c + 1] = lo;
}

// c*log2(c) as 64-bit fixed point with `shift` fraction bits (rounded to
// nearest from the same FP64 value the hi/lo table splits).
__global__ void xlogx_fx_kernel(int64_t *__restrict__ tbl, int64_t N, int shift) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > N) return;
    const double v = c > 1 ? (double)c * log2((double)c) : 0.0;
    tbl[c] = __double2ll_rn(ldexp(v, shift));
}

// ── per-variable preparation ───────────────────────────────────────────────
__device__ __forceinline__ double xlog2x_d(double c) { return c > 1.0 ? c * log2(c) : 0.0; }

__device__ double block_sum_d(double v, double *red) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    if (tid == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
    __syncthreads();
    return s;
}

__global__ __launch_bounds__(256) void gene_prep_kernel(
    const int8_t *__restrict__ bins, int64_t n_vars, int64_t T, int64_t ld, int32_t tau,
    const int32_t *__restrict__ npv, uint16_t *__restrict__ states, int64_t lds,
    double *__restrict__ S1, double *__restrict__ S2a, float *__restrict__ feat, int32_t bcap) {
    extern __shared__ uint32_t sh[];
    __shared__ double red[8];
    const int64_t g = blockIdx.x;
    const int tid = threadIdx.x;
    const int b = npv[g];
    const int64_t L = T - tau;
    if (b > bcap || b < 1) {  // level count above the caller's bound
        if (tid == 0) S1[g] = S2a[g] = __builtin_nan("");
        return;
    }
    uint32_t *c1 = sh;          // bcap
    uint32_t *c2 = sh + bcap;   // bcap * bcap
    for (int i = tid; i < b * b; i += 256) c2[i] = 0;
    __syncthreads();
    const int8_t *x = bins + g * ld;
    uint16_t *st = states + g * lds;
    // Only the joint (x[t+tau], x[t]) histogram is counted; its column sums
    // give the x[t] marginal and its diagonal the lag-tau agreement.  State
    // 0 (both at level 0, the bulk of zero-inflated data) is counted in a
    // register: one LDS address taking every lane's increment would
    // serialise the workgroup's atomics.
    // 16 consecutive samples per thread and step: the 16 x[t] bytes in one
    // 16-byte load, the 16 x[t+tau] bytes requested together before any is
    // used, the 16 states written as two 16-byte stores (rows are padded to
    // 64 elements, samples past L get state 0)
    uint32_t z00 = 0;
    for (int64_t t0 = 16 * (int64_t)tid; t0 < lds; t0 += 16 * 256) {
        const uint4 c16 = *reinterpret_cast<const uint4 *>(x + t0);
        const uint32_t cw[4] = {c16.x, c16.y, c16.z, c16.w};
        uint32_t av[16];
#pragma unroll
        for (int q = 0; q < 16; q++)
            av[q] = t0 + q < L ? (uint32_t)(uint8_t)x[t0 + q + tau] : 0u;
        uint32_t sw[8];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            uint32_t key = 0;
            if (t0 + q < L) {
                const uint32_t c = (cw[q >> 2] >> (8 * (q & 3))) & 0xFFu;
                key = av[q] * (uint32_t)b + c;
                if (key) atomicAdd(&c2[key], 1u);
                else z00++;
            }
            if (q & 1) sw[q >> 1] |= key << 16;
            else sw[q >> 1] = key;
        }
        uint4 *so = reinterpret_cast<uint4 *>(st + t0);
        so[0] = make_uint4(sw[0], sw[1], sw[2], sw[3]);
        so[1] = make_uint4(sw[4], sw[5], sw[6], sw[7]);
    }
    if (z00) atomicAdd(&c2[0], z00);
    __syncthreads();
    uint32_t same = 0;
    for (int i = tid; i < b; i += 256) {
        uint32_t col = 0;
        for (int a = 0; a < b; a++) col += c2[a * b + i];
        c1[i] = col;
        same += c2[i * b + i];
    }
    __syncthreads();
    double s1 = 0, s2 = 0, lvl = 0;
    uint32_t mode = 0;
    for (int i = tid; i < b; i += 256) {
        s1 += xlog2x_d((double)c1[i]);
        lvl += (double)i * (double)c1[i];
        mode = c1[i] > mode ? c1[i] : mode;
    }
    for (int i = tid; i < b * b; i += 256) s2 += xlog2x_d((double)c2[i]);
    double S1v = block_sum_d(s1, red);
    double S2v = block_sum_d(s2, red);
    if (!feat) {  // entropies only (shuffled surrogate copies)
        if (tid == 0) {
            S1[g] = S1v;
            S2a[g] = S2v;
        }
        return;
    }
    double lv = block_sum_d(lvl, red);
    double sm = block_sum_d((double)same, red);
    // block max of mode
    for (int off = 32; off >= 1; off >>= 1) {
        uint32_t o = __shfl_xor(mode, off, 64);
        mode = o > mode ? o : mode;
    }
    __shared__ uint32_t mred[4];
    if ((tid & 63) == 0) mred[tid >> 6] = mode;
    __syncthreads();
    if (tid == 0) {
        S1[g] = S1v;
        S2a[g] = S2v;
        {
            uint32_t md = mred[0];
            for (int i = 1; i < 4; i++) md = mred[i] > md ? mred[i] : md;
            double Ld = (double)L, lg = log2(Ld);
            float *f = feat + g * 8;
            f[0] = (float)b;
            f[1] = (float)(lg - S1v / Ld);          // H(X_t)
            f[2] = (float)((S1v - S2v) / Ld);       // H(X_t+1 | X_t)
            f[3] = (float)((double)c1[0] / Ld);     // fraction at the lowest level
            f[4] = (float)((double)md / Ld);        // fraction at the modal level
            f[5] = (float)log2((double)(b > 0 ? b : 1));
            f[6] = (float)(lv / Ld / (double)(b > 0 ? b : 1));
            f[7] = (float)(sm / Ld);                // lag-tau self agreement
        }
    }
}

}  // namespace tnx

using namespace tnx;

// ── binarization (tenex/binarize.py:51-200, kernels/binarize_kernel.py) ──
// (no contraction into FMA: the thresholds round as numpy's separate ops)
#pragma clang fp contract(off)
// Per-variable threshold thr[g], then bins = 1[x > thr[g]]:
//   0 nonzero  thr = 0
//   1 mean     numpy's float32 mean (pairwise sum over the row, / T)
//   2 median   numpy's median: the middle order statistic, or the float32
//              mean (a + b) / 2 of the two middle ones
//   3 quantile numpy's linear quantile in float32: v = (T - 1) q,
//              g = v - floor(v), a = x_(floor v), b = x_(floor v + 1),
//              g >= 0.5 ? b - (b - a)(1 - g) : a + (b - a) g
// Order statistics by an exact radix select on order-preserving keys, one
// workgroup per variable, four 8-bit passes over the row.
__device__ __forceinline__ uint32_t order_key(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_value(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

constexpr int kBinThreads = 256;

// Key of the k-th smallest element (0-based) of row[0..T).
__device__ uint32_t select_kth(const float *__restrict__ row, int64_t T, int64_t k,
                               uint32_t *hist, int64_t *s_k, int *s_bin) {
    const int tid = threadIdx.x;
    uint32_t prefix = 0, mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = tid; i < 256; i += kBinThreads) hist[i] = 0;
        __syncthreads();
        for (int64_t t = tid; t < T; t += kBinThreads) {
            const uint32_t key = order_key(row[t]);
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            int64_t c = 0;
            int b = 0;
            for (; b < 255; b++) {
                if (c + (int64_t)hist[b] > k) break;
                c += hist[b];
            }
            *s_bin = b;
            *s_k = k - c;
        }
        __syncthreads();
        prefix |= (uint32_t)(*s_bin) << shift;
        mask |= 255u << shift;
        k = *s_k;
        __syncthreads();
    }
    return prefix;
}

// Key of the (k+1)-th smallest element given the k-th smallest key kk.
__device__ uint32_t select_next(const float *__restrict__ row, int64_t T, int64_t k, uint32_t kk,
                                uint32_t *s_min, unsigned long long *s_le) {
    const int tid = threadIdx.x;
    if (tid == 0) {
        *s_min = 0xffffffffu;
        *s_le = 0ull;
    }
    __syncthreads();
    uint32_t mn = 0xffffffffu;
    unsigned long long le = 0;
    for (int64_t t = tid; t < T; t += kBinThreads) {
        const uint32_t key = order_key(row[t]);
        if (key <= kk) le++;
        else mn = key < mn ? key : mn;
    }
    atomicMin(s_min, mn);
    atomicAdd(s_le, le);
    __syncthreads();
    const uint32_t r = (*s_le > (unsigned long long)(k + 1)) ? kk : *s_min;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kBinThreads) void binarize_thr_kernel(
    const float *__restrict__ x, int64_t T, int64_t ldx, int32_t method, float q,
    float *__restrict__ thr) {
    __shared__ uint32_t hist[256];
    __shared__ int64_t s_k;
    __shared__ int s_bin;
    __shared__ uint32_t s_min;
    __shared__ unsigned long long s_le;
    const int64_t g = blockIdx.x;
    const float *row = x + g * ldx;
    float th = 0.0f;
    if (method == 1) {
        if (threadIdx.x == 0) {
            auto ident = [row](int64_t i) { return row[i]; };
            const float sum = np_rowsum(ident, T);
            thr[g] = (float)((double)sum / (double)T);
        }
        return;
    }
    if (method == 2) {
        const int64_t k = (T - 1) / 2;
        const uint32_t ka = select_kth(row, T, k, hist, &s_k, &s_bin);
        if (T % 2 == 1) {
            th = key_value(ka);
        } else {
            const uint32_t kb = select_next(row, T, k, ka, &s_min, &s_le);
            th = (key_value(ka) + key_value(kb)) / 2.0f;
        }
    } else if (method == 3) {
        const float v = (float)(T - 1) * q;
        int64_t lo = (int64_t)floorf(v);
        lo = lo < 0 ? 0 : (lo > T - 1 ? T - 1 : lo);
        const float gm = v - (float)lo;
        const uint32_t ka = select_kth(row, T, lo, hist, &s_k, &s_bin);
        const uint32_t kb = lo + 1 < T ? select_next(row, T, lo, ka, &s_min, &s_le) : ka;
        const float a = key_value(ka), b = key_value(kb);
        const float d = b - a;
        th = gm >= 0.5f ? b - d * (1.0f - gm) : a + d * gm;
    }
    if (threadIdx.x == 0) thr[g] = th;
}

// out[g, t] = 1[x[g, t] > thr[g]] for t < T, zero padding up to ldo.
__global__ void binarize_apply_kernel(const float *__restrict__ x, int64_t n, int64_t T,
                                      int64_t ldx, const float *__restrict__ thr,
                                      int8_t *__restrict__ out, int64_t ldo) {
    const int64_t g = blockIdx.y;
    const float th = thr[g];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ldo;
         t += (int64_t)gridDim.x * blockDim.x)
        out[g * ldo + t] = (t < T && x[g * ldx + t] > th) ? (int8_t)1 : (int8_t)0;
}
#pragma clang fp contract(on)

__global__ void fill_f32_kernel(float *__restrict__ p, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

extern "C" {

int tnx_abi_version(void) { return TNX_ABI_VERSION; }

const char *tnx_last_error(void) { return tnx::g_err; }

int tnx_device_props(int device, tnx_device_props_t *out) {
    clear_error();
    TNX_CHECK_ARG(out != nullptr, "tnx_device_props: out is NULL");
    hipDeviceProp_t p;
    TNX_CHECK_HIP(hipGetDeviceProperties(&p, device));
    memset(out, 0, sizeof(*out));
    out->lds_bytes_per_block = (int64_t)p.sharedMemPerBlock;
    out->n_cu = p.multiProcessorCount;
    out->wave_size = p.warpSize;
    out->clock_khz = p.clockRate;
    out->gcn_arch = p.major * 100 + p.minor * 10;
    strncpy(out->name, p.gcnArchName, sizeof(out->name) - 1);
    return TNX_OK;
}

int tnx_fsbw_l(const float *x, int64_t n_vars, int64_t T, float kp, int32_t *raw_out,
               int32_t *n_bins_out, float *std_out, float *smin_out, void *stream) {
    clear_error();
    TNX_CHECK_ARG(x && raw_out, "tnx_fsbw_l: null pointer");
    TNX_CHECK_ARG(n_vars >= 0 && T >= 2, "tnx_fsbw_l: need T >= 2 (got %lld)", (long long)T);
    TNX_CHECK_ARG(std_out && smin_out, "tnx_fsbw_l: std_out and smin_out are required scratch");
    if (n_vars == 0) return TNX_OK;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(fsbw_stats_kernel, dim3((unsigned)((n_vars + 63) / 64)), dim3(64), 0, s,
                       x, n_vars, T, kp, std_out, smin_out, n_bins_out);
    TNX_CHECK_LAUNCH("fsbw_stats_kernel");
    int64_t tot = n_vars * T;
    hipLaunchKernelGGL(fsbw_bins_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, x,
                       n_vars, T, std_out, smin_out, raw_out);
    TNX_CHECK_LAUNCH("fsbw_bins_kernel");
    return TNX_OK;
}

int tnx_remap_dense(const int32_t *raw, int64_t n_vars, int64_t T, int8_t *bins_out,
                    int64_t ld, int32_t *n_per_var_out, int32_t *status, void *stream) {
    clear_error();
    TNX_CHECK_ARG(raw && bins_out && n_per_var_out && status, "tnx_remap_dense: null pointer");
    TNX_CHECK_ARG(ld >= T && T >= 1, "tnx_remap_dense: need ld >= T >= 1");
    if (n_vars == 0) return TNX_OK;
    hipLaunchKernelGGL(remap_kernel<int8_t>, dim3((unsigned)n_vars), dim3(256), 0,
                       as_stream(stream), raw, n_vars, T, bins_out, ld, n_per_var_out, status);
    TNX_CHECK_LAUNCH("remap_kernel");
    return TNX_OK;
}

int tnx_remap_dense16(const int32_t *raw, int64_t n_vars, int64_t T, int16_t *bins_out,
                      int64_t ld, int32_t *n_per_var_out, int32_t *status, void *stream) {
    clear_error();
    TNX_CHECK_ARG(raw && bins_out && n_per_var_out && status, "tnx_remap_dense16: null pointer");
    TNX_CHECK_ARG(ld >= T && T >= 1, "tnx_remap_dense16: need ld >= T >= 1");
    if (n_vars == 0) return TNX_OK;
    hipLaunchKernelGGL(remap_kernel<int16_t>, dim3((unsigned)n_vars), dim3(256), 0,
                       as_stream(stream), raw, n_vars, T, bins_out, ld, n_per_var_out, status);
    TNX_CHECK_LAUNCH("remap_kernel");
    return TNX_OK;
}

int tnx_coarsen_bins16(const int16_t *bins16, int64_t n_vars, int64_t T, int64_t ld,
                       int32_t *n_per_var, int32_t limit, int8_t *bins_out, void *stream) {
    clear_error();
    TNX_CHECK_ARG(bins16 && n_per_var && bins_out, "tnx_coarsen_bins16: null pointer");
    TNX_CHECK_ARG(limit >= 2 && limit <= 127, "tnx_coarsen_bins16: limit must be in [2,127]");
    TNX_CHECK_ARG(ld >= T && n_vars <= 65535, "tnx_coarsen_bins16: need ld >= T, n <= 65535");
    if (n_vars == 0) return TNX_OK;
    hipStream_t s = as_stream(stream);
    const unsigned gx = (unsigned)std::min<int64_t>((ld + 255) / 256, 64);
    hipLaunchKernelGGL(coarsen16_kernel, dim3(gx, (unsigned)n_vars), dim3(256), 0, s, bins16, T,
                       ld, n_per_var, limit, bins_out);
    TNX_CHECK_LAUNCH("coarsen16_kernel");
    hipLaunchKernelGGL(coarsen_npv_kernel, dim3((unsigned)((n_vars + 255) / 256)), dim3(256), 0, s,
                       n_per_var, n_vars, limit);
    TNX_CHECK_LAUNCH("coarsen_npv_kernel");
    return TNX_OK;
}

int tnx_coarsen_bins(int8_t *bins, int64_t n_vars, int64_t T, int64_t ld, int32_t *n_per_var,
                     int32_t limit, void *stream) {
    clear_error();
    TNX_CHECK_ARG(bins && n_per_var, "tnx_coarsen_bins: null pointer");
    TNX_CHECK_ARG(limit >= 2 && limit <= 127, "tnx_coarsen_bins: limit must be in [2,127]");
    TNX_CHECK_ARG(n_vars <= 65535, "tnx_coarsen_bins: n_vars > 65535");
    if (n_vars == 0) return TNX_OK;
    hipStream_t s = as_stream(stream);
    unsigned gx = (unsigned)((T + 255) / 256);
    if (gx > 64) gx = 64;
    hipLaunchKernelGGL(coarsen_kernel, dim3(gx, (unsigned)n_vars), dim3(256), 0, s, bins, n_vars,
                       T, ld, n_per_var, limit);
    TNX_CHECK_LAUNCH("coarsen_kernel");
    hipLaunchKernelGGL(coarsen_npv_kernel, dim3((unsigned)((n_vars + 255) / 256)), dim3(256), 0,
                       s, n_per_var, n_vars, limit);
    TNX_CHECK_LAUNCH("coarsen_npv_kernel");
    return TNX_OK;
}

int tnx_xlogx_table(float *tbl, int64_t N, void *stream) {
    clear_error();
    TNX_CHECK_ARG(tbl && N >= 0, "tnx_xlogx_table: bad arguments");
    hipLaunchKernelGGL(xlogx_kernel, dim3((unsigned)((N + 256) / 256)), dim3(256), 0,
                       as_stream(stream), tbl, N);
    TNX_CHECK_LAUNCH("xlogx_kernel");
    return TNX_OK;
}

int tnx_xlogx_table_fx(int64_t *tbl, int64_t N, int32_t shift, void *stream) {
    clear_error();
    TNX_CHECK_ARG(tbl && N >= 0 && N < (1LL << 31), "tnx_xlogx_table_fx: bad arguments");
    // the largest sum of entries is at most N log2 N: it must stay below 2^62
    const double top = N > 1 ? (double)N * log2((double)N) : 1.0;
    TNX_CHECK_ARG(shift >= 0 && ldexp(top, shift) < ldexp(1.0, 62),
                  "tnx_xlogx_table_fx: shift %d overflows 64 bits for N = %lld", shift,
                  (long long)N);
    hipLaunchKernelGGL(xlogx_fx_kernel, dim3((unsigned)((N + 256) / 256)), dim3(256), 0,
                       as_stream(stream), tbl, N, (int)shift);
    TNX_CHECK_LAUNCH("xlogx_fx_kernel");
    return TNX_OK;
}

int tnx_gene_prep(const int8_t *bins, int64_t n_vars, int64_t T, int64_t ld, int32_t tau,
                  const int32_t *n_per_var, uint16_t *states, int64_t lds, double *S1,
                  double *S2a, float *feat, int32_t b_max, void *stream) {
    clear_error();
    TNX_CHECK_ARG(bins && n_per_var && states && S1 && S2a, "tnx_gene_prep: null pointer");
    TNX_CHECK_ARG(tau >= 1 && tau < T, "tnx_gene_prep: need 1 <= tau < T");
    TNX_CHECK_ARG(ld >= T && lds >= T - tau, "tnx_gene_prep: row strides too small");
    TNX_CHECK_ARG(ld % 16 == 0 && lds % 16 == 0 && lds <= ld,
                  "tnx_gene_prep: ld and lds must be multiples of 16 with lds <= ld");
    TNX_CHECK_ARG(b_max >= 0 && b_max <= 127, "tnx_gene_prep: b_max %d not in [0, 127]", b_max);
    if (n_vars == 0) return TNX_OK;
    // the histogram is sized for the largest level count: small b leaves LDS
    // room for many workgroups per CU
    const int bcap = b_max > 0 ? b_max : 127;
    size_t shm = (size_t)(bcap + bcap * bcap) * 4;
    hipStream_t s = as_stream(stream);
    TNX_CHECK_HIP(hipFuncSetAttribute((const void *)gene_prep_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(gene_prep_kernel, dim3((unsigned)n_vars), dim3(256), shm, s, bins, n_vars,
                       T, ld, tau, n_per_var, states, lds, S1, S2a, feat, bcap);
    TNX_CHECK_LAUNCH("gene_prep_kernel");
    return TNX_OK;
}

int tnx_binarize(const float *x, int64_t n_vars, int64_t T, int64_t ldx, int32_t method, float q,
                 int8_t *bins_out, int64_t ld_out, float *thr_out, void *stream) {
    clear_error();
    TNX_CHECK_ARG(x && bins_out && thr_out, "tnx_binarize: null pointer");
    TNX_CHECK_ARG(method >= 0 && method <= 3,
                  "tnx_binarize: method must be 0 (nonzero), 1 (mean), 2 (median) or 3 (quantile)");
    TNX_CHECK_ARG(method != 3 || (q >= 0.0f && q <= 1.0f),
                  "tnx_binarize: quantile q must be in [0, 1]");
    TNX_CHECK_ARG(n_vars >= 0 && T >= 1 && ldx >= T && ld_out >= T,
                  "tnx_binarize: need T >= 1, ldx >= T and ld_out >= T");
    TNX_CHECK_ARG(n_vars <= 65535, "tnx_binarize: at most 65535 variables per call");
    if (n_vars == 0) return TNX_OK;
    hipStream_t s = as_stream(stream);
    if (method == 0) {
        hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)((n_vars + 255) / 256)), dim3(256), 0, s,
                           thr_out, n_vars, 0.0f);
        TNX_CHECK_LAUNCH("fill_f32_kernel");
    } else {
        hipLaunchKernelGGL(binarize_thr_kernel, dim3((unsigned)n_vars), dim3(kBinThreads), 0, s, x,
                           T, ldx, method, q, thr_out);
        TNX_CHECK_LAUNCH("binarize_thr_kernel");
    }
    const unsigned gx = (unsigned)std::min<int64_t>((ld_out + 255) / 256, 64);
    hipLaunchKernelGGL(binarize_apply_kernel, dim3(gx, (unsigned)n_vars), dim3(256), 0, s, x,
                       n_vars, T, ldx, thr_out, bins_out, ld_out);
    TNX_CHECK_LAUNCH("binarize_apply_kernel");
    return TNX_OK;
}

}  // extern "C"
