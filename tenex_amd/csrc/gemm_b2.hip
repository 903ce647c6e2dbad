// gemm_b2.hip -- TE for binary data from exact INT8 MFMA contractions (gfx950).
//
// For binary bins (values 0/1), with f_i[t] = x_i[t+tau], p_i[t] = x_i[t],
// t < L = T - tau, the three contractions
//     C[i,j] = sum_t f_i p_i p_j,  D[i,j] = sum_t p_i p_j,  E[i,j] = sum_t f_i p_j
// plus per-gene marginals nf, np, nfp give all eight joint counts of
// (x_i[t+tau], x_i[t], x_j[t]) by inclusion-exclusion
// (tenex/kernels/gemm_b2.py:240-249).  The reference evaluates C, D, E as
// three FP32 torch.mm calls and a Triton epilogue (gemm_b2.py:389-391,
// :66-187); here one kernel contracts the 0/1 operands on the matrix cores
// and applies the epilogue in registers.  Two precisions, both exact:
//   INT8: v_mfma_i32_32x32x32_i8, INT32 accumulation (any L < 2^31);
//   FP4:  0/1 as e2m1 nibbles (1.0 = 0x2), v_mfma_scale_f32_32x32x64_f8f6f4
//         with scales 2^0 and FP32 accumulation (exact below 2^24: L <
//         2^24), half the operand bytes and twice the products per cycle.
// The epilogue:
//     TE(j -> i) = (S3 + S1[i] - S2a[i] - S2b) / L
// with S3 over the 8 joint counts and S2b over the 4 (x_i[t], x_j[t]) counts,
// c*log2(c) from the hi/lo table (as doubles in LDS with fp64 sums when the
// table fits, else compensated fp32 sums) and an fp64 final combination.
//
// Workspace (tnx_gemm_b2_workspace): F and P operand rows (n_pad x Lp int8
// bytes, zero padded; FP4 uses the first half), then nf, np, nfp (int32
// n_pad each).
#include <algorithm>
#include <type_traits>

#include "common.hpp"

namespace tnx {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kB2TileI = 64;    // rows (targets) per workgroup
constexpr int kB2TileJ = 128;   // cols (sources) per workgroup
constexpr int kB2K = 64;        // bytes of t per main-loop step (2 MFMAs)

__host__ __device__ inline int64_t b2_npad(int64_t n) { return (n + 127) / 128 * 128; }
// padded series length: whole 128-sample steps of the FP4 loop (two INT8 steps)
__host__ __device__ inline int64_t b2_lp(int64_t L) { return (L + 127) / 128 * 128; }

// Operand layout: 32-row x 32-byte blocks of 1 KB (row block major, then the
// 32-byte k chunk).  Inside a block, the 16-byte run h = (k % 32) / 16 of row
// r sits at byte (h * 32 + r) * 16: exactly where lane h * 32 + r of a
// v_mfma_i32_32x32x32_i8 operand load reads it, so each wave's operand load
// is one contiguous kilobyte.
__device__ __forceinline__ int64_t b2_tiled(int64_t row, int64_t k, int64_t Lp) {
    const int64_t blk = (row >> 5) * (Lp >> 5) + (k >> 5);
    return blk * 1024 + ((((k & 31) >> 4) * 32 + (row & 31)) << 4) + (k & 15);
}

// Samples t0 .. t0 + 15 of row `row` (t0 % 16 == 0): p = x[t], f =
// x[t + tau] as 0/1 bytes, zero at t >= L.
__device__ __forceinline__ void b2_load16(const int8_t *row, int64_t t0, int32_t tau, int64_t L,
                                          int64_t ld, uint32_t (&pw)[4], uint32_t (&fw)[4]) {
#pragma unroll
    for (int q = 0; q < 4; q++) pw[q] = fw[q] = 0u;
    const int64_t tf = t0 + tau, base = tf & ~(int64_t)15;
    if (t0 + 16 <= L && base + 32 <= ld) {
        // 16-byte loads: P from the aligned run at t0, F from the two
        // aligned runs around t0 + tau, shifted by (t0 + tau) % 16 bytes
        const uint4 pv = *reinterpret_cast<const uint4 *>(row + t0);
        const uint4 a = *reinterpret_cast<const uint4 *>(row + base);
        const uint4 b = *reinterpret_cast<const uint4 *>(row + base + 16);
        const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        const int sh = (int)(tf & 15), ws = sh >> 2, bs = 8 * (sh & 3);
        pw[0] = pv.x; pw[1] = pv.y; pw[2] = pv.z; pw[3] = pv.w;
        // word selection by compares (static register indices)
        auto pick = [&](int j) {
            return ws == 0 ? w[j] : ws == 1 ? w[j + 1] : ws == 2 ? w[j + 2] : w[j + 3];
        };
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t lo = pick(q), hi = pick(q + 1);
            fw[q] = bs ? ((lo >> bs) | (hi << (32 - bs))) : lo;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int64_t t = t0 + q;
            if (t < L) {
                fw[q >> 2] |= (uint32_t)(uint8_t)row[t + tau] << (8 * (q & 3));
                pw[q >> 2] |= (uint32_t)(uint8_t)row[t] << (8 * (q & 3));
            }
        }
    }
}

// Four 0/1 bytes to four e2m1 nibbles in the low 16 bits (sample q at
// nibble q; 1.0 is 0x2).
__device__ __forceinline__ uint32_t b2_nibbles(uint32_t w) {
    const uint32_t t = (w | (w >> 4)) & 0x00FF00FFu;
    return ((t & 0xFFu) | ((t >> 8) & 0xFF00u)) * 2u;
}

// One workgroup per variable: operand rows (tiled, 16 bytes per thread and
// step) and marginals.  INT8 layout: 32-row x 32-byte blocks, lane h * 32 +
// r of an operand load holds bytes 16 h .. of row r.  FP4 layout: 32-row x
// 64-sample blocks (1 KB), lane h * 32 + r holds samples 32 h .. 32 h + 31
// of row r as nibbles (sample q at nibble q of its 16 bytes); both
// operands of a product use the same map, so each sample of a row meets
// the same sample of a column.
template <bool FP4>
__global__ __launch_bounds__(256) void b2_pack_kernel(const int8_t *__restrict__ bins, int64_t n,
                                                      int64_t ld, int32_t tau, int64_t L,
                                                      int64_t Lp, int8_t *__restrict__ F,
                                                      int8_t *__restrict__ P,
                                                      int32_t *__restrict__ nf,
                                                      int32_t *__restrict__ np_,
                                                      int32_t *__restrict__ nfp) {
    __shared__ int32_t red[3][4];
    // Workgroup w runs on XCD w % 8; XCD x packs the x-th contiguous slice of
    // rows (the grid is a multiple of 128), so the eight rows whose 16-byte
    // runs share one 128-byte line of the tiled layout are written through
    // one L2
    const int64_t i = (int64_t)(blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
    const int tid = threadIdx.x;
    int32_t sf = 0, sp = 0, sfp = 0;
    // sum of the four 0/1 bytes of a word
    auto bsum = [](uint32_t w) { return (int32_t)((w * 0x01010101u) >> 24); };
    const int8_t *row = bins + i * ld;
    constexpr int S = FP4 ? 32 : 16;   // samples per thread and step
    for (int64_t t0 = S * (int64_t)tid; t0 < Lp; t0 += S * 256) {
        uint32_t fo[4], po[4];
#pragma unroll
        for (int g = 0; g < S / 16; g++) {
            uint32_t fw[4], pw[4];
            if (i < n) {
                b2_load16(row, t0 + 16 * g, tau, L, ld, pw, fw);
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) fw[q] = pw[q] = 0u;
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                sf += bsum(fw[q]);
                sp += bsum(pw[q]);
                sfp += bsum(fw[q] & pw[q]);
            }
            if constexpr (FP4) {
                fo[2 * g] = b2_nibbles(fw[0]) | (b2_nibbles(fw[1]) << 16);
                fo[2 * g + 1] = b2_nibbles(fw[2]) | (b2_nibbles(fw[3]) << 16);
                po[2 * g] = b2_nibbles(pw[0]) | (b2_nibbles(pw[1]) << 16);
                po[2 * g + 1] = b2_nibbles(pw[2]) | (b2_nibbles(pw[3]) << 16);
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    fo[q] = fw[q];
                    po[q] = pw[q];
                }
            }
        }
        int64_t o;
        if constexpr (FP4) {
            o = ((i >> 5) * (Lp >> 6) + (t0 >> 6)) * 1024 + ((((t0 & 63) >> 5) * 32 + (i & 31)) << 4);
        } else {
            o = b2_tiled(i, t0, Lp);
        }
        *reinterpret_cast<uint4 *>(F + o) = make_uint4(fo[0], fo[1], fo[2], fo[3]);
        *reinterpret_cast<uint4 *>(P + o) = make_uint4(po[0], po[1], po[2], po[3]);
    }
    for (int off = 32; off >= 1; off >>= 1) {
        sf += __shfl_xor(sf, off, 64);
        sp += __shfl_xor(sp, off, 64);
        sfp += __shfl_xor(sfp, off, 64);
    }
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = sf;
        red[1][tid >> 6] = sp;
        red[2][tid >> 6] = sfp;
    }
    __syncthreads();
    if (tid == 0) {
        nf[i] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        np_[i] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        nfp[i] = red[2][0] + red[2][1] + red[2][2] + red[2][3];
    }
}

// Entries of the c*log2(c) table the epilogue keeps in LDS (8 bytes each);
// longer series read it from memory.
constexpr int kB2LdsTbl = 8192;

// Workgroup: 4 waves as 2 (i) x 2 (j); wave tile 32 (i) x 64 (j).  LT: the
// table (N + 1 <= kB2LdsTbl entries) is copied to LDS while the MFMA loop
// runs, so the epilogue's twelve lookups per pair are LDS reads.
template <bool LT, bool FP4>
__global__ __launch_bounds__(256) void b2_mfma_kernel(
    const int8_t *__restrict__ F, const int8_t *__restrict__ P, const int32_t *__restrict__ nf,
    const int32_t *__restrict__ np_, const int32_t *__restrict__ nfp,
    const double *__restrict__ S1, const double *__restrict__ S2a,
    const float *__restrict__ xlogx, int64_t n, int64_t L, int64_t Lp,
    float *__restrict__ te_out, int64_t ld_out, int32_t *__restrict__ counts_dbg) {
    extern __shared__ unsigned long long b2_ltbl[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if constexpr (LT) {  // hi + lo of each entry as one double
        const float2 *gt = reinterpret_cast<const float2 *>(xlogx);
        double *lt = reinterpret_cast<double *>(b2_ltbl);
        for (int e = threadIdx.x; e <= (int)L; e += 256) {
            const float2 v = gt[e];
            lt[e] = (double)v.x + (double)v.y;
        }
    }
    const int64_t i0 = (int64_t)blockIdx.y * kB2TileI + (wave >> 1) * 32;
    const int64_t j0 = (int64_t)blockIdx.x * kB2TileJ + (wave & 1) * 64;
    // this lane's 16 operand bytes of each 32-row block (tiled layout: the
    // wave reads one contiguous kilobyte per block and k chunk; padded rows
    // are zero, so no masking)
    constexpr int KCH = FP4 ? 64 : 32;  // samples per 1 KB operand block
    constexpr int STEP = 2 * KCH;        // samples per main-loop step
    const int64_t kblk = Lp / KCH;       // k chunks per row block
    const int8_t *fa = F + (i0 >> 5) * kblk * 1024 + 16 * lane;
    const int8_t *pa = P + (i0 >> 5) * kblk * 1024 + 16 * lane;
    const int8_t *pb0 = P + (j0 >> 5) * kblk * 1024 + 16 * lane;
    const int8_t *pb1 = P + ((j0 + 32) >> 5) * kblk * 1024 + 16 * lane;
    using Acc = typename std::conditional<FP4, v16f, v16i>::type;
    Acc accC0 = {}, accC1 = {}, accD0 = {}, accD1 = {}, accE0 = {}, accE1 = {};
    // Two-stage software pipeline: the operands of step k + STEP are
    // requested before the MFMAs of step k are issued, so the memory latency
    // of one step hides behind the matrix work of the previous one.
    struct Ops {
        v4i f[2], p[2], q0[2], q1[2];
    };
    auto load = [&](Ops &o, int64_t k) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int64_t kk = (k / KCH + u) * 1024;  // this k chunk's block
            o.f[u] = *reinterpret_cast<const v4i *>(fa + kk);
            o.p[u] = *reinterpret_cast<const v4i *>(pa + kk);
            o.q0[u] = *reinterpret_cast<const v4i *>(pb0 + kk);
            o.q1[u] = *reinterpret_cast<const v4i *>(pb1 + kk);
        }
    };
    auto mma = [&](const Ops &o) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const v4i fp = o.f[u] & o.p[u];  // f p: a bitwise AND of 0/1 bytes or nibbles
            if constexpr (FP4) {
                constexpr int sc = 0x7F7F7F7F;  // E8M0 scale 2^0
                auto w = [](const v4i &x) { return v8i{x[0], x[1], x[2], x[3], 0, 0, 0, 0}; };
                accC0 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(w(fp), w(o.q0[u]), accC0,
                                                                       4, 4, 0, sc, 0, sc);
                accC1 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(w(fp), w(o.q1[u]), accC1,
                                                                       4, 4, 0, sc, 0, sc);
                accD0 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(w(o.p[u]), w(o.q0[u]),
                                                                       accD0, 4, 4, 0, sc, 0, sc);
                accD1 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(w(o.p[u]), w(o.q1[u]),
                                                                       accD1, 4, 4, 0, sc, 0, sc);
                accE0 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(w(o.f[u]), w(o.q0[u]),
                                                                       accE0, 4, 4, 0, sc, 0, sc);
                accE1 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(w(o.f[u]), w(o.q1[u]),
                                                                       accE1, 4, 4, 0, sc, 0, sc);
            } else {
                accC0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(fp, o.q0[u], accC0, 0, 0, 0);
                accC1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(fp, o.q1[u], accC1, 0, 0, 0);
                accD0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(o.p[u], o.q0[u], accD0, 0, 0, 0);
                accD1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(o.p[u], o.q1[u], accD1, 0, 0, 0);
                accE0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(o.f[u], o.q0[u], accE0, 0, 0, 0);
                accE1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(o.f[u], o.q1[u], accE1, 0, 0, 0);
            }
        }
    };
    // ping-pong, no register copies; a fetch past the last step re-reads it
    // (no branch around a load: the compiler waits only for the oldest)
    Ops a, b;
    const int64_t lastk = Lp - STEP;
    load(a, 0);
    int64_t k = 0;
    for (; k + 2 * STEP <= Lp; k += 2 * STEP) {
        load(b, k + STEP);
        mma(a);
        load(a, k + 2 * STEP < lastk ? k + 2 * STEP : lastk);
        mma(b);
    }
    if (k < Lp) mma(a);  // Lp / STEP odd: the last step was loaded into a
    if constexpr (LT) __syncthreads();  // the LDS table is complete
    // Epilogue.  32x32 accumulator layout: lane holds column (lane & 31),
    // register q holds row 8*(q/4) + 4*(lane/32) + (q%4).
    const float2 *tbl = reinterpret_cast<const float2 *>(xlogx);
    const int32_t N = (int32_t)L;
    const double invN = 1.0 / (double)L;
    const int64_t jc[2] = {j0 + (lane & 31), j0 + 32 + (lane & 31)};
    const int32_t npj[2] = {jc[0] < n ? np_[jc[0]] : 0, jc[1] < n ? np_[jc[1]] : 0};
    // S3 - S2b of one pair from its three counts and the marginals: the 8
    // joint cells and 4 (x_t, y_t) cells by inclusion-exclusion, c*log2(c)
    // each (counts outside [0, N] come only from non-binary input and are
    // ignored).  With the LDS table the terms add in FP64; the table in
    // memory keeps its hi/lo pairs and double-float sums.
    auto s3_minus_s2 = [&](int32_t Cc, int32_t Dd, int32_t Ee, int32_t nfi, int32_t npi,
                           int32_t nfpi, int32_t npjv) -> double {
        const int32_t c3[8] = {Cc, Dd - Cc, Ee - Cc, npjv - Dd - Ee + Cc, nfpi - Cc,
                               npi - Dd - nfpi + Cc, nfi - Ee - nfpi + Cc,
                               N - nfi - npi - npjv + nfpi + Dd + Ee - Cc};
        const int32_t c2[4] = {Dd, npjv - Dd, npi - Dd, N - npi - npjv + Dd};
        if constexpr (LT) {
            const double *lt = reinterpret_cast<const double *>(b2_ltbl);
            double a3 = 0.0, a2 = 0.0;
#pragma unroll
            for (int e = 0; e < 8; e++)
                if (c3[e] > 1 && c3[e] <= N) a3 += lt[c3[e]];
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (c2[e] > 1 && c2[e] <= N) a2 += lt[c2[e]];
            return a3 - a2;
        } else {
            DF a3 = {0.f, 0.f}, a2 = {0.f, 0.f};
#pragma unroll
            for (int e = 0; e < 8; e++)
                if (c3[e] > 1 && c3[e] <= N) df_add(a3, tbl[c3[e]].x, tbl[c3[e]].y);
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (c2[e] > 1 && c2[e] <= N) df_add(a2, tbl[c2[e]].x, tbl[c2[e]].y);
            return df_to_double(a3) - df_to_double(a2);
        }
    };
    // 32x32 accumulator layout: lane holds column (lane & 31), register q
    // holds row 8*(q/4) + 4*(lane/32) + (q%4); each row's marginals are
    // read once for both column halves
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int64_t i = i0 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
        if (i >= n) continue;
        const int32_t nfi = nf[i], npi = np_[i], nfpi = nfp[i];
        const double s12 = S1[i] - S2a[i];
#pragma unroll
        for (int half = 0; half < 2; half++) {
            const int64_t j = jc[half];
            if (j >= n) continue;
            const int32_t Cc = (int32_t)(half ? accC1[q] : accC0[q]);
            const int32_t Dd = (int32_t)(half ? accD1[q] : accD0[q]);
            const int32_t Ee = (int32_t)(half ? accE1[q] : accE0[q]);
            if (counts_dbg) {
                counts_dbg[i * n + j] = Cc;
                counts_dbg[n * n + i * n + j] = Dd;
                counts_dbg[2 * n * n + i * n + j] = Ee;
            }
            if (i == j) {
                te_out[i * ld_out + j] = 0.0f;
                continue;
            }
            const double d = s3_minus_s2(Cc, Dd, Ee, nfi, npi, nfpi, npj[half]);
            te_out[i * ld_out + j] = (float)((d + s12) * invN);
        }
    }
}

}  // namespace tnx

using namespace tnx;

extern "C" {

int64_t tnx_gemm_b2_workspace(int64_t n, int64_t T, int32_t tau) {
    if (n <= 0 || tau < 1 || tau >= T) return 0;
    const int64_t np = b2_npad(n), Lp = b2_lp(T - tau);
    return 2 * np * Lp + 3 * np * 4 + 256;
}

int tnx_te_gemm_b2(const int8_t *bins, int64_t n, int64_t T, int64_t ld, int32_t tau,
                   const double *S1, const double *S2a, const float *xlogx, void *workspace,
                   int64_t workspace_bytes, float *te_out, int64_t ld_out, int32_t *counts_dbg,
                   int32_t precision, void *stream) {
    clear_error();
    TNX_CHECK_ARG(bins && S1 && S2a && xlogx && workspace && te_out, "tnx_te_gemm_b2: null pointer");
    TNX_CHECK_ARG(tau >= 1 && tau < T && ld >= T && n >= 1 && ld_out >= n,
                  "tnx_te_gemm_b2: bad shape (n=%lld T=%lld ld=%lld tau=%d)", (long long)n,
                  (long long)T, (long long)ld, tau);
    TNX_CHECK_ARG(T - tau < (1LL << 31) - 128, "tnx_te_gemm_b2: L too large");
    TNX_CHECK_ARG(precision >= TNX_B2_AUTO && precision <= TNX_B2_FP4,
                  "tnx_te_gemm_b2: bad precision %d", precision);
    const int64_t need = tnx_gemm_b2_workspace(n, T, tau);
    TNX_CHECK_ARG(workspace_bytes >= need, "tnx_te_gemm_b2: workspace %lld < %lld bytes",
                  (long long)workspace_bytes, (long long)need);
    const int64_t np = b2_npad(n), L = T - tau, Lp = b2_lp(L);
    // FP32 sums of 0/1 products are exact below 2^24
    const bool fp4 = precision == TNX_B2_FP4 || (precision == TNX_B2_AUTO && L < (1 << 24));
    TNX_CHECK_ARG(!fp4 || L < (1 << 24), "tnx_te_gemm_b2: FP4 needs L < 2^24 (L = %lld)",
                  (long long)L);
    int8_t *F = reinterpret_cast<int8_t *>(workspace);
    int8_t *P = F + np * Lp;
    int32_t *nf = reinterpret_cast<int32_t *>(P + np * Lp);
    int32_t *npp = nf + np;
    int32_t *nfp = npp + np;
    hipStream_t s = as_stream(stream);
    if (fp4)
        hipLaunchKernelGGL(b2_pack_kernel<true>, dim3((unsigned)np), dim3(256), 0, s, bins, n, ld,
                           tau, L, Lp, F, P, nf, npp, nfp);
    else
        hipLaunchKernelGGL(b2_pack_kernel<false>, dim3((unsigned)np), dim3(256), 0, s, bins, n, ld,
                           tau, L, Lp, F, P, nf, npp, nfp);
    TNX_CHECK_LAUNCH("b2_pack_kernel");
    dim3 grid((unsigned)((n + kB2TileJ - 1) / kB2TileJ), (unsigned)((n + kB2TileI - 1) / kB2TileI));
    const bool lt = L + 1 <= kB2LdsTbl;
    const size_t shm = lt ? (size_t)(L + 1) * 8 : 0;
#define TNX_B2_LAUNCH(LT, FP)                                                                     \
    hipLaunchKernelGGL((b2_mfma_kernel<LT, FP>), grid, dim3(256), shm, s, F, P, nf, npp, nfp, S1, \
                       S2a, xlogx, n, L, Lp, te_out, ld_out, counts_dbg)
    if (lt && fp4) TNX_B2_LAUNCH(true, true);
    else if (lt) TNX_B2_LAUNCH(true, false);
    else if (fp4) TNX_B2_LAUNCH(false, true);
    else TNX_B2_LAUNCH(false, false);
#undef TNX_B2_LAUNCH
    TNX_CHECK_LAUNCH("b2_mfma_kernel");
    return TNX_OK;
}

}  // extern "C"
