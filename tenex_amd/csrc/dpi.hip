// dpi.hip -- data-processing-inequality trimming as a max-min product (gfx950).
//
// The FDR network's DPI trim (tenex/inference/fdr.py:61-133) keeps edge
// i -> j of the thresholded TE network A when
//     A[i,j] >= max(0, max_k min(A[i,k], A[k,j])) + threshold,
// the mediators k being the variables with outgoing weight.  The inner
// max_k min(.,.) is a GEMM over the (max, min) semiring: every product is an
// exact comparison, so any order of evaluation gives the same bits.
//
// tnx_dpi_maxmin evaluates it for a block of rows against a mediator set,
// on operands the caller has gathered:
//     A_IK (m x k):  A[rows, mediators]
//     B_K  (k x n):  A[mediators, :]
//     E    (m x n):  A[rows, :]            (the edges to keep or drop)
//     out[i,j] = E[i,j] if E[i,j] >= max(0, max_k min(A_IK[i,k], B_K[k,j])) + thr
//                else 0.
// 128 x 128 output tiles per 256-thread workgroup (8 x 8 per thread), K in
// steps of 32 through LDS with the next step's operands loaded into
// registers while the current one is reduced: per k each thread does 64
// min and 64 max (the compiler fuses pairs of max into v_max3) on four
// 16-byte LDS reads, so the kernel is VALU bound.
// Zero padding past m, k, n is neutral (min(x, 0) <= 0 never raises the
// max above its initial 0).
#include "common.hpp"

namespace tnx {

constexpr int kDpiTile = 128;
constexpr int kDpiK = 32;

__device__ __forceinline__ float4 ld4_or0(const float *__restrict__ p, int64_t row, int64_t rows,
                                          int64_t col, int64_t cols, int64_t ld) {
    if (row >= rows) return make_float4(0.f, 0.f, 0.f, 0.f);
    const float *q = p + row * ld + col;
    if (col + 3 < cols && ((reinterpret_cast<uintptr_t>(q) & 15) == 0))
        return *reinterpret_cast<const float4 *>(q);
    float4 v;
    v.x = col + 0 < cols ? q[0] : 0.f;
    v.y = col + 1 < cols ? q[1] : 0.f;
    v.z = col + 2 < cols ? q[2] : 0.f;
    v.w = col + 3 < cols ? q[3] : 0.f;
    return v;
}

// 128 x 128 output tile per workgroup; thread (ty, tx) of a 16 x 16 grid owns
// rows {4 ty .. 4 ty + 3, 64 + 4 ty ..} and the same split of columns, so its
// per-k operands are four 16-byte LDS reads (conflict-free: a wave reads 4
// distinct A vectors and 16 consecutive B vectors) for 64 min and 64 max.
__global__ __launch_bounds__(256) void dpi_maxmin_kernel(
    const float *__restrict__ A, int64_t m, int64_t kdim, int64_t lda,
    const float *__restrict__ B, int64_t n, int64_t ldb, const float *__restrict__ E,
    int64_t lde, float thr, float *__restrict__ out, int64_t ldo) {
    __shared__ float As[kDpiK][kDpiTile];      // [k][i]
    __shared__ float Bs[kDpiK][kDpiTile];      // [k][j]
    const int t = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.y * kDpiTile, j0 = (int64_t)blockIdx.x * kDpiTile;
    const int ty = t >> 4, tx = t & 15;
    // loaders: A tile 128 rows x 32 k (1024 float4: rows ar + 32 q), B tile
    // 32 k x 128 cols (1024 float4: k rows br + 8 q)
    const int ar = t >> 3, ac = (t & 7) * 4;
    const int br = t >> 5, bc = (t & 31) * 4;
    float acc[8][8];
#pragma unroll
    for (int u = 0; u < 8; u++)
#pragma unroll
        for (int v = 0; v < 8; v++) acc[u][v] = 0.f;
    float4 av[4], bv[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        av[q] = ld4_or0(A, i0 + ar + 32 * q, m, ac, kdim, lda);
        bv[q] = ld4_or0(B, br + 8 * q, kdim, j0 + bc, n, ldb);
    }
    for (int64_t k0 = 0; k0 < kdim; k0 += kDpiK) {
        __syncthreads();  // the previous step's reads are done
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = ar + 32 * q;
            As[ac + 0][r] = av[q].x; As[ac + 1][r] = av[q].y;
            As[ac + 2][r] = av[q].z; As[ac + 3][r] = av[q].w;
            *reinterpret_cast<float4 *>(&Bs[br + 8 * q][bc]) = bv[q];
        }
        __syncthreads();
        const int64_t kn = k0 + kDpiK;
        if (kn < kdim) {  // next step's operands in flight during this step
#pragma unroll
            for (int q = 0; q < 4; q++) {
                av[q] = ld4_or0(A, i0 + ar + 32 * q, m, kn + ac, kdim, lda);
                bv[q] = ld4_or0(B, kn + br + 8 * q, kdim, j0 + bc, n, ldb);
            }
        }
#pragma unroll 4
        for (int k = 0; k < kDpiK; k++) {
            const float4 a0 = *reinterpret_cast<const float4 *>(&As[k][4 * ty]);
            const float4 a1 = *reinterpret_cast<const float4 *>(&As[k][64 + 4 * ty]);
            const float4 b0 = *reinterpret_cast<const float4 *>(&Bs[k][4 * tx]);
            const float4 b1 = *reinterpret_cast<const float4 *>(&Bs[k][64 + 4 * tx]);
            const float a[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
            const float b[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
            for (int u = 0; u < 8; u++)
#pragma unroll
                for (int v = 0; v < 8; v++) acc[u][v] = fmaxf(acc[u][v], fminf(a[u], b[v]));
        }
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int64_t i = i0 + (u < 4 ? 4 * ty + u : 64 + 4 * ty + u - 4);
        if (i >= m) continue;
#pragma unroll
        for (int v = 0; v < 8; v++) {
            const int64_t j = j0 + (v < 4 ? 4 * tx + v : 64 + 4 * tx + v - 4);
            if (j >= n) continue;
            const float e = E[i * lde + j];
            out[i * ldo + j] = e >= acc[u][v] + thr ? e : 0.f;
        }
    }
}

}  // namespace tnx

using namespace tnx;

extern "C" {

int tnx_dpi_maxmin(const float *A_ik, int64_t m, int64_t k, int64_t lda, const float *B_kn,
                   int64_t n, int64_t ldb, const float *E, int64_t lde, float trim_threshold,
                   float *out, int64_t ldo, void *stream) {
    clear_error();
    TNX_CHECK_ARG(m >= 0 && k >= 0 && n >= 0, "tnx_dpi_maxmin: negative size");
    if (m == 0 || n == 0) return TNX_OK;
    TNX_CHECK_ARG(E && out && (k == 0 || (A_ik && B_kn)), "tnx_dpi_maxmin: null pointer");
    TNX_CHECK_ARG((k == 0 || (lda >= k && ldb >= n)) && lde >= n && ldo >= n,
                  "tnx_dpi_maxmin: leading dimensions too small");
    TNX_CHECK_ARG((m + kDpiTile - 1) / kDpiTile < 65536, "tnx_dpi_maxmin: m too large");
    const dim3 grid((unsigned)((n + kDpiTile - 1) / kDpiTile),
                    (unsigned)((m + kDpiTile - 1) / kDpiTile));
    hipLaunchKernelGGL(dpi_maxmin_kernel, grid, dim3(256), 0, as_stream(stream), A_ik, m, k,
                       lda, B_kn, n, ldb, E, lde, trim_threshold, out, ldo);
    TNX_CHECK_LAUNCH("dpi_maxmin_kernel");
    return TNX_OK;
}

}  // extern "C"
