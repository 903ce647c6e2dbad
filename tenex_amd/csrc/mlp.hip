// mlp.hip -- the learned null: a 3-layer MLP with BF16 weights and FP32
// accumulation (no reference counterpart).
//
//   h1 = relu(W1 x + b1)   W1 [H x D]
//   h2 = relu(W2 h1 + b2)  W2 [H x H]
//   y  = W3 h2 + b3        W3 [O x H]
//
// Weights are stored as bf16 (round-to-nearest-even from FP32 masters kept by
// the optimiser) and expanded to FP32 in LDS; every product and sum is FP32.
// Forward: one row per thread.  Backward: one wavefront per 64 rows; the
// per-row activations and deltas are kept in LDS and each lane then reduces
// whole rows of the weight gradients over the wave's 64 rows before one
// atomic add per gradient entry.
#include <math.h>

#include "common.hpp"

namespace tnx {

constexpr int kMlpH = 64;
constexpr int kMlpMaxD = 64;
constexpr int kMlpMaxO = 4;

__device__ __forceinline__ float bf16_to_f32(uint16_t v) {
    return __uint_as_float((uint32_t)v << 16);
}

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

struct MlpShared {
    float W1[kMlpH * kMlpMaxD];
    float W2[kMlpH * kMlpH];
    float W3[kMlpMaxO * kMlpH];
    float b1[kMlpH], b2[kMlpH], b3[kMlpMaxO];
};

__device__ void load_weights(MlpShared &s, const tnx_mlp_t &m) {
    for (int i = threadIdx.x; i < kMlpH * m.D; i += blockDim.x) s.W1[i] = bf16_to_f32(m.W1[i]);
    for (int i = threadIdx.x; i < kMlpH * kMlpH; i += blockDim.x) s.W2[i] = bf16_to_f32(m.W2[i]);
    for (int i = threadIdx.x; i < m.O * kMlpH; i += blockDim.x) s.W3[i] = bf16_to_f32(m.W3[i]);
    for (int i = threadIdx.x; i < kMlpH; i += blockDim.x) {
        s.b1[i] = m.b1[i];
        s.b2[i] = m.b2[i];
    }
    for (int i = threadIdx.x; i < m.O; i += blockDim.x) s.b3[i] = m.b3[i];
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(tnx_mlp_t m, const float *__restrict__ X,
                                                          int64_t n_rows, float *__restrict__ Y) {
    __shared__ MlpShared s;
    __shared__ float xs[256][kMlpMaxD + 1];
    load_weights(s, m);
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int D = m.D, O = m.O;
    const bool valid = r < n_rows;
    for (int k = 0; k < D; k++) xs[threadIdx.x][k] = valid ? X[r * D + k] : 0.f;
    __syncthreads();
    if (!valid) return;
    float h1[kMlpH];
#pragma unroll
    for (int i = 0; i < kMlpH; i++) {
        float z = s.b1[i];
        for (int k = 0; k < D; k++) z += s.W1[i * D + k] * xs[threadIdx.x][k];
        h1[i] = z > 0.f ? z : 0.f;
    }
    float y[kMlpMaxO];
#pragma unroll
    for (int o = 0; o < kMlpMaxO; o++) y[o] = o < O ? s.b3[o] : 0.f;
    for (int i = 0; i < kMlpH; i++) {
        float z = s.b2[i];
#pragma unroll
        for (int k = 0; k < kMlpH; k++) z += s.W2[i * kMlpH + k] * h1[k];
        const float h2 = z > 0.f ? z : 0.f;
#pragma unroll
        for (int o = 0; o < kMlpMaxO; o++)
            if (o < O) y[o] += s.W3[o * kMlpH + i] * h2;
    }
#pragma unroll
    for (int o = 0; o < kMlpMaxO; o++)
        if (o < O) Y[r * O + o] = y[o];
}

// One wavefront = 64 rows.  LDS per wave: x, h1, h2, d1, d2, dy per row.
struct MlpRows {
    float x[64][kMlpMaxD + 1];
    float h1[64][kMlpH + 1];
    float h2[64][kMlpH + 1];
    float d1[64][kMlpH + 1];
    float d2[64][kMlpH + 1];
    float dy[64][kMlpMaxO];
};

__global__ __launch_bounds__(64) void mlp_backward_kernel(tnx_mlp_t m, const float *__restrict__ X,
                                                          const float *__restrict__ tgt,
                                                          const float *__restrict__ w_out,
                                                          int64_t n_rows, float *__restrict__ grad,
                                                          float *__restrict__ loss_out) {
    extern __shared__ float dyn[];
    MlpShared &s = *reinterpret_cast<MlpShared *>(dyn);
    MlpRows &R = *reinterpret_cast<MlpRows *>(dyn + sizeof(MlpShared) / 4);
    load_weights(s, m);
    __syncthreads();
    const int lane = threadIdx.x;
    const int D = m.D, O = m.O;
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    const int64_t r = r0 + lane;
    const bool valid = r < n_rows;
    const float inv = 1.0f / (float)(n_rows * O);
    float loss = 0.f;
    // forward, keep activations; dy
    for (int k = 0; k < D; k++) R.x[lane][k] = valid ? X[r * D + k] : 0.f;
    for (int i = 0; i < kMlpH; i++) {
        float z = s.b1[i];
        for (int k = 0; k < D; k++) z += s.W1[i * D + k] * R.x[lane][k];
        R.h1[lane][i] = z > 0.f ? z : 0.f;
    }
    float y[kMlpMaxO];
#pragma unroll
    for (int o = 0; o < kMlpMaxO; o++) y[o] = o < O ? s.b3[o] : 0.f;
    for (int i = 0; i < kMlpH; i++) {
        float z = s.b2[i];
        for (int k = 0; k < kMlpH; k++) z += s.W2[i * kMlpH + k] * R.h1[lane][k];
        const float h2 = z > 0.f ? z : 0.f;
        R.h2[lane][i] = h2;
#pragma unroll
        for (int o = 0; o < kMlpMaxO; o++)
            if (o < O) y[o] += s.W3[o * kMlpH + i] * h2;
    }
#pragma unroll
    for (int o = 0; o < kMlpMaxO; o++) {
        if (o >= O) continue;
        const float w = w_out ? w_out[o] : 1.0f;
        const float e = valid ? (y[o] - tgt[r * O + o]) : 0.f;
        loss += 0.5f * w * e * e * inv;
        R.dy[lane][o] = w * e * inv;
    }
    // backward deltas
    for (int k = 0; k < kMlpH; k++) {
        float g = 0.f;
        for (int o = 0; o < O; o++) g += s.W3[o * kMlpH + k] * R.dy[lane][o];
        R.d2[lane][k] = R.h2[lane][k] > 0.f ? g : 0.f;
    }
    for (int k = 0; k < kMlpH; k++) {
        float g = 0.f;
        for (int i = 0; i < kMlpH; i++) g += s.W2[i * kMlpH + k] * R.d2[lane][i];
        R.d1[lane][k] = R.h1[lane][k] > 0.f ? g : 0.f;
    }
    __syncthreads();
    // weight gradients over the wave's 64 rows: lane i owns row i of each
    // weight matrix (and bias entry i)
    float *gW1 = grad, *gb1 = gW1 + kMlpH * D, *gW2 = gb1 + kMlpH, *gb2 = gW2 + kMlpH * kMlpH,
          *gW3 = gb2 + kMlpH, *gb3 = gW3 + O * kMlpH;
    const int i = lane;
    {
        float gb = 0.f;
        for (int rr = 0; rr < 64; rr++) gb += R.d2[rr][i];
        atomicAdd(gb2 + i, gb);
        for (int k = 0; k < kMlpH; k++) {
            float g = 0.f;
            for (int rr = 0; rr < 64; rr++) g += R.d2[rr][i] * R.h1[rr][k];
            atomicAdd(gW2 + i * kMlpH + k, g);
        }
        float gb1v = 0.f;
        for (int rr = 0; rr < 64; rr++) gb1v += R.d1[rr][i];
        atomicAdd(gb1 + i, gb1v);
        for (int k = 0; k < D; k++) {
            float g = 0.f;
            for (int rr = 0; rr < 64; rr++) g += R.d1[rr][i] * R.x[rr][k];
            atomicAdd(gW1 + i * D + k, g);
        }
        for (int o = 0; o < O; o++) {
            float g = 0.f;
            for (int rr = 0; rr < 64; rr++) g += R.dy[rr][o] * R.h2[rr][i];
            atomicAdd(gW3 + o * kMlpH + i, g);
        }
        if (i < O) {
            float g = 0.f;
            for (int rr = 0; rr < 64; rr++) g += R.dy[rr][i];
            atomicAdd(gb3 + i, g);
        }
    }
    for (int off = 32; off >= 1; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (lane == 0 && loss_out) atomicAdd(loss_out, loss);
}

__global__ void mlp_adam_kernel(float *__restrict__ master, const float *__restrict__ grad,
                                float *__restrict__ m1, float *__restrict__ m2, int64_t n,
                                float lr, float b1, float b2, float eps, float c1, float c2,
                                uint16_t *__restrict__ W1, uint16_t *__restrict__ W2,
                                uint16_t *__restrict__ W3, float *__restrict__ bb1,
                                float *__restrict__ bb2, float *__restrict__ bb3, int D, int O) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.f - b1) * g;
    const float v = b2 * m2[i] + (1.f - b2) * g * g;
    m1[i] = a;
    m2[i] = v;
    const float w = master[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    master[i] = w;
    // refresh the bf16 / bias copies (params layout W1, b1, W2, b2, W3, b3)
    const int64_t n1 = (int64_t)kMlpH * D, n2 = n1 + kMlpH, n3 = n2 + kMlpH * kMlpH,
                  n4 = n3 + kMlpH, n5 = n4 + (int64_t)O * kMlpH;
    if (i < n1) W1[i] = f32_to_bf16(w);
    else if (i < n2) bb1[i - n1] = w;
    else if (i < n3) W2[i - n2] = f32_to_bf16(w);
    else if (i < n4) bb2[i - n3] = w;
    else if (i < n5) W3[i - n4] = f32_to_bf16(w);
    else bb3[i - n5] = w;
}

static int check_mlp(const tnx_mlp_t *m) {
    if (!m || !m->W1 || !m->W2 || !m->W3 || !m->b1 || !m->b2 || !m->b3) {
        set_error("tnx_mlp: null weight pointer");
        return TNX_E_ARG;
    }
    if (m->H != kMlpH || m->D < 1 || m->D > kMlpMaxD || m->O < 1 || m->O > kMlpMaxO) {
        set_error("tnx_mlp: need H == 64, 1 <= D <= 64, 1 <= O <= 4 (got D=%d H=%d O=%d)", m->D,
                  m->H, m->O);
        return TNX_E_ARG;
    }
    return TNX_OK;
}

}  // namespace tnx

using namespace tnx;

extern "C" {

int tnx_mlp_forward(const tnx_mlp_t *m, const float *X, int64_t n_rows, float *Y, void *stream) {
    clear_error();
    int rc = check_mlp(m);
    if (rc) return rc;
    TNX_CHECK_ARG(X && Y && n_rows >= 0, "tnx_mlp_forward: bad arguments");
    if (n_rows == 0) return TNX_OK;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n_rows + 255) / 256)), dim3(256), 0,
                       as_stream(stream), *m, X, n_rows, Y);
    TNX_CHECK_LAUNCH("mlp_forward_kernel");
    return TNX_OK;
}

int tnx_mlp_backward(const tnx_mlp_t *m, const float *X, const float *tgt, const float *w_out,
                     int64_t n_rows, float *grad, float *loss_out, void *stream) {
    clear_error();
    int rc = check_mlp(m);
    if (rc) return rc;
    TNX_CHECK_ARG(X && tgt && grad && n_rows >= 0, "tnx_mlp_backward: bad arguments");
    if (n_rows == 0) return TNX_OK;
    const size_t shm = sizeof(MlpShared) + sizeof(MlpRows);
    TNX_CHECK_HIP(hipFuncSetAttribute((const void *)mlp_backward_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(mlp_backward_kernel, dim3((unsigned)((n_rows + 63) / 64)), dim3(64), shm,
                       as_stream(stream), *m, X, tgt, w_out, n_rows, grad, loss_out);
    TNX_CHECK_LAUNCH("mlp_backward_kernel");
    return TNX_OK;
}

int tnx_mlp_adam(const tnx_mlp_t *m, float *master, const float *grad, float *m1, float *m2,
                 int64_t n_params, float lr, float beta1, float beta2, float eps, int32_t t,
                 void *stream) {
    clear_error();
    int rc = check_mlp(m);
    if (rc) return rc;
    const int64_t need = (int64_t)kMlpH * m->D + kMlpH + kMlpH * kMlpH + kMlpH +
                         (int64_t)m->O * kMlpH + m->O;
    TNX_CHECK_ARG(master && grad && m1 && m2 && n_params == need && t >= 1,
                  "tnx_mlp_adam: bad arguments (n_params %lld, expected %lld)",
                  (long long)n_params, (long long)need);
    const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)((n_params + 255) / 256)), dim3(256), 0,
                       as_stream(stream), master, grad, m1, m2, n_params, lr, beta1, beta2, eps,
                       c1, c2, const_cast<uint16_t *>(m->W1), const_cast<uint16_t *>(m->W2),
                       const_cast<uint16_t *>(m->W3), const_cast<float *>(m->b1),
                       const_cast<float *>(m->b2), const_cast<float *>(m->b3), m->D, m->O);
    TNX_CHECK_LAUNCH("mlp_adam_kernel");
    return TNX_OK;
}

}  // extern "C"
