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

// Backward: a 256-thread workgroup walks chunks of kBwdRows = 16 rows.
// Thread t works on row r = t % 16 and on hidden units [4q, 4q + 4),
// q = t / 16, for the forward and delta passes (weights read from LDS as
// 16-byte vectors, activations and deltas exchanged through LDS).  For the
// weight gradients thread t owns a 4 x 4 block: output units i0 = 4 (t % 16)
// .. +3 and inputs k0 = 4 (t / 16) .. +3 of W2 and W1, accumulated in
// registers over every chunk the workgroup visits; the threads of column
// block 0..3 also own b2, b1, W3 and b3.  Short chunks keep a step of a few
// thousand rows spread over every CU.  Each workgroup writes one partial
// gradient row; a second kernel sums the rows in workgroup order, so the
// result does not depend on timing.
constexpr int kBwdThreads = 256;
constexpr int kBwdMaxWG = 256;
constexpr int kBwdRows = 16;                        // rows per chunk
constexpr int kBwdQ = kBwdThreads / kBwdRows;       // unit groups per row (16)
constexpr int kBwdU = kMlpH / kBwdQ;                // units per thread (4)
constexpr int kLd = kMlpH + 4;                      // row stride: 16-byte aligned rows

struct BwdShared {
    float W1T[kMlpMaxD * kMlpH];  // [k][j]
    float W2[kMlpH * kMlpH];      // [j][k]
    float W2T[kMlpH * kMlpH];     // [k][j]
    float W3[kMlpMaxO * kMlpH];
    float b1[kMlpH], b2[kMlpH], b3[kMlpMaxO];
    float x[kBwdRows][kLd];
    float h1[kBwdRows][kLd];
    float h2[kBwdRows][kLd];
    float d1[kBwdRows][kLd];
    float d2[kBwdRows][kLd];
    float yp[kBwdQ][kBwdRows][kMlpMaxO];
    float dy[kBwdRows][kMlpMaxO];
};

struct ParamLayout {
    int W1, b1, W2, b2, W3, b3, P;
    int part_ld;  // floats per workgroup row of partial gradients (P + loss, padded to 4)
    __host__ __device__ ParamLayout(int D, int O) {
        W1 = 0;
        b1 = kMlpH * D;
        W2 = b1 + kMlpH;
        b2 = W2 + kMlpH * kMlpH;
        W3 = b2 + kMlpH;
        b3 = W3 + O * kMlpH;
        P = b3 + O;
        part_ld = (P + 1 + 3) & ~3;
    }
};

__host__ __device__ inline int64_t bwd_groups(int64_t n_rows) {
    const int64_t chunks = (n_rows + kBwdRows - 1) / kBwdRows;
    return chunks < kBwdMaxWG ? chunks : kBwdMaxWG;
}

__device__ __forceinline__ float4 lds4(const float *p) {
    return *reinterpret_cast<const float4 *>(p);
}

__global__ __launch_bounds__(kBwdThreads) void mlp_backward_kernel(
    tnx_mlp_t m, const float *__restrict__ X, const float *__restrict__ tgt,
    const float *__restrict__ w_out, int64_t n_rows, float *__restrict__ part) {
    extern __shared__ float dyn[];
    BwdShared &S = *reinterpret_cast<BwdShared *>(dyn);
    const int D = m.D, O = m.O;
    const int t = threadIdx.x, r = t % kBwdRows, q = t / kBwdRows;
    const int u0 = kBwdU * q;
    const ParamLayout L(D, O);
    {
        // every global load is issued before the first LDS write
        uint16_t w1v[16], w2v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int i = t + kBwdThreads * u;
            w1v[u] = i < kMlpH * D ? m.W1[i] : (uint16_t)0;
            w2v[u] = m.W2[i];
        }
        const uint16_t w3v = t < O * kMlpH ? m.W3[t] : (uint16_t)0;
        const float b1v = t < kMlpH ? m.b1[t] : 0.f;
        const float b2v = t < kMlpH ? m.b2[t] : 0.f;
        const float b3v = t < O ? m.b3[t] : 0.f;
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int i = t + kBwdThreads * u;
            if (i < kMlpH * D) {
                const int j = i / D, k = i - j * D;
                S.W1T[k * kMlpH + j] = bf16_to_f32(w1v[u]);
            }
            const float v = bf16_to_f32(w2v[u]);
            S.W2[i] = v;
            S.W2T[(i & 63) * kMlpH + (i >> 6)] = v;
        }
        if (t < O * kMlpH) S.W3[t] = bf16_to_f32(w3v);
        if (t < kMlpH) {
            S.b1[t] = b1v;
            S.b2[t] = b2v;
        }
        if (t < O) S.b3[t] = b3v;
    }
    __syncthreads();

    const float inv = 1.0f / (float)(n_rows * O);
    const int ib = t & 15, kb = t >> 4;
    const int i0 = 4 * ib, k0 = 4 * kb;
    float gw2[4][4], gw1[4][4], gw3[kMlpMaxO][4], gb[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        gb[u] = 0.f;
#pragma unroll
        for (int v = 0; v < 4; v++) gw2[u][v] = gw1[u][v] = gw3[u][v] = 0.f;
    }
    float loss = 0.f;
    constexpr int kXLoads = (kBwdRows * kMlpMaxD + kBwdThreads - 1) / kBwdThreads;
    const int64_t n_chunks = (n_rows + kBwdRows - 1) / kBwdRows;
    for (int64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        const int64_t base = c * kBwdRows;
        const int rows = (int)(n_rows - base < kBwdRows ? n_rows - base : kBwdRows);
        float xv[kXLoads], tv[kMlpMaxO];
#pragma unroll
        for (int u = 0; u < kXLoads; u++) {
            const int e = t + kBwdThreads * u;
            xv[u] = e < rows * D ? X[base * D + e] : 0.f;
        }
#pragma unroll
        for (int o = 0; o < kMlpMaxO; o++)
            tv[o] = (q == 0 && o < O && r < rows) ? tgt[(base + r) * O + o] : 0.f;
#pragma unroll
        for (int u = 0; u < kXLoads; u++) {
            const int e = t + kBwdThreads * u;
            if (e < kBwdRows * D) {
                const int rr = e / D, k = e - rr * D;
                S.x[rr][k] = xv[u];
            }
        }
        __syncthreads();
        // layer 1: units u0 .. u0+3 of row r
        float4 h1r = lds4(&S.b1[u0]);
        for (int k = 0; k < D; k++) {
            const float xk = S.x[r][k];
            const float4 w = lds4(&S.W1T[k * kMlpH + u0]);
            h1r.x += w.x * xk; h1r.y += w.y * xk; h1r.z += w.z * xk; h1r.w += w.w * xk;
        }
        h1r.x = fmaxf(h1r.x, 0.f); h1r.y = fmaxf(h1r.y, 0.f);
        h1r.z = fmaxf(h1r.z, 0.f); h1r.w = fmaxf(h1r.w, 0.f);
        *reinterpret_cast<float4 *>(&S.h1[r][u0]) = h1r;
        __syncthreads();
        // layer 2 and this group's share of the outputs
        float4 h2r = lds4(&S.b2[u0]);
        for (int k = 0; k < kMlpH; k++) {
            const float hk = S.h1[r][k];
            const float4 w = lds4(&S.W2T[k * kMlpH + u0]);
            h2r.x += w.x * hk; h2r.y += w.y * hk; h2r.z += w.z * hk; h2r.w += w.w * hk;
        }
        h2r.x = fmaxf(h2r.x, 0.f); h2r.y = fmaxf(h2r.y, 0.f);
        h2r.z = fmaxf(h2r.z, 0.f); h2r.w = fmaxf(h2r.w, 0.f);
        *reinterpret_cast<float4 *>(&S.h2[r][u0]) = h2r;
        for (int o = 0; o < O; o++) {
            const float4 w = lds4(&S.W3[o * kMlpH + u0]);
            S.yp[q][r][o] = w.x * h2r.x + w.y * h2r.y + w.z * h2r.z + w.w * h2r.w;
        }
        __syncthreads();
        if (q == 0) {
            const bool valid = r < rows;
#pragma unroll
            for (int o = 0; o < kMlpMaxO; o++) {
                if (o >= O) continue;
                float y = S.b3[o];
#pragma unroll
                for (int qq = 0; qq < kBwdQ; qq++) y += S.yp[qq][r][o];
                const float w = w_out ? w_out[o] : 1.0f;
                const float e = valid ? (y - tv[o]) : 0.f;
                loss += 0.5f * w * e * e * inv;
                S.dy[r][o] = w * e * inv;
            }
        }
        __syncthreads();
        // deltas of layer 2 (units u0 ..)
        {
            float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int o = 0; o < O; o++) {
                const float dv = S.dy[r][o];
                const float4 w = lds4(&S.W3[o * kMlpH + u0]);
                g.x += w.x * dv; g.y += w.y * dv; g.z += w.z * dv; g.w += w.w * dv;
            }
            g.x = h2r.x > 0.f ? g.x : 0.f; g.y = h2r.y > 0.f ? g.y : 0.f;
            g.z = h2r.z > 0.f ? g.z : 0.f; g.w = h2r.w > 0.f ? g.w : 0.f;
            *reinterpret_cast<float4 *>(&S.d2[r][u0]) = g;
        }
        __syncthreads();
        // deltas of layer 1 (units u0 ..): d1[k] = sum_j W2[j][k] d2[j]
        {
            float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int j = 0; j < kMlpH; j++) {
                const float dv = S.d2[r][j];
                const float4 w = lds4(&S.W2[j * kMlpH + u0]);
                d.x += w.x * dv; d.y += w.y * dv; d.z += w.z * dv; d.w += w.w * dv;
            }
            d.x = h1r.x > 0.f ? d.x : 0.f; d.y = h1r.y > 0.f ? d.y : 0.f;
            d.z = h1r.z > 0.f ? d.z : 0.f; d.w = h1r.w > 0.f ? d.w : 0.f;
            *reinterpret_cast<float4 *>(&S.d1[r][u0]) = d;
        }
        __syncthreads();
        // weight gradients of the chunk (padding rows carry zero deltas)
        for (int rr = 0; rr < kBwdRows; rr++) {
            const float4 a2 = lds4(&S.d2[rr][i0]);
            const float4 a1 = lds4(&S.d1[rr][i0]);
            const float4 hh = lds4(&S.h1[rr][k0]);
            const float4 x4 = lds4(&S.x[rr][k0]);
            const float av2[4] = {a2.x, a2.y, a2.z, a2.w}, av1[4] = {a1.x, a1.y, a1.z, a1.w};
            const float hv[4] = {hh.x, hh.y, hh.z, hh.w};
            const float xx[4] = {k0 + 0 < D ? x4.x : 0.f, k0 + 1 < D ? x4.y : 0.f,
                                 k0 + 2 < D ? x4.z : 0.f, k0 + 3 < D ? x4.w : 0.f};
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    gw2[u][v] += av2[u] * hv[v];
                    gw1[u][v] += av1[u] * xx[v];
                }
        }
        if (kb == 0) {
            for (int rr = 0; rr < kBwdRows; rr++) {
                const float4 a = lds4(&S.d2[rr][i0]);
                gb[0] += a.x; gb[1] += a.y; gb[2] += a.z; gb[3] += a.w;
            }
        } else if (kb == 1) {
            for (int rr = 0; rr < kBwdRows; rr++) {
                const float4 a = lds4(&S.d1[rr][i0]);
                gb[0] += a.x; gb[1] += a.y; gb[2] += a.z; gb[3] += a.w;
            }
        } else if (kb == 2) {
            for (int rr = 0; rr < kBwdRows; rr++) {
                const float4 h = lds4(&S.h2[rr][i0]);
#pragma unroll
                for (int o = 0; o < kMlpMaxO; o++) {
                    if (o >= O) continue;
                    const float dv = S.dy[rr][o];
                    gw3[o][0] += dv * h.x; gw3[o][1] += dv * h.y;
                    gw3[o][2] += dv * h.z; gw3[o][3] += dv * h.w;
                }
            }
        } else if (kb == 3 && ib == 0) {
            for (int rr = 0; rr < kBwdRows; rr++)
#pragma unroll
                for (int o = 0; o < kMlpMaxO; o++)
                    if (o < O) gb[o] += S.dy[rr][o];
        }
        __syncthreads();  // the next chunk overwrites x, h1, h2, d1, d2, dy
    }
    float *pr = part + (int64_t)blockIdx.x * L.part_ld;
    // W2 rows as 16-byte stores (the row stride and W2's offset are multiples
    // of four floats when D is; W1 rows are D long)
    const bool w2_vec = ((L.W2 | L.part_ld) & 3) == 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int i = i0 + u;
        if (w2_vec) {
            *reinterpret_cast<float4 *>(pr + L.W2 + i * kMlpH + k0) =
                make_float4(gw2[u][0], gw2[u][1], gw2[u][2], gw2[u][3]);
        } else {
#pragma unroll
            for (int v = 0; v < 4; v++) pr[L.W2 + i * kMlpH + k0 + v] = gw2[u][v];
        }
#pragma unroll
        for (int v = 0; v < 4; v++)
            if (k0 + v < D) pr[L.W1 + i * D + k0 + v] = gw1[u][v];
    }
    if (kb == 0) {
        for (int u = 0; u < 4; u++) pr[L.b2 + i0 + u] = gb[u];
    } else if (kb == 1) {
        for (int u = 0; u < 4; u++) pr[L.b1 + i0 + u] = gb[u];
    } else if (kb == 2) {
        for (int o = 0; o < O; o++)
            for (int u = 0; u < 4; u++) pr[L.W3 + o * kMlpH + i0 + u] = gw3[o][u];
    } else if (kb == 3 && ib == 0) {
        for (int o = 0; o < O; o++) pr[L.b3 + o] = gb[o];
    }
    if (t < kWave) {  // the loss lives in threads q == 0 (t < kBwdRows) of wave 0
        for (int off = 32; off >= 1; off >>= 1) loss += __shfl_xor(loss, off, 64);
        if (t == 0) pr[L.P] = loss;
    }
}

__device__ __forceinline__ void adam_update(int i, float g, float *__restrict__ master,
                                            float *__restrict__ m1, float *__restrict__ m2,
                                            float lr, float b1, float b2, float eps, float c1,
                                            float c2, const tnx_mlp_t &m) {
    const float a = b1 * m1[i] + (1.f - b1) * g;
    const float v = b2 * m2[i] + (1.f - b2) * g * g;
    m1[i] = a;
    m2[i] = v;
    const float w = master[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    master[i] = w;
    // refresh the bf16 / bias copies (params layout W1, b1, W2, b2, W3, b3)
    const ParamLayout L(m.D, m.O);
    if (i < L.b1) const_cast<uint16_t *>(m.W1)[i] = f32_to_bf16(w);
    else if (i < L.W2) const_cast<float *>(m.b1)[i - L.b1] = w;
    else if (i < L.b2) const_cast<uint16_t *>(m.W2)[i - L.W2] = f32_to_bf16(w);
    else if (i < L.W3) const_cast<float *>(m.b2)[i - L.b2] = w;
    else if (i < L.b3) const_cast<uint16_t *>(m.W3)[i - L.W3] = f32_to_bf16(w);
    else const_cast<float *>(m.b3)[i - L.b3] = w;
}

// Sum of the per-workgroup partial rows: a 256-thread block covers 32
// parameters, eight threads per parameter each summing every eighth row,
// then the eight sums are added in a fixed order.  Entry P is the loss.
// With do_adam the Adam step follows for each parameter.
constexpr int kRedParams = 32;
constexpr int kRedSlices = 8;

__global__ __launch_bounds__(256) void mlp_reduce_adam_kernel(
    const float *__restrict__ part, int G, float *__restrict__ grad, float *__restrict__ loss,
    int do_adam, tnx_mlp_t m, float *__restrict__ master, float *__restrict__ m1,
    float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
    __shared__ float red[kRedSlices][kRedParams];
    const ParamLayout L(m.D, m.O);
    const int pl = threadIdx.x % kRedParams, gs = threadIdx.x / kRedParams;
    const int p = blockIdx.x * kRedParams + pl;
    float acc = 0.f;
    if (p <= L.P) {
#pragma unroll 4
        for (int g = gs; g < G; g += kRedSlices) acc += part[(int64_t)g * L.part_ld + p];
    }
    red[gs][pl] = acc;
    __syncthreads();
    if (gs != 0 || p > L.P) return;
    float sum = red[0][pl];
#pragma unroll
    for (int k = 1; k < kRedSlices; k++) sum += red[k][pl];
    if (p == L.P) {
        if (loss) *loss = sum;
        return;
    }
    grad[p] = sum;
    if (do_adam) adam_update(p, sum, master, m1, m2, lr, b1, b2, eps, c1, c2, m);
}

__global__ void mlp_adam_kernel(const float *__restrict__ grad, tnx_mlp_t m,
                                float *__restrict__ master, float *__restrict__ m1,
                                float *__restrict__ m2, int n, float lr, float b1, float b2,
                                float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) adam_update(i, grad[i], master, m1, m2, lr, b1, b2, eps, c1, c2, m);
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

int64_t tnx_mlp_workspace_bytes(int32_t D, int32_t O, int64_t n_rows) {
    if (D < 1 || D > kMlpMaxD || O < 1 || O > kMlpMaxO || n_rows < 0) return -1;
    const ParamLayout L(D, O);
    return bwd_groups(n_rows) * (int64_t)L.part_ld * (int64_t)sizeof(float);
}

static int launch_backward(const tnx_mlp_t *m, const float *X, const float *tgt,
                           const float *w_out, int64_t n_rows, float *part, hipStream_t st) {
    const size_t shm = sizeof(BwdShared);
    static_assert(sizeof(BwdShared) <= 160 * 1024, "backward LDS exceeds 160 KiB");
    TNX_CHECK_HIP(hipFuncSetAttribute((const void *)mlp_backward_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(mlp_backward_kernel, dim3((unsigned)bwd_groups(n_rows)),
                       dim3(kBwdThreads), shm, st, *m, X, tgt, w_out, n_rows, part);
    TNX_CHECK_LAUNCH("mlp_backward_kernel");
    return TNX_OK;
}

static int check_train_args(const tnx_mlp_t *m, int64_t n_rows, void *workspace,
                            int64_t ws_bytes, const char *who) {
    int rc = check_mlp(m);
    if (rc) return rc;
    TNX_CHECK_ARG(n_rows >= 1 && n_rows <= ((int64_t)1 << 31), "%s: n_rows must be in [1, 2^31]",
                  who);
    const int64_t need = tnx_mlp_workspace_bytes(m->D, m->O, n_rows);
    TNX_CHECK_ARG(workspace && ws_bytes >= need, "%s: workspace of %lld bytes needed (got %lld)",
                  who, (long long)need, (long long)ws_bytes);
    return TNX_OK;
}

int tnx_mlp_backward(const tnx_mlp_t *m, const float *X, const float *tgt, const float *w_out,
                     int64_t n_rows, float *grad, float *loss_out, void *workspace,
                     int64_t ws_bytes, void *stream) {
    clear_error();
    TNX_CHECK_ARG(X && tgt && grad, "tnx_mlp_backward: null pointer");
    int rc = check_train_args(m, n_rows, workspace, ws_bytes, "tnx_mlp_backward");
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    float *part = static_cast<float *>(workspace);
    rc = launch_backward(m, X, tgt, w_out, n_rows, part, st);
    if (rc) return rc;
    const ParamLayout L(m->D, m->O);
    hipLaunchKernelGGL(mlp_reduce_adam_kernel, dim3((unsigned)((L.P + kRedParams) / kRedParams)),
                       dim3(256), 0, st, part, (int)bwd_groups(n_rows), grad, loss_out, 0, *m,
                       nullptr, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f);
    TNX_CHECK_LAUNCH("mlp_reduce_adam_kernel");
    return TNX_OK;
}

int tnx_mlp_adam(const tnx_mlp_t *m, float *master, const float *grad, float *m1, float *m2,
                 int64_t n_params, float lr, float beta1, float beta2, float eps, int32_t t,
                 void *stream) {
    clear_error();
    int rc = check_mlp(m);
    if (rc) return rc;
    const ParamLayout L(m->D, m->O);
    TNX_CHECK_ARG(master && grad && m1 && m2 && n_params == L.P && t >= 1,
                  "tnx_mlp_adam: bad arguments (n_params %lld, expected %d)",
                  (long long)n_params, L.P);
    const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)((L.P + 255) / 256)), dim3(256), 0,
                       as_stream(stream), grad, *m, master, m1, m2, L.P, lr, beta1, beta2, eps,
                       c1, c2);
    TNX_CHECK_LAUNCH("mlp_adam_kernel");
    return TNX_OK;
}

int tnx_mlp_fit(const tnx_mlp_t *m, float *master, float *m1, float *m2, int64_t n_params,
                const float *X, const float *tgt, int64_t n_rows, const tnx_adam_t *opt,
                float *grad, void *workspace, int64_t ws_bytes, float *loss_hist, void *stream) {
    clear_error();
    TNX_CHECK_ARG(master && m1 && m2 && X && tgt && opt && grad, "tnx_mlp_fit: null pointer");
    TNX_CHECK_ARG(opt->batch >= 0, "tnx_mlp_fit: batch %d < 0", opt->batch);
    const int64_t rows = (opt->batch > 0 && opt->batch < n_rows) ? opt->batch : n_rows;
    const int64_t n_batches = rows > 0 ? n_rows / rows : 1;
    int rc = check_train_args(m, rows, workspace, ws_bytes, "tnx_mlp_fit");
    if (rc) return rc;
    const ParamLayout L(m->D, m->O);
    TNX_CHECK_ARG(n_params == L.P, "tnx_mlp_fit: n_params %lld, expected %d",
                  (long long)n_params, L.P);
    TNX_CHECK_ARG(opt->steps >= 0 && opt->t0 >= 0 && opt->lr > 0.f && opt->lr_min >= 0.f,
                  "tnx_mlp_fit: bad schedule (steps %d, t0 %d)", opt->steps, opt->t0);
    hipStream_t st = as_stream(stream);
    float *part = static_cast<float *>(workspace);
    const int G = (int)bwd_groups(rows);
    const double pi = 3.14159265358979323846;
    for (int s = 0; s < opt->steps; s++) {
        const float lr = (float)(opt->lr_min + (opt->lr - opt->lr_min) * 0.5 *
                                                   (1.0 + cos(pi * (double)s / opt->steps)));
        const int64_t r0 = (s % n_batches) * rows;
        rc = launch_backward(m, X + r0 * m->D, tgt + r0 * m->O, nullptr, rows, part, st);
        if (rc) return rc;
        const int t = opt->t0 + s + 1;
        const float c1 = 1.f - powf(opt->beta1, (float)t), c2 = 1.f - powf(opt->beta2, (float)t);
        hipLaunchKernelGGL(mlp_reduce_adam_kernel,
                           dim3((unsigned)((L.P + kRedParams) / kRedParams)), dim3(256), 0, st,
                           part, G, grad, loss_hist ? loss_hist + s : nullptr, 1, *m, master, m1,
                           m2, lr, opt->beta1, opt->beta2, opt->eps, c1, c2);
        TNX_CHECK_LAUNCH("mlp_reduce_adam_kernel");
    }
    return TNX_OK;
}

}  // extern "C"
