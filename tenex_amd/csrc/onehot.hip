// onehot.hip -- exact joint counts of pairs with few levels on FP4 MFMA.
//
// For target x (b_x levels) and source y (b_y levels) the joint counts of
// (x_{t+tau}, x_t, y_t) over t < L are inner products of 0/1 indicator rows:
//     N[a, b, c] = sum_t [x_{t+tau} = a][x_t = b] [y_t = c]
//     M[b, c]    = sum_t [x_t = b] [y_t = c]                  ((x_t, y_t))
// Target x contributes the rows (a, b) with a < b_x - 1 and the rows x_t = b
// with b < b_x - 1 (b_x^2 - 1 rows), source y its rows y_t = c: all b_y of
// them, or with src_ie the first b_y - 1 (at least one).  The rest follow
// from marginals:  M[b_x - 1, c] = n_y(c) - sum_b M[b, c],
// N[b_x - 1, b, c] = M[b, c] - sum_a N[a, b, c], and with src_ie the last
// source level's cells N[a, b, b_y - 1] = n_x(a, b) - sum_c N[a, b, c] and
// M[b, b_y - 1] = n_x(b) - sum_c M[b, c], all exact integers.
// Indicators are FP4 (e2m1: 1.0 is 0x2), products accumulate in FP32, which
// is exact for counts below 2^24: one v_mfma_scale_f32_32x32x64_f8f6f4
// (scales 2^0) counts 32 x 32 (row, column) cells over 64 samples.  A pair
// of b_x = b_y = 4 costs 15 x 4 cells instead of one LDS atomic per
// sample, so pairs whose (b_x^2 - 1) b_y is small count far faster here
// than in the per-sample histogram kernels (te_persist.hip).
//
// TE = (S3 + S1[x] - S2a[x] - S2b) / L with S3 - S2b the exact 64-bit
// fixed-point sum of c log2 c over the N and M cells (xlogx_fx), the same
// integers and final combination as the histogram kernels, so a pair gives
// the same float TE on either path.
//
// Layout: operand rows are grouped into 128-row groups (targets) and
// 128-column groups (sources) that never split a variable; a workgroup
// (four waves) takes two row groups x two column groups, each wave one
// 128 x 128 sub-tile (4 x 4 MFMA blocks, 256 accumulator registers).
// Operands stream through LDS in stages of NK 64-sample k chunks, fetched
// two stages ahead into registers.  After the K loop each wave stages its
// counts as u16 in LDS and one lane per pair sums the pair's cells.
#include <algorithm>

#include "common.hpp"
#include "te_args.hpp"

namespace tnx {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kOhGroup = 128;                 // rows (columns) of a group
constexpr int kOhNK = 2;                      // 64-sample k chunks (1 KB blocks) per stage
constexpr int kOhKChunk = 64;                 // samples per k chunk
constexpr int kOhFp4One = 0x2;                // 1.0 in e2m1
constexpr int kOhScale = 0x7F7F7F7F;          // E8M0 scale 2^0 in every byte
constexpr int kOhStageBytes = 16 * 1024 * kOhNK;  // 8 NK A + 8 NK B blocks of 1 KB
constexpr int kOhStr = 130;                   // u16 row stride of a staged sub-tile
constexpr int kOhSubBytes = kOhGroup * kOhStr * 2;
constexpr int kOhTbl = 2048;                  // c log2 c entries kept in LDS
constexpr int kOhLdsBytes = 4 * kOhSubBytes + kOhTbl * 8 + 4 * kOhGroup * 4;  // + n_y(c) by column
static_assert(2 * kOhStageBytes <= 4 * kOhSubBytes, "stages alias the staging area");
static_assert(kOhLdsBytes <= 160 * 1024, "LDS");


// 0x01 in each byte of w equal to v (v < 256), 0 elsewhere.
__device__ __forceinline__ uint32_t eq_bytes(uint32_t w, uint32_t v) {
    const uint32_t z = w ^ (v * 0x01010101u);
    const uint32_t t = ~(((z & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | z | 0x7F7F7F7Fu);
    return t >> 7;
}

// 16 bytes of row `row` from byte t0 on (t0 % 16 == 0): aligned loads where
// the row holds them, bytes otherwise; bytes at or past `lim` read zero.
__device__ __forceinline__ void load16(const int8_t *row, int64_t t0, int64_t lim, int64_t ld,
                                       uint32_t (&w)[4]) {
    if (t0 + 16 <= lim) {
        const uint4 v = *reinterpret_cast<const uint4 *>(row + t0);
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
        return;
    }
    w[0] = w[1] = w[2] = w[3] = 0u;
    for (int q = 0; q < 16; q++)
        if (t0 + q < lim) w[q >> 2] |= (uint32_t)(uint8_t)row[t0 + q] << (8 * (q & 3));
    (void)ld;
}

// 16 bytes from byte t0 + tau on (any alignment), zero at or past `lim`.
__device__ __forceinline__ void load16_shift(const int8_t *row, int64_t t0, int32_t tau,
                                             int64_t lim, int64_t ld, uint32_t (&w)[4]) {
    const int64_t tf = t0 + tau, base = tf & ~(int64_t)15;
    if (tf + 16 <= lim && base + 32 <= ld) {
        const uint4 a = *reinterpret_cast<const uint4 *>(row + base);
        const uint4 b = *reinterpret_cast<const uint4 *>(row + base + 16);
        const uint32_t u[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        const int sh = (int)(tf & 15), ws = sh >> 2, bs = 8 * (sh & 3);
        auto pick = [&](int j) {
            return ws == 0 ? u[j] : ws == 1 ? u[j + 1] : ws == 2 ? u[j + 2] : u[j + 3];
        };
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t lo = pick(q), hi = pick(q + 1);
            w[q] = bs ? ((lo >> bs) | (hi << (32 - bs))) : lo;
        }
        return;
    }
    w[0] = w[1] = w[2] = w[3] = 0u;
    for (int q = 0; q < 16; q++)
        if (tf + q < lim) w[q >> 2] |= (uint32_t)(uint8_t)row[tf + q] << (8 * (q & 3));
}

// Marginals of one variable per workgroup: n(a, b) of (x_{t+tau}, x_t) and
// n(c) of x_t over t < L, by LDS atomics.
__global__ __launch_bounds__(256) void oh_marg_kernel(const int8_t *__restrict__ bins, int64_t ld,
                                                      int32_t tau, int64_t L,
                                                      const int32_t *__restrict__ npv,
                                                      const int32_t *__restrict__ mvars,
                                                      const int64_t *__restrict__ moff,
                                                      int32_t *__restrict__ marg) {
    __shared__ uint32_t h[64 * 64 + 64];
    const int v = mvars[blockIdx.x];
    const int b = npv[v];
    if (b > 64) return;  // the planner lists variables of at most 64 levels
    const int nc = b * b + b;
    for (int i = threadIdx.x; i < nc; i += 256) h[i] = 0u;
    __syncthreads();
    const int8_t *row = bins + (int64_t)v * ld;
    for (int64_t t = threadIdx.x; t < L; t += 256) {
        const int xa = (uint8_t)row[t + tau], xb = (uint8_t)row[t];
        atomicAdd(&h[xa * b + xb], 1u);
        atomicAdd(&h[b * b + xb], 1u);
    }
    __syncthreads();
    int32_t *dst = marg + moff[v];
    for (int i = threadIdx.x; i < nc; i += 256) dst[i] = (int32_t)h[i];
}

// Four 0/1 bytes (one per sample) to four e2m1 nibbles in the low 16 bits,
// sample q at nibble q.
__device__ __forceinline__ uint32_t nibbles4(uint32_t w) {
    const uint32_t t = (w | (w >> 4)) & 0x00FF00FFu;  // bytes 0 and 2: two samples each
    return ((t & 0xFFu) | ((t >> 8) & 0xFF00u)) * (uint32_t)kOhFp4One;
}

// Operand rows, one 32-row block per wave and k chunk: lane (r, h) writes
// the 32 FP4 indicators of row r for samples kc*64 + 32h .. +31 (sample q
// at nibble q of its 16 bytes; zero past L and for padding rows), so each
// wave stores one contiguous kilobyte.  A and B use the same (lane, nibble)
// -> sample map, so every MFMA pairs each sample of a row with the same
// sample of a column.  TGT: rows (a, b) / x_t rows of targets; else source
// level rows.
template <bool TGT>
__global__ __launch_bounds__(256) void oh_pack_kernel(const int8_t *__restrict__ bins, int64_t ld,
                                                      int32_t tau, int64_t L, int64_t Kp,
                                                      const int32_t *__restrict__ rows,
                                                      int64_t n_blk, int8_t *__restrict__ out) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t blk = blockIdx.x;
    if (blk >= n_blk) return;
    const int64_t r = blk * 32 + (lane & 31);
    const int h = lane >> 5;
    const int var = rows[2 * r];
    const int code = rows[2 * r + 1];
    const int64_t kch = Kp / kOhKChunk;
    const int8_t *row = var >= 0 ? bins + (int64_t)var * ld : nullptr;
    for (int64_t kc = wave + 4 * (int64_t)blockIdx.y; kc < kch; kc += 4 * (int64_t)gridDim.y) {
        uint32_t o[4] = {0u, 0u, 0u, 0u};
        const int64_t tb = kc * kOhKChunk + 32 * h;
        if (var >= 0 && tb < L) {
#pragma unroll
            for (int half = 0; half < 2; half++) {
                const int64_t t0 = tb + 16 * half;
                uint32_t e[4] = {0u, 0u, 0u, 0u};  // 0x01 per sample
                if (t0 < L) {
                    uint32_t p[4];
                    load16(row, t0, L, ld, p);  // x_t (or y_t) for t < L
                    if (TGT) {
                        const uint32_t a = (uint32_t)(code >> 8), b = (uint32_t)(code & 0xFF);
                        if (a == 255u) {
#pragma unroll
                            for (int q = 0; q < 4; q++) e[q] = eq_bytes(p[q], b);
                        } else {
                            uint32_t f[4];
                            load16_shift(row, t0, tau, L + tau, ld, f);  // x_{t+tau}
#pragma unroll
                            for (int q = 0; q < 4; q++) e[q] = eq_bytes(p[q], b) & eq_bytes(f[q], a);
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; q++) e[q] = eq_bytes(p[q], (uint32_t)code);
                    }
                    // samples past L: load16 read zero bytes there, which equal level 0
                    const int64_t nv = L - t0;
                    if (nv < 16) {
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const int64_t vb = nv - 4 * q;
                            const uint32_t m =
                                vb >= 4 ? 0xFFFFFFFFu : vb <= 0 ? 0u : ((1u << (8 * vb)) - 1u);
                            e[q] &= m;
                        }
                    }
                }
                o[2 * half] = nibbles4(e[0]) | (nibbles4(e[1]) << 16);
                o[2 * half + 1] = nibbles4(e[2]) | (nibbles4(e[3]) << 16);
            }
        }
        *reinterpret_cast<uint4 *>(out + (blk * kch + kc) * 1024 + 16 * lane) =
            make_uint4(o[0], o[1], o[2], o[3]);
    }
}

// c log2 c (fixed point) of a count in [0, L]: LDS below kOhTbl, else the
// table in memory (an index outside [0, L] is clamped to L, so a wrong
// count can only give a wrong sum, never a read past the table).
__device__ __forceinline__ int64_t oh_f(const int64_t *lt, const int64_t *__restrict__ gt,
                                        int32_t c, uint32_t L) {
    const uint32_t u = (uint32_t)c;
    return u < (uint32_t)kOhTbl ? lt[u] : gt[u < L ? u : L];
}

// S3 - S2b of one pair (target of BX levels) from its staged counts: G
// points at the target's first row in the pair's first column (row stride
// kOhStr), nyc at the source's n_y(c).  The first C1 columns are stored
// (C1 = b_y, or b_y - 1 with the source's last level left out: then that
// column is the rows' totals, nt(a, b) and nt(b) of the target, less the
// stored columns).  Per column, the counts of a batch of x_t levels are
// read first, then all their c log2 c lookups, so a lane has many
// independent LDS reads in flight.
template <int BX>
__device__ __forceinline__ int64_t oh_pair(const uint16_t *G, const int32_t *nyc, int32_t ny_last,
                                           int by, int C1, const int32_t *__restrict__ nt,
                                           const int64_t *lfx, const int64_t *__restrict__ gfx,
                                           uint32_t uL, int32_t *cdbg) {
    constexpr int R1 = BX - 1;
    constexpr int BB = BX <= 4 ? BX : 4;  // x_t levels per batch
    int64_t s3 = 0, s2 = 0;
    for (int c = 0; c < C1; c++) {
        const int32_t ny = nyc[c];
        int32_t sumM = 0;  // M[b, c] over b < R1
#pragma unroll
        for (int b0 = 0; b0 < BX; b0 += BB) {
            int32_t v[BB][R1], M[BB];
#pragma unroll
            for (int u = 0; u < BB; u++) {
                const int b = b0 + u;
                if (b < BX) {
#pragma unroll
                    for (int aa = 0; aa < R1; aa++) v[u][aa] = G[(aa * BX + b) * kOhStr + c];
                    M[u] = b < R1 ? (int32_t)G[(R1 * BX + b) * kOhStr + c] : 0;
                }
            }
#pragma unroll
            for (int u = 0; u < BB; u++) {
                const int b = b0 + u;
                if (b < BX) {
                    if (b == R1) M[u] = ny - sumM;  // the last x_t level from n_y(c)
                    else sumM += M[u];
                    int32_t sa = 0;
                    int64_t f = 0;
#pragma unroll
                    for (int aa = 0; aa < R1; aa++) {
                        f += oh_f(lfx, gfx, v[u][aa], uL);
                        sa += v[u][aa];
                    }
                    s3 += f + oh_f(lfx, gfx, M[u] - sa, uL);
                    s2 += oh_f(lfx, gfx, M[u], uL);
                    if (cdbg) {
#pragma unroll
                        for (int aa = 0; aa < R1; aa++) cdbg[(aa * BX + b) * by + c] = v[u][aa];
                        cdbg[(R1 * BX + b) * by + c] = M[u] - sa;
                    }
                }
            }
        }
    }
    if (C1 < by) {
        // the source's last level, left out of the operands: each row's
        // total (n(a, b) for a state row, n(b) for an x_t row: nt[BX^2 + b])
        // less its stored columns
        const int c = C1;
        auto rest = [&](int r, int32_t tot) {
            int32_t s = 0;
            for (int cc = 0; cc < C1; cc++) s += (int32_t)G[r * kOhStr + cc];
            return tot - s;
        };
        int32_t sumM = 0;
        for (int b = 0; b < BX; b++) {
            const int32_t M = b < R1 ? rest(R1 * BX + b, nt[BX * BX + b]) : ny_last - sumM;
            sumM += M;
            int32_t sa = 0;
            for (int aa = 0; aa < R1; aa++) {
                const int32_t v = rest(aa * BX + b, nt[aa * BX + b]);
                s3 += oh_f(lfx, gfx, v, uL);
                sa += v;
                if (cdbg) cdbg[(aa * BX + b) * by + c] = v;
            }
            s3 += oh_f(lfx, gfx, M - sa, uL);
            s2 += oh_f(lfx, gfx, M, uL);
            if (cdbg) cdbg[(R1 * BX + b) * by + c] = M - sa;
        }
    }
    return s3 - s2;
}

// The K loop and the per-pair sums of one tile (see the file comment).
__global__ __launch_bounds__(256) void oh_mfma_kernel(tnx_onehot_args_t o, tnx_te_args_t a) {
    extern __shared__ __attribute__((aligned(16))) uint8_t oh_lds[];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wr = wave >> 1, wc = wave & 1;
    const int64_t tile = xcd_item(blockIdx.x, o.n_tiles);
    const int32_t *tl = o.tiles + tile * 8;
    const int rg0 = tl[0], rg1 = tl[1], cg0 = tl[2], cg1 = tl[3];
    const bool active = (tl[4] >> wave) & 1;
    const int64_t L = a.T - a.tau;
    int64_t *lfx = reinterpret_cast<int64_t *>(oh_lds + 4 * kOhSubBytes);
    for (int i = tid; i < kOhTbl; i += 256) lfx[i] = i <= L ? a.xlogx_fx[i] : 0;

    // Stage layout: block j (1 KB) at j * 1024; j < 8 NK: A (group g = j /
    // (4 NK), row block (j / NK) % 4, chunk j % NK); then B alike.  Wave w
    // fetches blocks w * 4 NK .. +4 NK.
    const int64_t kblk = o.Kp / kOhKChunk;  // 1 KB blocks per row block
    constexpr int PER = 4 * kOhNK;  // blocks per wave per stage
    // An absent group (-1) reads group 0 instead: every load is issued
    // unconditionally (the data of an absent group is never used), so the
    // compiler counts the loads in flight and waits only for the oldest
    // stage.
    const int8_t *src[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int j = wave * PER + i;
        const bool isB = j >= 8 * kOhNK;
        const int jj = isB ? j - 8 * kOhNK : j;
        const int g = jj / (4 * kOhNK), rb = (jj / kOhNK) % 4, kc = jj % kOhNK;
        const int grp = isB ? (g ? cg1 : cg0) : (g ? rg1 : rg0);
        const int8_t *base = isB ? o.B : o.A;
        src[i] = base + ((int64_t)(grp >= 0 ? grp : 0) * 4 + rb) * kblk * 1024 + (int64_t)kc * 1024 +
                 16 * lane;
    }
    const int64_t n_st = o.Kp / (kOhKChunk * kOhNK);
    auto fetch = [&](v4i (&r)[PER], int64_t s) {
#pragma unroll
        for (int i = 0; i < PER; i++)
            r[i] = *reinterpret_cast<const v4i *>(src[i] + s * kOhNK * 1024);
    };
    auto put = [&](const v4i (&r)[PER], int buf) {
#pragma unroll
        for (int i = 0; i < PER; i++)
            *reinterpret_cast<v4i *>(oh_lds + buf * kOhStageBytes + (wave * PER + i) * 1024 +
                                     16 * lane) = r[i];
    };
    v16f acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = v16f{};
    auto compute = [&](int buf) {
        if (!active) return;
        const uint8_t *st = oh_lds + buf * kOhStageBytes;
#pragma unroll
        for (int kc = 0; kc < kOhNK; kc++) {
            v8i fa[4], fb[4];  // FP4 operands: the low four dwords
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const v4i ua = *reinterpret_cast<const v4i *>(
                    st + ((wr * 4 + i) * kOhNK + kc) * 1024 + 16 * lane);
                const v4i ub = *reinterpret_cast<const v4i *>(
                    st + (8 * kOhNK + (wc * 4 + i) * kOhNK + kc) * 1024 + 16 * lane);
                fa[i] = v8i{ua[0], ua[1], ua[2], ua[3], 0, 0, 0, 0};
                fb[i] = v8i{ub[0], ub[1], ub[2], ub[3], 0, 0, 0, 0};
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                        fa[i], fb[j], acc[i][j], 4, 4, 0, kOhScale, 0, kOhScale);
        }
    };
    // two register sets, two LDS buffers: stage s is written to buffer s & 1
    // and stage s + 2 fetched into the freed registers before stage s is
    // computed; one barrier per stage
    // (fetches past the last stage re-read the last stage: no branch around
    // a load, so the loads of the newer stage stay in flight at each put)
    const int64_t last = n_st - 1;
    v4i r0[PER], r1[PER];
    fetch(r0, 0);
    __builtin_amdgcn_sched_barrier(0);  // keep stage 0's loads ahead of stage 1's
    fetch(r1, last < 1 ? last : 1);
    for (int64_t s = 0; s < n_st; s += 2) {
        put(r0, 0);
        fetch(r0, s + 2 < last ? s + 2 : last);
        __syncthreads();
        compute(0);
        put(r1, 1);
        fetch(r1, s + 3 < last ? s + 3 : last);
        __syncthreads();
        if (s + 1 < n_st) compute(1);
    }
    __syncthreads();  // the stages are dead: the staging area reuses them

    // Counts (exact in FP32, < 2^16: L <= 65535) to this wave's staging area.  32 x 32
    // accumulator layout: lane holds column (lane & 31), register q holds
    // row 8 (q / 4) + 4 (lane / 32) + q % 4.
    uint16_t *stg = reinterpret_cast<uint16_t *>(oh_lds + wave * kOhSubBytes);
    if (active) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const int row = i * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                    stg[row * kOhStr + j * 32 + (lane & 31)] = (uint16_t)(int)acc[i][j][q];
                }
    }
    __syncthreads();  // staging and table complete
    if (!active) return;
    const int grow = wr ? rg1 : rg0, gcol = wc ? cg1 : cg0;
    // n_y(c) of this column group's columns, indexed by column
    int32_t *mcol = reinterpret_cast<int32_t *>(oh_lds + 4 * kOhSubBytes + kOhTbl * 8) + wave * kOhGroup;
    for (int cc = lane; cc < kOhGroup; cc += kWave) {
        const int64_t br = (int64_t)gcol * kOhGroup + cc;
        const int v = o.brow[2 * br];
        mcol[cc] = v >= 0 ? o.marg[o.moff[v] + a.n_per_var[v] * a.n_per_var[v] + o.brow[2 * br + 1]]
                          : 0;
    }
    wave_lds_fence();
    const int tb = o.rg_off[grow], nt = o.rg_off[grow + 1] - tb;
    const int sb = o.cg_off[gcol], ns = o.cg_off[gcol + 1] - sb;
    const double invN = 1.0 / (double)L;
    const uint32_t uL = (uint32_t)L;
    for (int p = lane; p < nt * ns; p += kWave) {
        const int ti = tb + p / ns, sj = sb + p % ns;
        const int x = o.t_var[ti], y = o.s_var[sj];
        int64_t slot = 0;
        if (x == y || !out_slot(a, x, y, slot)) continue;
        const int bx = a.n_per_var[x], by = a.n_per_var[y];
        const uint16_t *G = stg + o.t_row[ti] * kOhStr + o.s_col[sj];
        // n_y(c) of the stored columns from LDS; the left-out last level's from memory
        const int C1 = (o.src_ie && by >= 2) ? by - 1 : by;
        const int32_t *nyc = mcol + o.s_col[sj];
        const int32_t *nt = o.marg + o.moff[x];   // the target's n(a, b), n(b)
        const int32_t ny_last = o.marg[o.moff[y] + by * by + by - 1];
        int32_t *cdbg = nullptr;  // test hook: this pair's joint counts N[a, b, c]
        if (a.counts_map) {
            const int r = a.counts_map[(int64_t)x * a.n_vars + y];
            if (r >= 0 && bx * bx * by <= a.counts_ld) cdbg = a.counts_dbg + (int64_t)r * a.counts_ld;
        }
        int64_t d = 0;
        switch (bx) {
#define TNX_OH_PAIR(BX) \
    case BX: d = oh_pair<BX>(G, nyc, ny_last, by, C1, nt, lfx, a.xlogx_fx, uL, cdbg); break;
            TNX_OH_PAIR(2) TNX_OH_PAIR(3) TNX_OH_PAIR(4) TNX_OH_PAIR(5) TNX_OH_PAIR(6)
            TNX_OH_PAIR(7) TNX_OH_PAIR(8) TNX_OH_PAIR(9) TNX_OH_PAIR(10) TNX_OH_PAIR(11)
#undef TNX_OH_PAIR
            default: break;  // the planner takes 2 <= b_x <= 11
        }
        const double s12 = a.S1[x] - a.S2a[x];
        emit_te(a, slot, (ldexp((double)d, -a.fx_shift) + s12) * invN);
    }
}

}  // namespace tnx

using namespace tnx;

extern "C" {

int tnx_onehot_pack(const tnx_onehot_args_t *o, const int8_t *bins, int64_t T, int64_t ld,
                    int32_t tau, const int32_t *n_per_var, void *stream) {
    clear_error();
    TNX_CHECK_ARG(o && bins && n_per_var, "tnx_onehot_pack: null pointer");
    TNX_CHECK_ARG(tau >= 1 && tau < T && ld >= T && ld % 16 == 0,
                  "tnx_onehot_pack: bad shape (T=%lld ld=%lld tau=%d)", (long long)T,
                  (long long)ld, tau);
    const int64_t L = T - tau;
    TNX_CHECK_ARG(o->Kp >= L && o->Kp % 128 == 0 && L < 65536,
                  "tnx_onehot_pack: Kp %lld must be a multiple of 128 >= L = %lld (L < 65536)",
                  (long long)o->Kp, (long long)L);
    TNX_CHECK_ARG(o->rows_a % kOhGroup == 0 && o->rows_b % kOhGroup == 0,
                  "tnx_onehot_pack: rows_a and rows_b must be multiples of %d", kOhGroup);
    hipStream_t s = as_stream(stream);
    if (o->n_mvars > 0) {
        TNX_CHECK_ARG(o->mvars && o->moff && o->marg, "tnx_onehot_pack: null marginal arrays");
        hipLaunchKernelGGL(oh_marg_kernel, dim3((unsigned)o->n_mvars), dim3(256), 0, s, bins, ld,
                           tau, L, n_per_var, o->mvars, o->moff, o->marg);
        TNX_CHECK_LAUNCH("oh_marg_kernel");
    }
    const int64_t kch = o->Kp / kOhKChunk;
    const unsigned gy = (unsigned)std::min<int64_t>(8, (kch + 3) / 4);
    if (o->rows_a > 0) {
        TNX_CHECK_ARG(o->A && o->arow, "tnx_onehot_pack: null A");
        hipLaunchKernelGGL(oh_pack_kernel<true>, dim3((unsigned)(o->rows_a / 32), gy), dim3(256), 0,
                           s, bins, ld, tau, L, o->Kp, o->arow, o->rows_a / 32, o->A);
        TNX_CHECK_LAUNCH("oh_pack_kernel<true>");
    }
    if (o->rows_b > 0) {
        TNX_CHECK_ARG(o->B && o->brow, "tnx_onehot_pack: null B");
        hipLaunchKernelGGL(oh_pack_kernel<false>, dim3((unsigned)(o->rows_b / 32), gy), dim3(256),
                           0, s, bins, ld, tau, L, o->Kp, o->brow, o->rows_b / 32, o->B);
        TNX_CHECK_LAUNCH("oh_pack_kernel<false>");
    }
    return TNX_OK;
}

int tnx_onehot_te(const tnx_onehot_args_t *o, const tnx_te_args_t *te, void *stream) {
    clear_error();
    TNX_CHECK_ARG(o && te, "tnx_onehot_te: null pointer");
    TNX_CHECK_ARG(o->A && o->B && o->tiles && o->rg_off && o->t_var && o->t_row && o->cg_off &&
                      o->s_var && o->s_col && o->marg && o->moff,
                  "tnx_onehot_te: null plan array");
    TNX_CHECK_ARG(te->n_per_var && te->S1 && te->S2a && te->xlogx_fx,
                  "tnx_onehot_te: te args need n_per_var, S1, S2a and xlogx_fx");
    TNX_CHECK_ARG(te->fx_shift >= 0 && te->fx_shift <= 62, "tnx_onehot_te: bad fx_shift %d",
                  te->fx_shift);
    const int64_t L = te->T - te->tau;
    TNX_CHECK_ARG(te->tau >= 1 && L >= 1 && L < 65536 && o->Kp >= L && o->Kp % 128 == 0,
                  "tnx_onehot_te: need 1 <= L < 65536 and Kp a multiple of 128 >= L");
    TNX_CHECK_ARG(te->out_mode >= TNX_OUT_PAIRFREE && te->out_mode <= TNX_OUT_STAGE,
                  "tnx_onehot_te: bad out_mode %d", te->out_mode);
    if (o->n_tiles <= 0) return TNX_OK;
    TNX_CHECK_HIP(hipFuncSetAttribute((const void *)oh_mfma_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, kOhLdsBytes));
    hipLaunchKernelGGL(oh_mfma_kernel, dim3((unsigned)o->n_tiles), dim3(256), kOhLdsBytes,
                       as_stream(stream), *o, *te);
    TNX_CHECK_LAUNCH("oh_mfma_kernel");
    return TNX_OK;
}

}  // extern "C"
