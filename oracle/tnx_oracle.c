/*
 * tnx_oracle.c -- CPU oracle for the transfer-entropy (TE) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (tenex_amd/) links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Every routine restates the reference algorithm it cites (paths are relative
 * to the cxinsys/tenex source tree) in plain C with exact integer counts and
 * FP64 arithmetic.  Built by oracle/Makefile into oracle/lib/libtnx_oracle.so.
 *
 * Parity pins (what checks each function independently of this file):
 *   orc_np_rowsum_f32 / orc_fsbw_l : bit-exact against numpy.std/min/max on
 *       random rows of many lengths (tests/test_oracle.py); numpy is the
 *       library the reference calls (preprocess.py:80-87).
 *   orc_remap            : against numpy.unique(return_inverse=True).
 *   orc_joint_counts     : sum of counts == N, marginal identities, brute
 *       force over a Python loop on tiny inputs.
 *   orc_te_pairs         : two independent FP64 forms (entropy S-form and the
 *       reference's direct log-ratio form, scatter_add.py:98-117) must agree;
 *       closed forms: constant source gives 0, planted copy gives
 *       H(X_{t+1}|X_t), 0 <= TE <= log2(b_t), relabel invariance.
 *   orc_te_b2            : the 8-state inclusion-exclusion (gemm_b2.py:240-284)
 *       against orc_te_pairs on binary data.
 *   orc_surrogate        : identity permutation gives te_k == te_obs; stats
 *       against numpy on the per-k TE values it returns.
 *   orc_mlp_*            : finite differences of orc_mlp_loss.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* numpy float32 pairwise summation (numpy/_core/src/umath/loops_utils.h.src, */
/* pairwise_sum_FLOAT), fed in buffers of 8192 elements along a contiguous   */
/* reduction axis, as np.add.reduce does for float32 rows.                   */
/* ------------------------------------------------------------------------- */
static float pw_sum_f32(const float *a, int64_t n) {
    if (n < 8) {
        float res = 0.0f;
        for (int64_t i = 0; i < n; i++) res += a[i];
        return res;
    } else if (n <= 128) {
        float r[8];
        int64_t i;
        for (int j = 0; j < 8; j++) r[j] = a[j];
        for (i = 8; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; j++) r[j] += a[i + j];
        float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; i++) res += a[i];
        return res;
    } else {
        int64_t n2 = n / 2;
        n2 -= n2 % 8;
        return pw_sum_f32(a, n2) + pw_sum_f32(a + n2, n - n2);
    }
}

EXPORT float orc_np_rowsum_f32(const float *a, int64_t n) {
    float res = 0.0f;
    for (int64_t s = 0; s < n; s += 8192) {
        int64_t m = (n - s < 8192) ? (n - s) : 8192;
        float p = pw_sum_f32(a + s, m);
        res = (s == 0) ? p : (res + p);
    }
    return res;
}

/*
 * FSBW-L discretisation, numpy variant: preprocess.py:61-88
 * (bin_fsbw_l_numpy) followed by the clamp(min=0) of discretize()
 * (preprocess.py:622).  std is numpy's float32 std with ddof=1
 * (numpy/_core/_methods.py _var: float32 pairwise sums, the two divisions
 * by an intp count done in float64 then rounded to float32).
 */
EXPORT int orc_fsbw_l(const float *x, int64_t n_vars, int64_t T, float kp,
                      int32_t *raw_out, int32_t *n_bins_out, float *std_out,
                      float *smin_out) {
    if (T < 2) return -1;
    float *tmp = (float *)malloc(sizeof(float) * (size_t)T);
    if (!tmp) return -2;
    for (int64_t g = 0; g < n_vars; g++) {
        const float *row = x + g * T;
        float s = orc_np_rowsum_f32(row, T);
        float mean = (float)((double)s / (double)T);
        for (int64_t t = 0; t < T; t++) {
            float d = row[t] - mean;
            tmp[t] = d * d;
        }
        float ss = orc_np_rowsum_f32(tmp, T);
        float var = (float)((double)ss / (double)(T - 1));
        float sd = sqrtf(var);
        if (sd == 0.0f || isnan(sd)) sd = 1.0f;
        float mn = row[0], mx = row[0];
        for (int64_t t = 1; t < T; t++) {
            if (row[t] < mn) mn = row[t];
            if (row[t] > mx) mx = row[t];
        }
        float smin = mn - kp * sd;
        if (n_bins_out) n_bins_out[g] = (int32_t)ceilf((mx - mn) / sd);
        if (std_out) std_out[g] = sd;
        if (smin_out) smin_out[g] = smin;
        for (int64_t t = 0; t < T; t++) {
            int32_t b = (int32_t)floorf((row[t] - smin) / sd);
            raw_out[g * T + t] = b < 0 ? 0 : b;
        }
    }
    free(tmp);
    return 0;
}

/* Dense per-gene relabelling: preprocess.py:401-459 (remap_bins), i.e. the
 * inverse of torch.unique(sorted=True, return_inverse=True) per gene. */
static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

EXPORT int orc_remap(const int32_t *raw, int64_t n_vars, int64_t TK,
                     int32_t *dense_out, int32_t *n_per_var_out) {
    int32_t *u = (int32_t *)malloc(sizeof(int32_t) * (size_t)TK);
    if (!u) return -2;
    for (int64_t g = 0; g < n_vars; g++) {
        const int32_t *r = raw + g * TK;
        memcpy(u, r, sizeof(int32_t) * (size_t)TK);
        qsort(u, (size_t)TK, sizeof(int32_t), cmp_i32);
        int64_t m = 0;
        for (int64_t i = 0; i < TK; i++)
            if (i == 0 || u[i] != u[m - 1]) u[m++] = u[i];
        n_per_var_out[g] = (int32_t)m;
        for (int64_t t = 0; t < TK; t++) {
            int64_t lo = 0, hi = m - 1;
            while (lo < hi) {
                int64_t mid = (lo + hi) / 2;
                if (u[mid] < r[t]) lo = mid + 1; else hi = mid;
            }
            dense_out[g * TK + t] = (int32_t)lo;
        }
    }
    free(u);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Joint counts and TE.                                                       */
/* Histogram of (a = x_{t+tau}, b = x_t, c = y_t), t in [0, L), L = T - tau,  */
/* full_smem.cu:62-84 / scatter_add.py:63-90; stored dense as                 */
/* cnt3[(a * bt + b) * bs + c] with bt = n_per_var[tgt], bs = n_per_var[src]. */
/* ------------------------------------------------------------------------- */
static inline double xlog2x(double c) { return c > 0.0 ? c * log2(c) : 0.0; }

EXPORT int orc_joint_counts(const int32_t *bins, int64_t n_vars, int64_t T, int tau,
                            const int32_t *npv, int64_t tgt, int64_t src,
                            int32_t *cnt3_out) {
    if (tgt < 0 || tgt >= n_vars || src < 0 || src >= n_vars || tau < 1 || tau >= T)
        return -1;
    int64_t bt = npv[tgt], bs = npv[src];
    memset(cnt3_out, 0, sizeof(int32_t) * (size_t)(bt * bt * bs));
    const int32_t *x = bins + tgt * T, *y = bins + src * T;
    for (int64_t t = 0; t + tau < T; t++) {
        int64_t a = x[t + tau], b = x[t], c = y[t];
        if (a < 0 || a >= bt || b < 0 || b >= bt || c < 0 || c >= bs) return -3;
        cnt3_out[(a * bt + b) * bs + c]++;
    }
    return 0;
}

/* Per-gene entropy sums S1 = sum_b c1 log2 c1 over x_t and
 * S2a = sum_{a,b} c2a log2 c2a over (x_{t+tau}, x_t), t in [0, L). */
EXPORT int orc_gene_entropies(const int32_t *bins, int64_t n_vars, int64_t T, int tau,
                              const int32_t *npv, double *S1, double *S2a) {
#pragma omp parallel for schedule(dynamic)
    for (int64_t g = 0; g < n_vars; g++) {
        int64_t b = npv[g];
        int64_t *c1 = (int64_t *)calloc((size_t)b, sizeof(int64_t));
        int64_t *c2 = (int64_t *)calloc((size_t)(b * b), sizeof(int64_t));
        const int32_t *x = bins + g * T;
        for (int64_t t = 0; t + tau < T; t++) {
            c1[x[t]]++;
            c2[x[t + tau] * b + x[t]]++;
        }
        double s1 = 0, s2 = 0;
        for (int64_t i = 0; i < b; i++) s1 += xlog2x((double)c1[i]);
        for (int64_t i = 0; i < b * b; i++) s2 += xlog2x((double)c2[i]);
        S1[g] = s1;
        S2a[g] = s2;
        free(c1);
        free(c2);
    }
    return 0;
}

/* TE(src -> tgt) in bits from a dense cnt3 (layout above).
 * te_s:      (S3 + S1 - S2a - S2b) / N  (entropy form)
 * te_direct: (1/N) sum c3 log2(c3 c1 / (c2a c2b)) (scatter_add.py:98-117,
 *            full_smem.cu:113-128), evaluated in FP64. */
static void te_from_counts(const int32_t *cnt3, int64_t bt, int64_t bs, double N,
                           double *te_s, double *te_direct) {
    int64_t B2a = bt * bt, B2b = bt * bs;
    double *c2a = (double *)calloc((size_t)B2a, sizeof(double));
    double *c2b = (double *)calloc((size_t)B2b, sizeof(double));
    double *c1 = (double *)calloc((size_t)bt, sizeof(double));
    double S3 = 0;
    for (int64_t a = 0; a < bt; a++)
        for (int64_t b = 0; b < bt; b++)
            for (int64_t c = 0; c < bs; c++) {
                double v = cnt3[(a * bt + b) * bs + c];
                c2a[a * bt + b] += v;
                c2b[b * bs + c] += v;
                c1[b] += v;
                S3 += xlog2x(v);
            }
    double S1 = 0, S2a = 0, S2b = 0, D = 0;
    for (int64_t i = 0; i < bt; i++) S1 += xlog2x(c1[i]);
    for (int64_t i = 0; i < B2a; i++) S2a += xlog2x(c2a[i]);
    for (int64_t i = 0; i < B2b; i++) S2b += xlog2x(c2b[i]);
    for (int64_t a = 0; a < bt; a++)
        for (int64_t b = 0; b < bt; b++)
            for (int64_t c = 0; c < bs; c++) {
                double v = cnt3[(a * bt + b) * bs + c];
                double den = c2a[a * bt + b] * c2b[b * bs + c];
                if (v > 0 && den > 0) D += v * log2(v * c1[b] / den);
            }
    *te_s = (S3 + S1 - S2a - S2b) / N;
    *te_direct = D / N;
    free(c2a);
    free(c2b);
    free(c1);
}

/* TE for an explicit pair list pairs[p] = (target, source) (full_smem.cu:55-56
 * convention).  Bins are dense per gene: 0 <= bins[g, t] < npv[g]. */
EXPORT int orc_te_pairs(const int32_t *bins, int64_t n_vars, int64_t T, int tau,
                        const int32_t *npv, const int64_t *pairs, int64_t n_pairs,
                        double *te_s, double *te_direct) {
    if (tau < 1 || tau >= T) return -1;
    double N = (double)(T - tau);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : bad)
    for (int64_t p = 0; p < n_pairs; p++) {
        int64_t tg = pairs[2 * p], sr = pairs[2 * p + 1];
        int64_t bt = npv[tg], bs = npv[sr];
        int32_t *cnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(bt * bt * bs));
        if (orc_joint_counts(bins, n_vars, T, tau, npv, tg, sr, cnt) != 0) {
            bad |= 1;
        } else {
            double a, b;
            te_from_counts(cnt, bt, bs, N, &a, &b);
            te_s[p] = a;
            if (te_direct) te_direct[p] = b;
        }
        free(cnt);
    }
    return bad ? -3 : 0;
}

/*
 * GEMM-B2 closed form (gemm_b2.py:210-285): for binary bins, the three
 * contractions C = fp.p^T, D = p.p^T, E = f.p^T (f = x_{t+tau}, p = x_t,
 * fp = f*p) plus per-gene marginals give the 8 joint counts by
 * inclusion-exclusion; TE from the 8-state log-ratio sum, here in FP64.
 * te[i * n + j] = TE(j -> i) (internal layout), diagonal 0.
 * Optionally returns C, D, E (int64 n*n) for bit-exact count checks.
 */
EXPORT int orc_te_b2(const int32_t *bins, int64_t n, int64_t T, int tau, double *te,
                     int64_t *Cout, int64_t *Dout, int64_t *Eout) {
    if (tau < 1 || tau >= T) return -1;
    int64_t L = T - tau;
    double N = (double)L;
    int64_t *nf = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    int64_t *np_ = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    int64_t *nfp = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++)
        for (int64_t t = 0; t < L; t++) {
            int32_t f = bins[i * T + t + tau], p = bins[i * T + t];
            nf[i] += f;
            np_[i] += p;
            nfp[i] += f * p;
        }
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < n; i++) {
        const int32_t *xi = bins + i * T;
        for (int64_t j = 0; j < n; j++) {
            const int32_t *xj = bins + j * T;
            int64_t C = 0, D = 0, E = 0;
            for (int64_t t = 0; t < L; t++) {
                int32_t f = xi[t + tau], p = xi[t], q = xj[t];
                C += f * p * q;
                D += p * q;
                E += f * q;
            }
            if (Cout) { Cout[i * n + j] = C; Dout[i * n + j] = D; Eout[i * n + j] = E; }
            double npj = (double)np_[j], nfi = (double)nf[i], npi = (double)np_[i],
                   nfpi = (double)nfp[i];
            double c[8] = {
                (double)C, (double)(D - C), (double)(E - C), npj - D - E + C,
                nfpi - C, npi - D - nfpi + C, nfi - E - nfpi + C,
                N - nfi - npi - npj + nfpi + D + E - C};
            double c1[8] = {npi, npi, N - npi, N - npi, npi, npi, N - npi, N - npi};
            double c2a_[4] = {nfpi, npi - nfpi, nfi - nfpi, N - nfi - npi + nfpi};
            double c2b_[4] = {(double)D, npj - D, npi - D, N - npi - npj + D};
            double s = 0;
            for (int k = 0; k < 8; k++) {
                double c2a = c2a_[k & 3];
                double c2b = c2b_[k >> 1];
                double den = c2a * c2b;
                if (den < 1) den = 1;
                if (c[k] > 0) s += c[k] * log2(c[k] * c1[k] / den);
            }
            te[i * n + j] = (i == j) ? 0.0 : s / N;
        }
    }
    free(nf);
    free(np_);
    free(nfp);
    return 0;
}

/*
 * Surrogate statistics (full_smem_surrogate_test.cu:94-246 semantics,
 * surrogate_test.py:107-148 shuffle): for surrogate k each variable v is
 * read through its own block permutation perm[k, v, :] over the head
 * region [0, n_blocks*block_length); the tail is untouched.  For pair
 * (tgt, src): x'_{t+tau} = x[map_t(t+tau)], x'_t = x[map_t(t)],
 * y'_t = y[map_s(t)].  Accumulates in FP64 per pair:
 *   sum += te_k, sumsq += te_k^2, count_ge += (te_k >= obs[p]).
 * te_k_out (optional, n_pairs * n_surr) receives each te_k.
 */
static inline int64_t perm_map(const int32_t *perm, int64_t t, int64_t head, int64_t bl) {
    if (t >= head) return t;
    int64_t bi = t / bl;
    return (int64_t)perm[bi] * bl + (t - bi * bl);
}

EXPORT int orc_surrogate(const int32_t *bins, int64_t n_vars, int64_t T, int tau,
                         const int32_t *npv, const int32_t *perm, int64_t n_surr,
                         int64_t n_blocks, int64_t block_length, const int64_t *pairs,
                         int64_t n_pairs, const float *obs, double *sum, double *sumsq,
                         int32_t *count_ge, double *te_k_out) {
    if (tau < 1 || tau >= T) return -1;
    int64_t L = T - tau, head = n_blocks * block_length;
    double N = (double)L;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
    for (int64_t p = 0; p < n_pairs; p++) {
        int64_t tg = pairs[2 * p], sr = pairs[2 * p + 1];
        int64_t bt = npv[tg], bs = npv[sr];
        int64_t cells = bt * bt * bs;
        int32_t *cnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)cells);
        const int32_t *x = bins + tg * T, *y = bins + sr * T;
        double s = 0, s2 = 0;
        int32_t cg = 0;
        for (int64_t k = 0; k < n_surr; k++) {
            const int32_t *pt = perm + (k * n_vars + tg) * n_blocks;
            const int32_t *ps = perm + (k * n_vars + sr) * n_blocks;
            memset(cnt, 0, sizeof(int32_t) * (size_t)cells);
            for (int64_t t = 0; t < L; t++) {
                int64_t a = x[perm_map(pt, t + tau, head, block_length)];
                int64_t b = x[perm_map(pt, t, head, block_length)];
                int64_t c = y[perm_map(ps, t, head, block_length)];
                if (a >= bt || b >= bt || c >= bs || a < 0 || b < 0 || c < 0) { bad |= 1; break; }
                cnt[(a * bt + b) * bs + c]++;
            }
            double te_s, te_d;
            te_from_counts(cnt, bt, bs, N, &te_s, &te_d);
            s += te_s;
            s2 += te_s * te_s;
            if ((float)te_s >= obs[p]) cg++;
            if (te_k_out) te_k_out[p * n_surr + k] = te_s;
        }
        sum[p] = s;
        sumsq[p] = s2;
        count_ge[p] = cg;
        free(cnt);
    }
    return bad ? -3 : 0;
}

/* ------------------------------------------------------------------------- */
/* Learned-null MLP (no reference counterpart): 3 dense layers               */
/*   h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3             */
/* with W1 [H x D], W2 [H x H], W3 [O x H] row-major; loss = mean over rows   */
/* and outputs of 0.5 * w_o * (y - target)^2.  FP64 throughout.               */
/* params layout: W1, b1, W2, b2, W3, b3 (contiguous).                        */
/* ------------------------------------------------------------------------- */
EXPORT int64_t orc_mlp_n_params(int64_t D, int64_t H, int64_t O) {
    return H * D + H + H * H + H + O * H + O;
}

EXPORT int orc_mlp_forward(const double *params, int64_t D, int64_t H, int64_t O,
                           const double *X, int64_t n_rows, double *Y) {
    const double *W1 = params, *b1 = W1 + H * D, *W2 = b1 + H, *b2 = W2 + H * H,
                 *W3 = b2 + H, *b3 = W3 + O * H;
#pragma omp parallel for
    for (int64_t r = 0; r < n_rows; r++) {
        double h1[1024], h2[1024];
        if (H > 1024) continue;
        for (int64_t i = 0; i < H; i++) {
            double s = b1[i];
            for (int64_t k = 0; k < D; k++) s += W1[i * D + k] * X[r * D + k];
            h1[i] = s > 0 ? s : 0;
        }
        for (int64_t i = 0; i < H; i++) {
            double s = b2[i];
            for (int64_t k = 0; k < H; k++) s += W2[i * H + k] * h1[k];
            h2[i] = s > 0 ? s : 0;
        }
        for (int64_t o = 0; o < O; o++) {
            double s = b3[o];
            for (int64_t k = 0; k < H; k++) s += W3[o * H + k] * h2[k];
            Y[r * O + o] = s;
        }
    }
    return H > 1024 ? -1 : 0;
}

/* loss and gradient (grad has the params layout). */
EXPORT double orc_mlp_loss_grad(const double *params, int64_t D, int64_t H, int64_t O,
                                const double *X, const double *Tgt, const double *wout,
                                int64_t n_rows, double *grad) {
    const double *W1 = params, *b1 = W1 + H * D, *W2 = b1 + H, *b2 = W2 + H * H,
                 *W3 = b2 + H, *b3 = W3 + O * H;
    int64_t P = orc_mlp_n_params(D, H, O);
    if (grad) memset(grad, 0, sizeof(double) * (size_t)P);
    double *gW1 = grad, *gb1 = gW1 ? gW1 + H * D : 0, *gW2 = gb1 ? gb1 + H : 0,
           *gb2 = gW2 ? gW2 + H * H : 0, *gW3 = gb2 ? gb2 + H : 0, *gb3 = gW3 ? gW3 + O * H : 0;
    double loss = 0;
    double inv = 1.0 / (double)(n_rows * O);
    double h1[1024], h2[1024], z1[1024], z2[1024], d2[1024], d1[1024], dy[64];
    if (H > 1024 || O > 64) return NAN;
    for (int64_t r = 0; r < n_rows; r++) {
        const double *x = X + r * D;
        for (int64_t i = 0; i < H; i++) {
            double s = b1[i];
            for (int64_t k = 0; k < D; k++) s += W1[i * D + k] * x[k];
            z1[i] = s;
            h1[i] = s > 0 ? s : 0;
        }
        for (int64_t i = 0; i < H; i++) {
            double s = b2[i];
            for (int64_t k = 0; k < H; k++) s += W2[i * H + k] * h1[k];
            z2[i] = s;
            h2[i] = s > 0 ? s : 0;
        }
        for (int64_t o = 0; o < O; o++) {
            double s = b3[o];
            for (int64_t k = 0; k < H; k++) s += W3[o * H + k] * h2[k];
            double e = s - Tgt[r * O + o];
            double w = wout ? wout[o] : 1.0;
            loss += 0.5 * w * e * e * inv;
            dy[o] = w * e * inv;
        }
        if (!grad) continue;
        for (int64_t o = 0; o < O; o++) {
            gb3[o] += dy[o];
            for (int64_t k = 0; k < H; k++) gW3[o * H + k] += dy[o] * h2[k];
        }
        for (int64_t k = 0; k < H; k++) {
            double s = 0;
            for (int64_t o = 0; o < O; o++) s += W3[o * H + k] * dy[o];
            d2[k] = z2[k] > 0 ? s : 0;
        }
        for (int64_t i = 0; i < H; i++) {
            gb2[i] += d2[i];
            for (int64_t k = 0; k < H; k++) gW2[i * H + k] += d2[i] * h1[k];
        }
        for (int64_t k = 0; k < H; k++) {
            double s = 0;
            for (int64_t i = 0; i < H; i++) s += W2[i * H + k] * d2[i];
            d1[k] = z1[k] > 0 ? s : 0;
        }
        for (int64_t i = 0; i < H; i++) {
            gb1[i] += d1[i];
            for (int64_t k = 0; k < D; k++) gW1[i * D + k] += d1[i] * x[k];
        }
    }
    return loss;
}
