/*
 * tenex_hip.h -- C ABI of libtenex_hip.so, the MI355X (gfx950) transfer-entropy
 * engine.
 *
 * Conventions
 *  - Every entry point returns int: 0 = OK, negative = error; the message is
 *    in tnx_last_error() (thread-local, valid until the next call on that
 *    thread).
 *  - All array arguments are device pointers unless the comment says "host".
 *  - The library never allocates or frees device memory and never
 *    synchronises: callers own every buffer and pass a hipStream_t as void*.
 *    Kernels run on that stream, on the calling thread's current device.
 *  - Bins are dense per variable (0 <= bins[v, t] < n_per_var[v] <= 127),
 *    int8, K = 1, row stride `ld` bytes (ld >= T, ld % 16 == 0); the entry
 *    points taking int16 rows (levels < 32768) say so.
 *  - Pair convention of the reference: a pair is (target, source); its TE is
 *    TE(source -> target).  Pair-free linear order:
 *       id = target * (n - 1) + s,   source = s + (s >= target)
 *    (tenex/csrc/full_smem.cu:190-195).
 *
 * Reference interfaces replaced (paths in the cxinsys/tenex tree):
 *  - get_smem_optin                      tenex/csrc/full_smem.cu:150-155
 *  - bin_fsbw_l_numpy + clamp            tenex/preprocess.py:61-88, :622
 *  - remap_bins / _compact_dtype         tenex/preprocess.py:401-459, :335-354
 *  - coarsen_bins_cuda                   tenex/csrc/adaptive_smem.cu:557-592
 *  - te_smem_launch(_pairfree)           tenex/csrc/full_smem.cu:319-381
 *  - te_adaptive_smem_launch(_pairfree)  tenex/csrc/adaptive_smem.cu:307-376
 *  - te_gmem_launch                      tenex/csrc/gmem.cu:190-230
 *  - compute_te_gemm_b2 (torch.mm x3 + Triton epilogue)
 *                                        tenex/kernels/gemm_b2.py:307-425
 *  - full/adaptive_smem_surrogate_test_launch
 *                                        tenex/csrc/full_smem_surrogate_test.cu:297-394
 *                                        tenex/csrc/adaptive_smem_surrogate_test.cu:306-400
 *  - fused_outte/inte_4way_launch        tenex/csrc/trace_sort_entropy.cu:191-304
 *  - the binarizers                      tenex/binarize.py:51-200
 *  - the DPI mediator loop of _trim      tenex/inference/fdr.py:104-110
 */
#ifndef TENEX_HIP_H
#define TENEX_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define TNX_API __attribute__((visibility("default")))
#else
#define TNX_API
#endif

#define TNX_ABI_VERSION 13

/* Status codes. */
#define TNX_OK 0
#define TNX_E_ARG (-1)     /* invalid argument / shape */
#define TNX_E_HIP (-2)     /* HIP runtime error */
#define TNX_E_RANGE (-3)   /* value outside the supported range */
#define TNX_E_CAP (-4)     /* caller-provided capacity too small */

/* Library ABI version (== TNX_ABI_VERSION). */
TNX_API int tnx_abi_version(void);

/* Message of the last failing call on this thread ("" if none). */
TNX_API const char *tnx_last_error(void);

typedef struct {
    int64_t lds_bytes_per_block; /* max dynamic LDS per workgroup (160 KiB on gfx950) */
    int32_t n_cu;                /* compute units */
    int32_t wave_size;           /* 64 */
    int32_t clock_khz;
    int32_t gcn_arch;            /* e.g. 950 */
    char name[64];               /* gcnArchName */
} tnx_device_props_t;

/* Device properties of `device` (host struct out).
 * Replaces get_smem_optin (tenex/csrc/full_smem.cu:150-155). */
TNX_API int tnx_device_props(int device, tnx_device_props_t *out);

/* ── Discretisation ───────────────────────────────────────────────────── */

/* FSBW-L with numpy-exact statistics: per variable, std (ddof=1, numpy's
 * float32 pairwise sums), min, max; n_bins = ceil((max-min)/std);
 * raw[v,t] = max(0, floor((x - (min - kp*std)) / std)), std == 0 -> 1.
 * x: float32 (n_vars, T) row-major.  raw_out int32 (n_vars, T).  Any of
 * n_bins_out / std_out / smin_out may be NULL.
 * Replaces bin_fsbw_l_numpy + clamp (tenex/preprocess.py:61-88, :622). */
TNX_API int tnx_fsbw_l(const float *x, int64_t n_vars, int64_t T, float kp, int32_t *raw_out,
               int32_t *n_bins_out, float *std_out, float *smin_out, void *stream);

/* Dense per-variable relabelling (inverse of sorted unique).  raw int32
 * (n_vars, T) with 0 <= raw < 65536; writes int8 bins (row stride ld) and
 * n_per_var.  status (device int32, caller-zeroed) gets bit 1 if a raw value
 * is >= 65536 or < 0, bit 2 if a variable has more than 127 levels.
 * Replaces remap_bins + _compact_dtype (tenex/preprocess.py:401-459, :335-354). */
TNX_API int tnx_remap_dense(const int32_t *raw, int64_t n_vars, int64_t T, int8_t *bins_out,
                    int64_t ld, int32_t *n_per_var_out, int32_t *status, void *stream);

/* tnx_remap_dense into int16 rows for variables of up to 32,767 levels
 * (the reference's int16 bins, tenex/preprocess.py:335-354): status bit 2
 * (4) when a variable has more than 32,767 levels. */
TNX_API int tnx_remap_dense16(const int32_t *raw, int64_t n_vars, int64_t T, int16_t *bins_out,
                              int64_t ld, int32_t *n_per_var_out, int32_t *status, void *stream);

/* tnx_coarsen_bins from int16 rows into int8 rows bins_out (same ld): the
 * reference's coarsening of variables above the LDS bin limit applied to
 * bins of up to 32,767 levels (tenex/preprocess.py:462-551). */
TNX_API int tnx_coarsen_bins16(const int16_t *bins16, int64_t n_vars, int64_t T, int64_t ld,
                               int32_t *n_per_var, int32_t limit, int8_t *bins_out,
                               void *stream);

/* In-place lossy coarsening: for variables with n_per_var > limit,
 * bin = bin * limit / n_per_var and n_per_var = limit.
 * Replaces coarsen_bins_cuda (tenex/csrc/adaptive_smem.cu:557-592). */
TNX_API int tnx_coarsen_bins(int8_t *bins, int64_t n_vars, int64_t T, int64_t ld,
                     int32_t *n_per_var, int32_t limit, void *stream);

/* ── Per-variable preparation ─────────────────────────────────────────── */

/* c*log2(c) for c in [0, N] as float hi/lo pairs (tbl[2c], tbl[2c+1]) with
 * hi + lo == the FP64 value to ~2^-48 relative.  tbl: float (2*(N+1)). */
TNX_API int tnx_xlogx_table(float *tbl, int64_t N, void *stream);
/* Binarization into {0, 1} bins (replaces the binarizers of tenex/binarize.py:
 * 51-200 and mean_binarize / threshold_binarize of
 * tenex/kernels/binarize_kernel.py): per variable a threshold thr_out[g]
 * (device float[n_vars]), then bins_out[g * ld_out + t] = x[g * ldx + t] >
 * thr_out[g] for t < T, zero padding up to ld_out.  method 0: thr = 0
 * ('nonzero'); 1: numpy's float32 mean ('mean'); 2: numpy's median
 * ('median'); 3: numpy's linear quantile q in float32 ('quantile').  x:
 * device float32 rows of stride ldx. */
TNX_API int tnx_binarize(const float *x, int64_t n_vars, int64_t T, int64_t ldx, int32_t method,
                         float q, int8_t *bins_out, int64_t ld_out, float *thr_out, void *stream);

/* c*log2(c) for c in [0, N] as 64-bit fixed point: tbl[c] = round(c log2 c
 * * 2^shift) (the FP64 value tnx_xlogx_table splits).  Sums of entries are
 * exact integers, independent of their order; shift must keep N log2 N *
 * 2^shift below 2^62 (the Python layer uses 55 - ceil(log2(N + 1)), 41 fraction
 * bits at N = 9,999).  The persistent TE launch reduces with it. */
TNX_API int tnx_xlogx_table_fx(int64_t *tbl, int64_t N, int32_t shift, void *stream);

/* Per variable v with b = n_per_var[v], L = T - tau:
 *   states[v, t] = x[t+tau] * b + x[t]          (uint16, row stride lds)
 *   S1[v]  = sum_b  c1 log2 c1   over x_t,             t < L   (FP64)
 *   S2a[v] = sum_ab c2a log2 c2a over (x_{t+tau}, x_t), t < L   (FP64)
 * feat (optional, float (n_vars, 8)): b, H(X_t), H(X_t+1|X_t),
 * fraction of x_t == 0, fraction of x_t == mode, log2 b, mean level / b,
 * lag-1 self agreement.  b_max: an upper bound of n_per_var (it sizes the
 * LDS histogram; 0 means 127); a variable above it gets NaN entropies.
 * ld and lds must be multiples of 16 with lds <= ld.
 * (Per-target marginals of every TE kernel, full_smem.cu:87-141.) */
TNX_API int tnx_gene_prep(const int8_t *bins, int64_t n_vars, int64_t T, int64_t ld, int32_t tau,
                  const int32_t *n_per_var, uint16_t *states, int64_t lds,
                  double *S1, double *S2a, float *feat, int32_t b_max, void *stream);

/* ── Pairwise TE (LDS histogram kernels) ─────────────────────────────── */

/* A work item: targets tgt_order[tgt_beg .. tgt_beg+tgt_cnt) against sources
 * src_order[src_beg .. src_end). */
typedef struct {
    int32_t tgt_beg, tgt_cnt, src_beg, src_end;
} tnx_work_t;

typedef struct {
    int64_t n_wave_items;   /* items for the wave-per-pair kernel */
    int64_t n_block_items;  /* items for the workgroup-per-pair kernel */
    int64_t n_pairs;        /* pairs covered (excludes target == source) */
    int32_t cap_wave;       /* histogram cells per wave */
    int32_t cap_block;      /* histogram cells per workgroup pass */
    int32_t max_block_cols; /* max b_t * b_s over block-kernel pairs */
    int32_t n_tgt, n_src;
    double est_cost_wave;   /* relative cost estimates */
    double est_cost_block;
} tnx_plan_info_t;

/* Host-only planner.  n_per_var (host, n_vars); tgt_list / src_list (host)
 * choose targets and sources (NULL = all variables).  Writes tgt_order
 * (n_tgt), src_order (n_src) and work (wave items first, then block items)
 * into host buffers; work_cap is the capacity of `work` in items.
 * cap_wave <= 0 selects the gfx950 default.  A pair (x, y) is counted by
 * the wave kernel iff b_x^2 * b_y <= cap_wave.  cap_block <= 0 sizes the
 * block kernel's histogram so that two workgroups share a CU's LDS; an
 * explicit cap_block must be a multiple of 4 and >= the widest block
 * pair's b_t * b_s (reported as max_block_cols).
 * src_skip (host, n_vars, or NULL): target v leaves out the first
 * src_skip[v] sources of src_order (sources are ordered by level count,
 * ties by list position), the pairs a one-hot plan (tnx_onehot_te) counts.
 * This is the scheduling half of te_(adaptive_)smem_launch_pairfree
 * (tenex/csrc/full_smem.cu:350-381, tenex/csrc/adaptive_smem.cu:351-376). */
TNX_API int tnx_te_plan(const int32_t *n_per_var, int64_t n_vars, int64_t L,
                const int32_t *tgt_list, int64_t n_tgt, const int32_t *src_list,
                int64_t n_src, int32_t cap_wave, int32_t cap_block,
                const int32_t *src_skip,
                int32_t *tgt_order, int32_t *src_order, tnx_work_t *work, int64_t work_cap,
                tnx_plan_info_t *info);

/* Output modes for tnx_te_run. */
#define TNX_OUT_PAIRFREE 0 /* out[id - pair_offset], id in [pair_offset, pair_offset+n_pairs_local) */
#define TNX_OUT_MATRIX 1   /* out[target * ld_out + source] (internal layout) */
#define TNX_OUT_STATS 2    /* surrogate statistics at the public slot source * ld_out + target:
                            * te = (float)TE; sum_te += te; sumsq_te += te^2 (FP64);
                            * count_ge += (te >= observed[slot]) */
#define TNX_OUT_STAGE 3    /* K surrogate copies in one launch: variable v = k * m + j
                            * (m = stage_vars) is copy k of variable j; out[k * stage_stride
                            * + (source % m) * ld_out + (target % m)] = (float)TE, for
                            * tnx_stats_accumulate to fold in copy order */

typedef struct {
    const int8_t *bins;       /* (n_vars, ld) */
    const uint16_t *states;   /* (n_vars, lds) from tnx_gene_prep */
    const int32_t *n_per_var;
    const double *S1, *S2a;   /* from tnx_gene_prep */
    const float *xlogx;       /* from tnx_xlogx_table(N = T - tau) */
    const int32_t *tgt_order, *src_order; /* device copies of the plan */
    const tnx_work_t *work;   /* device copy of the plan */
    int64_t n_vars, T, ld, lds;
    int32_t tau, out_mode;
    int64_t n_wave_items, n_block_items;
    int32_t cap_wave, cap_block, max_block_cols;
    float *out;
    int64_t ld_out, pair_offset, n_pairs_local;
    double *sum_te, *sumsq_te;  /* TNX_OUT_STATS only (accumulated) */
    int32_t *count_ge;
    const float *observed;
    int32_t hot_zero;  /* nonzero: key 0 (state 0, source level 0) is frequent
                        * (zero-inflated data); the kernels count it in
                        * per-lane words instead of one histogram cell.
                        * Results are identical either way; only speed differs. */
    int32_t stage_vars;    /* TNX_OUT_STAGE only: variables per copy (m) */
    int64_t stage_stride;  /* TNX_OUT_STAGE only: floats per copy (>= m * ld_out) */
    int32_t item_per_wave;     /* nonzero: each wave of a wave-kernel workgroup takes its own
                                * work item (suits plans of one-target items: pair lists);
                                * zero: the four waves split each item's pairs.  Results are
                                * identical either way; only speed differs. */
    /* One persistent launch (from ABI 7).  queue_items == NULL: the wave and
     * block items run as two grid-striding launches.  Otherwise one grid of
     * resident workgroups takes entries from 2 * TNX_QUEUES queues built by
     * tnx_te_queue: queue q holds queue_items[queue_off[q] .. queue_off[q +
     * 1]); queues 0 .. TNX_QUEUES - 1 hold wave entries (a wave takes one at
     * a time: entry e is wave item e / 4, its pairs e % 4, e % 4 + 4, ...,
     * or with item_per_wave wave item e whole), queues TNX_QUEUES .. hold
     * block entries (the workgroup takes one at a time: block item e).  A
     * workgroup drains its own XCD's queues first, then the others.  Needs
     * T - tau <= 12288, four cap_wave slices within 156 KiB and a block
     * histogram (the four slices) that holds max_block_cols twice plus 128
     * words; pairs give the counts of the two-launch path, and TE from
     * exact fixed-point sums (xlogx_fx below). */
    const int32_t *queue_items;
    const int32_t *queue_off;  /* device int32[2 * TNX_QUEUES + 1] */
    int32_t *queue_ctr;        /* device int32[TNX_QUEUE_CTR_INTS] scratch: the queues' take
                                * counters, cleared by tnx_te_run on the stream before the
                                * launch (a kernel, not a memset) */
    /* Test hook: when counts_map != NULL, each computed pair (x, y) with
     * r = counts_map[x * n_vars + y] >= 0 also writes its joint counts to
     * counts_dbg[r * counts_ld + ((a * b_x + b) * b_y + c)] from inside the
     * kernel that computes it (persistent launch only). */
    const int32_t *counts_map;
    int32_t *counts_dbg;
    int64_t counts_ld;
    /* The persistent launch's c*log2(c) table (tnx_xlogx_table_fx, N = T -
     * tau) and its fraction bits: per pair S3 - S2b is summed exactly in
     * 64-bit fixed point, then scaled and combined with S1 - S2a in FP64. */
    const int64_t *xlogx_fx;
    int32_t fx_shift;
    /* Per-variable series length (from ABI 8; persistent launch only):
     * var_L != NULL gives target variable v the L = var_L[v] samples
     * (1 <= var_L[v] <= T - tau; larger values are clamped) and 1 / var_L[v]
     * as its normaliser, so stacked copies prepared with different lags
     * (tnx_gene_prep per copy with its own tau, T - tau = the copy's L) run
     * in one launch; tau is then the smallest lag and xlogx_fx the table
     * for N = T - tau. */
    const int32_t *var_L;
    /* From ABI 9: cap_sort > 0 leaves every pair of more than cap_sort
     * b_t^2 b_s cells unwritten by the block role (their dense histograms
     * would scan mostly empty cells); the caller computes them with
     * tnx_te_sort_pairs.  Must exceed cap_wave.  0: every pair here. */
    int64_t cap_sort;
    /* From ABI 13, the persistent launch: hot_cw > 0 (4, 8 or 16) counts
     * the samples of each target's four hot states (x_{t+tau}, x_t) with
     * both levels below 2 (zero-inflated data put most samples there) in
     * lane-private copies: hot_cw copies per hot cell in each wave slice,
     * 16 in the block role, folded back before the reduction.  Every
     * variable needs n_per_var <= hot_maxb <= 32.  Each wave slice grows by
     * 4 * hot_cw * hot_maxb words after cap_wave (rounded up to 4).  Results
     * are identical either way; only speed differs.  0: off. */
    int32_t hot_cw;
    int32_t hot_maxb;
} tnx_te_args_t;

#define TNX_QUEUES 8            /* work queues of the persistent launch (one per XCD) */
#define TNX_QUEUE_CTR_INTS 256  /* ints of queue_ctr */

/* Host-only: the persistent launch's work queues for a plan (work, orders
 * from tnx_te_plan, all host).  Wave entries (see queue_items above) go to
 * queues 0 .. TNX_QUEUES - 1, block entries to the next TNX_QUEUES, each by
 * the source chunk of its item (the workgroups of one XCD then share the
 * source rows in their L2), chunks balanced over the queues by estimated
 * cost; inside a queue chunk by chunk, heaviest items first.  qitems
 * (capacity qcap) receives the entries, qoff (2 * TNX_QUEUES + 1 ints) the
 * queue bounds, *n_wave_entries the number of wave entries.  qitems == NULL:
 * size query (only *n_wave_entries is written).  No reference counterpart:
 * the reference launches one grid over all pairs (full_smem.cu:350-381). */
TNX_API int tnx_te_queue(const tnx_work_t *work, int64_t n_wave_items, int64_t n_block_items,
                         const int32_t *n_per_var, int64_t n_vars, const int32_t *tgt_order,
                         const int32_t *src_order, int64_t L, int32_t cap_wave,
                         int32_t item_per_wave, int32_t *qitems, int64_t qcap, int32_t *qoff,
                         int64_t *n_wave_entries);

/* Pairwise TE for every pair of the plan, (S3 + S1 - S2a - S2b) / N with
 * S = sum c log2 c: in the persistent launch S3 - S2b is an exact 64-bit
 * fixed-point sum (xlogx_fx), in the two-launch kernels FP32 terms with
 * compensated summation; the final combination is FP64.  Replaces te_smem_launch(_pairfree),
 * te_adaptive_smem_launch(_pairfree) and te_gmem_launch
 * (tenex/csrc/full_smem.cu:319-381, adaptive_smem.cu:307-376,
 * gmem.cu:190-230). */
TNX_API int tnx_te_run(const tnx_te_args_t *args, void *stream);

/* Transfer entropy of explicit [target, source] pairs by sorting their
 * occupied joint states (no dense histogram): pairs (device int32
 * [n_pairs][2]) on int16 bins (device, rows of stride ld, levels < 32768);
 * out[target * ld_out + source] (ld_out > 0) or out[p] (ld_out == 0) =
 * (2^-fx_shift (S3 - S2b) + S1[target] - S2a[target]) / (T - tau), S3 and
 * S2b exact fixed-point sums from xlogx_fx (tnx_xlogx_table_fx with N = T -
 * tau).  Counterpart of the reference's global-memory overflow kernel
 * te_gmem_launch (tenex/csrc/gmem.cu:190-230) for pairs whose b_t^2 b_s
 * dense counters dwarf the series, and the only exact path for variables
 * of more than 127 levels.  Needs T - tau <= 16384. */
TNX_API int tnx_te_sort_pairs(const int16_t *bins16, int64_t n_vars, int64_t T, int64_t ld,
                              int32_t tau, const int32_t *pairs, int64_t n_pairs,
                              const double *S1, const double *S2a, const int64_t *xlogx_fx,
                              int32_t fx_shift, float *out, int64_t ld_out, void *stream);

/* S1 and S2a (as tnx_gene_prep) of the variables vars[0..n_sel) from sorted
 * (x_t, x_{t+tau}) keys of their int16 rows: the per-variable terms of
 * variables whose levels exceed tnx_gene_prep's bound. */
TNX_API int tnx_gene_prep_sort(const int16_t *bins16, int64_t n_vars, int64_t T, int64_t ld,
                               int32_t tau, const int32_t *vars, int64_t n_sel,
                               const int64_t *xlogx_fx, int32_t fx_shift, double *S1,
                               double *S2a, void *stream);

/* TRACE OutTE / InTE for n_jobs variables (replaces fused_outte_4way_launch
 * and fused_inte_4way_launch, tenex/csrc/trace_sort_entropy.cu:191-304, and
 * the 4-entropy combination of TRACEMethod._compute_cuda,
 * tenex/inference/trace.py:913-966).  Job j: variable var_ids[j] with the d_j
 * connected variables conn[conn_off[j] .. conn_off[j + 1]) (mode 0 OutTE:
 * its network targets; mode 1 InTE: its network sources), keyed with the
 * reference's 64-bit polynomial state hash.  sums_out[4 j + e] (device
 * int64) receives sum over distinct keys of c*log2(c) in fixed point
 * (xlogx_fx from tnx_xlogx_table_fx with N = T - tau and fx_shift) for the
 * entropies e = 0..3 of the reference's order; te_out (device float, may be
 * NULL) receives max(0, (-S0 + S1 + S2 - S3) 2^-fx_shift / (T - tau)).
 * bins: device int8 rows of stride ld (a multiple of 16).  Needs
 * T - tau <= 16384 (the keys of one entropy are sorted in LDS). */
TNX_API int tnx_trace_entropy(const int8_t *bins, int64_t n_vars, int64_t T, int64_t ld,
                              int32_t tau, int32_t mode, const int32_t *var_ids,
                              const int32_t *conn, const int64_t *conn_off, int64_t n_jobs,
                              const int64_t *xlogx_fx, int32_t fx_shift, int64_t *sums_out,
                              float *te_out, void *stream);

/* tnx_te_run's persistent launch without any device memory of the caller's
 * beyond inputs and outputs: no counterpart in the reference; for callers
 * that cannot allocate (HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  args as for tnx_te_run with queue_items set and
 * queue_ctr NULL: the queue take counters live in a library-owned device
 * array of TNX_NOALLOC_SCRATCH_INTS int32 per GPU (1 KiB, held until the
 * process ends), cleared by a kernel before each launch.  Calls nothing
 * that allocates.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (concurrent streams or threads) race and give wrong
 * results; serialise them.  TNX_E_ARG without a queue or with queue_ctr
 * set, TNX_E_CAP if the queues need more counters than the array holds. */
#define TNX_NOALLOC_SCRATCH_INTS 256
TNX_API int tnx_te_run_noalloc(const tnx_te_args_t *args, void *stream);

/* Surrogate statistics from a TNX_OUT_STAGE run: for each public slot
 * slots[i] (ascending, < stage_stride) and copies k = 0..K-1 in order,
 * te = stage[k * stage_stride + slot]; sum_te[slot] += te, sumsq_te[slot] +=
 * te^2 (FP64), count_ge[slot] += (te >= observed[slot]): the same additions
 * in the same order as K TNX_OUT_STATS runs.  count_ge NULL: no counting
 * (observed may then be NULL too).  No reference counterpart: it
 * batches the per-surrogate accumulation of tenex/inference/
 * surrogate_test.py:890-929 (sums over the surrogate TE matrices). */
TNX_API int tnx_stats_accumulate(const float *stage, int32_t K, int64_t stage_stride,
                                 const int64_t *slots, int64_t n_slots, const float *observed,
                                 double *sum_te, double *sumsq_te, int32_t *count_ge,
                                 void *stream);

/* Debug: raw joint counts of explicit pairs (host-side count n_pairs,
 * pairs device int32 [target, source]) through the same accumulation code
 * as tnx_te_run; counts_out int32 (n_pairs, max_cells) in layout
 * ((a * b_t + b) * b_s + c).  Pairs with b_t^2 b_s > max_cells are an error.
 * Also writes te_out[p] (float) computed by the wave path. */
TNX_API int tnx_te_counts_debug(const int8_t *bins, const uint16_t *states, const int32_t *n_per_var,
                        const double *S1, const double *S2a, const float *xlogx,
                        int64_t n_vars, int64_t T, int64_t ld, int64_t lds, int32_t tau,
                        const int32_t *pairs, int64_t n_pairs, int32_t max_cells,
                        int32_t *counts_out, float *te_out, void *stream);

/* ── GEMM-B2: binary data on INT8 MFMA ───────────────────────────────── */

/* Workspace bytes for tnx_te_gemm_b2 at (n, T). */
TNX_API int64_t tnx_gemm_b2_workspace(int64_t n, int64_t T, int32_t tau);

/* Precision of tnx_te_gemm_b2's contractions (all exact). */
#define TNX_B2_AUTO 0  /* FP4 when L < 2^24, else INT8 */
#define TNX_B2_INT8 1  /* int8 x int8 -> int32 (v_mfma_i32_32x32x32_i8) */
#define TNX_B2_FP4 2   /* 0/1 as e2m1 nibbles -> fp32 (v_mfma_scale_f32_32x32x64_f8f6f4) */

/* All n*(n-1) TE values for binary bins (0/1) via exact MFMA contractions
 * (C = fp.p^T, D = p.p^T, E = f.p^T) and the fused 8-state epilogue.
 * te_out float (n, ld_out), internal layout te_out[i * ld_out + j] =
 * TE(j -> i), diagonal 0.  S1, S2a from tnx_gene_prep.  counts_dbg (NULL
 * or int32 (3, n, n)): C, D, E.  Replaces compute_te_gemm_b2
 * (tenex/kernels/gemm_b2.py:307-425, :66-187, :210-285). */
TNX_API int tnx_te_gemm_b2(const int8_t *bins, int64_t n, int64_t T, int64_t ld, int32_t tau,
                   const double *S1, const double *S2a, const float *xlogx,
                   void *workspace, int64_t workspace_bytes, float *te_out, int64_t ld_out,
                   int32_t *counts_dbg, int32_t precision, void *stream);

/* ── One-hot INT8 MFMA counts for pairs of few levels ───────────────── */

/* The joint counts of pair (x, y) are a product of one-hot indicator rows:
 * target x contributes b_x^2 - 1 rows over t < L (states (a, b) =
 * (x[t+tau], x[t]) with a < b_x - 1 at row a*b_x + b, then x_t = b for
 * b < b_x - 1 at row (b_x - 1)*b_x + b), source y its b_y rows y[t] = c
 * (with src_ie, b_y - 1 rows).  Counts with a = b_x - 1, of x_t = b_x - 1
 * (and with src_ie of c = b_y - 1) follow from the per-variable
 * marginals, so every joint count is an exact integer.
 * Rows are grouped in 128-row (target) and 128-column (source) groups that
 * never split a variable; a tile is two row groups x two column groups
 * (one 128 x 128 sub-tile per wave, mask bit w = 2*wr + wc).
 * Operands are 0/1 indicators as FP4 (e2m1, 1.0 = 0x2) in 1 KB blocks of
 * 32 rows x 64 samples in MFMA lane order (block (row / 32) * (Kp / 64) +
 * k / 64; inside it row r, sample k at byte ((k % 64) / 32 * 32 + r % 32)
 * * 16 + (k % 32) / 2, even samples in the low nibble), contracted by
 * v_mfma_scale_f32_32x32x64_f8f6f4 with FP32 sums (exact: every count is
 * below 2^24).  Each pair's TE is written
 * through the tnx_te_args_t output modes (out_mode, out, ld_out, ...),
 * with S3 - S2b summed in the exact fixed point of xlogx_fx.  No reference
 * counterpart: the reference counts every pair with per-sample atomics
 * (tenex/csrc/full_smem.cu:31-142, tenex/csrc/adaptive_smem.cu:30-149). */
typedef struct {
    int8_t *A, *B;                 /* operands (rows_a x Kp / 2, rows_b x Kp / 2 bytes), tiled */
    int64_t Kp;                    /* padded series length: multiple of 128, >= L (L <= 65535) */
    int64_t rows_a, rows_b;        /* multiples of 128 */
    const int32_t *arow;           /* (rows_a, 2): variable (-1: padding), (a << 8) | b
                                    * (a = 255: an x_t row) */
    const int32_t *brow;           /* (rows_b, 2): variable (-1: padding), level c */
    int32_t *marg;                 /* per marginal variable: n(a, b) (b^2), n(c) (b) */
    const int64_t *moff;           /* (n_vars): offset of the variable's marginals, -1: none */
    const int32_t *tiles;          /* (n_tiles, 8): rg0, rg1, cg0, cg1, mask, 0, 0, 0 */
    int64_t n_tiles;
    const int32_t *rg_off;         /* (n_rg + 1): targets of row group g: rg_off[g] .. */
    const int32_t *t_var, *t_row;  /* target variable, first row inside its group */
    const int32_t *cg_off;         /* (n_cg + 1) */
    const int32_t *s_var, *s_col;  /* source variable, first column inside its group */
    const int32_t *mvars;          /* (n_mvars): variables whose marginals are counted */
    int64_t n_mvars;
    int32_t src_ie;                /* 1: a source of b_y >= 2 levels has b_y - 1 rows (levels
                                    * 0 .. b_y - 2); its last level's counts follow from the
                                    * target's n(a, b) and n(b) */
    int32_t pad0;
} tnx_onehot_args_t;

/* Marginals of mvars and the A, B operands from bins (int8, K = 1) for the
 * series length L = T - tau (o->Kp >= L; b <= 64 for marginal variables). */
TNX_API int tnx_onehot_pack(const tnx_onehot_args_t *o, const int8_t *bins, int64_t T,
                            int64_t ld, int32_t tau, const int32_t *n_per_var, void *stream);

/* TE of every pair of the tiles (target != source) through te's output mode;
 * te supplies n_per_var, S1, S2a, xlogx_fx, fx_shift, T, tau and the
 * output fields (the histogram plan fields are not read). */
TNX_API int tnx_onehot_te(const tnx_onehot_args_t *o, const tnx_te_args_t *te, void *stream);

/* ── Surrogate test ──────────────────────────────────────────────────── */

/* One block-shuffled copy of the bins: out[v, t] = bins[v, perm[v, t / bl] * bl
 * + t % bl] for t < n_blocks * bl, the tail [n_blocks * bl, T) and the row
 * padding copied as they are.  perm: int32 (n_vars, n_blocks), one surrogate.
 * Replaces _block_shuffle_gpu / _random_shuffle_gpu
 * (tenex/inference/surrogate_test.py:107-162). */
TNX_API int tnx_block_shuffle(const int8_t *bins, int64_t n_vars, int64_t T, int64_t ld,
                              const int32_t *perm, int32_t n_blocks, int32_t block_length,
                              int8_t *out, void *stream);

/* K block-shuffled copies stacked: out row k * n_vars + v (ld bytes) is row
 * v of bins shuffled with the block order in perm row k * n_vars + v (perm
 * int32 (n_copies * n_vars, n_blocks)); n_vars * n_copies < 2^31.
 * tnx_block_shuffle is n_copies = 1.  No reference counterpart: the batched
 * form of _block_shuffle_gpu (tenex/inference/surrogate_test.py:107-148)
 * for the surrogate copies that go through one TE launch. */
TNX_API int tnx_block_shuffle_copies(const int8_t *bins, int64_t n_vars, int32_t n_copies,
                                     int64_t T, int64_t ld, const int32_t *perm,
                                     int32_t n_blocks, int32_t block_length, int8_t *out,
                                     void *stream);

typedef struct {
    const int8_t *bins;       /* (n_vars, ld) */
    const int32_t *n_per_var;
    const float *xlogx;       /* N = T - tau */
    const int32_t *perm;      /* (n_surr, n_vars, n_blocks) block order per surrogate */
    const int32_t *tgt_order, *src_order;
    const tnx_work_t *work;   /* plan from tnx_te_plan (same cap values) */
    const float *observed;    /* public layout M[src * n_vars + tgt] */
    int64_t n_vars, T, ld;
    int32_t tau, n_surr, n_blocks, block_length;
    int64_t n_wave_items, n_block_items;
    int32_t cap_wave, cap_block, max_block_cols;
    double *sum_te, *sumsq_te; /* (n_vars, n_vars) public layout, accumulated (+=) */
    int32_t *count_ge;         /* (n_vars, n_vars) public layout, accumulated (+=) */
    float *te_k_dbg;           /* optional (n_vars*n_vars, n_surr): te_k per slot */
    int32_t max_levels;        /* max n_per_var over the plan's targets if known, else 0;
                                * <= 16 lets the kernel keep shuffled states as bytes */
    int32_t max_tgt_cnt;       /* max tgt_cnt over the work items (0 = unknown: 4) */
} tnx_surr_args_t;

/* Block-shuffle surrogate statistics for every pair of the plan: for each
 * surrogate k, each variable's head [0, n_blocks*block_length) is read in
 * block order perm[k, v, :], the tail untouched; TE per (pair, k) with the
 * same arithmetic as tnx_te_run; FP64 sum / sum of squares and the count of
 * te_k >= observed (float32 compare).  Replaces
 * full_smem_surrogate_test_launch / adaptive_smem_surrogate_test_launch
 * (tenex/csrc/full_smem_surrogate_test.cu:297-394,
 * tenex/csrc/adaptive_smem_surrogate_test.cu:306-400). */
TNX_API int tnx_surrogate_run(const tnx_surr_args_t *args, void *stream);

/* ── FDR network: DPI trimming ───────────────────────────────────────── */

/* Max-min product for the data-processing-inequality trim of the FDR
 * network (replaces the mediator loop of _trim, tenex/inference/fdr.py:61-133,
 * :104-110): with A_ik = A[rows, mediators] (m x k), B_kn = A[mediators, :]
 * (k x n) and E = A[rows, :] (m x n), all float32 row-major,
 *     out[i,j] = E[i,j] if E[i,j] >= max(0, max_l min(A_ik[i,l], B_kn[l,j]))
 *                           + trim_threshold, else 0.
 * Exact (comparisons only). */
TNX_API int tnx_dpi_maxmin(const float *A_ik, int64_t m, int64_t k, int64_t lda,
                           const float *B_kn, int64_t n, int64_t ldb, const float *E,
                           int64_t lde, float trim_threshold, float *out, int64_t ldo,
                           void *stream);

/* ── Learned null: 3-layer MLP (no reference counterpart) ────────────── */

/* Network: h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3.
 * Weights bf16 row-major W1 [H x D], W2 [H x H], W3 [O x H]; biases fp32.
 * D <= 64, H == 64, O <= 4.  Accumulation in fp32. */
typedef struct {
    const uint16_t *W1, *W2, *W3; /* bf16 */
    const float *b1, *b2, *b3;
    int32_t D, H, O;
} tnx_mlp_t;

/* Y[r, o] for rows r < n_rows of X (float, (n_rows, D)). */
TNX_API int tnx_mlp_forward(const tnx_mlp_t *m, const float *X, int64_t n_rows, float *Y,
                    void *stream);

/* Bytes of device workspace tnx_mlp_backward / tnx_mlp_fit need for
 * n_rows training rows: one partial gradient per workgroup (at most 256
 * workgroups of 64-row chunks), summed in a fixed order. */
TNX_API int64_t tnx_mlp_workspace_bytes(int32_t D, int32_t O, int64_t n_rows);

/* Loss 0.5 * w_o * (y - tgt)^2 averaged over rows and outputs; gradients
 * (fp32, params layout W1, b1, W2, b2, W3, b3) written to grad and the loss
 * to loss_out (device float; may be NULL).  Deterministic: per-workgroup
 * partials in `workspace` summed in workgroup order. */
TNX_API int tnx_mlp_backward(const tnx_mlp_t *m, const float *X, const float *tgt,
                             const float *w_out, int64_t n_rows, float *grad, float *loss_out,
                             void *workspace, int64_t ws_bytes, void *stream);

/* Adam step on fp32 master weights `master` (params layout), refreshing the
 * bf16 copies in m; moments m1, m2 fp32; step t >= 1. */
TNX_API int tnx_mlp_adam(const tnx_mlp_t *m, float *master, const float *grad, float *m1,
                 float *m2, int64_t n_params, float lr, float beta1, float beta2,
                 float eps, int32_t t, void *stream);

/* Full-batch Adam training schedule for tnx_mlp_fit: step s in [0, steps)
 * uses lr_min + (lr - lr_min) * 0.5 * (1 + cos(pi * s / steps)) and Adam
 * step number t0 + s + 1. */
typedef struct {
    float lr, lr_min, beta1, beta2, eps;
    int32_t steps, t0;
    int32_t batch;  /* rows per step: 0 (or >= n_rows) full batch, else step s
                     * takes rows [(s mod nb) * batch, +batch), nb = n_rows / batch,
                     * of X and tgt, whose rows the caller has shuffled */
} tnx_adam_t;

/* `opt->steps` steps (backward + Adam) launched from the host loop with no
 * host synchronisation; loss_hist (device float[steps], may be NULL)
 * receives each step's loss.  grad: device float[n_params] scratch; the
 * workspace is sized for the rows of one step (tnx_mlp_workspace_bytes). */
TNX_API int tnx_mlp_fit(const tnx_mlp_t *m, float *master, float *m1, float *m2,
                        int64_t n_params, const float *X, const float *tgt, int64_t n_rows,
                        const tnx_adam_t *opt, float *grad, void *workspace, int64_t ws_bytes,
                        float *loss_hist, void *stream);

/* ── Hardware ceilings (roofline peaks) ────────────────────────────────── */

/* Chip-wide non-returning LDS atomic increments: n_blocks workgroups of 256
 * threads, each thread doing iters * 8 increments; mode 0 conflict-free,
 * mode 1 random keys in 2048 cells.  out: uint32 (n_blocks,) sink.
 * No reference counterpart (the reference states no LDS ceiling). */
TNX_API int tnx_bench_lds_atomics(int32_t mode, int64_t n_blocks, int32_t iters, uint32_t *out,
                                  void *stream);

/* Matrix-core throughput from registers (no memory in the loop): mode 0
 * FP4 (v_mfma_scale_f32_32x32x64_f8f6f4), 1 FP8, 2 INT8
 * (v_mfma_i32_32x32x32_i8); n_blocks workgroups of four waves, each wave
 * iters x 4 MFMAs; out: n_blocks x 256 floats (device).  The rate is the
 * roofline peak the MFMA kernels are measured against, next to the
 * published figure.  No reference counterpart. */
TNX_API int tnx_bench_mfma(int32_t mode, int64_t n_blocks, int32_t iters, float *out,
                           void *stream);

#ifdef __cplusplus
}
#endif
#endif /* TENEX_HIP_H */
