/*
 * flashsdkde.h - C ABI of the MI355X (gfx950) pairwise Gaussian-kernel hot path.
 *
 * Conventions shared by every entry point:
 *   - All array pointers are DEVICE pointers, owned by the caller. Point sets
 *     are row-major float32 (rows, 16) with an explicit row stride `ld`
 *     (in elements, ld >= 16, ld % 4 == 0, base 16-byte aligned); 1-D sets are
 *     contiguous float32.
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *     Every call only enqueues work on that stream; nothing synchronizes.
 *   - No entry point allocates device memory. Scratch comes from the caller
 *     through (ws, ws_bytes); fsk_workspace_bytes() gives the size needed.
 *   - Return value: FSK_OK (0) or a negative FSK_E* code. fsk_last_error()
 *     returns a thread-local message describing the last failure.
 *   - Accumulator outputs (pdf sums, weighted sums) are fully overwritten:
 *     callers never pre-zero them (the reference zero-fills with torch first,
 *     kernels/emp_score_16d_symmetric_atomic.py:189-190).
 *   - The score, debias and weighted-combine entry points take at most 2^27
 *     rows (FSK_EUNSUPPORTED above); the KDE entry points have no such limit.
 *   - Inputs must be finite. Device code clamps exponents with min/max that
 *     ignore NaN, so a NaN or inf coordinate is not propagated: it yields
 *     finite but meaningless sums. The entry points do not scan their inputs
 *     (that would need a device-to-host read per call); the Python layer
 *     refuses non-finite points with ValueError before launching.
 *
 * The reference (Flash-SD-KDE, Python + Triton) has no C interface; each entry
 * point below replaces the Python launch wrapper cited next to it, which is
 * where the reference crosses from host to device.
 */
#ifndef FLASHSDKDE_H
#define FLASHSDKDE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define FSK_OK 0
#define FSK_EINVAL -1       /* bad argument (shape, stride, bandwidth, pointer) */
#define FSK_EUNSUPPORTED -2 /* flag combination or size not implemented */
#define FSK_ELAUNCH -3      /* HIP launch / runtime error */
#define FSK_EWORKSPACE -4   /* workspace missing or too small */

/* ---- flags: low byte = cross-term precision ---- */
#define FSK_PREC_F32 0u      /* FP32 MFMA (v_mfma_f32_16x16x4_f32): the plain FP32 form */
#define FSK_PREC_F32_VALU 1u /* FP32 FMA on the vector ALU */
#define FSK_PREC_BF16 2u     /* BF16 MFMA, FP32 accumulate */
#define FSK_PREC_F16 3u      /* FP16 MFMA, FP32 accumulate */
#define FSK_PREC_BF16X3 4u   /* BF16 hi/lo split, 3 products, FP32 accumulate */
#define FSK_PREC_BF16X6 5u   /* BF16 three-piece split, 7 products (all terms above
                              * FP32 rounding), FP32 accumulate: FP32-class cross term */
#define FSK_PREC_F16X3 6u    /* FP16 hi/lo split cross term, 3 products (2^-22 terms
                              * dropped), FP32 accumulate; phi.X as in BF16X3 */
#define FSK_PREC_F16X3S 7u   /* cross term as F16X3; phi.X with a single FP16 phi
                              * against FP16 [Xh | Xl] (one MFMA, no phi split):
                              * score forms only (log-densities as F16X3) */
#define FSK_PREC_BF16X3S 8u  /* cross term as BF16X3 (BF16 range); phi.X as F16X3S
                              * (FP16 x - mu pieces: |x - mu| < 6e4); log-densities
                              * as BF16X3 */
/* The FP16 forms (_F16, _F16X3, _F16X3S) hold -2c (q - mu) and c |x - mu|^2, with
 * c = log2(e) / (2 h^2) and mu the data mean, as FP16 pieces: the caller
 * keeps c * max |x - mu|^2 within FP16's range (the Python layer refuses
 * inputs above 3e4). */
#define FSK_PREC_MASK 0xffu
#define FSK_FLAG_LAPLACIAN (1u << 8) /* phi * (1 + d/2 - |q-x|^2/(2h^2)) */
#define FSK_FLAG_PHI_S (1u << 9)     /* phi * |q-x|^2/h^2 alone: the Laplacian correction's
                                      * moment, for the two-pass (non-fused) form */
#define FSK_FLAG_ROWS_IN_DATA (1u << 11) /* fsk_score16_rows: every row is also one of the
                                          * data points, so each row's sums hold its self
                                          * term (lets F16X3S / BF16X3S leave out negligible
                                          * tile pairs as for rows = data) */
#define FSK_FLAG_DENSE (1u << 10)    /* evaluate every term: no negligible tile pairs left
                                      * out in the score forms whose phi.X takes phi rounded
                                      * to FP16 (F16, F16X3S, BF16X3S; fsk_score16_ordered,
                                      * fsk_score16_rows) or in the 16-bit log-densities
                                      * (fsk_kde16_logdensity) */

/* ---- operation ids for fsk_workspace_bytes ---- */
#define FSK_OP_KDE16_SPLITK 1
#define FSK_OP_KDE16_ATOMIC 2
#define FSK_OP_KDE1_SPLITK 3
#define FSK_OP_KDE1_ATOMIC 4
#define FSK_OP_SCORE16_SYM 5
#define FSK_OP_SCORE16_ORDERED 6
#define FSK_OP_KDE16_LOGDENSITY 7
#define FSK_OP_KDE1_SCORE 8
#define FSK_OP_LOO_GRID 9 /* profiling id only; workspace: fsk_loo_workspace_bytes */
#define FSK_OP_SCORE16_ROWS 10
#define FSK_OP_KDE_ND_SPLITK 11    /* workspace: fsk_nd_workspace_bytes */
#define FSK_OP_SCORE_ND_ORDERED 12 /* workspace: fsk_nd_workspace_bytes */
#define FSK_OP_SPATIAL_CODES16 13

const char *fsk_last_error(void);
int fsk_version(void);

/* Measurement hook (no reference counterpart): when enabled, each entry point
 * records a pair of HIP events on the caller's stream around its main pair
 * kernel (not the small prep / combine kernels). fsk_profile_enable resets the
 * counts; fsk_profile_last_ms waits for the recorded launches of `op` (the
 * last 64 at most) and returns their mean duration in ms (negative if none).
 * Single-threaded use; events belong to the current device. */
int fsk_profile_enable(int on);
double fsk_profile_last_ms(int op);

/* Issue-rate microbenchmark (no reference counterpart; used to measure the
 * roofline the pair kernels are judged against). which: 0 FP32 MFMA flop/s,
 * 1 BF16 MFMA flop/s, 2 v_exp_f32 per s, 3 v_fma_f32 per s, 4 seconds of the
 * MFMA+exp mix kernel, 6 seconds of the same mix with FMAs in place of the
 * exps, 7 / 8 seconds of 4 BF16 MFMAs with 16 exps / 16 FMAs per iteration,
 * 9 seconds of 8 exps with 8 FMAs per iteration,
 * 5 HBM copy bytes/s (read + write): the first half of
 * `out` copied to its second half by `iters` launches of a 4096-workgroup
 * copy kernel (blocks*512 bytes read and written per launch). `blocks`
 * workgroups of 256 threads, `iters` loop trips; `out` is a 16-byte aligned
 * device buffer of blocks*256 floats. Synchronizes the stream. */
int fsk_microbench(int which, int blocks, int iters, float *out, double *result, void *stream);

/* Bytes of scratch the operation needs for the given sizes (0 if none). */
size_t fsk_workspace_bytes(int op, int64_t n, int64_t m, unsigned flags);

/* 16-D Gaussian KDE sums, deterministic split-K.
 * out[i] = sum_j exp(-|q_i - x_j|^2 / (2 h^2))            (no normalisation)
 * (with FSK_FLAG_LAPLACIAN each term is multiplied by 1 + 8 - |q-x|^2/(2h^2); with
 * FSK_FLAG_PHI_S by |q-x|^2/h^2).
 * Per-split FP32 partials are combined in FP64.
 * Replaces kde_eval_16d_splitk, kernels/kde_eval_16d_stream_splitk.py:263-370
 * (pass A :14-81 and reduce_pdf_partials, kernels/reduce_partials.py:61-79). */
int fsk_kde16_splitk(const float *data, int64_t ld, int64_t n,
                     const float *query, int64_t ldq, int64_t m,
                     double bandwidth, unsigned flags,
                     void *ws, size_t ws_bytes, float *out, void *stream);

/* Same sums, one pass, per-tile atomics into out (zeroed by the call itself).
 * Replaces kde_eval_16d_atomic, kernels/kde_eval_16d_stream_splitk.py:373-452,
 * and gaussian_kde_triton_nd's launch, kernels/flash_sd_kde/kde.py:242-331. */
int fsk_kde16_atomic(const float *data, int64_t ld, int64_t n,
                     const float *query, int64_t ldq, int64_t m,
                     double bandwidth, unsigned flags,
                     void *ws, size_t ws_bytes, float *out, void *stream);

/* ---- any dimension d <= FSK_ND_MAX_DIM ----
 * Points are (rows, d) float32 with row stride ld >= d (no alignment
 * requirement). The library centres them on the data mean into a copy padded
 * with zero coordinates to 16 ceil(d/16) features (distances, hence every
 * kernel value, are unchanged) and runs the FP32 MFMA pair kernels over all
 * 16-wide slices at once. The reference fixes d = 16 (globals.py:25,
 * flash_sd_kde/kde.py:120); its paper describes other dimensions as padding
 * to the kernel width (paper/rebuttal.md:253,505), which is what this is. */
#define FSK_ND_MAX_DIM 64

/* Workspace of FSK_OP_KDE_ND_SPLITK (m queries) or FSK_OP_SCORE_ND_ORDERED. */
size_t fsk_nd_workspace_bytes(int op, int64_t n, int64_t m, int d);

/* KDE sums out[i] = sum_j exp(-|q_i - x_j|^2 / (2 h^2)) in d dimensions
 * (FSK_FLAG_LAPLACIAN: times 1 + d/2 - |q-x|^2/(2h^2)); split-K, FP64 combine.
 * Generalises fsk_kde16_splitk (kernels/kde_eval_16d_stream_splitk.py:263-370). */
int fsk_kde_nd_splitk(const float *data, int64_t ld, int64_t n, const float *query,
                      int64_t ldq, int64_t m, int d, double bandwidth, unsigned flags,
                      void *ws, size_t ws_bytes, float *out, void *stream);

/* Empirical-score accumulators in d dimensions: pdf[i] = sum_j phi_ij (j = i
 * included), w[i, f] = sum_j phi_ij x_j[f] written at w + i ldw_row + f ldw_col.
 * Deterministic split-K; generalises fsk_score16_ordered
 * (kernels/emp_score_16d_ordered_splitk.py:162-263). flags must be 0 (FP32). */
int fsk_score_nd_ordered(const float *data, int64_t ld, int64_t n, int d, double bandwidth,
                         unsigned flags, void *ws, size_t ws_bytes, float *pdf, float *w,
                         int64_t ldw_row, int64_t ldw_col, void *stream);

/* FP64 debias x + (h^2/2) (w/(pdf + eps) - x)/h^2 in d dimensions (as
 * fsk_sdkde_debias_f64; flash_sd_kde/kde.py:319-322). */
int fsk_sdkde_debias_nd_f64(const float *x, int64_t ldx, int64_t n, int d, const float *pdf,
                            const float *w, int64_t ldw_row, int64_t ldw_col, double bandwidth,
                            double eps, float *out, int64_t ldo, void *stream);

/* Size of the static device scratch behind fsk_kde16_splitk_noalloc (per GPU). */
#define FSK_NOALLOC_SCRATCH_BYTES (32u << 20)

/* fsk_kde16_splitk without a workspace argument: the same kernels and results
 * (bitwise), with their scratch in a file-scope device array of
 * FSK_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends. The
 * call allocates nothing (no hipMalloc / hipFree), for callers that cannot
 * allocate: code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation. No counterpart in the reference.
 *  - Sizes whose workspace (fsk_workspace_bytes(FSK_OP_KDE16_SPLITK, n, m, 0))
 *    exceeds FSK_NOALLOC_SCRATCH_BYTES return FSK_EUNSUPPORTED (e.g. 200 000
 *    points with 65 536 queries fit; 2^20 points do not).
 *  - The scratch of the current device is used (the stream must belong to it);
 *    its address is looked up on the first call on each device.
 *  - All calls on one GPU share that array: calls that overlap in time on one
 *    GPU (other streams, other host threads) race. Serialise them. */
int fsk_kde16_splitk_noalloc(const float *data, int64_t ld, int64_t n,
                             const float *query, int64_t ldq, int64_t m,
                             double bandwidth, unsigned flags, float *out, void *stream);

/* 16-D log-density by log-sum-exp (precision form):
 * out[i] = log( 1/(n (2 pi)^8 h^16) * sum_j exp(-|q_i - x_j|^2/(2h^2)) ),
 * computed as per-split (max, sum) pairs in base 2 and an FP64 combine.
 * Data are centred on the data mean before the cross term.
 * The 16-bit forms leave out negligible terms from n = 16384 on: a pass over
 * every (n / 16384)-th data point bounds each query's sum from below, and
 * terms below 2^-(ceil(log2 n) + 28) of that bound are not evaluated (pairs
 * of data tiles a box bound proves far are not read; other tiles whose every
 * term is below the wave's thresholds add nothing), so a sum loses at most
 * 2^-28 of itself. Data and queries in spatial order let most blocks skip.
 * FSK_FLAG_DENSE evaluates every term.
 * No reference counterpart (the reference returns log(density + eps),
 * flash_sd_kde/estimator.py:114); FP64 semantics follow
 * experiments/error_suite_16d/reference.py:32-82. */
int fsk_kde16_logdensity(const float *data, int64_t ld, int64_t n,
                         const float *query, int64_t ldq, int64_t m,
                         double bandwidth, unsigned flags,
                         void *ws, size_t ws_bytes, double *out, void *stream);

/* 1-D Gaussian KDE sums, split-K with FP64 combine.
 * out[i] = sum_j exp(-(q_i - x_j)^2 / (2 h^2))  (FSK_FLAG_LAPLACIAN: times 1.5 - (q-x)^2/(2h^2)).
 * All 1-D entry points sort the scaled data and the queries on the device
 * (radix sort in the workspace) and let each block of sorted queries visit
 * only the data within the distance beyond which every term is exactly 0 in
 * FP32. The plain sums (no flag) and fsk_kde1_score reach less far: a
 * query's terms more than 2^-(ceil(log2 n) + 32) below its largest (its
 * nearest data point's) are left out, at most 2^-32 of its sum. n and m are
 * at most 2^31 - 1, and their workspace size is queried
 * from the current device (fsk_workspace_bytes returns 0 without one). The
 * sorts record memset nodes, so the 1-D entry points refuse HIP graph
 * capture (FSK_EUNSUPPORTED); every other entry point is capturable.
 * Replaces kde_eval_1d_splitk, kernels/kde_eval_1d_splitk.py:141-219. */
int fsk_kde1_splitk(const float *data, int64_t n, const float *query, int64_t m,
                    double bandwidth, unsigned flags,
                    void *ws, size_t ws_bytes, float *out, void *stream);

/* 1-D sums with per-tile atomics (out zeroed by the call).
 * Replaces kde_eval_1d_atomic, kernels/kde_eval_1d_splitk.py:222-256, and the
 * launch in gaussian_kde_triton, kernels/flash_sd_kde/kde.py:147-225. */
int fsk_kde1_atomic(const float *data, int64_t n, const float *query, int64_t m,
                    double bandwidth, unsigned flags,
                    void *ws, size_t ws_bytes, float *out, void *stream);

/* 1-D KDE density-derivative sums: pdf_sum[i] = sum_j phi(z_ij),
 * deriv_sum[i] = sum_j (-z_ij) phi(z_ij), z = (q - x)/h, phi = exp(-z^2/2).
 * Replaces the launch in gaussian_kde_score_triton,
 * kernels/flash_sd_kde/kde.py:348-438 (kernel :59-100). */
int fsk_kde1_score(const float *data, int64_t n, const float *query, int64_t m,
                   double bandwidth, void *ws, size_t ws_bytes,
                   float *pdf_sum, float *deriv_sum, void *stream);

/* Empirical-score accumulators over all ordered pairs (j = i included):
 *   pdf[i] = sum_j exp(-|x_i - x_j|^2/(2h^2)),  w[i,:] = sum_j exp(...) x_j.
 * w is written at w[i*ldw_row + f*ldw_col].
 * Symmetric variant: each unordered tile pair is evaluated once and scattered
 * to both sides with atomics (non-deterministic summation order).
 * Replaces emp_score_16d_symmetric_atomic,
 * kernels/emp_score_16d_symmetric_atomic.py:163-234 (kernel :10-103), and the
 * default flash backend emp_score_16d_flash_sd_kde,
 * kernels/flash_sd_kde/empirical.py:134-223. */
int fsk_score16_sym_atomic(const float *data, int64_t ld, int64_t n,
                           double bandwidth, unsigned flags,
                           void *ws, size_t ws_bytes,
                           float *pdf, float *w, int64_t ldw_row, int64_t ldw_col,
                           void *stream);

/* Same accumulators, deterministic: bounded split-K over the data with
 * FP32 partials and an FP64 combine; bitwise reproducible for fixed n (and
 * fixed FSK_FLAG_DENSE).
 * Points are centred on their mean (FP64, fixed summation order) before the
 * pair terms are formed, and mu * pdf is added back in FP64.
 * flags may select a 16-bit cross term and phi.X (FSK_PREC_F16X3 / _BF16X6 /
 * _BF16X3 / _BF16 / _F16; the precision form); FSK_PREC_F32 is the plain
 * FP32 form (FP32 MFMA; rows whose norm terms span a narrow range are
 * evaluated as exp2(d_ij) * exp2(c|x_i - mu|^2 + K), which needs no per-pair
 * add, the others directly).
 * Negligible tiles: in the forms whose phi.X takes phi rounded to FP16, a
 * wave leaves out a row tile x tile pair (16 x 32 pairs) whose every exponent
 * is at or below a threshold: 2^-26 for F16 (those phi are exactly 0 in FP16
 * and the row sums add the same FP16 phi, so the results are bitwise the
 * dense ones); 2^-(ceil(log2 n) + 28) for F16X3S / BF16X3S (the centred
 * sums of phi.X bitwise; the FP32 row sums of the unrounded phi lose at most
 * 2^-28 of a row's self term, and w = centred sums + mu pdf carries that). Data in spatial order (nearby points in nearby rows) let most tile
 * pairs of a large set skip. FSK_FLAG_DENSE evaluates every tile.
 * Replaces emp_score_16d_ordered_splitk,
 * kernels/emp_score_16d_ordered_splitk.py:162-263 (pass A :14-92, reduce
 * kernels/reduce_partials.py:82-108). */
int fsk_score16_ordered(const float *data, int64_t ld, int64_t n,
                        double bandwidth, unsigned flags,
                        void *ws, size_t ws_bytes,
                        float *pdf, float *w, int64_t ldw_row, int64_t ldw_col,
                        void *stream);

/* Empirical-score accumulators of m query rows under the KDE of n data points:
 *   pdf[i] = sum_j exp(-|r_i - x_j|^2/(2h^2)),  w[i,:] = sum_j exp(...) x_j,
 * deterministic split-K with FP64 combine (workspace FSK_OP_SCORE16_ROWS with
 * (n, m, flags)). flags takes the precisions of fsk_score16_ordered: FSK_PREC_F32
 * or a 16-bit MFMA form (BF16, F16, BF16X3, BF16X6, F16X3, F16X3S, BF16X3S; the FP16
 * forms carry the range limit stated there, here over rows and data alike), and
 * FSK_FLAG_DENSE. Tiles are left out as in fsk_score16_ordered for F16 (bitwise);
 * F16X3S / BF16X3S leave them out only when rows is data (same pointer, stride
 * and count) or the caller sets FSK_FLAG_ROWS_IN_DATA (every row is one of the
 * data points), where every row holds its self term. With rows = data this is
 * fsk_score16_ordered; a row subset gives exact score labels for the learned
 * surrogate. The reference evaluates the score at separate points only in 1-D
 * (kde_score_eval, flash_sd_kde/reference.py:83-120). */
int fsk_score16_rows(const float *rows, int64_t ldr, int64_t m, const float *data, int64_t ld,
                     int64_t n, double bandwidth, unsigned flags, void *ws, size_t ws_bytes,
                     float *pdf, float *w, int64_t ldw_row, int64_t ldw_col, void *stream);

/* Morton (Z-order) codes of n 16-D points (no counterpart in the reference;
 * it puts data and queries in spatial order for the kernels above that leave
 * out negligible terms, without a host synchronization): the coordinates
 * along the points' three leading principal axes (block power iteration on
 * the FP64 covariance of a strided sample of at most 4096 centred points),
 * each quantized to `bits` bits (1..10) over the mean +- 4 standard
 * deviations along the axis (points beyond fall into the end cells); bit b
 * of the coordinate along the k-th leading axis goes to code bit 3b + 2 - k.
 * codes: n values (DEVICE). Workspace: fsk_workspace_bytes(
 * FSK_OP_SPATIAL_CODES16, n, 0, 0). */
int fsk_spatial_codes16(const float *x, int64_t ld, int64_t n, int bits, uint32_t *codes,
                        void *workspace, size_t workspace_bytes, void *stream);

/* FP64 split-K combine: out[i] = (float) sum_s (double) partials[s*ld_split + i].
 * Replaces reduce_pdf_partials, kernels/reduce_partials.py:61-79. */
int fsk_reduce_partials_f64(const float *partials, int64_t n_splits, int64_t m,
                            int64_t ld_split, float *out, void *stream);

/* out[i*ldo_row + f*ldo_col] = (float) sum_s (double) partials[s*ld_split + i*16 + f].
 * Replaces reduce_weighted_partials, kernels/reduce_partials.py:82-108. */
int fsk_reduce_weighted_partials_f64(const float *partials, int64_t n_splits,
                                     int64_t m, int64_t ld_split, float *out,
                                     int64_t ldo_row, int64_t ldo_col, void *stream);

/* One-step SD-KDE debias, in FP64:
 *   score = (w/(pdf + eps) - x) / h^2 ;  out = x + (h^2/2) score
 * out has row stride ldo; score_out (nullable) is (n,16) contiguous.
 * Replaces the torch epilogue of emp_sd_kde_fit_transform,
 * flash_sd_kde/kde.py:316-322, and of empirical_sd_kde_triton_nd,
 * kernels/flash_sd_kde/empirical.py:280-283. */
int fsk_sdkde_debias_f64(const float *x, int64_t ldx, int64_t n,
                         const float *pdf, const float *w, int64_t ldw_row,
                         int64_t ldw_col, double bandwidth, double eps,
                         float *out, int64_t ldo, float *score_out, void *stream);

/* Batched leave-one-out KDE log-likelihood over a bandwidth grid (no reference
 * counterpart; the batched pattern). Datasets are concatenated row-major in
 * `data` (row stride ld); `offsets` (HOST array, n_sets + 1 entries, offsets[0]
 * = 0) delimit them; `bandwidths` (DEVICE, float64, n_sets x n_bw, row-major)
 * holds each dataset's grid. out (DEVICE, float64, n_sets x n_bw):
 *   out[d, b] = 1/n_d sum_i log( 1/(n_d - 1) sum_{j != i} N(x_i; x_j, h_b^2 I) ).
 * One persistent launch over all datasets (work queue, largest first). The
 * tunables, chosen by the Python autotuner: waves_per_wg in {4, 8}, row_tiles
 * in {1, 2} (16 or 32 rows per wave: the task tile is 16 row_tiles
 * waves_per_wg rows) and wg_per_cu in 1..8. Terms below 2^-(ceil(log2 n_d) + 28) of
 * a row's nearest-neighbour term are skipped per data tile and group of 4
 * bandwidths (their sum is below 2^-28 of the row's sum). exps_done (DEVICE, nullable): the
 * count of exponentials evaluated is ADDED to it (the caller zeroes it).
 * Limits: n_sets <= 64, n_bw <= 32, every n_d >= 2.
 * Capturable: no allocation, no host synchronization, fills by kernel. */
size_t fsk_loo_workspace_bytes(const int64_t *offsets, int n_sets, int n_bw);
int fsk_loo_bandwidth_grid(const float *data, int64_t ld, const int64_t *offsets, int n_sets,
                           const double *bandwidths, int n_bw, int waves_per_wg, int row_tiles,
                           int wg_per_cu, void *ws, size_t ws_bytes, double *out,
                           unsigned long long *exps_done, void *stream);

/* ---- Learned score surrogate: MLP building blocks (no reference counterpart;
 * the learned pattern). BF16 weights (uint16 bit patterns), FP32 activations
 * and accumulation; K and N multiples of 8; row-major everything. ---- */

/* y[m, n] = act(sum_k x[m, k] wt[n, k] + bias[n]); act 0 = identity, 1 = SiLU;
 * z (nullable) keeps the pre-activation. With x = dZ and wt = W^T this is the
 * data gradient dX = dZ W. act 2 (bias must be null, z required and READ):
 * y[m, n] = (sum_k x[m, k] wt[n, k]) * silu'(z[m, n]), z with row stride N,
 * i.e. the previous layer's dZ in one pass. */
int fsk_mlp_linear_fwd(const float *x, int64_t ldx, int64_t M, int K, const uint16_t *wt_bf16,
                       const float *bias, int N, int act, float *z, float *y, int64_t ldy,
                       void *stream);
/* Output layer of a training step with the loss fused in: z = x wt^T + bias,
 * d = z - t (t row stride N), dy = scale d (row stride N) and *loss (nullable,
 * FP64) += sum d^2; z itself is not stored. K <= 256. Same values as
 * fsk_mlp_linear_fwd (act 0) followed by fsk_mlp_mse. */
int fsk_mlp_linear_fwd_mse(const float *x, int64_t ldx, int64_t M, int K, const uint16_t *wt_bf16,
                           const float *bias, int N, const float *t, double scale, float *dy,
                           double *loss, void *stream);
/* dw_part[s] (N x K) = sum_{m in split s} bf16(dz[m, :])^T bf16(x[m, :]),
 * db_part[s] (N) = sum_{m in split s} dz[m, :] (FP32), splits >= 1, split s
 * covering rows [s L, (s+1) L) with L = ceil(M / splits) rounded up to 64
 * (deterministic). dz and x 16-byte aligned, ldx a multiple of 4. */
int fsk_mlp_linear_bwd_weight(const float *dz, int64_t M, int N, const float *x, int64_t ldx,
                              int K, int splits, float *dw_part, float *db_part, void *stream);
/* Backward of a hidden layer in one launch: the weight-gradient partials of
 * fsk_mlp_linear_bwd_weight and the data gradient of fsk_mlp_linear_fwd act 2,
 * dx[m, k] = (sum_n dz[m, n] W[n, k]) * silu'(z_prev[m, k]) with wt_bf16 the
 * (K x N) BF16 copy of W^T and z_prev, dx of row stride K. N <= 256. */
int fsk_mlp_linear_bwd(const float *dz, int64_t M, int N, const float *x, int64_t ldx, int K,
                       int splits, float *dw_part, float *db_part, const uint16_t *wt_bf16,
                       const float *z_prev, float *dx, void *stream);
/* Weight gradients of several layers in one launch: fsk_mlp_linear_bwd_weight
 * for each entry (M batch rows shared by all). At most 4 layers. The layer
 * inputs are FP32 (rounded to BF16 as they are staged), or already BF16
 * (uint16) when x_bf16 is nonzero; the products are the same either way. */
typedef struct {
  const float *dz; /* M x N output gradients */
  const void *x;   /* M x K layer inputs, row stride ldx (elements) */
  int64_t ldx;
  int N, K, splits, x_bf16;
  float *dw_part, *db_part;
} fsk_mlp_wgrad_layer;
int fsk_mlp_wgrad_multi(const fsk_mlp_wgrad_layer *layers, int n_layers, int64_t M,
                        void *stream);
/* The row-local part of one training step of the 16-256-256-256-16 network in
 * one launch: with c = min(*counter, steps - 1), gather batch rows
 * x[idx[c*batch + r]] and t[...] (xb keeps the inputs), copy table[3c..3c+2]
 * to sched, run the four forward layers (BF16 operands, FP32 accumulate,
 * SiLU), the MSE gradient dz4 = mse_scale (y - t) (loss += sum (y - t)^2) and
 * the data gradients dz_{l-1} = (dz_l W_l) silu'(z_{l-1}), writing what the
 * weight gradients need: the activations h1..h3 as BF16 (the values the next
 * layer multiplies) and the gradients dz1..dz4 in FP32 (the bias gradients
 * sum them), row-major. Values equal those of the layer-by-layer entry points
 * (h rounded to BF16). Finish the step with fsk_mlp_wgrad_multi (x_bf16 = 1
 * for the h inputs) and fsk_mlp_adam_multi_step. */
typedef struct {
  const float *x, *t; /* normalised training inputs / targets, n x 16 */
  int64_t n;
  const int32_t *idx; /* steps x batch row indices in [0, n) */
  int64_t batch, steps;
  const float *table; /* steps x 3 schedule (lr, 1 - beta1^t, 1 - beta2^t) */
  const int64_t *counter;
  float *sched;
  int in_dim, hidden, out_dim, reserved; /* must be 16, 256, 16 */
  const uint16_t *w_bf16[4], *wt_bf16[4];
  const float *bias[4];
  float *xb;    /* batch x 16 */
  uint16_t *h[3]; /* batch x 256, BF16 */
  float *dz[4]; /* batch x 256 (layers 1..3), batch x 16 (layer 4) */
  double mse_scale;
  double *loss; /* nullable */
} fsk_mlp_rows_step_args;
int fsk_mlp_rows_step(const fsk_mlp_rows_step_args *args, void *stream);
/* Adam on FP32 master weights w (rows x cols) with moments m, v; the gradient
 * is the sum of `splits` partials; also writes w_bf16 (rows x cols) and
 * wt_bf16 (cols x rows) when non-null. step counts from 1. If `sched` (DEVICE,
 * 3 floats: lr, 1 - beta1^step, 1 - beta2^step) is non-null it overrides lr and
 * step, so one captured graph can replay every training step. */
int fsk_mlp_adam(float *w, float *m, float *v, const float *grad_part, int splits, int64_t count,
                 int rows, int cols, double lr, double beta1, double beta2, double eps, int step,
                 uint16_t *w_bf16, uint16_t *wt_bf16, const float *sched, void *stream);
/* One tensor of a batched Adam update (fsk_mlp_adam_multi). */
typedef struct {
  float *w, *m, *v;           /* master weights and moments, rows x cols */
  const float *grad_part;     /* splits partials of count values each */
  uint16_t *w_bf16, *wt_bf16; /* nullable BF16 copies (w, transposed w) */
  int64_t count;
  int rows, cols, splits, reserved;
} fsk_adam_tensor;
#define FSK_ADAM_MAX_TENSORS 16
/* The Adam update of fsk_mlp_adam for up to FSK_ADAM_MAX_TENSORS tensors in
 * one launch (the host array is read at the call; capturable). */
int fsk_mlp_adam_multi(const fsk_adam_tensor *tensors, int n_tensors, double lr, double beta1,
                       double beta2, double eps, int step, const float *sched, void *stream);
/* fsk_mlp_adam_multi that also adds 1 to *counter (the training step counter
 * of fsk_mlp_gather_batch), so a captured step needs no fsk_mlp_step_advance. */
int fsk_mlp_adam_multi_step(const fsk_adam_tensor *tensors, int n_tensors, double lr,
                            double beta1, double beta2, double eps, int step, const float *sched,
                            int64_t *counter, void *stream);
/* Training-step inputs read on the device, so a captured step replays
 * unchanged: with c = min(*counter, steps - 1), xb[r] = x[idx[c*batch + r]],
 * tb[r] = t[idx[c*batch + r]] (rows of dx and dt floats, multiples of 4,
 * 16-byte aligned) and sched[0..2] = table[3c .. 3c+2]; idx holds values in
 * [0, n). fsk_mlp_step_advance adds 1 to *counter (call it last in a step). */
int fsk_mlp_gather_batch(const float *x, const float *t, int64_t n, int dx, int dt,
                         const int32_t *idx, int64_t batch, const float *table, int64_t steps,
                         const int64_t *counter, float *xb, float *tb, float *sched,
                         void *stream);
int fsk_mlp_step_advance(int64_t *counter, void *stream);
/* BF16 copies of w (rows x cols) and of its transpose. */
int fsk_mlp_cast_bf16(const float *w, int rows, int cols, uint16_t *w_bf16, uint16_t *wt_bf16,
                      void *stream);
/* dy = scale (y - t); loss (nullable, FP64) += sum (y - t)^2. */
int fsk_mlp_mse(const float *y, const float *t, int64_t count, double scale, float *dy,
                double *loss, void *stream);

/* ---- FP64 reference kernels (measurement: the error of the reduced-precision
 * forms at sizes a CPU FP64 oracle cannot reach). Direct differences in FP64,
 * FP64 exp and accumulation; inputs are contiguous (rows, 16) float64. */

/* log p(q_i) by log-sum-exp (the FP64 semantics of
 * experiments/error_suite_16d/reference.py:32-82, reference_log_density).
 * Workspace: fsk_ref64_workspace_bytes(n, m). */
size_t fsk_ref64_workspace_bytes(int64_t n, int64_t m);
int fsk_ref64_logdensity16(const double *data, int64_t n, const double *query, int64_t m,
                           double bandwidth, void *ws, size_t ws_bytes, double *out,
                           void *stream);

/* pdf[i] = sum_j phi_ij, w[i,:] = sum_j phi_ij x_j (FP64; the quantities of
 * flash_sd_kde/reference.py:263-284 without rounding). */
int fsk_ref64_score16(const double *data, int64_t n, double bandwidth, double *pdf, double *w,
                      void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FLASHSDKDE_H */
