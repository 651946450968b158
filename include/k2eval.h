/*
 * k2eval.h -- C ABI of the MI355X (gfx950) K2 evaluator, libk2eval.so.
 *
 *   K2(v; T, k) = 2 sum_{j < n_zeros} e^{-k g_j^2} Re(w_j e^{i g_j v}),  w_j = 1/(rho_j zeta'(rho_j))
 *
 * plus dK2/dv and d2K2/dv2, for lists of candidate points v against a zero table.
 *
 * Encodings (little-endian uint64 word pairs):
 *   G[2j], G[2j+1]  lo, hi words of gamma_j / 2pi in Q18.110 fixed point; gamma_j is defined
 *                   as 2 pi G_j / 2^110, so the phase gamma_j v / 2pi is an exact integer product
 *   gamma[j]        float64 gamma_j (used for the Gaussian weight and the derivative factors)
 *   w[2j], w[2j+1]  real, imaginary part of w_j
 *   V[2c], V[2c+1]  lo (fraction), hi (integer) words of v_c in 64.64 fixed point
 *
 * Conventions: every entry point returns 0 on success or a K2_E* code, and k2_last_error()
 * describes the last failure of the calling thread.  Host buffers are borrowed for the duration
 * of the call.  Calls on one table handle are serialised by the caller (host threads); distinct
 * handles may be used from distinct threads.  The device work of calls on one handle (or one
 * batch plan) runs in call order even when the calls name different streams: a call whose stream
 * differs from the previous call's first makes its stream wait for that call's work, since the
 * calls share the handle's scratch buffers and weight cache.  Work of distinct handles may
 * overlap.  *_device entry points take device pointers and a hipStream_t
 * (as void*; NULL = the default stream) and are asynchronous with respect to the host, except
 * the first use of a new k, which computes that k's weight arrays and waits for them,
 * k2_polish_device, which reads the still-moving candidates back between Newton steps, and
 * k2_eval_batched_device, which waits for its one-call plan before freeing it.
 *
 * Reference interfaces replaced (reference repository bzyzz/mertens_bound):
 *   search/path1_eval_fast.cpp:15-31   compute_peak_AB            (same signature, below)
 *   search/path1_eval_fast.cpp:58-70   polish_peaks_with_weights  (same signature, below)
 *   search/path1_eval_fast.cpp:15-53 + :77-81 fused -> k2_eval (zero table + candidate list ->
 *   K2 values, without the n_peaks x n_zeros A/B matrices)
 */
#ifndef K2EVAL_H
#define K2EVAL_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* flags */
#define K2_VALUE 1u  /* K2 (always computed) */
#define K2_D1 2u     /* dK2/dv */
#define K2_D2 4u     /* d2K2/dv2 */
#define K2_MIXED 8u  /* FP32 sin/cos + weights after the exact removal of the integer turns,
                        compensated FP32 partial sums, FP64 head zeros and FP64 cross-block sum;
                        calls with at most 8 candidates run in FP64 throughout (faster there) */

/* error codes */
#define K2_OK 0
#define K2_E_ARG 1      /* invalid argument (null pointer, size out of range) */
#define K2_E_HIP 2      /* HIP runtime error */
#define K2_E_NOMEM 3    /* device allocation failed */

typedef struct k2_table k2_table;

/* one request of a batched call: the first n_zeros zeros of the table (gamma < T cutoff),
 * Gaussian constant k, candidates V[2*cand_offset .. 2*(cand_offset+n_cand)) */
typedef struct k2_request {
    uint32_t n_zeros;
    uint32_t n_cand;
    uint64_t cand_offset;
    double k;
} k2_request;

/* evaluation tiling (chosen by the autotuner; zero fields mean "library default") */
typedef struct k2_tiling {
    int32_t nc;            /* candidates per lane: 1, 2, 4 or 8 */
    int32_t target_blocks; /* workgroups per launch the short zero splits aim at (FP64
                              launches put long splits of 8 short ones first, so they run
                              about a fifth of this many) */
    int32_t mixed_head;    /* mixed form: zeros evaluated in FP64 before the FP32 form takes over;
                              0 = the shortest multiple of 4096 zeros after which the weight
                              magnitudes sum_j 2 e^{-k g_j^2} |w_j| add up to <= mixed_tail */
    double mixed_tail;     /* 0 = 0.16 (FP32 tail error about 7e-10 per unit of magnitude) */
    int32_t z_max_cand;    /* candidate counts up to this run as Z tiles (lanes over zeros, 8
                              candidates per workgroup); 0 = library default, -1 = never */
    int32_t wg_threads;    /* C-tile workgroup size: 256, or 512 (nc 1 or 2 only); 0 = 256 */
} k2_tiling;

const char *k2_last_error(void);
int k2_version(void);

/* Upload a zero table (host arrays).  The table keeps its own device copy. Replaces the
 * reference's per-call table arguments of compute_peak_AB (path1_eval_fast.cpp:20-28). */
int k2_table_create(k2_table **out, uint32_t n_zeros, const uint64_t *G, const double *gamma,
                    const double *w);
int k2_table_destroy(k2_table *t);
uint32_t k2_table_size(const k2_table *t);
int k2_set_tiling(k2_table *t, const k2_tiling *tiling);
int k2_get_tiling(const k2_table *t, k2_tiling *tiling);

/* K2 (+ derivatives) of n_cand candidates over the first n_zeros zeros (0 = all).
 * Host buffers.  Replaces compute_peak_AB + the weighted sum h0 = sum w_i A_i
 * (path1_eval_fast.cpp:15-53,77-81) with the reduction fused on the device. */
int k2_eval(k2_table *t, uint32_t n_cand, const uint64_t *V, double k, uint32_t n_zeros,
            uint32_t flags, double *K, double *dK, double *d2K);

/* Same with device buffers on a stream; dK / d2K may be NULL when not requested. */
int k2_eval_device(k2_table *t, uint32_t n_cand, const uint64_t *V_dev, double k,
                   uint32_t n_zeros, uint32_t flags, double *K_dev, double *dK_dev,
                   double *d2K_dev, void *stream);

/* k2_eval_device without any device allocation: no counterpart in the reference; for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops).  The partial sums go
 * to a static 64 MiB scratch array of the library (one per GPU, held until the process ends)
 * instead of the table's growable buffer; plans are cut to fewer zero splits where they would
 * not fit, and calls whose candidates need more than 64 MiB even so (above about 2.7 million
 * candidates, 1.3 million in the mixed form) return K2_E_ARG.  The weights of k must already be
 * resident in the table (a k2_mixed_head or k2_eval_device call with this k computes them; at
 * most 4 k per table), else K2_E_ARG.  All calls on one GPU share the scratch: calls on
 * different tables or streams that overlap in time race (calls on one table are ordered). */
int k2_eval_device_noalloc(k2_table *t, uint32_t n_cand, const uint64_t *V_dev, double k,
                           uint32_t n_zeros, uint32_t flags, double *K_dev, double *dK_dev,
                           double *d2K_dev, void *stream);

/* Uneven requests in one call (device buffers): requests share the table, each has its own
 * zero cutoff, k and candidate slice (slices may not overlap).  Outputs are indexed like V
 * (cand_offset + i); entries outside every request are left untouched.  One persistent launch
 * (k2_batch_create + k2_batch_run_direct + k2_batch_destroy).  No counterpart in the reference,
 * whose drivers loop over weight vectors and cutoffs one call at a time
 * (search/eval_tech1_600k.py:252-267). */
int k2_eval_batched_device(k2_table *t, uint32_t n_req, const k2_request *req_host,
                           const uint64_t *V_dev, uint64_t n_cand_total, uint32_t flags,
                           double *K_dev, double *dK_dev, double *d2K_dev, void *stream);

/* Batched plans: the work items of a request set, bound to fixed device buffers, with the
 * launches captured once in a HIP graph and replayed by k2_batch_run. */
typedef struct k2_batch k2_batch;

typedef struct k2_batch_tiling {
    int32_t nc;          /* candidates per lane in C tiles (256*nc per tile): 2 or 4 */
    int32_t cz;          /* candidates per Z tile: 8 or 16 */
    int32_t chunk;       /* zeros per C item; Z items take chunk*256*nc/cz */
    int32_t c_min_fill;  /* leftover candidates of a request get a C tile from this many on,
                            else Z tiles */
    int32_t wg_per_cu;   /* resident workgroups per CU of the persistent grid (0 = occupancy) */
} k2_batch_tiling;

typedef struct k2_batch_info {
    int64_t n_items;
    int32_t grid;
    int32_t n_k;
    double useful_evals;   /* sum over requests of n_cand * n_zeros */
    double padded_evals;   /* evaluations issued, idle lanes of partial tiles included */
    int64_t partial_bytes;
} k2_batch_info;

int k2_batch_create(k2_table *t, uint32_t n_req, const k2_request *req_host,
                    const uint64_t *V_dev, uint64_t n_cand_total, uint32_t flags,
                    const k2_batch_tiling *tiling, double *K_dev, double *dK_dev, double *d2K_dev,
                    k2_batch **out);
int k2_batch_run(k2_batch *b, void *stream);         /* HIP graph replay */
int k2_batch_run_direct(k2_batch *b, void *stream);  /* the same launches, no graph */
int k2_batch_get_info(const k2_batch *b, k2_batch_info *info);
int k2_batch_destroy(k2_batch *b);

/* Newton polish: n_steps Newton iterations on K2' = 0 from each V (device buffers, V updated in
 * place).  Step rule of path1_eval_fast.cpp:120-128 (|K''| > 1e-12 guard, step = K'/K''
 * clamped to +-clamp, stop once |step| < 1e-13); the FP64 step is applied exactly to the
 * 64.64 word.  With K2_MIXED the sums run in the precision form and a candidate stops once
 * |step| < 1e-11, the form's noise floor in v (FP32 rounding in K2' moves converged points by
 * ~5e-12).  K/dK/d2K receive the values at the final point; steps_dev the iterations taken;
 * K0_dev (may be NULL) K at the starting points, in the same form, taken from the first pass
 * (the reference's unpolished h0, path1_eval_fast.cpp:77-81, without a pass of its own). */
int k2_polish_device(k2_table *t, uint32_t n_cand, uint64_t *V_dev, double k, uint32_t n_zeros,
                     int n_steps, double clamp, uint32_t flags, double *K_dev, double *dK_dev,
                     double *d2K_dev, int32_t *steps_dev, double *K0_dev, void *stream);
/* Candidate evaluations (K, K', K'' passes x candidates) of the last k2_polish_device call on t:
 * after the first step only the candidates still moving are evaluated again. */
int k2_polish_stats(k2_table *t, int64_t *cand_evals);

/* Several real weight vectors in one pass (device buffers): for m < n_w (<= 8)
 *   K_m = sum_{j < n_zeros} wt[m*ld + j] 2 Re(w_j e^{i g_j v})   (K2_D1: K_m', K2_D2: K_m'')
 * written to K_dev[m*n_cand + c] (dK_dev, d2K_dev likewise).  The reference forms these as
 * h = A @ wt per weight vector after compute_peak_AB (search/eval_tech1_600k.py:243-267: Gaussian
 * at two k, the JP kernel, raw and a 10,000-zero truncation); here the exact phase and the sin/cos
 * of each candidate and zero are shared by all the vectors. */
int k2_eval_weights_device(k2_table *t, uint32_t n_cand, const uint64_t *V_dev, uint32_t n_w,
                           const double *wt_dev, uint32_t ld, uint32_t n_zeros, uint32_t flags,
                           double *K_dev, double *dK_dev, double *d2K_dev, void *stream);

/* Enclosure of K2 (and K2' with K2_D1) over the first n_zeros zeros (device buffers): the exact
 * sum over the table's definitions (gamma_j = 2 pi G_j / 2^110, w_j as stored, k and v as given)
 * lies in [K - K_rad, K + K_rad].  FP64 with error-free (TwoSum) accumulation and per-zero error
 * bounds; tail_dev (2 doubles, optional) receives bounds on sum_{j >= n_zeros} 2 e^{-k g^2} |w_j|
 * and the same with the factor g_j, i.e. on what the rest of the table can add to K2 and K2'.
 * The ball-sum semantics of certification/certify_K2.py:112-121 (S, absS; combined at :186-192),
 * evaluated on the GPU.  flags: K2_VALUE or K2_VALUE | K2_D1. */
int k2_enclose_device(k2_table *t, uint32_t n_cand, const uint64_t *V_dev, double k,
                      uint32_t n_zeros, uint32_t flags, double *K_dev, double *K_rad_dev,
                      double *dK_dev, double *dK_rad_dev, double *tail_dev, void *stream);
/* debug: the enclosure mode's e^{-x} (same operations as on the device), for checking its
 * error bound on the host */
int k2_exp_neg_host(uint32_t n, const double *x, double *out);
/* debug (host only): the zero split plan of a C-tile launch over nz zeros with n_ctiles
 * candidate tiles, target workgroups and long-split factor (1 = uniform splits): out[0] splits
 * in all, out[1] zeros per short split, out[2] long splits (first), out[3] zeros per long split,
 * out[4] workgroups of the launch (the split count padded to a multiple of 8 when there are long
 * splits).  Splits are contiguous in zero order and the last one may be shorter. */
int k2_split_plan(int32_t nz, int32_t n_ctiles, int32_t target, int32_t factor, int32_t *out);

/* The FP64 head (zeros) the mixed form uses for Gaussian constant k over the first n_zeros
 * zeros under the table's tiling (computes and caches k's weights on first use). */
int k2_mixed_head(k2_table *t, double k, uint32_t n_zeros, uint32_t *head);

/* Staged pre-screen on the first zeros (device buffers): for shifts m = -n_shift..n_shift the
 * sums h_m = K2 over the first n zeros at v + m*shift_step, stage value max_m |h_m|; stage s
 * (stage_zeros[s] zeros, increasing, <= 1024) continues only if its value reaches
 * stage_thresholds[s].  val_dev: [n_cand][4] stage values (-1 where not reached), best_shift_dev:
 * argmax m of the last completed stage, pass_dev: 1 if all stages passed; n_cand <= 2^30.  FP32
 * sums after the integer phase reduction of the FP32 form.  Replaces the leaf screen of
 * search/gpu_mertens_stream.cu:186-250 (64/128/256 zeros, 7 shifts of 0.0014, thresholds
 * 0.78/0.855/0.892). */
int k2_screen_device(k2_table *t, uint32_t n_cand, const uint64_t *V_dev, double k,
                     double shift_step, int n_shift, int n_stages, const int32_t *stage_zeros,
                     const float *stage_thresholds, float *val_dev, int32_t *best_shift_dev,
                     int32_t *pass_dev, void *stream);

/* Debug: phase words floor(G_j V_c / 2^110) mod 2^64 for zeros [z0, z0+nz) (host buffers,
 * out[c*nz + j]).  mode 0 = evaluator path, 1 = schoolbook product, 2 = the exact-column phase
 * window: bits [128,174) of the product's columns 3..5, i.e. floor(W / 2^18) or up to 4 below
 * it (mod 2^46); 3 = the phase the FP32 kernels use (columns 4..5 exact, column 3 summed in
 * FP32 and rounded to a unit here), in the same units, within 7.3e3 of floor(W / 2^18); 4 = the
 * one-launch FP64 kernel's remainder words reassembled into W with bit 0 (not used) clear. */
int k2_phase_words(k2_table *t, uint32_t n_cand, const uint64_t *V, uint32_t z0, uint32_t nz,
                   int mode, uint64_t *out);

/* Debug: the evaluator's sin/cos of 2 pi W / 2^64 (mode 0, FP64 with the 1,024-point table;
 * mode 2, FP64 with the 2,048-point table of the one-launch kernel) or the FP32 form's of the
 * window floor(W / 2^18) (mode 1, widened); out[2i] = sin, out[2i+1] = cos.  Host buffers. */
int k2_sincos_words(uint32_t n, const uint64_t *W, int mode, double *out);

/* Debug: measured issue rates of the VALU instruction classes of the evaluation kernels on the
 * current GPU, in lane operations per second: out[0] v_mad_u64_u32 (the exact phase's integer
 * multiply-add), out[1] v_fma_f64, out[2] v_fma_f32, out[3] v_sin_f32 (hardware transcendental).
 * The class peaks of the roofline; no counterpart in the reference. */
int k2_valu_rates(double *out);

/* Timing of the evaluation kernels on the stream they run on (HIP events around each launch).
 * enable != 0 starts collecting; k2_kernel_time returns the summed milliseconds and the launch
 * count since the last reset. */
int k2_profile(k2_table *t, int enable);
int k2_kernel_time(k2_table *t, double *ms, int64_t *launches);

/* ---- reference signatures (search/path1_eval_fast.cpp) -------------------------------------- */

/* search/path1_eval_fast.cpp:15-31.  Same arguments and output layout ([n_peaks][n_zeros],
 * row-major); evaluated on the GPU.  Returns void like the reference; failures are reported
 * through k2_last_error() and leave the outputs NaN. */
void compute_peak_AB(int n_peaks, const uint64_t *x_lo_arr, const uint64_t *x_hi_arr,
                     const double *dy0_arr, int n_zeros, const double *alpha, const double *gamma,
                     const uint64_t *ghi_lo, const uint64_t *ghi_hi, const uint64_t *glo_lo,
                     const uint64_t *glo_hi, const uint64_t *psi_lo, const uint64_t *psi_hi,
                     double *out_A, double *out_B);

/* search/path1_eval_fast.cpp:58-70 */
void polish_peaks_with_weights(int n_peaks, int n_zeros, const double *A, const double *B,
                               const double *gamma, const double *w, double grid_half_width,
                               int grid_steps, double *out_h_unpol, double *out_h_pol,
                               double *out_ddy_pol);

#ifdef __cplusplus
}
#endif
#endif /* K2EVAL_H */
