/*
 * haplo_hip.h — C ABI of libhaplo_hip.so, the MI355X (gfx950) likelihood engine for HaploADMIXTURE.
 *
 * The library replaces the reference's GPU plugin (src/cuda/structs.jl, kernels.jl, runners.jl; loaded
 * by src/HaploADMIXTURE.jl:13-21) behind plain-C entry points: no C++ or framework types cross the
 * boundary, all arrays are host pointers in the reference's column-major layouts (src/structs.jl:14-139):
 *     q      K x I           q[k + K*i]
 *     p      K x 4J          p[k + K*(4j + h)],  h = 0..3 = haplotypes 00, 01, 10, 11
 *     XtX_q  K x K x I       XtX[k2 + K*(k + K*i)]          (reference XtX_q[k2, k, i])
 *     Xtz_q  K x I
 *     XtX_p  K x K x 4J,  Xtz_p K x 4J
 * BED input is the PLINK 1 payload without its 3-byte magic: SNP-major, ceil(I_file/4) bytes per SNP,
 * individuals packed 4 per byte LSB-first (decoded as in src/loops.jl:55-83).
 *
 * Every function returns an hk_status (0 = success, < 0 = error); hk_last_error() gives the message.
 * A context owns all device memory and one HIP stream on the device that was current at creation.
 * Calls block until their results are on the host.  A context is not thread-safe; distinct contexts
 * are independent.
 */
#ifndef HAPLO_HIP_H
#define HAPLO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hk_ctx hk_ctx;

enum hk_status {
  HK_OK = 0,
  HK_EINVAL = -1,   /* bad argument (sizes, K out of range, null pointer)         */
  HK_EHIP = -2,     /* HIP runtime error (no device, launch failure, ...)         */
  HK_ENOMEM = -3,   /* device allocation failed                                   */
  HK_ENAN = -4,     /* NaN / Inf in parameters or results (algorithms_inner.jl:132-133) */
  HK_EQP = -5       /* a row QP failed (singular curvature)                       */
};

enum hk_precision {
  HK_PREC_FP64 = 0,  /* every per-pair product and every sum in FP64 (reference T = T2 = Float64) */
  HK_PREC_MIXED = 1, /* FP32 per-pair products, FP64 accumulation, FP64 log-likelihood, QP and QN */
  HK_PREC_FP32 = 2   /* the reference's mixed mode T2 = Float32 (driver.jl:149, algorithms_inner.jl:115-130):
                        FP32 per-pair products AND FP32 gradient / curvature accumulation; log-likelihood
                        sum, QP and QN in FP64.  A comparison point for HK_PREC_MIXED. */
};

/* Largest K.  K <= 8 runs the register-resident pass kernels; 9 <= K <= 16 runs passes over windows of the packed
 * K x K columns (slower per pair).  The four-lane P row QPs serve every K (spilling to scratch above 8). */
#define HK_MAX_K 16
#define HK_SKFR_MAX_K 16 /* largest SKFR cluster count */

/* _cu_admixture_base (src/cuda/structs.jl:34-44, CuAdmixData :17-32): upload the genotypes of the first
 * I individuals and first 2J SNPs of a BED payload with bytes_per_snp = ceil(I_file/4) bytes per SNP and
 * n_snp >= 2J SNPs, and allocate the parameter and pass buffers.  1 <= K <= HK_MAX_K. */
int hk_ctx_create(hk_ctx** out, int64_t I, int64_t J, int K, const uint8_t* bed, int64_t bytes_per_snp,
                  int64_t n_snp, int precision);
/* A further fit context (own K, precision, parameters, buffers and stream) over the genotype upload of
 * `base`, which must outlive it.  Used for several fits of one dataset (grid of K x restarts). */
int hk_ctx_create_shared(hk_ctx** out, hk_ctx* base, int K, int precision);
int hk_ctx_destroy(hk_ctx* ctx);
const char* hk_last_error(const hk_ctx* ctx);
/* Library-level error text for failures before a context exists (hk_ctx_create). */
const char* hk_create_error(void);

/* copyto_sync!([d_cu.q, d_cu.p], [q, p]) (src/algorithms_inner.jl:36,124,250): set the parameters the
 * passes read.  HK_ENAN if any value is NaN or infinite. */
int hk_set_params(hk_ctx* ctx, const double* q, const double* p);
int hk_get_params(hk_ctx* ctx, double* q, double* p);

/* loglikelihood(d_cu, g_cu, q, p) (src/cuda/runners.jl:2-16 -> loops.jl:236-258). rows (may be NULL)
 * receives the I per-individual sums. */
int hk_loglik(hk_ctx* ctx, double* ll, double* rows);

/* em!(d_cu, g_cu) / em_p!(d_cu, g_cu) (src/cuda/runners.jl:18-55 -> loops.jl:260-326): UNNORMALISED
 * EM sums q_next (K x I, skipped when fix_q != 0) and p_next (K x 4J). */
int hk_em(hk_ctx* ctx, int fix_q, double* q_next, double* p_next);

/* update_q_cuda!(d_cu, g_cu) (src/cuda/runners.jl:58-72 -> loops.jl:328-365): XtX_q, Xtz_q (not negated). */
int hk_update_q(hk_ctx* ctx, double* XtX, double* Xtz);

/* update_p_cuda!(d_cu, g_cu) (src/cuda/runners.jl:76-92 -> loops.jl:367-403): XtX_p, Xtz_p. */
int hk_update_p(hk_ctx* ctx, double* XtX, double* Xtz);

/* count_double_missing! (src/utils.jl:1-18): dm (length I), the counts the EM normaliser 2(J - dm)
 * (src/algorithms_inner.jl:66) uses.  After hk_ctx_create they cover the fit's J blocks; after
 * hk_count_double_missing_file, every SNP pair of the file. */
int hk_double_missing(hk_ctx* ctx, int64_t* dm);

/* count_double_missing!(doublemissing, g) as AdmixData2 calls it (src/structs.jl:169 -> src/utils.jl:1-4):
 * over the SNP pairs of the WHOLE file, ceil(n_snp / 2) in the reference, not only the J blocks of the fit,
 * so the EM normaliser matches the reference when the file holds more than 2J SNPs.  bed, bytes_per_snp and
 * n_snp describe the whole payload as for hk_ctx_create.  An odd last SNP has no partner and is not
 * counted (the reference's loop reads one column past the matrix there).  Replaces the context's counts;
 * contexts sharing its genotype upload see them too.  Must be called on the context that owns the upload. */
int hk_count_double_missing_file(hk_ctx* ctx, const uint8_t* bed, int64_t bytes_per_snp, int64_t n_snp);

/* Genotype codes as the device stream holds them: codes[s*I + i] = 2-bit BED code of individual i at
 * SNP s (s < 2J); bit-exact decode check of src/loops.jl:55-63. */
int hk_decode(hk_ctx* ctx, uint8_t* codes);

/* ---- Runner level: the device arrays of CuAdmixData (src/cuda/structs.jl:1-15) by name ----
 * So that src/algorithms_inner.jl and src/algorithms_outer.jl run unchanged on a HipAdmixData whose fields
 * proxy these arrays: copyto_sync! into / out of d_cu.q, d_cu.p, d_cu.q_next, d_cu.p_next, d_cu.XtX_q, ...
 * (algorithms_inner.jl:36,42,80,124,126,182,250,252,326; algorithms_outer.jl:50,122,149), d_cu.p .= d_cu.p_next
 * (algorithms_outer.jl:9-10) and loglikelihood(d_cu, g_cu, d_cu.q, d_cu.p) (algorithms_outer.jl:11,51,123,150;
 * algorithms_inner.jl:183,327).  Arrays hold FP64 in the reference's layouts; HK_BUF_Q / HK_BUF_P are the
 * parameters hk_set_params sets.  The curvature arrays are allocated on first use (zero-filled). */
enum hk_buffer {
  HK_BUF_Q = 0,      /* d_cu.q       K x I      */
  HK_BUF_P = 1,      /* d_cu.p       K x 4J     */
  HK_BUF_Q_NEXT = 2, /* d_cu.q_next  K x I      */
  HK_BUF_P_NEXT = 3, /* d_cu.p_next  K x 4J     */
  HK_BUF_XTX_Q = 4,  /* d_cu.XtX_q   K x K x I  */
  HK_BUF_XTZ_Q = 5,  /* d_cu.Xtz_q   K x I      */
  HK_BUF_XTX_P = 6,  /* d_cu.XtX_p   K x K x 4J */
  HK_BUF_XTZ_P = 7   /* d_cu.Xtz_p   K x 4J     */
};
/* copyto_sync!([d_cu.<buf>], [host]) and copyto_sync!([host], [d_cu.<buf>]); HK_ENAN for non-finite
 * parameters uploaded into q / p / q_next / p_next. */
int hk_buf_upload(hk_ctx* ctx, int buf, const double* host);
int hk_buf_download(hk_ctx* ctx, int buf, double* host);
/* d_cu.<dst> .= d_cu.<src> (algorithms_outer.jl:9-10); both arrays of the same shape. */
int hk_buf_copy(hk_ctx* ctx, int dst, int src);
/* loglikelihood(d_cu, g_cu, d_cu.<qbuf>, d_cu.<pbuf>) (src/cuda/runners.jl:57-60). */
int hk_loglik_buf(hk_ctx* ctx, int qbuf, int pbuf, double* ll);
/* em!(d_cu, g_cu) / em_p!(d_cu, g_cu) (src/cuda/runners.jl:89-99): unnormalised EM sums from d_cu.q, d_cu.p
 * into d_cu.q_next (skipped when fix_q != 0) and d_cu.p_next. */
int hk_em_buf(hk_ctx* ctx, int fix_q);
/* update_q_cuda!(d_cu, g_cu) / update_p_cuda!(d_cu, g_cu) (src/cuda/runners.jl:115-118, 133-136): curvature and
 * gradient from d_cu.q, d_cu.p into d_cu.XtX_q, d_cu.Xtz_q / d_cu.XtX_p, d_cu.Xtz_p (Xtz not negated). */
int hk_update_q_buf(hk_ctx* ctx);
int hk_update_p_buf(hk_ctx* ctx);

/* Device-resident row steps of update_q! / update_p! (src/algorithms_inner.jl:97-211, 224-355):
 * pass + per-row QP + step + projection on the current parameters; q_next / p_next are returned.
 * hk_q_step uses the current (q, p); hk_p_step uses (q_pass, p) where q_pass may be NULL (current q). */
int hk_q_step(hk_ctx* ctx, int penalty, double rho, double* q_next);
int hk_p_step(hk_ctx* ctx, const double* q_pass, int penalty, double rho, double* p_next);

/* Options of the device-resident fit: run_admixture kwargs (src/driver.jl:38-57) and admixture_qn!
 * defaults (src/algorithms_outer.jl:38-64). */
typedef struct {
  int em_iters;          /* admix_n_em_iters, 5; init_em! warm-up passes                         */
  int n_iter;            /* admix_n_iter, 1000                                                   */
  double rtol;           /* admix_rtol, 1e-7 (relative change of the log-likelihood)             */
  int Qn;                /* Q, 3 (secant pairs of the ZAL quasi-Newton step, 1..4)               */
  int fix_q, fix_p;      /* driver.jl:54-55                                                      */
  int q_penalty;         /* 0 (default): sum(dq) = 0 equality; 1: rho 11' penalty                */
  int p_penalty;         /* 1 (default): rho on every (l, l2) block diagonal; 0: K equalities    */
  double rho;            /* 1e7                                                                  */
  double eps;            /* simplex pseudocount 1e-5                                             */
  int verbose;           /* print per-iteration lines like @info in algorithms_outer.jl:18,162   */
  int profile;           /* time every pass with HIP events (hk_fit_result.t_*)                  */
  int use_graph;         /* replay each QN iteration from a captured hipGraph (ignored if profile) */
  int em_fix_q;          /* EM warm-up updates P only (em_p!, algorithms_inner.jl:40): keeps a given Q */
  int guarantee_increase; /* step halving of every Q/P step until the log-likelihood does not drop
                             (update_q!/update_p!, algorithms_inner.jl:170-199,314-343); eager path */
  int qn_mode;           /* admixture_qn!(mode=...) (algorithms_outer.jl:39,128-141): 0 = :ZAL (default),
                            1 = :LBQN, limited-memory Broyden of the second kind on the same secants     */
  int batch_rounds;      /* hk_fit_batch (read from the first fit's options): 1 (default) = every fit's QN
                            iteration in ONE captured graph per round, replayed with a device-side stopping
                            rule per fit; 0 = each fit's iteration enqueued on its own stream as soon as its
                            previous one completed.  Fits with guarantee_increase, profile or verbose, or
                            use_graph = 0, always take the second form. */
} hk_fit_opts;

typedef struct {
  double ll;             /* d.ll_new at exit                                                      */
  double ll_init;        /* log-likelihood entering the QN loop ("initial ll")                    */
  int iters;             /* QN iterations run                                                     */
  int converged;         /* 1 if the relative change fell below rtol                              */
  int qp_failures;       /* row QPs that failed and kept their row unchanged                      */
  double seconds;        /* wall time of hk_fit                                                   */
  /* profile != 0: summed device time (ms) and launch counts per kernel family */
  double t_qpass_ms, t_ppass_ms, t_ll_ms, t_qstep_ms, t_pstep_ms, t_em_ms, t_qn_ms;
  int n_qpass, n_ppass, n_ll, n_qstep, n_pstep, n_em, n_qn;
  int qp_iter_limit;     /* row QPs stopped at the active-set iteration limit (4n + 20 for n variables):
                            their last iterate, feasible by construction, was applied, not a solution of the QP */
} hk_fit_result;

void hk_fit_opts_default(hk_fit_opts* o);

/* init_em! + admixture_qn! (src/algorithms_outer.jl:1-172, ZAL mode) entirely on the device, starting
 * from the context's parameters (left holding the fit).  ll_start: the log-likelihood the QN loop starts
 * from when em_iters == 0 (NaN: evaluate it).  ll_trace (may be NULL, length n_iter) receives ll_new per
 * iteration. */
int hk_fit(hk_ctx* ctx, const hk_fit_opts* opts, double ll_start, double* ll_trace, hk_fit_result* res);
/* hk_fit plus t_trace (may be NULL, length n_iter): seconds since the call began at which iteration i's
 * ll_new was known on the host. */
int hk_fit_ex(hk_ctx* ctx, const hk_fit_opts* opts, double ll_start, double* ll_trace, double* t_trace,
              hk_fit_result* res);

/* Fits of several contexts (typically hk_ctx_create_shared over one upload, uneven K) run concurrently: the EM
 * warm-up in lockstep, then the QN iterations of all fits as one captured graph per round whose kernels each
 * fit's device-side stopping rule switches off once it stops (batch_rounds, the default), or each fit's
 * iteration enqueued as soon as its previous one completed.  Each fit's results are those of hk_fit_ex on its
 * context.  opts, ll_start (may be NULL) and res have n entries; seconds_at_finish
 * (may be NULL) receives the wall time at which each fit stopped. */
int hk_fit_batch(hk_ctx** ctxs, int n, const hk_fit_opts* opts, const double* ll_start, hk_fit_result* res,
                 double* seconds_at_finish);

/* Autotuner: time candidate chunk counts of the passes on the context's data (reps launches each; each pass
 * timed with the reduction over its chunks) and, where K and the precision have two curvature-pass variants
 * (hk_set_pass_shape), both variants at each count; keep the fastest.  The chosen q/p chunk counts are returned,
 * the variants are read with hk_get_pass_shape.  No counterpart in the reference (its kernels take fixed launch
 * shapes). */
int hk_autotune(hk_ctx* ctx, int reps, int* q_chunks, int* p_chunks);

/* Curvature-pass variant per side (q: the individual-parallel update_q pass, p: the block-parallel update_p
 * pass): 0 = the default for K, 1 = the other staging form (FP32 passes: K <= 6 on both sides, register
 * prefetch of the next stage off; K = 7 on the Q side, prefetch on).  Same arithmetic in the same order, so
 * results are bitwise the same; HK_EINVAL where K and the precision have no second variant. */
int hk_set_pass_shape(hk_ctx* ctx, int q_shape, int p_shape);
int hk_get_pass_shape(hk_ctx* ctx, int* q_shape, int* p_shape);

/* Pass geometry tuning: number of j chunks of the individual-parallel passes and i chunks of the
 * block-parallel passes (0 = automatic). */
int hk_set_tuning(hk_ctx* ctx, int q_chunks, int p_chunks);
int hk_get_tuning(hk_ctx* ctx, int* q_chunks, int* p_chunks);

/* Wait for all work of the context's stream. */
int hk_sync(hk_ctx* ctx);

/* No counterpart in the reference: hk_loglik for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  q (K x I), p (K x 4J) and ll (one
 * double) are DEVICE pointers; rows (I doubles, device) may be NULL.  The work is enqueued on the context's
 * stream and not waited for (hk_sync).  It calls no hipMalloc / hipFree: the staged parameters and partial
 * sums live in a fixed 64 MiB device array that each GPU holds for the life of the process, shared by every
 * call on that GPU -- calls that overlap on one GPU race.  A problem that needs more than that (about
 * (16 K + 128 ceil(6K / 16)) J bytes for the MIXED engine -- FP32 p and its BF16 terms, 480 J at K = 6 --,
 * 32 K J for FP64, plus one partial row of I doubles) is rejected with HK_EINVAL. */
int hk_loglik_noalloc(hk_ctx* ctx, const double* q, const double* p, double* ll, double* rows);

/* Time one pass kind on the current parameters: kind 0 = log-likelihood, 1 = update_q pass,
 * 2 = update_p pass, 3 = EM pass; reps launches timed with HIP events on the context stream.
 * ms_per_launch receives the mean device time of the pass's main kernel. */
int hk_time_pass(hk_ctx* ctx, int kind, int reps, double* ms_per_launch);

/* ---- Learned warm start (no counterpart in the reference: its init_q hook, driver.jl:56,184, is fed) ----
 * Features: f_i = S^-1/2 sum_{s observed} (g_is - mu_s) R_s over the first S = 2 n_blocks SNPs, R a seeded
 * Gaussian S x 256 matrix, mu_s the observed mean dosage; F is I x 256 row-major (host). */
int hk_features(hk_ctx* ctx, int n_blocks, uint64_t seed, float* F);

/* MLP din -> width -> width -> dout (ReLU, softmax), BF16 weights, FP32 accumulation, Adam; 2 <= dout <= HK_MAX_K. */
typedef struct hk_mlp hk_mlp;
int hk_mlp_create(hk_mlp** out, int din, int width, int dout, uint64_t seed);
int hk_mlp_destroy(hk_mlp* m);
const char* hk_mlp_last_error(const hk_mlp* m);
/* Parameter block: FP32 masters W1 (din x width) b1 W2 (width x width) b2 W3 (width x dout) b3, row-major,
 * each segment starting at offsets[t] (6 entries); n = total length. */
int hk_mlp_num_params(hk_mlp* m, int64_t* n, int64_t* offsets);
int hk_mlp_get_params(hk_mlp* m, float* w, float* grad);
int hk_mlp_set_params(hk_mlp* m, const float* w);
/* Keep a training set on the device: X (n x din), T (n x dout) targets (may be NULL for inference only). */
int hk_mlp_load_data(hk_mlp* m, const float* X, const float* T, int64_t n);
/* P (n x dout) = softmax(MLP(X)) for host X. */
int hk_mlp_forward(hk_mlp* m, const float* X, int64_t n, float* P);
/* P (I x dout, host) = softmax(MLP(features of ctx)): the hk_features projection of ctx's first n_blocks
 * blocks (seed), computed on the device and fed to the network without a host round trip; scratch is kept
 * with the network for the next call (din must be 256).  Errors are reported on ctx. */
int hk_mlp_predict(hk_mlp* m, hk_ctx* ctx, int n_blocks, uint64_t seed, float* P);
/* Forward + backward on resident rows [row0, row0+n): gradients stay on the device (hk_mlp_get_params);
 * match != 0 relabels the target columns by the cross-entropy-minimising permutation (every permutation
 * enumerated for dout <= 8, the first minimum kept; a minimum-cost assignment above). */
int hk_mlp_grad(hk_mlp* m, int64_t row0, int64_t n, int match, float* loss, int* perm);
/* One Adam step (bias-corrected, step >= 1) on the stored gradients; refreshes the BF16 weights. */
int hk_mlp_adam(hk_mlp* m, float lr, float beta1, float beta2, float eps, int step);

/* ---- SKFR marker pre-selection: _filter_SKFR (src/driver.jl:84-131) over SparseKmeansFeatureRanking's
 * sparsekmeans1 (mode 0, skfr_mode = :global) / sparsekmeans2 (mode 1, :local) and sparsekmeans_repeat
 * (tries > 1) on the context's J SNP blocks (blocksize = 2: SNPs 2b and 2b+1 form block b).
 * Sparse k-means with feature ranking on standardised allele-2 dosages: k clusters (1..HK_SKFR_MAX_K, <= I),
 * `sparsity` blocks kept (1..J; :local keeps that many per cluster, markers = their union), at most
 * max_iter (skfr_max_inner_iter) reassignments per try; the try with the smallest within-cluster sum of
 * squares over the kept SNPs wins.  seed drives the k-means++ starts; init (k individual indices, or NULL)
 * replaces them in the first try.  Outputs: clusters (I labels 0..k-1), aims (0-based SNP indices of the
 * kept blocks, pairs 2b, 2b+1, most important block first as in docs/src/index.md:287, ranked by
 * sum_c n_c |theta_c,b|^2 of the last center update; room for 2 * min(J, k * sparsity)), n_aims, and
 * (each may be NULL) wss and iters of the returned try, and prof[8]: device ms and launch counts of the
 * center-update, distance and reassignment kernels, then the whole call's device ms (HIP events; NULL =
 * no timing).  Runs synchronously on the context stream; fit parameters are untouched. */
int hk_skfr(hk_ctx* ctx, int k, int sparsity, int mode, int max_iter, int tries, uint64_t seed,
            const int32_t* init, int32_t* clusters, int32_t* aims, int* n_aims, double* wss, int* iters,
            double* prof);

/* sparsekmeans_path (SparseKmeansFeatureRanking, called by _filter_SKFR for a vector of sparsities, src/driver.jl:
 * 108-131): SKFR for each of n_s DECREASING sparsities, the first as hk_skfr with `tries` seeded starts, each next
 * one started from the previous one's cluster labels.  clusters: n_s x I labels (row t for sparsities[t]);
 * aims: the markers of sparsity t at offset sum_{u<t} 2 min(J, k sparsities[u]) (room for 2 min(J, k s) each),
 * n_aims[t] of them, most important block first; wss and iters (n_s each) may be NULL. */
int hk_skfr_path(hk_ctx* ctx, int k, int n_s, const int32_t* sparsities, int mode, int max_iter, int tries,
                 uint64_t seed, int32_t* clusters, int32_t* aims, int* n_aims, double* wss, int* iters);

/* Calibration: copy `bytes` device-to-device `reps` times with a plain streaming kernel (k_copy16, 16 B per
 * lane) on the context stream; gbps = (read + write bytes) / time.  Used to scale HBM byte counters. */
int hk_copy_bench(hk_ctx* ctx, int64_t bytes, int reps, double* gbps);

#ifdef __cplusplus
}
#endif

#endif /* HAPLO_HIP_H */
