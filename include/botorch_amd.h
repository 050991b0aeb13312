/*
 * botorch_amd C ABI - MI355X (gfx950) kernels for BoTorch's GP-posterior /
 * cached-root Sobol sampling / qLogNEI hot path and the exact-MLL gradient.
 *
 * Conventions
 *  - Every pointer argument is DEVICE memory unless its comment says "host".
 *  - Matrices are row-major with an explicit leading dimension (elements).
 *  - "Scaled" inputs are (x - center) / lengthscale, coordinate-wise.
 *  - Every call is asynchronous on `stream` (a hipStream_t; NULL = default
 *    stream) unless its comment says it synchronises; scratch comes from the
 *    caller through (ws, ws_bytes) sized by the matching *_workspace_size
 *    function.  The device allocations the library makes itself are two small
 *    tables, each made on the first call for an n_pad on a device and held
 *    until the process ends: the task plan of bta_potrf_inverse (also used by
 *    bta_gp_cache_build and bta_mll_value_grad) and the trace kernel's worker
 *    schedule of bta_mll_value_grad.
 *  - Return value: BTA_OK or a BTA_ERR_* code; bta_last_error() gives the text
 *    (thread-local).
 *
 * Each entry point names the reference interface (botorch 0.18.1 file:line) it
 * stands in for.  The reference's own arithmetic sits in gpytorch /
 * linear_operator, reached from the cited call sites.
 */
#ifndef BOTORCH_AMD_H
#define BOTORCH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BTA_ABI_VERSION 5

enum {
  BTA_OK = 0,
  BTA_ERR_INVALID = 1,     /* bad argument / shape */
  BTA_ERR_NOT_PSD = 2,     /* Cholesky failed after all jitter tries (NotPSDError) */
  BTA_ERR_NAN = 3,         /* NaN input to a factorisation (NanError) */
  BTA_ERR_UNSUPPORTED = 4, /* shape outside the implemented envelope (UnsupportedError) */
  BTA_ERR_HIP = 5,         /* HIP runtime error */
  BTA_ERR_WORKSPACE = 6    /* workspace too small */
};

enum { BTA_KERNEL_MATERN52 = 0, BTA_KERNEL_RBF = 1 };
/* FP64: everything FP64.  MIXED: the two L^-1 contractions on split-FP16 MFMA
 * (three FP16 products of hi/lo operand planes, FP64 segment sums).  MIXED_F32:
 * the contractions on FP32 MFMA with FP64 segment sums.  Factors, kernel
 * values, [alpha | A] products, posterior moments and samples are FP64 in every
 * mode; the mixed modes evaluate the per-sample log-improvement / fat-max
 * transcendentals in FP32. */
enum { BTA_PREC_FP64 = 0, BTA_PREC_MIXED = 1, BTA_PREC_MIXED_F32 = 2 };

/* per-restart status bits written by bta_qlognei_value_grad */
enum { BTA_STATUS_JITTER = 1, BTA_STATUS_NOT_PSD = 2, BTA_STATUS_NAN = 4 };

/* bta_gemm flags */
enum {
  BTA_GEMM_A_LOWER = 1,    /* A[m][k] == 0 for k > m            */
  BTA_GEMM_A_UPPER = 2,    /* A[m][k] == 0 for k < m            */
  BTA_GEMM_B_LOWER = 4,    /* element (n,k) of B == 0 for k > n */
  BTA_GEMM_B_UPPER = 8,    /* element (n,k) of B == 0 for k < n */
  BTA_GEMM_C_LOWER = 16,   /* only output tiles touching the lower triangle */
  BTA_GEMM_B_KN = 32,      /* B stored K x N row-major (else N x K row-major) */
  BTA_GEMM_F32 = 64,       /* A, B, C, C0 are float (else double) */
  BTA_GEMM_A_KM = 128,     /* A stored K x M row-major (else M x K row-major) */
  BTA_GEMM_ACC64 = 256     /* float GEMM: add each 256-wide K segment into FP64 sums */
};

typedef void* bta_stream_t; /* hipStream_t */

const char* bta_version(void);
int bta_abi_version(void);
const char* bta_last_error(void);

/* ------------------------------------------------------------------------ */
/* Covariance construction.                                                  */
/* Replaces gpytorch MaternKernel(nu=2.5) / RBFKernel / ScaleKernel forward,  */
/* reached from models/gp_regression.py:209-214 (SingleTaskGP.forward) and   */
/* models/utils/gpytorch_modules.py:121-133.                                  */
/* ------------------------------------------------------------------------ */

/* Xs[i][j] = (X[i][j] - center[j]) / lengthscale[j] for i < n, 0 for n <= i <
 * n_pad.  If center is NULL the column means of X are used and written to
 * center_out (d doubles) when that is not NULL. */
int bta_scale_inputs(const double* X, int64_t n, int32_t d, const double* center,
                     const double* lengthscale, double* center_out, double* Xs,
                     int64_t n_pad, bta_stream_t stream);

/* K[i][j] = outputscale * k(X1s[i], X2s[j]) (+ diag_add when i == j) for i <
 * n1, j < n2; zero in the padding, except K[i][i] = 1 for n1 <= i < min(n1_pad,
 * n2_pad) when diag_add != 0 (identity padding for factorisations).  K is
 * double, or float when out_f32 != 0. */
int bta_covar(int32_t kind, const double* X1s, int64_t n1, const double* X2s, int64_t n2,
              int32_t d, double outputscale, double diag_add, int32_t out_f32, void* K,
              int64_t ldk, int64_t n1_pad, int64_t n2_pad, bta_stream_t stream);

/* Adjoint of bta_covar w.r.t. X1 (unscaled coordinates):
 * dX1[i][:] = (sum_j G[i][j] dk(X1s[i], X2s[j])/dX1s[i] + extra[i][:]) / lengthscale.
 * G is double or float (g_f32); extra may be NULL. */
int bta_covar_grad(int32_t kind, const double* X1s, int64_t n1, const double* X2s,
                   int64_t n2, int32_t d, double outputscale, const void* G, int32_t g_f32,
                   int64_t ldg, const double* extra, const double* lengthscale, double* dX1,
                   bta_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Dense linear algebra (MFMA).                                              */
/* ------------------------------------------------------------------------ */

/* C = alpha * op(A) * op(B) + beta * C0 with op(A)[m][k] = A[m][k] (default)
 * or A[k][m] (BTA_GEMM_A_KM), op(B)[k][n] = B[n][k] (default) or B[k][n]
 * (BTA_GEMM_B_KN).  M, N, K must be multiples of 64.  batch copies via
 * element strides sA, sB, sC0, sC.  Triangular flags skip zero K-ranges.
 * C0 may alias C.  Stands in for the dense products torch issues inside
 * linear_operator's MatmulLinearOperator / root decompositions. */
int bta_gemm(int32_t flags, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
             int64_t lda, const void* B, int64_t ldb, double beta, const void* C0,
             int64_t ldc0, void* C, int64_t ldc, int64_t batch, int64_t sA, int64_t sB,
             int64_t sC0, int64_t sC, bta_stream_t stream);

/* FP16 hi/lo planes of scale * X (X double when src_f64, else float; rows x
 * cols, row stride ldx), or of scale * X^T when transpose != 0: hi = fp16(x),
 * lo = fp16(2^11 (x - hi)) with x = fp32(scale * X), so x ~ hi + 2^-11 lo with
 * lo in FP16's normal range for |x| >= 2^-14.  Output rows_pad x cols_pad
 * (cols_pad x rows_pad transposed), row stride ldo, zero in the padding.
 * Operand format of bta_gemm_f16x3; no counterpart in the reference (an
 * internal building block of the split-FP16 contractions). */
int bta_split_f16(int32_t src_f64, const void* X, int64_t ldx, int64_t rows, int64_t cols,
                  int64_t rows_pad, int64_t cols_pad, float scale, int32_t transpose,
                  void* hi, void* lo, int64_t ldo, bta_stream_t stream);

/* C = alpha (A_hi + 2^-11 A_lo)(B_hi + 2^-11 B_lo)^T / (a_scale a_row_scale[m] b_scale)
 * without the lo*lo term (planes as bta_split_f16 writes them), on FP16 MFMA.
 * A planes M x K (stride lda), B planes N x K (stride ldb), FP16; C float M x N
 * (stride ldc).  When c_hi is not NULL the planes of c_scale * C are also
 * written (stride ldc).  flags: BTA_GEMM_B_LOWER / BTA_GEMM_B_UPPER.  M
 * multiple of 128, N of 64, K of 32, lda/ldb of 8; a_row_scale may be NULL.
 * k_segment: 0 accumulates in FP32 over all of K; otherwise FP32 segments of
 * k_segment (a multiple of 32) are added into FP64 sums.  No counterpart in the
 * reference: the contraction kernel of BTA_PREC_MIXED. */
int bta_gemm_f16x3(int32_t flags, int64_t M, int64_t N, int64_t K, double alpha,
                   const void* a_hi, const void* a_lo, int64_t lda, const float* a_row_scale,
                   float a_scale, const void* b_hi, const void* b_lo, int64_t ldb,
                   float b_scale, float* C, int64_t ldc, void* c_hi, void* c_lo,
                   float c_scale, int32_t k_segment, bta_stream_t stream);

size_t bta_potrf_workspace_size(int64_t n_pad);

/* In-place lower Cholesky of the n_pad x n_pad matrix A (n_pad a multiple of
 * 64) with linear_operator psd_safe_cholesky semantics (reached from
 * utils/low_rank.py:140, acquisition/cached_cholesky.py:123; max_tries=6 from
 * botorch/__init__.py:54): factor; if it fails add jitter0*10^i to the
 * diagonal (i = 0..max_tries-1) and refactor; BTA_ERR_NOT_PSD after the last
 * try, BTA_ERR_NAN if A holds NaN.  Also writes the inverses of the 64x64
 * diagonal blocks of L into the diagonal blocks of Linv (ld ldinv; may be NULL)
 * and zeroes the strict upper triangle of A.  *jitter_used (host) receives the
 * jitter that was added (0 if none).  A must be 16-byte aligned with an even
 * lda (BTA_ERR_INVALID otherwise: the trailing updates load A directly into
 * LDS 16 bytes at a time).  SYNCHRONISES on `stream` (reads info), as the
 * reference does on any(info). */
int bta_potrf(double* A, int64_t n_pad, int64_t lda, double* Linv, int64_t ldinv,
              int32_t max_tries, double jitter0, double* jitter_used, void* ws,
              size_t ws_bytes, bta_stream_t stream);

/* bta_potrf followed by the whole lower inverse: Linv = L^-1 (strict upper
 * triangle zeroed), built by side tasks inside the factorisation's launches
 * instead of a separate bta_trtri_lower pass.  Linv must be 16-byte aligned
 * with an even ldinv and must not overlap A (BTA_ERR_INVALID otherwise).  The
 * jitter schedule, errors and workspace are bta_potrf's; like it, the call
 * SYNCHRONISES on `stream` (reads info).  The first call for a given n_pad on a
 * device builds a task plan and keeps a device copy of it (201 KB at
 * n_pad = 4096) until the process ends.  Replaces psd_safe_cholesky (as bta_potrf) plus the
 * triangular inverse behind fast_pred_var's R = L^-T (gpytorch
 * DefaultPredictionStrategy via models/gpytorch.py:585) and behind the exact
 * marginal log likelihood's K^-1 (gpytorch ExactMarginalLogLikelihood, called
 * from optim/closures/model_closures.py:133-142). */
int bta_potrf_inverse(double* A, int64_t n_pad, int64_t lda, double* Linv, int64_t ldinv,
                      int32_t max_tries, double jitter0, double* jitter_used, void* ws,
                      size_t ws_bytes, bta_stream_t stream);

/* Lower-triangular inverse: Linv = L^-1, where Linv already holds the inverted
 * 64x64 diagonal blocks (from bta_potrf).  Recursive doubling over block
 * pairs, two GEMMs per level.  Used for fast_pred_var's R = L^-T
 * (gpytorch DefaultPredictionStrategy via models/gpytorch.py:585). */
int bta_trtri_lower(const double* L, int64_t ldl, double* Linv, int64_t ldinv,
                    int64_t n_pad, void* ws, size_t ws_bytes, bta_stream_t stream);

/* Adjoint of the Cholesky factorisation of batch small n x n matrices (n <= 64,
 * matrix b at L + b stride): dA = (S + S^T) / 2 with S = L^-T Phi(L^T dL) L^-1,
 * Phi = lower triangle with the diagonal halved (torch's Cholesky backward, which
 * the reference's posterior samples go through: posteriors/gpytorch.py:86-127). */
int bta_chol_small_backward(const double* L, const double* dL, int32_t n, int64_t stride,
                            int64_t batch, double* dA, bta_stream_t stream);

/* Batched Cholesky of small n x n matrices (n <= 64) with per-matrix jitter
 * escalation; status[b] gets BTA_STATUS_* bits.  Strict upper triangle zeroed. */
int bta_chol_small_batched(double* A, int32_t n, int64_t lda, int64_t stride,
                           int64_t batch, int32_t max_tries, double jitter0,
                           int32_t* status, bta_stream_t stream);

/* ------------------------------------------------------------------------ */
/* GP model: SingleTaskGP exact prediction (fast_pred_var).                  */
/* Replaces models/gpytorch.py:544-610 (BatchedMultiOutputGPyTorchModel.     */
/* posterior) -> gpytorch ExactGP.__call__ / DefaultPredictionStrategy, with  */
/* Standardize.untransform_posterior (models/transforms/outcome.py:431-511). */
/* ------------------------------------------------------------------------ */
typedef struct bta_gp_model {
  int32_t kind;             /* BTA_KERNEL_* */
  int32_t d;
  int64_t n;                /* training points */
  int64_t n_pad;            /* multiple of 128 */
  const double* train_xs;   /* n_pad x d scaled training inputs (zero padding) */
  const double* center;     /* d */
  const double* lengthscale;/* d */
  double outputscale;
  double noise;             /* likelihood noise variance */
  double mean_constant;     /* ConstantMean */
  double y_mean, y_std;     /* Standardize statistics */
  const double* linv;       /* n_pad x n_pad  L^-1 (lower, zero padding); set by caller */
  const double* alpha;      /* n_pad  K^-1 (y_std - c) */
  double chol_jitter;       /* added to the diagonal for the factorisation only: L =
                               chol(K + chol_jitter I) (0: none) */
  int32_t refine_steps;     /* refinement steps of the solves alpha = K^-1 r and
                               A = K^-1 k(X, X_b) against K itself (residual from
                               kernel values evaluated on the fly; 0: none) */
} bta_gp_model;

size_t bta_gp_cache_workspace_size(int64_t n_pad);

/* Builds the prediction caches: K = k(X, X) + noise I, L = chol(K +
 * chol_jitter I) with psd_safe_cholesky's jitter on failure, Linv = L^-1,
 * alpha = Linv^T Linv (y - c), then refine_steps times alpha += Linv^T Linv
 * (y - c - K alpha).  y: n standardized targets.  L, Linv: n_pad x n_pad;
 * alpha: n_pad.  SYNCHRONISES (potrf info). */
int bta_gp_cache_build(const bta_gp_model* m, const double* y, double* L, double* Linv,
                       double* alpha, int32_t max_tries, double* jitter_used, void* ws,
                       size_t ws_bytes, bta_stream_t stream);

size_t bta_gp_posterior_workspace_size(const bta_gp_model* m, int64_t batch, int32_t mq);

/* Posterior over X (batch x mq x d, unscaled) in the ORIGINAL outcome scale:
 * mean (batch x mq), covar (batch x mq x mq, may be NULL), var (batch x mq,
 * may be NULL).  mean = y_mean + y_std (c + K*x alpha);
 * covar = y_std^2 (K** - (K*x L^-T)(K*x L^-T)^T). */
int bta_gp_posterior(const bta_gp_model* m, const double* X, int64_t batch, int32_t mq,
                     double* mean, double* covar, double* var, void* ws, size_t ws_bytes,
                     bta_stream_t stream);

size_t bta_gp_posterior_backward_workspace_size(const bta_gp_model* m, int64_t batch,
                                                int32_t mq);

/* Adjoint of bta_gp_posterior: dX (batch x mq x d) = d <dmean, mean> + <dcov,
 * covar> / dX for dmean (batch x mq) and dcov (batch x mq x mq), either of which
 * may be NULL.  With U = K*x K^-1: dK*x = y_std dmean alpha^T - y_std^2 (dcov +
 * dcov^T) U and dK** = y_std^2 dcov, taken through dk/dX.  Stands in for the
 * autograd backward through SingleTaskGP.posterior (models/gpytorch.py:544-610
 * -> gpytorch ExactGP prediction), which the reference keeps differentiable. */
int bta_gp_posterior_backward(const bta_gp_model* m, const double* X, int64_t batch, int32_t mq,
                              const double* dmean, const double* dcov, double* dX, void* ws,
                              size_t ws_bytes, bta_stream_t stream);

/* ------------------------------------------------------------------------ */
/* qLogNoisyExpectedImprovement (cache_root=True, fat=True, incremental,     */
/* IdentityMCObjective, no constraints).  Replaces the forward + autograd    */
/* backward of acquisition/logei.py:527-565 (_get_samples_and_objectives),   */
/* utils/low_rank.py:84-172 (sample_cached_cholesky), logei.py:376-391,688-  */
/* 715 (_sample_forward / _log_improvement) and monte_carlo.py:268-290       */
/* (fatmax over q, logmeanexp over samples), as called once per L-BFGS step  */
/* by generation/gen.py:423-485 (f_np_wrapper).                              */
/* ------------------------------------------------------------------------ */
typedef struct bta_qlognei_state {
  int32_t precision;        /* BTA_PREC_* */
  int32_t q;                /* 1..8 */
  int32_t n_b;              /* pruned baseline points, 1..256 */
  int32_t S;                /* MC samples */
  const double* xb_s;       /* n_b x d scaled baseline inputs */
  const void* linv;         /* L^-1 as prepared: n_pad x n_pad double (FP64), float
                               (MIXED_F32), or four n_pad x n_pad FP16 planes (MIXED):
                               [hi(cL^-1) | lo(cL^-1) | hi(cL^-T) | lo(cL^-T)],
                               c = linv_scale */
  const double* ext;        /* n_pad x (1 + n_b): [alpha | A], A = K^-1 k(X, X_b) */
  const double* Lb;         /* n_b x n_b cached baseline Cholesky (original scale) */
  const double* Zt;         /* (n_b + q) x S base samples, transposed */
  const double* best_f;     /* S per-sample best baseline objective */
  double tau_relu, tau_max;
  int32_t max_tries;        /* psd_safe_cholesky tries for the q x q block */
  float linv_scale;         /* power of two with max|L^-1| * linv_scale <= 1 (MIXED) */
  double jitter;            /* 1e-8 (FP64 default) */
  const double* Lb_levels;  /* NULL: cached-root sampling with Lb.  Otherwise joint
                               sampling (the reference's fallback when the cached-root
                               update fails, acquisition/cached_cholesky.py:143-158):
                               (max_tries + 1) x n_b x n_b, level t = chol(Sigma_bb +
                               j_t I) with j_0 = 0, j_t = jitter 10^(t-1); the jitter of
                               a restart's try t goes on its whole joint diagonal */
  int32_t lb_levels_mask;   /* bit t: level t of Lb_levels exists (its Cholesky succeeded) */
} bta_qlognei_state;

size_t bta_qlognei_prepare_workspace_size(const bta_gp_model* m, int32_t n_b, int32_t S);

/* Builds ext = [alpha | K^-1 k(X, X_b)] (n_pad x (1 + n_b), FP64), Lb
 * (psd_safe_cholesky of the baseline posterior covariance, original scale),
 * best_f (per-sample max of the baseline samples mu_b + Lb z_b, z_b = Zb (S x
 * n_b)) as acquisition/logei.py:393-459 (_init_baseline) does, L^-1 in the
 * state's precision into linv_out (MIXED_F32: n_pad^2 floats; MIXED: the four
 * FP16 planes of bta_qlognei_state.linv, 4 n_pad^2 halfs; may be NULL for FP64,
 * where the state uses m->linv) and the power-of-two scale c with max|c L^-1|
 * <= 1 into *linv_scale (host, may be NULL).  SYNCHRONISES. */
int bta_qlognei_prepare(const bta_gp_model* m, const bta_qlognei_state* st,
                        const double* Zb, void* linv_out, double* ext, double* Lb,
                        double* best_f, double* jitter_used, float* linv_scale, void* ws,
                        size_t ws_bytes, bta_stream_t stream);

size_t bta_qlognei_workspace_size(const bta_gp_model* m, const bta_qlognei_state* st,
                                  int64_t b);

/* Measured tuner (no counterpart in the reference: the project's own autotuner).
 * For BTA_PREC_MIXED, times every plan candidate (tile width, worker count, XCD
 * groups) of the step's forward and then adjoint split-FP16 product inside
 * whole value(+grad) steps on these arguments (HIP events), keeps the fastest for
 * these problem shapes, and writes them to chosen (host, 6 ints: bn, workers,
 * groups of the forward and of the adjoint product; zeros for other precisions,
 * whose products keep the cost model's plans).  The outputs hold a valid step
 * afterwards.  SYNCHRONISES; not for graph capture. */
int bta_qlognei_autotune(const bta_gp_model* m, const bta_qlognei_state* st, const double* X,
                         int64_t b, double* value, double* grad, int32_t* status, void* ws,
                         size_t ws_bytes, int32_t* chosen, bta_stream_t stream);

/* value[b] = qLogNEI(X[b]) for X: b x q x d (unscaled); grad (b x q x d, may be
 * NULL for value only) = d value[b] / d X[b].  status[b]: BTA_STATUS_* bits. */
int bta_qlognei_value_grad(const bta_gp_model* m, const bta_qlognei_state* st,
                           const double* X, int64_t b, double* value, double* grad,
                           int32_t* status, void* ws, size_t ws_bytes,
                           bta_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Many independent models in one step.  Stands in for running the per-step  */
/* evaluation above (generation/gen.py:423-485 -> acquisition/logei.py:527-  */
/* 565, monte_carlo.py:268-290) model after model, one optimize_acqf per     */
/* SingleTaskGP: the whole batch is five launches (one per stage of the      */
/* step, each covering every model), with no host synchronisation, so it   */
/* can be captured in a HIP graph.                                           */
/* ------------------------------------------------------------------------ */
typedef struct bta_qlognei_batch_item {
  const bta_gp_model* model;
  const bta_qlognei_state* state; /* prepared for this model, with q and Zt set */
  const double* X;                /* b x q x d query points (unscaled) */
  int64_t b;                      /* restarts of this model */
  double* value;                  /* b */
  double* grad;                   /* b x q x d, or NULL (the same for every item) */
  int32_t* status;                /* b, BTA_STATUS_* bits */
} bta_qlognei_batch_item;

/* Written by bta_qlognei_batch_prepare; offsets are bytes into ws. */
typedef struct bta_qlognei_batch_plan {
  void* ws;
  int32_t n_items, q, kind, kd, grad;
  int32_t covar_blocks, mc_blocks, fin_blocks, fwd_tiles, adj_tiles, fwd_workers, adj_workers;
  size_t off_cov, off_fwd, off_adj, off_mc, off_grad, off_fin, off_cov_start, off_mc_start,
      off_fin_start, off_fwd_tiles, off_adj_tiles;
} bta_qlognei_batch_plan;

/* Workspace of a batch: every item's bta_qlognei_workspace_size plus the launch
 * tables (0 for an invalid item list). */
size_t bta_qlognei_batch_workspace_size(const bta_qlognei_batch_item* items, int32_t n_items);

/* Checks the items and writes the per-model launch arguments and the tile
 * lists into ws; fills *plan (host).  Every item must use BTA_PREC_MIXED, and
 * the items must share q (1..5), the kernel kind and the presence of the
 * gradient; n, d, n_b, S and b may differ.  The buffers the items name (X,
 * outputs, model and state arrays) are baked into the tables: refill X in
 * place between steps.  SYNCHRONISES (host -> device copy of the tables). */
int bta_qlognei_batch_prepare(const bta_qlognei_batch_item* items, int32_t n_items,
                              void* ws, size_t ws_bytes, bta_qlognei_batch_plan* plan,
                              bta_stream_t stream);

/* One value(+grad) step of every item of the plan: each item's outputs equal
 * those of bta_qlognei_value_grad on its own. */
int bta_qlognei_batch_value_grad(const bta_qlognei_batch_plan* plan, bta_stream_t stream);

/* ------------------------------------------------------------------------ */
/* qLogEHVI / qLogNEHVI fused log-areas.  Replaces the reference's only       */
/* native op, csrc/logei_fused.cpp:184-373 (fused_log_areas_forward /        */
/* _backward, bound by _FusedLogAreas at acquisition/multi_objective/        */
/* logei.py:107-170): steps 1-4 of _compute_log_qehvi (logei.py:437-479).    */
/* ------------------------------------------------------------------------ */

/* out[b][c][s] = sum_k fatmin(fatmin_j(log_fatplus(obj[b][s][j][k] - cl[bc][c][k],
 * tau_relu), tau_max), log(min(cu[bc][c][k], clamp) - cl[bc][c][k]), tau_max) with
 * obj B x n_sub x i x m, cell bounds B_cells x nc x m (B_cells 1 or B, bc = 0 or b),
 * out B x nc x n_sub; clamp 1e10 (double) / 1e8 (float).  f32 != 0: every tensor is
 * float, else double.  1 <= i <= 32, 1 <= m <= 8. */
int bta_log_areas_forward(int32_t f32, const void* obj, int64_t B, int64_t n_sub, int32_t i,
                          int32_t m, const void* cell_lower, const void* cell_upper,
                          int64_t B_cells, int64_t nc, double tau_relu, double tau_max,
                          void* out, bta_stream_t stream);

/* grad_obj (B x n_sub x i x m) = d (sum grad_out * out) / d obj for grad_out
 * (B x nc x n_sub), the other arguments as in bta_log_areas_forward. */
int bta_log_areas_backward(int32_t f32, const void* grad_out, const void* obj, int64_t B,
                           int64_t n_sub, int32_t i, int32_t m, const void* cell_lower,
                           const void* cell_upper, int64_t B_cells, int64_t nc, double tau_relu,
                           double tau_max, void* grad_obj, bta_stream_t stream);

/* Bytes of the static device array bta_qlognei_value_grad_noalloc works in. */
#define BTA_NOALLOC_SCRATCH_BYTES (96u << 20)

/* bta_qlognei_value_grad without a workspace argument and without any device
 * allocation: the scratch is a static device array of
 * BTA_NOALLOC_SCRATCH_BYTES (96 MiB) in the library, one copy per GPU, held
 * until the process ends; its address is looked up once per device on the
 * first call there (the stream's device must be current).  Calls whose
 * bta_qlognei_workspace_size exceeds the array fail with BTA_ERR_WORKSPACE
 * before anything is launched (the 512-restart, q = 4, n = 2048 step needs
 * ~59 MB in the mixed precision, ~76 MB in FP64).  Calls that overlap on one GPU share the array and race; run them
 * on one stream.  No counterpart in the reference: an entry point for callers
 * that cannot allocate (code under HIP graph capture, serving loops, hosts
 * that forbid hipMalloc).  The first call for a new shape still fills the
 * library's host-side GEMM plan cache (host memory only). */
int bta_qlognei_value_grad_noalloc(const bta_gp_model* m, const bta_qlognei_state* st,
                                   const double* X, int64_t b, double* value, double* grad,
                                   int32_t* status, bta_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Lock-step batched L-BFGS-B on the device.  Replaces the per-restart SciPy  */
/* setulb loop of optim/batched_lbfgs_b.py:440-575 (fmin_l_bfgs_b_batched,    */
/* driven from generation/gen.py:236-249): b independent problems of n        */
/* variables, m correction pairs, L-BFGS-B 3.0 steps and the reference's      */
/* stopping rules (factr, pgtol, maxls, maxiter; maxfun as the library's     */
/* host driver).  All arrays are device memory, row-major b x n.            */
/* ------------------------------------------------------------------------ */

/* Bytes of one problem's state (multiply by b); 0 when n, m do not fit the
 * kernel's LDS budget (64 KiB per problem: n <= 224 at m = 10, 471 at m = 3). */
size_t bta_lbfgsb_state_size(int32_t n, int32_t m);

/* Starts b problems at x0 (projected onto the bounds): nbd[i] = 0 unbounded,
 * 1 lower only, 2 both, 3 upper only (SciPy's convention; lower/upper values
 * are ignored where absent).  Writes the first points to evaluate into x_eval
 * and zeroes *ndone (may be NULL). */
int bta_lbfgsb_init(void* state, int64_t b, int32_t n, int32_t m, const double* x0,
                    const double* lower, const double* upper, const int32_t* nbd, double factr,
                    double pgtol, int32_t maxls, int32_t maxiter, int32_t maxfun,
                    double* x_eval, int32_t* ndone, bta_stream_t stream);

/* One round: f (b) and g (b x n) are the objective and gradient at x_eval,
 * both scaled by fscale (-1 to minimise -acqf); step_status (b, may be NULL)
 * is OR-ed into each problem's record.  Advances every running problem to its
 * next evaluation or its end and writes the next points into x_eval; ended
 * problems keep their final x there and add 1 to *ndone once.  A NaN in a
 * gradient ends that problem (task 6). */
int bta_lbfgsb_step(void* state, int64_t b, int32_t n, int32_t m, const double* lower,
                    const double* upper, const int32_t* nbd, const double* f, const double* g,
                    double fscale, const int32_t* step_status, double* x_eval, int32_t* ndone,
                    bta_stream_t stream);

/* x (b x n), f (b, in the minimised scale) and info (b x 6: iterations,
 * evaluations, warnflag 0 converged / 1 maxiter or maxfun / 2 other, task
 * 1 pgtol / 2 factr / 3 abnormal line search / 4 maxiter / 5 maxfun /
 * 6 NaN gradient, OR of the step statuses, pair updates). */
int bta_lbfgsb_results(const void* state, int64_t b, int32_t n, int32_t m, double* x, double* f,
                       int32_t* info, bta_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Exact marginal log likelihood value + adjoint.  Replaces gpytorch          */
/* ExactMarginalLogLikelihood forward + torch autograd backward as driven by */
/* optim/closures/model_closures.py:125-145 and fit.py:76-289.               */
/* ------------------------------------------------------------------------ */
size_t bta_mll_workspace_size(int64_t n, int32_t d);

/* For K = outputscale k(X, X; lengthscale) + noise I and r = y - constant:
 * out (host, 4 + d doubles) = [ log N(y; c, K)  (un-normalised, no priors),
 *   d/d outputscale, d/d noise, d/d constant, d/d lengthscale[0..d) ].
 * The gradient is 1/2 tr((a a^T - K^-1) dK/dtheta), a = K^-1 r.  X: n x d,
 * y: n, lengthscale: host d doubles.  SYNCHRONISES. */
int bta_mll_value_grad(int32_t kind, const double* X, const double* y, int64_t n, int32_t d,
                       const double* lengthscale, double outputscale, double noise,
                       double constant, int32_t max_tries, double* out, void* ws,
                       size_t ws_bytes, bta_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* BOTORCH_AMD_H */
