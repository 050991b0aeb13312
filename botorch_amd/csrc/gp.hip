// SingleTaskGP prediction caches, posterior, and the qLogNEI baseline cache.
//
// Cache build (once per model, fast_pred_var form):
//   K = k(X,X) + noise I  ->  L = chol(K)  ->  L^-1  ->  alpha = L^-T L^-1 (y - c)
// Posterior over batch x mq points (original outcome scale):
//   V = k(X*, X) L^-T;  mean = y_mean + y_std (c + k(X*,X) alpha)
//   covar = y_std^2 (k(X*,X*) - V V^T)
// qLogNEI baseline cache (acquisition/logei.py:393-459):
//   A^T = V_b L^-1 (= (K^-1 k(X, X_b))^T), Lb = chol(Sigma_bb), best_f from the
//   baseline samples, and the extended operator [L^-1; alpha^T; A^T] used by
//   both GEMMs of every value+grad step.
#include "bta_common.h"
#include "gemm.h"
#include "kernels.h"

#include <algorithm>
#include <cmath>

namespace bta {

// rows [i*mq_pad + r] of Xs = scaled X[i*mq + r] (zero rows in the padding)
__global__ void scale_batched_kernel(const double* X, int64_t batch, int mq, int mq_pad, int d,
                                     const double* center, const double* ls, double* Xs) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tot = batch * mq_pad * d;
  if (idx >= tot) return;
  const int j = (int)(idx % d);
  const int64_t row = idx / d;
  const int64_t bi = row / mq_pad;
  const int r = (int)(row % mq_pad);
  Xs[idx] = r < mq ? (X[(bi * mq + r) * d + j] - center[j]) / ls[j] : 0.0;
}

// mean[i*mq+r] = y_mean + y_std (c + KA[(i*mq_pad + r)*lda]);
// covar/var from the padded batch blocks of Sg (already in the outcome scale)
__global__ void posterior_out_kernel(const double* KA, int64_t lda, const double* Sg,
                                     int64_t batch, int mq, int mq_pad, double y_mean,
                                     double y_std, double c, double* mean, double* covar,
                                     double* var) {
  const int64_t bi = blockIdx.y;
  const double* Sb = Sg + bi * mq_pad * mq_pad;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < mq * mq; e += gridDim.x * blockDim.x) {
    const int i = e / mq, j = e % mq;
    const double v = Sb[i * mq_pad + j];
    if (covar) covar[(bi * mq + i) * mq + j] = v;
    if (i == j) {
      if (var) var[bi * mq + i] = v;
      if (mean) mean[bi * mq + i] = y_mean + y_std * (c + KA[(bi * mq_pad + i) * lda]);
    }
  }
}

// out[blockIdx.x] = max |x[i]| over the block's grid-stride share
constexpr int MAXABS_BLOCKS = 256;
__global__ __launch_bounds__(256) void max_abs_kernel(const double* x, int64_t n, double* out) {
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    m = fmax(m, fabs(x[i]));
  __shared__ double red[4];
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

__global__ void build_ext_kernel(const double* alpha, const double* at, int64_t np, int n_b,
                                 double* E) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= np) return;
  E[c * (n_b + 1)] = alpha[c];
  for (int kb = 0; kb < n_b; ++kb) E[c * (n_b + 1) + 1 + kb] = at[kb * np + c];
}

__global__ void residual_kernel(const double* y, int64_t n, double c, double* r, int64_t n_pad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_pad) r[i] = i < n ? y[i] - c : 0.0;
}

// out[c][i] = Y[c][i] - sum_{j < n} B[c][j] Khat[j][i] for the nc row vectors c of
// B and Y (row stride ld, i < n_pad; Khat = s k(xs, xs) + diag I on the first n
// points, zero beyond), the residual of one refinement step of the solves.
// The kernel values come from direct coordinate differences, evaluated on the
// fly (nothing n x n is stored).  Block: 64 columns i, its 4 waves split each
// 64-point tile of j in four; KR_RHS right-hand sides per pass (grid.y).
constexpr int KR_RHS = 16;

__global__ __launch_bounds__(256) void kmat_residual_kernel(int kind, const double* __restrict__ xs,
                                                            int64_t n, int d, double s,
                                                            double diag,
                                                            const double* __restrict__ B,
                                                            const double* __restrict__ Y,
                                                            int64_t ld, int nc, double* out) {
  __shared__ double xj[64][33];
  __shared__ double bj[KR_RHS][64];
  __shared__ double part[4][KR_RHS][65];
  const int tid = threadIdx.x, il = tid & 63, chunk = tid >> 6;
  const int64_t i = (int64_t)blockIdx.x * 64 + il;
  const int c0 = blockIdx.y * KR_RHS;
  const int ncp = min(KR_RHS, nc - c0);
  double xi[32];
#pragma unroll
  for (int t = 0; t < 32; ++t) xi[t] = (t < d && i < n) ? xs[i * d + t] : 0.0;
  double acc[KR_RHS];
#pragma unroll
  for (int c = 0; c < KR_RHS; ++c) acc[c] = 0.0;
  for (int64_t j0 = 0; j0 < n; j0 += 64) {
    __syncthreads();
    for (int e = tid; e < 64 * d; e += 256) {
      const int jj = e / d, t = e % d;
      xj[jj][t] = j0 + jj < n ? xs[(j0 + jj) * d + t] : 0.0;
    }
    for (int e = tid; e < KR_RHS * 64; e += 256) {
      const int c = e / 64, jj = e % 64;
      bj[c][jj] = (c < ncp && j0 + jj < n) ? B[(int64_t)(c0 + c) * ld + j0 + jj] : 0.0;
    }
    __syncthreads();
    for (int jj = chunk * 16; jj < chunk * 16 + 16; ++jj) {
      const int64_t j = j0 + jj;
      if (j >= n) break;
      double r2 = 0.0;
#pragma unroll
      for (int t = 0; t < 32; ++t)
        if (t < d) {
          const double df = xi[t] - xj[jj][t];
          r2 = fma(df, df, r2);
        }
      const double k = kern_value(kind, r2, s) + (j == i ? diag : 0.0);
#pragma unroll
      for (int c = 0; c < KR_RHS; ++c) acc[c] = fma(bj[c][jj], k, acc[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < KR_RHS; ++c) part[chunk][c][il] = acc[c];
  __syncthreads();
  if (chunk == 0) {
    for (int c = 0; c < ncp; ++c) {
      const double sum = i < n ? ((part[0][c][il] + part[1][c][il]) +
                                  (part[2][c][il] + part[3][c][il]))
                               : 0.0;
      const int64_t o = (int64_t)(c0 + c) * ld + i;
      out[o] = Y[o] - sum;
    }
  }
}

// R = Y - Khat B for the nc rows of B and Y (n_pad columns, stride ld).
static int kmat_residual(const bta_gp_model* m, const double* B, const double* Y, int64_t ld,
                         int nc, double* R, hipStream_t s) {
  if (m->d > 32 || nc < 1) {
    set_error("refinement: d <= 32 and at least one right-hand side needed");
    return BTA_ERR_UNSUPPORTED;
  }
  const dim3 grid((unsigned)(m->n_pad / 64), (unsigned)((nc + KR_RHS - 1) / KR_RHS));
  hipLaunchKernelGGL(kmat_residual_kernel, grid, dim3(256), 0, s, m->kind, m->train_xs, m->n,
                     m->d, m->outputscale, m->noise, B, Y, ld, nc, R);
  BTA_LAUNCH_CHECK("kmat_residual_kernel");
  return BTA_OK;
}

__global__ void add_rows_kernel(double* X, const double* D, int64_t ld, int64_t cols) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < cols) X[(int64_t)blockIdx.y * ld + c] += D[(int64_t)blockIdx.y * ld + c];
}

__global__ void fill_double_kernel(double* p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

static int fill(double* p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return BTA_OK;
  hipLaunchKernelGGL(fill_double_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n,
                     v);
  BTA_LAUNCH_CHECK("fill_double_kernel");
  return BTA_OK;
}

// alpha = Linv^T Linv r, computed as two 64-row MFMA products
// (row 0 carries the vector, rows 1..63 are zero).
static int solve_alpha(const double* Linv, int64_t n_pad, const double* r, double* R1,
                       double* T1, double* alpha_row, hipStream_t s) {
  int rc = fill(R1, 64 * n_pad, 0.0, s);
  if (rc) return rc;
  BTA_HIP(hipMemcpyAsync(R1, r, n_pad * sizeof(double), hipMemcpyDeviceToDevice, s));
  // T1[0][i] = sum_k r[k] Linv[i][k]
  rc = gemm(BTA_GEMM_B_LOWER, 64, n_pad, n_pad, 1.0, R1, n_pad, Linv, n_pad, 0.0, nullptr, 0, T1,
            n_pad, s);
  if (rc) return rc;
  // alpha[j] = sum_i T1[0][i] Linv[i][j]
  return gemm(BTA_GEMM_B_KN | BTA_GEMM_B_UPPER, 64, n_pad, n_pad, 1.0, T1, n_pad, Linv, n_pad,
              0.0, nullptr, 0, alpha_row, n_pad, s);
}

size_t gp_cache_ws(int64_t n_pad) {
  WsPlan p;
  p.take(std::max(potrf_ws_bytes(n_pad), trtri_ws_bytes(n_pad)));
  p.take((size_t)n_pad * 8);
  p.take((size_t)64 * n_pad * 8 * 3);
  return p.off;
}

int gp_cache_build(const bta_gp_model* m, const double* y, double* L, double* Linv,
                   double* alpha, int max_tries, double* jitter_used, void* ws, size_t ws_bytes,
                   hipStream_t s) {
  if (!m || !y || !L || !Linv || !alpha || m->n < 1 || m->n_pad < m->n || m->n_pad % 128) {
    set_error("gp_cache_build: invalid model (n=%ld n_pad=%ld)", (long)m->n, (long)m->n_pad);
    return BTA_ERR_INVALID;
  }
  if (!ws || ws_bytes < gp_cache_ws(m->n_pad)) {
    set_error("gp_cache_build: workspace too small");
    return BTA_ERR_WORKSPACE;
  }
  const int64_t np = m->n_pad;
  WsPlan p;
  char* base = (char*)ws;
  const size_t fact_bytes = std::max(potrf_ws_bytes(np), trtri_ws_bytes(np));
  void* wf = base + p.take(fact_bytes);
  double* r = (double*)(base + p.take((size_t)np * 8));
  double* R1 = (double*)(base + p.take((size_t)64 * np * 8 * 3));
  double* T1 = R1 + 64 * np;
  double* A1 = T1 + 64 * np;
  // the factor is of Khat + chol_jitter I (Khat = s k(X, X) + noise I)
  int rc = covar(m->kind, m->train_xs, m->n, m->train_xs, m->n, m->d, m->outputscale,
                 m->noise + m->chol_jitter, false, L, np, np, np, s, /*pad_identity=*/1);
  if (rc) return rc;
  // L^-1 built inside the factorisation's launches when Linv suits the
  // direct-to-LDS loads (16-byte aligned), else by the separate inverse pass
  const bool fused = ((uintptr_t)Linv & 15) == 0;
  rc = potrf(L, np, np, Linv, np, max_tries, 1e-8, jitter_used, wf, fact_bytes, s,
             fused ? POTRF_INVERT : 0);
  if (rc) return rc;
  if (!fused) {
    rc = trtri_lower(L, np, Linv, np, np, wf, fact_bytes, s);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(residual_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, y,
                     m->n, m->mean_constant, r, np);
  BTA_LAUNCH_CHECK("residual_kernel");
  rc = solve_alpha(Linv, np, r, R1, T1, A1, s);
  if (rc) return rc;
  BTA_HIP(hipMemcpyAsync(alpha, A1, np * sizeof(double), hipMemcpyDeviceToDevice, s));
  // refinement against Khat: alpha += (L L^T)^-1 (r - Khat alpha)
  for (int it = 0; it < m->refine_steps; ++it) {
    rc = kmat_residual(m, alpha, r, np, 1, A1, s);
    if (rc) return rc;
    rc = solve_alpha(Linv, np, A1, R1, T1, A1, s);
    if (rc) return rc;
    hipLaunchKernelGGL(add_rows_kernel, dim3((unsigned)((np + 255) / 256), 1), dim3(256), 0, s,
                       alpha, A1, np, np);
    BTA_LAUNCH_CHECK("add_rows_kernel");
  }
  return BTA_OK;
}

struct PostWs {
  int mq_pad;
  int64_t R_pad;
  size_t xs, kx, vt, ka, kbb, sg, arow, total;
  PostWs(const bta_gp_model* m, int64_t batch, int mq) {
    mq_pad = (int)round_up(mq, 64);
    R_pad = batch * mq_pad;
    WsPlan p;
    xs = p.take((size_t)R_pad * m->d * 8);
    kx = p.take((size_t)R_pad * m->n_pad * 8);
    vt = p.take((size_t)R_pad * m->n_pad * 8);
    ka = p.take((size_t)R_pad * 64 * 8);
    kbb = p.take((size_t)batch * mq_pad * mq_pad * 8);
    sg = p.take((size_t)batch * mq_pad * mq_pad * 8);
    arow = p.take((size_t)64 * m->n_pad * 8);
    total = p.off;
  }
};

int gp_posterior(const bta_gp_model* m, const double* X, int64_t batch, int mq, double* mean,
                 double* covar_out, double* var, void* ws, size_t ws_bytes, hipStream_t s) {
  if (!m || !X || batch < 1 || mq < 1 || !m->linv || !m->alpha || batch > 65535) {
    set_error("gp_posterior: invalid arguments");
    return BTA_ERR_INVALID;
  }
  const PostWs w(m, batch, mq);
  if (!ws || ws_bytes < w.total) {
    set_error("gp_posterior: workspace too small");
    return BTA_ERR_WORKSPACE;
  }
  char* base = (char*)ws;
  double* xs = (double*)(base + w.xs);
  double* kx = (double*)(base + w.kx);
  double* vt = (double*)(base + w.vt);
  double* ka = (double*)(base + w.ka);
  double* kbb = (double*)(base + w.kbb);
  double* sg = (double*)(base + w.sg);
  double* arow = (double*)(base + w.arow);
  const int64_t np = m->n_pad;
  const int64_t tot = w.R_pad * m->d;
  hipLaunchKernelGGL(scale_batched_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                     X, batch, mq, w.mq_pad, m->d, m->center, m->lengthscale, xs);
  BTA_LAUNCH_CHECK("scale_batched_kernel");
  int rc = covar(m->kind, xs, w.R_pad, m->train_xs, m->n, m->d, m->outputscale, 0.0, false, kx,
                 np, w.R_pad, np, s);
  if (rc) return rc;
  rc = gemm(BTA_GEMM_B_LOWER, w.R_pad, np, np, 1.0, kx, np, m->linv, np, 0.0, nullptr, 0, vt, np,
            s);
  if (rc) return rc;
  rc = cast_pad<double>(m->alpha, np, 1, np, arow, np, 64, np, s);
  if (rc) return rc;
  rc = gemm(0, w.R_pad, 64, np, 1.0, kx, np, arow, np, 0.0, nullptr, 0, ka, 64, s);
  if (rc) return rc;
  const int64_t mp = w.mq_pad;
  rc = covar(m->kind, xs, mp, xs, mp, m->d, m->outputscale, 0.0, false, kbb, mp, mp, mp, s, 0,
             (int)batch, mp * m->d, mp * m->d, mp * mp);
  if (rc) return rc;
  const double v2 = m->y_std * m->y_std;
  rc = gemm(0, mp, mp, np, -v2, vt, np, vt, np, v2, kbb, mp, sg, mp, s, (int)batch, mp * np,
            mp * np, mp * mp, mp * mp);
  if (rc) return rc;
  dim3 g((unsigned)std::min<int64_t>(64, ((int64_t)mq * mq + 255) / 256), (unsigned)batch);
  hipLaunchKernelGGL(posterior_out_kernel, g, dim3(256), 0, s, ka, (int64_t)64, sg, batch, mq,
                     w.mq_pad, m->y_mean, m->y_std, m->mean_constant, mean, covar_out, var);
  BTA_LAUNCH_CHECK("posterior_out_kernel");
  return BTA_OK;
}

// ---------------------------------------------------------------------------
// Adjoint of the posterior: dX from d mean and d covar.
//   dK_x   = y_std dmean alpha^T - y_std^2 (dC + dC^T) U,   U = V L^-1 = K_x K^-1
//   dK_xx  = y_std^2 dC  (the prior covariance among the batch's own points)
//   dX     = through dk/dX of both, divided by the lengthscale
// ---------------------------------------------------------------------------
// S_b = y_std^2 (dC_b + dC_b^T), zero-padded to mq_pad x mq_pad; dm_pad = y_std
// dmean in the padded row layout (0 where dmean is NULL).
__global__ void post_bwd_prep_kernel(const double* dmean, const double* dcov, int64_t batch, int mq,
                                     int mq_pad, double y_std, double* S, double* dm_pad) {
  const int64_t bi = blockIdx.y;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < mq_pad * mq_pad;
       e += gridDim.x * blockDim.x) {
    const int i = e / mq_pad, j = e % mq_pad;
    double v = 0.0;
    if (dcov && i < mq && j < mq)
      v = y_std * y_std *
          (dcov[(bi * mq + i) * mq + j] + dcov[(bi * mq + j) * mq + i]);
    S[bi * mq_pad * mq_pad + e] = v;
    if (j == 0) dm_pad[bi * mq_pad + i] = (dmean && i < mq) ? y_std * dmean[bi * mq + i] : 0.0;
  }
}

// G[r][k] += dm_pad[r] alpha[k]  (the mean's part of dK_x)
__global__ void add_outer_kernel(double* G, int64_t ldg, const double* dm_pad,
                                 const double* alpha, int64_t n) {
  const int64_t r = blockIdx.y;
  const double a = dm_pad[r];
  if (a == 0.0) return;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    G[r * ldg + k] += a * alpha[k];
}

// extra[r][j] = sum_i2 S_b[i][i2] dk(xs_i, xs_i2)/dxs_i[j] = -sum_i2 S_b[i][i2] h (xs_i - xs_i2)_j
// (scaled coordinates; the row's own term has no gradient)
__global__ void post_direct_grad_kernel(int kind, const double* xs, int64_t batch, int mq,
                                        int mq_pad, int d, double s, const double* S,
                                        double* extra) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= batch * mq_pad * d) return;
  const int j = (int)(e % d);
  const int64_t row = e / d;
  const int64_t bi = row / mq_pad;
  const int i = (int)(row % mq_pad);
  double acc = 0.0;
  if (i < mq) {
    const double* xi = xs + row * d;
    for (int i2 = 0; i2 < mq; ++i2) {
      if (i2 == i) continue;
      const double sv = S[(bi * mq_pad + i) * mq_pad + i2];
      if (sv == 0.0) continue;
      const double* x2 = xs + (bi * mq_pad + i2) * d;
      double r2 = 0.0;
      for (int t = 0; t < d; ++t) {
        const double df = xi[t] - x2[t];
        r2 = fma(df, df, r2);
      }
      acc -= sv * kern_grad_factor(kind, r2, s, nullptr) * (xi[j] - x2[j]);
    }
  }
  extra[e] = acc;
}

// dX[(bi mq + r) d + j] = dXp[(bi mq_pad + r) d + j]
__global__ void unpad_rows_kernel(const double* dXp, int64_t batch, int mq, int mq_pad, int d,
                                  double* dX) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= batch * mq * d) return;
  const int j = (int)(e % d);
  const int64_t row = e / d;
  const int64_t bi = row / mq;
  const int r = (int)(row % mq);
  dX[e] = dXp[(bi * mq_pad + r) * d + j];
}

struct PostBwdWs {
  int mq_pad;
  int64_t R_pad;
  size_t xs, kx, vt, S, dm, extra, dxp, total;
  PostBwdWs(const bta_gp_model* m, int64_t batch, int mq) {
    mq_pad = (int)round_up(mq, 64);
    R_pad = batch * mq_pad;
    WsPlan p;
    xs = p.take((size_t)R_pad * m->d * 8);
    kx = p.take((size_t)R_pad * m->n_pad * 8);
    vt = p.take((size_t)R_pad * m->n_pad * 8);
    S = p.take((size_t)batch * mq_pad * mq_pad * 8);
    dm = p.take((size_t)R_pad * 8);
    extra = p.take((size_t)R_pad * m->d * 8);
    dxp = p.take((size_t)R_pad * m->d * 8);
    total = p.off;
  }
};

int gp_posterior_backward(const bta_gp_model* m, const double* X, int64_t batch, int mq,
                          const double* dmean, const double* dcov, double* dX, void* ws,
                          size_t ws_bytes, hipStream_t s) {
  if (!m || !X || !dX || batch < 1 || mq < 1 || !m->linv || !m->alpha || batch > 65535 ||
      (!dmean && !dcov)) {
    set_error("gp_posterior_backward: invalid arguments");
    return BTA_ERR_INVALID;
  }
  const PostBwdWs w(m, batch, mq);
  if (!ws || ws_bytes < w.total) {
    set_error("gp_posterior_backward: workspace too small");
    return BTA_ERR_WORKSPACE;
  }
  char* base = (char*)ws;
  double* xs = (double*)(base + w.xs);
  double* kx = (double*)(base + w.kx);
  double* vt = (double*)(base + w.vt);
  double* S = (double*)(base + w.S);
  double* dm = (double*)(base + w.dm);
  double* extra = (double*)(base + w.extra);
  double* dxp = (double*)(base + w.dxp);
  const int64_t np = m->n_pad, mp = w.mq_pad;
  const int64_t tot = w.R_pad * m->d;
  hipLaunchKernelGGL(scale_batched_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                     X, batch, mq, w.mq_pad, m->d, m->center, m->lengthscale, xs);
  BTA_LAUNCH_CHECK("scale_batched_kernel");
  int rc = covar(m->kind, xs, w.R_pad, m->train_xs, m->n, m->d, m->outputscale, 0.0, false, kx,
                 np, w.R_pad, np, s);
  if (rc) return rc;
  // V = K_x L^-T into vt, then U = V L^-1 into kx
  rc = gemm(BTA_GEMM_B_LOWER, w.R_pad, np, np, 1.0, kx, np, m->linv, np, 0.0, nullptr, 0, vt, np,
            s);
  if (rc) return rc;
  rc = gemm(BTA_GEMM_B_KN | BTA_GEMM_B_UPPER, w.R_pad, np, np, 1.0, vt, np, m->linv, np, 0.0,
            nullptr, 0, kx, np, s);
  if (rc) return rc;
  {
    const dim3 g((unsigned)std::min<int64_t>(64, (mp * mp + 255) / 256), (unsigned)batch);
    hipLaunchKernelGGL(post_bwd_prep_kernel, g, dim3(256), 0, s, dmean, dcov, batch, mq,
                       w.mq_pad, m->y_std, S, dm);
    BTA_LAUNCH_CHECK("post_bwd_prep_kernel");
  }
  // G = -S U per batch block (into vt), then + y_std dmean alpha^T
  rc = gemm(BTA_GEMM_B_KN, mp, np, mp, -1.0, S, mp, kx, np, 0.0, nullptr, 0, vt, np, s,
            (int)batch, mp * mp, mp * np, 0, mp * np);
  if (rc) return rc;
  if (dmean) {
    const dim3 g((unsigned)std::min<int64_t>(64, (np + 255) / 256), (unsigned)w.R_pad);
    hipLaunchKernelGGL(add_outer_kernel, g, dim3(256), 0, s, vt, np, dm, m->alpha, m->n);
    BTA_LAUNCH_CHECK("add_outer_kernel");
  }
  hipLaunchKernelGGL(post_direct_grad_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     s, m->kind, xs, batch, mq, w.mq_pad, m->d, m->outputscale, S, extra);
  BTA_LAUNCH_CHECK("post_direct_grad_kernel");
  rc = covar_grad(m->kind, xs, w.R_pad, m->train_xs, m->n, m->d, m->outputscale, vt, false, np,
                  extra, m->lengthscale, dxp, s);
  if (rc) return rc;
  const int64_t to = batch * mq * m->d;
  hipLaunchKernelGGL(unpad_rows_kernel, dim3((unsigned)((to + 255) / 256)), dim3(256), 0, s, dxp,
                     batch, mq, w.mq_pad, m->d, dX);
  BTA_LAUNCH_CHECK("unpad_rows_kernel");
  return BTA_OK;
}

struct PrepWs {
  int64_t nbp;
  size_t kb, vb, at, kbb, sg, lbp, arow, ka, mu, fact, mx, total;
  PrepWs(const bta_gp_model* m, int n_b, int S) {
    nbp = round_up(n_b, 64);
    const int64_t np = m->n_pad;
    WsPlan p;
    kb = p.take((size_t)nbp * np * 8);
    vb = p.take((size_t)nbp * np * 8);
    at = p.take((size_t)nbp * np * 8);
    kbb = p.take((size_t)nbp * nbp * 8);
    sg = p.take((size_t)nbp * nbp * 8);
    lbp = p.take((size_t)nbp * nbp * 8);
    arow = p.take((size_t)64 * np * 8);
    ka = p.take((size_t)nbp * 64 * 8);
    mu = p.take((size_t)nbp * 8);
    fact = p.take(potrf_ws_bytes(nbp));
    mx = p.take((MAXABS_BLOCKS + 1) * sizeof(double));
    total = p.off;
  }
};

int qlognei_prepare(const bta_gp_model* m, const bta_qlognei_state* st, const double* Zb,
                    void* linv_out, double* ext, double* Lb, double* best_f, double* jitter_used,
                    float* linv_scale, void* ws,
                    size_t ws_bytes, hipStream_t s) {
  if (!m || !st || !Zb || !ext || !Lb || !best_f || !m->linv || !m->alpha || !st->xb_s ||
      st->n_b < 1 || st->S < 1 || (st->precision != BTA_PREC_FP64 && !linv_out)) {
    set_error("qlognei_prepare: invalid arguments");
    return BTA_ERR_INVALID;
  }
  const PrepWs w(m, st->n_b, st->S);
  if (!ws || ws_bytes < w.total) {
    set_error("qlognei_prepare: workspace too small");
    return BTA_ERR_WORKSPACE;
  }
  char* base = (char*)ws;
  double* kb = (double*)(base + w.kb);
  double* vb = (double*)(base + w.vb);
  double* at = (double*)(base + w.at);
  double* kbb = (double*)(base + w.kbb);
  double* sg = (double*)(base + w.sg);
  double* arow = (double*)(base + w.arow);
  double* ka = (double*)(base + w.ka);
  const int64_t np = m->n_pad, nbp = w.nbp;
  const int n_b = st->n_b;
  int rc = covar(m->kind, st->xb_s, n_b, m->train_xs, m->n, m->d, m->outputscale, 0.0, false, kb,
                 np, nbp, np, s);
  if (rc) return rc;
  rc = gemm(BTA_GEMM_B_LOWER, nbp, np, np, 1.0, kb, np, m->linv, np, 0.0, nullptr, 0, vb, np, s);
  if (rc) return rc;
  rc = gemm(BTA_GEMM_B_KN | BTA_GEMM_B_UPPER, nbp, np, np, 1.0, vb, np, m->linv, np, 0.0,
            nullptr, 0, at, np, s);
  if (rc) return rc;
  rc = cast_pad<double>(m->alpha, np, 1, np, arow, np, 64, np, s);
  if (rc) return rc;
  rc = gemm(0, nbp, 64, np, 1.0, kb, np, arow, np, 0.0, nullptr, 0, ka, 64, s);
  if (rc) return rc;
  // Sigma_bb = y_std^2 (k(Xb,Xb) - Vb Vb^T), identity-padded for the factorisation
  rc = covar(m->kind, st->xb_s, n_b, st->xb_s, n_b, m->d, m->outputscale, 0.0, false, kbb, nbp,
             nbp, nbp, s, /*pad_identity=*/1);
  if (rc) return rc;
  const double v2 = m->y_std * m->y_std;
  rc = gemm(0, nbp, nbp, np, -v2, vb, np, vb, np, v2, kbb, nbp, sg, nbp, s);
  if (rc) return rc;
  // refinement of A^T = K(X_b, X) Khat^-1 against Khat (vb is free now, and kb
  // after the residual): A^T += (K(X_b, X) - A^T Khat) L^-T L^-1
  for (int it = 0; it < m->refine_steps; ++it) {
    if (it > 0) {
      rc = covar(m->kind, st->xb_s, n_b, m->train_xs, m->n, m->d, m->outputscale, 0.0, false, kb,
                 np, nbp, np, s);
      if (rc) return rc;
    }
    rc = kmat_residual(m, at, kb, np, n_b, vb, s);
    if (rc) return rc;
    rc = gemm(BTA_GEMM_B_LOWER, nbp, np, np, 1.0, vb, np, m->linv, np, 0.0, nullptr, 0, kb, np,
              s);
    if (rc) return rc;
    rc = gemm(BTA_GEMM_B_KN | BTA_GEMM_B_UPPER, nbp, np, np, 1.0, kb, np, m->linv, np, 0.0,
              nullptr, 0, vb, np, s);
    if (rc) return rc;
    hipLaunchKernelGGL(add_rows_kernel, dim3((unsigned)((np + 255) / 256), (unsigned)n_b),
                       dim3(256), 0, s, at, vb, np, np);
    BTA_LAUNCH_CHECK("add_rows_kernel");
  }
  rc = potrf(sg, nbp, nbp, nullptr, 0, st->max_tries, st->jitter, jitter_used, base + w.fact,
             potrf_ws_bytes(nbp), s);
  if (rc) return rc;
  rc = cast_pad<double>(sg, nbp, n_b, n_b, Lb, n_b, n_b, n_b, s);
  if (rc) return rc;
  // mu_b (original scale) into the workspace, then best_f
  double* mu = (double*)(base + w.mu);
  {
    // mu[j] = y_mean + y_std (c + ka[j*64])
    rc = cast_pad<double>(ka, 64, nbp, 1, mu, 1, nbp, 1, s);
    if (rc) return rc;
  }
  rc = launch_baseline_best_f(mu, m->y_mean, m->y_std, m->mean_constant, Lb, Zb, n_b, st->S,
                              best_f, s);
  if (rc) return rc;
  // E[c] = [alpha_c | A[c][0..n_b)]  (FP64, n_pad x (1 + n_b))
  hipLaunchKernelGGL(build_ext_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s,
                     m->alpha, at, np, n_b, ext);
  BTA_LAUNCH_CHECK("build_ext_kernel");
  // the power-of-two scale that puts L^-1 in the FP16 range
  float lscale = 1.0f;
  if (linv_scale || st->precision == BTA_PREC_MIXED) {
    double* mx = (double*)(base + w.mx);
    const unsigned nblk = MAXABS_BLOCKS;
    hipLaunchKernelGGL(max_abs_kernel, dim3(nblk), dim3(256), 0, s, m->linv, np * np, mx);
    BTA_LAUNCH_CHECK("max_abs_kernel");
    hipLaunchKernelGGL(max_abs_kernel, dim3(1), dim3(256), 0, s, mx, (int64_t)nblk,
                       mx + nblk);
    BTA_LAUNCH_CHECK("max_abs_kernel");
    double h = 0.0;
    BTA_HIP(hipMemcpyAsync(&h, mx + nblk, sizeof(double), hipMemcpyDeviceToHost, s));
    BTA_HIP(hipStreamSynchronize(s));
    lscale = (h > 0.0 && std::isfinite(h)) ? (float)exp2(-ceil(log2(h))) : 1.0f;
    if (linv_scale) *linv_scale = lscale;
  }
  // L^-1 in the state's precision: FP32, or the four FP16 planes
  // [hi(c L^-1) | lo(c L^-1) | hi(c L^-T) | lo(c L^-T)] with c = lscale
  if (st->precision == BTA_PREC_MIXED) {
    _Float16* pl = (_Float16*)linv_out;
    const size_t pn = (size_t)np * np;
    rc = split_f16(true, m->linv, np, np, np, np, np, lscale, false, pl, pl + pn, np, s);
    if (rc) return rc;
    rc = split_f16(true, m->linv, np, np, np, np, np, lscale, true, pl + 2 * pn, pl + 3 * pn, np,
                   s);
  } else if (st->precision == BTA_PREC_MIXED_F32) {
    rc = cast_pad<float>(m->linv, np, np, np, (float*)linv_out, np, np, np, s);
  } else if (linv_out && linv_out != (const void*)m->linv) {
    rc = cast_pad<double>(m->linv, np, np, np, (double*)linv_out, np, np, np, s);
  }
  return rc;
}

}  // namespace bta

extern "C" size_t bta_gp_cache_workspace_size(int64_t n_pad) { return bta::gp_cache_ws(n_pad); }

extern "C" int bta_gp_cache_build(const bta_gp_model* m, const double* y, double* L,
                                  double* Linv, double* alpha, int32_t max_tries,
                                  double* jitter_used, void* ws, size_t ws_bytes,
                                  bta_stream_t stream) {
  return bta::gp_cache_build(m, y, L, Linv, alpha, max_tries, jitter_used, ws, ws_bytes,
                             (hipStream_t)stream);
}

extern "C" size_t bta_gp_posterior_workspace_size(const bta_gp_model* m, int64_t batch,
                                                  int32_t mq) {
  if (!m || batch < 1 || mq < 1) return 0;
  return bta::PostWs(m, batch, mq).total;
}

extern "C" int bta_gp_posterior(const bta_gp_model* m, const double* X, int64_t batch,
                                int32_t mq, double* mean, double* covar, double* var, void* ws,
                                size_t ws_bytes, bta_stream_t stream) {
  return bta::gp_posterior(m, X, batch, mq, mean, covar, var, ws, ws_bytes,
                           (hipStream_t)stream);
}

extern "C" size_t bta_gp_posterior_backward_workspace_size(const bta_gp_model* m,
                                                          int64_t batch, int32_t mq) {
  if (!m || batch < 1 || mq < 1) return 0;
  return bta::PostBwdWs(m, batch, mq).total;
}

extern "C" int bta_gp_posterior_backward(const bta_gp_model* m, const double* X, int64_t batch,
                                         int32_t mq, const double* dmean, const double* dcov,
                                         double* dX, void* ws, size_t ws_bytes,
                                         bta_stream_t stream) {
  return bta::gp_posterior_backward(m, X, batch, mq, dmean, dcov, dX, ws, ws_bytes,
                                    (hipStream_t)stream);
}

extern "C" size_t bta_qlognei_prepare_workspace_size(const bta_gp_model* m, int32_t n_b,
                                                     int32_t S) {
  if (!m || n_b < 1 || S < 1) return 0;
  return bta::PrepWs(m, n_b, S).total;
}

extern "C" int bta_qlognei_prepare(const bta_gp_model* m, const bta_qlognei_state* st,
                                   const double* Zb, void* linv_out, double* ext, double* Lb,
                                   double* best_f, double* jitter_used, float* linv_scale,
                                   void* ws, size_t ws_bytes, bta_stream_t stream) {
  return bta::qlognei_prepare(m, st, Zb, linv_out, ext, Lb, best_f, jitter_used, linv_scale, ws,
                              ws_bytes, (hipStream_t)stream);
}
