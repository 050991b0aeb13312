// Exact marginal log likelihood and its hyperparameter adjoint.
//
//   K = s k(X,X; l) + noise I,  L = chol(K),  t = L^-1 (y - c),  a = L^-T t
//   log p(y) = -1/2 (|t|^2 + 2 sum log L_ii + n log 2 pi)
//   d log p / d theta = 1/2 sum_ij W_ij dK_ij/dtheta,   W = a a^T - K^-1
//
// K^-1 = L^-T L^-1 is never stored: mll_trace_kernel computes each lower
// tile of it on the MFMA GEMM mainloop (both operands read straight from L^-1,
// the K-range starting at the tile's larger block index), then the epilogue
// rebuilds k and dk/dtheta for the tile's point pairs from the scaled inputs
// staged in LDS and folds W_ij dK_ij/dtheta into per-worker sums.  The
// n^2 x (d+2) derivative tensor is never formed either.
// Reference loss: optim/closures/model_closures.py:125-145 (ExactMarginal-
// LogLikelihood, gpytorch); derivative: Rasmussen & Williams (2006) eq. 5.9.
#include "bta_common.h"
#include "gemm.h"
#include "gemm_tile.h"
#include "kernels.h"

#include <algorithm>
#include <vector>

namespace bta {

struct TraceArgs {
  const double* linv;
  int64_t ld;
  const double* xs;
  const double* alpha;
  double* partial;  // gridDim.x x (MD + 2)
  int n_pad, n, d, kind;
  double s;
};

template <int B, int BK, int MD>
__global__ __launch_bounds__(256) void mll_trace_kernel(TraceArgs a) {
  typedef Cfg<double, B, B, BK, true, true> C;
  static_assert(2 * C::STAGE >= 2 * B * MD, "xs staging must fit in the mainloop LDS");
  __shared__ __attribute__((aligned(16))) double lds[2 * C::STAGE];
  const int MT = a.n_pad / B;
  const int total = MT * (MT + 1) / 2;
  const int W = gridDim.x, w = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wy = wv >> 1, wx = wv & 1, l16 = lane & 15, gq = lane >> 4;
  const int d = a.d;
  double acc_os = 0.0, acc_noise = 0.0, acc_ls[MD];
#pragma unroll
  for (int j = 0; j < MD; ++j) acc_ls[j] = 0.0;

  for (int jj = 0;; ++jj) {
    const int t = (jj & 1) ? (jj + 1) * W - 1 - w : jj * W + w;
    if (t >= total) break;
    int tm, tn;
    decode_tile(t, BTA_GEMM_C_LOWER, MT, MT, tm, tn);
    const int kb = tm * B;  // tm >= tn: both L^-1 columns vanish above row tm*B
    v4d acc[C::MI][C::NI];
    gemm_tile<double, B, B, BK, true, true>(a.linv, a.ld, a.linv, a.ld, tm * B, tn * B, kb,
                                            a.n_pad, lds, acc);
    // stage the tile's scaled rows and columns (mainloop LDS is free now)
    double* xr = lds;
    double* xc = lds + B * MD;
    for (int e = threadIdx.x; e < B * d; e += 256) {
      const int r = e / d, j = e % d;
      const int gi = tm * B + r, gj = tn * B + r;
      xr[r * MD + j] = gi < a.n ? a.xs[(int64_t)gi * d + j] : 0.0;
      xc[r * MD + j] = gj < a.n ? a.xs[(int64_t)gj * d + j] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ni = 0; ni < C::NI; ++ni) {
      const int cl = wx * (B / 2) + ni * 16 + l16;
      const int gj = tn * B + cl;
      if (gj >= a.n) continue;
      const double aj = a.alpha[gj];
      double xj[MD];
#pragma unroll
      for (int j = 0; j < MD; ++j) xj[j] = j < d ? xc[cl * MD + j] : 0.0;
#pragma unroll
      for (int mi = 0; mi < C::MI; ++mi)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int rl = wy * (B / 2) + mi * 16 + Mfma<double>::row(gq, q);
          const int gi = tm * B + rl;
          if (gi >= a.n || gi < gj) continue;
          const double wij = a.alpha[gi] * aj - acc[mi][ni][q];
          const double f = gi == gj ? 1.0 : 2.0;
          double r2 = 0.0;
          double dl[MD];
#pragma unroll
          for (int j = 0; j < MD; ++j) {
            const double df = j < d ? xr[rl * MD + j] - xj[j] : 0.0;
            dl[j] = df * df;
            r2 += dl[j];
          }
          double kv;
          const double h = kern_grad_factor(a.kind, r2, a.s, &kv);
          const double fw = f * wij;
          acc_os = fma(fw, kv, acc_os);
          if (gi == gj) acc_noise += wij;
          const double fwh = fw * h;
#pragma unroll
          for (int j = 0; j < MD; ++j) acc_ls[j] = fma(fwh, dl[j], acc_ls[j]);
        }
    }
    __syncthreads();
  }
  // per-worker reduction
  double* red = lds;
  double v = wave_sum(acc_os);
  if (lane == 0) red[wv * (MD + 2)] = v;
  v = wave_sum(acc_noise);
  if (lane == 0) red[wv * (MD + 2) + 1] = v;
#pragma unroll
  for (int j = 0; j < MD; ++j) {
    v = wave_sum(acc_ls[j]);
    if (lane == 0) red[wv * (MD + 2) + 2 + j] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < MD + 2; e += 256)
    a.partial[(int64_t)w * (MD + 2) + e] =
        red[e] + red[(MD + 2) + e] + red[2 * (MD + 2) + e] + red[3 * (MD + 2) + e];
}

// out[0] = sum t_i^2, out[1] = sum_{i<n} log L_ii, out[2] = sum_{i<n} alpha_i,
// out[3..] = sum over workers of the trace partials (MD + 2 values).
__global__ __launch_bounds__(256) void mll_finish_kernel(const double* t, const double* L,
                                                         int64_t ld, const double* alpha,
                                                         int64_t n, const double* partial,
                                                         int W, int np2, double* out) {
  __shared__ double red[4 * 3];
  double q = 0.0, lg = 0.0, sa = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    q += t[i] * t[i];
    lg += log(L[i * ld + i]);
    sa += alpha[i];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  q = wave_sum(q);
  lg = wave_sum(lg);
  sa = wave_sum(sa);
  if (lane == 0) {
    red[wv * 3] = q;
    red[wv * 3 + 1] = lg;
    red[wv * 3 + 2] = sa;
  }
  __syncthreads();
  if (threadIdx.x < 3)
    out[threadIdx.x] = red[threadIdx.x] + red[3 + threadIdx.x] + red[6 + threadIdx.x] +
                       red[9 + threadIdx.x];
  for (int e = threadIdx.x; e < np2; e += 256) {
    double s = 0.0;
    for (int w = 0; w < W; ++w) s += partial[(int64_t)w * np2 + e];
    out[3 + e] = s;
  }
}

__global__ void residual_pad_kernel(const double* y, int64_t n, double c, double* R,
                                    int64_t n_pad) {
  // R is 64 x n_pad; row 0 = y - c (zero padded), other rows zero
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 64 * n_pad) return;
  R[i] = i < n ? y[i] - c : 0.0;
}

struct MllWs {
  int64_t n_pad;
  int W, MD;
  size_t ls, center, xs, L, Linv, fact, R, T, Aw, partial, out, total;
  MllWs(int64_t n, int d) {
    n_pad = round_up(n, 128);
    MD = d <= 16 ? 16 : 32;
    W = 2 * num_cus();
    WsPlan p;
    ls = p.take(d * 8);
    center = p.take(d * 8);
    xs = p.take((size_t)n_pad * d * 8);
    L = p.take((size_t)n_pad * n_pad * 8);
    Linv = p.take((size_t)n_pad * n_pad * 8);
    fact = p.take(std::max(potrf_ws_bytes(n_pad), trtri_ws_bytes(n_pad)));
    R = p.take((size_t)64 * n_pad * 8);
    T = p.take((size_t)64 * n_pad * 8);
    Aw = p.take((size_t)64 * n_pad * 8);
    partial = p.take((size_t)W * (MD + 2) * 8);
    out = p.take((size_t)(MD + 8) * 8);
    total = p.off;
  }
};

int mll_value_grad(int kind, const double* X, const double* y, int64_t n, int d,
                   const double* ls_host, double s, double noise, double c, int max_tries,
                   double* out_host, void* ws, size_t ws_bytes, hipStream_t st) {
  if (!X || !y || !ls_host || !out_host || n < 2 || d < 1 || d > 32 || !(s > 0) ||
      !(noise >= 0)) {
    set_error("mll_value_grad: invalid arguments");
    return BTA_ERR_INVALID;
  }
  const MllWs w(n, d);
  if (!ws || ws_bytes < w.total) {
    set_error("mll_value_grad: workspace %zu < %zu", ws_bytes, w.total);
    return BTA_ERR_WORKSPACE;
  }
  char* base = (char*)ws;
  double* ls = (double*)(base + w.ls);
  double* center = (double*)(base + w.center);
  double* xs = (double*)(base + w.xs);
  double* L = (double*)(base + w.L);
  double* Linv = (double*)(base + w.Linv);
  void* fact = base + w.fact;
  const size_t fact_bytes = std::max(potrf_ws_bytes(w.n_pad), trtri_ws_bytes(w.n_pad));
  double* R = (double*)(base + w.R);
  double* T = (double*)(base + w.T);
  double* Aw = (double*)(base + w.Aw);
  double* partial = (double*)(base + w.partial);
  double* out = (double*)(base + w.out);
  const int64_t np = w.n_pad;

  BTA_HIP(hipMemcpyAsync(ls, ls_host, d * sizeof(double), hipMemcpyHostToDevice, st));
  int rc = scale_inputs(X, n, d, nullptr, ls, center, xs, np, st);
  if (rc) return rc;
  rc = covar(kind, xs, n, xs, n, d, s, noise, false, L, np, np, np, st, /*pad_identity=*/1);
  if (rc) return rc;
  double jit = 0.0;
  rc = potrf(L, np, np, Linv, np, max_tries, 1e-8, &jit, fact, fact_bytes, st);
  if (rc) return rc;
  rc = trtri_lower(L, np, Linv, np, np, fact, fact_bytes, st);
  if (rc) return rc;
  hipLaunchKernelGGL(residual_pad_kernel, dim3((unsigned)((64 * np + 255) / 256)), dim3(256), 0,
                     st, y, n, c, R, np);
  BTA_LAUNCH_CHECK("residual_pad_kernel");
  rc = gemm(BTA_GEMM_B_LOWER, 64, np, np, 1.0, R, np, Linv, np, 0.0, nullptr, 0, T, np, st);
  if (rc) return rc;
  rc = gemm(BTA_GEMM_B_KN | BTA_GEMM_B_UPPER, 64, np, np, 1.0, T, np, Linv, np, 0.0, nullptr, 0,
            Aw, np, st);
  if (rc) return rc;
  TraceArgs ta;
  ta.linv = Linv;
  ta.ld = np;
  ta.xs = xs;
  ta.alpha = Aw;
  ta.partial = partial;
  ta.n_pad = (int)np;
  ta.n = (int)n;
  ta.d = d;
  ta.kind = kind;
  ta.s = s;
  const int MT128 = (int)(np / 128), MT64 = (int)(np / 64);
  int W;
  if (d <= 16) {
    const int total = MT128 * (MT128 + 1) / 2;
    W = std::min(w.W, total);
    hipLaunchKernelGGL((mll_trace_kernel<128, 8, 16>), dim3(W), dim3(256), 0, st, ta);
  } else {
    const int total = MT64 * (MT64 + 1) / 2;
    W = std::min(w.W, total);
    hipLaunchKernelGGL((mll_trace_kernel<64, 16, 32>), dim3(W), dim3(256), 0, st, ta);
  }
  BTA_LAUNCH_CHECK("mll_trace_kernel");
  hipLaunchKernelGGL(mll_finish_kernel, dim3(1), dim3(256), 0, st, T, L, np, Aw, n, partial, W,
                     w.MD + 2, out);
  BTA_LAUNCH_CHECK("mll_finish_kernel");
  std::vector<double> h(3 + w.MD + 2);
  BTA_HIP(hipMemcpyAsync(h.data(), out, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  BTA_HIP(hipStreamSynchronize(st));
  const double LOG2PI = 1.8378770664093454836;
  out_host[0] = -0.5 * (h[0] + 2.0 * h[1] + (double)n * LOG2PI);
  out_host[1] = 0.5 * h[3] / s;
  out_host[2] = 0.5 * h[4];
  out_host[3] = h[2];
  for (int j = 0; j < d; ++j) out_host[4 + j] = 0.5 * h[5 + j] / ls_host[j];
  return BTA_OK;
}

}  // namespace bta

extern "C" size_t bta_mll_workspace_size(int64_t n, int32_t d) {
  if (n < 2 || d < 1 || d > 32) return 0;
  return bta::MllWs(n, d).total;
}

extern "C" int bta_mll_value_grad(int32_t kind, const double* X, const double* y, int64_t n,
                                  int32_t d, const double* lengthscale, double outputscale,
                                  double noise, double constant, int32_t max_tries, double* out,
                                  void* ws, size_t ws_bytes, bta_stream_t stream) {
  return bta::mll_value_grad(kind, X, y, n, d, lengthscale, outputscale, noise, constant,
                             max_tries, out, ws, ws_bytes, (hipStream_t)stream);
}
