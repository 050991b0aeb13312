// ARD Matern-5/2 / RBF covariance construction and its adjoint.
//
// Inputs are pre-scaled ((x - center) / lengthscale, in FP64), so a covariance
// element is one d-term squared distance plus the radial function.  Distances
// come from direct coordinate differences, not the ||a||^2 + ||b||^2 - 2ab
// expansion: at d <= 32 the work is tiny next to the contraction that follows,
// and direct differences carry no cancellation near the diagonal.
//
// bta_covar: a 256-thread block covers ROWS query rows x 256 training columns;
// each thread keeps its training point in registers, the block's query rows
// sit in LDS (broadcast reads), and every store is a coalesced row segment.
#include "bta_common.h"
#include "kernels.h"

namespace bta {

constexpr int MAXD = 32;
constexpr int EXT_MAX = 257;  // 1 + n_b with n_b <= 256

__global__ __launch_bounds__(256) void col_mean_kernel(const double* X, int64_t n, int d,
                                                       double* mean) {
  const int j = blockIdx.x;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += X[i * d + j];
  __shared__ double red[4];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) mean[j] = (red[0] + red[1] + red[2] + red[3]) / (double)n;
}

__global__ void scale_kernel(const double* X, int64_t n, int d, const double* center,
                             const double* ls, double* Xs, int64_t n_pad) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_pad * d) return;
  int64_t i = idx / d;
  int j = (int)(idx % d);
  Xs[idx] = i < n ? (X[idx] - center[j]) / ls[j] : 0.0;
}

template <typename TO, int ROWS>
__global__ __launch_bounds__(256) void covar_kernel(const double* __restrict__ X1s, int64_t n1,
                                                    const double* __restrict__ X2s, int64_t n2,
                                                    int d, int kind, double s, double diag_add,
                                                    int pad_identity, TO* __restrict__ K,
                                                    int64_t ldk, int64_t n1_pad, int64_t n2_pad,
                                                    int64_t sX1, int64_t sX2, int64_t sK) {
  __shared__ double xr[ROWS * MAXD];
  X1s += blockIdx.z * sX1;
  X2s += blockIdx.z * sX2;
  K += blockIdx.z * sK;
  const int64_t r0 = (int64_t)blockIdx.y * ROWS;
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (int i = threadIdx.x; i < ROWS * d; i += 256) {
    int64_t r = r0 + i / d;
    xr[i] = r < n1 ? X1s[r * d + i % d] : 0.0;
  }
  __syncthreads();
  if (c >= n2_pad) return;
  double t[MAXD];
  const bool cvalid = c < n2;
#pragma unroll
  for (int j = 0; j < MAXD; ++j) t[j] = (j < d && cvalid) ? X2s[c * d + j] : 0.0;
#pragma unroll 2
  for (int rr = 0; rr < ROWS; ++rr) {
    const int64_t r = r0 + rr;
    if (r >= n1_pad) break;
    double v;
    if (r < n1 && cvalid) {
      double r2 = 0.0;
#pragma unroll
      for (int j = 0; j < MAXD; ++j) {
        if (j < d) {
          double df = xr[rr * d + j] - t[j];
          r2 = fma(df, df, r2);
        }
      }
      v = kern_value(kind, r2, s);
      if (r == c) v += diag_add;
    } else {
      v = (pad_identity && r == c) ? 1.0 : 0.0;
    }
    K[r * ldk + c] = (TO)v;
  }
}

// Symmetric case for the factorisation: K = s k(Xs, Xs) + diag_add I on and
// below the diagonal only (the Cholesky reads nothing above it), identity on
// the padding, kernel kind fixed at compile time.  Blocks wholly above the
// diagonal exit at once.  The block's ROWS row points are staged in LDS (read
// as wave-wide broadcasts); each lane loads its own column point's coordinates
// straight into registers (staging the 256 column points in LDS as well
// measured 64.5 against 59.3 us at n = 4096, d = 16).  When zero2 is set, the
// first block also zeroes those two ints (the status pair of the
// factorisation that follows).
template <int KIND, int MD, int ROWS>
__global__ __launch_bounds__(256) void covar_sym_lower_kernel(const double* __restrict__ Xs,
                                                              int64_t n, int d, double s,
                                                              double diag_add,
                                                              double* __restrict__ K,
                                                              int64_t ldk, int64_t n_pad,
                                                              int* zero2) {
  __shared__ double xr[ROWS * MD];
  if (zero2 && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 2) zero2[threadIdx.x] = 0;
  const int64_t r0 = (int64_t)blockIdx.y * ROWS;
  const int64_t c0 = (int64_t)blockIdx.x * 256;
  const int64_t c = c0 + threadIdx.x;
  if (c0 > r0 + ROWS - 1) return;  // block-uniform
  for (int i = threadIdx.x; i < ROWS * MD; i += 256) {
    const int64_t r = r0 + i / MD;
    const int j = i % MD;
    xr[i] = (r < n && j < d) ? Xs[r * d + j] : 0.0;
  }
  __syncthreads();
  if (c >= n_pad) return;
  double t[MD];
  const bool cvalid = c < n;
#pragma unroll
  for (int j = 0; j < MD; ++j) t[j] = (cvalid && j < d) ? Xs[c * d + j] : 0.0;
#pragma unroll 2
  for (int rr = 0; rr < ROWS; ++rr) {
    const int64_t r = r0 + rr;
    if (r >= n_pad) break;
    if (c > r) continue;
    double v;
    if (r < n && cvalid) {
      double r2 = 0.0;
#pragma unroll
      for (int j = 0; j < MD; ++j) {
        const double df = xr[rr * MD + j] - t[j];  // zero past d
        r2 = fma(df, df, r2);
      }
      v = kern_value_k<KIND>(r2, s);
      if (r == c) v += diag_add;
    } else {
      v = r == c ? 1.0 : 0.0;
    }
    K[r * ldk + c] = v;
  }
}

int covar_sym_lower(int kind, const double* Xs, int64_t n, int d, double s, double diag_add,
                    double* K, int64_t ldk, int64_t n_pad, hipStream_t st, int* zero2) {
  if (d < 1 || d > MAXD || n_pad < n || ldk < n_pad || n < 0) {
    set_error("covar_sym_lower: bad sizes (d=%d n=%ld)", d, (long)n);
    return BTA_ERR_INVALID;
  }
  if (n_pad == 0) return BTA_OK;
  constexpr int ROWS = 16;
  dim3 grid((unsigned)((n_pad + 255) / 256), (unsigned)((n_pad + ROWS - 1) / ROWS));
  if (grid.y > 65535u) {
    set_error("covar_sym_lower: too many rows");
    return BTA_ERR_UNSUPPORTED;
  }
  const bool m52 = kind == BTA_KERNEL_MATERN52;
  if (d <= 16) {
    if (m52)
      hipLaunchKernelGGL((covar_sym_lower_kernel<BTA_KERNEL_MATERN52, 16, ROWS>), grid,
                         dim3(256), 0, st, Xs, n, d, s, diag_add, K, ldk, n_pad, zero2);
    else
      hipLaunchKernelGGL((covar_sym_lower_kernel<BTA_KERNEL_RBF, 16, ROWS>), grid, dim3(256), 0,
                         st, Xs, n, d, s, diag_add, K, ldk, n_pad, zero2);
  } else {
    if (m52)
      hipLaunchKernelGGL((covar_sym_lower_kernel<BTA_KERNEL_MATERN52, MAXD, ROWS>), grid,
                         dim3(256), 0, st, Xs, n, d, s, diag_add, K, ldk, n_pad, zero2);
    else
      hipLaunchKernelGGL((covar_sym_lower_kernel<BTA_KERNEL_RBF, MAXD, ROWS>), grid, dim3(256),
                         0, st, Xs, n, d, s, diag_add, K, ldk, n_pad, zero2);
  }
  BTA_LAUNCH_CHECK("covar_sym_lower_kernel");
  return BTA_OK;
}

// dX1[r] = (sum_c G[r][c] * (-h(r,c)) (x1s_r - x2s_c) + extra[r]) / ls
template <typename TG, int RB, int MD>
__global__ __launch_bounds__(256) void covar_grad_kernel(
    const double* __restrict__ X1s, int64_t n1, const double* __restrict__ X2s, int64_t n2,
    int d, int kind, double s, const TG* __restrict__ G, int64_t ldg,
    const double* __restrict__ extra, const double* __restrict__ ls, double* __restrict__ dX1,
    const double* __restrict__ dext, const double* __restrict__ E, int ne) {
  __shared__ double xr[RB * MD];
  __shared__ double red[4][RB * MD];
  __shared__ double de[RB * EXT_MAX];
  const int64_t r0 = (int64_t)blockIdx.x * RB;
  for (int i = threadIdx.x; i < RB * d; i += 256) {
    int64_t r = r0 + i / d;
    xr[i] = r < n1 ? X1s[r * d + i % d] : 0.0;
  }
  for (int i = threadIdx.x; i < RB * ne; i += 256) {
    int64_t r = r0 + i / ne;
    de[i] = r < n1 ? dext[r * ne + i % ne] : 0.0;
  }
  __syncthreads();
  double acc[RB][MD];
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int j = 0; j < MD; ++j) acc[a][j] = 0.0;
  for (int64_t c = threadIdx.x; c < n2; c += 256) {
    double t[MD];
#pragma unroll
    for (int j = 0; j < MD; ++j) t[j] = j < d ? X2s[c * d + j] : 0.0;
#pragma unroll
    for (int a = 0; a < RB; ++a) {
      const int64_t r = r0 + a;
      if (r >= n1) break;
      double r2 = 0.0;
#pragma unroll
      for (int j = 0; j < MD; ++j) {
        if (j < d) {
          double df = xr[a * d + j] - t[j];
          r2 = fma(df, df, r2);
        }
      }
      double gv = (double)G[r * ldg + c];
      for (int j = 0; j < ne; ++j) gv = fma(de[a * ne + j], E[c * ne + j], gv);
      const double w = -gv * kern_grad_factor(kind, r2, s, nullptr);
#pragma unroll
      for (int j = 0; j < MD; ++j)
        if (j < d) acc[a][j] = fma(w, xr[a * d + j] - t[j], acc[a][j]);
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int j = 0; j < MD; ++j) {
      if (j < d) {
        double v = wave_sum(acc[a][j]);
        if (lane == 0) red[wv][a * MD + j] = v;
      }
    }
  __syncthreads();
  for (int i = threadIdx.x; i < RB * d; i += 256) {
    int a = i / d, j = i % d;
    int64_t r = r0 + a;
    if (r >= n1) continue;
    double v = red[0][a * MD + j] + red[1][a * MD + j] + red[2][a * MD + j] +
               red[3][a * MD + j];
    if (extra) v += extra[r * d + j];
    dX1[r * d + j] = v / ls[j];
  }
}


// P[r][j0 + jj] = sum_c k(X1s[r], X2s[c]) E[c][j0 + jj]  (FP64): the products of
// the cross-covariance rows with the few dense columns [alpha | K^-1 k(X, X_b)].
template <int RB, int NEC>
__global__ __launch_bounds__(256) void covar_ext_kernel(const double* __restrict__ X1s,
                                                        int64_t n1, const double* __restrict__ X2s,
                                                        int64_t n2, int d, int kind, double s,
                                                        const double* __restrict__ E, int ne,
                                                        double* __restrict__ P) {
  __shared__ double xr[RB * MAXD];
  __shared__ double red[4][RB * NEC];
  const int64_t r0 = (int64_t)blockIdx.x * RB;
  const int j0 = blockIdx.y * NEC;
  for (int i = threadIdx.x; i < RB * d; i += 256) {
    int64_t r = r0 + i / d;
    xr[i] = r < n1 ? X1s[r * d + i % d] : 0.0;
  }
  __syncthreads();
  double acc[RB][NEC];
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int j = 0; j < NEC; ++j) acc[a][j] = 0.0;
  for (int64_t c = threadIdx.x; c < n2; c += 256) {
    double t[MAXD];
#pragma unroll
    for (int j = 0; j < MAXD; ++j) t[j] = j < d ? X2s[c * d + j] : 0.0;
    double e[NEC];
#pragma unroll
    for (int j = 0; j < NEC; ++j) e[j] = j0 + j < ne ? E[c * ne + j0 + j] : 0.0;
#pragma unroll
    for (int a = 0; a < RB; ++a) {
      double r2 = 0.0;
#pragma unroll
      for (int j = 0; j < MAXD; ++j) {
        if (j < d) {
          double df = xr[a * d + j] - t[j];
          r2 = fma(df, df, r2);
        }
      }
      const double kv = kern_value(kind, r2, s);
#pragma unroll
      for (int j = 0; j < NEC; ++j) acc[a][j] = fma(kv, e[j], acc[a][j]);
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int j = 0; j < NEC; ++j) {
      double v = wave_sum(acc[a][j]);
      if (lane == 0) red[wv][a * NEC + j] = v;
    }
  __syncthreads();
  for (int i = threadIdx.x; i < RB * NEC; i += 256) {
    const int a = i / NEC, j = i % NEC;
    const int64_t r = r0 + a;
    if (r < n1 && j0 + j < ne)
      P[r * ne + j0 + j] = red[0][i] + red[1][i] + red[2][i] + red[3][i];
  }
}

int covar_ext(int kind, const double* X1s, int64_t n1, const double* X2s, int64_t n2, int d,
              double s, const double* E, int ne, double* P, hipStream_t st) {
  if (d < 1 || d > MAXD || n1 < 1 || n2 < 1 || ne < 1) {
    set_error("covar_ext: bad sizes");
    return BTA_ERR_INVALID;
  }
  constexpr int RB = 4, NEC = 4;
  dim3 grid((unsigned)((n1 + RB - 1) / RB), (unsigned)((ne + NEC - 1) / NEC));
  hipLaunchKernelGGL((covar_ext_kernel<RB, NEC>), grid, dim3(256), 0, st, X1s, n1, X2s, n2, d,
                     kind, s, E, ne, P);
  BTA_LAUNCH_CHECK("covar_ext_kernel");
  return BTA_OK;
}

int scale_inputs(const double* X, int64_t n, int d, const double* center, const double* ls,
                 double* center_out, double* Xs, int64_t n_pad, hipStream_t st) {
  if (d < 1 || d > MAXD || n < 1 || n_pad < n) {
    set_error("scale_inputs: need 1 <= d <= %d and n_pad >= n >= 1", MAXD);
    return BTA_ERR_INVALID;
  }
  const double* c = center;
  if (!c) {
    if (!center_out) {
      set_error("scale_inputs: center or center_out required");
      return BTA_ERR_INVALID;
    }
    hipLaunchKernelGGL(col_mean_kernel, dim3(d), dim3(256), 0, st, X, n, d, center_out);
    BTA_LAUNCH_CHECK("col_mean_kernel");
    c = center_out;
  } else if (center_out && center_out != center) {
    BTA_HIP(hipMemcpyAsync(center_out, center, d * sizeof(double), hipMemcpyDeviceToDevice, st));
  }
  int64_t tot = n_pad * d;
  hipLaunchKernelGGL(scale_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, X, n,
                     d, c, ls, Xs, n_pad);
  BTA_LAUNCH_CHECK("scale_kernel");
  return BTA_OK;
}

int covar(int kind, const double* X1s, int64_t n1, const double* X2s, int64_t n2, int d,
          double s, double diag_add, bool f32, void* K, int64_t ldk, int64_t n1_pad,
          int64_t n2_pad, hipStream_t st, int pad_identity, int batch, int64_t sX1,
          int64_t sX2, int64_t sK) {
  if (d < 1 || d > MAXD || n1_pad < n1 || n2_pad < n2 || ldk < n2_pad || n1 < 0 || n2 < 0) {
    set_error("covar: bad sizes (d=%d n1=%ld n2=%ld)", d, (long)n1, (long)n2);
    return BTA_ERR_INVALID;
  }
  if (n1_pad == 0 || n2_pad == 0) return BTA_OK;
  constexpr int ROWS = 8;
  dim3 grid((unsigned)((n2_pad + 255) / 256), (unsigned)((n1_pad + ROWS - 1) / ROWS),
            (unsigned)batch);
  if (grid.y > 65535u || batch < 1 || batch > 65535) {
    set_error("covar: too many rows");
    return BTA_ERR_UNSUPPORTED;
  }
  if (f32)
    hipLaunchKernelGGL((covar_kernel<float, ROWS>), grid, dim3(256), 0, st, X1s, n1, X2s, n2, d,
                       kind, s, diag_add, pad_identity, (float*)K, ldk, n1_pad, n2_pad, sX1,
                       sX2, sK);
  else
    hipLaunchKernelGGL((covar_kernel<double, ROWS>), grid, dim3(256), 0, st, X1s, n1, X2s, n2,
                       d, kind, s, diag_add, pad_identity, (double*)K, ldk, n1_pad, n2_pad,
                       sX1, sX2, sK);
  BTA_LAUNCH_CHECK("covar_kernel");
  return BTA_OK;
}

template <int RB, int MD>
int covar_grad_launch(int kind, const double* X1s, int64_t n1, const double* X2s, int64_t n2,
                      int d, double s, const void* G, bool g_f32, int64_t ldg,
                      const double* extra, const double* ls, double* dX1, hipStream_t st,
                      const double* dext, const double* E, int ne);

int covar_grad(int kind, const double* X1s, int64_t n1, const double* X2s, int64_t n2, int d,
               double s, const void* G, bool g_f32, int64_t ldg, const double* extra,
               const double* ls, double* dX1, hipStream_t st, const double* dext,
               const double* E, int ne) {
  if (d < 1 || d > MAXD || n1 < 0 || n2 < 0 || ldg < n2 || ne < 0 || ne > EXT_MAX ||
      (ne > 0 && (!dext || !E))) {
    set_error("covar_grad: bad sizes");
    return BTA_ERR_INVALID;
  }
  if (n1 == 0) return BTA_OK;
  if (d <= 8)
    return covar_grad_launch<4, 8>(kind, X1s, n1, X2s, n2, d, s, G, g_f32, ldg, extra, ls,
                                     dX1, st, dext, E, ne);
  if (d <= 16)
    return covar_grad_launch<2, 16>(kind, X1s, n1, X2s, n2, d, s, G, g_f32, ldg, extra, ls,
                                     dX1, st, dext, E, ne);
  return covar_grad_launch<1, 32>(kind, X1s, n1, X2s, n2, d, s, G, g_f32, ldg, extra, ls,
                                     dX1, st, dext, E, ne);
}

template <int RB, int MD>
int covar_grad_launch(int kind, const double* X1s, int64_t n1, const double* X2s, int64_t n2,
                      int d, double s, const void* G, bool g_f32, int64_t ldg,
                      const double* extra, const double* ls, double* dX1, hipStream_t st,
                      const double* dext, const double* E, int ne) {
  dim3 grid((unsigned)((n1 + RB - 1) / RB));
  if (g_f32)
    hipLaunchKernelGGL((covar_grad_kernel<float, RB, MD>), grid, dim3(256), 0, st, X1s, n1, X2s,
                       n2, d, kind, s, (const float*)G, ldg, extra, ls, dX1, dext, E, ne);
  else
    hipLaunchKernelGGL((covar_grad_kernel<double, RB, MD>), grid, dim3(256), 0, st, X1s, n1,
                       X2s, n2, d, kind, s, (const double*)G, ldg, extra, ls, dX1, dext, E, ne);
  BTA_LAUNCH_CHECK("covar_grad_kernel");
  return BTA_OK;
}

}  // namespace bta

extern "C" int bta_scale_inputs(const double* X, int64_t n, int32_t d, const double* center,
                                const double* lengthscale, double* center_out, double* Xs,
                                int64_t n_pad, bta_stream_t stream) {
  if (!X || !lengthscale || !Xs) {
    bta::set_error("bta_scale_inputs: null pointer");
    return BTA_ERR_INVALID;
  }
  return bta::scale_inputs(X, n, d, center, lengthscale, center_out, Xs, n_pad,
                           (hipStream_t)stream);
}

extern "C" int bta_covar(int32_t kind, const double* X1s, int64_t n1, const double* X2s,
                         int64_t n2, int32_t d, double outputscale, double diag_add,
                         int32_t out_f32, void* K, int64_t ldk, int64_t n1_pad, int64_t n2_pad,
                         bta_stream_t stream) {
  if (!X1s || !X2s || !K || (kind != BTA_KERNEL_MATERN52 && kind != BTA_KERNEL_RBF)) {
    bta::set_error("bta_covar: null pointer or unknown kernel");
    return BTA_ERR_INVALID;
  }
  return bta::covar(kind, X1s, n1, X2s, n2, d, outputscale, diag_add, out_f32 != 0, K, ldk,
                    n1_pad, n2_pad, (hipStream_t)stream, diag_add != 0.0, 1, 0, 0, 0);
}

extern "C" int bta_covar_grad(int32_t kind, const double* X1s, int64_t n1, const double* X2s,
                              int64_t n2, int32_t d, double outputscale, const void* G,
                              int32_t g_f32, int64_t ldg, const double* extra,
                              const double* lengthscale, double* dX1, bta_stream_t stream) {
  if (!X1s || !X2s || !G || !lengthscale || !dX1 ||
      (kind != BTA_KERNEL_MATERN52 && kind != BTA_KERNEL_RBF)) {
    bta::set_error("bta_covar_grad: null pointer or unknown kernel");
    return BTA_ERR_INVALID;
  }
  return bta::covar_grad(kind, X1s, n1, X2s, n2, d, outputscale, G, g_f32 != 0, ldg, extra,
                         lengthscale, dX1, (hipStream_t)stream, nullptr, nullptr, 0);
}
