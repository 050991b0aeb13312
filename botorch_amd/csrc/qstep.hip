// Elementwise stages of one qLogNEI value+grad step, tiled through LDS.
//
// qstep_covar_kernel (one 32-row x 64-column tile per block):
//   xs  = (X - center) / l                          (scaled query rows, FP64)
//   K_x = s k(xs, X_train)                          (stored FP64, FP32, or as FP16
//                                                    hi/lo planes of 2^e K_x)
//   Ppart[tile] = K_x[:, tile] [alpha | A][tile, :] (FP64 partial products)
// qstep_grad_kernel (same tiling), with U = K_x K^-1 and the q x q blocks gq:
//   g_rc = sum_j gq[r][j] U[q floor(r/q) + j][c] + dext[r] . [alpha|A][c]
//   w_rc = -g_rc h(|xs_r - t_c|)                    (dK_x/dxs = -h (xs - t))
//   Dpart[tile][r] = sum_{c in tile} w_rc (xs_r - t_c)
// qstep_grad_finish_kernel: dX = (sum_tiles Dpart + direct) / l.
//
// A block stages its rows, its 64 training points and their [alpha|A] rows
// (16 columns per pass) in LDS with coalesced loads; each thread evaluates 8
// kernel entries (one column, 8 rows) in FP64, and the row sums over the
// tile's columns run from an LDS copy of the tile (8 lanes per row,
// shuffle-reduced).  Partials over column tiles are summed in a fixed order by
// their consumer: no atomics, deterministic results.
#include "bta_common.h"
#include "kernels.h"

namespace bta {

constexpr int QS_RT = 32;           // rows per tile
constexpr int QS_CT = QS_COL_TILE;  // columns (training points) per tile
constexpr int QS_MAXQ = 8;

template <int MD>
struct QsStage {
  double xr[QS_RT][MD + 1];
  double tc[QS_CT][MD + 1];
  double ec[QS_CT][QS_NE_MAX];
  double kt[QS_RT][QS_CT + 1];
};

template <int MD>
__device__ __forceinline__ double qs_dist2(const double* a, const double* b) {
  double r2 = 0.0;
#pragma unroll
  for (int j = 0; j < MD; ++j) {
    const double df = a[j] - b[j];
    r2 = fma(df, df, r2);
  }
  return r2;
}

// sum over groups of 8 adjacent lanes
__device__ __forceinline__ double sum8(double v) {
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  return v;
}

template <int MD>
__device__ __forceinline__ void qs_stage_t(QsStage<MD>& sm, const double* T, int64_t n, int d,
                                           int64_t c0) {
  for (int e = threadIdx.x; e < QS_CT * MD; e += 256) {
    const int c = e / MD, j = e % MD;
    sm.tc[c][j] = (j < d && c0 + c < n) ? T[(c0 + c) * d + j] : 0.0;
  }
}

// columns [j0, j0 + nj) of [alpha | A] for the tile's training points
template <int MD>
__device__ __forceinline__ void qs_stage_e(QsStage<MD>& sm, const double* E, int64_t n, int ne,
                                           int j0, int nj, int64_t c0) {
  for (int e = threadIdx.x; e < QS_CT * nj; e += 256) {
    const int c = e / nj, j = e % nj;
    sm.ec[c][j] = c0 + c < n ? E[(c0 + c) * ne + j0 + j] : 0.0;
  }
}

// mode: 0 = FP64 K, 1 = FP32 K, 2 = FP16 hi/lo planes of K * kscale
template <int MD, int MODE>
__global__ __launch_bounds__(256) void qstep_covar_kernel(QsCovarArgs a) {
  __shared__ QsStage<MD> sm;
  const int tid = threadIdx.x;
  const int ct = blockIdx.x;
  const int d = a.d;
  const int64_t r0 = (int64_t)blockIdx.y * QS_RT, c0 = (int64_t)ct * QS_CT;
  for (int e = tid; e < QS_RT * MD; e += 256) {
    const int r = e / MD, j = e % MD;
    double v = 0.0;
    if (j < d && r0 + r < a.R) v = (a.X[(r0 + r) * d + j] - a.center[j]) / a.ls[j];
    sm.xr[r][j] = v;
    if (ct == 0 && j < d) a.xs_out[(r0 + r) * d + j] = v;
  }
  qs_stage_t<MD>(sm, a.T, a.n, d, c0);
  qs_stage_e<MD>(sm, a.E, a.n, a.ne, 0, min(a.ne, QS_NE_MAX), c0);
  __syncthreads();
  const int c = tid & 63, rg = tid >> 6;  // a wave covers the 64 columns of 8 rows
  const bool cvalid = c0 + c < a.n;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int r = rg + 4 * i;
    const int64_t gr = r0 + r;
    double kv = 0.0;
    if (gr < a.R && cvalid) kv = kern_value(a.kind, qs_dist2<MD>(sm.xr[r], sm.tc[c]), a.s);
    sm.kt[r][c] = kv;
    const int64_t o = gr * a.ldk + c0 + c;
    if (MODE == 0) {
      ((double*)a.K)[o] = kv;
    } else if (MODE == 1) {
      ((float*)a.K)[o] = (float)kv;
    } else {
      const float x = (float)(kv * a.kscale);
      const _Float16 h = (_Float16)x;
      ((_Float16*)a.K)[o] = h;
      a.Klo[o] = (_Float16)(x - (float)h);
    }
  }
  __syncthreads();
  // Ppart[ct][r][j] = sum_c kt[r][c] ec[c][j]: 8 lanes per row, 8 columns each
  const int r = tid >> 3, part = tid & 7;
  const int64_t gr = r0 + r;
  for (int j0 = 0; j0 < a.ne; j0 += QS_NE_MAX) {
    const int nj = min(a.ne - j0, QS_NE_MAX);
    if (j0 > 0) {
      __syncthreads();
      qs_stage_e<MD>(sm, a.E, a.n, a.ne, j0, nj, c0);
      __syncthreads();
    }
    for (int j = 0; j < nj; ++j) {
      double s = 0.0;
#pragma unroll
      for (int cc = 0; cc < 8; ++cc)
        s = fma(sm.kt[r][part * 8 + cc], sm.ec[part * 8 + cc][j], s);
      s = sum8(s);
      if (part == 0 && gr < a.R) a.Ppart[((int64_t)ct * a.R + gr) * a.ne + j0 + j] = s;
    }
  }
}

template <int MD, typename TU>
__global__ __launch_bounds__(256) void qstep_grad_kernel(QsGradArgs a) {
  __shared__ QsStage<MD> sm;
  __shared__ double de[QS_RT][QS_NE_MAX];
  __shared__ double gqs[QS_RT][QS_MAXQ];
  const int tid = threadIdx.x;
  const int ct = blockIdx.x;
  const int d = a.d, q = a.q;
  const int64_t r0 = (int64_t)blockIdx.y * QS_RT, c0 = (int64_t)ct * QS_CT;
  for (int e = tid; e < QS_RT * MD; e += 256) {
    const int r = e / MD, j = e % MD;
    sm.xr[r][j] = (j < d && r0 + r < a.R) ? a.xs[(r0 + r) * d + j] : 0.0;
  }
  for (int e = tid; e < QS_RT * q; e += 256) {
    const int r = e / q, j = e % q;
    gqs[r][j] = r0 + r < a.R ? a.gq[(r0 + r) * q + j] : 0.0;
  }
  qs_stage_t<MD>(sm, a.T, a.n, d, c0);
  __syncthreads();
  const TU* U = (const TU*)a.U;
  const int c = tid & 63, rg = tid >> 6;
  const bool cvalid = c0 + c < a.n;
  double g[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int r = rg + 4 * i;
    const int64_t gr = r0 + r;
    double v = 0.0;
    if (gr < a.R && cvalid) {
      const int64_t rb = gr - gr % q;
      for (int j = 0; j < q; ++j) v = fma(gqs[r][j], (double)U[(rb + j) * a.ldu + c0 + c], v);
    }
    g[i] = v;
  }
  for (int j0 = 0; j0 < a.ne; j0 += QS_NE_MAX) {
    const int nj = min(a.ne - j0, QS_NE_MAX);
    if (j0 > 0) __syncthreads();
    qs_stage_e<MD>(sm, a.E, a.n, a.ne, j0, nj, c0);
    for (int e = tid; e < QS_RT * nj; e += 256) {
      const int r = e / nj, j = e % nj;
      de[r][j] = r0 + r < a.R ? a.dext[(r0 + r) * a.ne + j0 + j] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = rg + 4 * i;
      double v = g[i];
      for (int j = 0; j < nj; ++j) v = fma(de[r][j], sm.ec[c][j], v);
      g[i] = v;
    }
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int r = rg + 4 * i;
    const int64_t gr = r0 + r;
    double w = 0.0;
    if (gr < a.R && cvalid)
      w = -g[i] * kern_grad_factor(a.kind, qs_dist2<MD>(sm.xr[r], sm.tc[c]), a.s, nullptr);
    sm.kt[r][c] = w;
  }
  __syncthreads();
  // Dpart[ct][r][j] = sum_c w[r][c] (xs_r[j] - t_c[j]): 8 lanes per row
  const int r = tid >> 3, part = tid & 7;
  const int64_t gr = r0 + r;
  double w[8];
#pragma unroll
  for (int cc = 0; cc < 8; ++cc) w[cc] = sm.kt[r][part * 8 + cc];
  for (int j = 0; j < d; ++j) {
    const double xj = sm.xr[r][j];
    double acc = 0.0;
#pragma unroll
    for (int cc = 0; cc < 8; ++cc) acc = fma(w[cc], xj - sm.tc[part * 8 + cc][j], acc);
    acc = sum8(acc);
    if (part == 0 && gr < a.R) a.Dpart[((int64_t)ct * a.R + gr) * d + j] = acc;
  }
}

__global__ void qstep_grad_finish_kernel(const double* __restrict__ Dpart, int nct, int64_t R,
                                         int d, const double* __restrict__ extra,
                                         const double* __restrict__ ls, double* __restrict__ dX) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= R * d) return;
  double acc = extra ? extra[e] : 0.0;
  for (int t = 0; t < nct; ++t) acc += Dpart[(int64_t)t * R * d + e];
  dX[e] = acc / ls[(int)(e % d)];
}

static int qs_check(const char* what, int d, int ne, int64_t n, int64_t n_pad, int64_t R,
                    int64_t R_pad) {
  if (d < 1 || d > 32 || ne < 1 || n_pad % QS_CT || R_pad % QS_RT || n > n_pad || R > R_pad ||
      R < 1 || n < 1) {
    set_error("%s: unsupported sizes (d=%d ne=%d n=%lld R=%lld)", what, d, ne, (long long)n,
              (long long)R);
    return BTA_ERR_UNSUPPORTED;
  }
  return BTA_OK;
}

int qstep_covar(const QsCovarArgs& a, int mode, hipStream_t st) {
  int rc = qs_check("qstep_covar", a.d, a.ne, a.n, a.n_pad, a.R, a.R_pad);
  if (rc) return rc;
  if (a.ldk < a.n_pad || (mode == 2 && !a.Klo) || mode < 0 || mode > 2) {
    set_error("qstep_covar: invalid output layout");
    return BTA_ERR_INVALID;
  }
  const dim3 grid((unsigned)(a.n_pad / QS_CT), (unsigned)(a.R_pad / QS_RT));
#define QS_LAUNCH(MD_)                                                                   \
  do {                                                                                   \
    if (mode == 0)                                                                       \
      hipLaunchKernelGGL((qstep_covar_kernel<MD_, 0>), grid, dim3(256), 0, st, a);       \
    else if (mode == 1)                                                                  \
      hipLaunchKernelGGL((qstep_covar_kernel<MD_, 1>), grid, dim3(256), 0, st, a);       \
    else                                                                                 \
      hipLaunchKernelGGL((qstep_covar_kernel<MD_, 2>), grid, dim3(256), 0, st, a);       \
  } while (0)
  if (a.d <= 8) QS_LAUNCH(8);
  else if (a.d <= 16) QS_LAUNCH(16);
  else QS_LAUNCH(32);
#undef QS_LAUNCH
  BTA_LAUNCH_CHECK("qstep_covar_kernel");
  return BTA_OK;
}

int qstep_grad(const QsGradArgs& a, bool u_f32, const double* extra, const double* ls,
               double* dX, hipStream_t st) {
  int rc = qs_check("qstep_grad", a.d, a.ne, a.n, a.n_pad, a.R, a.R_pad);
  if (rc) return rc;
  if (a.q < 1 || a.q > QS_MAXQ || a.R % a.q || a.ldu < a.n_pad) {
    set_error("qstep_grad: q must be in 1..8 and divide the row count");
    return BTA_ERR_INVALID;
  }
  const dim3 grid((unsigned)(a.n_pad / QS_CT), (unsigned)(a.R_pad / QS_RT));
#define QG_LAUNCH(MD_)                                                                      \
  do {                                                                                      \
    if (u_f32)                                                                              \
      hipLaunchKernelGGL((qstep_grad_kernel<MD_, float>), grid, dim3(256), 0, st, a);       \
    else                                                                                    \
      hipLaunchKernelGGL((qstep_grad_kernel<MD_, double>), grid, dim3(256), 0, st, a);      \
  } while (0)
  if (a.d <= 8) QG_LAUNCH(8);
  else if (a.d <= 16) QG_LAUNCH(16);
  else QG_LAUNCH(32);
#undef QG_LAUNCH
  BTA_LAUNCH_CHECK("qstep_grad_kernel");
  const int64_t tot = a.R * a.d;
  hipLaunchKernelGGL(qstep_grad_finish_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     st, a.Dpart, (int)(a.n_pad / QS_CT), a.R, a.d, extra, ls, dX);
  BTA_LAUNCH_CHECK("qstep_grad_finish_kernel");
  return BTA_OK;
}

}  // namespace bta
