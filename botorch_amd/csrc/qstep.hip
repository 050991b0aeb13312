// Elementwise stages of one qLogNEI value+grad step on FP64 MFMA.
//
// qstep_covar_kernel (one 64-row x 64-column tile per block, 16 rows a wave):
//   xs  = (X - center) / l                          (scaled query rows)
//   K_x = s k(xs, X_train)                          (stored FP64, FP32, or as FP16
//                                                    hi/lo planes of 2^e K_x)
//   Ppart[tile] = K_x[:, tile] [alpha | A][tile, :] (FP64 partial products)
// qstep_grad_kernel (same tiling), with U = K_x K^-1 and the q x q blocks gq:
//   g_rc = sum_j gq[r][j] U[q floor(r/q) + j][c] + dext[r] . [alpha|A][c]
//   w_rc = -g_rc h(|xs_r - t_c|)                    (dK_x/dxs = -h (xs - t))
//   Dpart[tile][r] = xs_r sum_c w_rc - sum_c w_rc t_c
// qstep_grad_finish_kernel: dX = (sum_tiles Dpart + direct) / l.
//
// Every contraction is a v_mfma_f64_16x16x4f64: the cross term t_c . x_r of
// r^2 = |t|^2 + |x|^2 - 2 t.x (the expansion gpytorch uses, clamped at 0), the
// [alpha|A] products, the restart-block mixing of U, dext [alpha|A]^T, and the
// contraction of w with [T | 1].  The cross term is formed transposed, D[c][r],
// so that each lane's four accumulator values K[r = lane%16][c = lane/16 + 4q]
// are already the A operand of the next product over c (k index = lane/16 + 4q);
// the VALU only evaluates the kernel functions.  Partials over column tiles are
// summed in a fixed order by their consumer: no atomics, deterministic results.
#include "bta_common.h"
#include "kernels.h"
#include "qstep_body.h"

namespace bta {

// One model's kernels: a tile per block.
template <int KD, int MODE, int KIND>
__global__ __launch_bounds__(256) void qstep_covar_kernel(QsCovarArgs a) {
  qstep_covar_body<KD, MODE, KIND>(a, blockIdx.x, blockIdx.y);
}

template <int KD, typename TU, int KIND>
__global__ __launch_bounds__(256) void qstep_grad_kernel(QsGradArgs a) {
  qstep_grad_body<KD, TU, KIND>(a, blockIdx.x, blockIdx.y);
}

__global__ void qstep_grad_finish_kernel(const double* __restrict__ Dpart, int nct, int64_t R,
                                         int d, const double* __restrict__ extra,
                                         const double* __restrict__ ls, double* __restrict__ dX) {
  qstep_finish_body(Dpart, nct, R, d, extra, ls, dX,
                    (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
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

template <int KD, int KIND>
static void covar_launch_k(const QsCovarArgs& a, int mode, dim3 grid, hipStream_t st) {
  if (mode == 0)
    hipLaunchKernelGGL((qstep_covar_kernel<KD, 0, KIND>), grid, dim3(256), 0, st, a);
  else if (mode == 1)
    hipLaunchKernelGGL((qstep_covar_kernel<KD, 1, KIND>), grid, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((qstep_covar_kernel<KD, 2, KIND>), grid, dim3(256), 0, st, a);
}

template <int KD>
static void covar_launch(const QsCovarArgs& a, int mode, dim3 grid, hipStream_t st) {
  if (a.kind == BTA_KERNEL_MATERN52)
    covar_launch_k<KD, BTA_KERNEL_MATERN52>(a, mode, grid, st);
  else
    covar_launch_k<KD, BTA_KERNEL_RBF>(a, mode, grid, st);
}

int qstep_covar(const QsCovarArgs& a, int mode, hipStream_t st) {
  int rc = qs_check("qstep_covar", a.d, a.ne, a.n, a.n_pad, a.R, a.R_pad);
  if (rc) return rc;
  if (a.ldk < a.n_pad || a.ldk % 8 || (mode == 2 && !a.Klo) || mode < 0 || mode > 2) {
    set_error("qstep_covar: invalid output layout");
    return BTA_ERR_INVALID;
  }
  const dim3 grid((unsigned)(a.n_pad / QS_CT), (unsigned)(a.R_pad / QS_RT));
  switch ((a.d + 3) / 4) {
    case 1: covar_launch<1>(a, mode, grid, st); break;
    case 2: covar_launch<2>(a, mode, grid, st); break;
    case 3: covar_launch<3>(a, mode, grid, st); break;
    case 4: covar_launch<4>(a, mode, grid, st); break;
    case 5: covar_launch<5>(a, mode, grid, st); break;
    case 6: covar_launch<6>(a, mode, grid, st); break;
    case 7: covar_launch<7>(a, mode, grid, st); break;
    default: covar_launch<8>(a, mode, grid, st); break;
  }
  BTA_LAUNCH_CHECK("qstep_covar_kernel");
  return BTA_OK;
}

template <int KD, int KIND>
static void grad_launch_k(const QsGradArgs& a, bool u_f32, dim3 grid, hipStream_t st) {
  if (u_f32)
    hipLaunchKernelGGL((qstep_grad_kernel<KD, float, KIND>), grid, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((qstep_grad_kernel<KD, double, KIND>), grid, dim3(256), 0, st, a);
}

template <int KD>
static void grad_launch(const QsGradArgs& a, bool u_f32, dim3 grid, hipStream_t st) {
  if (a.kind == BTA_KERNEL_MATERN52)
    grad_launch_k<KD, BTA_KERNEL_MATERN52>(a, u_f32, grid, st);
  else
    grad_launch_k<KD, BTA_KERNEL_RBF>(a, u_f32, grid, st);
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
  switch ((a.d + 3) / 4) {
    case 1: grad_launch<1>(a, u_f32, grid, st); break;
    case 2: grad_launch<2>(a, u_f32, grid, st); break;
    case 3: grad_launch<3>(a, u_f32, grid, st); break;
    case 4: grad_launch<4>(a, u_f32, grid, st); break;
    case 5: grad_launch<5>(a, u_f32, grid, st); break;
    case 6: grad_launch<6>(a, u_f32, grid, st); break;
    case 7: grad_launch<7>(a, u_f32, grid, st); break;
    default: grad_launch<8>(a, u_f32, grid, st); break;
  }
  BTA_LAUNCH_CHECK("qstep_grad_kernel");
  const int64_t tot = a.R * a.d;
  hipLaunchKernelGGL(qstep_grad_finish_kernel, dim3((unsigned)((4 * tot + 255) / 256)), dim3(256),
                     0, st, a.Dpart, (int)(a.n_pad / QS_CT), a.R, a.d, extra, ls, dX);
  BTA_LAUNCH_CHECK("qstep_grad_finish_kernel");
  return BTA_OK;
}

}  // namespace bta
