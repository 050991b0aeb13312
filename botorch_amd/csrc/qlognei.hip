// qLogNEI forward + backward, fused per restart.
//
// One step (R = b q query rows, n training points):
//   qstep_covar   xs, K_x = s k(xs, X_train) and the FP64 partial products
//                 P = K_x [alpha | A]  (A = K^-1 K(X_train, X_b))
//   GEMM          V = K_x L^-T         (FP64, FP32, or split-FP16 planes)
//   GEMM          U = V L^-1 = K_x K^-1 (gradient only; independent of the
//                 MC stage, so it runs before it)
//   qlognei_mc    per restart, below
//   qstep_grad    dX from U, gq, dext and dK/dxs (qstep.hip)
//
// Everything between the products and the input gradient is small and done
// in LDS by one 256-thread workgroup per restart b (qlognei_mc_kernel), or per
// two restarts, two waves each, for the mixed modes' FP32 sample arithmetic:
//
//   Sigma_qq = s^2 (K(X,X) - V V^T)        Sigma_qb = s^2 (K(X,X_b) - K_x A)
//   mu       = y_mean + s (c + K_x alpha)                       (s = y_std)
//   Lbl = Sigma_qb Lb^-T;  Lbr = chol(Sigma_qq - Lbl Lbl^T)  (jitter escalation)
//   Y[s,i] = mu_i + [Lbl | Lbr]_i . Z[s]                       (cached root)
//   v = log_fatplus(Y - best_f, tau_relu); u = fatmax_q(v, tau_max)
//   value = logmeanexp_S(u)
//
// followed by the closed-form adjoint of every step (softmax weights of the
// logmeanexp, the fat-max Pareto derivative with the amax tie split, the
// fat-softplus derivative, Z-contractions, the q x q Cholesky backward, the
// triangular solve with Lb).  The adjoint of V V^T enters as the q x q block
// gq = -2 s^2 gA of each restart (dK_x = gq U for its rows), the adjoint of P
// as dext[r] = [s dmu_r | -s^2 dSigma_qb[r]], and the direct input gradient
// from K(X,X) and K(X,X_b) as dxs_extra.  References: utils/low_rank.py:84-172,
// acquisition/logei.py:376-391,688-715, utils/safe_math.py:146-191,298-355,
// acquisition/monte_carlo.py:268-290.
#include "bta_common.h"
#include "gemm.h"
#include "gemm_h2_body.h"
#include "kernels.h"
#include "qlognei_step.h"

#include <cstdlib>
#include <mutex>

namespace bta {

template <int Q, typename TV, int NR = 1>
__global__ __launch_bounds__(256) void qlognei_mc_kernel(McArgs a) {
  extern __shared__ double sm_dyn[];
  mc_body<Q, TV, NR>(a, blockIdx.x, sm_dyn);
}

// One launch for the adjoint GEMM U = V L^-1 (workgroups 0 .. W-1, split-FP16,
// 64-column tiles) and the MC stage (workgroups W and on, NR restarts each):
// both need only V, and the MC workgroups take the CU slots the GEMM workers
// leave free, with no hand-off between queues.  The MC stage uses the first
// NR McLayout areas of the GEMM's LDS ring.

template <int Q, typename TV, int NR>
__global__ __launch_bounds__(256, 2) void adjoint_mc_kernel(H2Args g, int W, McArgs a) {
  __shared__ __attribute__((aligned(1024))) unsigned char lds[FUSED_LDS];
  if ((int)blockIdx.x < W)
    gemm_h2_body<64, false>(g, blockIdx.x, W, lds);
  else
    mc_body<Q, TV, NR>(a, (int)blockIdx.x - W, (double*)lds);
}

// BTA_SEQUENTIAL_ADJOINT=1 in the environment makes every call take the
// sequential path (adjoint GEMM, then the MC kernel): the fused launch runs the
// same two bodies on the same V, so both paths give identical results, which
// is how the fused launch is tested.
static bool fused_disabled() {
  const char* e = getenv("BTA_SEQUENTIAL_ADJOINT");
  return e && e[0] == '1';
}


// Launches the fused kernel when it applies (q <= 5, mixed sample arithmetic,
// the GEMM planned with 64-column tiles, the MC layout within MC_LDS_MAX);
// returns -1 (nothing launched) otherwise.
static int launch_adjoint_mc(const H2Args& args, const McArgs& a, int q, int64_t b, bool f32,
                             hipStream_t s) {
  int bn = 0, W = 0, G = 1;
  const int rc = gemm_h2_plan(args, &bn, &W, &G);
  if (rc) return rc;
  H2Args g = args;
  g.xcd_groups = G;
  const McLayout lay(q, a.n_b, a.S);
  if (!f32 || q > 5 || bn != 64 || g.seg_chunks > 0 || fused_disabled() ||
      (size_t)lay.total * sizeof(double) > MC_LDS_MAX || b < 1 || b > (1 << 24))
    return -1;
  if (a.nr == 2) {
    const dim3 grid((unsigned)(W + (b + 1) / 2));
    switch (q) {
      case 1: hipLaunchKernelGGL((adjoint_mc_kernel<1, float, 2>), grid, dim3(256), 0, s, g, W, a); break;
      case 2: hipLaunchKernelGGL((adjoint_mc_kernel<2, float, 2>), grid, dim3(256), 0, s, g, W, a); break;
      case 3: hipLaunchKernelGGL((adjoint_mc_kernel<3, float, 2>), grid, dim3(256), 0, s, g, W, a); break;
      case 4: hipLaunchKernelGGL((adjoint_mc_kernel<4, float, 2>), grid, dim3(256), 0, s, g, W, a); break;
      default: hipLaunchKernelGGL((adjoint_mc_kernel<5, float, 2>), grid, dim3(256), 0, s, g, W, a); break;
    }
  } else {
    const dim3 grid((unsigned)(W + b));
    switch (q) {
      case 1: hipLaunchKernelGGL((adjoint_mc_kernel<1, float, 1>), grid, dim3(256), 0, s, g, W, a); break;
      case 2: hipLaunchKernelGGL((adjoint_mc_kernel<2, float, 1>), grid, dim3(256), 0, s, g, W, a); break;
      case 3: hipLaunchKernelGGL((adjoint_mc_kernel<3, float, 1>), grid, dim3(256), 0, s, g, W, a); break;
      case 4: hipLaunchKernelGGL((adjoint_mc_kernel<4, float, 1>), grid, dim3(256), 0, s, g, W, a); break;
      default: hipLaunchKernelGGL((adjoint_mc_kernel<5, float, 1>), grid, dim3(256), 0, s, g, W, a); break;
    }
  }
  BTA_LAUNCH_CHECK("adjoint_mc_kernel");
  return BTA_OK;
}

template <int Q>
static int launch_mc_q(const McArgs& a, int64_t b, bool f32, bool split, hipStream_t s) {
  const McLayout lay(Q, a.n_b, a.S);
  const size_t bytes = (size_t)lay.total * sizeof(double);
  if (bytes > MC_LDS_MAX) {
    set_error("qlognei: LDS need %zu bytes exceeds 64 KiB (q=%d n_b=%d S=%d)", bytes, Q, a.n_b,
              a.S);
    return BTA_ERR_UNSUPPORTED;
  }
  if (a.nr == 2)
    hipLaunchKernelGGL((qlognei_mc_kernel<Q, float, 2>), dim3((unsigned)((b + 1) / 2)), dim3(256),
                       2 * bytes, s, a);
  else if (f32)
    hipLaunchKernelGGL((qlognei_mc_kernel<Q, float>), dim3((unsigned)b), dim3(256), bytes, s, a);
  else
    hipLaunchKernelGGL((qlognei_mc_kernel<Q, double>), dim3((unsigned)b), dim3(256), bytes, s,
                       a);
  BTA_LAUNCH_CHECK("qlognei_mc_kernel");
  return BTA_OK;
}

static int launch_mc(const McArgs& a, int q, int64_t b, bool f32, bool split, hipStream_t s) {
  switch (q) {
    case 1: return launch_mc_q<1>(a, b, f32, split, s);
    case 2: return launch_mc_q<2>(a, b, f32, split, s);
    case 3: return launch_mc_q<3>(a, b, f32, split, s);
    case 4: return launch_mc_q<4>(a, b, f32, split, s);
    case 5: return launch_mc_q<5>(a, b, f32, split, s);
    case 6: return launch_mc_q<6>(a, b, f32, split, s);
    case 7: return launch_mc_q<7>(a, b, f32, split, s);
    case 8: return launch_mc_q<8>(a, b, f32, split, s);
  }
  set_error("qlognei: q must be in 1..8 (got %d)", q);
  return BTA_ERR_UNSUPPORTED;
}

// ---------------------------------------------------------------------------
// Workspace plan for one value+grad call.
// ---------------------------------------------------------------------------
// kx holds K_x (FP64, FP32, or its two FP16 planes) and is then reused for
// U = V L^-1 (same element size); vh holds V's FP16 planes (split mode).
struct QlWs {
  int64_t R, R_pad;
  int ne, parts;  // parts: column-tile partials of P and of the input gradient
  size_t xs, kx, vt, vh, dxe, P, dext, gq, Dpart, total;
  QlWs(const bta_gp_model* m, const bta_qlognei_state* st, int64_t b) {
    const size_t es = st->precision == BTA_PREC_FP64 ? 8 : 4;
    R = b * st->q;
    R_pad = round_up(R > 0 ? R : 1, 128);
    ne = st->n_b + 1;
    parts = (int)(m->n_pad / QS_COL_TILE);
    const size_t nn = (size_t)R_pad * m->n_pad;
    WsPlan p;
    xs = p.take((size_t)R_pad * m->d * 8);
    kx = p.take(nn * es);
    vt = p.take(nn * es);
    vh = p.take(st->precision == BTA_PREC_MIXED ? nn * 4 : 0);
    dxe = p.take((size_t)R_pad * m->d * 8);
    P = p.take((size_t)parts * R_pad * ne * 8);
    dext = p.take((size_t)R_pad * ne * 8);
    gq = p.take((size_t)R_pad * st->q * 8);
    Dpart = p.take((size_t)parts * R_pad * m->d * 8);
    total = p.off;
  }
};

int qlognei_step_args(const bta_gp_model* m, const bta_qlognei_state* st, const double* X,
                      int64_t b, double* value, double* grad, int* status, void* ws,
                      size_t ws_bytes, StepArgs& A) {
  if (!m || !st || !X || !value || !status || b < 1 || b > (1 << 24)) {
    set_error("qlognei_value_grad: invalid arguments");
    return BTA_ERR_INVALID;
  }
  if (st->n_b < 1 || st->n_b > 64 * MC_MAXNB_SLOTS || st->S < 1 || st->q < 1 || st->q > 8 ||
      !st->linv || !st->ext || m->n_pad % 128 || m->d > MC_MAXD) {
    set_error("qlognei_value_grad: unsupported state (n_b=%d S=%d q=%d)", st->n_b, st->S,
              st->q);
    return BTA_ERR_UNSUPPORTED;
  }
  const QlWs w(m, st, b);
  if (!ws || ws_bytes < w.total) {
    set_error("qlognei_value_grad: workspace %zu < %zu", ws_bytes, w.total);
    return BTA_ERR_WORKSPACE;
  }
  char* base = (char*)ws;
  double* xs = (double*)(base + w.xs);
  void* kx = base + w.kx;
  void* vt = base + w.vt;
  double* dxe = (double*)(base + w.dxe);
  double* P = (double*)(base + w.P);
  double* dext = (double*)(base + w.dext);
  double* gq = (double*)(base + w.gq);
  const bool f32 = st->precision != BTA_PREC_FP64;
  const bool split = st->precision == BTA_PREC_MIXED;
  const int64_t np = m->n_pad;
  const size_t plane = (size_t)w.R_pad * np;  // halfs per FP16 plane of K_x / V
  // FP16 ranges: K_x in [0, s]; |V_rj| <= sqrt(s) since V V^T <= K(X, X)
  const float kx_scale = (float)exp2(-ceil(log2(m->outputscale)));
  const float v_scale = (float)exp2(-ceil(0.5 * log2(m->outputscale)));
  _Float16* kh = (_Float16*)kx;
  _Float16* vh = (_Float16*)(base + w.vh);
  // L^-1 planes of the split mode: [hi(L^-1) | lo(L^-1) | hi(L^-T) | lo(L^-T)]
  const _Float16* lh = (const _Float16*)st->linv;
  const size_t lplane = (size_t)np * np;

  A = StepArgs{};
  A.b = b;
  A.R = w.R;
  A.R_pad = w.R_pad;
  A.np = np;
  A.q = st->q;
  A.f32 = f32;
  A.split = split;
  A.grad = grad != nullptr;
  A.covar_mode = split ? 2 : (f32 ? 1 : 0);
  A.gemm_flags = f32 ? (BTA_GEMM_F32 | BTA_GEMM_ACC64) : 0;
  A.kx = kx;
  A.vt = vt;
  A.linv = st->linv;

  // xs, K_x and the column-tile partials of P = K_x [alpha | A] in one pass
  QsCovarArgs& c = A.c;
  c.X = X;
  c.center = m->center;
  c.ls = m->lengthscale;
  c.T = m->train_xs;
  c.E = st->ext;
  c.K = kx;
  c.Klo = split ? kh + plane : nullptr;
  c.xs_out = xs;
  c.Ppart = P;
  c.R = w.R;
  c.R_pad = w.R_pad;
  c.n = m->n;
  c.n_pad = np;
  c.ldk = np;
  c.d = m->d;
  c.kind = m->kind;
  c.ne = w.ne;
  c.s = m->outputscale;
  c.kscale = kx_scale;
  // V = K_x L^-T  (L^-1 lower: the K-range of column tile j stops at its end)
  if (split) {
    H2Args& h = A.fwd;
    h.Ah = kh;
    h.Al = kh + plane;
    h.Bh = lh;
    h.Bl = lh + lplane;
    h.C = (float*)vt;
    h.Ch = grad ? vh : nullptr;
    h.Cl = grad ? vh + plane : nullptr;
    h.a_row_scale = nullptr;
    h.a_scale = kx_scale;
    h.b_scale = st->linv_scale;
    h.c_scale = v_scale;
    h.lda = np;
    h.ldb = np;
    h.ldc = np;
    h.M = (int)w.R_pad;
    h.N = (int)np;
    h.K = (int)np;
    h.flags = BTA_GEMM_B_LOWER;
    h.alpha = 1.0;
    // U = V L^-1 = K_x K^-1 into kx (K x N view of L^-1: rows k >= column n)
    H2Args& hu = A.adj;
    hu.Ah = vh;
    hu.Al = vh + plane;
    hu.Bh = lh + 2 * lplane;
    hu.Bl = lh + 3 * lplane;
    hu.C = (float*)kx;
    hu.Ch = nullptr;
    hu.Cl = nullptr;
    hu.a_row_scale = nullptr;
    hu.a_scale = v_scale;
    hu.b_scale = st->linv_scale;
    hu.c_scale = 1.0f;
    hu.lda = np;
    hu.ldb = np;
    hu.ldc = np;
    hu.M = (int)w.R_pad;
    hu.N = (int)np;
    hu.K = (int)np;
    hu.flags = BTA_GEMM_B_UPPER;
    hu.alpha = 1.0;
  }
  McArgs& a = A.mc;
  a.vt = vt;
  a.ldv = np;
  a.gq = gq;
  a.xs = xs;
  a.bs = st->xb_s;
  a.Lb = st->Lb;
  a.Lb_levels = st->Lb_levels;
  a.lb_levels_mask = st->lb_levels_mask;
  a.Zt = st->Zt;
  a.best_f = st->best_f;
  a.P = P;
  a.p_parts = w.parts;
  a.dext = dext;
  a.value = value;
  a.dxs_extra = dxe;
  a.status = status;
  a.d = m->d;
  a.n_pad = (int)np;
  a.n_b = st->n_b;
  a.S = st->S;
  a.nrest = (int)b;
  a.kind = m->kind;
  a.max_tries = st->max_tries;
  a.grad = grad != nullptr;
  a.outputscale = m->outputscale;
  a.y_mean = m->y_mean;
  a.y_std = m->y_std;
  a.c = m->mean_constant;
  a.tau_relu = st->tau_relu;
  a.tau_max = st->tau_max;
  a.jitter = st->jitter;
  a.nr = mc_paired(a, st->q, f32, split) ? 2 : 1;
  // dK_x = gq U per restart block, contracted with dK/dxs on the fly
  QsGradArgs& g = A.g;
  g.xs = xs;
  g.T = m->train_xs;
  g.E = st->ext;
  g.U = kx;
  g.gq = gq;
  g.dext = dext;
  g.Dpart = (double*)(base + w.Dpart);
  g.R = w.R;
  g.R_pad = w.R_pad;
  g.n = m->n;
  g.n_pad = np;
  g.ldu = np;
  g.d = m->d;
  g.kind = m->kind;
  g.ne = w.ne;
  g.q = st->q;
  g.s = m->outputscale;
  FinishArgs& fin = A.fin;
  fin.Dpart = g.Dpart;
  fin.nct = (int)(np / QS_COL_TILE);
  fin.R = w.R;
  fin.d = m->d;
  fin.extra = dxe;
  fin.ls = m->lengthscale;
  fin.dX = grad;
  return BTA_OK;
}

int qlognei_value_grad(const bta_gp_model* m, const bta_qlognei_state* st, const double* X,
                       int64_t b, double* value, double* grad, int* status, void* ws,
                       size_t ws_bytes, hipStream_t s) {
  StepArgs A;
  int rc = qlognei_step_args(m, st, X, b, value, grad, status, ws, ws_bytes, A);
  if (rc) return rc;
  rc = qstep_covar(A.c, A.covar_mode, s);
  if (rc) return rc;
  if (A.split) {
    rc = gemm_h2(A.fwd, s);
  } else {
    rc = gemm(A.gemm_flags | BTA_GEMM_B_LOWER, A.R_pad, A.np, A.np, 1.0, A.kx, A.np, A.linv,
              A.np, 0.0, nullptr, 0, A.vt, A.np, s);
  }
  if (rc) return rc;
  if (A.grad && !A.split) {
    rc = gemm(A.gemm_flags | BTA_GEMM_B_KN | BTA_GEMM_B_UPPER, A.R_pad, A.np, A.np, 1.0, A.vt,
              A.np, A.linv, A.np, 0.0, nullptr, 0, A.kx, A.np, s);
    if (rc) return rc;
  }
  // split mode with the gradient: U and the MC stage in one launch where it
  // applies, else the GEMM and then the MC kernel
  rc = A.grad && A.split ? launch_adjoint_mc(A.adj, A.mc, A.q, A.b, A.f32, s) : -1;
  if (rc == -1) {
    if (A.grad && A.split) {
      rc = gemm_h2(A.adj, s);
      if (rc) return rc;
    }
    rc = launch_mc(A.mc, A.q, A.b, A.f32, A.split, s);
  }
  if (rc || !A.grad) return rc;
  return qstep_grad(A.g, A.f32, A.fin.extra, A.fin.ls, A.fin.dX, s);
}

// ---------------------------------------------------------------------------
// Baseline preparation helpers
// ---------------------------------------------------------------------------
// best_f[s] = max_j (mu_b[j] + sum_k Lb[j][k] Zb[s][k]) with
// mu_b = y_mean + y_std (c + ka)   (acquisition/utils.py:134-138, logei.py:431-458)
__global__ void baseline_best_f_kernel(const double* ka, double y_mean, double y_std, double c,
                                       const double* Lb, const double* Zb, int n_b, int S,
                                       double* best_f) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double best = -INFINITY;
  for (int j = 0; j < n_b; ++j) {
    double y = y_mean + y_std * (c + ka[j]);
    for (int k = 0; k <= j; ++k) y = fma(Lb[j * n_b + k], Zb[(int64_t)s * n_b + k], y);
    best = fmax(best, y);
  }
  best_f[s] = best;
}

// out[i][j] = (TO) in[i][j] for a rows x cols block, zero elsewhere up to the
// padded shape.
template <typename TO>
__global__ void cast_pad_kernel(const double* in, int64_t ldi, int64_t rows, int64_t cols,
                                TO* out, int64_t ldo, int64_t rows_pad, int64_t cols_pad) {
  const int64_t i = blockIdx.y;
  if (i >= rows_pad) return;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < cols_pad;
       j += (int64_t)gridDim.x * blockDim.x)
    out[i * ldo + j] = (i < rows && j < cols) ? (TO)in[i * ldi + j] : (TO)0;
}

template <typename TO>
int cast_pad(const double* in, int64_t ldi, int64_t rows, int64_t cols, TO* out, int64_t ldo,
             int64_t rows_pad, int64_t cols_pad, hipStream_t s) {
  if (rows_pad <= 0 || cols_pad <= 0) return BTA_OK;
  dim3 g((unsigned)((cols_pad + 255) / 256 > 32 ? 32 : (cols_pad + 255) / 256),
         (unsigned)rows_pad);
  hipLaunchKernelGGL((cast_pad_kernel<TO>), g, dim3(256), 0, s, in, ldi, rows, cols, out, ldo,
                     rows_pad, cols_pad);
  BTA_LAUNCH_CHECK("cast_pad_kernel");
  return BTA_OK;
}
template int cast_pad<double>(const double*, int64_t, int64_t, int64_t, double*, int64_t,
                              int64_t, int64_t, hipStream_t);
template int cast_pad<float>(const double*, int64_t, int64_t, int64_t, float*, int64_t, int64_t,
                             int64_t, hipStream_t);

int launch_baseline_best_f(const double* ka, double y_mean, double y_std, double c,
                           const double* Lb, const double* Zb, int n_b, int S, double* best_f,
                           hipStream_t s) {
  hipLaunchKernelGGL(baseline_best_f_kernel, dim3((S + 255) / 256), dim3(256), 0, s, ka, y_mean,
                     y_std, c, Lb, Zb, n_b, S, best_f);
  BTA_LAUNCH_CHECK("baseline_best_f_kernel");
  return BTA_OK;
}

size_t qlognei_ws_bytes(const bta_gp_model* m, const bta_qlognei_state* st, int64_t b) {
  return QlWs(m, st, b).total;
}

// ---------------------------------------------------------------------------
// Measured tuning of the step's split-FP16 products: every plan candidate of
// the forward product (tile width, worker count, XCD groups) is timed inside
// the whole step with HIP events, the fastest kept, then the same for the
// adjoint product (which decides whether the adjoint and the MC stage share a
// launch: they do for 64-column plans).  The winners replace the cost model's
// plans for these problem shapes.
// ---------------------------------------------------------------------------
int qlognei_autotune(const bta_gp_model* m, const bta_qlognei_state* st, const double* X,
                     int64_t b, double* value, double* grad, int* status, void* ws,
                     size_t ws_bytes, int* chosen, hipStream_t s) {
  StepArgs A;
  int rc = qlognei_step_args(m, st, X, b, value, grad, status, ws, ws_bytes, A);
  if (rc) return rc;
  for (int i = 0; i < 6; ++i) chosen[i] = 0;
  if (!A.split) return BTA_OK;  // the FP64 and FP32 products keep the cost model's plans
  hipEvent_t e0, e1;
  BTA_HIP(hipEventCreate(&e0));
  BTA_HIP(hipEventCreate(&e1));
  auto time_step = [&](float* ms) -> int {
    int r = qlognei_value_grad(m, st, X, b, value, grad, status, ws, ws_bytes, s);  // warm
    if (r) return r;
    if (hipEventRecord(e0, s) != hipSuccess) return BTA_ERR_HIP;
    for (int k = 0; k < 3 && !r; ++k)
      r = qlognei_value_grad(m, st, X, b, value, grad, status, ws, ws_bytes, s);
    if (r) return r;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(ms, e0, e1) != hipSuccess)
      return BTA_ERR_HIP;
    return BTA_OK;
  };
  const H2Args* prob[2] = {&A.fwd, &A.adj};
  for (int p = 0; p < (A.grad ? 2 : 1) && !rc; ++p) {
    int cand[3 * 32];
    const int nc = gemm_h2_candidates(*prob[p], cand, 32);
    float best = 1e30f;
    int bi = -1;
    for (int c = 0; c < nc && !rc; ++c) {
      gemm_h2_set_plan(*prob[p], cand[3 * c], cand[3 * c + 1], cand[3 * c + 2]);
      float ms = 0.0f;
      rc = time_step(&ms);
      if (!rc && ms < best) {
        best = ms;
        bi = c;
      }
    }
    if (!rc && bi >= 0) {
      gemm_h2_set_plan(*prob[p], cand[3 * bi], cand[3 * bi + 1], cand[3 * bi + 2]);
      for (int i = 0; i < 3; ++i) chosen[3 * p + i] = cand[3 * bi + i];
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return rc;
}

// ---------------------------------------------------------------------------
// Allocation-free variant of the step: its workspace is `scratch`, a static
// device array (one copy on every GPU the library runs on, held until the
// process ends).  Every kernel of the step writes the parts of its workspace
// it reads before reading them (the split-K zeroing of the FP64 GEMM is a
// kernel), so nothing in it has to start at zero, and the contents a call
// leaves behind never reach the next call's results.  Calls that overlap on
// one GPU share the array and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[BTA_NOALLOC_SCRATCH_BYTES];

// Device address of the current GPU's copy of `scratch`: looked up once per
// device, with that device current, and kept.  Host code passes this address,
// never the array's name (which on the host is a stand-in, not device memory).
static int scratch_address(void** out) {
  static std::mutex& mu = *new std::mutex;  // never destroyed
  static void* addr[64] = {nullptr};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= 64) {
    set_error("qlognei_value_grad_noalloc: device id %d out of range", dev);
    return BTA_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) {
      set_error("qlognei_value_grad_noalloc: no device address for the scratch array");
      return BTA_ERR_HIP;
    }
    addr[dev] = p;
  }
  *out = addr[dev];
  return BTA_OK;
}

int qlognei_value_grad_noalloc(const bta_gp_model* m, const bta_qlognei_state* st,
                               const double* X, int64_t b, double* value, double* grad,
                               int* status, hipStream_t s) {
  if (!m || !st || b < 1 || b > (1 << 24) || st->q < 1 || st->q > 8) {
    set_error("qlognei_value_grad_noalloc: invalid arguments");
    return BTA_ERR_INVALID;
  }
  const size_t need = qlognei_ws_bytes(m, st, b);
  if (need > (size_t)BTA_NOALLOC_SCRATCH_BYTES) {
    set_error("qlognei_value_grad_noalloc: this call needs %zu bytes of scratch, above the "
              "%zu-byte static array (use bta_qlognei_value_grad with a workspace, or fewer "
              "restarts per call)",
              need, (size_t)BTA_NOALLOC_SCRATCH_BYTES);
    return BTA_ERR_WORKSPACE;
  }
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc) return rc;
  return qlognei_value_grad(m, st, X, b, value, grad, status, ws,
                            (size_t)BTA_NOALLOC_SCRATCH_BYTES, s);
}

}  // namespace bta

extern "C" size_t bta_qlognei_workspace_size(const bta_gp_model* m, const bta_qlognei_state* st,
                                             int64_t b) {
  if (!m || !st || b < 1) return 0;
  return bta::qlognei_ws_bytes(m, st, b);
}

extern "C" int bta_qlognei_value_grad_noalloc(const bta_gp_model* m,
                                              const bta_qlognei_state* st, const double* X,
                                              int64_t b, double* value, double* grad,
                                              int32_t* status, bta_stream_t stream) {
  return bta::qlognei_value_grad_noalloc(m, st, X, b, value, grad, status,
                                         (hipStream_t)stream);
}

extern "C" int bta_qlognei_autotune(const bta_gp_model* m, const bta_qlognei_state* st,
                                    const double* X, int64_t b, double* value, double* grad,
                                    int32_t* status, void* ws, size_t ws_bytes, int32_t* chosen,
                                    bta_stream_t stream) {
  if (!chosen) {
    bta::set_error("qlognei_autotune: chosen (6 host ints) is NULL");
    return BTA_ERR_INVALID;
  }
  return bta::qlognei_autotune(m, st, X, b, value, grad, status, ws, ws_bytes, chosen,
                               (hipStream_t)stream);
}

extern "C" int bta_qlognei_value_grad(const bta_gp_model* m, const bta_qlognei_state* st,
                                      const double* X, int64_t b, double* value, double* grad,
                                      int32_t* status, void* ws, size_t ws_bytes,
                                      bta_stream_t stream) {
  return bta::qlognei_value_grad(m, st, X, b, value, grad, status, ws, ws_bytes,
                                 (hipStream_t)stream);
}
