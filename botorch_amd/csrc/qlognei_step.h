// The per-restart MC stage of the qLogNEI step and its adjoint (mc_body), the
// per-model launch arguments of the whole step (StepArgs), shared by the
// one-model step (qlognei.hip) and the step over many models (qbatch.hip).
// The algorithm is described at the top of qlognei.hip.
#pragma once
#include <cstdlib>

#include "bta_common.h"
#include "gemm.h"
#include "gemm_h2_body.h"
#include "kernels.h"

namespace bta {

constexpr int MC_MAXD = 32;
constexpr int MC_MAXNB_SLOTS = 4;  // n_b <= 256
constexpr int MC_LB_STAGE = 16;    // Lb staged in LDS for n_b up to this
// LDS bound of the MC stage, the same for the standalone and the fused
// launches so that a state is accepted by both or by neither (the value-only
// call runs the standalone kernel, value+grad may run the fused one).
constexpr size_t MC_LDS_MAX = 64 * 1024;

struct McArgs {
  const void* vt;
  int64_t ldv;
  double* gq;         // R x q: dV = gq V per restart block (-2 s^2 gA)
  const double* xs;
  const double* bs;
  const double* Lb;
  const double* Lb_levels;  // joint mode: chol(Sigma_bb + j_t I), t = 0..max_tries, or NULL
  int lb_levels_mask;       // bit t: Lb_levels[t] exists
  const double* Zt;
  const double* best_f;
  const double* P;    // R x ne x p_parts partial sums of K_x [alpha | A] (FP64)
  int p_parts;
  double* dext;       // R x ne: adjoint of P
  double* value;
  double* dxs_extra;
  int* status;
  int d, n_pad, n_b, S, kind, max_tries, grad;
  int nrest;  // restarts (b)
  int nr;     // restarts per MC workgroup (1 or 2; mc_paired)
  double outputscale, y_mean, y_std, c, tau_relu, tau_max, jitter;
};

// LDS layout (doubles) shared by host sizing and the kernel.
struct McLayout {
  int xq, red, mu, dmu, sqq, lbr, ga, lnew, dlnew, dsqb, u, G, scal, pv, vv, hh, lb, lb_n, total;
  __host__ __device__ McLayout(int Q, int n_b, int S) {
    const int m = n_b + Q;
    int o = 0;
    xq = o; o += Q * MC_MAXD;
    red = o; o += 4 * 64;
    pv = o; o += Q * (n_b + 1);
    vv = o; o += Q * (Q + 1) / 2;
    hh = o; o += Q * m;
    mu = o; o += Q;
    dmu = o; o += Q;
    sqq = o; o += Q * Q;
    lbr = o; o += Q * Q;
    ga = o; o += Q * Q;
    lnew = o; o += Q * m;
    dlnew = o; o += Q * m;
    dsqb = o; o += Q * (n_b > 0 ? n_b : 1);
    u = o; o += S;
    G = o; o += S * Q;
    scal = o; o += 8;
    // Lb (n_b x n_b) for the triangular solves, when n_b <= MC_LB_STAGE and it
    // keeps the layout within MC_LDS_MAX
    lb = o;
    lb_n = n_b <= MC_LB_STAGE && (size_t)(o + n_b * n_b) * sizeof(double) <= MC_LDS_MAX
               ? n_b * n_b : 0;
    o += lb_n;
    total = o;
  }
};

// Two restarts per MC workgroup (two waves each) in the split-FP16 mixed mode
// with q <= 5, whose MC workgroups share the CUs with the adjoint GEMM's
// workers in one launch (one round of paired workgroups in the slots the
// workers leave free is shorter than two rounds of single ones), when both
// layouts fit MC_LDS_MAX and the cached root is in use (not the joint mode).
// Every launch of that mode's MC stage (fused or alone, value-only or with the
// gradient, one model or many) takes the same choice, kept in McArgs::nr when
// the step's arguments are built, and a restart's results depend only on it,
// not on the partner restart, so value-only and value+grad steps, chunked or
// whole, one model or many, keep giving identical values.  Run alone over the
// whole GPU (mixed_f32), single workgroups fill both slots of every CU and
// stay.  BTA_MC_PAIR=0 in the environment keeps one restart per workgroup (how
// the two are compared in the tests).
inline bool mc_paired(const McArgs& a, int q, bool f32, bool split) {
  const char* e = getenv("BTA_MC_PAIR");
  const McLayout lay(q, a.n_b, a.S);
  return f32 && split && q <= 5 && !a.Lb_levels && !(e && e[0] == '0') &&
         (size_t)2 * lay.total * sizeof(double) <= MC_LDS_MAX;
}

// a / b in TM: FP32 takes the hardware reciprocal (1 ulp) and a product, FP64
// divides exactly
template <typename TM>
__device__ __forceinline__ TM div_tm(TM a, TM b) {
  if constexpr (sizeof(TM) == 4) return a * __builtin_amdgcn_rcpf(b);
  else return a / b;
}

// log(tau * (softplus(x/tau) + 0.1 / (1 + (x/tau)^2))) and its derivative
// (utils/safe_math.py:298-325), evaluated in TM (FP64, or FP32 for the mixed
// modes' sample arithmetic) from the FP64 difference x; rtau = 1 / tau.  In
// FP32, log1p(e) follows Goldberg's rule log(u) e / (u - 1), u = 1 + e (a few
// ulp; "What every computer scientist should know about floating-point
// arithmetic", 1991, thm. 4), and the divisions use div_tm.
template <typename TM>
__device__ __forceinline__ void log_fatplus_grad(double x, double tau, double rtau, double* v,
                                                 double* dv) {
  constexpr bool f32 = sizeof(TM) == 4;
  const TM z = f32 ? (TM)(x * rtau) : (TM)(x / tau);
  TM sp, sig;
  if (z > (TM)20) {  // torch softplus threshold
    sp = z;
    sig = (TM)1;
  } else {
    const TM e = exp(z);
    const TM u = (TM)1 + e;
    if constexpr (f32) sp = u == (TM)1 ? e : log(u) * div_tm(e, u - (TM)1);
    else sp = log1p(e);
    sig = div_tm(e, u);
  }
  const TM z2 = (TM)1 + z * z;
  const TM cz = div_tm((TM)1, z2);
  const TM f = sp + (TM)0.1 * cz;
  const TM tt = (TM)tau;
  *v = (double)log(tt * f);
  // dv = (sig - 0.2 z cz^2) / (f tau), with cz / f = 1 / (sp (1 + z^2) + 0.1) so
  // that no intermediate underflows for large |z|
  const TM cf = sp > (TM)0 ? div_tm((TM)1, sp * z2 + (TM)0.1) : (TM)10;
  const TM g = div_tm(sig, f) - (TM)0.2 * z * cz * cf;
  *dv = (double)(f32 ? g * (TM)rtau : g / tt);
}

// Reductions over the WPR waves of one restart (wave w of the restart writes
// red[w * 64 ...]); every thread of the restart gets the result.  The sums
// take the waves in order, after a fixed-tree sum within each wave.  Every
// thread of the workgroup must call them (they hold workgroup barriers).
template <int NV, int WPR>
__device__ __forceinline__ void group_sum(double (&v)[NV], double* red, double (&out)[NV], int wv) {
  if (NV <= 32) {
    int e;
    const double s =
        wave_sum_spread<(NV <= 32 ? NV : 1)>(*(const double(*)[NV <= 32 ? NV : 1])v, &e);
    if (e >= 0) red[wv * 64 + e] = s;
  } else {
    double s[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) s[i] = v[i];
    wave_sum_n<NV>(s);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
      for (int i = 0; i < NV; ++i) red[wv * 64 + i] = s[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double t = red[i];
#pragma unroll
    for (int w = 1; w < WPR; ++w) t += red[w * 64 + i];
    out[i] = t;
  }
  __syncthreads();
}

template <int WPR>
__device__ __forceinline__ double group_max1(double v, double* red, int wv) {
  const double s = wave_max(v);
  if ((threadIdx.x & 63) == 0) red[wv * 64] = s;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int w = 1; w < WPR; ++w) r = fmax(r, red[w * 64]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ double sq_dist(const double* a, const double* b, int d) {
  double r2 = 0.0;
  for (int j = 0; j < d; ++j) {
    double df = a[j] - b[j];
    r2 = fma(df, df, r2);
  }
  return r2;
}

// In-register Cholesky of a Q x Q SPD matrix (row-major, lower result).
template <int Q>
__device__ bool small_chol(const double* A, double jit, double* L) {
  for (int i = 0; i < Q * Q; ++i) L[i] = 0.0;
  for (int j = 0; j < Q; ++j) {
    double s = A[j * Q + j] + jit;
    for (int k = 0; k < j; ++k) s -= L[j * Q + k] * L[j * Q + k];
    if (!(s > 0.0)) return false;
    const double dj = sqrt(s), rdj = 1.0 / dj;
    L[j * Q + j] = dj;
    for (int i = j + 1; i < Q; ++i) {
      double t = A[i * Q + j];
      for (int k = 0; k < j; ++k) t -= L[i * Q + k] * L[j * Q + k];
      L[i * Q + j] = t * rdj;
    }
  }
  return true;
}

// The MC stage and its adjoint for NR restarts per workgroup (NR = 1 or 2):
// restart bw * NR + h runs on threads [h TPR, (h + 1) TPR) with the McLayout
// area sm_all + h * total in LDS.  With NR = 2 the two restarts pass the same
// barriers, so their serial phases overlap; a half past the last restart
// repeats the last one and stores nothing.
template <int Q, typename TV, int NR = 1>
__device__ __forceinline__ void mc_body(const McArgs& a, const int bw, double* sm_all) {
  static_assert(NR == 1 || NR == 2, "one or two restarts per workgroup");
  constexpr int TPR = 256 / NR, WPR = 4 / NR;  // threads and waves per restart
  const McLayout lay(Q, a.n_b, a.S);
  const int half = (int)threadIdx.x / TPR;
  const int tid = (int)threadIdx.x % TPR, lane = tid & 63, wv = tid >> 6;
  const int bt = bw * NR + half;
  const bool live = bt < a.nrest;
  const int b = live ? bt : a.nrest - 1;
  double* sm = sm_all + half * lay.total;
  const int d = a.d, n_b = a.n_b, S = a.S, m = n_b + Q;
  const double ys = a.y_std, ys2 = a.y_std * a.y_std;
  double* xq = sm + lay.xq;
  double* red = sm + lay.red;
  double* mu = sm + lay.mu;
  double* dmu = sm + lay.dmu;
  double* sqq = sm + lay.sqq;
  double* lbr = sm + lay.lbr;
  double* gA = sm + lay.ga;
  double* lnew = sm + lay.lnew;
  double* dlnew = sm + lay.dlnew;
  double* dsqb = sm + lay.dsqb;
  double* uu = sm + lay.u;
  double* G = sm + lay.G;
  double* scal = sm + lay.scal;
  const TV* vt = (const TV*)a.vt;
  const int64_t r0 = (int64_t)b * Q;

  // the query coordinates (the first TPR one per thread, the rest, for q d >
  // TPR, after), then the restart's V rows for V_b V_b^T (16-byte loads): their
  // first two column passes are issued here, ahead of the other loads of the
  // stage
  const bool has_x = tid < Q * d;
  const double xr = has_x ? a.xs[r0 * d + tid] : 0.0;
  constexpr int VW = 16 / (int)sizeof(TV);  // elements per 16-byte load
  typedef TV vec_t __attribute__((ext_vector_type(VW)));
  vec_t vpre[2][Q];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int k = (tid + TPR * h) * VW;
#pragma unroll
    for (int i = 0; i < Q; ++i)
      vpre[h][i] = k < a.n_pad ? *(const vec_t*)(vt + (r0 + i) * a.ldv + k) : vec_t{};
  }
  // In the cached-root mode Lb goes to LDS (lay.lb_n entries) for the two
  // triangular solves, whose column loops would otherwise wait on a global load
  // per step; the first TPR entries load here, with the V rows.
  const bool joint = a.Lb_levels != nullptr;
  const bool lb_lds = !joint && lay.lb_n > 0;
  double* lbs = sm + lay.lb;
  const double lbv = lb_lds && tid < lay.lb_n ? a.Lb[tid] : 0.0;
  if (has_x) xq[(tid / d) * MC_MAXD + tid % d] = xr;
  for (int i = tid + TPR; i < Q * d; i += TPR) xq[(i / d) * MC_MAXD + i % d] = a.xs[r0 * d + i];
  if (tid == 0) scal[1] = 0.0;
  __syncthreads();
  // Kernel values K(x_i, x_j) into sqq and K(x_i, X_b) into lnew[:, :n_b], and
  // with the gradient the radial factors of the direct input gradient into hh
  // (Q x m: h(x_i, x_j) for j < Q, 0 on the diagonal, h(x_i, b_kb) after); they
  // need only the coordinates, so they run (from the last wave down) while the
  // V rows and the P partials are in flight.
  double* hh = sm + lay.hh;
  for (int t = TPR - 1 - tid; t < Q * Q + Q * n_b; t += TPR) {
    if (t < Q * Q) {
      const int i = t / Q, j = t % Q;
      const double r2 = sq_dist(xq + i * MC_MAXD, xq + j * MC_MAXD, d);
      sqq[t] = kern_value(a.kind, r2, a.outputscale);
      if (a.grad)
        hh[i * m + j] = j == i ? 0.0 : kern_grad_factor(a.kind, r2, a.outputscale, nullptr);
    } else {
      const int e = t - Q * Q, i = e / n_b, kb = e % n_b;
      const double r2 = sq_dist(xq + i * MC_MAXD, a.bs + kb * d, d);
      lnew[i * m + kb] = kern_value(a.kind, r2, a.outputscale);
      if (a.grad) hh[i * m + Q + kb] = kern_grad_factor(a.kind, r2, a.outputscale, nullptr);
    }
  }

  // ---- P rows (fixed-order sums of their column-tile partials, one value per
  // wave task, partials across the lanes) and V_b V_b^T over the n_pad
  // columns of V (16-byte loads, all issued before the products) ----------
  constexpr int NQ = Q * (Q + 1) / 2;
  double* pv = sm + lay.pv;  // Q x (n_b + 1): P[r0 + i][j]
  double* vv = sm + lay.vv;  // packed lower triangle of V_b V_b^T
  {
    // 4 NR tasks per wave per pass (16 per restart), their loads issued
    // together; each task's sum is the same whatever the batch
    constexpr int TB = 4 * NR;
    const int ne = n_b + 1;
    for (int t0 = wv * TB; t0 < Q * ne; t0 += TB * WPR) {
      double s[TB];
#pragma unroll
      for (int u = 0; u < TB; ++u) {
        const int t = t0 + u;
        s[u] = 0.0;
        if (t < Q * ne) {
          const double* p = a.P + ((r0 + t / ne) * ne + t % ne) * a.p_parts;
          for (int pt = lane; pt < a.p_parts; pt += 64) s[u] += p[pt];
        }
      }
      int e;
      const double tot = wave_sum_spread<TB>(s, &e);
      if (e >= 0 && t0 + e < Q * ne) pv[t0 + e] = tot;
    }
    double acc[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) acc[i] = 0.0;
    auto vv_pass = [&](const vec_t(&v)[Q]) {
#pragma unroll
      for (int e = 0; e < VW; ++e) {
        int idx = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i)
#pragma unroll
          for (int j = 0; j <= i; ++j) {
            acc[idx] = fma((double)v[i][e], (double)v[j][e], acc[idx]);
            ++idx;
          }
      }
    };
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if ((tid + TPR * h) * VW < a.n_pad) vv_pass(vpre[h]);
    for (int k = (tid + 2 * TPR) * VW; k < a.n_pad; k += TPR * VW) {
      vec_t v[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) v[i] = *(const vec_t*)(vt + (r0 + i) * a.ldv + k);
      vv_pass(v);
    }
    double tot[NQ];
    group_sum<NQ, WPR>(acc, red, tot, wv);
    if (tid < NQ) vv[tid] = tot[tid];
  }
  __syncthreads();
  if (tid < Q * Q) {
    const int i = tid / Q, j = tid % Q;
    const int ii = i > j ? i : j, jj = i > j ? j : i;
    sqq[i * Q + j] = ys2 * (sqq[i * Q + j] - vv[ii * (ii + 1) / 2 + jj]);
  }
  if (tid < Q) mu[tid] = a.y_mean + ys * (a.c + pv[tid * (n_b + 1)]);
  // ---- Sigma_qb into lnew[:, :n_b] (and a copy in dsqb for the joint mode)
  if (lb_lds) {
    if (tid < lay.lb_n) lbs[tid] = lbv;
    for (int e = tid + TPR; e < lay.lb_n; e += TPR) lbs[e] = a.Lb[e];
  }
  for (int e = tid; e < Q * n_b; e += TPR) {
    const int i = e / n_b, kb = e % n_b;
    const double sqb = ys2 * (lnew[i * m + kb] - pv[i * (n_b + 1) + 1 + kb]);
    lnew[i * m + kb] = sqb;
    if (joint) dsqb[e] = sqb;
  }
  __syncthreads();

  // Cached root (a.Lb): one substitution, then the q x q Schur block's
  // Cholesky with psd_safe_cholesky's jitter escalation.  Joint mode (the
  // reference's fallback to sampling from the joint posterior's Cholesky): the
  // jitter j_t of try t goes on the whole (n_b + q) x (n_b + q) diagonal, so try
  // t uses the baseline factor Lb_t = chol(Sigma_bb + j_t I) (a.Lb_levels, where
  // bit t of a.lb_levels_mask says it exists) and redoes the substitution.
  const double* Lb = a.Lb;
  int st_bits = 0;
  for (int t = 0; t <= (joint ? a.max_tries : 0); ++t) {
    if (joint) {
      if (!((a.lb_levels_mask >> t) & 1)) continue;  // the baseline block fails too
      Lb = a.Lb_levels + (int64_t)t * n_b * n_b;
      if (t > 0) {
        for (int e = tid; e < Q * n_b; e += TPR) lnew[(e / n_b) * m + e % n_b] = dsqb[e];
        __syncthreads();
      }
    }
    // ---- Lbl = Sigma_qb Lb^-T: forward substitution, one wave per row ----
    auto lbl_rows = [&](const double* L) {
      for (int i = wv; i < Q; i += WPR) {
        double rhs[MC_MAXNB_SLOTS];
#pragma unroll
        for (int t2 = 0; t2 < MC_MAXNB_SLOTS; ++t2) {
          const int j = lane + 64 * t2;
          rhs[t2] = j < n_b ? lnew[i * m + j] : 0.0;
        }
        for (int kb = 0; kb < n_b; ++kb) {
          const int slot = kb >> 6, owner = kb & 63;
          double mine = 0.0;
#pragma unroll
          for (int t2 = 0; t2 < MC_MAXNB_SLOTS; ++t2)
            if (t2 == slot) mine = rhs[t2];
          const double y = __shfl(mine, owner, 64) / L[kb * n_b + kb];
#pragma unroll
          for (int t2 = 0; t2 < MC_MAXNB_SLOTS; ++t2) {
            const int j = lane + 64 * t2;
            if (j == kb) rhs[t2] = y;
            else if (j > kb && j < n_b) rhs[t2] -= L[j * n_b + kb] * y;
          }
        }
#pragma unroll
        for (int t2 = 0; t2 < MC_MAXNB_SLOTS; ++t2) {
          const int j = lane + 64 * t2;
          if (j < n_b) lnew[i * m + j] = rhs[t2];
        }
      }
    };
    if (lb_lds) lbl_rows(lbs);
    else lbl_rows(Lb);
    __syncthreads();

    // ---- Schur complement and its Cholesky --------------------------------
    if (tid < Q * Q) {
      const int i = tid / Q, j = tid % Q;
      double s = 0.0;
      for (int kb = 0; kb < n_b; ++kb) s = fma(lnew[i * m + kb], lnew[j * m + kb], s);
      gA[i * Q + j] = sqq[i * Q + j] - s;  // gA used as scratch for the Schur block
    }
    __syncthreads();
    if (tid == 0) {
      int st = 0;
      bool ok = false;
      bool has_nan = false;
      for (int i = 0; i < Q * Q; ++i) has_nan |= (gA[i] != gA[i]);
      const int first = joint ? t : 0, last = joint ? t : a.max_tries;
      for (int attempt = first; attempt <= last && !has_nan; ++attempt) {
        const double jit = attempt == 0 ? 0.0 : a.jitter * pow(10.0, (double)(attempt - 1));
        ok = small_chol<Q>(gA, jit, lbr);
        if (ok) {
          if (attempt > 0) st |= BTA_STATUS_JITTER;
          break;
        }
      }
      if (has_nan) st |= BTA_STATUS_NAN;
      else if (!ok) st |= BTA_STATUS_NOT_PSD;
      if (!ok)
        for (int i = 0; i < Q * Q; ++i) lbr[i] = NAN;
      scal[0] = (double)st;
    }
    __syncthreads();
    st_bits = (int)scal[0];
    if (!joint || !(st_bits & BTA_STATUS_NOT_PSD) || (st_bits & BTA_STATUS_NAN)) break;
    __syncthreads();  // every thread has read scal[0] before the next try rewrites it
  }
  if (joint && a.lb_levels_mask == 0 && tid == 0) {  // no baseline factor at any jitter
    scal[0] = (double)BTA_STATUS_NOT_PSD;
    for (int i = 0; i < Q * Q; ++i) lbr[i] = NAN;
  }
  __syncthreads();
  for (int e = tid; e < Q * Q; e += TPR) {
    const int i = e / Q, j = e % Q;
    lnew[i * m + n_b + j] = lbr[i * Q + j];
  }
  __syncthreads();

  // ---- samples, log improvement, fat-max over q ---------------------------
  // The per-sample transcendental arithmetic runs in TM: FP64 in the FP64
  // mode, FP32 in the mixed modes (samples and best_f stay FP64).
  typedef TV TM;
  const double tr = a.tau_relu, tm = a.tau_max, rtr = 1.0 / tr, rtm = 1.0 / tm;
  bool bad = false;
  double lmax = -INFINITY;  // this thread's max of u over its samples
  // y -> log improvement -> fat-max -> u[s], and the per-sample adjoint G[s]
  auto sample_tail = [&](const int s, const double bf, const double(&y)[Q]) {
    double v[Q], dv[Q];
    double M = -INFINITY;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      bad |= !isfinite(y[i]);
      log_fatplus_grad<TM>(y[i] - bf, tr, rtr, &v[i], &dv[i]);
      M = fmax(M, v[i]);
    }
    TM P = 0, sdp = 0, dp[Q];
    int nties = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const TM z = sizeof(TM) == 4 ? (TM)((M - v[i]) * rtm) : (TM)((M - v[i]) / tm);
      const TM den = (TM)2 + (TM)2 * z + z * z;
      P += div_tm((TM)2, den);
      dp[i] = div_tm((TM)-2 * ((TM)2 + (TM)2 * z), den * den);
      sdp += dp[i];
      nties += (v[i] == M);
    }
    const double u = M + tm * (double)log(P);
    uu[s] = u;
    lmax = fmax(lmax, u);
    const TM cm = div_tm((TM)1 + div_tm(sdp, P), (TM)nties);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const TM du = (v[i] == M ? cm : (TM)0) - div_tm(dp[i], P);
      G[s * Q + i] = (double)du * dv[i];
    }
  };
  // Each thread carries the samples s0 and s0 + TPR through the products: the
  // rows of [Lbl | Lbr] (LDS) and the base samples are read eight columns at a
  // time at clamped indices, so that every read issues before the products;
  // the products skip the columns past m.
  for (int s0 = tid; s0 < S; s0 += 2 * TPR) {
    const bool two = s0 + TPR < S;
    const int s1 = two ? s0 + TPR : s0;
    const double bf0 = a.best_f[s0], bf1 = a.best_f[s1];
    double y0[Q], y1[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) y0[i] = y1[i] = mu[i];
    for (int k0 = 0; k0 < m; k0 += 8) {
      double lc[Q][8], z0[8], z1[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = k0 + u < m ? k0 + u : m - 1;
        z0[u] = a.Zt[(int64_t)k * S + s0];
        z1[u] = a.Zt[(int64_t)k * S + s1];
#pragma unroll
        for (int i = 0; i < Q; ++i) lc[i][u] = lnew[i * m + k];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k0 + u < m)
#pragma unroll
          for (int i = 0; i < Q; ++i) {
            y0[i] = fma(lc[i][u], z0[u], y0[i]);
            y1[i] = fma(lc[i][u], z1[u], y1[i]);
          }
    }
    sample_tail(s0, bf0, y0);
    if (two) sample_tail(s0 + TPR, bf1, y1);
  }
  if (bad) scal[1] = 1.0;  // every writer stores the same value

  // ---- logmeanexp over samples (each thread reads back only its own u[s]
  // and G[s] until the next barrier) ----------------------------------------
  const double Mu = group_max1<WPR>(lmax, red, wv);
  double lsum[1] = {0.0};
  for (int s = tid; s < S; s += TPR) {  // uu keeps exp(u - max) for the adjoint
    const double e = exp(uu[s] - Mu);
    uu[s] = e;
    lsum[0] += e;
  }
  double tsum[1];
  group_sum<1, WPR>(lsum, red, tsum, wv);
  if (tid == 0 && live) {
    a.value[b] = Mu + log(tsum[0]) - log((double)S);
    int st = (int)scal[0];
    if (scal[1] != 0.0) st |= BTA_STATUS_NAN;
    a.status[b] = st;
  }
  if (!a.grad) return;

  // ---- dmu, d[Lbl | Lbr] = (w G)^T Z with the logmeanexp weights w --------
  constexpr int MREG = 32 / Q - 1;  // per-thread accumulators while m <= MREG
  if (m <= MREG) {
    constexpr int NV = Q * (MREG + 1);
    double acc[NV];
#pragma unroll
    for (int e = 0; e < NV; ++e) acc[e] = 0.0;
    auto accum = [&](const int s, const double(&z)[MREG]) {
      const double w = uu[s] / tsum[0];
      double gw[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        gw[i] = G[s * Q + i] * w;
        acc[i * (MREG + 1) + MREG] += gw[i];
      }
#pragma unroll
      for (int k = 0; k < MREG; ++k) {
        if (k < m) {
#pragma unroll
          for (int i = 0; i < Q; ++i)
            acc[i * (MREG + 1) + k] = fma(gw[i], z[k], acc[i * (MREG + 1) + k]);
        }
      }
    };
    // samples s and s + TPR per pass, the base samples of both loaded ahead of
    // the products (uniform branches); the sums run in sample order as before
    for (int s = tid; s < S; s += 2 * TPR) {
      const bool two = s + TPR < S;
      const int s2 = two ? s + TPR : s;
      double z[MREG], z2[MREG];
#pragma unroll
      for (int k = 0; k < MREG; ++k)
        if (k < m) {
          z[k] = a.Zt[(int64_t)k * S + s];
          z2[k] = a.Zt[(int64_t)k * S + s2];
        }
      accum(s, z);
      if (two) accum(s2, z2);
    }
    {
      int e;
      const double sw = wave_sum_spread<NV>(acc, &e);
      if (e >= 0) red[wv * 64 + e] = sw;
    }
    __syncthreads();
    for (int e = tid; e < Q * (m + 1); e += TPR) {
      const int i = e / (m + 1), k = e % (m + 1);
      const int idx = i * (MREG + 1) + (k == m ? MREG : k);
      double tot = red[idx];
#pragma unroll
      for (int w = 1; w < WPR; ++w) tot += red[w * 64 + idx];
      if (k == m) dmu[i] = tot;
      else dlnew[i * m + k] = k < n_b + i + 1 ? tot : 0.0;
    }
    __syncthreads();
  } else {
  for (int s = tid; s < S; s += TPR) {
    const double w = uu[s] / tsum[0];
#pragma unroll
    for (int i = 0; i < Q; ++i) G[s * Q + i] *= w;
  }
  __syncthreads();

  // one wave per (i, k) for wide baselines
  for (int task = wv; task < Q * (m + 1); task += WPR) {
    const int i = task / (m + 1), k = task % (m + 1);
    if (k < m && k >= n_b + i + 1) {  // upper part of Lbr: no dependence
      if (lane == 0) dlnew[i * m + k] = 0.0;
      continue;
    }
    double acc = 0.0;
    if (k == m) {
      for (int s = lane; s < S; s += 64) acc += G[s * Q + i];
    } else {
      const double* zr = a.Zt + (int64_t)k * S;
      for (int s = lane; s < S; s += 64) acc = fma(G[s * Q + i], zr[s], acc);
    }
    acc = wave_sum(acc);
    if (lane == 0) {
      if (k == m) dmu[i] = acc;
      else dlnew[i * m + k] = acc;
    }
  }
  __syncthreads();
  }

  // ---- Cholesky backward of Lbr (torch's formula), gA symmetric ----------
  if (tid == 0) {
    double T[Q * Q], Mm[Q * Q], Y[Q * Q];
    for (int i = 0; i < Q; ++i)
      for (int j = 0; j < Q; ++j) {
        double t = 0.0;
        if (j <= i)
          for (int k = 0; k < Q; ++k) t += lbr[k * Q + i] * dlnew[k * m + n_b + j];
        T[i * Q + j] = t;  // tril(L^T dL)
      }
    for (int i = 0; i < Q; ++i)
      for (int j = 0; j < Q; ++j)
        Mm[i * Q + j] = 0.5 * (i >= j ? T[i * Q + j] : T[j * Q + i]);
    // reciprocal diagonal once: the two solves below multiply by it
    double rd[Q];
    for (int i = 0; i < Q; ++i) rd[i] = 1.0 / lbr[i * Q + i];
    // Y = L^-T Mm : solve L^T Y = Mm (upper, back substitution per column)
    for (int j = 0; j < Q; ++j)
      for (int i = Q - 1; i >= 0; --i) {
        double s = Mm[i * Q + j];
        for (int k = i + 1; k < Q; ++k) s -= lbr[k * Q + i] * Y[k * Q + j];
        Y[i * Q + j] = s * rd[i];
      }
    // gA = Y L^-1 : solve gA L = Y, i.e. for each row r: x L = y
    for (int r = 0; r < Q; ++r)
      for (int j = Q - 1; j >= 0; --j) {
        double s = Y[r * Q + j];
        for (int k = j + 1; k < Q; ++k) s -= gA[r * Q + k] * lbr[k * Q + j];
        gA[r * Q + j] = s * rd[j];
      }
  }
  __syncthreads();
  // dLbl -= 2 gA Lbl  (Schur = Sigma_qq - Lbl Lbl^T)
  for (int e = tid; e < Q * n_b; e += TPR) {
    const int i = e / n_b, kb = e % n_b;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < Q; ++j) s += gA[i * Q + j] * lnew[j * m + kb];
    dlnew[i * m + kb] -= 2.0 * s;
  }
  __syncthreads();
  // ---- dSigma_qb = dLbl Lb^-1 : back substitution, one wave per row -------
  auto dsqb_rows = [&](const double* L) {
    for (int i = wv; i < Q; i += WPR) {
      double rhs[MC_MAXNB_SLOTS];
#pragma unroll
      for (int t = 0; t < MC_MAXNB_SLOTS; ++t) {
        const int j = lane + 64 * t;
        rhs[t] = j < n_b ? dlnew[i * m + j] : 0.0;
      }
      for (int kb = n_b - 1; kb >= 0; --kb) {
        const int slot = kb >> 6, owner = kb & 63;
        double mine = 0.0;
#pragma unroll
        for (int t = 0; t < MC_MAXNB_SLOTS; ++t)
          if (t == slot) mine = rhs[t];
        const double zz = __shfl(mine, owner, 64) / L[kb * n_b + kb];
#pragma unroll
        for (int t = 0; t < MC_MAXNB_SLOTS; ++t) {
          const int j = lane + 64 * t;
          if (j == kb) rhs[t] = zz;
          else if (j < kb) rhs[t] -= L[kb * n_b + j] * zz;
        }
      }
#pragma unroll
      for (int t = 0; t < MC_MAXNB_SLOTS; ++t) {
        const int j = lane + 64 * t;
        if (j < n_b) dsqb[i * n_b + j] = rhs[t];
      }
    }
  };
  if (lb_lds) dsqb_rows(lbs);
  else dsqb_rows(Lb);
  __syncthreads();

  // ---- dV = gq V with gq = -2 s^2 gA: the q x q block for the gradient stage
  if (!live) return;
  for (int e = tid; e < Q * Q; e += TPR) a.gq[r0 * Q + e] = -2.0 * ys2 * gA[e];
  for (int e = tid; e < Q * (n_b + 1); e += TPR) {
    const int i = e / (n_b + 1), k = e % (n_b + 1);
    a.dext[(r0 + i) * (n_b + 1) + k] = k == 0 ? ys * dmu[i] : -ys2 * dsqb[i * n_b + k - 1];
  }
  // ---- direct input gradient through K(X, X) and K(X, X_b) (scaled), from
  // the radial factors hh computed with the kernel values ------------------
  for (int e = tid; e < Q * d; e += TPR) {
    const int i = e / d, j = e % d;
    const double xij = xq[i * MC_MAXD + j];
    double acc = 0.0;
    for (int i2 = 0; i2 < Q; ++i2)
      acc -= 2.0 * ys2 * gA[i * Q + i2] * hh[i * m + i2] * (xij - xq[i2 * MC_MAXD + j]);
    for (int kb = 0; kb < n_b; ++kb)
      acc -= ys2 * dsqb[i * n_b + kb] * hh[i * m + Q + kb] * (xij - a.bs[kb * d + j]);
    a.dxs_extra[(r0 + i) * d + j] = acc;
  }
}

// One launch for the adjoint GEMM U = V L^-1 and the MC stage uses the first
// NR McLayout areas of the GEMM's LDS ring (64-column tiles; NR areas stay
// within MC_LDS_MAX).
constexpr int FUSED_LDS = H2Cfg<64, false>::STAGES * H2Cfg<64, false>::STAGE;
static_assert(MC_LDS_MAX <= (size_t)FUSED_LDS, "MC layout must fit the fused ring");

// dX = (sum of the column-tile partials + direct term) / lengthscale.
struct FinishArgs {
  const double* Dpart;
  int nct;
  int64_t R;
  int d;
  const double* extra;
  const double* ls;
  double* dX;
};

// Every launch argument of one model's value(+grad) step, and its workspace
// plan.  Built by qlognei_step_args from the caller's buffers.
struct StepArgs {
  int64_t b, R, R_pad, np;
  int q, covar_mode;            // covar_mode: 0 FP64, 1 FP32, 2 FP16 planes
  bool f32, split, grad;
  QsCovarArgs c;
  H2Args fwd, adj;              // split mode: V = K_x L^-T, U = V L^-1
  int gemm_flags;               // FP64 / mixed_f32 products through gemm()
  void* kx;
  void* vt;
  const void* linv;
  McArgs mc;
  QsGradArgs g;
  FinishArgs fin;
};

size_t qlognei_ws_bytes(const bta_gp_model* m, const bta_qlognei_state* st, int64_t b);
// Checks the arguments of one step and fills A (nothing is launched).
int qlognei_step_args(const bta_gp_model* m, const bta_qlognei_state* st, const double* X,
                      int64_t b, double* value, double* grad, int* status, void* ws,
                      size_t ws_bytes, StepArgs& A);

}  // namespace bta
