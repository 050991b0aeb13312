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
// by one 256-thread workgroup per restart b in LDS (qlognei_mc_kernel):
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

#include <cstdlib>

namespace bta {

constexpr int MC_MAXD = 32;
constexpr int MC_MAXNB_SLOTS = 4;  // n_b <= 256
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
  const double* Zt;
  const double* best_f;
  const double* P;    // p_parts x (R x ne) partial sums of K_x [alpha | A] (FP64)
  int64_t p_stride;   // elements between partials
  int p_parts;
  double* dext;       // R x ne: adjoint of P
  double* value;
  double* dxs_extra;
  int* status;
  int d, n_pad, n_b, S, kind, max_tries, grad;
  double outputscale, y_mean, y_std, c, tau_relu, tau_max, jitter;
};

// LDS layout (doubles) shared by host sizing and the kernel.
struct McLayout {
  int xq, red, mu, dmu, sqq, lbr, ga, lnew, dlnew, dsqb, u, G, scal, pv, vv, hh, total;
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
    total = o;
  }
};

// log(tau * (softplus(x/tau) + 0.1 / (1 + (x/tau)^2))) and its derivative
// (utils/safe_math.py:298-325), evaluated in TM (FP64, or FP32 for the mixed
// modes' sample arithmetic) from the FP64 difference x.
template <typename TM>
__device__ __forceinline__ void log_fatplus_grad(double x, double tau, double* v, double* dv) {
  const TM z = (TM)(x / tau);
  TM sp, sig;
  if (z > (TM)20) {  // torch softplus threshold
    sp = z;
    sig = (TM)1;
  } else {
    const TM e = exp(z);
    sp = log1p(e);
    sig = e / ((TM)1 + e);
  }
  const TM z2 = (TM)1 + z * z;
  const TM cz = (TM)1 / z2;
  const TM f = sp + (TM)0.1 * cz;
  const TM tt = (TM)tau;
  *v = (double)log(tt * f);
  // dv = (sig - 0.2 z cz^2) / (f tau), with cz / f = 1 / (sp (1 + z^2) + 0.1) so
  // that no intermediate underflows for large |z|
  const TM cf = sp > (TM)0 ? (TM)1 / (sp * z2 + (TM)0.1) : (TM)10;
  *dv = (double)((sig / f - (TM)0.2 * z * cz * cf) / tt);
}

// block-wide reduction of NV values (NV <= 64); every thread gets the result
// in out[]; uses red (4*64 doubles).  All threads must call.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red, double (&out)[NV]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
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
    if (lane == 0)
#pragma unroll
      for (int i = 0; i < NV; ++i) red[wv * 64 + i] = s[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) out[i] = red[i] + red[64 + i] + red[128 + i] + red[192 + i];
  __syncthreads();
}

template <int NV>
__device__ __forceinline__ void block_max(double (&v)[NV], double* red, double (&out)[NV]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = wave_max(v[i]);
    if (lane == 0) red[wv * 64 + i] = s;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i)
    out[i] = fmax(fmax(red[i], red[64 + i]), fmax(red[128 + i], red[192 + i]));
  __syncthreads();
}

__device__ __forceinline__ double block_max1(double v, double* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double s = wave_max(v);
  if (lane == 0) red[wv * 64] = s;
  __syncthreads();
  double r = fmax(fmax(red[0], red[64]), fmax(red[128], red[192]));
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

// Restart b's MC stage and its adjoint, with sm the McLayout area in LDS.
template <int Q, typename TV>
__device__ __forceinline__ void mc_body(const McArgs& a, const int b, double* sm) {
  const McLayout lay(Q, a.n_b, a.S);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
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

  for (int i = tid; i < Q * d; i += 256) xq[(i / d) * MC_MAXD + i % d] = a.xs[(r0 + i / d) * d + i % d];
  if (tid == 0) scal[1] = 0.0;

  // ---- P rows (fixed-order sums of their column-tile partials, one value per
  // wave task, partials across the lanes) and V_b V_b^T over the n_pad
  // columns of V (16-byte loads, all issued before the products) ----------
  constexpr int NQ = Q * (Q + 1) / 2;
  double* pv = sm + lay.pv;  // Q x (n_b + 1): P[r0 + i][j]
  double* vv = sm + lay.vv;  // packed lower triangle of V_b V_b^T
  {
    // four tasks per wave per pass, their loads issued together
    const int ne = n_b + 1;
    for (int t0 = wv * 4; t0 < Q * ne; t0 += 16) {
      double s[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = t0 + u;
        s[u] = 0.0;
        if (t < Q * ne) {
          const double* p = a.P + (r0 + t / ne) * ne + t % ne;
          for (int pt = lane; pt < a.p_parts; pt += 64) s[u] += p[pt * a.p_stride];
        }
      }
      int e;
      const double tot = wave_sum_spread<4>(s, &e);
      if (e >= 0 && t0 + e < Q * ne) pv[t0 + e] = tot;
    }
    constexpr int VW = 16 / (int)sizeof(TV);  // elements per 16-byte load
    typedef TV vec_t __attribute__((ext_vector_type(VW)));
    double acc[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) acc[i] = 0.0;
#pragma unroll 2
    for (int k = tid * VW; k < a.n_pad; k += 256 * VW) {
      vec_t v[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) v[i] = *(const vec_t*)(vt + (r0 + i) * a.ldv + k);
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
    }
    double tot[NQ];
    block_sum<NQ>(acc, red, tot);
    if (tid < NQ) vv[tid] = tot[tid];
  }
  __syncthreads();
  if (tid < Q * Q) {
    const int i = tid / Q, j = tid % Q;
    const int ii = i > j ? i : j, jj = i > j ? j : i;
    const double kqq = kern_value(a.kind, sq_dist(xq + i * MC_MAXD, xq + j * MC_MAXD, d),
                                  a.outputscale);
    sqq[i * Q + j] = ys2 * (kqq - vv[ii * (ii + 1) / 2 + jj]);
  }
  if (tid < Q) mu[tid] = a.y_mean + ys * (a.c + pv[tid * (n_b + 1)]);
  // ---- Sigma_qb into lnew[:, :n_b] ----------------------------------------
  for (int e = tid; e < Q * n_b; e += 256) {
    const int i = e / n_b, kb = e % n_b;
    const double kqb = kern_value(a.kind, sq_dist(xq + i * MC_MAXD, a.bs + kb * d, d),
                                  a.outputscale);
    lnew[i * m + kb] = ys2 * (kqb - pv[i * (n_b + 1) + 1 + kb]);
  }
  __syncthreads();

  // ---- Lbl = Sigma_qb Lb^-T: forward substitution, one wave per row ------
  for (int i = wv; i < Q; i += 4) {
    double rhs[MC_MAXNB_SLOTS];
#pragma unroll
    for (int t = 0; t < MC_MAXNB_SLOTS; ++t) {
      const int j = lane + 64 * t;
      rhs[t] = j < n_b ? lnew[i * m + j] : 0.0;
    }
    for (int kb = 0; kb < n_b; ++kb) {
      const int slot = kb >> 6, owner = kb & 63;
      double mine = 0.0;
#pragma unroll
      for (int t = 0; t < MC_MAXNB_SLOTS; ++t)
        if (t == slot) mine = rhs[t];
      const double y = __shfl(mine, owner, 64) / a.Lb[kb * n_b + kb];
#pragma unroll
      for (int t = 0; t < MC_MAXNB_SLOTS; ++t) {
        const int j = lane + 64 * t;
        if (j == kb) rhs[t] = y;
        else if (j > kb && j < n_b) rhs[t] -= a.Lb[j * n_b + kb] * y;
      }
    }
#pragma unroll
    for (int t = 0; t < MC_MAXNB_SLOTS; ++t) {
      const int j = lane + 64 * t;
      if (j < n_b) lnew[i * m + j] = rhs[t];
    }
  }
  __syncthreads();

  // ---- Schur complement and its Cholesky (jitter escalation) -------------
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
    for (int attempt = 0; attempt <= a.max_tries && !has_nan; ++attempt) {
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
  for (int e = tid; e < Q * Q; e += 256) {
    const int i = e / Q, j = e % Q;
    lnew[i * m + n_b + j] = lbr[i * Q + j];
  }
  __syncthreads();

  // ---- samples, log improvement, fat-max over q ---------------------------
  // The per-sample transcendental arithmetic runs in TM: FP64 in the FP64
  // mode, FP32 in the mixed modes (samples and best_f stay FP64).
  typedef TV TM;
  const double tr = a.tau_relu, tm = a.tau_max;
  bool bad = false;
#pragma unroll 2
  for (int s = tid; s < S; s += 256) {
    double y[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) y[i] = mu[i];
    for (int k = 0; k < m; ++k) {
      const double z = a.Zt[(int64_t)k * S + s];
#pragma unroll
      for (int i = 0; i < Q; ++i) y[i] = fma(lnew[i * m + k], z, y[i]);
    }
    const double bf = a.best_f[s];
    double v[Q], dv[Q];
    double M = -INFINITY;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      bad |= !isfinite(y[i]);
      log_fatplus_grad<TM>(y[i] - bf, tr, &v[i], &dv[i]);
      M = fmax(M, v[i]);
    }
    TM P = 0, sdp = 0, dp[Q];
    int nties = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const TM z = (TM)((M - v[i]) / tm);
      const TM den = (TM)2 + (TM)2 * z + z * z;
      P += (TM)2 / den;
      dp[i] = (TM)-2 * ((TM)2 + (TM)2 * z) / (den * den);
      sdp += dp[i];
      nties += (v[i] == M);
    }
    uu[s] = M + tm * (double)log(P);
    const TM cm = ((TM)1 + sdp / P) / (TM)nties;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const TM du = (v[i] == M ? cm : (TM)0) - dp[i] / P;
      G[s * Q + i] = (double)du * dv[i];
    }
  }
  if (bad) scal[1] = 1.0;  // every writer stores the same value
  __syncthreads();

  // ---- logmeanexp over samples --------------------------------------------
  double lmax = -INFINITY;
  for (int s = tid; s < S; s += 256) lmax = fmax(lmax, uu[s]);
  const double Mu = block_max1(lmax, red);
  double lsum[1] = {0.0};
  for (int s = tid; s < S; s += 256) lsum[0] += exp(uu[s] - Mu);
  double tsum[1];
  block_sum<1>(lsum, red, tsum);
  if (tid == 0) {
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
    for (int s = tid; s < S; s += 256) {
      const double w = exp(uu[s] - Mu) / tsum[0];
      double gw[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        gw[i] = G[s * Q + i] * w;
        acc[i * (MREG + 1) + MREG] += gw[i];
      }
#pragma unroll
      for (int k = 0; k < MREG; ++k) {
        if (k < m) {
          const double z = a.Zt[(int64_t)k * S + s];
#pragma unroll
          for (int i = 0; i < Q; ++i)
            acc[i * (MREG + 1) + k] = fma(gw[i], z, acc[i * (MREG + 1) + k]);
        }
      }
    }
    {
      int e;
      const double sw = wave_sum_spread<NV>(acc, &e);
      if (e >= 0) red[wv * 64 + e] = sw;
    }
    __syncthreads();
    for (int e = tid; e < Q * (m + 1); e += 256) {
      const int i = e / (m + 1), k = e % (m + 1);
      const int idx = i * (MREG + 1) + (k == m ? MREG : k);
      const double tot = red[idx] + red[64 + idx] + red[128 + idx] + red[192 + idx];
      if (k == m) dmu[i] = tot;
      else dlnew[i * m + k] = k < n_b + i + 1 ? tot : 0.0;
    }
    __syncthreads();
  } else {
  for (int s = tid; s < S; s += 256) {
    const double w = exp(uu[s] - Mu) / tsum[0];
#pragma unroll
    for (int i = 0; i < Q; ++i) G[s * Q + i] *= w;
  }
  __syncthreads();

  // one wave per (i, k) for wide baselines
  for (int task = wv; task < Q * (m + 1); task += 4) {
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
  for (int e = tid; e < Q * n_b; e += 256) {
    const int i = e / n_b, kb = e % n_b;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < Q; ++j) s += gA[i * Q + j] * lnew[j * m + kb];
    dlnew[i * m + kb] -= 2.0 * s;
  }
  __syncthreads();
  // ---- dSigma_qb = dLbl Lb^-1 : back substitution, one wave per row -------
  for (int i = wv; i < Q; i += 4) {
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
      const double zz = __shfl(mine, owner, 64) / a.Lb[kb * n_b + kb];
#pragma unroll
      for (int t = 0; t < MC_MAXNB_SLOTS; ++t) {
        const int j = lane + 64 * t;
        if (j == kb) rhs[t] = zz;
        else if (j < kb) rhs[t] -= a.Lb[kb * n_b + j] * zz;
      }
    }
#pragma unroll
    for (int t = 0; t < MC_MAXNB_SLOTS; ++t) {
      const int j = lane + 64 * t;
      if (j < n_b) dsqb[i * n_b + j] = rhs[t];
    }
  }
  __syncthreads();

  // ---- dV = gq V with gq = -2 s^2 gA: the q x q block for the gradient stage
  for (int e = tid; e < Q * Q; e += 256) a.gq[r0 * Q + e] = -2.0 * ys2 * gA[e];
  for (int e = tid; e < Q * (n_b + 1); e += 256) {
    const int i = e / (n_b + 1), k = e % (n_b + 1);
    a.dext[(r0 + i) * (n_b + 1) + k] = k == 0 ? ys * dmu[i] : -ys2 * dsqb[i * n_b + k - 1];
  }
  // ---- direct input gradient through K(X, X) and K(X, X_b) (scaled) ------
  // radial factors first (one thread each), then the d coordinates
  double* hh = sm + lay.hh;  // Q x m: h(x_i, x_i2) for i2 < Q, h(x_i, b_kb) after
  for (int e = tid; e < Q * m; e += 256) {
    const int i = e / m, j2 = e % m;
    const double* xi = xq + i * MC_MAXD;
    const double* x2 = j2 < Q ? xq + j2 * MC_MAXD : a.bs + (j2 - Q) * d;
    hh[e] = j2 == i ? 0.0
                    : kern_grad_factor(a.kind, sq_dist(xi, x2, d), a.outputscale, nullptr);
  }
  __syncthreads();
  for (int e = tid; e < Q * d; e += 256) {
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

template <int Q, typename TV>
__global__ __launch_bounds__(256) void qlognei_mc_kernel(McArgs a) {
  extern __shared__ double sm_dyn[];
  mc_body<Q, TV>(a, blockIdx.x, sm_dyn);
}

// One launch for the adjoint GEMM U = V L^-1 (workgroups 0 .. W-1, split-FP16,
// 64-column tiles) and the MC stage (workgroups W .. W+b-1, one per restart):
// both need only V, and the MC workgroups fill the CUs as the GEMM's light
// tiles finish, with no hand-off between queues.  The MC stage uses the first
// McLayout bytes of the GEMM's LDS ring.
constexpr int FUSED_LDS = H2Cfg<64, false>::STAGES * H2Cfg<64, false>::STAGE;

template <int Q, typename TV>
__global__ __launch_bounds__(256, 2) void adjoint_mc_kernel(H2Args g, int W, McArgs a) {
  __shared__ __attribute__((aligned(1024))) unsigned char lds[FUSED_LDS];
  if ((int)blockIdx.x < W)
    gemm_h2_body<64, false>(g, blockIdx.x, W, lds);
  else
    mc_body<Q, TV>(a, (int)blockIdx.x - W, (double*)lds);
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
static int launch_adjoint_mc(const H2Args& g, const McArgs& a, int q, int64_t b, bool f32,
                             hipStream_t s) {
  int bn = 0, W = 0;
  const int rc = gemm_h2_plan(g, &bn, &W);
  if (rc) return rc;
  const McLayout lay(q, a.n_b, a.S);
  static_assert(MC_LDS_MAX <= (size_t)FUSED_LDS, "MC layout must fit the fused ring");
  if (!f32 || q > 5 || bn != 64 || g.seg_chunks > 0 || fused_disabled() ||
      (size_t)lay.total * sizeof(double) > MC_LDS_MAX || b < 1 || b > (1 << 24))
    return -1;
  const dim3 grid((unsigned)(W + b));
  switch (q) {
    case 1: hipLaunchKernelGGL((adjoint_mc_kernel<1, float>), grid, dim3(256), 0, s, g, W, a); break;
    case 2: hipLaunchKernelGGL((adjoint_mc_kernel<2, float>), grid, dim3(256), 0, s, g, W, a); break;
    case 3: hipLaunchKernelGGL((adjoint_mc_kernel<3, float>), grid, dim3(256), 0, s, g, W, a); break;
    case 4: hipLaunchKernelGGL((adjoint_mc_kernel<4, float>), grid, dim3(256), 0, s, g, W, a); break;
    default: hipLaunchKernelGGL((adjoint_mc_kernel<5, float>), grid, dim3(256), 0, s, g, W, a); break;
  }
  BTA_LAUNCH_CHECK("adjoint_mc_kernel");
  return BTA_OK;
}

template <int Q>
static int launch_mc_q(const McArgs& a, int64_t b, bool f32, hipStream_t s) {
  const McLayout lay(Q, a.n_b, a.S);
  const size_t bytes = (size_t)lay.total * sizeof(double);
  if (bytes > MC_LDS_MAX) {
    set_error("qlognei: LDS need %zu bytes exceeds 64 KiB (q=%d n_b=%d S=%d)", bytes, Q, a.n_b,
              a.S);
    return BTA_ERR_UNSUPPORTED;
  }
  if (f32)
    hipLaunchKernelGGL((qlognei_mc_kernel<Q, float>), dim3((unsigned)b), dim3(256), bytes, s, a);
  else
    hipLaunchKernelGGL((qlognei_mc_kernel<Q, double>), dim3((unsigned)b), dim3(256), bytes, s,
                       a);
  BTA_LAUNCH_CHECK("qlognei_mc_kernel");
  return BTA_OK;
}

static int launch_mc(const McArgs& a, int q, int64_t b, bool f32, hipStream_t s) {
  switch (q) {
    case 1: return launch_mc_q<1>(a, b, f32, s);
    case 2: return launch_mc_q<2>(a, b, f32, s);
    case 3: return launch_mc_q<3>(a, b, f32, s);
    case 4: return launch_mc_q<4>(a, b, f32, s);
    case 5: return launch_mc_q<5>(a, b, f32, s);
    case 6: return launch_mc_q<6>(a, b, f32, s);
    case 7: return launch_mc_q<7>(a, b, f32, s);
    case 8: return launch_mc_q<8>(a, b, f32, s);
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

int qlognei_value_grad(const bta_gp_model* m, const bta_qlognei_state* st, const double* X,
                       int64_t b, double* value, double* grad, int* status, void* ws,
                       size_t ws_bytes, hipStream_t s) {
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
  const int fl = f32 ? (BTA_GEMM_F32 | BTA_GEMM_ACC64) : 0;
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

  // xs, K_x and the column-tile partials of P = K_x [alpha | A] in one pass
  QsCovarArgs c;
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
  int rc = qstep_covar(c, split ? 2 : (f32 ? 1 : 0), s);
  if (rc) return rc;
  // V = K_x L^-T  (L^-1 lower: the K-range of column tile j stops at its end)
  if (split) {
    H2Args h{};
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
    rc = gemm_h2(h, s);
  } else {
    rc = gemm(fl | BTA_GEMM_B_LOWER, w.R_pad, np, np, 1.0, kx, np, st->linv, np, 0.0, nullptr, 0,
              vt, np, s);
  }
  if (rc) return rc;
  // U = V L^-1 = K_x K^-1 into kx (K x N view of L^-1: rows k >= column n)
  H2Args hu{};
  if (grad && split) {
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
  } else if (grad) {
    rc = gemm(fl | BTA_GEMM_B_KN | BTA_GEMM_B_UPPER, w.R_pad, np, np, 1.0, vt, np, st->linv, np,
              0.0, nullptr, 0, kx, np, s);
    if (rc) return rc;
  }
  McArgs a;
  a.vt = vt;
  a.ldv = np;
  a.gq = gq;
  a.xs = xs;
  a.bs = st->xb_s;
  a.Lb = st->Lb;
  a.Zt = st->Zt;
  a.best_f = st->best_f;
  a.P = P;
  a.p_parts = w.parts;
  a.p_stride = w.R * w.ne;
  a.dext = dext;
  a.value = value;
  a.dxs_extra = dxe;
  a.status = status;
  a.d = m->d;
  a.n_pad = (int)np;
  a.n_b = st->n_b;
  a.S = st->S;
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
  // split mode with the gradient: U and the MC stage in one launch where it
  // applies, else the GEMM and then the MC kernel
  rc = grad && split ? launch_adjoint_mc(hu, a, st->q, b, f32, s) : -1;
  if (rc == -1) {
    if (grad && split) {
      rc = gemm_h2(hu, s);
      if (rc) return rc;
    }
    rc = launch_mc(a, st->q, b, f32, s);
  }
  if (rc || !grad) return rc;
  // dK_x = gq U per restart block, contracted with dK/dxs on the fly
  QsGradArgs g;
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
  return qstep_grad(g, f32, dxe, m->lengthscale, grad, s);
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

}  // namespace bta

extern "C" size_t bta_qlognei_workspace_size(const bta_gp_model* m, const bta_qlognei_state* st,
                                             int64_t b) {
  if (!m || !st || b < 1) return 0;
  return bta::qlognei_ws_bytes(m, st, b);
}

extern "C" int bta_qlognei_value_grad(const bta_gp_model* m, const bta_qlognei_state* st,
                                      const double* X, int64_t b, double* value, double* grad,
                                      int32_t* status, void* ws, size_t ws_bytes,
                                      bta_stream_t stream) {
  return bta::qlognei_value_grad(m, st, X, b, value, grad, status, ws, ws_bytes,
                                 (hipStream_t)stream);
}
