// Device bodies of the qLogNEI step's elementwise stages (the algorithm is
// described at the top of qstep.hip).  Each body works on the tile or lanes it
// is given, so the same code runs as one model's kernel (qstep.hip) and as the
// per-model share of a launch over many models (the batched step in
// qlognei.hip).
#pragma once
#include "bta_common.h"
#include "kernels.h"

#include <type_traits>

namespace bta {

constexpr int QS_RT = 64;           // rows per block (16 per wave)
constexpr int QS_CT = QS_COL_TILE;  // columns (training points) per block
constexpr int QS_MAXQ = 8;

__device__ __forceinline__ v4d mfma64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// training points of the tile and their squared norms
template <int KD>
struct QsT {
  double tc[QS_CT][4 * KD + 1];
  double tn2[QS_CT];
};

template <int KD>
__device__ __forceinline__ void qs_stage_t(QsT<KD>& sm, const double* T, int64_t n, int d,
                                           int64_t c0) {
  constexpr int MD = 4 * KD;
  for (int e = threadIdx.x; e < QS_CT * MD; e += 256) {
    const int c = e / MD, j = e % MD;
    sm.tc[c][j] = (j < d && c0 + c < n) ? T[(c0 + c) * d + j] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x < QS_CT) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < MD; ++j) s = fma(sm.tc[threadIdx.x][j], sm.tc[threadIdx.x][j], s);
    sm.tn2[threadIdx.x] = s;
  }
  __syncthreads();
}

// sum over the four 16-lane groups (same lane % 16)
__device__ __forceinline__ double sum_groups(double v) {
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

template <typename TO>
struct alignas(16) QsKStage {
  TO k[QS_RT][QS_CT + 16 / (int)sizeof(TO)];
};

// mode: 0 = FP64 K, 1 = FP32 K, 2 = FP16 hi/lo planes of K * kscale
// One 64-row x 64-column tile (column tile ct, row tile rt) of the covariance
// stage; 256 threads.
template <int KD, int MODE, int KIND>
__device__ __forceinline__ void qstep_covar_body(const QsCovarArgs& a, const int ct,
                                                 const int rt) {
  typedef typename std::conditional<MODE == 0, double,
                                    typename std::conditional<MODE == 1, float,
                                                              _Float16>::type>::type TO;
  __shared__ QsT<KD> sm;
  __shared__ QsKStage<TO> ks[MODE == 2 ? 2 : 1];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int d = a.d;
  const int64_t c0 = (int64_t)ct * QS_CT, r0 = (int64_t)rt * QS_RT;
  const int64_t r0w = r0 + wv * 16;
  const int64_t rl = r0w + l16;  // this lane's row as B operand / D column
  // Every global load of the block is issued before the staging barrier: the
  // raw row coordinates and the first 16 columns of E (all of them for
  // ne <= 16), so that their latency overlaps the staging of T.
  double xr[KD], xc[KD], xl[KD];
#pragma unroll
  for (int s = 0; s < KD; ++s) {
    const int j = 4 * s + g;
    const bool ok = j < d && rl < a.R;
    xr[s] = ok ? a.X[rl * d + j] : 0.0;
    xc[s] = ok ? a.center[j] : 0.0;
    xl[s] = ok ? a.ls[j] : 1.0;
  }
  double ev[4][4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t c = c0 + cb * 16 + g + 4 * q;
      ev[cb][q] = (l16 < a.ne && c < a.n) ? a.E[c * a.ne + l16] : 0.0;
    }
  qs_stage_t<KD>(sm, a.T, a.n, d, c0);

  // scaled row coordinates: lane (g, l16) holds x[rl][4s + g]
  double xb[KD];
  double xn2 = 0.0;
#pragma unroll
  for (int s = 0; s < KD; ++s) {
    const int j = 4 * s + g;
    const double v = (xr[s] - xc[s]) / xl[s];
    xb[s] = v;
    xn2 = fma(v, v, xn2);
    if (ct == 0 && j < d) a.xs_out[rl * d + j] = v;
  }
  xn2 = sum_groups(xn2);

  // K^T blocks: kv[cb][q] = K[rl][c0 + cb*16 + g + 4q]
  double kv[4][4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    v4d cr = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KD; ++s) cr = mfma64(sm.tc[cb * 16 + l16][4 * s + g], xb[s], cr);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int cl = cb * 16 + g + 4 * q;
      const double r2 = fmax(sm.tn2[cl] + xn2 - 2.0 * cr[q], 0.0);
      double k = kern_value_k<KIND>(r2, a.s);
      k = (c0 + cl < a.n && rl < a.R) ? k : 0.0;
      kv[cb][q] = k;
      if (MODE == 2) {
        // kscale is a power of two: scaling after the FP32 rounding is exact
        const float x = (float)k * a.kscale;
        const _Float16 h = (_Float16)x;
        ks[0].k[wv * 16 + l16][cl] = h;
        ks[MODE == 2 ? 1 : 0].k[wv * 16 + l16][cl] = (_Float16)((x - (float)h) * 2048.0f);
      } else {
        ks[0].k[wv * 16 + l16][cl] = (TO)k;
      }
    }
  }
  __syncthreads();
  // coalesced 16-byte row stores of the K tile (planes)
  {
    constexpr int V = 16 / (int)sizeof(TO);  // elements per 16-byte chunk
    constexpr int CPR = QS_CT / V;           // chunks per row
    for (int e = tid; e < QS_RT * CPR; e += 256) {
      const int r = e / CPR, cc = (e % CPR) * V;
      const int64_t o = (r0 + r) * a.ldk + c0 + cc;
      *(int4*)((TO*)a.K + o) = *(const int4*)&ks[0].k[r][cc];
      if (MODE == 2) *(int4*)(a.Klo + o) = *(const int4*)&ks[MODE == 2 ? 1 : 0].k[r][cc];
    }
  }
  // Ppart[r][j][ct] = sum_c K[r][c] E[c][j] in chunks of 16 columns j
  const int64_t nct = a.n_pad / QS_CT;
  for (int j0 = 0; j0 < a.ne; j0 += 16) {
    const int jj = j0 + l16;
    v4d p = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t c = c0 + cb * 16 + g + 4 * q;
        const double e =
            j0 == 0 ? ev[cb][q] : ((jj < a.ne && c < a.n) ? a.E[c * a.ne + jj] : 0.0);
        p = mfma64(kv[cb][q], e, p);
      }
    // D layout: lane (g, l16) reg q = P[r0w + g + 4q][jj]
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t r = r0w + g + 4 * q;
      if (jj < a.ne && r < a.R) a.Ppart[(r * a.ne + jj) * nct + ct] = p[q];
    }
  }
}

// One tile of the gradient stage (column tile ct, row tile rt); 256 threads.
template <int KD, typename TU, int KIND>
__device__ __forceinline__ void qstep_grad_body(const QsGradArgs& a, const int ct,
                                                const int rt) {
  constexpr int NB = (4 * KD + 1 + 15) / 16;  // 16-column blocks of [T | 1]
  __shared__ QsT<KD> sm;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int d = a.d, q = a.q, ne = a.ne;
  const int64_t c0 = (int64_t)ct * QS_CT, r0 = (int64_t)rt * QS_RT;
  const int64_t r0w = r0 + wv * 16;
  const int64_t rl = r0w + l16;
  const bool rvalid = rl < a.R;
  qs_stage_t<KD>(sm, a.T, a.n, d, c0);

  double xb[KD];
  double xn2 = 0.0;
#pragma unroll
  for (int s = 0; s < KD; ++s) {
    const int j = 4 * s + g;
    const double v = (j < d && rvalid) ? a.xs[rl * d + j] : 0.0;
    xb[s] = v;
    xn2 = fma(v, v, xn2);
  }
  xn2 = sum_groups(xn2);
  // U rows feeding this wave's rows: whole restarts [kr0, kr1)
  const int64_t kr0 = r0w - r0w % q;
  int64_t kr1 = ((r0w + 15) / q + 1) * q;
  if (kr1 > a.R) kr1 = a.R;
  const int ku = kr1 > kr0 ? (int)((kr1 - kr0 + 3) / 4) : 0;
  const int64_t rblk = rl - rl % q;  // first row of this lane's restart
  const TU* U = (const TU*)a.U;

  v4d dacc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) dacc[nb] = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const int64_t ca = c0 + cb * 16 + l16;  // A-operand column
    v4d cr = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KD; ++s) cr = mfma64(sm.tc[cb * 16 + l16][4 * s + g], xb[s], cr);
    // G^T[c][r] = sum_j E[c][j] dext[r][j] + sum_u U[u][c] M[u][r]
    v4d ga = {0.0, 0.0, 0.0, 0.0};
    for (int j0 = 0; j0 < ne; j0 += 4) {
      const int j = j0 + g;
      const double ea = (j < ne && ca < a.n) ? a.E[ca * ne + j] : 0.0;
      const double db = (j < ne && rvalid) ? a.dext[rl * ne + j] : 0.0;
      ga = mfma64(ea, db, ga);
    }
    for (int s = 0; s < ku; ++s) {
      const int64_t u = kr0 + 4 * s + g;
      const double ua = (u < kr1 && ca < a.n) ? (double)U[u * a.ldu + ca] : 0.0;
      // u in this lane's restart [rblk, rblk + q) (rows beyond R never are)
      const bool same = rvalid && u >= rblk && u < rblk + q;
      const double mb = same ? a.gq[rl * q + (u - rblk)] : 0.0;
      ga = mfma64(ua, mb, ga);
    }
    // w^T[c][r] = -G h(r^2); then [T | 1]-contraction over c
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int cl = cb * 16 + g + 4 * qq;
      const double r2 = fmax(sm.tn2[cl] + xn2 - 2.0 * cr[qq], 0.0);
      double w = -ga[qq] * kern_grad_factor_k<KIND>(r2, a.s);
      w = (c0 + cl < a.n && rvalid) ? w : 0.0;
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int j = nb * 16 + l16;
        const double tb = j < d ? sm.tc[cl][j] : (j == d ? 1.0 : 0.0);
        dacc[nb] = mfma64(w, tb, dacc[nb]);
      }
    }
  }
  // D layout: lane (g, l16) reg qq = S[r0w + g + 4qq][nb*16 + l16], S = w [T | 1]
  const int src = g * 16 + (d & 15);
#pragma unroll
  for (int qq = 0; qq < 4; ++qq) {
    double sw = 0.0;  // sum_c w[r][c], held in column d
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const double v = __shfl(dacc[nb][qq], src, 64);
      if (nb == d / 16) sw = v;
    }
    const int64_t r = r0w + g + 4 * qq;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int j = nb * 16 + l16;
      if (j < d && r < a.R)
        a.Dpart[((int64_t)ct * a.R + r) * d + j] = a.xs[r * d + j] * sw - dacc[nb][qq];
    }
  }
}

// Four lanes per element, each summing every fourth partial (8 loads in
// flight), then combined in a fixed order: deterministic.  gid: the thread's
// index over the 4 R d lanes of the stage.
__device__ __forceinline__ void qstep_finish_body(const double* __restrict__ Dpart, int nct,
                                                  int64_t R, int d,
                                                  const double* __restrict__ extra,
                                                  const double* __restrict__ ls,
                                                  double* __restrict__ dX, const int64_t gid) {
  const int64_t e = gid >> 2;
  const int part = (int)(gid & 3);
  const int64_t tot = R * d;
  double acc = 0.0;
  if (e < tot) {
#pragma unroll 8
    for (int t = part; t < nct; t += 4) acc += Dpart[(int64_t)t * tot + e];
  }
  acc += __shfl_xor(acc, 1, 64);
  acc += __shfl_xor(acc, 2, 64);
  if (part == 0 && e < tot) dX[e] = (acc + (extra ? extra[e] : 0.0)) / ls[(int)(e % d)];
}

}  // namespace bta
