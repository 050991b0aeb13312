// Exact marginal log likelihood and its hyperparameter adjoint.
//
//   K = s k(X,X; l) + noise I,  L = chol(K),  t = L^-1 (y - c),  a = L^-T t
//   log p(y) = -1/2 (|t|^2 + 2 sum log L_ii + n log 2 pi)
//   d log p / d theta = 1/2 sum_ij W_ij dK_ij/dtheta,   W = a a^T - K^-1
//
// K^-1 = X^T X with X = L^-1 is never stored: mll_trace_kernel forms each
// lower 128 x 64 tile of it on the FP64 MFMA main loop (d_mainloop, both
// operands read K-major straight from X), the K range of a tile,
// [max(128 tm, 64 tn), n_pad), cut into chunks of KC so that no task grows
// without bound at large n; a host-built schedule deals the tasks, longest
// first, to the least loaded CU (trace_schedule).  Each task's epilogue
// rebuilds k and dk/dtheta for the tile's point pairs from the scaled inputs
// staged in LDS and folds -P_ij dK_ij/dtheta
// (plus a_i a_j dK_ij/dtheta for the tile's first chunk; the sum is linear in
// K^-1) into per-worker sums, so the n^2 x (d + 2) derivative tensor is never
// formed either.  a comes from two triangular matrix-vector products with X.
// Reference loss: optim/closures/model_closures.py:125-145 (ExactMarginal-
// LogLikelihood, gpytorch); derivative: Rasmussen & Williams (2006) eq. 5.9.
#include "bta_common.h"
#include "gemm.h"
#include "gemm_d.h"
#include "gemm_tile.h"
#include "kernels.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <array>
#include <map>
#include <mutex>
#include <vector>

namespace bta {

constexpr int TR_B = 128;   // trace tile
constexpr int TR_KC = 4096;  // K chunk per task

struct TraceArgs {
  const double* X;  // L^-1 (lower), n_pad x n_pad, row stride ld
  int64_t ld;
  const double* xs;     // n x d scaled inputs
  const double* alpha;  // n_pad
  double* partial;      // gridDim.x x (MD + 2)
  const int* sched;     // worker w's tasks: sched[W + 1 + i], i in [sched[w], sched[w + 1])
  int n_pad, n, d, MT, CB;
  double s;
};

// Tiles are 128 rows x 64 columns; tile (tm, tn) holds lower elements when
// tn <= 2 tm + 1 and its K range starts at max(128 tm, 64 tn).  Tasks of chunk
// c cover the tile rows tm < MT - c CB (those whose K range reaches chunk c),
// T (T + 1) tiles for T such rows, enumerated row by row.
__host__ __device__ inline int trace_tasks(int MT, int CB) {
  int total = 0;
  for (int c = 0; MT - c * CB > 0; ++c) {
    const int T = MT - c * CB;
    total += T * (T + 1);
  }
  return total;
}
__host__ __device__ inline bool trace_task(int t, int MT, int CB, int& tm, int& tn, int& c) {
  for (c = 0;; ++c) {
    const int T = MT - c * CB;
    if (T <= 0) return false;
    const int cnt = T * (T + 1);
    if (t < cnt) break;
    t -= cnt;
  }
  int r = (int)((__builtin_sqrt(4.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((r + 1) * (r + 2) <= t) ++r;
  while (r * (r + 1) > t) --r;
  tm = r;
  tn = t - r * (r + 1);
  return true;
}

template <int MD, int KIND>
__global__ __launch_bounds__(256, (MD == 16 ? 2 : 1)) void mll_trace_kernel(TraceArgs a) {
  constexpr int TN = 64;  // tile columns
  constexpr int RING = DCfg<TN>::STAGES * DCfg<TN>::STAGE;
  constexpr int STG = (TR_B + TN) * MD * 8;
  __shared__ __attribute__((aligned(1024))) unsigned char lds[RING > STG ? RING : STG];
  const int total = trace_tasks(a.MT, a.CB);
  const int W = gridDim.x, w = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wy = wv >> 1, wx = wv & 1, l16 = lane & 15, gq = lane >> 4;
  const int d = a.d;
  double acc_os = 0.0, acc_noise = 0.0, acc_ls[MD];
#pragma unroll
  for (int j = 0; j < MD; ++j) acc_ls[j] = 0.0;

  const int* list = a.sched + W + 1;
  for (int i = a.sched[w]; i < a.sched[w + 1]; ++i) {
    const int t = list[i];
    int tm, tn, c;
    if (t >= total || !trace_task(t, a.MT, a.CB, tm, tn, c)) continue;
    const int kb = max(tm * TR_B, tn * TN) + c * TR_KC;
    const int ke = min(a.n_pad, kb + TR_KC);
    v4d acc[4][2];
    d_mainloop<TN, true, true>(a.X, a.ld, a.X, a.ld, tm * TR_B, tn * TN, a.n_pad, a.n_pad, kb, ke,
                               lds, acc);
    // the ring is free: stage the tile's scaled rows and columns (xc follows xr)
    double* xr = (double*)lds;
    double* xc = xr + TR_B * MD;
    for (int e = threadIdx.x; e < (TR_B + TN) * MD; e += 256) {
      const int r = e / MD, j = e % MD;
      const int gi = r < TR_B ? tm * TR_B + r : tn * TN + r - TR_B;
      xr[e] = (j < d && gi < a.n) ? a.xs[(int64_t)gi * d + j] : 0.0;
    }
    __syncthreads();
    const bool first = c == 0;
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) {
      const int cl = wx * 32 + ni * 16 + l16;
      const int gj = tn * TN + cl;
      const bool jok = gj < a.n;
      const double aj = jok ? a.alpha[gj] : 0.0;
      double xj[MD];
#pragma unroll
      for (int j = 0; j < MD; ++j) xj[j] = xc[cl * MD + j];
#pragma unroll
      for (int mi = 0; mi < 4; ++mi)
#pragma unroll 1
        for (int q = 0; q < 4; ++q) {
          // (a runtime q keeps one element in flight: the epilogue then fits
          // the register file next to the accumulators)
          const v4d& av = acc[mi][ni];
          const double pij = q == 0 ? av[0] : q == 1 ? av[1] : q == 2 ? av[2] : av[3];
          const int rl = wy * 64 + mi * 16 + Mfma<double>::row(gq, q);
          const int gi = tm * TR_B + rl;
          const bool ok = jok && gi < a.n && gi >= gj;
          const double wij = (first ? a.alpha[min(gi, a.n_pad - 1)] * aj : 0.0) - pij;
          const double f = ok ? (gi == gj ? 1.0 : 2.0) : 0.0;
          double r2 = 0.0, dsq[MD];
#pragma unroll
          for (int j = 0; j < MD; ++j) {
            const double df = xr[rl * MD + j] - xj[j];
            dsq[j] = df * df;
            r2 += dsq[j];
          }
          double kv, h;
          if (KIND == BTA_KERNEL_MATERN52) {
            const double r = sqrt_pos(fmax(r2, 1e-30));
            const double sr = 2.2360679774997896964 * r;
            const double e = a.s * exp_nonpos(-sr);
            kv = (1.0 + sr + (5.0 / 3.0) * r2) * e;
            h = (5.0 / 3.0) * (1.0 + sr) * e;
          } else {
            kv = a.s * exp_nonpos(-0.5 * r2);
            h = kv;
          }
          const double fw = f * wij;
          acc_os = fma(fw, kv, acc_os);
          acc_noise += (ok && gi == gj) ? wij : 0.0;
          const double fwh = fw * h;
#pragma unroll
          for (int j = 0; j < MD; ++j) acc_ls[j] = fma(fwh, dsq[j], acc_ls[j]);
        }
    }
    __syncthreads();
  }
  // per-worker reduction
  double* red = (double*)lds;
  double v[MD + 2];
  v[0] = acc_os;
  v[1] = acc_noise;
#pragma unroll
  for (int j = 0; j < MD; ++j) v[2 + j] = acc_ls[j];
  if (MD + 2 <= 32) {
    int e;
    const double t = wave_sum_spread<(MD + 2 <= 32 ? MD + 2 : 1)>(
        *(const double(*)[MD + 2 <= 32 ? MD + 2 : 1])v, &e);
    if (e >= 0) red[wv * (MD + 2) + e] = t;
  } else {
    wave_sum_n<MD + 2>(v);
    if (lane == 0)
#pragma unroll
      for (int e = 0; e < MD + 2; ++e) red[wv * (MD + 2) + e] = v[e];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < MD + 2; e += 256)
    a.partial[(int64_t)w * (MD + 2) + e] =
        red[e] + red[(MD + 2) + e] + red[2 * (MD + 2) + e] + red[3 * (MD + 2) + e];
}

// t = X r for the lower-triangular X (one wave per row, double2 loads) with the
// residual r_k = y_k - c (0 in the padding) formed on the fly, and
// lgd_i = log L_ii on the side
__global__ __launch_bounds__(256) void trmv_lower_kernel(const double* __restrict__ X, int64_t ld,
                                                         const double* __restrict__ y, double c,
                                                         int64_t ny, int64_t n,
                                                         double* __restrict__ t,
                                                         const double* __restrict__ L,
                                                         double* __restrict__ lgd) {
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const double* row = X + i * ld;
  double acc = 0.0;
  for (int64_t k = 2 * lane; k <= i; k += 128) {
    const v2d x = *(const v2d*)(row + k);
    const double r0 = k < ny ? y[k] - c : 0.0;
    acc = fma(x.x, r0, acc);
    if (k + 1 <= i) acc = fma(x.y, k + 1 < ny ? y[k + 1] - c : 0.0, acc);
  }
  acc = wave_sum(acc);
  if (lane == 0) {
    t[i] = acc;
    lgd[i] = log(L[i * ld + i]);
  }
}

// part[rb][j] = sum over the 64 rows i of block rb (i >= j) of X_ij t_i
__global__ __launch_bounds__(256) void trmv_lower_t_part_kernel(const double* __restrict__ X,
                                                                int64_t ld,
                                                                const double* __restrict__ t,
                                                                int64_t n_pad,
                                                                double* __restrict__ part) {
  __shared__ double ts[64];
  const int rb = blockIdx.y;
  const int64_t i0 = (int64_t)rb * 64;
  if (threadIdx.x < 64) ts[threadIdx.x] = t[i0 + threadIdx.x];
  __syncthreads();
  const int64_t jend = i0 + 64;
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < jend; j += (int64_t)gridDim.x * 256) {
    double acc = 0.0;
#pragma unroll 8
    for (int r = 0; r < 64; ++r) {
      const int64_t i = i0 + r;
      acc = fma(i >= j ? X[i * ld + j] : 0.0, ts[r], acc);
    }
    part[(int64_t)rb * n_pad + j] = acc;
  }
}

// a_j = sum_{rb >= j / 64} part[rb][j]
__global__ void trmv_lower_t_sum_kernel(const double* __restrict__ part, int64_t n_pad,
                                        double* __restrict__ a) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_pad) return;
  // sixteen partials loaded together per round, added in row-block order
  double s = 0.0;
  const int64_t nb = n_pad / 64;
  for (int64_t rb0 = j / 64; rb0 < nb; rb0 += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = rb0 + u < nb ? part[(rb0 + u) * n_pad + j] : 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) s += v[u];
  }
  a[j] = s;
}

// out[0] = sum t_i^2, out[1] = sum_{i<n} log L_ii, out[2] = sum_{i<n} a_i,
// out[3 + e] = sum over the W workers of partial[w][e], e < np2.
template <int NP2>
__global__ __launch_bounds__(256) void mll_finish_kernel(const double* t, const double* lgd,
                                                         const double* alpha, int64_t n,
                                                         const double* partial, int W,
                                                         double* out, const int* stat,
                                                         int* stat_out) {
  __shared__ double red[8 * 40];
  if (threadIdx.x < 2) stat_out[threadIdx.x] = stat[threadIdx.x];
  __shared__ double red3[4 * 3];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double q = 0.0, lg = 0.0, sa = 0.0;
  // eight elements per thread loaded together per round
  for (int64_t i0 = tid; i0 < n; i0 += 256 * 8) {
    double tv[8], lv[8], av[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t i = i0 + 256 * u;
      tv[u] = i < n ? t[i] : 0.0;
      lv[u] = i < n ? lgd[i] : 0.0;
      av[u] = i < n ? alpha[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      q = fma(tv[u], tv[u], q);
      lg += lv[u];
      sa += av[u];
    }
  }
  q = wave_sum(q);
  lg = wave_sum(lg);
  sa = wave_sum(sa);
  if (lane == 0) {
    red3[wv * 3] = q;
    red3[wv * 3 + 1] = lg;
    red3[wv * 3 + 2] = sa;
  }
  // worker sums: thread tid takes workers w = tid (mod 256), all of its loads
  // in flight together, then a fixed-order reduction over lanes and waves
  static_assert(NP2 <= 40, "red holds 40 values per wave");
  double acc[NP2];
#pragma unroll
  for (int e = 0; e < NP2; ++e) acc[e] = 0.0;
  for (int w = tid; w < W; w += 256) {
#pragma unroll
    for (int e = 0; e < NP2; ++e) acc[e] += partial[(int64_t)w * NP2 + e];
  }
  if (NP2 <= 32) {
    int e;
    const double v = wave_sum_spread<(NP2 <= 32 ? NP2 : 1)>(
        *(const double(*)[NP2 <= 32 ? NP2 : 1])acc, &e);
    if (e >= 0) red[wv * 40 + e] = v;
  } else {
    wave_sum_n<NP2>(acc);
    if (lane == 0)
#pragma unroll
      for (int e = 0; e < NP2; ++e) red[wv * 40 + e] = acc[e];
  }
  __syncthreads();
  if (tid < 3) out[tid] = red3[tid] + red3[3 + tid] + red3[6 + tid] + red3[9 + tid];
  for (int e = tid; e < NP2; e += 256)
    out[3 + e] = red[e] + red[40 + e] + red[80 + e] + red[120 + e];
}


// Static schedule of the trace tasks over W workers.  Tasks, longest first
// (cost: the K length plus TRACE_EPI for the epilogue's kernel values), go one
// by one to the least loaded CU and there to its less loaded worker; with
// W = 2 x CUs the dispatcher places workers w and w + W/2 on one CU, which
// then runs their sum of work, otherwise every worker has a CU of its own.
// (The h2 heavy/light dealing it replaces left the busiest CU 16 % above the
// mean at n = 4096; this schedule, 3 %.)  Built on the host once per (device,
// n_pad, W) and kept on the device until the process ends.
constexpr int TRACE_EPI = 160;

static std::vector<int> trace_schedule_host(int MT, int CB, int n_pad, int W, int ncu) {
  const int total = trace_tasks(MT, CB);
  std::vector<std::pair<int, int>> cost(total);  // (cost, task)
  for (int t = 0; t < total; ++t) {
    int tm, tn, c;
    trace_task(t, MT, CB, tm, tn, c);
    const int kb = std::max(tm * TR_B, tn * 64) + c * TR_KC;
    const int ke = std::min(n_pad, kb + TR_KC);
    cost[t] = {std::max(ke - kb, 0) + TRACE_EPI, t};
  }
  std::stable_sort(cost.begin(), cost.end(),
                   [](const std::pair<int, int>& x, const std::pair<int, int>& y) {
                     return x.first > y.first;
                   });
  const bool paired = W == 2 * ncu;
  const int bins = paired ? W / 2 : W;
  std::vector<int64_t> bl(bins, 0), wl(W, 0);
  std::vector<std::vector<int>> lists(W);
  for (const auto& ct : cost) {
    int b = 0;
    for (int i = 1; i < bins; ++i)
      if (bl[i] < bl[b]) b = i;
    int w = b;
    if (paired && wl[b + bins] < wl[b]) w = b + bins;
    bl[b] += ct.first;
    wl[w] += ct.first;
    lists[w].push_back(ct.second);
  }
  std::vector<int> out(W + 1 + total);
  int k = W + 1;
  for (int w = 0; w < W; ++w) {
    out[w] = k - (W + 1);
    for (int t : lists[w]) out[k++] = t;
  }
  out[W] = total;
  return out;
}

static int trace_schedule(int MT, int CB, int n_pad, int W, const int** out) {
  static std::mutex& mu = *new std::mutex;  // never destroyed
  static std::map<std::array<int, 3>, const int*>& cache =
      *new std::map<std::array<int, 3>, const int*>;
  int dev = 0;
  BTA_HIP(hipGetDevice(&dev));
  const std::array<int, 3> key = {dev, n_pad, W};
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it == cache.end()) {
    const std::vector<int> h = trace_schedule_host(MT, CB, n_pad, W, num_cus());
    void* p = nullptr;
    BTA_HIP(hipMalloc(&p, h.size() * sizeof(int)));
    const hipError_t e = hipMemcpy(p, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(p);
      return hip_fail(e, "hipMemcpy(trace schedule)");
    }
    it = cache.emplace(key, (const int*)p).first;
  }
  *out = it->second;
  return BTA_OK;
}

struct MllWs {
  int64_t n_pad;
  int W, MD;
  size_t ls, center, xs, L, Linv, fact, t, lgd, alpha, part, partial, out, total;
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
    t = p.take((size_t)n_pad * 8);
    lgd = p.take((size_t)n_pad * 8);
    alpha = p.take((size_t)n_pad * 8);
    part = p.take((size_t)(n_pad / 64) * n_pad * 8);
    partial = p.take((size_t)W * (MD + 2) * 8);
    out = p.take((size_t)(MD + 8) * 8);
    total = p.off;
  }
};

template <int MD>
static void launch_trace(int kind, const TraceArgs& ta, int W, hipStream_t st) {
  if (kind == BTA_KERNEL_MATERN52)
    hipLaunchKernelGGL((mll_trace_kernel<MD, BTA_KERNEL_MATERN52>), dim3(W), dim3(256), 0, st, ta);
  else
    hipLaunchKernelGGL((mll_trace_kernel<MD, BTA_KERNEL_RBF>), dim3(W), dim3(256), 0, st, ta);
}

// One attempt of the MLL value+grad on the stream: covariance with the given
// diagonal, factorisation (enqueued only), inversion, residual products, the
// trace terms and the final sums into out[0 .. MD + 4]; the factorisation's
// status pair is copied to out[MD + 5].
static int enqueue_mll(const MllWs& w, int kind, int64_t n, int d, double s, double diag,
                       const double* y, double c, hipStream_t st, char* base, void* fact,
                       size_t fact_bytes) {
  double* xs = (double*)(base + w.xs);
  double* L = (double*)(base + w.L);
  double* Linv = (double*)(base + w.Linv);
  double* t = (double*)(base + w.t);
  double* lgd = (double*)(base + w.lgd);
  double* alpha = (double*)(base + w.alpha);
  double* part = (double*)(base + w.part);
  double* partial = (double*)(base + w.partial);
  double* out = (double*)(base + w.out);
  const int64_t np = w.n_pad;
  int rc = covar_sym_lower(kind, xs, n, d, s, diag, L, np, np, st,
                           const_cast<int*>(potrf_status(fact, np)));
  if (rc) return rc;
  // the factorisation's launches also build X = L^-1 (zero only in the 64 x 64
  // blocks right of the diagonal blocks, all the 128-row tiles below read)
  rc = potrf(L, np, np, Linv, np, 0, 1e-8, nullptr, fact, fact_bytes, st,
             POTRF_NO_SAVE | POTRF_KEEP_UPPER | POTRF_BAND_ZERO | POTRF_ASYNC |
                 POTRF_INFO_ZEROED | POTRF_INVERT);
  if (rc) return rc;
  hipLaunchKernelGGL(trmv_lower_kernel, dim3((unsigned)((np + 3) / 4)), dim3(256), 0, st, Linv,
                     np, y, c, n, np, t, L, lgd);
  BTA_LAUNCH_CHECK("trmv_lower_kernel");
  {
    const unsigned gx = (unsigned)std::min<int64_t>((np + 255) / 256, 16);
    hipLaunchKernelGGL(trmv_lower_t_part_kernel, dim3(gx, (unsigned)(np / 64)), dim3(256), 0, st,
                       Linv, np, t, np, part);
    BTA_LAUNCH_CHECK("trmv_lower_t_part_kernel");
    hipLaunchKernelGGL(trmv_lower_t_sum_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0,
                       st, part, np, alpha);
    BTA_LAUNCH_CHECK("trmv_lower_t_sum_kernel");
  }
  TraceArgs ta;
  ta.X = Linv;
  ta.ld = np;
  ta.xs = xs;
  ta.alpha = alpha;
  ta.partial = partial;
  ta.n_pad = (int)np;
  ta.n = (int)n;
  ta.d = d;
  ta.MT = (int)(np / TR_B);
  ta.CB = TR_KC / TR_B;
  ta.s = s;
  const int tasks = trace_tasks(ta.MT, ta.CB);
  const int W = std::min(w.W, tasks);
  rc = trace_schedule(ta.MT, ta.CB, (int)np, W, &ta.sched);
  if (rc) return rc;
  if (d <= 16)
    launch_trace<16>(kind, ta, W, st);
  else
    launch_trace<32>(kind, ta, W, st);
  BTA_LAUNCH_CHECK("mll_trace_kernel");
  // the factorisation's status pair goes next to the sums (read back with them)
  const int* stat = potrf_status(fact, np);
  int* stat_out = (int*)(out + 3 + w.MD + 2);
  if (w.MD == 16)
    hipLaunchKernelGGL(mll_finish_kernel<18>, dim3(1), dim3(256), 0, st, t, lgd, alpha, n, partial,
                       W, out, stat, stat_out);
  else
    hipLaunchKernelGGL(mll_finish_kernel<34>, dim3(1), dim3(256), 0, st, t, lgd, alpha, n, partial,
                       W, out, stat, stat_out);
  BTA_LAUNCH_CHECK("mll_finish_kernel");
  return BTA_OK;
}

int mll_value_grad(int kind, const double* X, const double* y, int64_t n, int d,
                   const double* ls_host, double s, double noise, double c, int max_tries,
                   double* out_host, void* ws, size_t ws_bytes, hipStream_t st) {
  if (!X || !y || !ls_host || !out_host || n < 2 || d < 1 || d > 32 || !(s > 0) ||
      !(noise >= 0) || (kind != BTA_KERNEL_MATERN52 && kind != BTA_KERNEL_RBF)) {
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
  void* fact = base + w.fact;
  const size_t fact_bytes = std::max(potrf_ws_bytes(w.n_pad), trtri_ws_bytes(w.n_pad));
  double* out = (double*)(base + w.out);
  const int64_t np = w.n_pad;

  // small host copies go through this thread's pinned buffer when there is one:
  // [0, 32) the length scales, [64, 64 + MD + 6) the sums read back
  double* pin = pinned_host_buffer();
  if (pin) {
    std::memcpy(pin, ls_host, d * sizeof(double));
    BTA_HIP(hipMemcpyAsync(ls, pin, d * sizeof(double), hipMemcpyHostToDevice, st));
  } else {
    BTA_HIP(hipMemcpyAsync(ls, ls_host, d * sizeof(double), hipMemcpyHostToDevice, st));
  }
  int rc = scale_inputs(X, n, d, nullptr, ls, center, xs, np, st);
  if (rc) return rc;
  // psd_safe_cholesky's schedule: no jitter, then jitter0 * 10^(i-1); K is
  // rebuilt for each attempt instead of being saved.  The whole call is
  // enqueued before the factorisation's outcome is known (one host
  // synchronisation at the end, so the GPU never waits for the host between
  // the factorisation and the inversion); after a failure the results are
  // discarded and the next attempt runs.
  const double jitter0 = 1e-8;
  std::vector<double> h(3 + w.MD + 2 + 1);
  for (int attempt = 0;; ++attempt) {
    const double jit = attempt == 0 ? 0.0 : jitter0 * std::pow(10.0, (double)(attempt - 1));
    rc = enqueue_mll(w, kind, n, d, s, noise + jit, y, c, st, base, fact, fact_bytes);
    if (rc) return rc;
    double* hb = pin ? pin + 64 : h.data();
    BTA_HIP(hipMemcpyAsync(hb, out, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    BTA_HIP(hipStreamSynchronize(st));
    if (pin) std::memcpy(h.data(), hb, h.size() * sizeof(double));
    int pst[2];
    std::memcpy(pst, &h[3 + w.MD + 2], sizeof(pst));
    if (pst[0] == 0) break;
    if (pst[1]) {
      set_error("mll_value_grad: covariance contains NaN (failing column %d)", pst[0]);
      return BTA_ERR_NAN;
    }
    if (attempt >= max_tries) {
      set_error("mll_value_grad: matrix not positive definite after adding jitter up to %.1e",
                jit);
      return BTA_ERR_NOT_PSD;
    }
  }
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
