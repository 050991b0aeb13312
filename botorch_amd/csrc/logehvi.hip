// Fused log-areas of qLogEHVI / qLogNEHVI: the reference's only native op
// (csrc/logei_fused.cpp:184-373, bound at acquisition/multi_objective/
// logei.py:107-170), as HIP kernels.
//
// For objective subsets obj[b][s][j][k] (b < B, s < n_sub subsets of size i,
// j < i, k < m objectives) and hyper-rectangle cells [cl, cu] (nc cells, shared
// or one set per b):
//   li[j][k]  = log_fatplus(obj[b][s][j][k] - cl[c][k], tau_relu)
//   lim[k]    = fatmin_j(li[j][k], tau_max)
//   lcl[k]    = log(min(cu[c][k], clamp) - cl[c][k])
//   out[b][c][s] = sum_k fatmin(lim[k], lcl[k], tau_max)
// (acquisition/multi_objective/logei.py:437-479, steps 1-4 of
// _compute_log_qehvi).  fatmin(x) = mn - tau log sum_j pareto((x_j - mn) / tau)
// with pareto(z) = 2 / (2 + 2z + z^2) and mn = the first minimum; its gradient
// goes to every x_j through the Pareto weights and to the first minimum through
// the 1 + ... term (csrc/logei_fused.cpp:117-177 has the same closed form).
//
// Forward: one thread per output element (b, c, s), the objective loop inside.
// Backward: the objectives do not mix, so one thread per (b, s, k) walks the
// cells and accumulates d out / d obj[b][s][:][k] for its k in registers
// (at most I_MAX = 32 subset members); nothing is accumulated atomically.
#include "bta_common.h"

namespace bta {

constexpr int LA_MAX_I = 32;
constexpr int LA_MAX_M = 8;
constexpr double LA_ALPHA = 0.1;  // Cauchy weight of the fat softplus

template <typename T>
__device__ __forceinline__ T la_softplus(T y) {
  if (y > T(20)) return y;
  if (y < T(-20)) return exp(y);
  return log1p(exp(y));
}

template <typename T>
__device__ __forceinline__ T la_sigmoid(T y) {
  if (y >= T(0)) return T(1) / (T(1) + exp(-y));
  const T e = exp(y);
  return e / (T(1) + e);
}

// log(tau (softplus(x / tau) + 0.1 / (1 + (x / tau)^2))) and, when dv is not
// null, its derivative in x; a non-positive argument of the log gives -1e30.
template <typename T>
__device__ __forceinline__ T la_log_fatplus(T x, T tau, T inv_tau, T* dv) {
  const T y = x * inv_tau;
  const T cy = T(1) / (T(1) + y * y);
  const T f = la_softplus(y) + T(LA_ALPHA) * cy;
  const T tf = tau * f;
  if (!(tf > T(0))) {
    if (dv) *dv = T(0);
    return T(-1e30);
  }
  if (dv) *dv = (la_sigmoid(y) - T(2 * LA_ALPHA) * y * cy * cy) / tf;
  return log(tf);
}

// fatmin over x[0..n); gw (n values, may be null) receives d fatmin / d x.
template <typename T, int N>
__device__ __forceinline__ T la_fatmin(const T (&x)[N], int n, T tau, T inv_tau, T (&gw)[N],
                                       bool want_grad) {
  T mn = x[0];
  int ami = 0;
#pragma unroll
  for (int j = 1; j < N; ++j)
    if (j < n && x[j] < mn) {
      mn = x[j];
      ami = j;
    }
  if (n == 1 || mn < T(-1e29)) {
    if (want_grad)
#pragma unroll
      for (int j = 0; j < N; ++j) gw[j] = j == ami ? T(1) : T(0);
    return mn;
  }
  T S = T(0), Sd = T(0);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    if (j < n) {
      const T z = (x[j] - mn) * inv_tau;
      const T den = T(2) + T(2) * z + z * z;
      S += T(2) / den;
      if (want_grad) {
        const T d = T(-2) * (T(2) + T(2) * z) / (den * den);
        gw[j] = d;  // the Pareto derivative, turned into the weight below
        Sd += d;
      }
    }
  }
  if (want_grad) {
    const T rS = T(1) / S;
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (j < n) gw[j] = j == ami ? T(1) + (Sd + T(1)) * rS : -gw[j] * rS;
  }
  return mn - tau * log(S);
}

template <typename T>
struct LaArgs {
  const T* obj;  // B x n_sub x i x m
  const T* cl;   // B_cells x nc x m
  const T* cu;
  const T* g_out;  // B x nc x n_sub (backward)
  T* out;          // forward: B x nc x n_sub; backward: B x n_sub x i x m
  int64_t B, n_sub, B_cells, nc;
  int i, m;
  T tau_relu, tau_max, clamp;
};

template <typename T, int I>
__global__ __launch_bounds__(256) void log_areas_fwd_kernel(LaArgs<T> a) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = a.B * a.nc * a.n_sub;
  if (e >= total) return;
  const int64_t s = e % a.n_sub, c = (e / a.n_sub) % a.nc, b = e / (a.n_sub * a.nc);
  const int64_t bc = a.B_cells == 1 ? 0 : b;
  const T* o = a.obj + (b * a.n_sub + s) * a.i * a.m;
  const T* l = a.cl + (bc * a.nc + c) * a.m;
  const T* u = a.cu + (bc * a.nc + c) * a.m;
  const T itr = T(1) / a.tau_relu, itm = T(1) / a.tau_max;
  T area = T(0);
  for (int k = 0; k < a.m; ++k) {
    const T lk = l[k];
    T li[I], unused[I];
#pragma unroll
    for (int j = 0; j < I; ++j)
      li[j] = j < a.i ? la_log_fatplus<T>(o[j * a.m + k] - lk, a.tau_relu, itr, nullptr) : T(0);
    T pair[2] = {la_fatmin<T, I>(li, a.i, a.tau_max, itm, unused, false),
                 log(fmin(u[k], a.clamp) - lk)};
    T g2[2];
    area += la_fatmin<T, 2>(pair, 2, a.tau_max, itm, g2, false);
  }
  a.out[e] = area;
}

template <typename T, int I>
__global__ __launch_bounds__(256) void log_areas_bwd_kernel(LaArgs<T> a) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = a.B * a.n_sub * a.m;
  if (e >= total) return;
  const int k = (int)(e % a.m);
  const int64_t s = (e / a.m) % a.n_sub, b = e / (a.m * a.n_sub);
  const int64_t bc = a.B_cells == 1 ? 0 : b;
  const T* o = a.obj + (b * a.n_sub + s) * a.i * a.m;
  const T itr = T(1) / a.tau_relu, itm = T(1) / a.tau_max;
  T ok[I], g[I];
#pragma unroll
  for (int j = 0; j < I; ++j) {
    ok[j] = j < a.i ? o[j * a.m + k] : T(0);
    g[j] = T(0);
  }
  for (int64_t c = 0; c < a.nc; ++c) {
    const T lk = a.cl[(bc * a.nc + c) * a.m + k];
    const T uk = a.cu[(bc * a.nc + c) * a.m + k];
    const T go = a.g_out[(b * a.nc + c) * a.n_sub + s];
    T li[I], lg[I], fw[I];
#pragma unroll
    for (int j = 0; j < I; ++j) {
      T d = T(0);
      li[j] = j < a.i ? la_log_fatplus<T>(ok[j] - lk, a.tau_relu, itr, &d) : T(0);
      lg[j] = d;
    }
    T pair[2] = {la_fatmin<T, I>(li, a.i, a.tau_max, itm, fw, true),
                 log(fmin(uk, a.clamp) - lk)};
    T g2[2];
    la_fatmin<T, 2>(pair, 2, a.tau_max, itm, g2, true);
    const T gl = go * g2[0];
#pragma unroll
    for (int j = 0; j < I; ++j) g[j] += gl * fw[j] * lg[j];
  }
  T* go = a.out + (b * a.n_sub + s) * a.i * a.m;
#pragma unroll
  for (int j = 0; j < I; ++j)
    if (j < a.i) go[j * a.m + k] = g[j];
}

template <typename T>
static int log_areas_launch(bool bwd, const LaArgs<T>& a, hipStream_t st) {
  const int64_t total = bwd ? a.B * a.n_sub * a.m : a.B * a.nc * a.n_sub;
  if (total == 0) return BTA_OK;
  const dim3 grid((unsigned)((total + 255) / 256));
  const int I = a.i <= 4 ? 4 : a.i <= 8 ? 8 : a.i <= 16 ? 16 : 32;
#define LA_GO(II)                                                                  \
  if (bwd) hipLaunchKernelGGL((log_areas_bwd_kernel<T, II>), grid, dim3(256), 0, st, a); \
  else hipLaunchKernelGGL((log_areas_fwd_kernel<T, II>), grid, dim3(256), 0, st, a);
  switch (I) {
    case 4: LA_GO(4) break;
    case 8: LA_GO(8) break;
    case 16: LA_GO(16) break;
    default: LA_GO(32) break;
  }
#undef LA_GO
  BTA_LAUNCH_CHECK(bwd ? "log_areas_bwd_kernel" : "log_areas_fwd_kernel");
  return BTA_OK;
}

static int log_areas(bool bwd, int f32, const void* g_out, const void* obj, int64_t B,
                     int64_t n_sub, int i, int m, const void* cl, const void* cu, int64_t B_cells,
                     int64_t nc, double tau_relu, double tau_max, void* out, hipStream_t st) {
  if (!obj || !cl || !cu || !out || (bwd && !g_out) || B < 0 || n_sub < 0 || nc < 0 ||
      i < 1 || i > LA_MAX_I || m < 1 || m > LA_MAX_M || !(B_cells == 1 || B_cells == B) ||
      !(tau_relu > 0.0) || !(tau_max > 0.0) || B * nc * n_sub > ((int64_t)1 << 40)) {
    set_error("log_areas: invalid arguments (need 1 <= i <= %d, 1 <= m <= %d, B_cells 1 or B, "
              "positive temperatures)",
              LA_MAX_I, LA_MAX_M);
    return BTA_ERR_INVALID;
  }
  if (f32) {
    LaArgs<float> a{(const float*)obj, (const float*)cl, (const float*)cu, (const float*)g_out,
                    (float*)out, B, n_sub, B_cells, nc, i, m, (float)tau_relu,
                    (float)tau_max, 1e8f};
    return log_areas_launch<float>(bwd, a, st);
  }
  LaArgs<double> a{(const double*)obj, (const double*)cl, (const double*)cu,
                   (const double*)g_out, (double*)out, B, n_sub, B_cells, nc, i, m, tau_relu,
                   tau_max, 1e10};
  return log_areas_launch<double>(bwd, a, st);
}

}  // namespace bta

extern "C" int bta_log_areas_forward(int32_t f32, const void* obj, int64_t B, int64_t n_sub,
                                     int32_t i, int32_t m, const void* cell_lower,
                                     const void* cell_upper, int64_t B_cells, int64_t nc,
                                     double tau_relu, double tau_max, void* out,
                                     bta_stream_t stream) {
  return bta::log_areas(false, f32, nullptr, obj, B, n_sub, i, m, cell_lower, cell_upper,
                        B_cells, nc, tau_relu, tau_max, out, (hipStream_t)stream);
}

extern "C" int bta_log_areas_backward(int32_t f32, const void* grad_out, const void* obj,
                                      int64_t B, int64_t n_sub, int32_t i, int32_t m,
                                      const void* cell_lower, const void* cell_upper,
                                      int64_t B_cells, int64_t nc, double tau_relu,
                                      double tau_max, void* grad_obj, bta_stream_t stream) {
  return bta::log_areas(true, f32, grad_out, obj, B, n_sub, i, m, cell_lower, cell_upper,
                        B_cells, nc, tau_relu, tau_max, grad_obj, (hipStream_t)stream);
}
