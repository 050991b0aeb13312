// Shared definitions for the botorch_amd HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/botorch_amd.h"

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

namespace bta {

// Thread-local last-error text, returned by bta_last_error().
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
int num_cus();
double* pinned_host_buffer();  // 512 doubles, per thread; nullptr if unavailable

#define BTA_HIP(call)                                                   \
  do {                                                                  \
    hipError_t _e = (call);                                             \
    if (_e != hipSuccess) return ::bta::hip_fail(_e, #call);            \
  } while (0)

#define BTA_LAUNCH_CHECK(what)                                          \
  do {                                                                  \
    hipError_t _e = hipGetLastError();                                  \
    if (_e != hipSuccess) return ::bta::hip_fail(_e, what);             \
  } while (0)

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Matern-5/2 / RBF value and the radial factor of its gradient, in scaled
// coordinates (x / lengthscale).  For r^2 = |xs - ts|^2:
//   Matern-5/2: k = s (1 + sqrt5 r + 5/3 r^2) exp(-sqrt5 r)
//               dk/dxs = -s 5/3 (1 + sqrt5 r) exp(-sqrt5 r) (xs - ts) = -h (xs - ts)
//   RBF:        k = s exp(-r^2/2),  dk/dxs = -s exp(-r^2/2) (xs - ts)
__device__ __forceinline__ double kern_value(int kind, double r2, double s) {
  if (kind == BTA_KERNEL_MATERN52) {
    double r = sqrt(r2);
    double sr = 2.2360679774997896964 * r;
    return s * (1.0 + sr + (5.0 / 3.0) * r2) * exp(-sr);
  }
  return s * exp(-0.5 * r2);
}

// returns h with dk/dxs = -h (xs - ts); also the value through *k
__device__ __forceinline__ double kern_grad_factor(int kind, double r2, double s, double* k) {
  if (kind == BTA_KERNEL_MATERN52) {
    double r = sqrt(r2);
    double sr = 2.2360679774997896964 * r;
    double e = exp(-sr);
    if (k) *k = s * (1.0 + sr + (5.0 / 3.0) * r2) * e;
    return s * (5.0 / 3.0) * (1.0 + sr) * e;
  }
  double e = exp(-0.5 * r2);
  if (k) *k = s * e;
  return s * e;
}

// e^x for x <= 0 (the kernels' arguments): Cody-Waite reduction x = k ln2 + f
// with ln2 split in two (|f| <= ln2/2), then the degree-12 Taylor polynomial
// of e^f (remainder < 2e-16 relative) and an exact power-of-two scaling.
// Below -745 the result underflows to 0; NaN propagates.
__host__ __device__ __forceinline__ double exp_nonpos(double x) {
  const bool under = x < -745.2;
  x = under ? 0.0 : x;
  const double kf = rint(x * 1.4426950408889634074);
  double f = fma(kf, -6.93147180369123816490e-01, x);
  f = fma(kf, -1.90821492927058770002e-10, f);
  double p = 1.0 / 479001600.0;  // 1/12!
  p = fma(p, f, 1.0 / 39916800.0);
  p = fma(p, f, 1.0 / 3628800.0);
  p = fma(p, f, 1.0 / 362880.0);
  p = fma(p, f, 1.0 / 40320.0);
  p = fma(p, f, 1.0 / 5040.0);
  p = fma(p, f, 1.0 / 720.0);
  p = fma(p, f, 1.0 / 120.0);
  p = fma(p, f, 1.0 / 24.0);
  p = fma(p, f, 1.0 / 6.0);
  p = fma(p, f, 0.5);
  p = fma(p, f, 1.0);
  p = fma(p, f, 1.0);
  return under ? 0.0 : ldexp(p, (int)kf);
}

// sqrt(r2) for r2 >= 1e-30: v_rsq_f64 seed, one Goldschmidt step for the
// root and half-reciprocal pair, two Newton corrections of the root.
__device__ __forceinline__ double sqrt_pos(double r2) {
  const double y = __builtin_amdgcn_rsq(r2);
  double r = r2 * y, h = 0.5 * y;
  const double e = fma(-r, h, 0.5);
  r = fma(r, e, r);
  h = fma(h, e, h);
  double d = fma(-r, r, r2);
  r = fma(d, h, r);
  d = fma(-r, r, r2);
  return fma(d, h, r);
}

// Compile-time kernel kind: one code path per kernel instantiation.  The
// Matern distance is clamped at 1e-30 as gpytorch's MaternKernel does.
template <int KIND>
__device__ __forceinline__ double kern_value_k(double r2, double s) {
  if (KIND == BTA_KERNEL_MATERN52) {
    const double r = sqrt_pos(fmax(r2, 1e-30));
    const double sr = 2.2360679774997896964 * r;
    return s * (1.0 + sr + (5.0 / 3.0) * r2) * exp_nonpos(-sr);
  }
  return s * exp_nonpos(-0.5 * r2);
}

template <int KIND>
__device__ __forceinline__ double kern_grad_factor_k(double r2, double s) {
  if (KIND == BTA_KERNEL_MATERN52) {
    const double r = sqrt_pos(fmax(r2, 1e-30));
    const double sr = 2.2360679774997896964 * r;
    return s * (5.0 / 3.0) * (1.0 + sr) * exp_nonpos(-sr);
  }
  return s * exp_nonpos(-0.5 * r2);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// NV independent wave sums with the shuffles of each butterfly step issued
// together: one cross-lane latency per step instead of one per value and step.
// Each value is summed in the same order as wave_sum.
template <int NV>
__device__ __forceinline__ void wave_sum_n(double (&v)[NV]) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double t[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) t[i] = __shfl_xor(v[i], off, 64);
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] += t[i];
  }
}

// NV <= 32 wave sums with P - 1 + (6 - log2 P) shuffles instead of 6 NV
// (P = NV rounded up to a power of two): the first log2 P butterfly steps
// halve the values each lane carries (the lane with the step's bit set keeps
// the upper half, its partner the lower), the remaining steps are ordinary
// butterflies.  The total of value e ends in lanes e * 64/P .. (e + 1) * 64/P - 1;
// the first of them gets *e_out = e, every other lane -1 (also lanes past NV).
// Every value is summed over the same tree as wave_sum: bit-identical totals.
template <int NV>
__device__ __forceinline__ double wave_sum_spread(const double (&v)[NV], int* e_out) {
  static_assert(NV >= 1 && NV <= 32, "wave_sum_spread: 1..32 values");
  constexpr int P = NV <= 1 ? 1 : NV <= 2 ? 2 : NV <= 4 ? 4 : NV <= 8 ? 8 : NV <= 16 ? 16 : 32;
  constexpr int LG = P == 1 ? 0 : P == 2 ? 1 : P == 4 ? 2 : P == 8 ? 3 : P == 16 ? 4 : 5;
  const int lane = threadIdx.x & 63;
  double a[P];
#pragma unroll
  for (int i = 0; i < P; ++i) a[i] = i < NV ? v[i] : 0.0;
#pragma unroll
  for (int s = 0; s < LG; ++s) {
    const int h = P >> (s + 1), off = 32 >> s;
    const bool up = (lane & off) != 0;
#pragma unroll
    for (int i = 0; i < h; ++i) {
      // The empty asm (no instruction) makes lo and hi opaque values: without
      // it the compiler turns the lane-dependent choice below into a select of
      // the element index and extracts a[index] through compare/select chains
      // over the whole array (about 7x the instructions at NV = 10).
      double lo = a[i], hi = a[h + i];
      __asm__("" : "+v"(lo), "+v"(hi));
      const double keep = up ? hi : lo;
      const double send = up ? lo : hi;
      a[i] = keep + __shfl_xor(send, off, 64);
    }
  }
  double t = a[0];
#pragma unroll
  for (int off = 32 >> LG; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  const int e = lane >> (6 - LG);
  *e_out = ((lane & ((64 >> LG) - 1)) == 0 && e < NV) ? e : -1;
  return t;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

}  // namespace bta
