// Device-resident lock-step batched L-BFGS-B: the optimiser of optimize_acqf
// (optim/batched_lbfgs_b.py:169-636, one SciPy setulb state per restart,
// generation/gen.py:236-249) kept on the GPU.
//
// Each problem (restart) owns a slab in HBM holding its scalar state, x, g, the
// saved iterate and gradient, the search direction, the Cauchy point, the
// correction pairs S and Y (n x m, one column per pair) and the small m x m /
// 2m x 2m matrices of the compact representation.  One round
// (bta_lbfgsb_step) takes f and g of every problem at the point it asked for,
// advances each problem's reverse-communication machine until it needs a new
// evaluation or ends, and writes the next points to x_eval.  Caller loop:
//   init -> x_eval;  repeat { f, g = evaluate(x_eval); step -> x_eval }
// with no host round trip: a problem that ended keeps its final x in x_eval
// and ignores later f, g; *ndone counts the ended problems.
//
// One 64-lane workgroup per problem stages the slab into LDS, runs the machine
// (loops over variables and matrix entries across lanes, the breakpoint heap,
// the factorisations and the line search uniformly on every lane) and writes
// the slab back.  Ended problems return after reading one flag.
#include "bta_common.h"
#include "lbfgsb_core.h"

namespace bta {
namespace {

using lbfgsb::Layout;
using lbfgsb::Scal;
using lbfgsb::View;

constexpr int LB_LANES = 64;
constexpr int64_t LB_MAX_LDS = 64 * 1024;

struct WaveLanes {
  static constexpr int W = LB_LANES;
  __device__ int lane() const { return (int)threadIdx.x; }
  __device__ double sum(double v) const {
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
  }
  __device__ double max(double v) const {
    for (int o = W / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, W));
    return v;
  }
  __device__ double min(double v) const {
    for (int o = W / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, W));
    return v;
  }
  __device__ int sum_int(int v) const {
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
  }
  __device__ int min_int(int v) const {
    for (int o = W / 2; o > 0; o >>= 1) v = ::min(v, __shfl_xor(v, o, W));
    return v;
  }
  __device__ int any(int v) const { return __ballot(v != 0) != 0ull; }
  // position of this lane among the lanes with flag set, and their count
  __device__ int ballot_rank(int flag, int& total) const {
    const unsigned long long b = __ballot(flag != 0);
    total = __popcll(b);
    return __popcll(b & ((1ull << threadIdx.x) - 1ull));
  }
  __device__ void sync() const { __syncthreads(); }
  __device__ void mark(int) const {}
  // v of lane j (j uniform)
  __device__ double bcast(double v, int j) const {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
    return __hiloint2double(hi, lo);
  }
  // upper Cholesky row by row: row j's diagonal from the column above it, then
  // the rest of row j across lanes (serial_chol_upper's operations and order)
  __device__ int chol_upper(double* A, int ld, int k) const {
    const int l0 = lane();
    for (int j = 0; j < k; ++j) {
      double s = 0.0;
      for (int q = 0; q < j; ++q) s += A[q * ld + j] * A[q * ld + j];
      s = A[j * ld + j] - s;
      __syncthreads();
      if (!(s > 0.0)) return j + 1;
      const double ujj = sqrt(s);
      if (l0 == 0) A[j * ld + j] = ujj;
      for (int l = j + 1 + l0; l < k; l += W) {
        double t = A[j * ld + l];
        for (int i = 0; i < j; ++i) t -= A[i * ld + j] * A[i * ld + l];
        A[j * ld + l] = t / ujj;
      }
      __syncthreads();
    }
    return 0;
  }
  // R^T y = b: lane i holds b_i; step j finishes y_j and removes it from the
  // later entries (serial_solve_rt's operations and order)
  __device__ int solve_rt(const double* R, int ld, int k, double* b) const {
    if (k > W) return lbfgsb::serial_solve_rt(R, ld, k, b);
    for (int j = 0; j < k; ++j)
      if (R[j * ld + j] == 0.0) return j + 1;
    const int l0 = lane();
    double bi = l0 < k ? b[l0] : 0.0;
    for (int j = 0; j < k; ++j) {
      const double yj = bcast(bi, j) / R[j * ld + j];
      if (l0 == j) bi = yj;
      else if (l0 > j && l0 < k) bi -= R[j * ld + l0] * yj;
    }
    __syncthreads();
    if (l0 < k) b[l0] = bi;
    __syncthreads();
    return 0;
  }
  // R y = b from the last column (serial_solve_r's operations and order)
  __device__ int solve_r(const double* R, int ld, int k, double* b) const {
    if (k > W) return lbfgsb::serial_solve_r(R, ld, k, b);
    for (int j = 0; j < k; ++j)
      if (R[j * ld + j] == 0.0) return j + 1;
    const int l0 = lane();
    double bi = l0 < k ? b[l0] : 0.0;
    for (int j = k - 1; j >= 0; --j) {
      const double yj = bcast(bi, j) / R[j * ld + j];
      if (l0 == j) bi = yj;
      else if (l0 < j) bi -= R[l0 * ld + j] * yj;
    }
    __syncthreads();
    if (l0 < k) b[l0] = bi;
    __syncthreads();
    return 0;
  }
};

struct LbArgs {
  double* state;  // b slabs of slab doubles
  int64_t b;
  int n, m, slab;
  const double* lo;
  const double* up;
  const int32_t* nbd;
  const double* fval;  // b (step)
  const double* gval;  // b x n (step)
  double fscale;
  const int32_t* step_status;  // b, optional
  double* x_eval;              // b x n
  int32_t* ndone;              // optional
  const double* x0;            // b x n (init)
  double factr, pgtol;
  int maxls, maxiter, maxfun, init;
};

__global__ __launch_bounds__(LB_LANES) void lbfgsb_kernel(LbArgs a) {
  extern __shared__ double lds[];
  const int64_t r = blockIdx.x;
  const int lane = (int)threadIdx.x;
  double* slab = a.state + r * (int64_t)a.slab;
  const Layout L(a.n, a.m);
  if (!a.init && ((const Scal*)slab)->done) return;
  if (a.init) {
    for (int k = lane; k < a.slab; k += LB_LANES) lds[k] = 0.0;
    __syncthreads();
    if (lane == 0) lbfgsb::init_scal(*(Scal*)lds, a.factr, a.pgtol, a.maxls, a.maxiter, a.maxfun);
    for (int i = lane; i < a.n; i += LB_LANES) lds[L.o_x + i] = a.x0[r * a.n + i];
  } else {
    for (int k = lane; k < a.slab; k += LB_LANES) lds[k] = slab[k];
  }
  __syncthreads();
  WaveLanes c;
  View w = lbfgsb::make_view(lds, L, a.lo, a.up, a.nbd);
  const double* f = a.init ? nullptr : a.fval + r;
  const double* g = a.init ? nullptr : a.gval + r * a.n;
  const int32_t st = (a.init || !a.step_status) ? 0 : a.step_status[r];
  lbfgsb::advance(c, w, f, g, a.fscale, st);
  __syncthreads();
  for (int i = lane; i < a.n; i += LB_LANES) a.x_eval[r * a.n + i] = w.x[i];
  for (int k = lane; k < a.slab; k += LB_LANES) slab[k] = lds[k];
  if (lane == 0 && w.s->done && a.ndone) atomicAdd(a.ndone, 1);
}

__global__ void lbfgsb_zero_kernel(int32_t* p) { *p = 0; }

// x (b x n), f (b), info (b x 6: nit, nfev, warnflag, task, step_status,
// number of pair updates)
__global__ __launch_bounds__(LB_LANES) void lbfgsb_results_kernel(const double* state, int n,
                                                                  int m, int slab, double* x,
                                                                  double* f, int32_t* info) {
  const int64_t r = blockIdx.x;
  const double* s0 = state + r * (int64_t)slab;
  const Layout L(n, m);
  const Scal* s = (const Scal*)s0;
  for (int i = threadIdx.x; i < n; i += LB_LANES) x[r * n + i] = s0[L.o_x + i];
  if (threadIdx.x == 0) {
    f[r] = s->f;
    int32_t* o = info + r * 6;
    o[0] = s->nit;
    o[1] = s->nfev;
    o[2] = s->warnflag;
    o[3] = s->task;
    o[4] = s->step_status;
    o[5] = s->iupdat;
  }
}

int lb_check(int64_t b, int n, int m, const char* what) {
  if (b < 0 || n < 1 || m < 1) {
    set_error("%s: need b >= 0, n >= 1, m >= 1", what);
    return BTA_ERR_INVALID;
  }
  const Layout L(n, m);
  if ((int64_t)L.total * 8 > LB_MAX_LDS) {
    set_error("%s: n = %d variables with m = %d pairs need %lld bytes of LDS per problem "
              "(limit %lld)",
              what, n, m, (long long)L.total * 8, (long long)LB_MAX_LDS);
    return BTA_ERR_UNSUPPORTED;
  }
  return BTA_OK;
}

int lb_launch(const LbArgs& a, hipStream_t st) {
  if (a.b == 0) return BTA_OK;
  const Layout L(a.n, a.m);
  hipLaunchKernelGGL(lbfgsb_kernel, dim3((unsigned)a.b), dim3(LB_LANES),
                     (size_t)L.total * sizeof(double), st, a);
  BTA_LAUNCH_CHECK("lbfgsb_kernel");
  return BTA_OK;
}

}  // namespace
}  // namespace bta

extern "C" size_t bta_lbfgsb_state_size(int32_t n, int32_t m) {
  if (bta::lb_check(1, n, m, "lbfgsb_state_size") != BTA_OK) return 0;
  return (size_t)bta::lbfgsb::Layout(n, m).slab * sizeof(double);
}

extern "C" int bta_lbfgsb_init(void* state, int64_t b, int32_t n, int32_t m, const double* x0,
                               const double* lower, const double* upper, const int32_t* nbd,
                               double factr, double pgtol, int32_t maxls, int32_t maxiter,
                               int32_t maxfun, double* x_eval, int32_t* ndone,
                               bta_stream_t stream) {
  int rc = bta::lb_check(b, n, m, "lbfgsb_init");
  if (rc != BTA_OK) return rc;
  if (!state || !x0 || !lower || !upper || !nbd || !x_eval || !(factr >= 0.0) ||
      !(pgtol >= 0.0) || maxls < 1 || maxiter < 0 || maxfun < 0) {
    bta::set_error("lbfgsb_init: invalid arguments (null pointer, negative tolerance or "
                   "maxls < 1)");
    return BTA_ERR_INVALID;
  }
  hipStream_t st = (hipStream_t)stream;
  if (ndone) {
    hipLaunchKernelGGL(bta::lbfgsb_zero_kernel, dim3(1), dim3(1), 0, st, ndone);
    BTA_LAUNCH_CHECK("lbfgsb_zero_kernel");
  }
  bta::LbArgs a{};
  a.state = (double*)state;
  a.b = b;
  a.n = n;
  a.m = m;
  a.slab = bta::lbfgsb::Layout(n, m).slab;
  a.lo = lower;
  a.up = upper;
  a.nbd = nbd;
  a.fscale = 1.0;
  a.x_eval = x_eval;
  a.ndone = ndone;
  a.x0 = x0;
  a.factr = factr;
  a.pgtol = pgtol;
  a.maxls = maxls;
  a.maxiter = maxiter;
  a.maxfun = maxfun;
  a.init = 1;
  return bta::lb_launch(a, st);
}

extern "C" int bta_lbfgsb_step(void* state, int64_t b, int32_t n, int32_t m, const double* lower,
                               const double* upper, const int32_t* nbd, const double* f,
                               const double* g, double fscale, const int32_t* step_status,
                               double* x_eval, int32_t* ndone, bta_stream_t stream) {
  int rc = bta::lb_check(b, n, m, "lbfgsb_step");
  if (rc != BTA_OK) return rc;
  if (!state || !lower || !upper || !nbd || !f || !g || !x_eval) {
    bta::set_error("lbfgsb_step: null pointer argument");
    return BTA_ERR_INVALID;
  }
  bta::LbArgs a{};
  a.state = (double*)state;
  a.b = b;
  a.n = n;
  a.m = m;
  a.slab = bta::lbfgsb::Layout(n, m).slab;
  a.lo = lower;
  a.up = upper;
  a.nbd = nbd;
  a.fval = f;
  a.gval = g;
  a.fscale = fscale;
  a.step_status = step_status;
  a.x_eval = x_eval;
  a.ndone = ndone;
  a.init = 0;
  return bta::lb_launch(a, (hipStream_t)stream);
}

extern "C" int bta_lbfgsb_results(const void* state, int64_t b, int32_t n, int32_t m, double* x,
                                  double* f, int32_t* info, bta_stream_t stream) {
  int rc = bta::lb_check(b, n, m, "lbfgsb_results");
  if (rc != BTA_OK) return rc;
  if (!state || !x || !f || !info) {
    bta::set_error("lbfgsb_results: null pointer argument");
    return BTA_ERR_INVALID;
  }
  if (b == 0) return BTA_OK;
  hipLaunchKernelGGL(bta::lbfgsb_results_kernel, dim3((unsigned)b), dim3(bta::LB_LANES), 0,
                     (hipStream_t)stream, (const double*)state, n, m,
                     bta::lbfgsb::Layout(n, m).slab, x, f, info);
  BTA_LAUNCH_CHECK("lbfgsb_results_kernel");
  return BTA_OK;
}
