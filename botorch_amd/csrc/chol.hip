// Blocked FP64 Cholesky with psd_safe_cholesky's jitter schedule, the
// triangular inverse, and a batched small Cholesky for q x q / n_b x n_b
// blocks.
//
// bta_potrf is right-looking with 64-column panels:
//   leaf   : one 256-thread block factors the 64 x 64 diagonal block in LDS and
//            inverts it (W = L_jj^-1, written into Linv's diagonal block);
//   panel  : L_ij = A_ij W^T for the blocks below - one MFMA GEMM (B lower);
//   update : A_ik -= L_ij L_kj^T on the lower trailing tiles - one MFMA GEMM.
// The trailing update dominates the flops and runs on the MFMA GEMM; the leaf
// is latency-bound.
//
// bta_trtri_lower inverts L bottom-up by recursive doubling: with the 64x64
// diagonal inverses in place, each level joins pairs of inverted diagonal
// blocks [A 0; C B] -> [A^-1 0; -B^-1 C A^-1 B^-1] with two MFMA GEMMs batched
// over all pairs of the level, so the whole inverse is ~4 log2(n/64) launches.
#include "bta_common.h"
#include "gemm.h"
#include "kernels.h"

namespace bta {

constexpr int LEAF = 64;
constexpr int LP = LEAF + 1;  // LDS pitch (odd: column walks hit distinct banks)

// Factor the n x n (n <= 64) matrix held in LDS (pitch LP) in place; returns
// -1 on success or the failing column; *nan_seen set when the pivot is NaN.
// All 256 threads must call it.
__device__ int chol_lds(double* Ls, int n, int* nan_seen) {
  __shared__ int fail;
  if (threadIdx.x == 0) fail = -1;
  const int t = threadIdx.x;
  for (int c = 0; c < n; ++c) {
    __syncthreads();
    const double p = Ls[c * LP + c];
    if (!(p > 0.0)) {  // uniform: every thread read the same pivot
      if (t == 0) {
        fail = c;
        if (p != p) *nan_seen = 1;
      }
      break;
    }
    const double dg = sqrt(p);
    if (t > c && t < n) Ls[t * LP + c] /= dg;
    __syncthreads();
    // every thread has read the pivot by now; nothing below reads L[c][c]
    if (t == c) Ls[c * LP + c] = dg;
    const int i = c + 1 + (t >> 2);
    if (i < n) {
      const double lic = Ls[i * LP + c];
      for (int j = c + 1 + (t & 3); j <= i; j += 4) Ls[i * LP + j] -= lic * Ls[j * LP + c];
    }
  }
  __syncthreads();
  return fail;
}

// Inverse of the lower-triangular n x n (n <= 64) LDS matrix: thread j < n
// produces column j of L^-1 by forward substitution and writes it to out.
__device__ void trinv_lds_store(const double* Ls, int n, double* out, int64_t ldo) {
  const int j = threadIdx.x;
  if (j >= LEAF) return;
  double x[LEAF];
#pragma unroll
  for (int i = 0; i < LEAF; ++i) {
    double acc = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < i; ++k) acc -= Ls[i * LP + k] * x[k];
    x[i] = (i < n) ? acc / Ls[i * LP + i] : 0.0;
  }
  if (j < n) {
#pragma unroll
    for (int i = 0; i < LEAF; ++i)
      if (i < n) out[(int64_t)i * ldo + j] = x[i];
  }
}

// info[0]: first failing column + 1 (0 = ok); info[1]: NaN pivot seen.
__global__ __launch_bounds__(256) void potrf_leaf_kernel(double* A, int64_t lda, double* W,
                                                         int64_t ldw, int j0, int* info) {
  __shared__ double Ls[LEAF * LP];
  __shared__ int nan_seen;
  if (info[0] != 0) return;  // an earlier panel already failed
  if (threadIdx.x == 0) nan_seen = 0;
  for (int e = threadIdx.x; e < LEAF * LEAF; e += 256) {
    int i = e / LEAF, j = e % LEAF;
    Ls[i * LP + j] = j <= i ? A[(int64_t)i * lda + j] : 0.0;
  }
  const int fail = chol_lds(Ls, LEAF, &nan_seen);
  if (fail >= 0) {
    if (threadIdx.x == 0) {
      info[0] = j0 + fail + 1;
      info[1] = nan_seen;
    }
    return;
  }
  for (int e = threadIdx.x; e < LEAF * LEAF; e += 256) {
    int i = e / LEAF, j = e % LEAF;
    A[(int64_t)i * lda + j] = j <= i ? Ls[i * LP + j] : 0.0;
  }
  trinv_lds_store(Ls, LEAF, W, ldw);
}

__global__ void fill_int_kernel(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void add_diag_kernel(double* A, int64_t lda, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[i * lda + i] += v;
}

// zero A[i][j] for j > i (whole matrix, n x n)
__global__ void zero_upper_kernel(double* A, int64_t lda, int64_t n) {
  int64_t i = blockIdx.y;
  for (int64_t j = i + 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x)
    A[i * lda + j] = 0.0;
}

__global__ void copy_lower_kernel(const double* src, int64_t lds_, double* dst, int64_t ldd,
                                  int64_t n) {
  int64_t i = blockIdx.y;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= i && j < n;
       j += (int64_t)gridDim.x * blockDim.x)
    dst[i * ldd + j] = src[i * lds_ + j];
}

// Batched small Cholesky with jitter escalation (psd_safe_cholesky semantics
// per batch element).  Strict upper triangle zeroed on output.
__global__ __launch_bounds__(256) void chol_small_kernel(double* A, int n, int64_t lda,
                                                         int64_t stride, int max_tries,
                                                         double jitter0, int* status) {
  __shared__ double Ls[LEAF * LP];
  __shared__ int nan_seen;
  double* Ab = A + (int64_t)blockIdx.x * stride;
  int st = 0;
  int fail = 0;
  for (int attempt = 0; attempt <= max_tries; ++attempt) {
    const double jit = attempt == 0 ? 0.0 : jitter0 * pow(10.0, (double)(attempt - 1));
    if (threadIdx.x == 0) nan_seen = 0;
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += 256) {
      int i = e / n, j = e % n;
      Ls[i * LP + j] = j <= i ? Ab[(int64_t)i * lda + j] + (i == j ? jit : 0.0) : 0.0;
    }
    fail = chol_lds(Ls, n, &nan_seen);
    if (fail < 0) {
      if (attempt > 0) st |= BTA_STATUS_JITTER;
      break;
    }
    if (nan_seen) {
      st |= BTA_STATUS_NAN;
      break;
    }
  }
  if (fail >= 0) st |= BTA_STATUS_NOT_PSD;
  for (int e = threadIdx.x; e < n * n; e += 256) {
    int i = e / n, j = e % n;
    Ab[(int64_t)i * lda + j] = (j <= i && fail < 0) ? Ls[i * LP + j] : (j <= i ? NAN : 0.0);
  }
  if (threadIdx.x == 0 && status) status[blockIdx.x] = st;
}

size_t potrf_ws_bytes(int64_t n_pad) {
  // saved input (n_pad^2) + one 64x64 inverse block + info words
  return (size_t)(n_pad * n_pad + LEAF * LEAF) * sizeof(double) + 256;
}

int potrf(double* A, int64_t n_pad, int64_t lda, double* Linv, int64_t ldinv, int max_tries,
          double jitter0, double* jitter_used, void* ws, size_t ws_bytes, hipStream_t s) {
  if (n_pad <= 0 || n_pad % LEAF || lda < n_pad || (Linv && ldinv < n_pad)) {
    set_error("potrf: n_pad must be a positive multiple of 64");
    return BTA_ERR_INVALID;
  }
  if (!ws || ws_bytes < potrf_ws_bytes(n_pad)) {
    set_error("potrf: workspace too small (%zu < %zu)", ws_bytes, potrf_ws_bytes(n_pad));
    return BTA_ERR_WORKSPACE;
  }
  double* saved = (double*)ws;
  double* wblk = saved + n_pad * n_pad;
  int* info = (int*)(wblk + LEAF * LEAF);
  if (jitter_used) *jitter_used = 0.0;
  {
    dim3 g((unsigned)((n_pad + 255) / 256 > 64 ? 64 : (n_pad + 255) / 256), (unsigned)n_pad);
    hipLaunchKernelGGL(copy_lower_kernel, g, dim3(256), 0, s, A, lda, saved, n_pad, n_pad);
    BTA_LAUNCH_CHECK("copy_lower_kernel");
  }
  for (int attempt = 0; attempt <= max_tries; ++attempt) {
    double jit = 0.0;
    if (attempt > 0) {
      jit = jitter0 * pow(10.0, (double)(attempt - 1));
      dim3 g((unsigned)((n_pad + 255) / 256 > 64 ? 64 : (n_pad + 255) / 256), (unsigned)n_pad);
      hipLaunchKernelGGL(copy_lower_kernel, g, dim3(256), 0, s, saved, n_pad, A, lda, n_pad);
      BTA_LAUNCH_CHECK("copy_lower_kernel");
      hipLaunchKernelGGL(add_diag_kernel, dim3((unsigned)((n_pad + 255) / 256)), dim3(256), 0,
                         s, A, lda, n_pad, jit);
      BTA_LAUNCH_CHECK("add_diag_kernel");
    }
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, info, 2, 0);
    BTA_LAUNCH_CHECK("fill_int_kernel");
    for (int64_t j0 = 0; j0 < n_pad; j0 += LEAF) {
      double* Ajj = A + j0 * lda + j0;
      double* W = Linv ? Linv + j0 * ldinv + j0 : wblk;
      const int64_t ldw = Linv ? ldinv : LEAF;
      hipLaunchKernelGGL(potrf_leaf_kernel, dim3(1), dim3(256), 0, s, Ajj, lda, W, ldw, (int)j0,
                         info);
      BTA_LAUNCH_CHECK("potrf_leaf_kernel");
      const int64_t rem = n_pad - j0 - LEAF;
      if (rem <= 0) break;
      double* P = A + (j0 + LEAF) * lda + j0;
      int rc = gemm(BTA_GEMM_B_LOWER, rem, LEAF, LEAF, 1.0, P, lda, W, ldw, 0.0, nullptr, 0, P,
                    lda, s);
      if (rc) return rc;
      double* T = A + (j0 + LEAF) * lda + j0 + LEAF;
      rc = gemm(BTA_GEMM_C_LOWER, rem, rem, LEAF, -1.0, P, lda, P, lda, 1.0, T, lda, T, lda, s);
      if (rc) return rc;
    }
    int hinfo[2];
    BTA_HIP(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, s));
    BTA_HIP(hipStreamSynchronize(s));
    if (hinfo[0] == 0) {
      dim3 g((unsigned)((n_pad + 255) / 256 > 64 ? 64 : (n_pad + 255) / 256), (unsigned)n_pad);
      hipLaunchKernelGGL(zero_upper_kernel, g, dim3(256), 0, s, A, lda, n_pad);
      BTA_LAUNCH_CHECK("zero_upper_kernel");
      if (jitter_used) *jitter_used = jit;
      return BTA_OK;
    }
    if (hinfo[1]) {
      set_error("potrf: matrix contains NaN (failing column %d)", hinfo[0]);
      return BTA_ERR_NAN;
    }
  }
  set_error("potrf: matrix not positive definite after adding jitter up to %.1e",
            jitter0 * pow(10.0, (double)(max_tries - 1)));
  return BTA_ERR_NOT_PSD;
}

size_t trtri_ws_bytes(int64_t n_pad) { return (size_t)(n_pad * n_pad / 4 + 64) * sizeof(double); }

int trtri_lower(const double* L, int64_t ldl, double* X, int64_t ldx, int64_t n_pad, void* ws,
                size_t ws_bytes, hipStream_t s) {
  if (n_pad <= 0 || n_pad % LEAF) {
    set_error("trtri: n_pad must be a positive multiple of 64");
    return BTA_ERR_INVALID;
  }
  if (!ws || ws_bytes < trtri_ws_bytes(n_pad)) {
    set_error("trtri: workspace too small");
    return BTA_ERR_WORKSPACE;
  }
  double* T = (double*)ws;
  {
    dim3 g((unsigned)((n_pad + 255) / 256 > 64 ? 64 : (n_pad + 255) / 256), (unsigned)n_pad);
    hipLaunchKernelGGL(zero_upper_kernel, g, dim3(256), 0, s, X, ldx, n_pad);
    BTA_LAUNCH_CHECK("zero_upper_kernel");
  }
  const int64_t NB = n_pad / LEAF;
  for (int64_t sb = 1; sb < NB; sb *= 2) {
    const int64_t S = sb * LEAF;
    const int64_t pf = NB / (2 * sb);  // pairs with a full second block
    if (pf > 0) {
      const int64_t strideL = 2 * S * (ldl + 1), strideX = 2 * S * (ldx + 1);
      // T_p = C_p X_A,p   (C_p = L[b0, a0], X_A = X[a0, a0] lower)
      int rc = gemm(BTA_GEMM_B_KN | BTA_GEMM_B_UPPER, S, S, S, 1.0, L + S * ldl, ldl, X, ldx,
                    0.0, nullptr, 0, T, S, s, (int)pf, strideL, strideX, 0, S * S);
      if (rc) return rc;
      // X_C,p = -X_B,p T_p
      rc = gemm(BTA_GEMM_B_KN | BTA_GEMM_A_LOWER, S, S, S, -1.0, X + S * ldx + S, ldx, T, S, 0.0,
                nullptr, 0, X + S * ldx, ldx, s, (int)pf, strideX, S * S, 0, strideX);
      if (rc) return rc;
    }
    const int64_t a0 = pf * 2 * sb, b0 = a0 + sb;
    if (b0 < NB) {  // ragged last pair: second block of sbr < sb leaves
      const int64_t SB = (NB - b0) * LEAF;
      const int64_t ra = a0 * LEAF, rb = b0 * LEAF;
      int rc = gemm(BTA_GEMM_B_KN | BTA_GEMM_B_UPPER, SB, S, S, 1.0, L + rb * ldl + ra, ldl,
                    X + ra * ldx + ra, ldx, 0.0, nullptr, 0, T, S, s);
      if (rc) return rc;
      rc = gemm(BTA_GEMM_B_KN | BTA_GEMM_A_LOWER, SB, S, SB, -1.0, X + rb * ldx + rb, ldx, T, S,
                0.0, nullptr, 0, X + rb * ldx + ra, ldx, s);
      if (rc) return rc;
    }
  }
  return BTA_OK;
}

int chol_small_batched(double* A, int n, int64_t lda, int64_t stride, int64_t batch,
                       int max_tries, double jitter0, int* status, hipStream_t s) {
  if (n < 1 || n > LEAF || lda < n || batch < 0) {
    set_error("chol_small_batched: need 1 <= n <= 64");
    return BTA_ERR_UNSUPPORTED;
  }
  if (batch == 0) return BTA_OK;
  hipLaunchKernelGGL(chol_small_kernel, dim3((unsigned)batch), dim3(256), 0, s, A, n, lda,
                     stride, max_tries, jitter0, status);
  BTA_LAUNCH_CHECK("chol_small_kernel");
  return BTA_OK;
}

}  // namespace bta

extern "C" size_t bta_potrf_workspace_size(int64_t n_pad) { return bta::potrf_ws_bytes(n_pad); }

extern "C" int bta_potrf(double* A, int64_t n_pad, int64_t lda, double* Linv, int64_t ldinv,
                         int32_t max_tries, double jitter0, double* jitter_used, void* ws,
                         size_t ws_bytes, bta_stream_t stream) {
  if (!A) {
    bta::set_error("bta_potrf: null pointer");
    return BTA_ERR_INVALID;
  }
  return bta::potrf(A, n_pad, lda, Linv, ldinv, max_tries, jitter0, jitter_used, ws, ws_bytes,
                    (hipStream_t)stream);
}

extern "C" int bta_trtri_lower(const double* L, int64_t ldl, double* Linv, int64_t ldinv,
                               int64_t n_pad, void* ws, size_t ws_bytes, bta_stream_t stream) {
  if (!L || !Linv) {
    bta::set_error("bta_trtri_lower: null pointer");
    return BTA_ERR_INVALID;
  }
  return bta::trtri_lower(L, ldl, Linv, ldinv, n_pad, ws, ws_bytes, (hipStream_t)stream);
}

extern "C" int bta_chol_small_batched(double* A, int32_t n, int64_t lda, int64_t stride,
                                      int64_t batch, int32_t max_tries, double jitter0,
                                      int32_t* status, bta_stream_t stream) {
  if (!A) {
    bta::set_error("bta_chol_small_batched: null pointer");
    return BTA_ERR_INVALID;
  }
  return bta::chol_small_batched(A, n, lda, stride, batch, max_tries, jitter0, status,
                                 (hipStream_t)stream);
}
