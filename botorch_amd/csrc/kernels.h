// Internal declarations shared between the library's translation units.
#pragma once
#include "bta_common.h"

namespace bta {

int scale_inputs(const double* X, int64_t n, int d, const double* center, const double* ls,
                 double* center_out, double* Xs, int64_t n_pad, hipStream_t st);
int covar(int kind, const double* X1s, int64_t n1, const double* X2s, int64_t n2, int d,
          double s, double diag_add, bool f32, void* K, int64_t ldk, int64_t n1_pad,
          int64_t n2_pad, hipStream_t st, int pad_identity = 0, int batch = 1,
          int64_t sX1 = 0, int64_t sX2 = 0, int64_t sK = 0);
int covar_grad(int kind, const double* X1s, int64_t n1, const double* X2s, int64_t n2, int d,
               double s, const void* G, bool g_f32, int64_t ldg, const double* extra,
               const double* ls, double* dX1, hipStream_t st, const double* dext = nullptr,
               const double* E = nullptr, int ne = 0);
int covar_ext(int kind, const double* X1s, int64_t n1, const double* X2s, int64_t n2, int d,
              double s, const double* E, int ne, double* P, hipStream_t st);
size_t potrf_ws_bytes(int64_t n_pad);
int potrf(double* A, int64_t n_pad, int64_t lda, double* Linv, int64_t ldinv, int max_tries,
          double jitter0, double* jitter_used, void* ws, size_t ws_bytes, hipStream_t s);
size_t trtri_ws_bytes(int64_t n_pad);
int trtri_lower(const double* L, int64_t ldl, double* X, int64_t ldx, int64_t n_pad, void* ws,
                size_t ws_bytes, hipStream_t s);
template <typename TO>
int cast_pad(const double* in, int64_t ldi, int64_t rows, int64_t cols, TO* out, int64_t ldo,
             int64_t rows_pad, int64_t cols_pad, hipStream_t s);
int launch_baseline_best_f(const double* ka, double y_mean, double y_std, double c,
                           const double* Lb, const double* Zb, int n_b, int S, double* best_f,
                           hipStream_t s);

int gemm_x3(int flags, int64_t M, int64_t N, int64_t K, double alpha, const float* A,
            int64_t lda, const float* a_row_scale, float a_scale, const float* B, int64_t ldb,
            float b_scale, float* C, int64_t ldc, hipStream_t s);

// Bump allocator over a caller workspace (256-byte aligned slices).
struct WsPlan {
  size_t off = 0;
  size_t take(size_t bytes) {
    size_t r = off;
    off += (bytes + 255) / 256 * 256;
    return r;
  }
};

}  // namespace bta
