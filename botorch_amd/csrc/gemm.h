// Internal interface of the MFMA GEMM (gemm.hip) for the other kernels.
#pragma once
#include "bta_common.h"

namespace bta {

struct GemmArgs {
  const void* A;
  const void* B;
  const void* C0;
  void* C;
  int64_t lda, ldb, ldc0, ldc;
  int64_t sA, sB, sC0, sC;
  int M, N, K;
  int flags;
  double alpha, beta;
};

int launch_gemm(const GemmArgs& g, int batch, hipStream_t s);

// Convenience: C = alpha A op(B) + beta C0 (single batch).
inline int gemm(int flags, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
                int64_t lda, const void* B, int64_t ldb, double beta, const void* C0,
                int64_t ldc0, void* C, int64_t ldc, hipStream_t s, int batch = 1,
                int64_t sA = 0, int64_t sB = 0, int64_t sC0 = 0, int64_t sC = 0) {
  GemmArgs g;
  g.A = A; g.B = B; g.C0 = C0; g.C = C;
  g.lda = lda; g.ldb = ldb; g.ldc0 = ldc0; g.ldc = ldc;
  g.sA = sA; g.sB = sB; g.sC0 = sC0; g.sC = sC;
  g.M = (int)M; g.N = (int)N; g.K = (int)K;
  g.flags = flags; g.alpha = alpha; g.beta = beta;
  return launch_gemm(g, batch, s);
}

}  // namespace bta
