// Split-FP16 GEMM for FP32 operands on gfx950 ("x3"):
//   a = a_hi + a_lo, b = b_hi + b_lo with a_hi = fp16(a), a_lo = fp16(a - a_hi),
//   a b ~= a_hi b_hi + a_hi b_lo + a_lo b_hi        (22-bit operands)
// three v_mfma_f32_16x16x32_f16 per 16 x 16 x 32 block, FP32 accumulation,
// flushed into FP64 sums every 256 K (as BTA_GEMM_ACC64).  The dropped a_lo b_lo
// term is 2^-22 relative.  Operands are scaled by powers of two (exact) so they
// sit in the FP16 range: a per-row scale for A (or one scalar) and one scalar
// for B; the epilogue divides them out.
//
// FP16 MFMA on CDNA4 runs at 16x the FP32-MFMA rate, so the three products
// cost ~5x less MFMA time than one FP32 product.  Operands are split once, while
// staging into LDS (hi and lo planes), so every wave reads ready fragments.
//
// Tile: 128 x 64 outputs, 256 threads = 4 waves (2 x 2), K chunk 32, two LDS
// stages.  Tile order, triangular K-ranges and the worker schedule are those
// of gemm_kernel.
#include "bta_common.h"
#include "gemm.h"
#include "gemm_tile.h"

#include <algorithm>
#include <vector>

namespace bta {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int X3_BM = 128, X3_BN = 64, X3_BK = 32;
constexpr int X3_PITCH = X3_BK + 8;  // halfs per LDS row (80 bytes: 16-byte aligned rows)
constexpr int X3_A = X3_BM * X3_PITCH, X3_B = X3_BN * X3_PITCH;
constexpr int X3_STAGE = 2 * (X3_A + X3_B);  // hi and lo planes
constexpr int X3_SEG = 256;                   // K per FP64 flush
constexpr int X3_MI = X3_BM / 32, X3_NI = X3_BN / 32;

struct X3Args {
  const float* A;
  const float* B;
  float* C;
  const float* a_row_scale;  // optional, per row of A (powers of two)
  float a_scale, b_scale;    // scalars (powers of two)
  int64_t lda, ldb, ldc;
  int M, N, K, flags;
  double alpha;
};

__device__ __forceinline__ void split4(float4v v, float s, half4& hi, half4& lo) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float x = v[i] * s;
    const _Float16 h = (_Float16)x;
    hi[i] = h;
    lo[i] = (_Float16)(x - (float)h);
  }
}

template <bool BKN>
__global__ __launch_bounds__(256) void gemm_x3_kernel(X3Args g) {
  __shared__ __attribute__((aligned(16))) _Float16 lds[2 * X3_STAGE];
  const int MT = g.M / X3_BM, NT = g.N / X3_BN;
  const int total = MT * NT;
  const int W = gridDim.x, w = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wy = wv >> 1, wx = wv & 1, l16 = lane & 15, gq = lane >> 4;

  for (int jj = 0;; ++jj) {
    const int t = (jj & 1) ? (jj + 1) * W - 1 - w : jj * W + w;
    if (t >= total) break;
    int tm, tn;
    decode_tile(t, g.flags, MT, NT, tm, tn);
    int kb = 0, ke = g.K;
    if (g.flags & BTA_GEMM_B_LOWER) ke = min(ke, (tn + 1) * X3_BN);
    if (g.flags & BTA_GEMM_B_UPPER) kb = max(kb, tn * X3_BN);
    kb = kb / X3_BK * X3_BK;
    ke = (ke + X3_BK - 1) / X3_BK * X3_BK;
    const int row0 = tm * X3_BM, col0 = tn * X3_BN;

    // per-thread A rows for staging: 4 float4 per thread (128 x 32 floats)
    float sa[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int v = tid + i * 256;
      const int r = v / (X3_BK / 4);
      sa[i] = g.a_scale * (g.a_row_scale ? g.a_row_scale[row0 + r] : 1.0f);
    }
    float4v ra[4], rb[2];
    auto gload = [&](int k0) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int v = tid + i * 256;
        const int r = v / (X3_BK / 4), kv = v % (X3_BK / 4);
        ra[i] = *(const float4v*)(g.A + (int64_t)(row0 + r) * g.lda + k0 + kv * 4);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int v = tid + i * 256;
        if (BKN) {
          const int kr = v / (X3_BN / 4), nv = v % (X3_BN / 4);
          rb[i] = *(const float4v*)(g.B + (int64_t)(k0 + kr) * g.ldb + col0 + nv * 4);
        } else {
          const int r = v / (X3_BK / 4), kv = v % (X3_BK / 4);
          rb[i] = *(const float4v*)(g.B + (int64_t)(col0 + r) * g.ldb + k0 + kv * 4);
        }
      }
    };
    auto sstore = [&](_Float16* s) {
      _Float16* ahi = s;
      _Float16* alo = s + X3_A;
      _Float16* bhi = s + 2 * X3_A;
      _Float16* blo = bhi + X3_B;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int v = tid + i * 256;
        const int r = v / (X3_BK / 4), kv = v % (X3_BK / 4);
        half4 h, l;
        split4(ra[i], sa[i], h, l);
        *(half4*)(ahi + r * X3_PITCH + kv * 4) = h;
        *(half4*)(alo + r * X3_PITCH + kv * 4) = l;
      }
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int v = tid + i * 256;
        half4 h, l;
        split4(rb[i], g.b_scale, h, l);
        if (BKN) {  // transpose K x N into the N-major LDS plane
          const int kr = v / (X3_BN / 4), nv = v % (X3_BN / 4);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            bhi[(nv * 4 + e) * X3_PITCH + kr] = h[e];
            blo[(nv * 4 + e) * X3_PITCH + kr] = l[e];
          }
        } else {
          const int r = v / (X3_BK / 4), kv = v % (X3_BK / 4);
          *(half4*)(bhi + r * X3_PITCH + kv * 4) = h;
          *(half4*)(blo + r * X3_PITCH + kv * 4) = l;
        }
      }
    };

    v4f acc[X3_MI][X3_NI];
    double acc64[X3_MI][X3_NI][4];
#pragma unroll
    for (int mi = 0; mi < X3_MI; ++mi)
#pragma unroll
      for (int ni = 0; ni < X3_NI; ++ni)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc64[mi][ni][q] = 0.0;

    const int nk = (ke - kb) / X3_BK;
    if (nk > 0) {
      gload(kb);
      sstore(lds);
      __syncthreads();
    }
    for (int it = 0; it < nk; ++it) {
      if ((it * X3_BK) % X3_SEG == 0) {
#pragma unroll
        for (int mi = 0; mi < X3_MI; ++mi)
#pragma unroll
          for (int ni = 0; ni < X3_NI; ++ni) acc[mi][ni] = (v4f)(0.0f);
      }
      const _Float16* cur = lds + (it & 1) * X3_STAGE;
      _Float16* nxt = lds + ((it + 1) & 1) * X3_STAGE;
      if (it + 1 < nk) gload(kb + (it + 1) * X3_BK);
      half8 ah[X3_MI], al[X3_MI], bh[X3_NI], bl[X3_NI];
#pragma unroll
      for (int mi = 0; mi < X3_MI; ++mi) {
        const int r = wy * (X3_BM / 2) + mi * 16 + l16;
        ah[mi] = *(const half8*)(cur + r * X3_PITCH + gq * 8);
        al[mi] = *(const half8*)(cur + X3_A + r * X3_PITCH + gq * 8);
      }
#pragma unroll
      for (int ni = 0; ni < X3_NI; ++ni) {
        const int r = wx * (X3_BN / 2) + ni * 16 + l16;
        bh[ni] = *(const half8*)(cur + 2 * X3_A + r * X3_PITCH + gq * 8);
        bl[ni] = *(const half8*)(cur + 2 * X3_A + X3_B + r * X3_PITCH + gq * 8);
      }
#pragma unroll
      for (int mi = 0; mi < X3_MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < X3_NI; ++ni) {
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[mi], bh[ni], acc[mi][ni], 0,
                                                               0, 0);
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[mi], bl[ni], acc[mi][ni], 0,
                                                               0, 0);
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[mi], bh[ni], acc[mi][ni], 0,
                                                               0, 0);
        }
      if (it + 1 < nk) sstore(nxt);
      if (((it + 1) * X3_BK) % X3_SEG == 0 || it + 1 == nk) {
#pragma unroll
        for (int mi = 0; mi < X3_MI; ++mi)
#pragma unroll
          for (int ni = 0; ni < X3_NI; ++ni)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc64[mi][ni][q] += (double)acc[mi][ni][q];
      }
      __syncthreads();
    }
#pragma unroll
    for (int mi = 0; mi < X3_MI; ++mi)
#pragma unroll
      for (int ni = 0; ni < X3_NI; ++ni)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int64_t row = row0 + wy * (X3_BM / 2) + mi * 16 + Mfma<float>::row(gq, q);
          const int64_t col = col0 + wx * (X3_BN / 2) + ni * 16 + l16;
          const double srow =
              (double)g.a_scale * (g.a_row_scale ? (double)g.a_row_scale[row] : 1.0);
          g.C[row * g.ldc + col] =
              (float)(acc64[mi][ni][q] * g.alpha / (srow * (double)g.b_scale));
        }
  }
}

// Same snake-schedule cost model as the FP32/FP64 GEMM, for the 128 x 64 tile.
static int x3_workers(const X3Args& g) {
  const int MT = g.M / X3_BM, NT = g.N / X3_BN, total = MT * NT;
  const int ncu = num_cus();
  std::vector<double> cu(ncu);
  double best = 1e300;
  int bestW = 1;
  const int cands[] = {2 * ncu, ncu, 3 * ncu / 4, ncu / 2};
  for (int c : cands) {
    const int W = std::max(1, std::min(c, total));
    std::fill(cu.begin(), cu.end(), 0.0);
    for (int w = 0; w < W; ++w) {
      double load = 0.0;
      for (int j = 0;; ++j) {
        const int t = (j & 1) ? (j + 1) * W - 1 - w : j * W + w;
        if (t >= total) break;
        int tm, tn;
        decode_tile(t, g.flags, MT, NT, tm, tn);
        int kb = 0, ke = g.K;
        if (g.flags & BTA_GEMM_B_LOWER) ke = std::min(ke, (tn + 1) * X3_BN);
        if (g.flags & BTA_GEMM_B_UPPER) kb = std::max(kb, tn * X3_BN);
        load += (ke > kb ? ke - kb : 0) + 2 * X3_BK;
      }
      cu[w % ncu] += load;
    }
    const double t = *std::max_element(cu.begin(), cu.end());
    if (t < best) {
      best = t;
      bestW = W;
    }
  }
  return bestW;
}

int gemm_x3(int flags, int64_t M, int64_t N, int64_t K, double alpha, const float* A,
            int64_t lda, const float* a_row_scale, float a_scale, const float* B, int64_t ldb,
            float b_scale, float* C, int64_t ldc, hipStream_t s) {
  if (M % X3_BM || N % X3_BN || K % X3_BK || M <= 0 || N <= 0 || K < 0 ||
      (flags & (BTA_GEMM_A_LOWER | BTA_GEMM_A_UPPER | BTA_GEMM_C_LOWER | BTA_GEMM_A_KM))) {
    set_error("gemm_x3: M%%128, N%%64, K%%32 must be 0 and A untransposed, full output");
    return BTA_ERR_INVALID;
  }
  X3Args g;
  g.A = A;
  g.B = B;
  g.C = C;
  g.a_row_scale = a_row_scale;
  g.a_scale = a_scale;
  g.b_scale = b_scale;
  g.lda = lda;
  g.ldb = ldb;
  g.ldc = ldc;
  g.M = (int)M;
  g.N = (int)N;
  g.K = (int)K;
  g.flags = flags;
  g.alpha = alpha;
  const int W = x3_workers(g);
  if (flags & BTA_GEMM_B_KN)
    hipLaunchKernelGGL(gemm_x3_kernel<true>, dim3(W), dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL(gemm_x3_kernel<false>, dim3(W), dim3(256), 0, s, g);
  BTA_LAUNCH_CHECK("gemm_x3_kernel");
  return BTA_OK;
}

}  // namespace bta
