// Split-FP16 GEMM on pre-split operands ("h2"):
//   C = alpha * (A_hi + A_lo)(B_hi + B_lo)^T / (a_scale * a_row_scale[r] * b_scale)
// with the a_lo b_lo term dropped (2^-22 relative).  A (M x K) and B (N x K) are
// two FP16 planes each, K-contiguous, written by their producers (the
// covariance kernel, this GEMM's own split epilogue, the factor set-up), so the
// main loop is pure data movement plus three v_mfma_f32_16x16x32_f16 per
// fragment pair.  FP32 accumulators are flushed into FP64 sums every 256 K.
//
// Tile 128 x 128 x 32, 256 threads = 4 waves of 64 x 64, two LDS stages of the
// four planes (80 KiB, two workgroups per CU), one barrier per K chunk, the
// next chunk's global loads in flight during the current chunk's MFMAs.  The
// epilogue writes C in FP32 and, optionally, C's own hi/lo planes (scaled by
// c_scale) for a following split GEMM.  Triangular K ranges and the worker
// schedule follow gemm_kernel.
#include "bta_common.h"
#include "gemm.h"
#include "gemm_tile.h"
#include "kernels.h"

#include <algorithm>
#include <vector>

namespace bta {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef int int4v __attribute__((ext_vector_type(4)));

constexpr int H2_BM = 128, H2_BN = 128, H2_BK = 64;
constexpr int H2_PITCH = H2_BK + 8;  // halfs per LDS row: 144 bytes, conflict-free b128 reads
constexpr int H2_PLANE = H2_BM * H2_PITCH;
constexpr int H2_STAGE = 4 * H2_PLANE;  // A_hi, A_lo, B_hi, B_lo
constexpr int H2_SEG = 256;             // K per FP64 flush
constexpr int H2_F = 4;                 // 16 x 16 fragments per wave side (64 / 16)
constexpr int H2_ROWS = 256 / (H2_BK / 8);  // plane rows staged per pass (32)
constexpr int H2_PASS = H2_BM / H2_ROWS;    // passes per plane (4)

// One workgroup per CU (one wave per SIMD): the FP32 + FP64 accumulators of a
// 64 x 64 wave tile plus fragments and staging need ~350 registers.
__global__ __launch_bounds__(256, 1) void gemm_h2_kernel(H2Args g) {
  __shared__ __attribute__((aligned(16))) _Float16 lds[2 * H2_STAGE];
  const int MT = g.M / H2_BM, NT = g.N / H2_BN;
  const int total = MT * NT;
  const int W = gridDim.x, w = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wy = wv >> 1, wx = wv & 1, l16 = lane & 15, gq = lane >> 4;
  // staging: thread -> (row, 16-byte column) of each plane, H2_PASS rows per plane
  const int sr = tid / (H2_BK / 8), sk = (tid % (H2_BK / 8)) * 8;

  for (int jj = 0;; ++jj) {
    const int t = (jj & 1) ? (jj + 1) * W - 1 - w : jj * W + w;
    if (t >= total) break;
    int tm, tn;
    decode_tile(t, g.flags, MT, NT, tm, tn);
    int kb = 0, ke = g.K;
    if (g.flags & BTA_GEMM_B_LOWER) ke = min(ke, (tn + 1) * H2_BN);
    if (g.flags & BTA_GEMM_B_UPPER) kb = max(kb, tn * H2_BN);
    kb = kb / H2_BK * H2_BK;
    ke = (ke + H2_BK - 1) / H2_BK * H2_BK;
    const int row0 = tm * H2_BM, col0 = tn * H2_BN;
    const _Float16* src[4] = {g.Ah + (int64_t)(row0 + sr) * g.lda + sk,
                              g.Al + (int64_t)(row0 + sr) * g.lda + sk,
                              g.Bh + (int64_t)(col0 + sr) * g.ldb + sk,
                              g.Bl + (int64_t)(col0 + sr) * g.ldb + sk};
    const int64_t step[4] = {H2_ROWS * g.lda, H2_ROWS * g.lda, H2_ROWS * g.ldb,
                             H2_ROWS * g.ldb};
    int4v st[4 * H2_PASS];
    auto gload = [&](int k0) {
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int i = 0; i < H2_PASS; ++i)
          st[p * H2_PASS + i] = *(const int4v*)(src[p] + i * step[p] + k0);
    };
    auto sstore = [&](_Float16* s) {
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int i = 0; i < H2_PASS; ++i)
          *(int4v*)(s + p * H2_PLANE + (sr + i * H2_ROWS) * H2_PITCH + sk) = st[p * H2_PASS + i];
    };

    v4f acc[H2_F][H2_F];
    double acc64[H2_F][H2_F][4];
#pragma unroll
    for (int mi = 0; mi < H2_F; ++mi)
#pragma unroll
      for (int ni = 0; ni < H2_F; ++ni)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc64[mi][ni][q] = 0.0;

    const int nk = (ke - kb) / H2_BK;
    if (nk > 0) {
      gload(kb);
      sstore(lds);
      __syncthreads();
    }
    for (int it = 0; it < nk; ++it) {
      if ((it * H2_BK) % H2_SEG == 0) {
#pragma unroll
        for (int mi = 0; mi < H2_F; ++mi)
#pragma unroll
          for (int ni = 0; ni < H2_F; ++ni) acc[mi][ni] = (v4f)(0.0f);
      }
      const _Float16* cur = lds + (it & 1) * H2_STAGE;
      _Float16* nxt = lds + ((it + 1) & 1) * H2_STAGE;
      if (it + 1 < nk) gload(kb + (it + 1) * H2_BK);
#pragma unroll
      for (int ks = 0; ks < H2_BK / 32; ++ks) {
        half8 ah[H2_F], al[H2_F], bh[H2_F], bl[H2_F];
        const int ko = ks * 32 + gq * 8;
#pragma unroll
        for (int f = 0; f < H2_F; ++f) {
          const int ra = wy * 64 + f * 16 + l16, rb = wx * 64 + f * 16 + l16;
          ah[f] = *(const half8*)(cur + ra * H2_PITCH + ko);
          al[f] = *(const half8*)(cur + H2_PLANE + ra * H2_PITCH + ko);
          bh[f] = *(const half8*)(cur + 2 * H2_PLANE + rb * H2_PITCH + ko);
          bl[f] = *(const half8*)(cur + 3 * H2_PLANE + rb * H2_PITCH + ko);
        }
        // three passes over the 16 fragments: no back-to-back dependent MFMAs
#pragma unroll
        for (int mi = 0; mi < H2_F; ++mi)
#pragma unroll
          for (int ni = 0; ni < H2_F; ++ni)
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[mi], bh[ni], acc[mi][ni], 0,
                                                                 0, 0);
#pragma unroll
        for (int mi = 0; mi < H2_F; ++mi)
#pragma unroll
          for (int ni = 0; ni < H2_F; ++ni)
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[mi], bl[ni], acc[mi][ni], 0,
                                                                 0, 0);
#pragma unroll
        for (int mi = 0; mi < H2_F; ++mi)
#pragma unroll
          for (int ni = 0; ni < H2_F; ++ni)
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[mi], bh[ni], acc[mi][ni], 0,
                                                                 0, 0);
      }
      if (it + 1 < nk) sstore(nxt);
      if (((it + 1) * H2_BK) % H2_SEG == 0 || it + 1 == nk) {
#pragma unroll
        for (int mi = 0; mi < H2_F; ++mi)
#pragma unroll
          for (int ni = 0; ni < H2_F; ++ni)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc64[mi][ni][q] += (double)acc[mi][ni][q];
      }
      __syncthreads();
    }
    const double bs = (double)g.b_scale;
#pragma unroll
    for (int mi = 0; mi < H2_F; ++mi)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t row = row0 + wy * 64 + mi * 16 + Mfma<float>::row(gq, q);
        const double f =
            g.alpha / ((double)g.a_scale * (g.a_row_scale ? (double)g.a_row_scale[row] : 1.0) * bs);
#pragma unroll
        for (int ni = 0; ni < H2_F; ++ni) {
          const int64_t col = col0 + wx * 64 + ni * 16 + l16;
          const float v = (float)(acc64[mi][ni][q] * f);
          g.C[row * g.ldc + col] = v;
          if (g.Ch) {
            const float x = v * g.c_scale;
            const _Float16 h = (_Float16)x;
            g.Ch[row * g.ldc + col] = h;
            g.Cl[row * g.ldc + col] = (_Float16)(x - (float)h);
          }
        }
      }
  }
}

// Snake-schedule cost model (as gemm_kernel's) for the 128 x 128 tile.
static int h2_workers(const H2Args& g) {
  const int MT = g.M / H2_BM, NT = g.N / H2_BN, total = MT * NT;
  const int ncu = num_cus();
  std::vector<double> cu(ncu);
  double best = 1e300;
  int bestW = 1;
  const int cands[] = {2 * ncu, 3 * ncu / 2, ncu, 3 * ncu / 4, ncu / 2};
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
        if (g.flags & BTA_GEMM_B_LOWER) ke = std::min(ke, (tn + 1) * H2_BN);
        if (g.flags & BTA_GEMM_B_UPPER) kb = std::max(kb, tn * H2_BN);
        load += (ke > kb ? ke - kb : 0) + 4 * H2_BK;
      }
      // two workgroups share a CU: a CU's time is the larger of its two
      // workers' loads, or their sum at half rate when both are busy
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

int gemm_h2(const H2Args& a, hipStream_t s) {
  const uintptr_t al16 = (uintptr_t)a.Ah | (uintptr_t)a.Al | (uintptr_t)a.Bh | (uintptr_t)a.Bl;
  if (a.M % H2_BM || a.N % H2_BN || a.K % H2_BK || a.M <= 0 || a.N <= 0 || a.K < 0 ||
      a.lda % 8 || a.ldb % 8 || a.lda < a.K || a.ldb < a.K || a.ldc < a.N || !a.Ah || !a.Al ||
      !a.Bh || !a.Bl || !a.C || (a.Ch && !a.Cl) || (al16 & 15) ||
      (a.flags & ~(BTA_GEMM_B_LOWER | BTA_GEMM_B_UPPER))) {
    set_error("gemm_h2: M%%128, N%%128, K%%64, lda/ldb%%8 must be 0, operand planes 16-byte "
              "aligned; flags B_LOWER/B_UPPER only");
    return BTA_ERR_INVALID;
  }
  if (a.K == 0) {
    set_error("gemm_h2: K must be positive");
    return BTA_ERR_INVALID;
  }
  const int W = h2_workers(a);
  hipLaunchKernelGGL(gemm_h2_kernel, dim3(W), dim3(256), 0, s, a);
  BTA_LAUNCH_CHECK("gemm_h2_kernel");
  return BTA_OK;
}

// hi/lo FP16 planes of scale * X (X FP32 or FP64, rows x cols, row-major), or of
// scale * X^T when transpose is set (output cols x rows).  Zero-fills the padded
// region [rows, rows_pad) x [cols, cols_pad) of the (untransposed) index space.
template <typename T>
__global__ void split_f16_kernel(const T* __restrict__ X, int64_t ldx, int64_t rows,
                                 int64_t cols, int64_t rows_pad, int64_t cols_pad, float scale,
                                 int transpose, _Float16* __restrict__ hi,
                                 _Float16* __restrict__ lo, int64_t ldo) {
  __shared__ float tile[32][33];
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int i = ty; i < 32; i += 8) {
    const int64_t r = r0 + i, c = c0 + tx;
    float v = 0.0f;
    if (r < rows && c < cols) v = (float)((double)X[r * ldx + c] * (double)scale);
    tile[i][tx] = v;
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    int64_t orow, ocol;
    float x;
    if (transpose) {
      orow = c0 + i;
      ocol = r0 + tx;
      x = tile[tx][i];
      if (orow >= cols_pad || ocol >= rows_pad) continue;
    } else {
      orow = r0 + i;
      ocol = c0 + tx;
      x = tile[i][tx];
      if (orow >= rows_pad || ocol >= cols_pad) continue;
    }
    const _Float16 h = (_Float16)x;
    hi[orow * ldo + ocol] = h;
    lo[orow * ldo + ocol] = (_Float16)(x - (float)h);
  }
}

int split_f16(bool f64, const void* X, int64_t ldx, int64_t rows, int64_t cols,
              int64_t rows_pad, int64_t cols_pad, float scale, bool transpose, _Float16* hi,
              _Float16* lo, int64_t ldo, hipStream_t s) {
  if (rows < 0 || cols < 0 || rows_pad < rows || cols_pad < cols || !X || !hi || !lo ||
      ldo < (transpose ? rows_pad : cols_pad)) {
    set_error("split_f16: invalid sizes");
    return BTA_ERR_INVALID;
  }
  if (rows_pad == 0 || cols_pad == 0) return BTA_OK;
  dim3 grid((unsigned)((cols_pad + 31) / 32), (unsigned)((rows_pad + 31) / 32));
  if (f64)
    hipLaunchKernelGGL(split_f16_kernel<double>, grid, dim3(256), 0, s, (const double*)X, ldx,
                       rows, cols, rows_pad, cols_pad, scale, (int)transpose, hi, lo, ldo);
  else
    hipLaunchKernelGGL(split_f16_kernel<float>, grid, dim3(256), 0, s, (const float*)X, ldx,
                       rows, cols, rows_pad, cols_pad, scale, (int)transpose, hi, lo, ldo);
  BTA_LAUNCH_CHECK("split_f16_kernel");
  return BTA_OK;
}

}  // namespace bta

extern "C" int bta_split_f16(int32_t src_f64, const void* X, int64_t ldx, int64_t rows,
                             int64_t cols, int64_t rows_pad, int64_t cols_pad, float scale,
                             int32_t transpose, void* hi, void* lo, int64_t ldo,
                             bta_stream_t stream) {
  return bta::split_f16(src_f64 != 0, X, ldx, rows, cols, rows_pad, cols_pad, scale,
                        transpose != 0, (_Float16*)hi, (_Float16*)lo, ldo,
                        (hipStream_t)stream);
}

extern "C" int bta_gemm_f16x3(int32_t flags, int64_t M, int64_t N, int64_t K, double alpha,
                              const void* a_hi, const void* a_lo, int64_t lda,
                              const float* a_row_scale, float a_scale, const void* b_hi,
                              const void* b_lo, int64_t ldb, float b_scale, float* C,
                              int64_t ldc, void* c_hi, void* c_lo, float c_scale,
                              bta_stream_t stream) {
  if (M > (1 << 30) || N > (1 << 30) || K > (1 << 30)) {
    bta::set_error("bta_gemm_f16x3: sizes beyond 2^30");
    return BTA_ERR_INVALID;
  }
  bta::H2Args a;
  a.Ah = (const _Float16*)a_hi;
  a.Al = (const _Float16*)a_lo;
  a.Bh = (const _Float16*)b_hi;
  a.Bl = (const _Float16*)b_lo;
  a.C = C;
  a.Ch = (_Float16*)c_hi;
  a.Cl = (_Float16*)c_lo;
  a.a_row_scale = a_row_scale;
  a.a_scale = a_scale;
  a.b_scale = b_scale;
  a.c_scale = c_scale;
  a.lda = lda;
  a.ldb = ldb;
  a.ldc = ldc;
  a.M = (int)M;
  a.N = (int)N;
  a.K = (int)K;
  a.flags = flags;
  a.alpha = alpha;
  return bta::gemm_h2(a, (hipStream_t)stream);
}
