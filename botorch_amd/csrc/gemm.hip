// MFMA GEMM for gfx950: C = alpha * A * op(B) + beta * C0.
//
// One kernel template serves every dense product on the path (posterior
// contraction with L^-1, its adjoint, Cholesky panel/trailing updates,
// triangular inversion, the MLL's K^-1 tiles):
//   * 256 threads = 4 waves in a 2 x 2 layout; each wave owns a (BM/2) x (BN/2)
//     sub-tile made of 16 x 16 MFMA blocks (v_mfma_f64_16x16x4f64 or
//     v_mfma_f32_16x16x4f32, accumulation in the instruction's own precision).
//   * A and B chunks of BK columns are staged global -> registers -> LDS with a
//     two-stage LDS ring, so the next chunk's global loads overlap the MFMAs.
//   * Inside a chunk, lane group g = lane/16 consumes the contiguous K-slice
//     [g*BK/4, (g+1)*BK/4); the K order of a dot product does not change its
//     value, and each lane then reads its fragment as one contiguous run.
//   * Triangular operands skip their zero K-range per tile, and tiles are
//     enumerated longest-first and dealt to a fixed grid of workers in a
//     snake order, which balances the uneven tile lengths of a triangular
//     product without atomics or a host-side schedule.
#include "bta_common.h"
#include "gemm.h"
#include "gemm_d.h"
#include "gemm_tile.h"
#include "kernels.h"

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

namespace bta {

// ACC64 (float only): each K-segment of ACC_SEG columns is accumulated by the
// FP32 MFMA and then added into FP64 accumulators, so the rounding error of a
// dot product grows with the segment length instead of the full K.
constexpr int ACC_SEG = 256;

template <typename T, int BM, int BN, int BK, bool BKN, bool AKM, bool ACC64>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  typedef Cfg<T, BM, BN, BK, BKN, AKM> C;
  __shared__ __attribute__((aligned(16))) T lds[2 * C::STAGE];
  const int64_t batch = blockIdx.y;
  const T* A = (const T*)g.A + batch * g.sA;
  const T* B = (const T*)g.B + batch * g.sB;
  T* Cp = (T*)g.C + batch * g.sC;
  const T* C0 = g.C0 ? (const T*)g.C0 + batch * g.sC0 : nullptr;
  const int MT = g.M / BM, NT = g.N / BN;
  const int total = (g.flags & BTA_GEMM_C_LOWER) ? MT * (MT + 1) / 2 : MT * NT;
  const int W = gridDim.x, w = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wy = wv >> 1, wx = wv & 1, l16 = lane & 15, gq = lane >> 4;

  for (int j = 0;; ++j) {
    const int t = (j & 1) ? (j + 1) * W - 1 - w : j * W + w;
    if (t >= total) break;
    int tm, tn;
    decode_tile(t, g.flags, MT, NT, tm, tn);
    int kb = 0, ke = g.K;
    if (g.flags & BTA_GEMM_A_LOWER) ke = min(ke, (tm + 1) * BM);
    if (g.flags & BTA_GEMM_A_UPPER) kb = max(kb, tm * BM);
    if (g.flags & BTA_GEMM_B_LOWER) ke = min(ke, (tn + 1) * BN);
    if (g.flags & BTA_GEMM_B_UPPER) kb = max(kb, tn * BN);
    kb = kb / BK * BK;
    ke = (ke + BK - 1) / BK * BK;
    typename Mfma<T>::acc_t acc[C::MI][C::NI];
    double acc64[ACC64 ? C::MI : 1][ACC64 ? C::NI : 1][4];
    if (ACC64) {
#pragma unroll
      for (int mi = 0; mi < (ACC64 ? C::MI : 1); ++mi)
#pragma unroll
        for (int ni = 0; ni < (ACC64 ? C::NI : 1); ++ni)
#pragma unroll
          for (int q = 0; q < 4; ++q) acc64[mi][ni][q] = 0.0;
      for (int k0 = kb; k0 < ke; k0 += ACC_SEG) {
        const int k1 = min(ke, k0 + ACC_SEG);
        gemm_tile<T, BM, BN, BK, BKN, AKM>(A, g.lda, B, g.ldb, tm * BM, tn * BN, k0, k1, lds,
                                           acc);
#pragma unroll
        for (int mi = 0; mi < (ACC64 ? C::MI : 1); ++mi)
#pragma unroll
          for (int ni = 0; ni < (ACC64 ? C::NI : 1); ++ni)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc64[mi][ni][q] += (double)acc[mi][ni][q];
      }
    } else {
      gemm_tile<T, BM, BN, BK, BKN, AKM>(A, g.lda, B, g.ldb, tm * BM, tn * BN, kb, ke, lds, acc);
    }
#pragma unroll
    for (int mi = 0; mi < C::MI; ++mi)
#pragma unroll
      for (int ni = 0; ni < C::NI; ++ni)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int64_t row = tm * BM + wy * (BM / 2) + mi * 16 + Mfma<T>::row(gq, q);
          const int64_t col = tn * BN + wx * (BN / 2) + ni * 16 + l16;
          double v = ACC64 ? acc64[ACC64 ? mi : 0][ACC64 ? ni : 0][q] : (double)acc[mi][ni][q];
          v *= g.alpha;
          if (C0) v += g.beta * (double)C0[row * g.ldc0 + col];
          Cp[row * g.ldc + col] = (T)v;
        }
  }
}

// Estimated makespan (in MAC units, over CUs) of the snake schedule for a
// tile shape and worker count.  Tiles are dealt to workers as in gemm_kernel;
// workgroup i of the grid lands on CU i % n_cus.
template <int BM, int BN, int BK>
static double schedule_cost(const GemmArgs& g, int batch, int W, std::vector<double>& cu) {
  const int MT = g.M / BM, NT = g.N / BN;
  const int total = (g.flags & BTA_GEMM_C_LOWER) ? MT * (MT + 1) / 2 : MT * NT;
  const int ncu = (int)cu.size();
  std::fill(cu.begin(), cu.end(), 0.0);
  for (int w = 0; w < W; ++w) {
    double load = 0.0;
    for (int j = 0;; ++j) {
      const int t = (j & 1) ? (j + 1) * W - 1 - w : j * W + w;
      if (t >= total) break;
      int tm, tn;
      decode_tile(t, g.flags, MT, NT, tm, tn);
      int kb = 0, ke = g.K;
      if (g.flags & BTA_GEMM_A_LOWER) ke = std::min(ke, (tm + 1) * BM);
      if (g.flags & BTA_GEMM_A_UPPER) kb = std::max(kb, tm * BM);
      if (g.flags & BTA_GEMM_B_LOWER) ke = std::min(ke, (tn + 1) * BN);
      if (g.flags & BTA_GEMM_B_UPPER) kb = std::max(kb, tn * BN);
      // K chunks plus a fixed per-tile cost (prologue, epilogue) of ~2 chunks
      load += (double)BM * BN * ((ke > kb ? ke - kb : 0) + 2 * BK);
    }
    for (int b = 0; b < batch; ++b) cu[(b * W + w) % ncu] += load;
  }
  return *std::max_element(cu.begin(), cu.end());
}

struct GemmPlan {
  int big;  // 1: 128 x 128 tiles
  int W;
};

template <int BM, int BN, int BK>
static void best_workers(const GemmArgs& g, int batch, double eff, double& best, int& bestW) {
  const int MT = g.M / BM, NT = g.N / BN;
  const int total = (g.flags & BTA_GEMM_C_LOWER) ? MT * (MT + 1) / 2 : MT * NT;
  const int ncu = num_cus();
  std::vector<double> cu(ncu);
  const int cands[] = {2 * ncu, ncu, 3 * ncu / 4, ncu / 2, ncu / 4, ncu / 8};
  for (int c : cands) {
    int W = batch > 1 ? (c + batch - 1) / batch : c;
    W = std::max(1, std::min(W, total));
    const double t = schedule_cost<BM, BN, BK>(g, batch, W, cu) / eff;
    if (t < best) {
      best = t;
      bestW = W;
    }
  }
}

// Picks the tile shape and worker count from the cost model above; plans are
// cached per problem signature.
static GemmPlan plan_gemm(const GemmArgs& g, int batch) {
  // Created once and never destroyed: no destructor of this library runs at
  // process exit, when the HIP runtime or a profiler may already be torn down.
  static std::mutex& mu = *new std::mutex;
  static auto& cache = *new std::map<std::array<int64_t, 6>, GemmPlan>;
  const bool f32 = g.flags & BTA_GEMM_F32;
  std::array<int64_t, 6> key = {g.M, g.N, g.K, g.flags, batch, num_cus()};
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  double best = 1e300;
  int bestW = 1, big = 0;
  // 64x64 tiles sustain a lower MFMA rate (half the operand reuse per byte).
  if (f32)
    best_workers<64, 64, 16>(g, batch, 0.7, best, bestW);
  else
    best_workers<64, 64, 16>(g, batch, 0.8, best, bestW);
  if (g.M % 128 == 0 && g.N % 128 == 0) {
    double b2 = best;
    int w2 = bestW;
    if (f32)
      best_workers<128, 128, 16>(g, batch, 1.0, b2, w2);
    else
      best_workers<128, 128, 8>(g, batch, 1.0, b2, w2);
    if (b2 < best) {
      best = b2;
      bestW = w2;
      big = 1;
    }
  }
  GemmPlan p{big, bestW};
  std::lock_guard<std::mutex> lk(mu);
  cache[key] = p;
  return p;
}

template <typename T, int BM, int BN, int BK, bool ACC64>
static int launch_cfg2(const GemmArgs& g, int batch, int W, hipStream_t s) {
  dim3 grid(W, batch);
  const bool bkn = g.flags & BTA_GEMM_B_KN, akm = g.flags & BTA_GEMM_A_KM;
  if (bkn && akm)
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, true, true, ACC64>), grid, dim3(256), 0, s, g);
  else if (bkn)
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, true, false, ACC64>), grid, dim3(256), 0, s,
                       g);
  else if (akm)
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, false, true, ACC64>), grid, dim3(256), 0, s,
                       g);
  else
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, false, false, ACC64>), grid, dim3(256), 0, s,
                       g);
  BTA_LAUNCH_CHECK("gemm_kernel");
  return BTA_OK;
}

template <typename T, int BM, int BN, int BK>
static int launch_cfg(const GemmArgs& g, int batch, int W, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (g.flags & BTA_GEMM_ACC64) return launch_cfg2<T, BM, BN, BK, true>(g, batch, W, s);
  }
  return launch_cfg2<T, BM, BN, BK, false>(g, batch, W, s);
}

int launch_gemm(const GemmArgs& g, int batch, hipStream_t s) {
  if (g.M % 64 || g.N % 64 || g.K % 64 || g.M <= 0 || g.N <= 0 || g.K < 0) {
    set_error("bta_gemm: M, N, K must be positive multiples of 64 (got %d %d %d)", g.M, g.N,
              g.K);
    return BTA_ERR_INVALID;
  }
  if ((g.flags & BTA_GEMM_C_LOWER) && g.M != g.N) {
    set_error("bta_gemm: C_LOWER needs M == N");
    return BTA_ERR_INVALID;
  }
  if (batch == 0) return BTA_OK;
  if (!(g.flags & BTA_GEMM_F32) && gemm_d_supported(g)) {
    const int rc = gemm_d(g, batch, s);
    if (rc >= 0) return rc;
  }
  const GemmPlan p = plan_gemm(g, batch);
  if (g.flags & BTA_GEMM_F32) {
    if (p.big) return launch_cfg<float, 128, 128, 16>(g, batch, p.W, s);
    return launch_cfg<float, 64, 64, 16>(g, batch, p.W, s);
  }
  if (p.big) return launch_cfg<double, 128, 128, 8>(g, batch, p.W, s);
  return launch_cfg<double, 64, 64, 16>(g, batch, p.W, s);
}

}  // namespace bta

extern "C" int bta_gemm(int32_t flags, int64_t M, int64_t N, int64_t K, double alpha,
                        const void* A, int64_t lda, const void* B, int64_t ldb, double beta,
                        const void* C0, int64_t ldc0, void* C, int64_t ldc, int64_t batch,
                        int64_t sA, int64_t sB, int64_t sC0, int64_t sC,
                        bta_stream_t stream) {
  bta::GemmArgs g;
  g.A = A;
  g.B = B;
  g.C0 = C0;
  g.C = C;
  g.lda = lda;
  g.ldb = ldb;
  g.ldc0 = ldc0;
  g.ldc = ldc;
  g.sA = sA;
  g.sB = sB;
  g.sC0 = sC0;
  g.sC = sC;
  g.M = (int)M;
  g.N = (int)N;
  g.K = (int)K;
  g.flags = flags;
  g.alpha = alpha;
  g.beta = beta;
  if (!A || !B || !C || M > (1 << 30) || N > (1 << 30) || K > (1 << 30) || batch < 0 ||
      batch > 65535) {
    bta::set_error("bta_gemm: invalid pointers or sizes");
    return BTA_ERR_INVALID;
  }
  const int32_t known = BTA_GEMM_A_LOWER | BTA_GEMM_A_UPPER | BTA_GEMM_B_LOWER |
                        BTA_GEMM_B_UPPER | BTA_GEMM_C_LOWER | BTA_GEMM_B_KN | BTA_GEMM_F32 |
                        BTA_GEMM_A_KM | BTA_GEMM_ACC64;
  if (flags & ~known) {
    bta::set_error("bta_gemm: unknown flag bits 0x%x", flags & ~known);
    return BTA_ERR_INVALID;
  }
  return bta::launch_gemm(g, (int)batch, (hipStream_t)stream);
}
