// FP64 MFMA GEMM with direct global -> LDS loads ("d" path), used for every
// FP64 product whose operands are 16-byte aligned: the Cholesky trailing
// updates, the triangular inversion and the MLL's K^-1 tiles.
//
//   C = alpha * op(A) op(B) + beta * C0,  op(A) M x K, op(B) K x N
//
// Tile 128 x BN (BN = 128 or 64), 256 threads = 4 waves in a 2 x 2 layout,
// each wave 64 x BN/2 of 16 x 16 v_mfma_f64_16x16x4f64 blocks.  K moves in
// chunks of 8 through a ring of 4 LDS stages (3 chunks in flight) filled by
// global_load_lds_dwordx4 (no registers, no ds_write), one barrier per chunk.
// An operand is either K-contiguous (A M x K / B N x K rows; 64-byte rows per
// stage with XOR-swizzled 16-byte slots, read as one b128 per fragment) or
// K-major (A K x M / B K x N; stage rows are k, read as two b64 per
// fragment).  Lane group g of a fragment holds k = 2g and 2g + 1 of the chunk;
// both operands use the same assignment, so each MFMA pair sums the chunk.
// M and N need only be multiples of 64: a tile's rows past the edge load
// in-bounds rows 64 earlier and are not stored.  Triangular flags trim each
// tile's K range; tiles are enumerated longest-first and dealt to workers so
// that the two workers sharing a CU hold a heavy and a light tile.
#include "bta_common.h"
#include "gemm.h"
#include "gemm_d.h"
#include "gemm_tile.h"
#include "kernels.h"

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <vector>

namespace bta {

template <int BN, bool AKM, bool BKN>
__global__ __launch_bounds__(256, 2) void gemm_d_kernel(GemmArgs g) {
  using C = DCfg<BN>;
  __shared__ __attribute__((aligned(1024))) unsigned char lds[C::STAGES * C::STAGE];
  const int64_t batch = blockIdx.y;
  const double* A = (const double*)g.A + batch * g.sA;
  const double* B = (const double*)g.B + batch * g.sB;
  double* Cp = (double*)g.C + batch * g.sC;
  const double* C0 = g.C0 ? (const double*)g.C0 + batch * g.sC0 : nullptr;
  const int MT = (g.M + D_BM - 1) / D_BM, NT = (g.N + BN - 1) / BN;
  const int total = (g.flags & BTA_GEMM_C_LOWER) ? MT * (MT + 1) / 2 : MT * NT;
  const int W = gridDim.x, w = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wy = wv >> 1, wx = wv & 1, l16 = lane & 15, gq = lane >> 4;

  for (int jj = 0; jj * W < total; ++jj) {
    const int t = h2_tile_of(jj, w, W);
    if (t >= total) continue;
    int tm, tn;
    decode_tile(t, g.flags, MT, NT, tm, tn);
    int kb = 0, ke = g.K;
    if (g.flags & BTA_GEMM_A_LOWER) ke = min(ke, (tm + 1) * D_BM);
    if (g.flags & BTA_GEMM_A_UPPER) kb = max(kb, tm * D_BM);
    if (g.flags & BTA_GEMM_B_LOWER) ke = min(ke, (tn + 1) * BN);
    if (g.flags & BTA_GEMM_B_UPPER) kb = max(kb, tn * BN);
    kb = kb / D_BK * D_BK;
    ke = (ke + D_BK - 1) / D_BK * D_BK;
    v4d acc[C::FM][C::FN];
    d_mainloop<BN, AKM, BKN>(A, g.lda, B, g.ldb, tm * D_BM, tn * BN, g.M, g.N, kb, ke, lds,
                             acc);
    // C0 may alias C: per 16-row block, its C0 loads go before its stores, so
    // they overlap instead of each waiting behind the previous store
#pragma unroll
    for (int mi = 0; mi < C::FM; ++mi) {
      double c0v[C::FN][4];
      if (C0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int64_t row = tm * D_BM + wy * 64 + mi * 16 + Mfma<double>::row(gq, q);
#pragma unroll
          for (int ni = 0; ni < C::FN; ++ni) {
            const int64_t col = tn * BN + wx * C::WN + ni * 16 + l16;
            c0v[ni][q] = (row < g.M && col < g.N) ? C0[row * g.ldc0 + col] : 0.0;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t row = tm * D_BM + wy * 64 + mi * 16 + Mfma<double>::row(gq, q);
        if (row >= g.M) continue;
#pragma unroll
        for (int ni = 0; ni < C::FN; ++ni) {
          const int64_t col = tn * BN + wx * C::WN + ni * 16 + l16;
          if (col >= g.N) continue;
          double v = g.alpha * acc[mi][ni][q];
          if (C0) v += g.beta * c0v[ni][q];
          Cp[row * g.ldc + col] = v;
        }
      }
    }
  }
}

// Cost model of the worker schedule (chunks per worker plus a per-tile fixed
// cost); a CU's time is the sum over the workers it hosts, workgroup i of the
// launch on CU i % ncu.
static double d_cost(const GemmArgs& g, int batch, int BN, int W, std::vector<double>& cu) {
  const int MT = (g.M + D_BM - 1) / D_BM, NT = (g.N + BN - 1) / BN;
  const int total = (g.flags & BTA_GEMM_C_LOWER) ? MT * (MT + 1) / 2 : MT * NT;
  const int ncu = (int)cu.size();
  const double per_chunk = BN == 128 ? 1.0 : 0.55;
  std::fill(cu.begin(), cu.end(), 0.0);
  for (int w = 0; w < W; ++w) {
    double load = 0.0;
    for (int j = 0; j * W < total; ++j) {
      const int t = h2_tile_of(j, w, W);
      if (t >= total) continue;
      int tm, tn;
      decode_tile(t, g.flags, MT, NT, tm, tn);
      int kb = 0, ke = g.K;
      if (g.flags & BTA_GEMM_A_LOWER) ke = std::min(ke, (tm + 1) * D_BM);
      if (g.flags & BTA_GEMM_A_UPPER) kb = std::max(kb, tm * D_BM);
      if (g.flags & BTA_GEMM_B_LOWER) ke = std::min(ke, (tn + 1) * BN);
      if (g.flags & BTA_GEMM_B_UPPER) kb = std::max(kb, tn * BN);
      load += per_chunk * (ke > kb ? (ke - kb) / D_BK : 0) + 6.0;
    }
    for (int b = 0; b < batch; ++b) cu[((int64_t)b * W + w) % ncu] += load;
  }
  return *std::max_element(cu.begin(), cu.end());
}

struct DPlan {
  int bn, workers;
};

static DPlan d_plan(const GemmArgs& g, int batch) {
  static std::mutex mu;
  static std::map<std::array<int64_t, 6>, DPlan> cache;
  const int ncu = num_cus();
  const std::array<int64_t, 6> key = {g.M, g.N, g.K, g.flags, batch, ncu};
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  std::vector<double> cu(ncu);
  DPlan best{128, 1};
  double bt = 1e300;
  for (int bn : {128, 64}) {
    if (bn == 64 && (g.flags & BTA_GEMM_C_LOWER)) continue;  // square tiles only
    const int MT = (g.M + D_BM - 1) / D_BM, NT = (g.N + bn - 1) / bn;
    const int total = (g.flags & BTA_GEMM_C_LOWER) ? MT * (MT + 1) / 2 : MT * NT;
    const int slots = 2 * ncu;
    for (int c : {slots, slots / 2, 3 * slots / 8, slots / 4, slots / 8}) {
      int Wc = batch > 1 ? (c + batch - 1) / batch : c;
      Wc = std::max(1, std::min(Wc, total));
      const double t = d_cost(g, batch, bn, Wc, cu);
      if (t < bt * 0.999) {
        bt = t;
        best = {bn, Wc};
      }
    }
  }
  std::lock_guard<std::mutex> lk(mu);
  cache[key] = best;
  return best;
}

bool gemm_d_supported(const GemmArgs& g) {
  const uintptr_t al = (uintptr_t)g.A | (uintptr_t)g.B;
  if (al & 15) return false;
  if ((g.lda | g.ldb | g.sA | g.sB) & 1) return false;
  if (g.M % 64 || g.N % 64 || g.K % D_BK || g.M < 128 || g.N < 64) return false;
  if ((g.flags & BTA_GEMM_C_LOWER) && g.M != g.N) return false;
  return true;
}

template <int BN>
static void d_launch(const GemmArgs& g, int batch, int W, hipStream_t s) {
  const bool akm = g.flags & BTA_GEMM_A_KM, bkn = g.flags & BTA_GEMM_B_KN;
  dim3 grid(W, batch);
  if (akm && bkn)
    hipLaunchKernelGGL((gemm_d_kernel<BN, true, true>), grid, dim3(256), 0, s, g);
  else if (akm)
    hipLaunchKernelGGL((gemm_d_kernel<BN, true, false>), grid, dim3(256), 0, s, g);
  else if (bkn)
    hipLaunchKernelGGL((gemm_d_kernel<BN, false, true>), grid, dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL((gemm_d_kernel<BN, false, false>), grid, dim3(256), 0, s, g);
}

int gemm_d(const GemmArgs& g, int batch, hipStream_t s) {
  if (batch == 0 || g.K == 0) return -1;  // caller's generic path handles these
  const DPlan p = d_plan(g, batch);
  if (p.bn == 128)
    d_launch<128>(g, batch, p.workers, s);
  else
    d_launch<64>(g, batch, p.workers, s);
  BTA_LAUNCH_CHECK("gemm_d_kernel");
  return BTA_OK;
}

}  // namespace bta
