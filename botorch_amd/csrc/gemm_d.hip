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

// K range of tile t; with nsplit > 0 the first nsplit tiles (the longest:
// tiles are enumerated longest-first) are cut in two at a chunk boundary,
// piece 0 being task t and piece 1 task total + t.
__host__ __device__ inline void d_task_range(const GemmArgs& g, int BN, int MT, int NT,
                                             int total, int nsplit, int task, int& tm, int& tn,
                                             int& kb, int& ke) {
  const int t = task < total ? task : task - total;
  decode_tile(t, g.flags, MT, NT, tm, tn);
  kb = 0;
  ke = g.K;
  if (g.flags & BTA_GEMM_A_LOWER) ke = min(ke, (tm + 1) * D_BM);
  if (g.flags & BTA_GEMM_A_UPPER) kb = max(kb, tm * D_BM);
  if (g.flags & BTA_GEMM_B_LOWER) ke = min(ke, (tn + 1) * BN);
  if (g.flags & BTA_GEMM_B_UPPER) kb = max(kb, tn * BN);
  kb = kb / D_BK * D_BK;
  ke = (ke + D_BK - 1) / D_BK * D_BK;
  if (t < nsplit) {
    const int mid = kb + (ke - kb) / (2 * D_BK) * D_BK;
    if (task < total)
      ke = mid;
    else
      kb = mid;
  }
}

// SPLIT: C was zeroed and every piece adds alpha * its product with an FP64
// atomic.  Each element then receives at most two addends on top of an exact
// zero, and a sum of two terms does not depend on their order: the result is
// the same on every run.
template <int BN, bool AKM, bool BKN, bool SPLIT>
__global__ __launch_bounds__(256, 2) void gemm_d_kernel(GemmArgs g, int nsplit) {
  using C = DCfg<BN>;
  __shared__ __attribute__((aligned(1024))) unsigned char lds[C::STAGES * C::STAGE];
  const int64_t batch = blockIdx.y;
  const double* A = (const double*)g.A + batch * g.sA;
  const double* B = (const double*)g.B + batch * g.sB;
  double* Cp = (double*)g.C + batch * g.sC;
  const double* C0 = g.C0 ? (const double*)g.C0 + batch * g.sC0 : nullptr;
  const int MT = (g.M + D_BM - 1) / D_BM, NT = (g.N + BN - 1) / BN;
  const int total = (g.flags & BTA_GEMM_C_LOWER) ? MT * (MT + 1) / 2 : MT * NT;
  const int tasks = SPLIT ? total + nsplit : total;
  const int W = gridDim.x, w = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wy = wv >> 1, wx = wv & 1, l16 = lane & 15, gq = lane >> 4;

  for (int jj = 0; jj * W < tasks; ++jj) {
    const int task = h2_tile_of(jj, w, W);
    if (task >= tasks) continue;
    int tm, tn, kb, ke;
    d_task_range(g, BN, MT, NT, total, SPLIT ? nsplit : 0, task, tm, tn, kb, ke);
    v4d acc[C::FM][C::FN];
    d_mainloop<BN, AKM, BKN>(A, g.lda, B, g.ldb, tm * D_BM, tn * BN, g.M, g.N, kb, ke, lds,
                             acc);
    if (SPLIT) {
#pragma unroll
      for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int64_t row = tm * D_BM + wy * 64 + mi * 16 + Mfma<double>::row(gq, q);
          if (row >= g.M) continue;
#pragma unroll
          for (int ni = 0; ni < C::FN; ++ni) {
            const int64_t col = tn * BN + wx * C::WN + ni * 16 + l16;
            if (col < g.N) unsafeAtomicAdd(Cp + row * g.ldc + col, g.alpha * acc[mi][ni][q]);
          }
        }
      continue;
    }
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

// C[b][row][col] = 0 over M x N per batch (N even, C 16-byte aligned rows)
__global__ void d_zero_kernel(double* C, int64_t ldc, int N, int64_t sC) {
  double* row = C + (int64_t)blockIdx.z * sC + (int64_t)blockIdx.y * ldc;
  for (int c = 2 * (blockIdx.x * 256 + threadIdx.x); c < N; c += 2 * gridDim.x * 256)
    *(v2d*)(row + c) = (v2d)(0.0);
}

// Time of one CU running workgroups with the given loads from the start: with
// two or more active it is saturated (rate 1, shared equally), a workgroup
// alone runs at 0.65 of that.
static double cu_time(std::vector<double>& loads) {
  std::sort(loads.begin(), loads.end());
  double t = 0.0, done = 0.0;  // done: work finished by every still-active workgroup
  const int k = (int)loads.size();
  for (int i = 0; i < k; ++i) {
    const int active = k - i;
    const double rate = active >= 2 ? 1.0 / active : 0.65;
    t += (loads[i] - done) / rate;
    done = loads[i];
  }
  return t;
}

// Cost model of the worker schedule (chunks per worker plus a per-task fixed
// cost, workgroup i of the launch on CU i % ncu).  Units: about 0.8 us.
static double d_cost(const GemmArgs& g, int batch, int BN, int W, int nsplit,
                     std::vector<double>& cu) {
  const int MT = (g.M + D_BM - 1) / D_BM, NT = (g.N + BN - 1) / BN;
  const int total = (g.flags & BTA_GEMM_C_LOWER) ? MT * (MT + 1) / 2 : MT * NT;
  const int tasks = total + nsplit;
  const int ncu = (int)cu.size();
  const double per_chunk = BN == 128 ? 1.0 : 0.55;
  std::vector<std::vector<double>> per_cu(ncu);
  for (int w = 0; w < W; ++w) {
    double load = 0.0;
    for (int j = 0; j * W < tasks; ++j) {
      const int task = h2_tile_of(j, w, W);
      if (task >= tasks) continue;
      int tm, tn, kb, ke;
      d_task_range(g, BN, MT, NT, total, nsplit, task, tm, tn, kb, ke);
      load += per_chunk * (ke > kb ? (ke - kb) / D_BK : 0) + 6.0;
    }
    for (int b = 0; b < batch; ++b) per_cu[((int64_t)b * W + w) % ncu].push_back(load);
  }
  double t = 0.0;
  for (auto& v : per_cu) t = std::max(t, cu_time(v));
  // zeroing C before a split launch: one more launch (about 5 us measured for
  // the small ones) plus M N batch doubles at about 3 TB/s
  if (nsplit > 0) t += 6.5 + (double)g.M * g.N * batch * 8.0 / 3e12 / 0.8e-6;
  return t;
}

struct DPlan {
  int bn, workers, nsplit;
};

// Tiles whose K range is longer than half of K may be cut in two when the
// output has no C0 term and exactly one triangular flag orders the tiles by
// length (the longest first).
static int d_split_candidates(const GemmArgs& g, int BN) {
  const int tri = g.flags & (BTA_GEMM_A_LOWER | BTA_GEMM_A_UPPER | BTA_GEMM_B_LOWER |
                             BTA_GEMM_B_UPPER);
  if (g.C0 || (g.flags & BTA_GEMM_C_LOWER) || tri == 0 || (tri & (tri - 1)) || g.K < 256)
    return 0;
  if (((uintptr_t)g.C & 15) || ((g.ldc | g.sC) & 1)) return 0;  // zeroed 16 bytes at a time
  const int MT = (g.M + D_BM - 1) / D_BM, NT = (g.N + BN - 1) / BN, total = MT * NT;
  int n = 0;
  for (int t = 0; t < total; ++t) {
    int tm, tn, kb, ke;
    d_task_range(g, BN, MT, NT, total, 0, t, tm, tn, kb, ke);
    if (ke - kb <= g.K / 2) break;
    ++n;
  }
  return n;
}

static DPlan d_plan(const GemmArgs& g, int batch) {
  // never destroyed (no exit-time destructor; see plan_gemm)
  static std::mutex& mu = *new std::mutex;
  static auto& cache = *new std::map<std::array<int64_t, 7>, DPlan>;
  const int ncu = num_cus();
  const std::array<int64_t, 7> key = {g.M, g.N, g.K, g.flags, batch, ncu, g.C0 != nullptr};
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  std::vector<double> cu(ncu);
  DPlan best{128, 1, 0};
  double bt = 1e300;
  for (int bn : {128, 64}) {
    if (bn == 64 && (g.flags & BTA_GEMM_C_LOWER)) continue;  // square tiles only
    const int MT = (g.M + D_BM - 1) / D_BM, NT = (g.N + bn - 1) / bn;
    const int total = (g.flags & BTA_GEMM_C_LOWER) ? MT * (MT + 1) / 2 : MT * NT;
    const int ns = d_split_candidates(g, bn);
    for (int split : {0, ns}) {
      const int slots = 2 * ncu;
      for (int c : {slots, slots / 2, 3 * slots / 8, slots / 4, slots / 8}) {
        int Wc = batch > 1 ? (c + batch - 1) / batch : c;
        Wc = std::max(1, std::min(Wc, total + split));
        const double t = d_cost(g, batch, bn, Wc, split, cu);
        if (t < bt * 0.999) {
          bt = t;
          best = {bn, Wc, split};
        }
      }
      if (ns == 0) break;
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

template <int BN, bool SPLIT>
static void d_launch(const GemmArgs& g, int batch, int W, int nsplit, hipStream_t s) {
  const bool akm = g.flags & BTA_GEMM_A_KM, bkn = g.flags & BTA_GEMM_B_KN;
  dim3 grid(W, batch);
  if (akm && bkn)
    hipLaunchKernelGGL((gemm_d_kernel<BN, true, true, SPLIT>), grid, dim3(256), 0, s, g, nsplit);
  else if (akm)
    hipLaunchKernelGGL((gemm_d_kernel<BN, true, false, SPLIT>), grid, dim3(256), 0, s, g, nsplit);
  else if (bkn)
    hipLaunchKernelGGL((gemm_d_kernel<BN, false, true, SPLIT>), grid, dim3(256), 0, s, g, nsplit);
  else
    hipLaunchKernelGGL((gemm_d_kernel<BN, false, false, SPLIT>), grid, dim3(256), 0, s, g,
                       nsplit);
}

int gemm_d(const GemmArgs& g, int batch, hipStream_t s) {
  if (batch == 0 || g.K == 0) return -1;  // caller's generic path handles these
  const DPlan p = d_plan(g, batch);
  if (p.nsplit > 0) {
    dim3 zg((unsigned)std::min((g.N + 511) / 512, 8), (unsigned)g.M, (unsigned)batch);
    hipLaunchKernelGGL(d_zero_kernel, zg, dim3(256), 0, s, (double*)g.C, g.ldc, g.N, g.sC);
    BTA_LAUNCH_CHECK("d_zero_kernel");
    if (p.bn == 128)
      d_launch<128, true>(g, batch, p.workers, p.nsplit, s);
    else
      d_launch<64, true>(g, batch, p.workers, p.nsplit, s);
  } else if (p.bn == 128) {
    d_launch<128, false>(g, batch, p.workers, 0, s);
  } else {
    d_launch<64, false>(g, batch, p.workers, 0, s);
  }
  BTA_LAUNCH_CHECK("gemm_d_kernel");
  return BTA_OK;
}

}  // namespace bta
