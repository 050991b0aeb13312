// Split-FP16 GEMM on pre-split operands ("h2"):
//   C = alpha * (A_hi + A_lo/2^11)(B_hi + B_lo/2^11)^T / (a_scale a_row_scale[r] b_scale)
// with the lo x lo term dropped (2^-22 relative).  A plane pair holds
// hi = fp16(x) and lo = fp16(2^11 (x - hi)): the lo plane stays in FP16's normal
// range for small x, so every entry keeps ~22 significant bits.  A (M x K) and
// B (N x K) are two FP16 planes each, K-contiguous, written by their producers
// (the covariance kernel, this GEMM's own split epilogue, the factor set-up),
// so the main loop is data movement plus three v_mfma_f32_16x16x32_f16 per
// fragment pair.  hi x hi and the two cross products accumulate in separate
// FP32 sets (the cross set weighted by 2^-11 when combined), either over the
// whole K range (default) or in FP32 segments of k_segment added into FP64
// sums.  With the lo planes in range, the operand representation dominates
// the error: FP32 accumulation over K = 2048 leaves V V^T as accurate as FP64
// segment sums do, so the default path keeps no FP64 accumulators and runs
// two workgroups (two waves per SIMD) on each CU.
//
// Tile 128 x BN (BN = 128 or 64), 256 threads = 4 waves of 64 x BN/2, K chunks
// of 32.  Operands move global -> LDS with direct LDS loads
// (global_load_lds_dwordx4: no registers, no ds_write) through a ring of LDS
// stages with STAGES - 1 chunks in flight; one barrier per chunk.  Each plane
// row (64 bytes) holds its four 16-byte chunks in XOR-swizzled slots
// (h2_swizzle), which keeps the fragment reads free of bank conflicts; the
// 1 KiB wave loads write contiguous LDS.  The epilogue writes C in FP32 and,
// optionally, C's own hi/lo planes (scaled by c_scale) for a following split
// GEMM.  Triangular K ranges follow gemm_kernel; a cost model picks BN and the
// worker count, and workers pair heavy with light tiles on each CU.
#include "bta_common.h"
#include "gemm.h"
#include "gemm_d.h"
#include "gemm_tile.h"
#include "kernels.h"
#include "gemm_h2_body.h"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace bta {

template <int BN, bool ACC64>
__global__ __launch_bounds__(256, (H2Cfg<BN, ACC64>::WG_PER_CU)) void gemm_h2_kernel(H2Args g) {
  __shared__ __attribute__((aligned(1024)))
  unsigned char lds[H2Cfg<BN, ACC64>::STAGES * H2Cfg<BN, ACC64>::STAGE];
  gemm_h2_body<BN, ACC64>(g, blockIdx.x, gridDim.x, lds);
}

// Cost model of the worker schedule: chunk counts per worker (plus a per-tile
// fixed cost); a CU's time is the sum over the workers it hosts, worker w on
// CU w % ncu.
static double h2_cost(const H2Args& g, int BN, int W) {
  const int MT = g.M / H2_BM, NT = g.N / BN, total = MT * NT;
  const int ncu = num_cus();
  std::vector<double> cu(ncu, 0.0);
  const double per_chunk = BN == 128 ? 1.0 : 0.55;  // MFMA + LDS time of one K chunk
  for (int w = 0; w < W; ++w) {
    double load = 0.0;
    for (int j = 0; j * W < total; ++j) {
      const int t = h2_tile_of(j, w, W);
      if (t >= total) continue;
      int tm, tn;
      decode_tile(t, g.flags, MT, NT, tm, tn);
      int kb = 0, ke = g.K;
      if (g.flags & BTA_GEMM_B_LOWER) ke = std::min(ke, (tn + 1) * BN);
      if (g.flags & BTA_GEMM_B_UPPER) kb = std::max(kb, tn * BN);
      load += per_chunk * (ke > kb ? (ke - kb) / H2_BK : 0) + 6.0;
    }
    cu[w % ncu] += load;
  }
  return *std::max_element(cu.begin(), cu.end());
}

struct H2Plan {
  int bn, workers, xcd;
};

// Worker groups for gemm_h2_body's strip dealing: one per XCD (workgroups go
// to the XCDs round-robin by id; 32 CUs per XCD), when the tile order is strip
// by strip (triangular B), the strips pair up evenly across the groups (a
// multiple of 2 G of them) and the worker count divides evenly.  Off unless
// BTA_H2_XCD=1 is set: it lets one L2 serve each L^-1 strip, but on the
// 2048 x 2048 x 2048 step products it measured no faster (50.7-52.6 us against
// 50.2-52.9 us per forward product, step 0.189-0.193 ms against 0.185-0.187 ms).
static int h2_xcd_allowed(const H2Args& g, int BN, int W) {
  const int G = num_cus() / 32;
  if (G < 2 || !(g.flags & (BTA_GEMM_B_LOWER | BTA_GEMM_B_UPPER)) || W % G ||
      (g.N / BN) % (2 * G))
    return 1;
  return G;
}

static int h2_xcd_groups(const H2Args& g, int BN, int W) {
  static const int on = [] {
    const char* e = getenv("BTA_H2_XCD");
    return e && e[0] == '1';
  }();
  return on ? h2_xcd_allowed(g, BN, W) : 1;
}

// The plans per problem signature: filled by the cost model below on a shape's
// first call, or by the measured tuner (gemm_h2_set_plan).  Never destroyed (no
// exit-time destructor; see plan_gemm).
typedef std::tuple<int, int, int, int, int> H2Key;
static std::mutex& h2_mu() {
  static std::mutex& mu = *new std::mutex;
  return mu;
}
static std::map<H2Key, H2Plan>& h2_cache() {
  static auto& cache = *new std::map<H2Key, H2Plan>;
  return cache;
}

static H2Plan h2_plan(const H2Args& g, bool acc64) {
  std::mutex& mu = h2_mu();
  auto& cache = h2_cache();
  const auto key = std::make_tuple(g.M, g.N, g.K, g.flags, (int)acc64);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  const int ncu = num_cus();
  const int slots = acc64 ? ncu : 2 * ncu;  // resident workgroups
  H2Plan best{64, 1, 1};
  double bt = 1e300;
  for (int bn : {64, 128}) {
    if (g.N % bn) continue;
    const int total = (g.M / H2_BM) * (g.N / bn);
    for (int c : {slots, slots / 2, 3 * slots / 8, slots / 4}) {
      const int W = std::max(1, std::min(c, total));
      const double t = h2_cost(g, bn, W);
      if (t < bt) {
        bt = t;
        best = {bn, W, h2_xcd_groups(g, bn, W)};
      }
    }
  }
  std::lock_guard<std::mutex> lk(mu);
  cache[key] = best;
  return best;
}

static int h2_check(const H2Args& a) {
  const uintptr_t al16 = (uintptr_t)a.Ah | (uintptr_t)a.Al | (uintptr_t)a.Bh | (uintptr_t)a.Bl;
  if (a.M % H2_BM || a.N % 64 || a.K % H2_BK || a.M <= 0 || a.N <= 0 || a.K <= 0 ||
      a.lda % 8 || a.ldb % 8 || a.lda < a.K || a.ldb < a.K || a.ldc < a.N || !a.Ah || !a.Al ||
      !a.Bh || !a.Bl || !a.C || (a.Ch && !a.Cl) || (al16 & 15) || a.seg_chunks < 0 ||
      (a.flags & ~(BTA_GEMM_B_LOWER | BTA_GEMM_B_UPPER))) {
    set_error("gemm_h2: M%%128, N%%64, K%%32, lda/ldb%%8 must be 0, operand planes 16-byte "
              "aligned, K > 0; flags B_LOWER/B_UPPER only");
    return BTA_ERR_INVALID;
  }
  return BTA_OK;
}

int gemm_h2_candidates(const H2Args& g, int* out, int max) {
  if (h2_check(g)) return 0;
  const int ncu = num_cus();
  const bool acc64 = g.seg_chunks > 0;
  const int slots = acc64 ? ncu : 2 * ncu;
  int k = 0;
  for (int bn : {64, 128}) {
    if (g.N % bn) continue;
    const int total = (g.M / H2_BM) * (g.N / bn);
    int last = -1;
    for (int c : {slots, 3 * slots / 4, slots / 2, 3 * slots / 8, slots / 4}) {
      const int W = std::max(1, std::min(c, total));
      if (W == last) continue;
      last = W;
      for (int G : {1, h2_xcd_allowed(g, bn, W)}) {
        if (G == 1 && k > 0 && out[3 * (k - 1)] == bn && out[3 * (k - 1) + 1] == W &&
            out[3 * (k - 1) + 2] == 1)
          continue;
        if (k >= max) return k;
        out[3 * k] = bn;
        out[3 * k + 1] = W;
        out[3 * k + 2] = G;
        ++k;
      }
    }
  }
  return k;
}

void gemm_h2_set_plan(const H2Args& g, int bn, int workers, int xcd_groups) {
  const auto key = std::make_tuple(g.M, g.N, g.K, g.flags, (int)(g.seg_chunks > 0));
  std::lock_guard<std::mutex> lk(h2_mu());
  h2_cache()[key] = H2Plan{bn, workers, xcd_groups};
}

int gemm_h2_plan(const H2Args& a, int* bn, int* workers, int* xcd_groups) {
  const int rc = h2_check(a);
  if (rc) return rc;
  const H2Plan p = h2_plan(a, a.seg_chunks > 0);
  *bn = p.bn;
  *workers = p.workers;
  if (xcd_groups) *xcd_groups = p.xcd;
  return BTA_OK;
}

int gemm_h2(const H2Args& args, hipStream_t s) {
  const int rc = h2_check(args);
  if (rc) return rc;
  const bool acc64 = args.seg_chunks > 0;
  const H2Plan p = h2_plan(args, acc64);
  H2Args a = args;
  a.xcd_groups = p.xcd;
  if (p.bn == 128) {
    if (acc64) hipLaunchKernelGGL((gemm_h2_kernel<128, true>), dim3(p.workers), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((gemm_h2_kernel<128, false>), dim3(p.workers), dim3(256), 0, s, a);
  } else {
    if (acc64) hipLaunchKernelGGL((gemm_h2_kernel<64, true>), dim3(p.workers), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((gemm_h2_kernel<64, false>), dim3(p.workers), dim3(256), 0, s, a);
  }
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
    lo[orow * ldo + ocol] = (_Float16)((x - (float)h) * 2048.0f);
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
                              int32_t k_segment, bta_stream_t stream) {
  if (M > (1 << 30) || N > (1 << 30) || K > (1 << 30) || k_segment < 0 || k_segment % 32) {
    bta::set_error("bta_gemm_f16x3: sizes beyond 2^30 or k_segment not a multiple of 32");
    return BTA_ERR_INVALID;
  }
  bta::H2Args a{};
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
  a.seg_chunks = k_segment / 32;
  return bta::gemm_h2(a, (hipStream_t)stream);
}
