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

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace bta {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));

constexpr int H2_BM = 128;
constexpr int H2_BK = 32;        // K per chunk (one MFMA K-step)
constexpr int H2_ROWB = 64;      // bytes per plane row per stage (32 halfs)
constexpr double H2_LO = 1.0 / 2048.0;  // weight of the lo planes (stored as 2^11 (x - hi))

template <int BN, bool ACC64>
struct H2Cfg {
  static constexpr int WN = BN / 2;                 // wave tile columns
  static constexpr int FM = 4, FN = WN / 16;        // 16 x 16 fragments per wave
  static constexpr int PA = H2_BM * H2_ROWB;        // bytes per A plane per stage
  static constexpr int PB = BN * H2_ROWB;           // bytes per B plane per stage
  static constexpr int STAGE = 2 * PA + 2 * PB;
  static constexpr int LA = PA / 1024, LB = PB / 1024;  // 1 KiB wave loads per plane
  static constexpr int LOADS_W = (2 * LA + 2 * LB) / 4;  // per wave per chunk
  // FP64 sums: one workgroup per CU, 4 stages; FP32 only: two per CU
  static constexpr int STAGES = ACC64 ? 4 : (BN == 64 ? 3 : 2);
  static constexpr int WG_PER_CU = ACC64 ? 1 : 2;
};

// Slot of 16-byte chunk c of plane row r is c ^ h2_swizzle(r % 16).  For a
// fixed c, rows 0..7 (and 8..15) land in distinct 16-byte slots of a 128-byte
// bank line and rows 0..15 in distinct slots of a 256-byte line, so fragment
// reads are conflict free whether the LDS serves 8 or 16 lanes per cycle.
__device__ __forceinline__ int h2_swizzle(int r) { return ((r >> 1) & 3) ^ ((r >> 3) & 1); }

// s_waitcnt on the vector-memory counter only (expcnt, lgkmcnt left open)
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

template <int OFF>
__device__ __forceinline__ half8 h2_read(uint32_t a) {
  half8 v;
  __asm__ __volatile__("ds_read_b128 %0, %1 offset:%2" : "=v"(v) : "v"(a), "i"(OFF));
  return v;
}
__device__ __forceinline__ void h2_tie(half8& v) { __asm__ __volatile__("" : "+v"(v)); }

// F fragments (16 rows, 1 KiB apart) of a plane pair: hi at a, lo at a + P
template <int F, int P>
__device__ __forceinline__ void h2_frag(uint32_t a, half8 (&hi)[F], half8 (&lo)[F]) {
  if (F > 0) { hi[0] = h2_read<0>(a); lo[0] = h2_read<P>(a); }
  if (F > 1) { hi[1] = h2_read<1024>(a); lo[1] = h2_read<P + 1024>(a); }
  if (F > 2) { hi[2] = h2_read<2048>(a); lo[2] = h2_read<P + 2048>(a); }
  if (F > 3) { hi[3] = h2_read<3072>(a); lo[3] = h2_read<P + 3072>(a); }
}

template <int BN, bool ACC64>
__global__ __launch_bounds__(256, (H2Cfg<BN, ACC64>::WG_PER_CU)) void gemm_h2_kernel(H2Args g) {
  using C = H2Cfg<BN, ACC64>;
  constexpr int S = C::STAGES;
  __shared__ __attribute__((aligned(1024))) unsigned char lds[S * C::STAGE];
  const int MT = g.M / H2_BM, NT = g.N / BN;
  const int total = MT * NT;
  const int W = gridDim.x, w = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wy = wv >> 1, wx = wv & 1, l16 = lane & 15, gq = lane >> 4;
  // fragment reads: row l16 of a 16-row block, chunk gq in its swizzled slot
  const int foff = l16 * H2_ROWB + ((gq ^ h2_swizzle(l16)) << 4);
  // direct loads: lane -> (row lane/4 of a 16-row block, slot lane%4, source chunk)
  const int lrow = lane >> 2;
  const int lchunk = (lane & 3) ^ h2_swizzle(lrow);
  const int seg = ACC64 ? (g.seg_chunks > 0 ? g.seg_chunks : 8) : 0;  // chunks per FP64 add

  for (int jj = 0;; ++jj) {
    const int t = h2_tile_of(jj, w, W);
    if (jj * W >= total) break;
    if (t >= total) continue;
    int tm, tn;
    decode_tile(t, g.flags, MT, NT, tm, tn);
    int kb = 0, ke = g.K;
    if (g.flags & BTA_GEMM_B_LOWER) ke = min(ke, (tn + 1) * BN);
    if (g.flags & BTA_GEMM_B_UPPER) kb = max(kb, tn * BN);
    kb = kb / H2_BK * H2_BK;
    ke = (ke + H2_BK - 1) / H2_BK * H2_BK;
    const int row0 = tm * H2_BM, col0 = tn * BN;
    const int nk = ke > kb ? (ke - kb) / H2_BK : 0;

    // the LOADS_W wave loads of chunk ci into stage ci % S
    auto issue = [&](int ci) {
      const int k0 = kb + ci * H2_BK + lchunk * 8;
      unsigned char* st = lds + (ci % S) * C::STAGE;
#pragma unroll
      for (int j = 0; j < C::LOADS_W; ++j) {
        const int gi = wv * C::LOADS_W + j;  // wave-uniform
        const _Float16* src;
        int off;
        if (gi < 2 * C::LA) {
          const int p = gi / C::LA, rb = gi % C::LA;
          src = (p ? g.Al : g.Ah) + (int64_t)(row0 + rb * 16 + lrow) * g.lda + k0;
          off = p * C::PA + rb * 1024;
        } else {
          const int gb = gi - 2 * C::LA;
          const int p = gb / C::LB, rb = gb % C::LB;
          src = (p ? g.Bl : g.Bh) + (int64_t)(col0 + rb * 16 + lrow) * g.ldb + k0;
          off = 2 * C::PA + p * C::PB + rb * 1024;
        }
        __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(st + off), 16, 0, 0);
      }
    };

    v4f acc[C::FM][C::FN], accx[C::FM][C::FN];
    double acc64[ACC64 ? C::FM : 1][ACC64 ? C::FN : 1][4];
#pragma unroll
    for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
      for (int ni = 0; ni < C::FN; ++ni) {
        acc[mi][ni] = (v4f)(0.0f);
        accx[mi][ni] = (v4f)(0.0f);
        if (ACC64)
#pragma unroll
          for (int q = 0; q < 4; ++q) acc64[ACC64 ? mi : 0][ACC64 ? ni : 0][q] = 0.0;
      }

#pragma unroll
    for (int ci = 0; ci < S - 1; ++ci)
      if (ci < nk) issue(ci);
    for (int it = 0; it < nk; ++it) {
      // this wave's loads of chunk it have landed once at most
      // min(S - 2, nk - 1 - it) later chunks remain outstanding
      const int later = min(S - 2, nk - 1 - it);
      if (S > 3 && later >= 2) wait_vmcnt<(S > 3 ? 2 : 0) * C::LOADS_W>();
      else if (S > 2 && later >= 1) wait_vmcnt<(S > 2 ? 1 : 0) * C::LOADS_W>();
      else wait_vmcnt<0>();
      // every wave's loads of chunk it landed and every wave is done with chunk
      // it - 1.  A bare s_barrier: the fence of __syncthreads would also wait for
      // the loads still in flight; the empty asm statements keep the compiler
      // from moving LDS accesses across it.
      // sched_barrier(0): the compiler moves nothing across, in particular no
      // LDS read of this chunk above the wait and the barrier (another wave's
      // direct loads fill part of the stage, and only the barrier orders them)
      __builtin_amdgcn_sched_barrier(0);
      __asm__ __volatile__("" ::: "memory");
      __builtin_amdgcn_s_barrier();
      __asm__ __volatile__("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
      if (it + S - 1 < nk) issue(it + S - 1);
      if (ACC64 && it % seg == 0 && it > 0) {
#pragma unroll
        for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
          for (int ni = 0; ni < C::FN; ++ni) {
            acc[mi][ni] = (v4f)(0.0f);
            accx[mi][ni] = (v4f)(0.0f);
          }
      }
      const unsigned char* cur = lds + (it % S) * C::STAGE;
      // fragment reads as inline asm (see lds_read_b128): no compiler-added
      // vmcnt(0) for the direct loads in flight; lgkmcnt waited below
      half8 ah[C::FM], al[C::FM], bh[C::FN], bl[C::FN];
      const uint32_t aa = lds_addr(cur + wy * 64 * H2_ROWB + foff);
      const uint32_t ba = lds_addr(cur + 2 * C::PA + wx * C::WN * H2_ROWB + foff);
      h2_frag<C::FM, C::PA>(aa, ah, al);
      h2_frag<C::FN, C::PB>(ba, bh, bl);
      lds_frag_wait();
#pragma unroll
      for (int f = 0; f < C::FM; ++f) {
        h2_tie(ah[f]);
        h2_tie(al[f]);
      }
#pragma unroll
      for (int f = 0; f < C::FN; ++f) {
        h2_tie(bh[f]);
        h2_tie(bl[f]);
      }
      // acc += A_hi B_hi, accx += A_lo B_hi + A_hi B_lo; three passes over the
      // fragments, so no MFMA waits on the one before it
#pragma unroll
      for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < C::FN; ++ni)
          accx[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[mi], bh[ni], accx[mi][ni], 0,
                                                                0, 0);
#pragma unroll
      for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < C::FN; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[mi], bh[ni], acc[mi][ni], 0,
                                                               0, 0);
#pragma unroll
      for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < C::FN; ++ni)
          accx[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[mi], bl[ni], accx[mi][ni], 0,
                                                                0, 0);
      if (ACC64 && ((it + 1) % seg == 0 || it + 1 == nk)) {
#pragma unroll
        for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
          for (int ni = 0; ni < C::FN; ++ni)
#pragma unroll
            for (int q = 0; q < 4; ++q)
              acc64[ACC64 ? mi : 0][ACC64 ? ni : 0][q] +=
                  (double)acc[mi][ni][q] + H2_LO * (double)accx[mi][ni][q];
      }
    }
    // all waves done reading the ring before the next tile's loads overwrite it
    __syncthreads();
    const double bs = (double)g.b_scale;
#pragma unroll
    for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t row = row0 + wy * 64 + mi * 16 + Mfma<float>::row(gq, q);
        const double f =
            g.alpha / ((double)g.a_scale * (g.a_row_scale ? (double)g.a_row_scale[row] : 1.0) * bs);
#pragma unroll
        for (int ni = 0; ni < C::FN; ++ni) {
          const int64_t col = col0 + wx * C::WN + ni * 16 + l16;
          const double sum = ACC64 ? acc64[ACC64 ? mi : 0][ACC64 ? ni : 0][q]
                                   : (double)acc[mi][ni][q] + H2_LO * (double)accx[mi][ni][q];
          const float v = (float)(sum * f);
          g.C[row * g.ldc + col] = v;
          if (g.Ch) {
            const float x = v * g.c_scale;
            const _Float16 h = (_Float16)x;
            g.Ch[row * g.ldc + col] = h;
            g.Cl[row * g.ldc + col] = (_Float16)((x - (float)h) * 2048.0f);
          }
        }
      }
  }
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
  int bn, workers;
};

static H2Plan h2_plan(const H2Args& g, bool acc64) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int, int>, H2Plan> cache;
  const auto key = std::make_tuple(g.M, g.N, g.K, g.flags, (int)acc64);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  const int ncu = num_cus();
  const int slots = acc64 ? ncu : 2 * ncu;  // resident workgroups
  H2Plan best{64, 1};
  double bt = 1e300;
  for (int bn : {64, 128}) {
    if (g.N % bn) continue;
    const int total = (g.M / H2_BM) * (g.N / bn);
    for (int c : {slots, slots / 2, 3 * slots / 8, slots / 4}) {
      const int W = std::max(1, std::min(c, total));
      const double t = h2_cost(g, bn, W);
      if (t < bt) {
        bt = t;
        best = {bn, W};
      }
    }
  }
  std::lock_guard<std::mutex> lk(mu);
  cache[key] = best;
  return best;
}

int gemm_h2(const H2Args& a, hipStream_t s) {
  const uintptr_t al16 = (uintptr_t)a.Ah | (uintptr_t)a.Al | (uintptr_t)a.Bh | (uintptr_t)a.Bl;
  if (a.M % H2_BM || a.N % 64 || a.K % H2_BK || a.M <= 0 || a.N <= 0 || a.K <= 0 ||
      a.lda % 8 || a.ldb % 8 || a.lda < a.K || a.ldb < a.K || a.ldc < a.N || !a.Ah || !a.Al ||
      !a.Bh || !a.Bl || !a.C || (a.Ch && !a.Cl) || (al16 & 15) || a.seg_chunks < 0 ||
      (a.flags & ~(BTA_GEMM_B_LOWER | BTA_GEMM_B_UPPER))) {
    set_error("gemm_h2: M%%128, N%%64, K%%32, lda/ldb%%8 must be 0, operand planes 16-byte "
              "aligned, K > 0; flags B_LOWER/B_UPPER only");
    return BTA_ERR_INVALID;
  }
  const bool acc64 = a.seg_chunks > 0;
  const H2Plan p = h2_plan(a, acc64);
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
