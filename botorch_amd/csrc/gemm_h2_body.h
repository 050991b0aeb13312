// Split-FP16 GEMM worker body shared by gemm_h2_kernel and the fused
// adjoint-GEMM + MC launch of the qLogNEI step (the algorithm is described at
// the top of gemm_h2.hip).
#pragma once
#include "bta_common.h"
#include "gemm.h"
#include "gemm_d.h"
#include "gemm_tile.h"
#include "kernels.h"

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

// The main loop and epilogue of one 128 x BN output tile (tm, tn), with the
// ring in lds (H2Cfg<BN, ACC64>::STAGES * STAGE bytes of LDS, 1 KiB aligned).
template <int BN, bool ACC64>
__device__ __forceinline__ void gemm_h2_tile(const H2Args& g, const int tm, const int tn,
                                             unsigned char* lds) {
  using C = H2Cfg<BN, ACC64>;
  constexpr int S = C::STAGES;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wy = wv >> 1, wx = wv & 1, l16 = lane & 15, gq = lane >> 4;
  // fragment reads: row l16 of a 16-row block, chunk gq in its swizzled slot
  const int foff = l16 * H2_ROWB + ((gq ^ h2_swizzle(l16)) << 4);
  // direct loads: lane -> (row lane/4 of a 16-row block, slot lane%4, source chunk)
  const int lrow = lane >> 2;
  const int lchunk = (lane & 3) ^ h2_swizzle(lrow);
  const int seg = ACC64 ? (g.seg_chunks > 0 ? g.seg_chunks : 8) : 0;  // chunks per FP64 add
  {
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

// Worker w of W: the tiles h2_tile_of(jj, w, W), jj = 0, 1, ... of the
// problem's longest-first tile order.  With g.xcd_groups = G > 1 (triangular
// B, tiles ordered column strip by column strip, NT a multiple of 2 G), the
// workers form G groups by w % G - the XCD a workgroup is dispatched to - and
// group x takes the strips x, 2G - 1 - x, 2G + x, 4G - 1 - x, ... of that order
// (a snake, so that every group's strips add up to the same K length), so all
// row tiles of a strip of B run on one XCD and read the strip through one L2;
// within the group the tiles are dealt as above over its W / G workers (W a
// multiple of G), still longest first.
template <int BN, bool ACC64>
__device__ __forceinline__ void gemm_h2_body(const H2Args& g, int w, int W, unsigned char* lds) {
  const int MT = g.M / H2_BM, NT = g.N / BN;
  const int G = g.xcd_groups;
  if (G > 1) {
    const int x = w % G, lw = w / G, Wx = W / G;
    const int total = (NT / G) * MT;  // MT row tiles of each of the group's NT / G strips
    for (int jj = 0;; ++jj) {
      const int u = h2_tile_of(jj, lw, Wx);
      if (jj * Wx >= total) break;
      if (u >= total) continue;
      const int j = u / MT;
      const int strip = j * G + ((j & 1) ? G - 1 - x : x);
      const int t = strip * MT + u % MT;
      int tm, tn;
      decode_tile(t, g.flags, MT, NT, tm, tn);
      gemm_h2_tile<BN, ACC64>(g, tm, tn, lds);
    }
    return;
  }
  const int total = MT * NT;
  for (int jj = 0;; ++jj) {
    const int t = h2_tile_of(jj, w, W);
    if (jj * W >= total) break;
    if (t >= total) continue;
    int tm, tn;
    decode_tile(t, g.flags, MT, NT, tm, tn);
    gemm_h2_tile<BN, ACC64>(g, tm, tn, lds);
  }
}

// Worker w of W over a list of tiles of many problems: tiles[t] = (problem,
// tm, tn, -), listed longest K range first; problem p's arguments are tab[p].
template <int BN>
__device__ __forceinline__ void gemm_h2_list_body(const H2Args* __restrict__ tab,
                                                  const int4* __restrict__ tiles, int total,
                                                  int w, int W, unsigned char* lds) {
  for (int jj = 0;; ++jj) {
    const int t = h2_tile_of(jj, w, W);
    if (jj * W >= total) break;
    if (t >= total) continue;
    const int4 e = tiles[t];
    const H2Args g = tab[e.x];
    gemm_h2_tile<BN, false>(g, e.y, e.z, lds);
  }
}

}  // namespace bta
