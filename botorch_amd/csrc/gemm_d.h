// FP64 MFMA main loop with direct global -> LDS loads, shared by the FP64
// GEMM (gemm_d.hip) and the MLL trace kernel.
#pragma once
#include "bta_common.h"
#include "gemm.h"
#include "gemm_tile.h"

namespace bta {

typedef __attribute__((address_space(3))) void* lds_ptr_t;

constexpr int D_BM = 128;  // tile rows
constexpr int D_BK = 8;    // K per chunk (64 bytes of a K-contiguous row)

template <int BN>
struct DCfg {
  static constexpr int WN = BN / 2;              // wave tile columns
  static constexpr int FM = 4, FN = WN / 16;     // 16 x 16 blocks per wave
  static constexpr int PA = D_BM * D_BK * 8;     // bytes of A per stage
  static constexpr int PB = BN * D_BK * 8;       // bytes of B per stage
  static constexpr int STAGE = PA + PB;
  static constexpr int LA = PA / 1024, LB = PB / 1024;  // 1 KiB wave loads
  static constexpr int LOADS_W = (LA + LB) / 4;          // per wave per chunk
  static constexpr int STAGES = 4;
  static_assert((LA + LB) % 4 == 0, "wave loads must split over 4 waves");
};

// Slot of 16-byte chunk c of a K-contiguous stage row r is c ^ d_swizzle(r % 16):
// for a fixed c, rows 0..7 (8..15) fall in distinct slots of a 128-byte line.
__device__ __forceinline__ int d_swizzle(int r) { return ((r >> 1) & 3) ^ ((r >> 3) & 1); }

// s_waitcnt on the vector-memory counter only (expcnt, lgkmcnt left open)
template <int N>
__device__ __forceinline__ void d_wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// v from the first active lane: a value the compiler then treats as scalar.
// (A non-template function: the builtin exists only in the device pass.)
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// LDS fragment reads as inline asm.  The compiler then sees no LDS read after
// the direct loads, so it adds no s_waitcnt vmcnt(0) in front of every chunk's
// fragment reads (which would drain the whole ring each iteration); the
// explicit vmcnt wait and the barrier in the main loop order the data, and
// lds_frag_wait() waits for the reads themselves.
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const void*)p;
}
template <int OFF>
__device__ __forceinline__ v2d lds_read_b128(uint32_t a) {
  v2d v;
  __asm__ __volatile__("ds_read_b128 %0, %1 offset:%2" : "=v"(v) : "v"(a), "i"(OFF));
  return v;
}
template <int OFF>
__device__ __forceinline__ double lds_read_b64(uint32_t a) {
  double v;
  __asm__ __volatile__("ds_read_b64 %0, %1 offset:%2" : "=v"(v) : "v"(a), "i"(OFF));
  return v;
}
__device__ __forceinline__ void lds_frag_wait() {
  __asm__ __volatile__("s_waitcnt lgkmcnt(0)" ::: "memory");
}
// ties v to the wait above: nothing using v is scheduled before it
__device__ __forceinline__ void lds_tie(double& v) { __asm__ __volatile__("" : "+v"(v)); }

// Tile of position p in round j of W workers: rounds alternate direction, and
// within a round the second half of the workers takes the tiles in reverse, so
// workers p and p + W/2 (one CU when two workgroups share it) pair a heavy and
// a light tile.
__host__ __device__ inline int h2_tile_of(int j, int p, int W) {
  const int h = W / 2;
  int q = (W > 1 && p >= h) ? W - 1 - (p - h) : p;
  if (j & 1) q = W - 1 - q;
  return j * W + q;
}

// F fragments of 16 rows each from a K-contiguous stage (64-byte rows): one
// b128 per fragment at a0 + f * 1024.
template <int F>
__device__ __forceinline__ void frag_nk(uint32_t a0, double (&x)[F][2]) {
  v2d v[F];
  if (F > 0) v[0] = lds_read_b128<0>(a0);
  if (F > 1) v[1] = lds_read_b128<1024>(a0);
  if (F > 2) v[2] = lds_read_b128<2048>(a0);
  if (F > 3) v[3] = lds_read_b128<3072>(a0);
#pragma unroll
  for (int f = 0; f < F; ++f) {
    x[f][0] = v[f].x;
    x[f][1] = v[f].y;
  }
}
// F fragments from a K-major stage of W doubles per k row: k = 2g and 2g + 1
// (a0 points at row 2g), fragment f at +16 f doubles.
template <int F, int W>
__device__ __forceinline__ void frag_mn(uint32_t a0, double (&x)[F][2]) {
  if (F > 0) { x[0][0] = lds_read_b64<0>(a0); x[0][1] = lds_read_b64<W * 8>(a0); }
  if (F > 1) { x[1][0] = lds_read_b64<128>(a0); x[1][1] = lds_read_b64<W * 8 + 128>(a0); }
  if (F > 2) { x[2][0] = lds_read_b64<256>(a0); x[2][1] = lds_read_b64<W * 8 + 256>(a0); }
  if (F > 3) { x[3][0] = lds_read_b64<384>(a0); x[3][1] = lds_read_b64<W * 8 + 384>(a0); }
}

// acc = op(A)[row0 : row0+128, kb:ke] op(B)[kb:ke, col0 : col0+BN] for the
// calling wave's 64 x BN/2 part.  A is M x K (lda) or, with AKM, K x M; B is
// N x K (ldb) or, with BKN, K x N.  Rows/columns at or past M/N read the
// in-bounds ones 64 earlier.  All 256 threads must call it; it ends with a
// barrier, so the caller may reuse the LDS afterwards.
template <int BN, bool AKM, bool BKN>
__device__ __forceinline__ void d_mainloop(const double* __restrict__ A, int64_t lda,
                                           const double* __restrict__ B, int64_t ldb, int row0,
                                           int col0, int M, int N, int kb, int ke,
                                           unsigned char* lds,
                                           v4d (&acc)[DCfg<BN>::FM][DCfg<BN>::FN]) {
  using C = DCfg<BN>;
  constexpr int S = C::STAGES;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wy = wv >> 1, wx = wv & 1, l16 = lane & 15, gq = lane >> 4;
  const int lrow = lane >> 2;
  const int lchunk = (lane & 3) ^ d_swizzle(lrow);
  const int nk = ke > kb ? (ke - kb) / D_BK : 0;

#pragma unroll
  for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
    for (int ni = 0; ni < C::FN; ++ni) acc[mi][ni] = (v4d)(0.0);

  // Each wave load j of a chunk reads one 1 KiB piece: its global address at
  // chunk 0, the per-chunk advance (D_BK columns, or D_BK rows of a K-major
  // operand) and its LDS offset are fixed, so the loop only adds ci * step.
  // The wave index is made scalar so the A/B choice is a scalar branch.
  const int wvs = wave_uniform(wv);
  const double* gsrc[C::LOADS_W];
  int64_t gstep[C::LOADS_W];
  int goff[C::LOADS_W];
#pragma unroll
  for (int j = 0; j < C::LOADS_W; ++j) {
    const int gi = wvs * C::LOADS_W + j;
    if (gi < C::LA) {
      goff[j] = gi * 1024;
      if (AKM) {  // stage rows are k, 128 doubles each: one k row per wave load
        int m = row0 + lane * 2;
        if (m >= M) m -= 64;
        gsrc[j] = A + (int64_t)(kb + gi) * lda + m;
        gstep[j] = (int64_t)D_BK * lda;
      } else {
        int r = row0 + gi * 16 + lrow;
        if (r >= M) r -= 64;
        gsrc[j] = A + (int64_t)r * lda + kb + lchunk * 2;
        gstep[j] = D_BK;
      }
    } else {
      const int gb = gi - C::LA;
      goff[j] = C::PA + gb * 1024;
      if (BKN) {  // stage rows are k, BN doubles each
        const int e = gb * 128 + lane * 2;
        const int kr = e / BN;
        int n = col0 + e % BN;
        if (n >= N) n -= 64;
        gsrc[j] = B + (int64_t)(kb + kr) * ldb + n;
        gstep[j] = (int64_t)D_BK * ldb;
      } else {
        int r = col0 + gb * 16 + lrow;
        if (r >= N) r -= 64;
        gsrc[j] = B + (int64_t)r * ldb + kb + lchunk * 2;
        gstep[j] = D_BK;
      }
    }
  }
  auto issue = [&](int ci) {
    unsigned char* st = lds + (ci % S) * C::STAGE;
#pragma unroll
    for (int j = 0; j < C::LOADS_W; ++j)
    {
      // (a named pointer: the builtin rejects the expression in the host pass)
      const double* src = gsrc[j] + ci * gstep[j];
      __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(st + goff[j]), 16, 0, 0);
    }
  };

  const int foff = l16 * 64 + ((gq ^ d_swizzle(l16)) << 4);
  // fragment reads of chunk ci into (a, b), not waited for
  auto read = [&](int ci, double (&a)[C::FM][2], double (&b)[C::FN][2]) {
    const unsigned char* cur = lds + (ci % S) * C::STAGE;
    if (AKM) {
      const uint32_t a0 = lds_addr((const double*)cur + (wy * 64 + l16) + 2 * gq * D_BM);
      frag_mn<C::FM, D_BM>(a0, a);
    } else {
      const uint32_t a0 = lds_addr(cur + wy * 64 * 64 + foff);
      frag_nk<C::FM>(a0, a);
    }
    if (BKN) {
      const uint32_t b0 = lds_addr((const double*)(cur + C::PA) + (wx * C::WN + l16) + 2 * gq * BN);
      frag_mn<C::FN, BN>(b0, b);
    } else {
      const uint32_t b0 = lds_addr(cur + C::PA + wx * C::WN * 64 + foff);
      frag_nk<C::FN>(b0, b);
    }
  };
  auto wait_ties = [&](double (&a)[C::FM][2], double (&b)[C::FN][2]) {
    lds_frag_wait();
#pragma unroll
    for (int f = 0; f < C::FM; ++f) {
      lds_tie(a[f][0]);
      lds_tie(a[f][1]);
    }
#pragma unroll
    for (int f = 0; f < C::FN; ++f) {
      lds_tie(b[f][0]);
      lds_tie(b[f][1]);
    }
  };
  // every wave's loads of chunk ci landed, and every wave finished the reads
  // it issued before: a wait for this wave's own loads (all but `later`
  // chunks), then a bare s_barrier (the fence of __syncthreads would wait for
  // the loads still in flight).  sched_barrier(0): the compiler moves nothing
  // across, in particular no LDS read above the wait and the barrier (another
  // wave's direct loads fill part of the stage, and only the barrier orders
  // them).
  auto arrive = [&](int later) {
    if (later >= 2)
      d_wait_vmcnt<2 * C::LOADS_W>();
    else if (later == 1)
      d_wait_vmcnt<C::LOADS_W>();
    else
      d_wait_vmcnt<0>();
    __builtin_amdgcn_sched_barrier(0);
    __asm__ __volatile__("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ __volatile__("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
  };
  static_assert(S == 4, "the vmcnt levels of arrive() assume 4 stages");

  auto mfma = [&](double (&a)[C::FM][2], double (&b)[C::FN][2]) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < C::FN; ++ni)
          acc[mi][ni] =
              __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi][s], b[ni][s], acc[mi][ni], 0, 0, 0);
  };

#pragma unroll
  for (int ci = 0; ci < S - 1; ++ci)
    if (ci < nk) issue(ci);
  // Chunk it's fragments are read one iteration ahead, so its MFMAs overlap
  // the reads of chunk it + 1.  Chunks in flight: it + 2 .. it + S (the
  // issue at iteration it reuses chunk it's stage, whose reads finished at
  // it - 1).
  double af[C::FM][2], bf[C::FN][2], an[C::FM][2], bn[C::FN][2];
  if (nk > 0) {
    arrive(min(S - 2, nk - 1));
    if (S - 1 < nk) issue(S - 1);
    read(0, af, bf);
    wait_ties(af, bf);
  }
  for (int it = 0; it < nk; ++it) {
    const bool more = it + 1 < nk;
    if (more) {
      arrive(min(S - 2, nk - 2 - it));
      if (it + S < nk) issue(it + S);
      read(it + 1, an, bn);
    }
    __builtin_amdgcn_sched_barrier(0);
    mfma(af, bf);
    __builtin_amdgcn_sched_barrier(0);
    if (more) {
      wait_ties(an, bn);
#pragma unroll
      for (int f = 0; f < C::FM; ++f) {
        af[f][0] = an[f][0];
        af[f][1] = an[f][1];
      }
#pragma unroll
      for (int f = 0; f < C::FN; ++f) {
        bf[f][0] = bn[f][0];
        bf[f][1] = bn[f][1];
      }
    }
  }
  __syncthreads();
}

bool gemm_d_supported(const GemmArgs& g);
// Returns BTA_OK after launching, a negative value when the shape is left to
// the generic kernel, or an error code.
int gemm_d(const GemmArgs& g, int batch, hipStream_t s);

}  // namespace bta
