// Blocked FP64 Cholesky with psd_safe_cholesky's jitter schedule, the
// triangular inverse, and a batched small Cholesky for q x q / n_b x n_b
// blocks.
//
// bta_potrf is right-looking with 64-column block columns, one launch per
// block column j (potrf_step_kernel), the update of column j-1 fused with the
// factorisation of column j:
//   strip : one workgroup per row block b >= j applies column j-1's update to
//           the diagonal block and to block b, factors the diagonal block in
//           LDS (FP64 MFMA on 16-column panels) and solves L_bj;
//   bulk  : the other workgroups apply column j-1's rank-64 update to the
//           trailing lower 128 x 64 tiles right of column j (direct-to-LDS
//           FP64 MFMA main loop of the GEMM).
// Early steps are bound by the bulk update, late steps by the latency of the
// diagonal factorisation.  trinv_diag_kernel then inverts every 64 x 64
// diagonal block into Linv.
//
// bta_potrf_inverse (POTRF_INVERT) builds all of Linv = L^-1 in the same
// launches: the non-strip workgroups first run the step's side tasks (partial
// products of the row-block recurrence of L^-1, accumulated in place in Linv,
// and the finishing products by the inverted diagonal blocks), placed by the
// host plan of inv_plan.h so that every operand is final before its launch,
// then the trailing tiles; two more launches carry the last rows' tasks.
//
// bta_trtri_lower inverts L bottom-up by recursive doubling: with the 64x64
// diagonal inverses in place, each level joins pairs of inverted diagonal
// blocks [A 0; C B] -> [A^-1 0; -B^-1 C A^-1 B^-1] with two MFMA GEMMs batched
// over all pairs of the level, so the whole inverse is ~4 log2(n/64) launches.
#include "bta_common.h"
#include "gemm.h"
#include "gemm_d.h"
#include "inv_plan.h"
#include "kernels.h"

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <vector>

namespace bta {

constexpr int LEAF = 64;
constexpr int LP = LEAF + 1;  // LDS pitch (odd: column walks hit distinct banks)

// Factor the n x n (n <= 64) matrix held in LDS (pitch LP) in place; returns
// -1 on success or the failing column; *nan_seen set when the pivot is NaN.
// All 256 threads must call it.
__device__ int chol_lds(double* Ls, int n, int* nan_seen) {
  __shared__ int fail;
  if (threadIdx.x == 0) fail = -1;
  const int t = threadIdx.x;
  for (int c = 0; c < n; ++c) {
    __syncthreads();
    const double p = Ls[c * LP + c];
    if (!(p > 0.0)) {  // uniform: every thread read the same pivot
      if (t == 0) {
        fail = c;
        if (p != p) *nan_seen = 1;
      }
      break;
    }
    const double dg = sqrt(p);
    if (t > c && t < n) Ls[t * LP + c] /= dg;
    __syncthreads();
    // every thread has read the pivot by now; nothing below reads L[c][c]
    if (t == c) Ls[c * LP + c] = dg;
    const int i = c + 1 + (t >> 2);
    if (i < n) {
      const double lic = Ls[i * LP + c];
      for (int j = c + 1 + (t & 3); j <= i; j += 4) Ls[i * LP + j] -= lic * Ls[j * LP + c];
    }
  }
  __syncthreads();
  return fail;
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const int2 p = __builtin_bit_cast(int2, v);
  int2 r;
  r.x = __builtin_amdgcn_readlane(p.x, l);
  r.y = __builtin_amdgcn_readlane(p.y, l);
  return __builtin_bit_cast(double, r);
}

// v from the lane given by the quad permutation CTRL (DPP, no LDS traffic)
template <int CTRL>
__device__ __forceinline__ double quad_perm_d(double v) {
  const int2 p = __builtin_bit_cast(int2, v);
  int2 r;
  r.x = __builtin_amdgcn_update_dpp(0, p.x, CTRL, 0xF, 0xF, false);
  r.y = __builtin_amdgcn_update_dpp(0, p.y, CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, r);
}
constexpr int QP_XOR1 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int QP_XOR2 = 0x4E;  // quad_perm [2, 3, 0, 1]

typedef __attribute__((address_space(3))) double* lds_dptr;

// Waits for this wave's outstanding LDS operations (a cheap guard where lanes
// of one wave exchange data through LDS outside the hot loops).
__device__ __forceinline__ void lds_wait() { __asm__ __volatile__("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// One 16 x 16 x 16 FP64 MFMA product on LDS operands X (pitch PX), Y (pitch
// PY): acc += sgn * X[xr + i][xk + k] B[k][n] with B[k][n] = Y[yr + n][yk + k]
// (BT, a product with Y^T) or Y[yk + k][yr + n] (!BT).  Accumulator layout:
// acc[q] = C[g + 4q][l16].
template <bool BT, int PX = LP, int PY = LP>
__device__ __forceinline__ void mm16(const double* X, int xr, int xk, const double* Y, int yr,
                                     int yk, double sgn, v4d& acc) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const double av = sgn * X[(xr + l16) * PX + xk + 4 * s + g];
    const double bv = BT ? Y[(yr + l16) * PY + yk + 4 * s + g] : Y[(yk + 4 * s + g) * PY + yr + l16];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
}

template <int P = LP>
__device__ __forceinline__ v4d ld16(const double* X, int r0, int c0) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
  v4d v;
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = X[(r0 + g + 4 * q) * P + c0 + l16];
  return v;
}

template <int P = LP>
__device__ __forceinline__ void st16(double* X, int r0, int c0, const v4d& v) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) X[(r0 + g + 4 * q) * P + c0 + l16] = v[q];
}

// 64 x 64 global block -> LDS (pitch LP): all 16 loads of a thread are issued
// before its LDS stores.
__device__ __forceinline__ void stage64(const double* __restrict__ src, int64_t ld, double* dst) {
  const int tid = threadIdx.x;
  double v[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int e = tid + 256 * t;
    v[t] = src[(int64_t)(e >> 6) * ld + (e & 63)];
  }
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int e = tid + 256 * t;
    dst[(e >> 6) * LP + (e & 63)] = v[t];
  }
}

// Two 64 x 64 blocks at once (the second only when `two`): all loads, then
// all LDS stores.
__device__ __forceinline__ void stage64x2(const double* __restrict__ s1, double* d1,
                                          const double* __restrict__ s2, double* d2,
                                          int64_t ld, bool two) {
  const int tid = threadIdx.x;
  double v[16], w[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int e = tid + 256 * t;
    v[t] = s1[(int64_t)(e >> 6) * ld + (e & 63)];
    if (two) w[t] = s2[(int64_t)(e >> 6) * ld + (e & 63)];
  }
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int e = tid + 256 * t;
    d1[(e >> 6) * LP + (e & 63)] = v[t];
    if (two) d2[(e >> 6) * LP + (e & 63)] = w[t];
  }
}

// Work the other three waves do while wave 0 factors a panel.
struct NoBackground {
  __device__ void operator()(int) const {}
};

// Cholesky of the 64 x 64 block in S (pitch LP), right-looking over four
// 16-column panels.  Wave 0 factors a panel (all 64 rows, lane i = row i, the
// panel's 16 entries in registers): the pivot of column c comes from lane c
// by readlane, 1/sqrt(pivot) from v_rsq_f64 with one third-order correction
// (sqrt as p * rsqrt with one correction, off the chain).  The next column's
// multiplier also goes by readlane, which keeps the pivot chain short; the
// later columns' come from the column's LDS copy colv in one batch of
// broadcast reads, applied during the next column's chain so their latency is
// hidden.  Meanwhile waves 1-3 run bg(p).  Then the four waves apply the panel to the trailing
// 16 x 16 blocks with FP64 MFMA (rank 16 per block).  No branches on the
// data: the first non-positive or NaN pivot is recorded (what follows it is
// meaningless and the caller discards it).  Leaves L
// in S (upper part zeroed) and 1 / L_ii in rdiag; *fail_s / *nan_s (LDS) get
// the first failing column (or -1) and whether its pivot was NaN.
template <typename BG = NoBackground>
__device__ __forceinline__ void chol64_blocked(double* S, double* rdiag, double* colv,
                                               int* fail_s, int* nan_s, BG bg = BG()) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int fail = -1, nanf = 0;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    if (wv == 0) {
      double a[16], cv[16];
#pragma unroll
      for (int m = 0; m < 16; ++m) a[m] = S[lane * LP + 16 * p + m];
      // colv through an opaque LDS base: the broadcast reads then share one
      // address register and differ in their immediate offsets
      uint32_t cb = (uint32_t)(uintptr_t)(lds_dptr)(colv + 16 * p);
      __asm__ __volatile__("" : "+v"(cb));
      const lds_dptr cvb = (lds_dptr)(uintptr_t)cb;
      double lprev = 0.0, pvl = 1.0;
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        const int cc = 16 * p + c;
        // pivot chain: pv -> y = 1/sqrt(pv) -> l -> next column's pivot
        const double pv = readlane_d(a[c], cc);
        pvl = lane == cc ? pv : pvl;  // the lane's own pivot, for L_ii and the checks
        // v_rsq_f64 refined by one third-order step: with e = 1 - pv y0^2,
        // 1/sqrt(pv) = y0 (1 + e/2 + 3e^2/8 + O(e^3))
        const double y0 = __builtin_amdgcn_rsq(pv);
        const double e = fma(-pv * y0, y0, 1.0);
        const double y = fma(y0 * e, fma(e, 0.375, 0.5), y0);
        const double l = a[c] * y;
        if (c + 1 < 16) {
          // column c - 1's broadcast (read during the chain) for column c + 1
          if (c >= 1) a[c + 1] = fma(-lprev, cv[c + 1], a[c + 1]);
          a[c + 1] = fma(-l, readlane_d(l, cc + 1), a[c + 1]);
          // and for the columns after it
          if (c >= 1) {
#pragma unroll
            for (int m = c + 2; m < 16; ++m) a[m] = fma(-lprev, cv[m], a[m]);
          }
          if (c + 2 < 16) {
            // a wave's LDS accesses are performed in issue order: the reads
            // see this column's colv and the previous column's reads are done
            colv[lane] = l;
            __asm__ __volatile__("" ::: "memory");
#pragma unroll
            for (int m = (c + 2) & ~1; m < 16; m += 2) {
              const v2d v = *(const __attribute__((address_space(3))) v2d*)(cvb + m);
              cv[m] = v.x;
              cv[m + 1] = v.y;
            }
          }
        }
        lprev = l;
        a[c] = l;  // rows above the diagonal are garbage, masked when stored
      }
#pragma unroll
      for (int m = 0; m < 16; ++m) S[lane * LP + 16 * p + m] = 16 * p + m < lane ? a[m] : 0.0;
      // L_ii = sqrt(pivot_i) = pv y, rounded by one correction; the first
      // non-positive or NaN pivot of the panel is the failing column
      const bool mine = lane >= 16 * p && lane < 16 * p + 16;
      const double yl = __builtin_amdgcn_rsq(pvl);
      double dg = pvl * fma(yl * fma(-pvl * yl, yl, 1.0), 0.5, yl);
      dg = fma(fma(-dg, dg, pvl), 0.5 * yl, dg);
      if (mine) S[lane * LP + lane] = dg;
      const uint64_t badm = __ballot(mine && !(pvl > 0.0));
      if (fail < 0 && badm) {
        fail = __builtin_ctzll(badm);
        nanf = __builtin_amdgcn_readlane((int)(pvl != pvl), fail);
      }
    } else {
      bg(p);
    }
    __syncthreads();
    if (p < 3) {
      // trailing blocks (ib, jb), p < jb <= ib, dealt to the waves
      const int nt = 4 - p - 1;  // block rows below the panel
      for (int bi = wv; bi < nt * (nt + 1) / 2; bi += 4) {
        int ib = 0;
        while ((ib + 1) * (ib + 2) / 2 <= bi) ++ib;
        const int jb = bi - ib * (ib + 1) / 2;
        const int r0 = 16 * (p + 1 + ib), c0 = 16 * (p + 1 + jb);
        v4d acc = ld16(S, r0, c0);
        mm16<true>(S, r0, 16 * p, S, c0, 16 * p, -1.0, acc);
        st16(S, r0, c0, acc);
      }
      __syncthreads();
    }
  }
  if (wv == 0) {
    rdiag[lane] = 1.0 / S[lane * LP + lane];
    if (lane == 0) {
      *fail_s = __builtin_amdgcn_readfirstlane(fail);
      *nan_s = __builtin_amdgcn_readfirstlane(nanf);
    }
  }
  __syncthreads();
}

// V = L_bb^-1 for diagonal 16-block b of L (S, pitch LP), by the calling
// wave: forward substitution with four lanes per column splitting each row's
// dot product by k mod 4 and combining it by two DPP quad exchanges.  V goes to
// V[(16 b + i) * PV + (DIAG ? 16 b : 0) + j].
template <int PV, bool DIAG>
__device__ __forceinline__ void inv16_block(const double* S, int blk, double* V) {
  const int lane = threadIdx.x & 63;
  const int jl = lane >> 2, r = lane & 3, b0 = 16 * blk;
  // this lane's operands L[b0 + i][b0 + 4t + r] (k < i), loaded up front
  double lv[16][4];
#pragma unroll
  for (int i = 0; i < 16; ++i)
#pragma unroll
    for (int t = 0; 4 * t < i; ++t) {
      const int k = 4 * t + r;
      lv[i][t] = k < i ? S[(b0 + i) * LP + b0 + k] : 0.0;
    }
  double rd[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) rd[i] = 1.0 / S[(b0 + i) * LP + b0 + i];
  double x[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    double pp = 0.0;
#pragma unroll
    for (int t = 0; 4 * t < i; ++t) pp = fma(lv[i][t], x[t], pp);
    pp += quad_perm_d<QP_XOR1>(pp);
    pp += quad_perm_d<QP_XOR2>(pp);
    const double xi = ((i == jl ? 1.0 : 0.0) - pp) * rd[i];
    if (r == (i & 3)) x[i >> 2] = xi;
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) V[(b0 + 4 * t + r) * PV + (DIAG ? b0 : 0) + jl] = x[t];
}

// the four diagonal 16-blocks, one per wave
template <int PV, bool DIAG>
__device__ __forceinline__ void inv16_diag(const double* S, double* V) {
  inv16_block<PV, DIAG>(S, threadIdx.x >> 6, V);
}

// W = L^-1 (pitch LP) for the lower-triangular L in S, with rdiag = 1 / L_ii:
// the diagonal 16-blocks by inv16_diag, the off-diagonal blocks by recursive
// doubling on FP64 MFMA:  W10 = -V1 L10 V0,  W32 = -V3 L32 V2, then
// W[23][01] = -W[23][23] (L[23][01] W[01][01]).  All 256 threads call it.
__device__ __forceinline__ void trinv64_blocked(const double* S, const double* rdiag, double* Wl) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  inv16_diag<LP, true>(S, Wl);
  {
    // zero this wave's block row right of its diagonal block
    const int b0 = 16 * wv;
    for (int e = lane; e < 16 * 48; e += 64) {
      const int rr = e / 48, cc = e % 48;
      const int col = b0 + 16 + cc;
      if (col < LEAF) Wl[(b0 + rr) * LP + col] = 0.0;
    }
  }
  __syncthreads();
  if (wv < 2) {  // W10 (wave 0), W32 (wave 1)
    const int i = 2 * wv + 1, j = 2 * wv;
    v4d t = (v4d)(0.0);
    mm16<false>(S, 16 * i, 16 * j, Wl, 16 * j, 16 * j, 1.0, t);  // L_ij V_j
    st16(Wl, 16 * i, 16 * j, t);
    lds_wait();
    v4d w = (v4d)(0.0);
    mm16<false>(Wl, 16 * i, 16 * i, Wl, 16 * j, 16 * i, -1.0, w);  // -V_i T
    st16(Wl, 16 * i, 16 * j, w);
  }
  __syncthreads();
  // T2_ij = sum_k L_ik W_kj, i in {2,3}, j, k in {0,1}: wave -> (i, j)
  const int i2 = 2 + (wv >> 1), j2 = wv & 1;
  v4d t2 = (v4d)(0.0);
  mm16<false>(S, 16 * i2, 16, Wl, 16 * j2, 16, 1.0, t2);            // L_i1 W_1j
  if (j2 == 0) mm16<false>(S, 16 * i2, 0, Wl, 0, 0, 1.0, t2);        // L_i0 W_00
  st16(Wl, 16 * i2, 16 * j2, t2);
  __syncthreads();
  // W_ij = -sum_{k in {2,3}, k <= i} W_ik T2_kj
  v4d w2 = (v4d)(0.0);
  mm16<false>(Wl, 16 * i2, 32, Wl, 16 * j2, 32, -1.0, w2);          // W_i2 T2_2j
  if (i2 == 3) mm16<false>(Wl, 48, 48, Wl, 16 * j2, 48, -1.0, w2);   // W_33 T2_3j
  __syncthreads();
  st16(Wl, 16 * i2, 16 * j2, w2);
  __syncthreads();
}

// One step of the blocked factorisation, as one launch.  Workgroups
// [0, nstrip) are the strip of block column j: workgroup s owns row block
// b = j + s.  Each applies column j-1's update to the diagonal block (all of
// them, redundantly, so none waits for another) and to its own block, and
// factors the diagonal block (chol64_blocked).  Workgroup 0 (b = j) writes
// L_jj; the others solve their block, L_bj L_jj^T = A_bj - L_b,j-1 L_j,j-1^T,
// by blocked forward substitution: wave w takes rows [16w, 16w + 16) and the
// four 16-column blocks in turn, X_k = (T_k - sum_{m<k} X_m L_km^T) V_k^T
// with V_k = L_kk^-1 (inv16_block), so no wave waits for another.  T is
// formed before the factorisation; waves 1-3 invert V_0..V_2 while wave 0
// factors the panels.  The
// remaining workgroups apply column j-1's update to the trailing lower tiles
// right of column j (the strip covers column j): A_bc -= L_b,j-1 L_c,j-1^T on
// 128 x 64 FP64 MFMA tiles.
struct StepArgs {
  double* A;
  int64_t lda;
  int* info;
  double* Linv;  // when set: the diagonal workgroup also writes L_jj^-1 into
  int64_t ldinv;  // Linv's diagonal block and zeroes the block right of it
  int j, NB, nstrip;
  const int4* side;  // this step's inverse side tasks (POTRF_INVERT): nside of
  int nside, npart;  // them, the first npart partial tasks, then finish tasks
};

constexpr int VP = 17;  // pitch of the 16 x 16 diagonal inverses (64 x 16 array)
constexpr int STEP_LDS_STRIP = (2 * LEAF * LP + LEAF * VP + 2 * LEAF) * 8 + 16;
constexpr int STEP_LDS = STEP_LDS_STRIP > DCfg<128>::STAGES * DCfg<128>::STAGE
                             ? STEP_LDS_STRIP
                             : DCfg<128>::STAGES * DCfg<128>::STAGE;

__device__ __forceinline__ void potrf_strip(const StepArgs& a, unsigned char* smem) {
  double* S = (double*)smem;     // L_b,j-1 staging, then the diagonal block -> L_jj
  double* Bf = S + LEAF * LP;    // L_j,j-1 staging, then this block's T -> X
  double* Vs = Bf + LEAF * LP;   // V_k, k = 0..3 (rows 16k.., pitch VP)
  double* rdiag = Vs + LEAF * VP;
  double* colv = rdiag + LEAF;
  int* flags = (int*)(colv + LEAF);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, l16 = lane & 15, g = lane >> 4;
  const int j = a.j, b = j + blockIdx.x;
  const bool own = b != j;  // workgroup-uniform
  const int64_t lda = a.lda;
  double* Ajj = a.A + (int64_t)j * LEAF * lda + (int64_t)j * LEAF;
  double* Abj = a.A + (int64_t)b * LEAF * lda + (int64_t)j * LEAF;
  const double* Q = Abj - LEAF;  // L_b,j-1 (j > 0)
  // D blocks (ib, jb), jb <= ib (10 of them), dealt round-robin over the waves
  v4d dacc[3];
  int dib[3], djb[3];
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int bi = wv + 4 * u;
    int ib = 0;
    while ((ib + 1) * (ib + 2) / 2 <= bi) ++ib;
    dib[u] = ib;
    djb[u] = bi - ib * (ib + 1) / 2;
    if (bi < 10)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        dacc[u][q] = Ajj[(int64_t)(16 * dib[u] + g + 4 * q) * lda + 16 * djb[u] + l16];
  }
  // T = A_bj - L_b,j-1 L_j,j-1^T: wave wv holds column block wv, row blocks
  // u = 0..3, formed before the factorisation (L_b,j-1 staged in S, which
  // takes D afterwards)
  v4d tacc[4];
  if (own)
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        tacc[u][q] = Abj[(int64_t)(16 * u + g + 4 * q) * lda + 16 * wv + l16];
  if (j > 0) {
    stage64x2(Ajj - LEAF, Bf, Q, S, lda, own);  // L_j,j-1 -> Bf, L_b,j-1 -> S
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 3; ++u)
      if (wv + 4 * u < 10)
#pragma unroll
        for (int s = 0; s < 16; ++s)
          dacc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              -Bf[(16 * dib[u] + l16) * LP + 4 * s + g], Bf[(16 * djb[u] + l16) * LP + 4 * s + g],
              dacc[u], 0, 0, 0);
    if (own)
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int s = 0; s < 16; ++s)
          tacc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              -S[(16 * u + l16) * LP + 4 * s + g], Bf[(16 * wv + l16) * LP + 4 * s + g], tacc[u],
              0, 0, 0);
    __syncthreads();  // S (L_b,j-1) read by every wave
  }
#pragma unroll
  for (int u = 0; u < 3; ++u)
    if (wv + 4 * u < 10) st16(S, 16 * dib[u], 16 * djb[u], dacc[u]);
  __syncthreads();
  // while wave 0 factors panel p, wave p (p >= 1) inverts the diagonal block
  // p - 1 finished before
  auto bg = [&](int p) {
    if (own && p >= 1 && wv == p) inv16_block<VP, false>(S, p - 1, Vs);
  };
  chol64_blocked(S, rdiag, colv, flags, flags + 1, bg);
  if (flags[0] >= 0) {
    if (!own && tid == 0) {
      a.info[0] = j * LEAF + flags[0] + 1;
      a.info[1] = flags[1];
    }
    return;
  }
  if (!own) {
    for (int e = tid; e < LEAF * LEAF; e += 256) {
      const int r = e >> 6, c = e & 63;
      if (c <= r) Ajj[(int64_t)r * lda + c] = S[r * LP + c];
    }
    if (a.Linv) {
      // the diagonal workgroup finishes before the others: invert here (Bf,
      // the staged L_j,j-1, is no longer needed)
      trinv64_blocked(S, rdiag, Bf);
      double* W = a.Linv + (int64_t)j * LEAF * a.ldinv + (int64_t)j * LEAF;
      const bool zr = j + 1 < a.NB;
      for (int e = tid; e < LEAF * LEAF; e += 256) {
        const int r = e >> 6, c = e & 63;
        W[(int64_t)r * a.ldinv + c] = Bf[r * LP + c];
        if (zr) W[(int64_t)r * a.ldinv + LEAF + c] = 0.0;
      }
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) st16(Bf, 16 * u, 16 * wv, tacc[u]);
  if (wv == 0) inv16_block<VP, false>(S, 3, Vs);
  __syncthreads();
  const int r0 = 16 * wv;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v4d acc = ld16(Bf, r0, 16 * k);
#pragma unroll
    for (int m = 0; m < k; ++m) mm16<true>(Bf, r0, 16 * m, S, 16 * k, 16 * m, -1.0, acc);
    st16(Bf, r0, 16 * k, acc);
    lds_wait();
    v4d x = (v4d)(0.0);
    mm16<true, LP, VP>(Bf, r0, 16 * k, Vs, 16 * k, 0, 1.0, x);
    st16(Bf, r0, 16 * k, x);
    lds_wait();
  }
#pragma unroll
  for (int t = 0; t < 16; ++t) Abj[(int64_t)(r0 + t) * lda + lane] = Bf[(r0 + t) * LP + lane];
}

// Rank-64 update of one trailing lower 128 x 64 tile (tile row tm holds the
// 2 tm + 2 tiles tn <= 2 tm + 1; rem = rows of the trailing matrix).  The
// wave's part of C is loaded before the products: its latency overlaps the
// first chunks of the main loop instead of following it.
__device__ __forceinline__ void potrf_bulk_tile(const StepArgs& a, int t, int rem,
                                                unsigned char* smem) {
  const int j = a.j;
  const int64_t lda = a.lda;
  const double* P = a.A + (int64_t)(j + 1) * LEAF * lda + (int64_t)(j - 1) * LEAF;
  double* Cm = a.A + (int64_t)(j + 1) * LEAF * lda + (int64_t)(j + 1) * LEAF;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wy = wv >> 1, wx = wv & 1, l16 = lane & 15, gq = lane >> 4;
  int tm = 0;
  while ((tm + 1) * (tm + 2) <= t) ++tm;  // tile rows before tm hold tm (tm + 1) tiles
  const int tn = t - tm * (tm + 1);
  double cv[4][2][4];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = tm * D_BM + wy * 64 + mi * 16 + Mfma<double>::row(gq, q);
#pragma unroll
      for (int ni = 0; ni < 2; ++ni) {
        const int col = tn * 64 + wx * 32 + ni * 16 + l16;
        cv[mi][ni][q] = (row < rem && col <= row) ? Cm[(int64_t)row * lda + col] : 0.0;
      }
    }
  v4d acc[4][2];
  d_mainloop<64, false, false>(P, lda, P, lda, tm * D_BM, tn * 64, rem, rem, 0, LEAF, smem, acc);
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = tm * D_BM + wy * 64 + mi * 16 + Mfma<double>::row(gq, q);
#pragma unroll
      for (int ni = 0; ni < 2; ++ni) {
        const int col = tn * 64 + wx * 32 + ni * 16 + l16;
        if (row < rem && col <= row) Cm[(int64_t)row * lda + col] = cv[mi][ni][q] - acc[mi][ni][q];
      }
    }
}

// Inverse side tasks (POTRF_INVERT).  X = L^-1 is built row block by row
// block alongside the factorisation:
//   X_rc = -X_rr S_rc,  S_rc = sum_{k=c}^{r-1} L_rk X_kc   (c < r)
// with X_rr = L_rr^-1 written by step r's diagonal workgroup.  The sums of the
// row pair (2q, 2q+1) over k < 2q - 1 are accumulated in place in X's own
// storage, in K chunks of 128 x 64 tiles (partial tasks) that a host-built
// plan places in the steps whose launches start after the chunk's operands
// are final: L_rk final after step k, X row block k after step k + 2.  The
// finish task of row r (step r + 2) adds the remaining terms k in
// [max(c, 2q - 1), r) and multiplies by -X_rr.  Every read and write of a
// task is ordered against its producers by launch order alone.
//
// Partial task: x = 0 | first << 4, y = q, z = c, w = k0 | k1 << 16 (leaves).
__device__ __forceinline__ void inv_partial(const StepArgs& a, int4 t, unsigned char* smem) {
  const int q = t.y, c = t.z, k0 = t.w & 0xffff, k1 = t.w >> 16;
  const bool first = (t.x >> 4) & 1;
  const int np = a.NB * LEAF;
  double* X = a.Linv;
  const int64_t ldx = a.ldinv;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wy = wv >> 1, wx = wv & 1, l16 = lane & 15, gq = lane >> 4;
  const int row0 = q * D_BM, col0 = c * LEAF;
  double cv[4][2][4];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int row = row0 + wy * 64 + mi * 16 + Mfma<double>::row(gq, qq);
#pragma unroll
      for (int ni = 0; ni < 2; ++ni) {
        const int col = col0 + wx * 32 + ni * 16 + l16;
        cv[mi][ni][qq] = (!first && row < np) ? X[(int64_t)row * ldx + col] : 0.0;
      }
    }
  v4d acc[4][2];
  d_mainloop<64, false, true>(a.A, a.lda, X, ldx, row0, col0, np, np, k0 * LEAF, k1 * LEAF, smem,
                              acc);
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int row = row0 + wy * 64 + mi * 16 + Mfma<double>::row(gq, qq);
#pragma unroll
      for (int ni = 0; ni < 2; ++ni) {
        const int col = col0 + wx * 32 + ni * 16 + l16;
        if (row < np) X[(int64_t)row * ldx + col] = cv[mi][ni][qq] + acc[mi][ni][qq];
      }
    }
}

// Finish task of X_rc: x = 1, y = r, z = c.  Wave w holds rows [16w, 16w + 16)
// of the 64 x 64 block, its four 16-column blocks in turn.
__device__ __forceinline__ void inv_finish(const StepArgs& a, int4 t, unsigned char* smem) {
  const int r = t.y, c = t.z, q = r >> 1;
  const bool accumulated = q >= 1 && c <= 2 * q - 2;  // partial tasks wrote S_rc's first terms
  double* X = a.Linv;
  const int64_t ldx = a.ldinv, lda = a.lda;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, l16 = lane & 15, g = lane >> 4;
  double* La = (double*)smem;  // L_rk, then X_rr
  double* Xb = La + LEAF * LP;  // X_kc, then S_rc
  double* Xrc = X + (int64_t)r * LEAF * ldx + (int64_t)c * LEAF;
  v4d acc[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int qq = 0; qq < 4; ++qq)
      acc[cb][qq] = accumulated ? Xrc[(int64_t)(16 * wv + g + 4 * qq) * ldx + 16 * cb + l16] : 0.0;
  // acc[cb] += sgn * La[16 wv.., 16 s..] Xb[16 s.., 16 cb..] for s < s_end, one
  // 16-column slab of k at a time (the four accumulators share the A operand)
  auto rows_times = [&](int s_end, double sgn, v4d(&o)[4]) {
#pragma unroll 1
    for (int s = 0; s < s_end; ++s)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int k = 16 * s + 4 * kk + g;
        const double av = sgn * La[(16 * wv + l16) * LP + k];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
          o[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Xb[k * LP + 16 * cb + l16], o[cb], 0, 0,
                                                      0);
      }
  };
  for (int k = max(c, 2 * q - 1); k < r; ++k) {
    stage64(a.A + (int64_t)r * LEAF * lda + (int64_t)k * LEAF, lda, La);
    stage64(X + (int64_t)k * LEAF * ldx + (int64_t)c * LEAF, ldx, Xb);
    __syncthreads();
    rows_times(4, 1.0, acc);
    __syncthreads();
  }
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) st16(Xb, 16 * wv, 16 * cb, acc[cb]);
  stage64(X + (int64_t)r * LEAF * ldx + (int64_t)r * LEAF, ldx, La);
  __syncthreads();
  v4d o[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) o[cb] = (v4d)(0.0);
  rows_times(wv + 1, -1.0, o);  // X_rr is lower triangular: its blocks s > wv are zero
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int qq = 0; qq < 4; ++qq)
      Xrc[(int64_t)(16 * wv + g + 4 * qq) * ldx + 16 * cb + l16] = o[cb][qq];
  __syncthreads();  // the next task may overwrite La / Xb
}

// The non-strip workgroups of a step: the step's side tasks (when the plan
// has any) and then the trailing update's tiles, dealt in a plain stride.
__device__ __forceinline__ void potrf_work(const StepArgs& a, unsigned char* smem) {
  const int j = a.j;
  int rem = 0, ntile = 0;
  if (j > 0 && a.NB - j - 1 > 0) {
    rem = (a.NB - j - 1) * LEAF;
    const int MT = (rem + D_BM - 1) / D_BM;
    ntile = MT * (MT + 1);
  }
  const int total = a.nside + ntile;
  const int W = gridDim.x - a.nstrip, w = blockIdx.x - a.nstrip;
  // Worker w takes task k W + w in even rounds k and k W + (W - 1 - w) in odd
  // ones, so the workers that drew the last (lightest) tasks of a round draw
  // the first of the next.  Tasks run in order: the partial tasks (the plan's
  // first npart, longest first), the finish tasks, then the tiles, each kind in
  // a loop of its own (a worker's task indices increase).
  int k = 0;
  auto task = [&](int kk) { return kk * W + ((kk & 1) ? W - 1 - w : w); };
  int t = task(k);
  for (; t < a.npart; t = task(++k)) inv_partial(a, a.side[t], smem);
  for (; t < a.nside; t = task(++k)) inv_finish(a, a.side[t], smem);
  for (; t < total; t = task(++k)) potrf_bulk_tile(a, t - a.nside, rem, smem);
}

__global__ __launch_bounds__(256, 2) void potrf_step_kernel(StepArgs a) {
  __shared__ __attribute__((aligned(1024))) unsigned char smem[STEP_LDS];
  __shared__ int stop;
  if (threadIdx.x == 0) stop = a.info[0];
  __syncthreads();
  if (stop != 0) return;  // an earlier step failed
  if ((int)blockIdx.x < a.nstrip)
    potrf_strip(a, smem);
  else
    potrf_work(a, smem);
}

// W_jj = L_jj^-1 for every 64 x 64 diagonal block of L (one workgroup per
// block), into Linv's diagonal blocks; the rest of each block row (the upper
// triangle of L^-1) is zeroed.
__global__ __launch_bounds__(256) void trinv_diag_kernel(const double* L, int64_t ldl,
                                                         double* Linv, int64_t ldinv,
                                                         int64_t n_cols) {
  __shared__ double S[LEAF * LP], Wl[LEAF * LP], rdiag[LEAF];
  const int64_t o = (int64_t)blockIdx.x * LEAF;
  stage64(L + o * ldl + o, ldl, S);
  __syncthreads();
  if (threadIdx.x < LEAF) rdiag[threadIdx.x] = 1.0 / S[threadIdx.x * LP + threadIdx.x];
  __syncthreads();
  trinv64_blocked(S, rdiag, Wl);
  for (int e = threadIdx.x; e < LEAF * LEAF; e += 256) {
    const int r = e >> 6, c = e & 63;
    Linv[(o + r) * ldinv + o + c] = Wl[r * LP + c];
  }
  // the block row right of the diagonal block is zero in L^-1
  const int64_t rest = (int64_t)n_cols - o - LEAF;
  for (int r = 0; r < LEAF; ++r)
    for (int64_t c = threadIdx.x; c < rest; c += 256) Linv[(o + r) * ldinv + o + LEAF + c] = 0.0;
}

__global__ void fill_int_kernel(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void add_diag_kernel(double* A, int64_t lda, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[i * lda + i] += v;
}

// zero A[i][j] for j > i (whole matrix, n x n)
__global__ void zero_upper_kernel(double* A, int64_t lda, int64_t n) {
  int64_t i = blockIdx.y;
  for (int64_t j = i + 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x)
    A[i * lda + j] = 0.0;
}

__global__ void copy_lower_kernel(const double* src, int64_t lds_, double* dst, int64_t ldd,
                                  int64_t n) {
  int64_t i = blockIdx.y;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= i && j < n;
       j += (int64_t)gridDim.x * blockDim.x)
    dst[i * ldd + j] = src[i * lds_ + j];
}

// Batched small Cholesky with jitter escalation (psd_safe_cholesky semantics
// per batch element).  Strict upper triangle zeroed on output.
__global__ __launch_bounds__(256) void chol_small_kernel(double* A, int n, int64_t lda,
                                                         int64_t stride, int max_tries,
                                                         double jitter0, int* status) {
  __shared__ double Ls[LEAF * LP];
  __shared__ int nan_seen;
  double* Ab = A + (int64_t)blockIdx.x * stride;
  int st = 0;
  int fail = 0;
  for (int attempt = 0; attempt <= max_tries; ++attempt) {
    const double jit = attempt == 0 ? 0.0 : jitter0 * pow(10.0, (double)(attempt - 1));
    if (threadIdx.x == 0) nan_seen = 0;
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += 256) {
      int i = e / n, j = e % n;
      Ls[i * LP + j] = j <= i ? Ab[(int64_t)i * lda + j] + (i == j ? jit : 0.0) : 0.0;
    }
    fail = chol_lds(Ls, n, &nan_seen);
    if (fail < 0) {
      if (attempt > 0) st |= BTA_STATUS_JITTER;
      break;
    }
    if (nan_seen) {
      st |= BTA_STATUS_NAN;
      break;
    }
  }
  if (fail >= 0) st |= BTA_STATUS_NOT_PSD;
  for (int e = threadIdx.x; e < n * n; e += 256) {
    int i = e / n, j = e % n;
    Ab[(int64_t)i * lda + j] = (j <= i && fail < 0) ? Ls[i * LP + j] : (j <= i ? NAN : 0.0);
  }
  if (threadIdx.x == 0 && status) status[blockIdx.x] = st;
}

size_t potrf_ws_bytes(int64_t n_pad) {
  // saved input (n_pad^2) + one 64x64 inverse block + info words
  return (size_t)(n_pad * n_pad + LEAF * LEAF) * sizeof(double) + 256;
}

// Device copy of the plan for (device, NB, CU count), built and uploaded on
// first use and kept until the process ends.
struct InvPlanDev {
  const int4* task = nullptr;
  std::vector<int> off, units, npart;
};
static int inv_plan(int NB, int ncu, const InvPlanDev** out) {
  static std::mutex& mu = *new std::mutex;  // never destroyed
  static std::map<std::array<int, 3>, InvPlanDev>& cache =
      *new std::map<std::array<int, 3>, InvPlanDev>;
  int dev = 0;
  BTA_HIP(hipGetDevice(&dev));
  const std::array<int, 3> key = {dev, NB, ncu};
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it == cache.end()) {
    const invplan::Plan P = invplan::build(NB, ncu);
    if (P.off.empty()) {
      set_error("potrf: inverse plan for %d block rows missed a deadline", NB);
      return BTA_ERR_INVALID;
    }
    InvPlanDev D;
    D.off = P.off;
    D.units = P.units;
    D.npart = P.npart;
    if (!P.task.empty()) {
      static_assert(sizeof(invplan::Task) == sizeof(int4), "task layout");
      void* p = nullptr;
      BTA_HIP(hipMalloc(&p, P.task.size() * sizeof(int4)));
      const hipError_t e = hipMemcpy(p, P.task.data(), P.task.size() * sizeof(int4),
                                     hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        (void)hipFree(p);
        return hip_fail(e, "hipMemcpy(inverse plan)");
      }
      D.task = (const int4*)p;
    }
    it = cache.emplace(key, std::move(D)).first;
  }
  *out = &it->second;
  return BTA_OK;
}

// Step j of the factorisation of the NB x NB block matrix A on a stream whose
// launches reach ncu CUs, with the plan's side tasks when plan is set.
static int launch_step(double* A, int64_t lda, int* info, double* Linv, int64_t ldinv, int j,
                       int NB, int ncu, const InvPlanDev* plan, hipStream_t s) {
  StepArgs sa;
  sa.A = A;
  sa.lda = lda;
  sa.info = info;
  sa.Linv = Linv;
  sa.ldinv = ldinv;
  sa.j = j;
  sa.NB = NB;
  sa.nstrip = std::max(NB - j, 0);
  sa.side = nullptr;
  sa.nside = sa.npart = 0;
  int units = invplan::bulk_tiles(NB, j);
  if (plan) {
    sa.side = plan->task + plan->off[j];
    sa.nside = plan->off[j + 1] - plan->off[j];
    sa.npart = plan->npart[j];
    units += plan->units[j];
  }
  const int tasks = invplan::bulk_tiles(NB, j) + sa.nside;
  int work = 0;
  if (tasks > 0) {
    // The strip workgroups (dispatched first, one per CU) are the step's
    // critical path once the update is small: if at most two rounds of work
    // fit on the CUs they leave free, keep every strip workgroup alone on its
    // CU; otherwise fill the second slot of every CU.
    const int w1 = std::max(ncu - sa.nstrip, 1), w2 = std::max(2 * ncu - sa.nstrip, 1);
    work = std::min(tasks, (units + w1 - 1) / w1 <= 2 ? w1 : w2);
  }
  if (sa.nstrip + work == 0) return BTA_OK;
  hipLaunchKernelGGL(potrf_step_kernel, dim3((unsigned)(sa.nstrip + work)), dim3(256), 0, s, sa);
  BTA_LAUNCH_CHECK("potrf_step_kernel");
  return BTA_OK;
}

int potrf(double* A, int64_t n_pad, int64_t lda, double* Linv, int64_t ldinv, int max_tries,
          double jitter0, double* jitter_used, void* ws, size_t ws_bytes, hipStream_t s,
          int opts) {
  const bool save = !(opts & POTRF_NO_SAVE);
  if (!save) max_tries = 0;
  if ((opts & POTRF_ASYNC) && save) {
    set_error("potrf: POTRF_ASYNC needs POTRF_NO_SAVE");
    return BTA_ERR_INVALID;
  }
  if (n_pad <= 0 || n_pad % LEAF || lda < n_pad || (Linv && ldinv < n_pad)) {
    set_error("potrf: n_pad must be a positive multiple of 64");
    return BTA_ERR_INVALID;
  }
  if (((uintptr_t)A & 15) || (lda & 1)) {
    set_error("potrf: A must be 16-byte aligned with an even leading dimension");
    return BTA_ERR_INVALID;
  }
  if (!ws || ws_bytes < potrf_ws_bytes(n_pad)) {
    set_error("potrf: workspace too small (%zu < %zu)", ws_bytes, potrf_ws_bytes(n_pad));
    return BTA_ERR_WORKSPACE;
  }
  const bool invert = (opts & POTRF_INVERT) != 0;
  if (invert && (!Linv || ((uintptr_t)Linv & 15) || (ldinv & 1))) {
    set_error("potrf: the inverse needs Linv 16-byte aligned with an even leading dimension");
    return BTA_ERR_INVALID;
  }
  if (invert && (const char*)Linv < (const char*)(A + (n_pad - 1) * lda + n_pad) &&
      (const char*)A < (const char*)(Linv + (n_pad - 1) * ldinv + n_pad)) {
    set_error("potrf: Linv overlaps A");
    return BTA_ERR_INVALID;
  }
  double* saved = (double*)ws;
  double* wblk = saved + n_pad * n_pad;
  int* info = (int*)(wblk + LEAF * LEAF);
  if (jitter_used) *jitter_used = 0.0;
  const int NB = (int)(n_pad / LEAF);
  const int ncu = num_cus();
  const InvPlanDev* plan = nullptr;
  if (invert) {
    const int rc = inv_plan(NB, ncu, &plan);
    if (rc) return rc;
  }
  // with the inverse: the diagonal blocks and the zero blocks right of them
  // are written by the steps, the rest of X by the side tasks
  const bool band = invert || (opts & POTRF_BAND_ZERO);
  if (save) {
    dim3 g((unsigned)((n_pad + 255) / 256 > 64 ? 64 : (n_pad + 255) / 256), (unsigned)n_pad);
    hipLaunchKernelGGL(copy_lower_kernel, g, dim3(256), 0, s, A, lda, saved, n_pad, n_pad);
    BTA_LAUNCH_CHECK("copy_lower_kernel");
  }
  for (int attempt = 0; attempt <= max_tries; ++attempt) {
    double jit = 0.0;
    if (attempt > 0) {
      jit = jitter0 * pow(10.0, (double)(attempt - 1));
      dim3 g((unsigned)((n_pad + 255) / 256 > 64 ? 64 : (n_pad + 255) / 256), (unsigned)n_pad);
      hipLaunchKernelGGL(copy_lower_kernel, g, dim3(256), 0, s, saved, n_pad, A, lda, n_pad);
      BTA_LAUNCH_CHECK("copy_lower_kernel");
      hipLaunchKernelGGL(add_diag_kernel, dim3((unsigned)((n_pad + 255) / 256)), dim3(256), 0,
                         s, A, lda, n_pad, jit);
      BTA_LAUNCH_CHECK("add_diag_kernel");
    }
    if (attempt > 0 || !(opts & POTRF_INFO_ZEROED)) {
      hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, info, 2, 0);
      BTA_LAUNCH_CHECK("fill_int_kernel");
    }
    for (int j = 0; j < (invert ? NB + 2 : NB); ++j) {
      const int rc = launch_step(A, lda, info, band ? Linv : nullptr, ldinv, j, NB, ncu, plan, s);
      if (rc) return rc;
    }
    if (opts & POTRF_ASYNC) {  // the caller reads info after its own synchronisation
      dim3 g((unsigned)((n_pad + 255) / 256 > 64 ? 64 : (n_pad + 255) / 256), (unsigned)n_pad);
      if (!(opts & POTRF_KEEP_UPPER)) {
        hipLaunchKernelGGL(zero_upper_kernel, g, dim3(256), 0, s, A, lda, n_pad);
        BTA_LAUNCH_CHECK("zero_upper_kernel");
      }
      if (Linv && !band) {
        hipLaunchKernelGGL(trinv_diag_kernel, dim3((unsigned)NB), dim3(256), 0, s, A, lda, Linv,
                           ldinv, n_pad);
        BTA_LAUNCH_CHECK("trinv_diag_kernel");
      }
      if (invert && !(opts & POTRF_BAND_ZERO)) {
        hipLaunchKernelGGL(zero_upper_kernel, g, dim3(256), 0, s, Linv, ldinv, n_pad);
        BTA_LAUNCH_CHECK("zero_upper_kernel");
      }
      return BTA_OK;
    }
    int hinfo[2];
    BTA_HIP(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, s));
    BTA_HIP(hipStreamSynchronize(s));
    if (hinfo[0] == 0) {
      dim3 g((unsigned)((n_pad + 255) / 256 > 64 ? 64 : (n_pad + 255) / 256), (unsigned)n_pad);
      if (!(opts & POTRF_KEEP_UPPER)) {
        hipLaunchKernelGGL(zero_upper_kernel, g, dim3(256), 0, s, A, lda, n_pad);
        BTA_LAUNCH_CHECK("zero_upper_kernel");
      }
      if (Linv && !band) {  // (with it, the steps wrote them)
        hipLaunchKernelGGL(trinv_diag_kernel, dim3((unsigned)NB), dim3(256), 0, s, A, lda, Linv,
                           ldinv, n_pad);
        BTA_LAUNCH_CHECK("trinv_diag_kernel");
      }
      if (invert && !(opts & POTRF_BAND_ZERO)) {
        hipLaunchKernelGGL(zero_upper_kernel, g, dim3(256), 0, s, Linv, ldinv, n_pad);
        BTA_LAUNCH_CHECK("zero_upper_kernel");
      }
      if (jitter_used) *jitter_used = jit;
      return BTA_OK;
    }
    if (hinfo[1]) {
      set_error("potrf: matrix contains NaN (failing column %d)", hinfo[0]);
      return BTA_ERR_NAN;
    }
  }
  set_error("potrf: matrix not positive definite after adding jitter up to %.1e",
            jitter0 * pow(10.0, (double)(max_tries - 1)));
  return BTA_ERR_NOT_PSD;
}

const int* potrf_status(const void* ws, int64_t n_pad) {
  return (const int*)((const double*)ws + n_pad * n_pad + LEAF * LEAF);
}

size_t trtri_ws_bytes(int64_t n_pad) { return (size_t)(n_pad * n_pad / 4 + 64) * sizeof(double); }

int trtri_lower(const double* L, int64_t ldl, double* X, int64_t ldx, int64_t n_pad, void* ws,
                size_t ws_bytes, hipStream_t s, int opts) {
  if (n_pad <= 0 || n_pad % LEAF) {
    set_error("trtri: n_pad must be a positive multiple of 64");
    return BTA_ERR_INVALID;
  }
  if (!ws || ws_bytes < trtri_ws_bytes(n_pad)) {
    set_error("trtri: workspace too small");
    return BTA_ERR_WORKSPACE;
  }
  double* T = (double*)ws;
  if (!(opts & TRTRI_UPPER_ZERO)) {
    dim3 g((unsigned)((n_pad + 255) / 256 > 64 ? 64 : (n_pad + 255) / 256), (unsigned)n_pad);
    hipLaunchKernelGGL(zero_upper_kernel, g, dim3(256), 0, s, X, ldx, n_pad);
    BTA_LAUNCH_CHECK("zero_upper_kernel");
  }
  const int64_t NB = n_pad / LEAF;
  for (int64_t sb = 1; sb < NB; sb *= 2) {
    const int64_t S = sb * LEAF;
    const int64_t pf = NB / (2 * sb);  // pairs with a full second block
    if (pf > 0) {
      const int64_t strideL = 2 * S * (ldl + 1), strideX = 2 * S * (ldx + 1);
      // T_p = C_p X_A,p   (C_p = L[b0, a0], X_A = X[a0, a0] lower)
      int rc = gemm(BTA_GEMM_B_KN | BTA_GEMM_B_UPPER, S, S, S, 1.0, L + S * ldl, ldl, X, ldx,
                    0.0, nullptr, 0, T, S, s, (int)pf, strideL, strideX, 0, S * S);
      if (rc) return rc;
      // X_C,p = -X_B,p T_p
      rc = gemm(BTA_GEMM_B_KN | BTA_GEMM_A_LOWER, S, S, S, -1.0, X + S * ldx + S, ldx, T, S, 0.0,
                nullptr, 0, X + S * ldx, ldx, s, (int)pf, strideX, S * S, 0, strideX);
      if (rc) return rc;
    }
    const int64_t a0 = pf * 2 * sb, b0 = a0 + sb;
    if (b0 < NB) {  // ragged last pair: second block of sbr < sb leaves
      const int64_t SB = (NB - b0) * LEAF;
      const int64_t ra = a0 * LEAF, rb = b0 * LEAF;
      int rc = gemm(BTA_GEMM_B_KN | BTA_GEMM_B_UPPER, SB, S, S, 1.0, L + rb * ldl + ra, ldl,
                    X + ra * ldx + ra, ldx, 0.0, nullptr, 0, T, S, s);
      if (rc) return rc;
      rc = gemm(BTA_GEMM_B_KN | BTA_GEMM_A_LOWER, SB, S, SB, -1.0, X + rb * ldx + rb, ldx, T, S,
                0.0, nullptr, 0, X + rb * ldx + ra, ldx, s);
      if (rc) return rc;
    }
  }
  return BTA_OK;
}

// Adjoint of the Cholesky factorisation of one n x n matrix per workgroup
// (n <= 64): with P = Phi(L^T dL) (lower triangle, halved diagonal),
// S = L^-T P L^-1 and dA = (S + S^T) / 2, the form torch's Cholesky backward
// takes (the reference differentiates its posterior samples through it).  L and
// dL are read from their lower triangles; dA is written in full.
__global__ __launch_bounds__(256) void chol_small_backward_kernel(const double* L, const double* dL,
                                                                  int n, int64_t stride,
                                                                  double* dA) {
  __shared__ double Ls[LEAF][LP];
  __shared__ double W[LEAF][LP];
  const int t = threadIdx.x;
  const double* Lb = L + blockIdx.x * stride;
  const double* dLb = dL + blockIdx.x * stride;
  for (int e = t; e < n * n; e += 256) {
    const int i = e / n, j = e % n;
    Ls[i][j] = j <= i ? Lb[e] : 0.0;
  }
  __syncthreads();
  // W = Phi(L^T dL): W[i][j] = sum_{k >= i} L[k][i] dL[k][j] for j <= i (halved at j == i)
  for (int e = t; e < n * n; e += 256) {
    const int i = e / n, j = e % n;
    double v = 0.0;
    if (j <= i) {
      for (int k = i; k < n; ++k) v = fma(Ls[k][i], dLb[k * n + j], v);
      if (j == i) v *= 0.5;
    }
    W[i][j] = v;
  }
  __syncthreads();
  // W <- L^-T W: back substitution on each column j (thread j)
  if (t < n) {
    const int j = t;
    for (int i = n - 1; i >= 0; --i) {
      double v = W[i][j];
      for (int k = i + 1; k < n; ++k) v -= Ls[k][i] * W[k][j];
      W[i][j] = v / Ls[i][i];
    }
  }
  __syncthreads();
  // W <- W L^-1: for each row i (thread i) solve x L = w from the right
  if (t < n) {
    const int i = t;
    for (int j = n - 1; j >= 0; --j) {
      double v = W[i][j];
      for (int k = j + 1; k < n; ++k) v -= W[i][k] * Ls[k][j];
      W[i][j] = v / Ls[j][j];
    }
  }
  __syncthreads();
  double* out = dA + blockIdx.x * stride;
  for (int e = t; e < n * n; e += 256) {
    const int i = e / n, j = e % n;
    out[e] = 0.5 * (W[i][j] + W[j][i]);
  }
}

int chol_small_backward(const double* L, const double* dL, int n, int64_t stride, int64_t batch,
                        double* dA, hipStream_t s) {
  if (!L || !dL || !dA || n < 1 || n > LEAF || stride < (int64_t)n * n || batch < 0 ||
      batch > (1 << 30)) {
    set_error("chol_small_backward: need 1 <= n <= 64, stride >= n^2");
    return BTA_ERR_INVALID;
  }
  if (batch == 0) return BTA_OK;
  hipLaunchKernelGGL(chol_small_backward_kernel, dim3((unsigned)batch), dim3(256), 0, s, L, dL, n,
                     stride, dA);
  BTA_LAUNCH_CHECK("chol_small_backward_kernel");
  return BTA_OK;
}

int chol_small_batched(double* A, int n, int64_t lda, int64_t stride, int64_t batch,
                       int max_tries, double jitter0, int* status, hipStream_t s) {
  if (n < 1 || n > LEAF || lda < n || batch < 0) {
    set_error("chol_small_batched: need 1 <= n <= 64");
    return BTA_ERR_UNSUPPORTED;
  }
  if (batch == 0) return BTA_OK;
  hipLaunchKernelGGL(chol_small_kernel, dim3((unsigned)batch), dim3(256), 0, s, A, n, lda,
                     stride, max_tries, jitter0, status);
  BTA_LAUNCH_CHECK("chol_small_kernel");
  return BTA_OK;
}

}  // namespace bta

extern "C" size_t bta_potrf_workspace_size(int64_t n_pad) { return bta::potrf_ws_bytes(n_pad); }

extern "C" int bta_potrf(double* A, int64_t n_pad, int64_t lda, double* Linv, int64_t ldinv,
                         int32_t max_tries, double jitter0, double* jitter_used, void* ws,
                         size_t ws_bytes, bta_stream_t stream) {
  if (!A) {
    bta::set_error("bta_potrf: null pointer");
    return BTA_ERR_INVALID;
  }
  return bta::potrf(A, n_pad, lda, Linv, ldinv, max_tries, jitter0, jitter_used, ws, ws_bytes,
                    (hipStream_t)stream, 0);
}

extern "C" int bta_potrf_inverse(double* A, int64_t n_pad, int64_t lda, double* Linv,
                                 int64_t ldinv, int32_t max_tries, double jitter0,
                                 double* jitter_used, void* ws, size_t ws_bytes,
                                 bta_stream_t stream) {
  if (!A || !Linv) {
    bta::set_error("bta_potrf_inverse: null pointer");
    return BTA_ERR_INVALID;
  }
  return bta::potrf(A, n_pad, lda, Linv, ldinv, max_tries, jitter0, jitter_used, ws, ws_bytes,
                    (hipStream_t)stream, bta::POTRF_INVERT);
}

extern "C" int bta_trtri_lower(const double* L, int64_t ldl, double* Linv, int64_t ldinv,
                               int64_t n_pad, void* ws, size_t ws_bytes, bta_stream_t stream) {
  if (!L || !Linv) {
    bta::set_error("bta_trtri_lower: null pointer");
    return BTA_ERR_INVALID;
  }
  return bta::trtri_lower(L, ldl, Linv, ldinv, n_pad, ws, ws_bytes, (hipStream_t)stream, 0);
}

extern "C" int bta_chol_small_batched(double* A, int32_t n, int64_t lda, int64_t stride,
                                      int64_t batch, int32_t max_tries, double jitter0,
                                      int32_t* status, bta_stream_t stream) {
  if (!A) {
    bta::set_error("bta_chol_small_batched: null pointer");
    return BTA_ERR_INVALID;
  }
  return bta::chol_small_batched(A, n, lda, stride, batch, max_tries, jitter0, status,
                                 (hipStream_t)stream);
}

extern "C" int bta_chol_small_backward(const double* L, const double* dL, int32_t n,
                                       int64_t stride, int64_t batch, double* dA,
                                       bta_stream_t stream) {
  return bta::chol_small_backward(L, dL, n, stride, batch, dA, (hipStream_t)stream);
}
