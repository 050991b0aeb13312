// MFMA GEMM tile mainloop shared by the GEMM kernel and the fused MLL
// trace kernel.  See gemm.hip for the design notes.
#pragma once
#include "bta_common.h"

namespace bta {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef v4d acc_t;
  // row (within the 16x16 block) of accumulator element q held by lane group g
  __host__ __device__ static constexpr int row(int g, int q) { return g + 4 * q; }
  __device__ static __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  typedef v4f acc_t;
  __host__ __device__ static constexpr int row(int g, int q) { return 4 * g + q; }
  __device__ static __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};

template <typename T, int BM, int BN, int BK, bool BKN, bool AKM = false>
struct Cfg {
  static constexpr int VEC = 16 / sizeof(T);
  static constexpr int PITCH = BK + VEC;     // rows of K-contiguous tiles
  static constexpr int PITCH_KN = BN + VEC;  // rows of a K x N B tile
  static constexpr int PITCH_KM = BM + VEC;  // rows of a K x M A tile
  static constexpr int A_ELEMS = AKM ? BK * PITCH_KM : BM * PITCH;
  static constexpr int B_ELEMS = BKN ? BK * PITCH_KN : BN * PITCH;
  static constexpr int STAGE = A_ELEMS + B_ELEMS;
  static constexpr int MI = BM / 32;
  static constexpr int NI = BN / 32;
  static constexpr int KS = BK / 4;
  static constexpr int A_VECS = BM * BK / VEC / 256;
  static constexpr int B_VECS = BN * BK / VEC / 256;
  static_assert(A_VECS >= 1 && B_VECS >= 1, "tile too small for 256 threads");
  static_assert(BM * BK % (VEC * 256) == 0 && BN * BK % (VEC * 256) == 0, "tile shape");
};

__host__ __device__ inline void decode_tile(int t, int flags, int MT, int NT, int& tm,
                                            int& tn) {
  if (flags & BTA_GEMM_C_LOWER) {
    // lower tiles (tm >= tn), enumerated by ascending tm
    int r = (int)((__builtin_sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    tm = r;
    tn = t - r * (r + 1) / 2;
  } else if (flags & BTA_GEMM_B_LOWER) {  // cost grows with tn
    tn = NT - 1 - t / MT;
    tm = t % MT;
  } else if (flags & BTA_GEMM_B_UPPER) {  // cost falls with tn
    tn = t / MT;
    tm = t % MT;
  } else if (flags & BTA_GEMM_A_LOWER) {
    tm = MT - 1 - t / NT;
    tn = t % NT;
  } else if (flags & BTA_GEMM_A_UPPER) {
    tm = t / NT;
    tn = t % NT;
  } else {
    tm = t % MT;
    tn = t / MT;
  }
}

template <typename T, int BM, int BN, int BK, bool BKN, bool AKM = false>
__device__ __forceinline__ void gemm_tile(
    const T* __restrict__ A, int64_t lda, const T* __restrict__ B, int64_t ldb, int row0,
    int col0, int kb, int ke, T* lds,
    typename Mfma<T>::acc_t (&acc)[Cfg<T, BM, BN, BK, BKN, AKM>::MI]
                                  [Cfg<T, BM, BN, BK, BKN, AKM>::NI]) {
  typedef Cfg<T, BM, BN, BK, BKN, AKM> C;
  typedef typename Vec16<T>::type V;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  const int wy = wv >> 1, wx = wv & 1;
  const int l16 = lane & 15, g = lane >> 4;

#pragma unroll
  for (int mi = 0; mi < C::MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < C::NI; ++ni) acc[mi][ni] = (typename Mfma<T>::acc_t)(0);
  if (kb >= ke) return;

  V ra[C::A_VECS], rb[C::B_VECS];
  auto gload = [&](int k0) {
#pragma unroll
    for (int i = 0; i < C::A_VECS; ++i) {
      int v = tid + i * 256;
      if (AKM) {
        int kr = v / (BM / C::VEC), mv = v % (BM / C::VEC);
        ra[i] = *(const V*)(A + (int64_t)(k0 + kr) * lda + row0 + mv * C::VEC);
      } else {
        int r = v / (BK / C::VEC), kv = v % (BK / C::VEC);
        ra[i] = *(const V*)(A + (int64_t)(row0 + r) * lda + k0 + kv * C::VEC);
      }
    }
#pragma unroll
    for (int i = 0; i < C::B_VECS; ++i) {
      int v = tid + i * 256;
      if (BKN) {
        int kr = v / (BN / C::VEC), nv = v % (BN / C::VEC);
        rb[i] = *(const V*)(B + (int64_t)(k0 + kr) * ldb + col0 + nv * C::VEC);
      } else {
        int r = v / (BK / C::VEC), kv = v % (BK / C::VEC);
        rb[i] = *(const V*)(B + (int64_t)(col0 + r) * ldb + k0 + kv * C::VEC);
      }
    }
  };
  auto sstore = [&](T* s) {
#pragma unroll
    for (int i = 0; i < C::A_VECS; ++i) {
      int v = tid + i * 256;
      if (AKM) {
        int kr = v / (BM / C::VEC), mv = v % (BM / C::VEC);
        *(V*)(s + kr * C::PITCH_KM + mv * C::VEC) = ra[i];
      } else {
        int r = v / (BK / C::VEC), kv = v % (BK / C::VEC);
        *(V*)(s + r * C::PITCH + kv * C::VEC) = ra[i];
      }
    }
    T* sb = s + C::A_ELEMS;
#pragma unroll
    for (int i = 0; i < C::B_VECS; ++i) {
      int v = tid + i * 256;
      if (BKN) {
        int kr = v / (BN / C::VEC), nv = v % (BN / C::VEC);
        *(V*)(sb + kr * C::PITCH_KN + nv * C::VEC) = rb[i];
      } else {
        int r = v / (BK / C::VEC), kv = v % (BK / C::VEC);
        *(V*)(sb + r * C::PITCH + kv * C::VEC) = rb[i];
      }
    }
  };

  gload(kb);
  sstore(lds);
  __syncthreads();
  const int nk = (ke - kb) / BK;
  for (int it = 0; it < nk; ++it) {
    const T* cur = lds + (it & 1) * C::STAGE;
    T* nxt = lds + ((it + 1) & 1) * C::STAGE;
    if (it + 1 < nk) gload(kb + (it + 1) * BK);
    T af[C::MI][C::KS], bf[C::NI][C::KS];
    const T* As = cur;
    const T* Bs = cur + C::A_ELEMS;
#pragma unroll
    for (int mi = 0; mi < C::MI; ++mi) {
      if (AKM) {
        const T* p = As + (g * C::KS) * C::PITCH_KM + wy * (BM / 2) + mi * 16 + l16;
#pragma unroll
        for (int s = 0; s < C::KS; ++s) af[mi][s] = p[s * C::PITCH_KM];
      } else {
        const T* p = As + (wy * (BM / 2) + mi * 16 + l16) * C::PITCH + g * C::KS;
#pragma unroll
        for (int s = 0; s < C::KS; ++s) af[mi][s] = p[s];
      }
    }
#pragma unroll
    for (int ni = 0; ni < C::NI; ++ni) {
      if (BKN) {
        const T* p = Bs + (g * C::KS) * C::PITCH_KN + wx * (BN / 2) + ni * 16 + l16;
#pragma unroll
        for (int s = 0; s < C::KS; ++s) bf[ni][s] = p[s * C::PITCH_KN];
      } else {
        const T* p = Bs + (wx * (BN / 2) + ni * 16 + l16) * C::PITCH + g * C::KS;
#pragma unroll
        for (int s = 0; s < C::KS; ++s) bf[ni][s] = p[s];
      }
    }
#pragma unroll
    for (int s = 0; s < C::KS; ++s)
#pragma unroll
      for (int mi = 0; mi < C::MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < C::NI; ++ni)
          acc[mi][ni] = Mfma<T>::op(af[mi][s], bf[ni][s], acc[mi][ni]);
    if (it + 1 < nk) sstore(nxt);
    __syncthreads();
  }
}

}  // namespace bta
