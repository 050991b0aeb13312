// Internal declarations shared between the library's translation units.
#pragma once
#include "bta_common.h"

namespace bta {

int scale_inputs(const double* X, int64_t n, int d, const double* center, const double* ls,
                 double* center_out, double* Xs, int64_t n_pad, hipStream_t st);
int covar(int kind, const double* X1s, int64_t n1, const double* X2s, int64_t n2, int d,
          double s, double diag_add, bool f32, void* K, int64_t ldk, int64_t n1_pad,
          int64_t n2_pad, hipStream_t st, int pad_identity = 0, int batch = 1,
          int64_t sX1 = 0, int64_t sX2 = 0, int64_t sK = 0);
// Lower triangle (diagonal included) of s k(Xs, Xs) + diag_add I, identity
// on the padding rows/columns; nothing above the diagonal is written.
int covar_sym_lower(int kind, const double* Xs, int64_t n, int d, double s, double diag_add,
                    double* K, int64_t ldk, int64_t n_pad, hipStream_t st,
                    int* zero2 = nullptr);
int covar_grad(int kind, const double* X1s, int64_t n1, const double* X2s, int64_t n2, int d,
               double s, const void* G, bool g_f32, int64_t ldg, const double* extra,
               const double* ls, double* dX1, hipStream_t st, const double* dext = nullptr,
               const double* E = nullptr, int ne = 0);
int covar_ext(int kind, const double* X1s, int64_t n1, const double* X2s, int64_t n2, int d,
              double s, const double* E, int ne, double* P, hipStream_t st);
size_t potrf_ws_bytes(int64_t n_pad);
// potrf options: POTRF_NO_SAVE keeps no copy of the input (a failure is
// returned at once; the caller rebuilds the matrix with jitter);
// POTRF_KEEP_UPPER leaves the strict upper triangle as it was.  With Linv, the
// 64 x 64 diagonal blocks of L^-1 are written and the rest of its upper
// triangle zeroed; POTRF_BAND_ZERO zeroes only the 64 x 64 blocks right of
// the diagonal blocks (enough for trtri_lower and for products on 128-row
// tiles, at any block offset, that treat Linv as lower triangular), and the
// diagonal blocks are inverted inside the factorisation steps.
// POTRF_ASYNC (with POTRF_NO_SAVE) only enqueues the factorisation: no host
// synchronisation and no jitter retry.  Its outcome is the pair of ints at
// potrf_status(ws, n_pad) once the stream has passed the factorisation
// ({0, 0} = factorised; else {failing column + 1, NaN seen}); the caller checks
// it and discards what it enqueued after a failure.  trtri_lower may reuse the
// same workspace: its scratch ends before the pair.  POTRF_INFO_ZEROED: the
// caller has already zeroed that pair on the stream, so no fill is launched.
// POTRF_INVERT (with Linv, 16-byte aligned with an even ldinv): Linv receives
// all of L^-1, built by side tasks of the factorisation steps and two more
// launches after them (no trtri_lower); its strict upper triangle is zeroed
// unless POTRF_BAND_ZERO asks for the band only.
constexpr int POTRF_NO_SAVE = 1, POTRF_KEEP_UPPER = 2, POTRF_BAND_ZERO = 4, POTRF_ASYNC = 8,
              POTRF_INFO_ZEROED = 16, POTRF_INVERT = 32;
int potrf(double* A, int64_t n_pad, int64_t lda, double* Linv, int64_t ldinv, int max_tries,
          double jitter0, double* jitter_used, void* ws, size_t ws_bytes, hipStream_t s,
          int opts = 0);
const int* potrf_status(const void* ws, int64_t n_pad);
size_t trtri_ws_bytes(int64_t n_pad);
// TRTRI_UPPER_ZERO: X's strict upper triangle is already zero.
constexpr int TRTRI_UPPER_ZERO = 1;
int trtri_lower(const double* L, int64_t ldl, double* X, int64_t ldx, int64_t n_pad, void* ws,
                size_t ws_bytes, hipStream_t s, int opts = 0);
template <typename TO>
int cast_pad(const double* in, int64_t ldi, int64_t rows, int64_t cols, TO* out, int64_t ldo,
             int64_t rows_pad, int64_t cols_pad, hipStream_t s);
int launch_baseline_best_f(const double* ka, double y_mean, double y_std, double c,
                           const double* Lb, const double* Zb, int n_b, int S, double* best_f,
                           hipStream_t s);

// Split-FP16 GEMM on pre-split operands (gemm_h2.hip).
struct H2Args {
  const _Float16 *Ah, *Al;  // M x K planes (row stride lda halfs)
  const _Float16 *Bh, *Bl;  // N x K planes (row stride ldb halfs)
  float* C;                 // M x N (row stride ldc)
  _Float16 *Ch, *Cl;        // optional planes of c_scale * C (row stride ldc)
  const float* a_row_scale; // optional per-row scale of A
  float a_scale, b_scale, c_scale;
  int64_t lda, ldb, ldc;
  int M, N, K, flags;       // flags: BTA_GEMM_B_LOWER / BTA_GEMM_B_UPPER (N x K view)
  double alpha;
  int seg_chunks;           // 0: FP32 sums over all of K; > 0: FP64 sums of FP32
                            // segments of seg_chunks x 32 K
  int xcd_groups;           // set by the plan: > 1 deals column strips to G worker
                            // groups (one per XCD), see gemm_h2_body
};
int gemm_h2(const H2Args& a, hipStream_t s);
// The tile width, worker count and worker groups gemm_h2 would launch with
// (after its checks).
int gemm_h2_plan(const H2Args& a, int* bn, int* workers, int* xcd_groups = nullptr);
// The plans the measured tuner tries for a problem: up to max (bn, workers,
// xcd_groups) triples into out; returns their number.
int gemm_h2_candidates(const H2Args& a, int* out, int max);
// Makes (bn, workers, xcd_groups) the plan of the problem's signature.
void gemm_h2_set_plan(const H2Args& a, int bn, int workers, int xcd_groups);
int split_f16(bool f64, const void* X, int64_t ldx, int64_t rows, int64_t cols,
              int64_t rows_pad, int64_t cols_pad, float scale, bool transpose, _Float16* hi,
              _Float16* lo, int64_t ldo, hipStream_t s);

// Fused elementwise stages of the qLogNEI step (qstep.hip).
constexpr int QS_NE_MAX = 16;  // columns of [alpha | A] staged per pass
constexpr int QS_COL_TILE = 64;

struct QsCovarArgs {
  const double* X;        // R x d raw query rows
  const double* center;   // d
  const double* ls;       // d
  const double* T;        // n x d scaled training inputs
  const double* E;        // n x ne  [alpha | A]
  void* K;                // R_pad x ldk (FP64 / FP32 / FP16 hi plane)
  _Float16* Klo;          // FP16 lo plane (mode 2)
  double* xs_out;         // R_pad x d scaled rows (zero beyond R)
  double* Ppart;          // R x ne x (n_pad / 64) partial products: one (row, column)'s
                          // column-tile partials are contiguous
  int64_t R, R_pad, n, n_pad, ldk;
  int d, kind, ne;
  double s;
  float kscale;
};

// The adjoint of K_x is dK_x[r] = sum_{j<q} gq[r][j] U[q floor(r/q) + j] with
// U = K_x K^-1 (= V L^-1): the q rows of a restart mix through its q x q block.
struct QsGradArgs {
  const double* xs;       // R x d scaled query rows
  const double* T;        // n x d scaled training inputs
  const double* E;        // n x ne
  const void* U;          // R x ldu, FP64 or FP32
  const double* gq;       // R x q
  const double* dext;     // R x ne adjoint of P
  double* Dpart;          // (n_pad / 64) x R x d partial input gradients
  int64_t R, R_pad, n, n_pad, ldu;
  int d, kind, ne, q;
  double s;
};

int qstep_covar(const QsCovarArgs& a, int mode, hipStream_t st);
int qstep_grad(const QsGradArgs& a, bool u_f32, const double* extra, const double* ls,
               double* dX, hipStream_t st);

// Bump allocator over a caller workspace (256-byte aligned slices).
struct WsPlan {
  size_t off = 0;
  size_t take(size_t bytes) {
    size_t r = off;
    off += (bytes + 255) / 256 * 256;
    return r;
  }
};

}  // namespace bta
