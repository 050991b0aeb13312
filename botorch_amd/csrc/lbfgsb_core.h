// L-BFGS-B (bound-constrained limited-memory BFGS), version 3.0 of Byrd, Lu,
// Nocedal and Zhu: R. H. Byrd, P. Lu, J. Nocedal, C. Zhu, "A limited memory
// algorithm for bound constrained optimization", SIAM J. Sci. Comput. 16 (1995);
// C. Zhu, R. H. Byrd, P. Lu, J. Nocedal, "Algorithm 778: L-BFGS-B", ACM TOMS 23
// (1997); J. L. Morales, J. Nocedal, "Remark on Algorithm 778", ACM TOMS 38
// (2011) (the projected subspace step).  The line search is the one of J. J.
// More and D. J. Thuente, "Line search algorithms with guaranteed sufficient
// decrease", ACM TOMS 20 (1994) (MINPACK-2 dcsrch / dcstep).
//
// This is the per-problem optimiser that the reference drives through SciPy's
// setulb once per restart and evaluation (optim/batched_lbfgs_b.py:440-575,
// called from generation/gen.py:236-249 with maxcor 10, factr 1e7, pgtol 1e-5,
// maxls 20).  The same reverse-communication steps are written here once,
// templated on a "lanes" context C:
//   - on the GPU, C is one 64-lane wavefront per problem: loops over the n
//     variables and over small-matrix entries are split across lanes, the
//     serial parts (the breakpoint heap of the Cauchy search, the 2m x 2m
//     factorisations, the line search) run uniformly on every lane;
//   - on the host (a test harness), C has one lane and every loop is serial.
// All arrays of a problem (View) live in one contiguous region: on the GPU the
// kernel stages the problem's slab from HBM into LDS, runs advance(), and
// writes the slab back.
//
// Steps follow the published algorithm: projected gradient norm, generalized
// Cauchy point along the projected steepest-descent path (breakpoints taken in
// heap order), free/active split at the Cauchy point, direct primal subspace
// minimisation through the 2col x 2col LEL^T factorisation of the compact
// representation B = theta I - W M W^T, projection of the subspace step (with
// backtracking to the feasible set when the projected step is not a descent
// direction), More-Thuente line search with ftol 1e-3, gtol 0.9, xtol 0.1, and
// the skip / refresh rules for the correction pairs.  The products of the pairs
// over the free and active variables (WN1) persist from one iteration to the
// next and are updated, as in the published code, only when the subspace step
// runs: a new row for the newest pair plus the contributions of the variables
// that entered or left the free set since the previous free/active split.
// That keeps its behaviour where an iteration had no free variable and so no
// update of WN1 (the next factorisation can then fail and refresh the memory).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace bta {
namespace lbfgsb {

constexpr double EPSMCH = 2.220446049250313e-16;  // double machine epsilon
constexpr double STPMAX_FREE = 1e10;              // step bound of unconstrained problems
constexpr double LS_FTOL = 1e-3, LS_GTOL = 0.9, LS_XTOL = 0.1;

// where the reverse-communication machine resumes
enum Pc : int32_t {
  PC_START = 0,     // nothing evaluated yet
  PC_FG0 = 1,       // waiting for f, g at the (projected) starting point
  PC_ITER = 2,      // start of an iteration
  PC_LS_START = 3,  // line search about to start from the new direction
  PC_LS_FG = 4,     // waiting for f, g at a line-search trial point
  PC_NEWX = 5,      // a new iterate was accepted
  PC_DONE = 6
};

// how a problem ended
enum Task : int32_t {
  TASK_RUNNING = 0,
  TASK_CONV_PGTOL = 1,   // projected gradient sup-norm <= pgtol
  TASK_CONV_FACTR = 2,   // relative reduction of f <= factr * eps
  TASK_ABNORMAL = 3,     // line search failed with no correction pairs left to drop
  TASK_STOP_MAXITER = 4,
  TASK_STOP_MAXFUN = 5,
  TASK_NAN_GRADIENT = 6  // the supplied gradient holds a NaN
};

enum LsTask : int32_t { LS_START = 0, LS_FG = 1, LS_CONV = 2, LS_WARN = 3, LS_ERROR = 4 };

// phases a lanes context may time through mark() (both library contexts ignore
// the marks): the generalized Cauchy point, the free/active split, the
// subspace factorisation, the reduced gradient and subspace step, the line
// search and the pair update
enum Phase : int32_t {
  PH_OTHER = 0, PH_CAUCHY, PH_FREEV, PH_FORMK, PH_SUBSM, PH_LNSRCH, PH_UPDATE, PH_COUNT
};

// the scalar state of one problem (first bytes of its slab)
struct Scal {
  double f, fold, theta, stp, stpmx, gd, gdold, dtd, dnorm, sbgnrm;
  double factr, pgtol;
  // More-Thuente state
  double ls_finit, ls_ginit, ls_gtest, ls_stx, ls_fx, ls_gx, ls_sty, ls_fy, ls_gy;
  double ls_stmin, ls_stmax, ls_width, ls_width1;
  int32_t ls_brackt, ls_stage, ls_task;
  int32_t maxls, maxiter, maxfun;
  int32_t col, head, itail, iupdat, updatd, iter, nfgv, ifun, iback, nskip, nfree;
  int32_t cnstnd, boxed, prjctd, info;
  int32_t pc, task, nit, nfev, warnflag, step_status, done, nenter, nleave, pad;
};

// offsets (in doubles) of one problem's arrays: the persistent slab first, the
// per-call temporaries after it
struct Layout {
  int n, m;
  int o_x, o_g, o_t, o_r, o_d, o_z, o_ws, o_wy, o_sy, o_ss, o_wt, o_wn1, o_iw, o_free, slab;
  int o_wn, o_p, o_c, o_v, o_wbp, o_wv, o_rd, o_xp, o_bt, o_bi, o_lists, total;
  __host__ __device__ Layout(int n_, int m_) : n(n_), m(m_) {
    int o = (int)((sizeof(Scal) + 7) / 8);
    const int ni = (n + 1) / 2;  // n int32 in doubles
    o_x = o; o += n;
    o_g = o; o += n;
    o_t = o; o += n;
    o_r = o; o += n;
    o_d = o; o += n;
    o_z = o; o += n;
    o_ws = o; o += n * m;
    o_wy = o; o += n * m;
    o_sy = o; o += m * m;
    o_ss = o; o += m * m;
    o_wt = o; o += m * m;
    o_wn1 = o; o += 4 * m * m;
    o_iw = o; o += ni;
    o_free = o; o += ni;
    slab = o;
    o_wn = o; o += 4 * m * m;
    o_p = o; o += 2 * m;
    o_c = o; o += 2 * m;
    o_v = o; o += 2 * m;
    o_wbp = o; o += 2 * m;
    o_wv = o; o += 2 * m;
    o_rd = o; o += n;
    o_xp = o; o += n;
    o_bt = o; o += n;
    o_bi = o; o += ni;
    o_lists = o; o += 4 * ni;
    total = o;
  }
};

// the arrays of one problem; bounds are shared by all problems
struct View {
  Scal* s;
  double *x, *g, *t, *r, *d, *z, *ws, *wy, *sy, *ss, *wt, *wn1;
  int32_t *iwhere, *freeprev;
  double *wn, *p, *c, *v, *wbp, *wv, *rd, *xp, *bt;
  int32_t *bi;
  // variables at the Cauchy point, increasing: free, active, entered and left
  // the free set since the previous split (counts nfree, n - nfree, nenter, nleave)
  int32_t *lfree, *lact, *lent, *llv;
  const double* lo;
  const double* up;
  const int32_t* nbd;
  int n, m;
};

__host__ __device__ inline View make_view(double* base, const Layout& L, const double* lo,
                                          const double* up, const int32_t* nbd) {
  View w;
  w.s = (Scal*)base;
  w.x = base + L.o_x; w.g = base + L.o_g; w.t = base + L.o_t; w.r = base + L.o_r;
  w.d = base + L.o_d; w.z = base + L.o_z; w.ws = base + L.o_ws; w.wy = base + L.o_wy;
  w.sy = base + L.o_sy; w.ss = base + L.o_ss; w.wt = base + L.o_wt; w.wn1 = base + L.o_wn1;
  w.iwhere = (int32_t*)(base + L.o_iw); w.freeprev = (int32_t*)(base + L.o_free);
  w.wn = base + L.o_wn; w.p = base + L.o_p; w.c = base + L.o_c; w.v = base + L.o_v;
  w.wbp = base + L.o_wbp; w.wv = base + L.o_wv; w.rd = base + L.o_rd; w.xp = base + L.o_xp;
  w.bt = base + L.o_bt; w.bi = (int32_t*)(base + L.o_bi);
  w.lfree = (int32_t*)(base + L.o_lists);
  w.lact = w.lfree + L.n;
  w.lent = w.lact + L.n;
  w.llv = w.lent + L.n;
  w.lo = lo; w.up = up; w.nbd = nbd;
  w.n = L.n; w.m = L.m;
  return w;
}

// ---------------------------------------------------------------------------
// small dense helpers, serial forms (the host context runs these; the device
// context has lane-parallel forms with the same operation order)

// in-place upper Cholesky A = R^T R of the leading k x k block (row-major,
// leading dimension ld), column by column; returns the 1-based index of the
// first non-positive pivot, 0 on success
__host__ __device__ inline int serial_chol_upper(double* A, int ld, int k) {
  for (int j = 0; j < k; ++j) {
    double s = 0.0;
    for (int q = 0; q < j; ++q) {
      double t = A[q * ld + j];
      for (int i = 0; i < q; ++i) t -= A[i * ld + q] * A[i * ld + j];
      t /= A[q * ld + q];
      A[q * ld + j] = t;
    }
    for (int q = 0; q < j; ++q) s += A[q * ld + j] * A[q * ld + j];
    s = A[j * ld + j] - s;
    if (!(s > 0.0)) return j + 1;
    A[j * ld + j] = sqrt(s);
  }
  return 0;
}

// solve R^T y = b in place (R upper k x k): y_j = (b_j - sum_{i<j} R_ij y_i) / R_jj,
// the terms subtracted in increasing i
__host__ __device__ inline int serial_solve_rt(const double* R, int ld, int k, double* b) {
  for (int j = 0; j < k; ++j)
    if (R[j * ld + j] == 0.0) return j + 1;
  for (int j = 0; j < k; ++j) {
    double t = b[j];
    for (int i = 0; i < j; ++i) t -= R[i * ld + j] * b[i];
    b[j] = t / R[j * ld + j];
  }
  return 0;
}

// solve R y = b in place (R upper k x k) column by column from the last: y_j =
// b_j / R_jj, then b_i -= R_ij y_j for i < j (LINPACK dtrsl / LAPACK dtrsm order)
__host__ __device__ inline int serial_solve_r(const double* R, int ld, int k, double* b) {
  for (int j = 0; j < k; ++j)
    if (R[j * ld + j] == 0.0) return j + 1;
  for (int j = k - 1; j >= 0; --j) {
    b[j] = b[j] / R[j * ld + j];
    const double yj = b[j];
    for (int i = 0; i < j; ++i) b[i] -= R[i * ld + j] * yj;
  }
  return 0;
}

// the one-lane context: every loop serial, the small dense operations in their
// serial forms (the host build of the machine)
struct SerialLanes {
  static constexpr int W = 1;
  __host__ __device__ int lane() const { return 0; }
  __host__ __device__ double sum(double v) const { return v; }
  __host__ __device__ double max(double v) const { return v; }
  __host__ __device__ double min(double v) const { return v; }
  __host__ __device__ int sum_int(int v) const { return v; }
  __host__ __device__ int min_int(int v) const { return v; }
  __host__ __device__ int any(int v) const { return v != 0; }
  __host__ __device__ int ballot_rank(int flag, int& total) const {
    total = flag ? 1 : 0;
    return 0;
  }
  __host__ __device__ void sync() const {}
  __host__ __device__ void mark(int) const {}
  __host__ __device__ int chol_upper(double* A, int ld, int k) const {
    return serial_chol_upper(A, ld, k);
  }
  __host__ __device__ int solve_rt(const double* R, int ld, int k, double* b) const {
    return serial_solve_rt(R, ld, k, b);
  }
  __host__ __device__ int solve_r(const double* R, int ld, int k, double* b) const {
    return serial_solve_r(R, ld, k, b);
  }
};

// p = M v for the 2col x 2col middle matrix M of B = theta I - W M W^T,
// W = [Y, theta S]: M^{-1} = [[-D, L^T], [L, theta S^T S]] with D = diag(S^T Y)
// and L its strictly lower part.  wt holds the upper factor J^T of
// T = theta S^T S + L D^{-1} L^T = J J^T.  v and p are shared arrays.
// Nonzero return: singular factor.
template <class C>
__host__ __device__ inline int bmv(C& c, const View& w, int col, const double* v, double* p) {
  if (col == 0) return 0;
  const int m = w.m;
  const double* sy = w.sy;
  // p2 = v2 + L D^{-1} v1 and p1 = D^{-1/2} v1
  for (int i = c.lane(); i < col; i += C::W) {
    double sum = 0.0;
    for (int k = 0; k < i; ++k) sum += sy[i * m + k] * v[k] / sy[k * m + k];
    p[col + i] = i == 0 ? v[col] : v[col + i] + sum;
    p[i] = v[i] / sqrt(sy[i * m + i]);
  }
  c.sync();
  // p2 = J^{-1} p2, then J^{-T} p2
  if (c.solve_rt(w.wt, m, col, p + col)) return 1;
  if (c.solve_r(w.wt, m, col, p + col)) return 1;
  // p1 = -D^{-1/2} p1 + D^{-1} L^T p2
  for (int i = c.lane(); i < col; i += C::W) {
    const double pi = -p[i] / sqrt(sy[i * m + i]);
    double sum = 0.0;
    for (int k = i + 1; k < col; ++k) sum += sy[k * m + i] * p[col + k] / sy[i * m + i];
    p[i] = pi + sum;
  }
  c.sync();
  return 0;
}

// one step of heap sort on t[1..n] (1-based, with the indices iorder): when
// iheap == 0 the array is first arranged as a min-heap; then the least element
// is moved to position n and t[1..n-1] is made a heap again
__host__ __device__ inline void hpsolb(int n, double* t0, int32_t* io0, int iheap) {
  double* t = t0 - 1;
  int32_t* io = io0 - 1;
  if (iheap == 0) {
    for (int k = 2; k <= n; ++k) {
      const double ddum = t[k];
      const int32_t indxin = io[k];
      int i = k;
      while (i > 1) {
        const int j = i / 2;
        if (ddum < t[j]) {
          t[i] = t[j];
          io[i] = io[j];
          i = j;
        } else {
          break;
        }
      }
      t[i] = ddum;
      io[i] = indxin;
    }
  }
  if (n > 1) {
    int i = 1;
    const double out = t[1];
    const int32_t indxou = io[1];
    const double ddum = t[n];
    const int32_t indxin = io[n];
    for (;;) {
      int j = i + i;
      if (j <= n - 1) {
        if (t[j + 1] < t[j]) j = j + 1;
        if (t[j] < ddum) {
          t[i] = t[j];
          io[i] = io[j];
          i = j;
          continue;
        }
      }
      break;
    }
    t[i] = ddum;
    io[i] = indxin;
    t[n] = out;
    io[n] = indxou;
  }
}

// ---------------------------------------------------------------------------
// More-Thuente safeguarded step (dcstep): updates the interval [stx, sty] that
// holds a minimiser of the line-search function and computes the next trial
// step inside [stpmin, stpmax]
__host__ __device__ inline void dcstep(double& stx, double& fx, double& dx, double& sty,
                                       double& fy, double& dy, double& stp, double fp, double dp,
                                       int& brackt, double stpmin, double stpmax) {
  const double sgnd = dp * (dx / fabs(dx));
  double stpf;
  if (fp > fx) {
    // higher function value: the minimum is bracketed
    const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
    const double s = fmax(fmax(fabs(theta), fabs(dx)), fabs(dp));
    double gamma = s * sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
    if (stp < stx) gamma = -gamma;
    const double p = (gamma - dx) + theta;
    const double q = ((gamma - dx) + gamma) + dp;
    const double r = p / q;
    const double stpc = stx + r * (stp - stx);
    const double stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
    if (fabs(stpc - stx) < fabs(stpq - stx))
      stpf = stpc;
    else
      stpf = stpc + (stpq - stpc) / 2.0;
    brackt = 1;
  } else if (sgnd < 0.0) {
    // lower function value, derivatives of opposite sign
    const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
    const double s = fmax(fmax(fabs(theta), fabs(dx)), fabs(dp));
    double gamma = s * sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
    if (stp > stx) gamma = -gamma;
    const double p = (gamma - dp) + theta;
    const double q = ((gamma - dp) + gamma) + dx;
    const double r = p / q;
    const double stpc = stp + r * (stx - stp);
    const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
    if (fabs(stpc - stp) > fabs(stpq - stp))
      stpf = stpc;
    else
      stpf = stpq;
    brackt = 1;
  } else if (fabs(dp) < fabs(dx)) {
    // lower function value, same-sign derivatives, decreasing magnitude
    const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
    const double s = fmax(fmax(fabs(theta), fabs(dx)), fabs(dp));
    double gamma = s * sqrt(fmax(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
    if (stp > stx) gamma = -gamma;
    const double p = (gamma - dp) + theta;
    const double q = (gamma + (dx - dp)) + gamma;
    const double r = p / q;
    double stpc;
    if (r < 0.0 && gamma != 0.0)
      stpc = stp + r * (stx - stp);
    else if (stp > stx)
      stpc = stpmax;
    else
      stpc = stpmin;
    const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
    if (brackt) {
      stpf = fabs(stpc - stp) < fabs(stpq - stp) ? stpc : stpq;
      if (stp > stx)
        stpf = fmin(stp + 0.66 * (sty - stp), stpf);
      else
        stpf = fmax(stp + 0.66 * (sty - stp), stpf);
    } else {
      stpf = fabs(stpc - stp) > fabs(stpq - stp) ? stpc : stpq;
      stpf = fmin(stpmax, stpf);
      stpf = fmax(stpmin, stpf);
    }
  } else {
    // lower function value, same-sign derivatives, magnitude not decreasing
    if (brackt) {
      const double theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
      const double s = fmax(fmax(fabs(theta), fabs(dy)), fabs(dp));
      double gamma = s * sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
      if (stp > sty) gamma = -gamma;
      const double p = (gamma - dp) + theta;
      const double q = ((gamma - dp) + gamma) + dy;
      const double r = p / q;
      stpf = stp + r * (sty - stp);
    } else if (stp > stx) {
      stpf = stpmax;
    } else {
      stpf = stpmin;
    }
  }
  // update the interval that contains a minimiser
  if (fp > fx) {
    sty = stp;
    fy = fp;
    dy = dp;
  } else {
    if (sgnd < 0.0) {
      sty = stx;
      fy = fx;
      dy = dx;
    }
    stx = stp;
    fx = fp;
    dx = dp;
  }
  stp = stpf;
}

// More-Thuente line search (dcsrch) in reverse communication: given f and its
// derivative g along the direction at stp, sets s.ls_task and the next stp
__host__ __device__ inline void dcsrch(Scal& s, double f, double g, double& stp, double stpmin,
                                       double stpmax) {
  const double xtrapl = 1.1, xtrapu = 4.0;
  if (s.ls_task == LS_START) {
    if (stp < stpmin || stp > stpmax || g >= 0.0) {
      s.ls_task = LS_ERROR;
      return;
    }
    s.ls_brackt = 0;
    s.ls_stage = 1;
    s.ls_finit = f;
    s.ls_ginit = g;
    s.ls_gtest = LS_FTOL * s.ls_ginit;
    s.ls_width = stpmax - stpmin;
    s.ls_width1 = s.ls_width / 0.5;
    s.ls_stx = 0.0;
    s.ls_fx = s.ls_finit;
    s.ls_gx = s.ls_ginit;
    s.ls_sty = 0.0;
    s.ls_fy = s.ls_finit;
    s.ls_gy = s.ls_ginit;
    s.ls_stmin = 0.0;
    s.ls_stmax = stp + xtrapu * stp;
    s.ls_task = LS_FG;
    return;
  }
  const double ftest = s.ls_finit + stp * s.ls_gtest;
  if (s.ls_stage == 1 && f <= ftest && g >= 0.0) s.ls_stage = 2;
  int warn = 0;
  if (s.ls_brackt && (stp <= s.ls_stmin || stp >= s.ls_stmax)) warn = 1;
  if (s.ls_brackt && s.ls_stmax - s.ls_stmin <= LS_XTOL * s.ls_stmax) warn = 1;
  if (stp == stpmax && f <= ftest && g <= s.ls_gtest) warn = 1;
  if (stp == stpmin && (f > ftest || g >= s.ls_gtest)) warn = 1;
  if (warn) s.ls_task = LS_WARN;
  if (f <= ftest && fabs(g) <= LS_GTOL * (-s.ls_ginit)) s.ls_task = LS_CONV;
  if (s.ls_task == LS_WARN || s.ls_task == LS_CONV) return;
  if (s.ls_stage == 1 && f <= s.ls_fx && f > ftest) {
    // the modified function predicts the step while the decrease is insufficient
    double fm = f - stp * s.ls_gtest;
    double fxm = s.ls_fx - s.ls_stx * s.ls_gtest;
    double fym = s.ls_fy - s.ls_sty * s.ls_gtest;
    double gm = g - s.ls_gtest;
    double gxm = s.ls_gx - s.ls_gtest;
    double gym = s.ls_gy - s.ls_gtest;
    dcstep(s.ls_stx, fxm, gxm, s.ls_sty, fym, gym, stp, fm, gm, s.ls_brackt, s.ls_stmin,
           s.ls_stmax);
    s.ls_fx = fxm + s.ls_stx * s.ls_gtest;
    s.ls_fy = fym + s.ls_sty * s.ls_gtest;
    s.ls_gx = gxm + s.ls_gtest;
    s.ls_gy = gym + s.ls_gtest;
  } else {
    dcstep(s.ls_stx, s.ls_fx, s.ls_gx, s.ls_sty, s.ls_fy, s.ls_gy, stp, f, g, s.ls_brackt,
           s.ls_stmin, s.ls_stmax);
  }
  if (s.ls_brackt) {
    if (fabs(s.ls_sty - s.ls_stx) >= 0.66 * s.ls_width1)
      stp = s.ls_stx + 0.5 * (s.ls_sty - s.ls_stx);
    s.ls_width1 = s.ls_width;
    s.ls_width = fabs(s.ls_sty - s.ls_stx);
  }
  if (s.ls_brackt) {
    s.ls_stmin = fmin(s.ls_stx, s.ls_sty);
    s.ls_stmax = fmax(s.ls_stx, s.ls_sty);
  } else {
    s.ls_stmin = stp + xtrapl * (stp - s.ls_stx);
    s.ls_stmax = stp + xtrapu * (stp - s.ls_stx);
  }
  stp = fmax(stp, stpmin);
  stp = fmin(stp, stpmax);
  if ((s.ls_brackt && (stp <= s.ls_stmin || stp >= s.ls_stmax)) ||
      (s.ls_brackt && s.ls_stmax - s.ls_stmin <= LS_XTOL * s.ls_stmax))
    stp = s.ls_stx;
  s.ls_task = LS_FG;
}

// ---------------------------------------------------------------------------
// the steps of one iteration (C: lanes context)

template <class C>
__host__ __device__ inline double projgr(C& c, const View& w) {
  double mx = 0.0;
  for (int i = c.lane(); i < w.n; i += C::W) {
    double gi = w.g[i];
    const int nb = w.nbd[i];
    if (nb != 0) {
      if (gi < 0.0) {
        if (nb >= 2) gi = fmax(w.x[i] - w.up[i], gi);
      } else {
        if (nb <= 2) gi = fmin(w.x[i] - w.lo[i], gi);
      }
    }
    mx = fmax(mx, fabs(gi));
  }
  return c.max(mx);
}

// projects x onto the box and classifies the variables (-1 unbounded, 3 fixed,
// 0 otherwise)
template <class C>
__host__ __device__ inline void active(C& c, View& w) {
  int prj = 0, cn = 0, notbox = 0;
  for (int i = c.lane(); i < w.n; i += C::W) {
    const int nb = w.nbd[i];
    if (nb > 0) {
      if (nb <= 2 && w.x[i] <= w.lo[i]) {
        if (w.x[i] < w.lo[i]) {
          prj = 1;
          w.x[i] = w.lo[i];
        }
      } else if (nb >= 2 && w.x[i] >= w.up[i]) {
        if (w.x[i] > w.up[i]) {
          prj = 1;
          w.x[i] = w.up[i];
        }
      }
    }
    if (nb != 2) notbox = 1;
    if (nb == 0) {
      w.iwhere[i] = -1;
    } else {
      cn = 1;
      w.iwhere[i] = (nb == 2 && w.up[i] - w.lo[i] <= 0.0) ? 3 : 0;
    }
  }
  prj = c.any(prj);
  cn = c.any(cn);
  notbox = c.any(notbox);
  c.sync();
  w.s->prjctd = prj;
  w.s->cnstnd = cn;
  w.s->boxed = !notbox;
  c.sync();
}

// generalized Cauchy point z along the projected steepest-descent path; on
// return w.c = W^T (z - x).  Nonzero return: singular middle matrix.
template <class C>
__host__ __device__ inline int cauchy(C& c, View& w) {
  Scal& s = *w.s;
  const int n = w.n, m = w.m, col = s.col, col2 = 2 * col;
  const double theta = s.theta;
  for (int k = c.lane(); k < col2; k += C::W) w.c[k] = 0.0;
  if (s.sbgnrm <= 0.0) {
    for (int i = c.lane(); i < n; i += C::W) w.z[i] = w.x[i];
    c.sync();
    return 0;
  }
  double f1p = 0.0;
  int unb = 0, nfm = 0, nbreak = 0;
  for (int base = 0; base < n; base += C::W) {
    const int i = base + c.lane();
    int isbrk = 0;
    double tb = 0.0;
    if (i < n) {
      const double neggi = -w.g[i];
      const int nb = w.nbd[i];
      double tl = 0.0, tu = 0.0;
      int iw = w.iwhere[i];
      if (iw != 3 && iw != -1) {
        if (nb <= 2) tl = w.x[i] - w.lo[i];
        if (nb >= 2) tu = w.up[i] - w.x[i];
        const bool xlower = nb <= 2 && tl <= 0.0;
        const bool xupper = nb >= 2 && tu <= 0.0;
        iw = 0;
        if (xlower) {
          if (neggi <= 0.0) iw = 1;
        } else if (xupper) {
          if (neggi >= 0.0) iw = 2;
        } else {
          if (fabs(neggi) <= 0.0) iw = -3;
        }
        w.iwhere[i] = iw;
      }
      if (iw != 0 && iw != -1) {
        w.d[i] = 0.0;
      } else {
        w.d[i] = neggi;
        f1p -= neggi * neggi;
        if (nb <= 2 && nb != 0 && neggi < 0.0) {
          isbrk = 1;
          tb = tl / (-neggi);
        } else if (nb >= 2 && neggi > 0.0) {
          isbrk = 1;
          tb = tu / neggi;
        } else {
          nfm++;
          if (fabs(neggi) > 0.0) unb = 1;
        }
      }
    }
    int tot = 0;
    const int pos = c.ballot_rank(isbrk, tot);
    if (isbrk) {
      w.bt[nbreak + pos] = tb;
      w.bi[nbreak + pos] = i;
    }
    nbreak += tot;
  }
  double f1 = c.sum(f1p);
  const int bnded = !c.any(unb);
  nfm = c.sum_int(nfm);
  c.sync();
  // p = W^T d over the variables that move, W = [Y, theta S]
  for (int k = c.lane(); k < col2; k += C::W) {
    const int j = k < col ? k : k - col;
    const int slot = (s.head + j) % m;
    const double* wc = (k < col ? w.wy : w.ws) + (size_t)slot * n;
    double acc = 0.0;
    for (int i = 0; i < n; ++i) acc += wc[i] * w.d[i];
    w.p[k] = k < col ? acc : theta * acc;
  }
  for (int i = c.lane(); i < n; i += C::W) w.z[i] = w.x[i];
  c.sync();
  if (nbreak == 0 && nfm == 0) return 0;
  // the breakpoint walk (uniform)
  double f2 = -theta * f1;
  const double f2_org = f2;
  if (col > 0) {
    if (bmv(c, w, col, w.p, w.v)) return -1;
    double vp = 0.0;
    for (int k = c.lane(); k < col2; k += C::W) vp += w.v[k] * w.p[k];
    f2 -= c.sum(vp);
  }
  double dtm = -f1 / f2, tsum = 0.0;
  bool all_fixed = false;
  if (nbreak > 0) {
    int nleft = nbreak, it = 1, ibkmin = 0;
    double tj = 0.0, bkmin = w.bt[0];
    for (int k = 1; k < nbreak; ++k)
      if (w.bt[k] < bkmin) {
        bkmin = w.bt[k];
        ibkmin = k;
      }
    for (;;) {
      const double tj0 = tj;
      int ibp;
      if (it == 1) {
        tj = bkmin;
        ibp = w.bi[ibkmin];
      } else {
        if (it == 2 && ibkmin != nbreak - 1) {
          // the least breakpoint is passed: the last one takes its place
          w.bt[ibkmin] = w.bt[nbreak - 1];
          w.bi[ibkmin] = w.bi[nbreak - 1];
        }
        hpsolb(nleft, w.bt, w.bi, it - 2);
        tj = w.bt[nleft - 1];
        ibp = w.bi[nleft - 1];
      }
      const double dt = tj - tj0;
      if (dtm < dt) break;  // the minimiser lies in this segment
      tsum += dt;
      nleft--;
      it++;
      const double dibp = w.d[ibp];
      w.d[ibp] = 0.0;
      double zibp;
      if (dibp > 0.0) {
        zibp = w.up[ibp] - w.x[ibp];
        w.z[ibp] = w.up[ibp];
        w.iwhere[ibp] = 2;
      } else {
        zibp = w.lo[ibp] - w.x[ibp];
        w.z[ibp] = w.lo[ibp];
        w.iwhere[ibp] = 1;
      }
      if (nleft == 0 && nbreak == n) {
        // every variable is at a bound
        dtm = dt;
        all_fixed = true;
        break;
      }
      const double dibp2 = dibp * dibp;
      f1 = f1 + dt * f2 + dibp2 - theta * dibp * zibp;
      f2 = f2 - theta * dibp2;
      if (col > 0) {
        // c += dt p; wbp = the row of W = [Y, theta S] of the breakpoint's variable
        for (int k = c.lane(); k < col2; k += C::W) {
          w.c[k] += dt * w.p[k];
          const int slot = (s.head + (k < col ? k : k - col)) % m;
          w.wbp[k] = k < col ? w.wy[(size_t)slot * n + ibp]
                             : theta * w.ws[(size_t)slot * n + ibp];
        }
        c.sync();
        if (bmv(c, w, col, w.wbp, w.v)) return -1;
        double wmc = 0.0, wmp = 0.0, wmw = 0.0;
        for (int k = c.lane(); k < col2; k += C::W) {
          wmc += w.v[k] * w.c[k];
          wmp += w.v[k] * w.p[k];
          wmw += w.v[k] * w.wbp[k];
          w.p[k] -= dibp * w.wbp[k];
        }
        wmc = c.sum(wmc);
        wmp = c.sum(wmp);
        wmw = c.sum(wmw);
        c.sync();
        f1 += dibp * wmc;
        f2 += 2.0 * dibp * wmp - dibp2 * wmw;
      }
      f2 = fmax(EPSMCH * f2_org, f2);
      if (nleft > 0) {
        dtm = -f1 / f2;
        continue;
      }
      if (bnded) {
        f1 = 0.0;
        f2 = 0.0;
        dtm = 0.0;
      } else {
        dtm = -f1 / f2;
      }
      break;
    }
  }
  c.sync();
  if (!all_fixed) {
    if (dtm <= 0.0) dtm = 0.0;
    tsum += dtm;
    for (int i = c.lane(); i < n; i += C::W) w.z[i] += tsum * w.d[i];
  }
  for (int k = c.lane(); k < col2; k += C::W) w.c[k] += dtm * w.p[k];
  c.sync();
  return 0;
}

// free set at the Cauchy point (iwhere <= 0) and its complement as index
// lists, and the variables that entered or left the free set since the
// previous split (counted from the second iteration on, constrained problems
// only); all lists in increasing index
template <class C>
__host__ __device__ inline void freev(C& c, View& w) {
  Scal& s = *w.s;
  const bool count = s.iter > 0 && s.cnstnd;
  int nf = 0, na = 0, ne = 0, nl = 0;
  for (int base = 0; base < w.n; base += C::W) {
    const int i = base + c.lane();
    const bool in = i < w.n;
    const int fr = in && w.iwhere[i] <= 0;
    int ent = 0, lv = 0;
    if (in) {
      if (count) {
        lv = w.freeprev[i] && !fr;
        ent = !w.freeprev[i] && fr;
      }
      w.freeprev[i] = fr;
    }
    int t;
    int pos = c.ballot_rank(fr, t);
    if (fr) w.lfree[nf + pos] = i;
    nf += t;
    pos = c.ballot_rank(in && !fr, t);
    if (in && !fr) w.lact[na + pos] = i;
    na += t;
    pos = c.ballot_rank(ent, t);
    if (ent) w.lent[ne + pos] = i;
    ne += t;
    pos = c.ballot_rank(lv, t);
    if (lv) w.llv[nl + pos] = i;
    nl += t;
  }
  c.sync();
  s.nfree = nf;
  s.nenter = ne;
  s.nleave = nl;
  c.sync();
}

// sum over the listed variables of a[k] b[k], in increasing index (up) or
// decreasing index (the order in which the published code walks its lists of
// active, entering and leaving variables)
__host__ __device__ inline double list_dot(const double* a, const double* b,
                                           const int32_t* list, int cnt, bool up) {
  double acc = 0.0;
  if (up) {
    for (int t = 0; t < cnt; ++t) acc += a[list[t]] * b[list[t]];
  } else {
    for (int t = cnt - 1; t >= 0; --t) acc += a[list[t]] * b[list[t]];
  }
  return acc;
}

// WN1 (2m x 2m, lower triangles of the diagonal blocks and the full (2,1)
// block, rows and columns by pair position, oldest first):
//   (1,1) Y^T Z Z^T Y,  (2,2) S^T A A^T S,  (2,1) L_a + R_z
// with L_a the strictly lower part of S^T A A^T Y and R_z the upper part of
// S^T Z Z^T Y.  Then the LEL^T factorisation of
//   K = [[-D - Y^T Z Z^T Y / theta, L_a^T - R_z^T], [L_a - R_z, theta S^T A A^T S]]
// into wn: the upper triangle of [[L^T, L^{-1}(-L_a^T + R_z^T)], [0, J^T]] with
// L L^T = D + Y^T Z Z^T Y / theta and J J^T = theta S^T A A^T S + (..)^T (..).
// Returns -1 / -2 when a block is not positive definite.
template <class C>
__host__ __device__ inline int formk(C& c, View& w) {
  Scal& s = *w.s;
  const int n = w.n, m = w.m, col = s.col, ld = 2 * m;
  const double theta = s.theta;
  double* wn1 = w.wn1;
  const int nf = s.nfree, na = n - s.nfree, ne = s.nenter, nl = s.nleave;
  int upcl = col;
  if (s.updatd) {
    if (s.iupdat > m) {
      // the oldest pair was dropped: entry (r, q) of the old part of WN1 - the
      // lower triangles of the diagonal blocks and the (2,1) block, r, q below
      // the last row / column of their block - takes entry (r + 1, q + 1)
      // (every source is read before any target is written)
      // (wn, rewritten below, holds the sources)
      const int m1 = m - 1;
      double* tmp = w.wn;
      for (int e = c.lane(); e < 3 * m1 * m1; e += C::W) {
        const int blk = e / (m1 * m1), rq = e % (m1 * m1), r = rq / m1, q = rq % m1;
        const int R = blk == 0 ? r : m + r, Q = blk == 1 ? m + q : q;
        tmp[e] = (blk == 2 || r >= q) ? wn1[(R + 1) * ld + Q + 1] : 0.0;
      }
      c.sync();
      for (int e = c.lane(); e < 3 * m1 * m1; e += C::W) {
        const int blk = e / (m1 * m1), rq = e % (m1 * m1), r = rq / m1, q = rq % m1;
        const int R = blk == 0 ? r : m + r, Q = blk == 1 ? m + q : q;
        if (blk == 2 || r >= q) wn1[R * ld + Q] = tmp[e];
      }
    }
    c.sync();
    // the newest pair's row in blocks (1,1), (2,2), (2,1) and its column in (2,1)
    const int iy = col - 1, pnew = (s.head + col - 1) % m;
    const double* ynew = w.wy + (size_t)pnew * n;
    const double* snew = w.ws + (size_t)pnew * n;
    for (int e = c.lane(); e < 4 * col; e += C::W) {
      const int kind = e / col, jy = e % col, pj = (s.head + jy) % m;
      const double* yj = w.wy + (size_t)pj * n;
      const double* sj = w.ws + (size_t)pj * n;
      if (kind == 0) {
        wn1[iy * ld + jy] = list_dot(ynew, yj, w.lfree, nf, true);
      } else if (kind == 1) {
        wn1[(m + iy) * ld + m + jy] = list_dot(snew, sj, w.lact, na, false);
      } else if (kind == 2) {
        if (jy < col - 1) wn1[(m + iy) * ld + jy] = list_dot(snew, yj, w.lact, na, false);
      } else {
        wn1[(m + jy) * ld + iy] = list_dot(sj, ynew, w.lfree, nf, true);
      }
    }
    c.sync();
    upcl = col - 1;
  }
  if (s.nenter > 0 || s.nleave > 0) {
    // the old rows: add the entering and remove the leaving variables
    for (int e = c.lane(); e < upcl * upcl; e += C::W) {
      const int iy = e / upcl, jy = e % upcl;
      const int pi = (s.head + iy) % m, pj = (s.head + jy) % m;
      const double* yi = w.wy + (size_t)pi * n;
      const double* yj = w.wy + (size_t)pj * n;
      const double* si = w.ws + (size_t)pi * n;
      const double* sj = w.ws + (size_t)pj * n;
      if (jy <= iy) {
        const double t1 = list_dot(yi, yj, w.lent, ne, false);
        const double t2 = list_dot(si, sj, w.lent, ne, false);
        const double t3 = list_dot(yi, yj, w.llv, nl, false);
        const double t4 = list_dot(si, sj, w.llv, nl, false);
        wn1[iy * ld + jy] = wn1[iy * ld + jy] + t1 - t3;
        wn1[(m + iy) * ld + m + jy] = wn1[(m + iy) * ld + m + jy] - t2 + t4;
      }
      // block (2,1): row iy (S), column jy (Y)
      const double t1 = list_dot(si, yj, w.lent, ne, false);
      const double t3 = list_dot(si, yj, w.llv, nl, false);
      if (iy <= jy)
        wn1[(m + iy) * ld + jy] = wn1[(m + iy) * ld + jy] + t1 - t3;
      else
        wn1[(m + iy) * ld + jy] = wn1[(m + iy) * ld + jy] - t1 + t3;
    }
    c.sync();
  }
  // WN = [[D + Y^T Z Z^T Y / theta, -L_a^T + R_z^T], [-L_a + R_z, theta S^T A A^T S]]
  const int c2 = 2 * col;
  for (int e = c.lane(); e < c2 * c2; e += C::W) {
    const int i = e / c2, j = e % c2;
    if (i > j) continue;
    double val;
    if (j < col) {
      val = wn1[j * ld + i] / theta;
      if (i == j) val = val + w.sy[i * m + i];
    } else if (i < col) {
      const int iy = j - col, jy = i;
      val = jy < iy ? -wn1[(m + iy) * ld + jy] : wn1[(m + iy) * ld + jy];
    } else {
      val = wn1[(m + j - col) * ld + m + i - col] * theta;
    }
    w.wn[i * ld + j] = val;
  }
  c.sync();
  const int info1 = c.chol_upper(w.wn, ld, col);
  c.sync();
  if (info1) return -1;
  // (1,2) block: L^{-1} times each of its columns
  for (int js = col + c.lane(); js < c2; js += C::W) {
    for (int j = 0; j < col; ++j) {
      double t = w.wn[j * ld + js];
      for (int k = 0; k < j; ++k) t -= w.wn[k * ld + j] * w.wn[k * ld + js];
      w.wn[j * ld + js] = t / w.wn[j * ld + j];
    }
  }
  c.sync();
  // (2,2) block += (1,2)^T (1,2)
  const int cc = col * col;
  for (int e = c.lane(); e < cc; e += C::W) {
    const int is = col + e / col, js = col + e % col;
    if (is > js) continue;
    double acc = 0.0;
    for (int k = 0; k < col; ++k) acc += w.wn[k * ld + is] * w.wn[k * ld + js];
    w.wn[is * ld + js] += acc;
  }
  c.sync();
  const int info2 = c.chol_upper(w.wn + col * ld + col, ld, col);
  c.sync();
  if (info2) return -2;
  return 0;
}

// reduced gradient at the Cauchy point over the free variables:
// rd = -Z^T (B (z - x) + g)
template <class C>
__host__ __device__ inline int cmprlb(C& c, View& w) {
  Scal& s = *w.s;
  const int n = w.n, m = w.m, col = s.col;
  if (!s.cnstnd && col > 0) {
    for (int i = c.lane(); i < n; i += C::W) w.rd[i] = -w.g[i];
    c.sync();
    return 0;
  }
  const int info = bmv(c, w, col, w.c, w.v);
  c.sync();
  if (info) return -8;
  const double theta = s.theta;
  for (int i = c.lane(); i < n; i += C::W) {
    if (w.iwhere[i] > 0) continue;
    double r = -theta * (w.z[i] - w.x[i]) - w.g[i];
    for (int j = 0; j < col; ++j) {
      const int slot = (s.head + j) % m;
      r += w.wy[(size_t)slot * n + i] * w.v[j] + w.ws[(size_t)slot * n + i] * (theta * w.v[col + j]);
    }
    w.rd[i] = r;
  }
  c.sync();
  return 0;
}

// subspace minimisation over the free variables from the Cauchy point z, then
// projection onto the box (backtracking along the step when the projected step
// is not a descent direction).  Updates z.  Nonzero return: singular factor.
template <class C>
__host__ __device__ inline int subsm(C& c, View& w) {
  Scal& s = *w.s;
  const int n = w.n, m = w.m, col = s.col, c2 = 2 * col, ld = 2 * m;
  const double theta = s.theta;
  if (s.nfree <= 0) return 0;
  // wv = W^T Z rd
  for (int k = c.lane(); k < c2; k += C::W) {
    const int j = k < col ? k : k - col;
    const int slot = (s.head + j) % m;
    const double* wc = (k < col ? w.wy : w.ws) + (size_t)slot * n;
    const double acc = list_dot(wc, w.rd, w.lfree, s.nfree, true);
    w.wv[k] = k < col ? acc : theta * acc;
  }
  c.sync();
  // wv = K^{-1} wv
  int info = c.solve_rt(w.wn, ld, c2, w.wv);
  if (!info) {
    for (int k = c.lane(); k < col; k += C::W) w.wv[k] = -w.wv[k];
    c.sync();
    info = c.solve_r(w.wn, ld, c2, w.wv);
  }
  c.sync();
  if (info) return info;
  // rd = (rd + Z^T W wv / theta-scaled) / theta
  for (int i = c.lane(); i < n; i += C::W) {
    if (w.iwhere[i] > 0) continue;
    double di = w.rd[i];
    for (int j = 0; j < col; ++j) {
      const int slot = (s.head + j) % m;
      di += w.wy[(size_t)slot * n + i] * w.wv[j] / theta + w.ws[(size_t)slot * n + i] * w.wv[col + j];
    }
    w.rd[i] = di * (1.0 / theta);
  }
  // projection of z + rd onto the box
  int hit = 0;
  for (int i = c.lane(); i < n; i += C::W) {
    w.xp[i] = w.z[i];
    if (w.iwhere[i] > 0) continue;
    const double dk = w.rd[i], xk = w.z[i];
    const int nb = w.nbd[i];
    double xn;
    if (nb == 1) {
      xn = fmax(w.lo[i], xk + dk);
      if (xn == w.lo[i]) hit = 1;
    } else if (nb == 2) {
      xn = fmin(w.up[i], fmax(w.lo[i], xk + dk));
      if (xn == w.lo[i] || xn == w.up[i]) hit = 1;
    } else if (nb == 3) {
      xn = fmin(w.up[i], xk + dk);
      if (xn == w.up[i]) hit = 1;
    } else {
      xn = xk + dk;
    }
    w.z[i] = xn;
  }
  hit = c.any(hit);
  if (!hit) {
    c.sync();
    return 0;
  }
  double ddp = 0.0;
  for (int i = c.lane(); i < n; i += C::W) ddp += (w.z[i] - w.x[i]) * w.g[i];
  ddp = c.sum(ddp);
  if (!(ddp > 0.0)) {
    c.sync();
    return 0;
  }
  // the projected step ascends: back to the Cauchy point and truncate the step
  // at the first bound it meets
  double amin = 1.0;
  for (int i = c.lane(); i < n; i += C::W) {
    w.z[i] = w.xp[i];
    if (w.iwhere[i] > 0) continue;
    const double dk = w.rd[i];
    const int nb = w.nbd[i];
    if (nb == 0) continue;
    double cand = 1.0;
    if (dk < 0.0 && nb <= 2) {
      const double t2 = w.lo[i] - w.z[i];
      if (t2 >= 0.0)
        cand = 0.0;
      else if (dk < t2)
        cand = t2 / dk;
    } else if (dk > 0.0 && nb >= 2) {
      const double t2 = w.up[i] - w.z[i];
      if (t2 <= 0.0)
        cand = 0.0;
      else if (dk > t2)
        cand = t2 / dk;
    }
    amin = fmin(amin, cand);
  }
  const double alpha = c.min(amin);
  int ibd = n;
  if (alpha < 1.0) {
    for (int i = c.lane(); i < n; i += C::W) {
      if (w.iwhere[i] > 0 || w.nbd[i] == 0) continue;
      const double dk = w.rd[i];
      double cand = 1.0;
      if (dk < 0.0 && w.nbd[i] <= 2) {
        const double t2 = w.lo[i] - w.z[i];
        cand = t2 >= 0.0 ? 0.0 : (dk < t2 ? t2 / dk : 1.0);
      } else if (dk > 0.0 && w.nbd[i] >= 2) {
        const double t2 = w.up[i] - w.z[i];
        cand = t2 <= 0.0 ? 0.0 : (dk > t2 ? t2 / dk : 1.0);
      }
      if (cand == alpha && i < ibd) ibd = i;
    }
    ibd = c.min_int(ibd);
  }
  c.sync();
  if (alpha < 1.0 && ibd < n) {
    const double dk = w.rd[ibd];
    c.sync();
    if (dk > 0.0) {
      w.z[ibd] = w.up[ibd];
      w.rd[ibd] = 0.0;
    } else if (dk < 0.0) {
      w.z[ibd] = w.lo[ibd];
      w.rd[ibd] = 0.0;
    }
  }
  c.sync();
  for (int i = c.lane(); i < n; i += C::W)
    if (w.iwhere[i] <= 0) w.z[i] += alpha * w.rd[i];
  c.sync();
  return 0;
}

// start of the line search along d = z - x: step bound and first trial step
template <class C>
__host__ __device__ inline void lnsrch_start(C& c, View& w) {
  Scal& s = *w.s;
  const int n = w.n;
  double dtd = 0.0, smx = STPMAX_FREE;
  for (int i = c.lane(); i < n; i += C::W) {
    const double a1 = w.d[i];
    dtd += a1 * a1;
    w.t[i] = w.x[i];
    w.r[i] = w.g[i];
    const int nb = w.nbd[i];
    if (!s.cnstnd || s.iter == 0 || nb == 0) continue;
    if (a1 < 0.0 && nb <= 2) {
      const double a2 = w.lo[i] - w.x[i];
      if (a2 >= 0.0)
        smx = 0.0;
      else if (a1 * smx < a2)
        smx = a2 / a1;
    } else if (a1 > 0.0 && nb >= 2) {
      const double a2 = w.up[i] - w.x[i];
      if (a2 <= 0.0)
        smx = 0.0;
      else if (a1 * smx > a2)
        smx = a2 / a1;
    }
  }
  dtd = c.sum(dtd);
  smx = c.min(smx);
  c.sync();
  s.dtd = dtd;
  s.dnorm = sqrt(dtd);
  s.stpmx = s.cnstnd ? (s.iter == 0 ? 1.0 : smx) : STPMAX_FREE;
  s.stp = (s.iter == 0 && !s.boxed) ? fmin(1.0 / s.dnorm, s.stpmx) : 1.0;
  s.fold = s.f;
  s.ifun = 0;
  s.iback = 0;
  s.ls_task = LS_START;
  c.sync();
}

// one line-search step at the current f, g: returns 1 when a trial point was
// written to x (evaluate it), 0 when the search ended (ls_task CONV / WARN),
// -4 when d is not a descent direction
template <class C>
__host__ __device__ inline int lnsrch_step(C& c, View& w) {
  Scal& s = *w.s;
  const int n = w.n;
  double gd = 0.0;
  for (int i = c.lane(); i < n; i += C::W) gd += w.g[i] * w.d[i];
  gd = c.sum(gd);
  c.sync();
  s.gd = gd;
  if (s.ifun == 0) {
    s.gdold = gd;
    if (gd >= 0.0) {
      c.sync();
      return -4;
    }
  }
  double stp = s.stp;
  dcsrch(s, s.f, gd, stp, 0.0, s.stpmx);
  s.stp = stp;
  c.sync();
  if (s.ls_task == LS_CONV || s.ls_task == LS_WARN) return 0;
  s.ifun += 1;
  s.nfgv += 1;
  s.iback = s.ifun - 1;
  c.sync();
  if (stp == 1.0) {
    for (int i = c.lane(); i < n; i += C::W) w.x[i] = w.z[i];
  } else {
    for (int i = c.lane(); i < n; i += C::W) w.x[i] = stp * w.d[i] + w.t[i];
  }
  c.sync();
  return 1;
}

// appends the pair (s, y) = (d, r) to the limited memory and updates S^T S, S^T Y
template <class C>
__host__ __device__ inline void matupd(C& c, View& w, double rr, double dr) {
  Scal& s = *w.s;
  const int n = w.n, m = w.m;
  if (s.iupdat <= m) {
    s.col = s.iupdat;
    s.itail = (s.head + s.iupdat - 1) % m;
  } else {
    s.itail = (s.itail + 1) % m;
    s.head = (s.head + 1) % m;
  }
  const int col = s.col, itail = s.itail;
  c.sync();
  for (int i = c.lane(); i < n; i += C::W) {
    w.ws[(size_t)itail * n + i] = w.d[i];
    w.wy[(size_t)itail * n + i] = w.r[i];
  }
  s.theta = rr / dr;
  if (s.iupdat > m) {
    // drop the oldest pair: entry (i, j) of the upper triangle of S^T S and of
    // the lower triangle of S^T Y takes entry (i + 1, j + 1), for i, j < col - 1
    // (every source is read before any target is written)
    // (wn, free until the next subspace step, holds the sources)
    const int c1 = col - 1;
    double* tmp = w.wn;
    for (int e = c.lane(); e < c1 * c1; e += C::W) {
      const int i = e / c1, j = e % c1;
      tmp[e] = i <= j ? w.ss[(i + 1) * m + j + 1] : 0.0;
      tmp[c1 * c1 + e] = i >= j ? w.sy[(i + 1) * m + j + 1] : 0.0;
    }
    c.sync();
    for (int e = c.lane(); e < c1 * c1; e += C::W) {
      const int i = e / c1, j = e % c1;
      if (i <= j) w.ss[i * m + j] = tmp[e];
      if (i >= j) w.sy[i * m + j] = tmp[c1 * c1 + e];
    }
  }
  c.sync();
  // the new last row of S^T Y and last column of S^T S
  for (int k = c.lane(); k < 2 * (col - 1); k += C::W) {
    const int j = k < col - 1 ? k : k - (col - 1);
    const int slot = (s.head + j) % m;
    double acc = 0.0;
    if (k < col - 1) {
      for (int i = 0; i < n; ++i) acc += w.d[i] * w.wy[(size_t)slot * n + i];
      w.sy[(col - 1) * m + j] = acc;
    } else {
      for (int i = 0; i < n; ++i) acc += w.ws[(size_t)slot * n + i] * w.d[i];
      w.ss[j * m + (col - 1)] = acc;
    }
  }
  c.sync();
  w.ss[(col - 1) * m + (col - 1)] = s.stp == 1.0 ? s.dtd : s.stp * s.stp * s.dtd;
  w.sy[(col - 1) * m + (col - 1)] = dr;
  c.sync();
}

// T = theta S^T S + L D^{-1} L^T in the upper triangle of wt, then its upper
// Cholesky factor J^T; nonzero return: T not positive definite
template <class C>
__host__ __device__ inline int formt(C& c, View& w) {
  Scal& s = *w.s;
  const int m = w.m, col = s.col;
  const double theta = s.theta;
  for (int e = c.lane(); e < col * col; e += C::W) {
    const int i = e / col, j = e % col;
    if (i > j) continue;
    double ddum = 0.0;
    for (int k = 0; k < i; ++k) ddum += w.sy[i * m + k] * w.sy[j * m + k] / w.sy[k * m + k];
    w.wt[i * m + j] = ddum + theta * w.ss[i * m + j];
  }
  c.sync();
  const int info = c.chol_upper(w.wt, m, col);
  c.sync();
  return info ? -3 : 0;
}

__host__ __device__ inline void refresh(Scal& s) {
  s.info = 0;
  s.col = 0;
  s.head = 0;
  s.theta = 1.0;
  s.iupdat = 0;
  s.updatd = 0;
}

enum MachineResult { MR_FG = 1, MR_NEWX = 2, MR_DONE = 3 };

// runs the machine from s.pc until it needs f, g at w.x (MR_FG), accepted a
// new iterate (MR_NEWX) or ended (MR_DONE, s.task set)
template <class C>
__host__ __device__ inline int machine(C& c, View& w) {
  Scal& s = *w.s;
  const int n = w.n;
  for (;;) {
    const int pc = s.pc;
    c.sync();
    if (pc == PC_START) {
      s.col = 0;
      s.head = 0;
      s.theta = 1.0;
      s.iupdat = 0;
      s.updatd = 0;
      s.iback = 0;
      s.itail = 0;
      s.ifun = 0;
      s.iter = 0;
      s.nfgv = 0;
      s.nskip = 0;
      s.nfree = n;
      s.info = 0;
      s.fold = 0.0;
      s.stp = 0.0;
      s.gd = 0.0;
      s.gdold = 0.0;
      s.dtd = 0.0;
      s.sbgnrm = 0.0;
      c.sync();
      active(c, w);
      s.pc = PC_FG0;
      c.sync();
      return MR_FG;
    }
    if (pc == PC_FG0) {
      s.nfgv = 1;
      const double sb = projgr(c, w);
      c.sync();
      s.sbgnrm = sb;
      if (sb <= s.pgtol) {
        s.task = TASK_CONV_PGTOL;
        s.pc = PC_DONE;
        c.sync();
        return MR_DONE;
      }
      s.pc = PC_ITER;
      c.sync();
      continue;
    }
    if (pc == PC_ITER) {
      int info = 0;
      if (!s.cnstnd && s.col > 0) {
        for (int i = c.lane(); i < n; i += C::W) {
          w.z[i] = w.x[i];
          w.lfree[i] = i;
        }
        c.sync();
        s.nfree = n;
        s.nenter = 0;
        s.nleave = 0;
        c.sync();
      } else {
        c.mark(PH_CAUCHY);
        info = cauchy(c, w);
        c.sync();
        if (info) {
          refresh(s);
          c.sync();
          continue;
        }
        c.mark(PH_FREEV);
        freev(c, w);
      }
      if (s.nfree != 0 && s.col != 0) {
        c.mark(PH_FORMK);
        info = formk(c, w);
        c.mark(PH_SUBSM);
        if (!info) info = cmprlb(c, w);
        if (!info) info = subsm(c, w);
        c.sync();
        if (info) {
          refresh(s);
          c.sync();
          continue;
        }
      }
      for (int i = c.lane(); i < n; i += C::W) w.d[i] = w.z[i] - w.x[i];
      c.sync();
      c.mark(PH_LNSRCH);
      lnsrch_start(c, w);
      s.pc = PC_LS_FG;
      c.sync();
      // fall through to the first line-search step at the current f, g
    }
    if (s.pc == PC_LS_FG) {
      c.mark(PH_LNSRCH);
      const int r = lnsrch_step(c, w);
      c.sync();
      const int info = r < 0 ? r : 0;
      if (info != 0 || (r == 1 && s.iback >= s.maxls)) {
        // restore the previous iterate
        for (int i = c.lane(); i < n; i += C::W) {
          w.x[i] = w.t[i];
          w.g[i] = w.r[i];
        }
        c.sync();
        s.f = s.fold;
        if (s.col == 0) {
          if (info == 0) {
            s.nfgv -= 1;
            s.ifun -= 1;
            s.iback -= 1;
          }
          s.info = info == 0 ? -9 : info;
          s.task = TASK_ABNORMAL;
          s.iter += 1;
          s.pc = PC_DONE;
          c.sync();
          return MR_DONE;
        }
        if (info == 0) s.nfgv -= 1;
        refresh(s);
        s.pc = PC_ITER;
        c.sync();
        continue;
      }
      if (r == 1) return MR_FG;
      s.iter += 1;
      const double sb = projgr(c, w);
      c.sync();
      s.sbgnrm = sb;
      s.pc = PC_NEWX;
      c.sync();
      return MR_NEWX;
    }
    if (pc == PC_NEWX) {
      if (s.sbgnrm <= s.pgtol) {
        s.task = TASK_CONV_PGTOL;
        s.pc = PC_DONE;
        c.sync();
        return MR_DONE;
      }
      const double ddum0 = fmax(fmax(fabs(s.fold), fabs(s.f)), 1.0);
      if (s.fold - s.f <= s.factr * EPSMCH * ddum0) {
        s.task = TASK_CONV_FACTR;
        if (s.iback >= 10) s.info = -5;
        s.pc = PC_DONE;
        c.sync();
        return MR_DONE;
      }
      double rr = 0.0;
      for (int i = c.lane(); i < n; i += C::W) {
        const double yi = w.g[i] - w.r[i];
        w.r[i] = yi;
        rr += yi * yi;
      }
      rr = c.sum(rr);
      c.sync();
      double dr, ddum;
      if (s.stp == 1.0) {
        dr = s.gd - s.gdold;
        ddum = -s.gdold;
      } else {
        dr = (s.gd - s.gdold) * s.stp;
        for (int i = c.lane(); i < n; i += C::W) w.d[i] *= s.stp;
        ddum = -s.gdold * s.stp;
      }
      c.sync();
      if (dr <= EPSMCH * ddum) {
        // skip the update
        s.nskip += 1;
        s.updatd = 0;
        s.pc = PC_ITER;
        c.sync();
        continue;
      }
      s.updatd = 1;
      s.iupdat += 1;
      c.sync();
      c.mark(PH_UPDATE);
      matupd(c, w, rr, dr);
      const int info = formt(c, w);
      if (info) refresh(s);
      s.pc = PC_ITER;
      c.sync();
      continue;
    }
    // PC_DONE
    return MR_DONE;
  }
}

// the lock-step driver around the machine for one problem (the per-restart
// body of optim/batched_lbfgs_b.py:440-533): consumes f, g when the machine
// waits for them (nullptr: nothing to consume), counts iterations and
// evaluations, stops at maxiter / maxfun.  Returns 1 when w.x is to be
// evaluated, 0 when the problem is done.
template <class C>
__host__ __device__ inline int advance(C& c, View& w, const double* fval, const double* gval,
                                       double fscale, int32_t step_status) {
  Scal& s = *w.s;
  const int n = w.n;
  if (s.done) return 0;
  if ((s.pc == PC_FG0 || s.pc == PC_LS_FG) && fval) {
    int bad = 0;
    for (int i = c.lane(); i < n; i += C::W) {
      const double gi = fscale * gval[i];
      w.g[i] = gi;
      if (gi != gi) bad = 1;
    }
    bad = c.any(bad);
    c.sync();
    s.f = fscale * fval[0];
    s.nfev += 1;
    s.step_status |= step_status;
    if (bad) {
      s.task = TASK_NAN_GRADIENT;
      s.warnflag = 2;
      s.pc = PC_DONE;
      s.done = 1;
      c.sync();
      return 0;
    }
  }
  c.sync();
  for (;;) {
    const int r = machine(c, w);
    c.sync();
    if (r == MR_FG) {
      if (s.nfev >= s.maxfun) {
        s.task = TASK_STOP_MAXFUN;
        s.warnflag = 1;
        s.pc = PC_DONE;
        s.done = 1;
        c.sync();
        return 0;
      }
      return 1;
    }
    if (r == MR_NEWX) {
      s.nit += 1;
      if (s.nit >= s.maxiter) {
        s.task = TASK_STOP_MAXITER;
        s.warnflag = 1;
        s.pc = PC_DONE;
        s.done = 1;
        c.sync();
        return 0;
      }
      c.sync();
      continue;
    }
    s.warnflag = (s.task == TASK_CONV_PGTOL || s.task == TASK_CONV_FACTR) ? 0 : 2;
    s.done = 1;
    c.sync();
    return 0;
  }
}

// fresh state: settings, x = x0 (the caller copies x0 into w.x)
__host__ __device__ inline void init_scal(Scal& s, double factr, double pgtol, int maxls,
                                          int maxiter, int maxfun) {
  s.f = 0.0;
  s.factr = factr;
  s.pgtol = pgtol;
  s.maxls = maxls;
  s.maxiter = maxiter;
  s.maxfun = maxfun;
  s.pc = PC_START;
  s.task = TASK_RUNNING;
  s.nit = 0;
  s.nfev = 0;
  s.warnflag = 0;
  s.step_status = 0;
  s.done = 0;
  s.ls_task = LS_START;
  s.ls_brackt = 0;
  s.ls_stage = 1;
  s.ls_finit = s.ls_ginit = s.ls_gtest = 0.0;
  s.ls_stx = s.ls_fx = s.ls_gx = s.ls_sty = s.ls_fy = s.ls_gy = 0.0;
  s.ls_stmin = s.ls_stmax = s.ls_width = s.ls_width1 = 0.0;
}

}  // namespace lbfgsb
}  // namespace bta
