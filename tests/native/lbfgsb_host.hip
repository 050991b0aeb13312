// Host build of the L-BFGS-B machine of botorch_amd/csrc/lbfgsb_core.h with a
// one-lane context, for tests that drive it step by step next to SciPy's
// setulb on the same functions.  Test infrastructure only: the library runs the
// same code on the GPU, one wavefront per problem.
#include <stdlib.h>
#include <string.h>

#include "../../botorch_amd/csrc/lbfgsb_core.h"

namespace {
using HostLanes = bta::lbfgsb::SerialLanes;

struct Problem {
  bta::lbfgsb::Layout L;
  double* mem;
  double *lo, *up;
  int32_t* nbd;
};
}  // namespace

extern "C" void* lbh_create(int32_t n, int32_t m, const double* x0, const double* lo,
                            const double* up, const int32_t* nbd, double factr, double pgtol,
                            int32_t maxls, int32_t maxiter, int32_t maxfun) {
  Problem* p = new Problem{bta::lbfgsb::Layout(n, m), nullptr, nullptr, nullptr, nullptr};
  p->mem = (double*)calloc((size_t)p->L.total, sizeof(double));
  p->lo = (double*)malloc(sizeof(double) * n);
  p->up = (double*)malloc(sizeof(double) * n);
  p->nbd = (int32_t*)malloc(sizeof(int32_t) * n);
  memcpy(p->lo, lo, sizeof(double) * n);
  memcpy(p->up, up, sizeof(double) * n);
  memcpy(p->nbd, nbd, sizeof(int32_t) * n);
  bta::lbfgsb::View w = bta::lbfgsb::make_view(p->mem, p->L, p->lo, p->up, p->nbd);
  bta::lbfgsb::init_scal(*w.s, factr, pgtol, maxls, maxiter, maxfun);
  memcpy(w.x, x0, sizeof(double) * n);
  return p;
}

extern "C" void lbh_destroy(void* h) {
  Problem* p = (Problem*)h;
  free(p->mem);
  free(p->lo);
  free(p->up);
  free(p->nbd);
  delete p;
}

// f, g: the values at the last point returned (ignored on the first call);
// writes the next point to evaluate into x; returns 1 while a point is to be
// evaluated, 0 when the problem is done
extern "C" int32_t lbh_advance(void* h, int32_t first, double f, const double* g, double* x) {
  Problem* p = (Problem*)h;
  HostLanes c;
  bta::lbfgsb::View w = bta::lbfgsb::make_view(p->mem, p->L, p->lo, p->up, p->nbd);
  const int r = bta::lbfgsb::advance(c, w, first ? nullptr : &f, first ? nullptr : g, 1.0, 0);
  memcpy(x, w.x, sizeof(double) * p->L.n);
  return r;
}

// info: nit, nfev, warnflag, task, iter, nfgv, nskip, col
extern "C" void lbh_result(void* h, double* f, int32_t* info) {
  Problem* p = (Problem*)h;
  const bta::lbfgsb::Scal& s = *(const bta::lbfgsb::Scal*)p->mem;
  *f = s.f;
  info[0] = s.nit;
  info[1] = s.nfev;
  info[2] = s.warnflag;
  info[3] = s.task;
  info[4] = s.iter;
  info[5] = s.nfgv;
  info[6] = s.nskip;
  info[7] = s.col;
}
