// Host-side check of exp_nonpos (botorch_amd/csrc/bta_common.h) against the C
// library's exp over the whole non-positive range the kernels use.  Prints the
// largest relative error (normal results) and absolute error (subnormal
// results); exits 1 if either exceeds its bound.
#include <cmath>
#include <cstdio>
#include <limits>

#include "../../botorch_amd/csrc/bta_common.h"

int main() {
  double max_rel = 0.0, max_abs_sub = 0.0;
  const int N = 2000000;
  for (int i = 0; i <= N; ++i) {
    const double x = -745.0 * (double)i / N;
    const double ref = std::exp(x), got = bta::exp_nonpos(x);
    if (ref >= 2.2250738585072014e-308) {
      const double rel = std::fabs(got - ref) / ref;
      if (rel > max_rel) max_rel = rel;
    } else {
      const double ab = std::fabs(got - ref);
      if (ab > max_abs_sub) max_abs_sub = ab;
    }
  }
  // small arguments, where f is tiny and the polynomial must return ~1 + x
  for (double x = -1e-300; x > -1.0; x *= 1.37) {
    const double rel = std::fabs(bta::exp_nonpos(x) - std::exp(x)) / std::exp(x);
    if (rel > max_rel) max_rel = rel;
  }
  const bool special = bta::exp_nonpos(0.0) == 1.0 && bta::exp_nonpos(-746.0) == 0.0 &&
                       bta::exp_nonpos(-1e308) == 0.0 &&
                       std::isnan(bta::exp_nonpos(std::numeric_limits<double>::quiet_NaN()));
  std::printf("max_rel %.3e max_abs_subnormal %.3e special %d\n", max_rel, max_abs_sub,
              (int)special);
  return (max_rel <= 4.5e-16 && max_abs_sub <= 1.5e-323 && special) ? 0 : 1;
}
