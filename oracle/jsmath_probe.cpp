// Test shim (oracle side): exposes the jsmath.h restatements so the math
// pinning tests can compare them with the C library.
#include <stdint.h>
#include "jsmath.h"

extern "C" void orc_math_eval(int fn, const double* x, double* out, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    switch (fn) {
      case 0: out[i] = orcm::log(x[i]); break;
      case 1: out[i] = orcm::acos(x[i]); break;
      case 2: out[i] = orcm::cbrt(x[i]); break;
      default: out[i] = orcm::cos(x[i]); break;
    }
  }
}

extern "C" double orc_cr_dot(const double* a, const double* b, int64_t n) {
  orcm::dd s = orcm::dd_zero();
  for (int64_t i = 0; i < n; i++) orcm::dd_add_prod(s, a[i], b[i]);
  return orcm::dd_value(s);
}
