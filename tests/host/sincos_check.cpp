// Host check of the small-angle sine/cosine of quantum-jax_amd/csrc/qj_common.h
// (the kick phase factor exp(-i c V) after range reduction, |r| <= 1/2):
// against long-double sinl/cosl the error, in ulps of the correctly rounded
// value, stays below 1.1 for float (K = 4) and double (K = 7) over edge
// values and random arguments (the evaluation's rounding; the truncation is
// below half an ulp).  Prints the largest errors and exits 1 beyond 1.1.
#include <math.h>
#include <stdio.h>

#include <random>

#include "qj_common.h"

template <class T, int K>
static double worst(std::mt19937_64& rng, int n) {
  std::uniform_real_distribution<double> u(-0.5, 0.5);
  double w = 0.0;
  auto check = [&](T r) {
    T s, c;
    qj::sincos_small<T, K>(r, s, c);
    const long double rs = sinl((long double)r), rc = cosl((long double)r);
    const T ts = (T)rs, tc = (T)rc;
    const double ulp_s = (ts == 0) ? 0.0 : (double)fabsl(((long double)s - rs)) /
                         (double)((T)nextafter(fabs(ts), (T)INFINITY) - fabs(ts));
    const double ulp_c = (double)fabsl(((long double)c - rc)) /
                         (double)((T)nextafter(fabs(tc), (T)INFINITY) - fabs(tc));
    if (ulp_s > w) w = ulp_s;
    if (ulp_c > w) w = ulp_c;
  };
  const double edges[] = {0.0, -0.0, 0.5, -0.5, 1e-30, -1e-30, 1e-8, 0.25, 0.125, 0.49999999};
  for (double e : edges) check((T)e);
  for (int i = 0; i < n; ++i) check((T)u(rng));
  return w;
}

int main() {
  std::mt19937_64 rng(7);
  const double wf = worst<float, 4>(rng, 2000000);
  const double wd = worst<double, 7>(rng, 2000000);
  printf("float K=4: max %.3f ulp; double K=7: max %.3f ulp\n", wf, wd);
  return (wf <= 1.1 && wd <= 1.1) ? 0 : 1;
}
