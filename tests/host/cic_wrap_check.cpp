// Host check of the CIC wrap helpers of quantum-jax_amd/csrc/qj_cic.h
// (compiled for the host, no GPU needed):
//  * fmod_floor's fast path (x or x - L for -L < x < 2L) against the
//    library fmod followed by the same floored correction, bit for bit;
//  * imod's low-bit form for power-of-two n against the 64-bit remainder.
// Prints the number of mismatches and exits 1 if there is any.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <random>

#include "qj_cic.h"

static double generic_floor_mod(double x, double L) {
  double r = fmod(x, L);
  if (r != 0.0) {
    if ((r < 0.0) != (L < 0.0)) r += L;
  } else {
    r = copysign(0.0, L);
  }
  return r;
}

static int generic_imod(double fi, int n) {
  long long ii = (long long)fi;
  long long r = ii % n;
  return (int)(r < 0 ? r + n : r);
}

static bool same_bits(double a, double b) {
  if (a != a && b != b) return true;
  uint64_t x, y;
  memcpy(&x, &a, 8);
  memcpy(&y, &b, 8);
  return x == y;
}

int main() {
  long long bad = 0, n = 0;
  std::mt19937_64 rng(12345);
  const double Ls[] = {1.0, 0.5, 2.0, 3.0, 0.1, 1e-3, 7.25, 1e5};
  for (double L : Ls) {
    const double edges[] = {0.0, -0.0, L, -L, 2 * L, nextafter(L, 0.0), nextafter(L, 4 * L),
                            nextafter(-L, 0.0), nextafter(-L, -4 * L), nextafter(2 * L, 0.0),
                            nextafter(0.0, 1.0), nextafter(0.0, -1.0), 1e-300, -1e-300,
                            3 * L, -3 * L, 1e30, -1e30, NAN, INFINITY, -INFINITY, 0.5 * L};
    for (double x : edges) {
      ++n;
      if (!same_bits(qj::fmod_floor(x, L), generic_floor_mod(x, L))) {
        ++bad;
        printf("fmod_floor mismatch x=%a L=%a\n", x, L);
      }
    }
    std::uniform_real_distribution<double> u(-1.5 * L, 2.5 * L);
    for (int i = 0; i < 2000000; ++i) {
      double x = u(rng);
      if (i % 4 == 1) x = L + (x - L) * 1e-12;   // near L
      if (i % 4 == 2) x = x * 1e-12;             // near 0
      ++n;
      if (!same_bits(qj::fmod_floor(x, L), generic_floor_mod(x, L))) {
        if (++bad < 10) printf("fmod_floor mismatch x=%a L=%a\n", x, L);
      }
    }
  }
  for (int nn = 8; nn <= 512; nn *= 2) {
    for (long long k = -5000; k <= 5000; ++k) {
      ++n;
      if (qj::imod((double)k, nn) != generic_imod((double)k, nn)) ++bad;
    }
    const double big[] = {-4.0e18 + 4096, 3.9e18, -1.0e15, 1.0e15, -(double)(1LL << 53)};
    for (double f : big) {
      ++n;
      if (qj::imod(f, nn) != generic_imod(f, nn)) ++bad;
    }
  }
  for (int nn : {12, 100, 7}) {  // the general form
    for (long long k = -1000; k <= 1000; ++k) {
      ++n;
      if (qj::imod((double)k, nn) != generic_imod((double)k, nn)) ++bad;
    }
  }
  printf("%lld checks, %lld mismatches\n", n, bad);
  return bad ? 1 : 0;
}
