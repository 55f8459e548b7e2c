// Host-side check of the device remainder used by the minimum image
// (rem_exact / f_modulo in tgx_device.h) against the C library's fmod:
// random operands, exact and near multiples of the period, and both signs.
// Prints the number of mismatches (0 expected) and the number of cases.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "tgx_device.h"

static uint64_t state = 0x9e3779b97f4a7c15ull;
static double uni() {  // splitmix64 -> [0, 1)
  uint64_t z = (state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

static bool same(double x, double y) { return std::memcmp(&x, &y, 8) == 0; }

static double f_modulo_ref(double a, double p) {
  double r = std::fmod(a, p);
  if (r != 0.0) {
    if ((a < 0.0) != (p < 0.0)) r = r + p;
  } else {
    r = std::copysign(0.0, p);
  }
  return r;
}

int main() {
  long bad = 0, n = 0;
  for (int t = 0; t < 2000000; t++) {
    const double p = 1.0 + 199.0 * uni();
    double a;
    switch (t % 6) {
      case 0: a = (uni() - 0.5) * 2000.0 * p; break;               // many periods
      case 1: a = (uni() - 0.5) * 2.0 * p; break;                  // within one period
      case 2: a = std::floor((uni() - 0.5) * 2e4) * p; break;      // rounded multiples
      case 3: {                                                    // ulps off a multiple
        double m = std::floor((uni() - 0.5) * 2e4) * p;
        int k = (int)(uni() * 9) - 4;
        for (; k > 0; k--) m = std::nextafter(m, INFINITY);
        for (; k < 0; k++) m = std::nextafter(m, -INFINITY);
        a = m;
        break;
      }
      case 4: a = (uni() - 0.5) * 1e-12 * p; break;               // tiny
      default: a = (uni() - 0.5) * 2e6 * p; break;                 // far away
    }
    n += 2;
    if (!same(tgx::rem_exact(a, p), std::fmod(a, p))) bad++;
    if (!same(tgx::f_modulo(a, p), f_modulo_ref(a, p))) bad++;
  }
  std::printf("%ld %ld\n", bad, n);
  return bad != 0;
}
