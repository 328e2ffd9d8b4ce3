// Checks qdiv (lepic_qdiv.h) against IEEE division, bit for bit, on the
// inputs where a cheaper quotient would go wrong: coordinates spread over a
// box, next to grid nodes (where int(x/h) flips), at the midpoints between
// two consecutive quotients (where rounding flips), and over a wide exponent
// range including zeros, subnormals, infinities and NaN.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../essaitemporaire_lepic_module_amd/csrc/lepic_qdiv.h"

static uint64_t st = 88172645463325252ull;
static uint64_t next64() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; }
static double u01() { return (double)(next64() >> 11) * (1.0 / 9007199254740992.0); }

static bool same(double a, double b) {
  if (std::isnan(a) && std::isnan(b)) return true;
  return std::memcmp(&a, &b, 8) == 0;
}

int main(int argc, char **argv) {
  long n = argc > 1 ? std::atol(argv[1]) : 1000000;
  long bad = 0, checked = 0;
  const double special[] = {0.0, -0.0, 1e-320, -1e-320, 1e-300, INFINITY, -INFINITY, NAN, 1e300,
                            -1e300, 5e-324, 1.0, -1.0};
  for (long t = 0; t < n; ++t) {
    double h;
    switch (t % 4) {
      case 0: h = 3.125e-4; break;
      case 1: h = 0.01 / 32.0 * (1.0 + 1e-3 * (double)(t % 7)); break;
      case 2: h = std::exp(std::log(1e-6) + u01() * std::log(1e6)); break;
      default: h = 1.0 / 3.0;
    }
    if (!qdiv_divisor_ok(h)) continue;
    const QDiv d = qdiv_make(h);
    double x;
    switch ((t / 4) % 6) {
      case 0: x = (u01() * 2.0 - 0.5) * 200.0 * h; break;
      case 1: {
        long k = (long)(u01() * 300.0) - 20;
        x = (double)k * h + (u01() - 0.5) * 1e-13 * h;
        break;
      }
      case 2: {
        long k = (long)(u01() * 300.0) - 20;
        x = std::nextafter((double)k * h, (next64() & 1) ? 1e9 : -1e9);
        break;
      }
      case 3: {
        double q = u01() * 300.0;
        double qn = std::nextafter(q, 1e9);
        double mid = q + (qn - q) / 2.0;
        x = mid * h;
        if (next64() & 1) x = std::nextafter(x, (next64() & 1) ? 1e9 : -1e9);
        break;
      }
      case 4: x = std::ldexp(u01() - 0.5, (int)(next64() % 2000) - 1000); break;
      default: x = special[t % (sizeof(special) / sizeof(special[0]))];
    }
    ++checked;
    if (!same(qdiv(x, d), x / h)) {
      if (++bad < 10) std::printf("mismatch x=%.17g h=%.17g\n", x, h);
    }
    // the per-particle form: one range test for three coordinates, then the
    // quotients x/h and (i h - x)/h of the kernels' stencils
    const double y = special[next64() % (sizeof(special) / sizeof(special[0]))] * (next64() & 1) +
                     (u01() - 0.5) * 300.0 * h;
    const double z = (t % 5 == 0) ? 0.0 : (u01() * 300.0) * h;
    const bool ok = qdiv_coords_ok(x, y, z);
    if (!same(qdiv_pre(x, d, ok), x / h)) {
      if (++bad < 10) std::printf("mismatch (pre) x=%.17g h=%.17g\n", x, h);
    }
    const double q = x / h;
    if (std::isfinite(q) && std::fabs(q) < 1e9) {
      const double r = (double)((long)q + 1) * h - x;
      if (!same(qdiv_pre(r, d, ok), r / h)) {
        if (++bad < 10) std::printf("mismatch (frac) x=%.17g h=%.17g\n", x, h);
      }
    }
  }
  std::printf("checked %ld mismatches %ld\n", checked, bad);
  return bad != 0;
}
